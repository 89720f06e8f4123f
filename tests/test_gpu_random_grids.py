"""Seeded random small grids through every kernel family, against the oracle.

Each case draws a grid (L0 in [8, 60), L1 in [10, 40), L2 around the port slices), PML
widths (x/y 0-3, z 2-4), metal and lossy cells, wrap weights iBc in {-1, 0, 1}, and:
- FP64 plan: fields, monitors and envelope sums bit for bit the oracle's;
- FP32 one-step kernels (a drawn tile shape and z chunk, with or without the two-step
  pass): relative L2 within 2e-4 of the oracle;
- FP32 two-step pass (split slots, a drawn z chunk, a DFT on or off): within 2e-4 of the
  oracle, and -- on grids without lossy/metal cells -- bit for bit the same with the y
  tiles from row 0;
- a fused batch of 2-4 drawn plans (some with a DFT of a drawn chunk, odd or even step
  counts): each plan bit for bit the plan alone;
- the adjoint with a drawn kernel, checkpoint interval, design box and objective: within
  1e-4 (J) and 1e-3 (gradient) of the oracle's FP64 adjoint."""
import os

import numpy as np
import pytest

import oracle
from smatranem_amd import casegen, fdtd
from tests.cases import rel_l2, small_waveguide

pytestmark = pytest.mark.gpu

ENV = ("SMX_F32_TWO", "SMX_F32_TILE", "SMX_F32_KERNEL", "SMX_F32_PML_ORIGIN")


def _kw(rng, spec=True):
    L0 = int(rng.integers(8, 60))
    L1 = int(rng.integers(10, 40))
    outer = int(rng.integers(7, 10))
    L2 = int(rng.integers(2 * outer + 8, 2 * outer + 36))
    return dict(L0=L0, L1=L1, L2=L2, outer=outer, pml_xy=int(rng.integers(0, 4)), pml_z=int(rng.integers(2, 5)),
                metal_cells=int(rng.integers(0, 4)) if spec else 0, lossy_cells=int(rng.integers(0, 3)) if spec else 0,
                iBc=tuple(int(v) for v in rng.integers(-1, 2, 3)), core=(max(L0 // 6, 1), max(L1 // 6, 1)),
                time_period=(1, 1, 3))


def _oracle(prep):
    c = prep.case
    st = oracle.setup_from_media(c.L, c.mu, c.eps, iR=prep.invtau_R, iI=prep.invtau_I)
    return oracle.run(c.L, st, c.iBc, (c.iB, c.iT, c.iI, c.inc_mode), c.Hinc, c.Einc, prep.N, c.time_period)


@pytest.fixture(autouse=True)
def _clean_env(monkeypatch):
    for k in ENV:
        monkeypatch.delenv(k, raising=False)


@pytest.mark.parametrize("seed", range(8))
def test_random_f64_bitwise(seed):
    rng = np.random.default_rng(1000 + seed)
    prep = casegen.prepare(small_waveguide(seed=1000 + seed, **_kw(rng)))
    Ho, Eo, mo, so = _oracle(prep)
    with prep.plan(fdtd.F64) as p:
        sums = p.run_steps(0, prep.case.time_period[2] * prep.N)
        H, E = p.read_fields()
        mons = p.read_monitors()
    assert np.abs(Eo).max() > 0
    assert np.array_equal(H, Ho) and np.array_equal(E, Eo)
    for a, b in zip(mons, mo):
        assert np.array_equal(a.view(np.float64), b)
    assert np.allclose(sums, so, rtol=1e-14, atol=0)


@pytest.mark.parametrize("seed", range(8))
def test_random_f32_one_step_kernels(seed, monkeypatch):
    rng = np.random.default_rng(2000 + seed)
    prep = casegen.prepare(small_waveguide(seed=2000 + seed, **_kw(rng)))
    tile, fz = int(rng.integers(0, 8)), int(rng.choice([8, 16, 32, 64]))
    monkeypatch.setenv("SMX_F32_KERNEL", "fused")
    monkeypatch.setenv("SMX_F32_TILE", f"{tile},{fz}")
    monkeypatch.setenv("SMX_F32_TWO", "off")
    Ho, Eo, _, _ = _oracle(prep)
    steps = prep.case.time_period[2] * prep.N
    with prep.plan(fdtd.F32) as p:
        p.run_steps(0, steps)
        H, E = p.read_fields()
    assert np.abs(Eo).max() > 0
    assert rel_l2(H, Ho) < 2e-4 and rel_l2(E, Eo) < 2e-4, (tile, fz, rel_l2(H, Ho), rel_l2(E, Eo))


@pytest.mark.parametrize("seed,spec", [(s, s % 2 == 0) for s in range(10)])
def test_random_f32_two_step(seed, spec, monkeypatch):
    rng = np.random.default_rng(3000 + seed)
    prep = casegen.prepare(small_waveguide(seed=3000 + seed, **_kw(rng, spec)))
    fz = int(rng.choice([5, 8, 16, 64]))
    dft = bool(rng.random() < 0.5)
    steps = prep.case.time_period[2] * prep.N
    L2 = int(prep.case.L[2])
    out = []
    for origin in ("1", "0"):
        monkeypatch.setenv("SMX_F32_KERNEL", "fused")
        monkeypatch.setenv("SMX_F32_TWO", f"2,{fz}")
        monkeypatch.setenv("SMX_F32_PML_ORIGIN", origin)
        with prep.plan(fdtd.F32) as p:
            if dft:
                p.set_dft([prep.case.iI + 2, L2 // 2], [0.3, 2 * np.pi / prep.N], chunk=8)
            p.run_steps(0, steps - 1)
            p.run_steps(steps - 1, 1)
            out.append(p.read_fields() + ((p.read_dft(),) if dft else ()))
    Ho, Eo, _, _ = _oracle(prep)
    assert np.abs(Eo).max() > 0
    assert rel_l2(out[0][0], Ho) < 2e-4 and rel_l2(out[0][1], Eo) < 2e-4
    for a, b in zip(out[0], out[1]):
        if spec:
            assert rel_l2(a, b) < 1e-5
        else:
            assert np.array_equal(a, b)


@pytest.mark.parametrize("seed", range(6))
def test_random_fused_batch(seed, monkeypatch):
    rng = np.random.default_rng(4000 + seed)
    m = int(rng.integers(2, 5))
    preps = [casegen.prepare(small_waveguide(seed=4000 + 10 * seed + j, **_kw(rng, spec=False))) for j in range(m)]
    monkeypatch.setenv("SMX_F32_TWO", "2,16")
    steps = min(pr.case.time_period[2] * pr.N for pr in preps) - int(rng.integers(0, 2))
    dft = [bool(rng.random() < 0.5) for _ in preps]
    chunks = [int(rng.choice([2, 7, 8, 16])) for _ in preps]

    def setup(p, pr, j):
        if dft[j]:
            p.set_dft([pr.case.iI + 2, int(pr.case.L[2]) // 2], [0.3, 2 * np.pi / pr.N], chunk=chunks[j])
    alone = []
    for j, pr in enumerate(preps):
        with pr.plan(fdtd.F32) as p:
            setup(p, pr, j)
            sums = p.run_steps(0, steps)
            alone.append((p.read_fields(), p.read_monitors(), p.read_dft() if dft[j] else None, sums))
    plans = [pr.plan(fdtd.F32) for pr in preps]
    try:
        for j, (p, pr) in enumerate(zip(plans, preps)):
            setup(p, pr, j)
        info = {}
        sums = fdtd.run_batch(plans, 0, steps, mode=fdtd.BATCH_FUSED, info=info)
        together = [(p.read_fields(), p.read_monitors(), p.read_dft() if dft[j] else None)
                    for j, p in enumerate(plans)]
    finally:
        for p in plans:
            p.close()
    assert info["graph_launches"] == steps // 2
    for j, (a, b) in enumerate(zip(alone, together)):
        assert np.abs(a[0][1]).max() > 0
        assert np.array_equal(a[0][0], b[0][0]) and np.array_equal(a[0][1], b[0][1])
        for x, y in zip(a[1], b[1]):
            assert np.array_equal(x, y)
        if dft[j]:
            assert np.array_equal(a[2], b[2])
        assert np.array_equal(a[3], sums[j])


@pytest.mark.parametrize("seed", range(6))
def test_random_adjoint(seed, monkeypatch):
    """The FP32 adjoint on a drawn grid (no lossy/metal cells), step kernel, checkpoint
    interval (0: design-box history), design box, objective plane and frequency: J within
    1e-4 and dJ/d eps within 1e-3 (relative L2) of the oracle's FP64 adjoint."""
    rng = np.random.default_rng(5000 + seed)
    L0, L1, outer = int(rng.integers(12, 48)), int(rng.integers(12, 36)), int(rng.integers(7, 10))
    L2 = int(rng.integers(2 * outer + 10, 2 * outer + 40))
    pml_xy, pml_z = int(rng.integers(0, 4)), int(rng.integers(2, 5))
    iBc = tuple(int(v) for v in rng.integers(-1, 2, 3)) if pml_xy == 0 else (0, 0, 0)
    case = small_waveguide(seed=5000 + seed, L0=L0, L1=L1, L2=L2, outer=outer, pml_xy=pml_xy, pml_z=pml_z, iBc=iBc,
                           metal_cells=0, lossy_cells=0, core=(max(L0 // 6, 1), max(L1 // 6, 1)),
                           time_period=(1, int(rng.integers(1, 3)), 3))
    prep = casegen.prepare(case)
    kern = rng.choice(["fused", "twopass", "staged", "two-step"])
    ckpt = int(rng.choice([0, 1, 5, 8, 10, 1000]))
    monkeypatch.setenv("SMX_F32_KERNEL", "twopass" if kern == "twopass" else "fused")
    if kern == "staged":
        monkeypatch.setenv("SMX_F32_TILE", "3,16")
    if kern == "two-step":
        monkeypatch.setenv("SMX_F32_TWO", "2,%d" % int(rng.choice([8, 16, 64])))
    m = max(pml_xy, 1) + 1
    x0, y0 = int(rng.integers(m, L0 // 2)), int(rng.integers(m, L1 // 2))
    x1, y1 = int(rng.integers(x0 + 1, L0 - m + 1)), int(rng.integers(y0 + 1, L1 - m + 1))
    z0 = int(rng.integers(outer, L2 // 2))
    box = (x0, x1, y0, y1, z0, int(rng.integers(z0 + 1, L2 - outer)))
    zT = int(rng.integers(case.iI + 3, L2 - pml_z - 1))
    w = float(rng.uniform(0.8, 1.1)) * 2 * np.pi / prep.N
    K = case.time_period[2] * prep.N
    with prep.plan(fdtd.F32) as p:
        J = p.adjoint_forward(K, ckpt, zT, w, box)
        g = p.adjoint_backward()
    N, iR, iI, _ = oracle.set_tau(case.omega, case.min_dt)
    st = oracle.setup_from_media(case.L, case.mu, case.eps, iR=iR, iI=iI)
    Jo, go = oracle.adjoint_run(case.L, st, case.iBc, (-1, -1, case.iI, 1), case.Hinc, case.Einc, N,
                                case.time_period, zT, w, iR, iI)
    assert Jo > 0 and abs(J - Jo) <= 1e-4 * Jo
    go3 = go.reshape(L2, L1, L0, 3)[box[4]:box[5], y0:y1, x0:x1, :].transpose(3, 0, 1, 2)
    assert np.linalg.norm(g - go3) <= 1e-3 * np.linalg.norm(go3), (kern, ckpt, box)
