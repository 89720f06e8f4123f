"""Multi-frequency port DFT (smx_plan_set_dft), against the reference monitors and a
direct DFT of the oracle's per-step plane samples."""
import numpy as np
import pytest

import oracle
from smatranem_amd import casegen, fdtd
from tests.cases import rel_l2, small_waveguide

pytestmark = pytest.mark.gpu

CASE = dict(pml_xy=0, pml_z=3, iBc=(1, -1, 0), metal_cells=4, lossy_cells=3)


def _prep():
    return casegen.prepare(small_waveguide(seed=6, **CASE))


@pytest.mark.parametrize("chunk", [1, 7, 32, 64, 200, 256])
def test_dft_at_omega_equals_reference_monitors(chunk):
    prep = _prep()
    c = prep.case
    w = 2 * np.pi / prep.N
    with prep.plan(fdtd.F64) as p:
        p.set_dft([c.iB, c.iT], [w, 0.5 * w], chunk=chunk)
        p.run(0, c.time_period[2])
        HB, EB, HT, ET = p.read_monitors()
        D = p.read_dft()
    for m, (Hm, Em) in enumerate(((HB, EB), (HT, ET))):
        ref = np.stack([Hm[0::2], Hm[1::2], Em[0::2], Em[1::2]])
        assert np.linalg.norm(ref) > 0
        assert rel_l2(D[m, :, 0, :], ref) < 1e-12


def _oracle_samples(prep, nsteps):
    """Per-step main-slot samples (Hx, Hy, Ex, Ey) of the monitor planes from the oracle."""
    c = prep.case
    N, iR, iI, _ = oracle.set_tau(c.omega, c.min_dt)
    st = oracle.setup_from_media(c.L, c.mu, c.eps, iR=iR, iI=iI)
    n, plane = c.n, c.plane
    flen = 3 * (n + st["Lnotn"] + 1)
    H, E = np.zeros(flen), np.zeros(flen)
    mons = [np.zeros(4 * plane) for _ in range(4)]
    ExpN = oracle.expn_table(N)
    Hinc = np.ascontiguousarray(c.Hinc).view(np.float64)
    Einc = np.ascontiguousarray(c.Einc).view(np.float64)
    out = np.zeros((nsteps, 2, 4, plane))
    for k in range(nsteps):
        t, P = k % N, k // N
        tn = (t + 1) % N
        fE = 2.0 * oracle.set_env(N, c.time_period, 1, t, P)
        fH = 2.0 * oracle.set_env(N, c.time_period, 0, t + 1, P)
        Ei = fE * (ExpN[4 * t + 2] * Einc[0::2] + ExpN[4 * t + 3] * Einc[1::2])
        Hi = fH * (ExpN[4 * tn] * Hinc[0::2] + ExpN[4 * tn + 1] * Hinc[1::2])
        oracle.main_calc(oracle.load_oracle().orc_main_calc, c.L, st["Ah"], st["Bh"], st["Ae"], st["Be"], c.iBc,
                         H, E, (c.iB, c.iT, c.iI, c.inc_mode), np.ascontiguousarray(Hi),
                         np.ascontiguousarray(Ei), ExpN[4 * tn:4 * tn + 4].copy(), *mons, st["Cont"], st["addr"])
        for m, z in enumerate((c.iB, c.iT)):
            base = 3 * plane * z
            out[k, m, 0] = H[base + 0:base + 3 * plane:3]
            out[k, m, 1] = H[base + 1:base + 3 * plane:3]
            out[k, m, 2] = E[base + 0:base + 3 * plane:3]
            out[k, m, 3] = E[base + 1:base + 3 * plane:3]
    return out


@pytest.mark.parametrize("prec,tol", [(fdtd.F64, 1e-12), (fdtd.F32, 2e-4)])
def test_dft_multifrequency_vs_direct(prec, tol, monkeypatch):
    prep = _prep()
    c = prep.case
    nsteps = 3 * prep.N + 5
    ws = np.array([0.1, 2 * np.pi / prep.N, 0.37, 1.1])
    S = _oracle_samples(prep, nsteps)
    k = np.arange(nsteps)
    ref = np.zeros((2, 4, len(ws), c.plane), dtype=np.complex128)
    for fi, w in enumerate(ws):
        pH = np.exp(1j * w * (k + 1))
        pE = np.exp(1j * w * (k + 1.5))
        for v in range(4):
            ph = pH if v < 2 else pE
            ref[:, v, fi, :] = np.einsum("k,kmq->mq", ph, S[:, :, v, :])
    with prep.plan(prec) as p:
        p.set_dft([c.iB, c.iT], ws, chunk=16)
        p.run_steps(0, nsteps)
        D = p.read_dft()
    assert rel_l2(D, ref) < tol


def test_dft_argument_checks():
    prep = _prep()
    with prep.plan(fdtd.F32) as p:
        for planes, ws, chunk in (([prep.case.L[2]], [0.1], 4), ([0], [0.1] * 40, 4), ([0], [0.1], 0), ([0], [0.1], 257)):
            with pytest.raises(fdtd.SmxError):
                p.set_dft(planes, ws, chunk)
        p._dft = (1, 1)
        with pytest.raises(fdtd.SmxError):
            p.read_dft()


LOSSLESS = dict(pml_xy=0, pml_z=3, iBc=(1, -1, 0), metal_cells=0, lossy_cells=0)


@pytest.mark.parametrize("two,chunk", [(None, 16), ("2,16", 16), ("2,16", 7), ("2,16", 64), ("2,8", 1), ("0,16", 16),
                                       ("3,16", 16), ("2,16", 30), (None, 256), ("2,16", 256)])
def test_dft_with_two_step_passes(two, chunk, monkeypatch):
    """With the two-step kernel selected the DFT samples of both steps of a pass come
    from the kernel itself (split-slot variant) or the steps run one at a time (the
    variants with region sweeps, or a ring of one slot); an odd step count ends on a
    single step.  Against a direct DFT of the oracle's per-step samples."""
    if two:
        monkeypatch.setenv("SMX_F32_TWO", two)
    prep = casegen.prepare(small_waveguide(seed=8, **LOSSLESS))
    c = prep.case
    nsteps = 3 * prep.N + 5
    ws = np.array([0.1, 2 * np.pi / prep.N, 0.9])
    S = _oracle_samples(prep, nsteps)
    k = np.arange(nsteps)
    ref = np.zeros((2, 4, len(ws), c.plane), dtype=np.complex128)
    for fi, w in enumerate(ws):
        for v in range(4):
            ph = np.exp(1j * w * (k + 1 + (0.5 if v >= 2 else 0.0)))
            ref[:, v, fi, :] = np.einsum("k,kmq->mq", ph, S[:, :, v, :])
    with prep.plan(fdtd.F32) as p:
        p.set_dft([c.iB, c.iT], ws, chunk=chunk)
        # the split-slot variant with a ring of two or more slots samples in the kernel
        fused = two is not None and two.startswith("2,") and chunk >= 2
        assert p.kernels.endswith(",true,false>") == fused, p.kernels
        p.run_steps(0, 7)
        p.run_steps(7, nsteps - 7)
        D = p.read_dft()
    assert rel_l2(D, ref) < 2e-4


@pytest.mark.parametrize("kernel,planes", [
    ({"SMX_F32_TWO": "2,16"}, "other"),          # two-step pass, monitor planes appended to the ring
    ({"SMX_F32_TWO": "2,16"}, "monitors"),       # DFT planes are the monitor planes
    ({"SMX_F32_TWO": "2,16"}, "one"),            # one monitor plane among the DFT planes
    ({"SMX_F32_KERNEL": "fused", "SMX_F32_TILE": "3,16"}, "other"),   # one step per pass
])
def test_monitors_from_ring(kernel, planes, monkeypatch):
    """FP32 plans with the DFT on take the reference monitors from the DFT ring, folded
    with the reference's phases (e^{i dphi ((k+1) mod N)}, and + 1/2 for E): against the
    same plan without the DFT, whose kernels accumulate them per step.  The run mixes
    smx_plan_run_steps, one smx_plan_step with the caller's phases (accumulated in the
    kernel, not folded twice), a read in the middle of a ring chunk, a write of the
    monitors, and an odd step count; fields agree bit for bit, monitors to rounding."""
    for k, v in kernel.items():
        monkeypatch.setenv(k, v)
    prep = casegen.prepare(small_waveguide(seed=8, **LOSSLESS))
    c = prep.case
    dft_planes = {"other": [c.iB + 1, c.iT - 1], "monitors": [c.iB, c.iT], "one": [c.iT, c.iB + 3]}[planes]
    rng = np.random.default_rng(3)
    Hi, Ei = rng.normal(size=2 * c.plane), rng.normal(size=2 * c.plane)
    ExpN2 = np.exp(1j * np.array([0.3, 0.7]))
    nsteps = 3 * prep.N + 5
    out = []
    for dft in (False, True):
        with prep.plan(fdtd.F32) as p:
            if dft:
                p.set_dft(dft_planes, [0.2, 2 * np.pi / prep.N], chunk=12)
            p.run_steps(0, 9)
            p.step(Hi, Ei, ExpN2)                  # step 9 with the caller's phases
            p.run_steps(10, 11)
            mid = p.read_monitors()               # inside a ring chunk
            p.write_monitors(*[m * 0.5 for m in mid])
            p.run_steps(21, nsteps - 21)
            out.append((p.read_fields(), mid, p.read_monitors(), p.kernels))
    (F0, mid0, mon0, k0), (F1, mid1, mon1, k1) = out
    print(k0, "|", k1, flush=True)
    assert np.array_equal(F0[0], F1[0]) and np.array_equal(F0[1], F1[1])
    for a, b in list(zip(mid0, mid1)) + list(zip(mon0, mon1)):
        assert np.abs(a).max() > 0
        assert np.abs(a - b).max() <= 1e-12 * np.abs(a).max(), np.abs(a - b).max() / np.abs(a).max()


def test_dft_f32_samples_metal_cell_in_pml_as_f64():
    """A metal cell in a PML layer is class 'e' (the metal class overrides 'a', F:188-233)
    and its DFT sample is the main slot, not main + aux: the FP32 sample follows the
    cell's class, as the FP64 one does.  The cell sits on the first cell of the +x PML
    layer, where sigma_x is zero at the integer sample and positive at the half one, so
    remove_sigma (F:822-853) accepts a lossy metal eps_x there."""
    case = small_waveguide(seed=6, pml_xy=2, pml_z=3, iBc=(0, 0, 0), metal_cells=0, lossy_cells=0)
    L0, L1, L2 = (int(v) for v in case.L[:3])
    zd, x, y = L2 // 2, L0 - 2, L1 // 2
    q = x + L0 * y
    l = q + L0 * L1 * zd
    case.eps[3 * l] = -10.0 + 1.5j
    prep = casegen.prepare(case)
    assert chr(prep.Cont[l]) == "e"
    assert prep.sigma[2 * x] == 0.0 and prep.sigma[2 * x + 1] > 0.0
    w = 2 * np.pi / prep.N
    D = {}
    for prec in (fdtd.F64, fdtd.F32):
        with prep.plan(prec) as p:
            p.set_dft([zd], [w, 0.7 * w], chunk=16)
            p.run(0, case.time_period[2])
            D[prec] = p.read_dft()
    ref = D[fdtd.F64]
    scale = np.abs(ref).max()
    assert scale > 0 and np.abs(ref[..., q]).max() > 1e-3 * scale
    assert rel_l2(D[fdtd.F32], ref) < 2e-4
    assert np.abs(D[fdtd.F32][..., q] - ref[..., q]).max() <= 2e-4 * scale
