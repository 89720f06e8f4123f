"""GPU adjoint (FP32) against the oracle's FP64 adjoint and against central finite
differences of the FP64 GPU forward."""
import numpy as np
import pytest

import oracle
from smatranem_amd import casegen, fdtd
from tests.cases import small_waveguide

pytestmark = pytest.mark.gpu


def _case(**kw):
    args = dict(seed=3, pml_xy=2, pml_z=3, iBc=(0, 0, 0), metal_cells=0, lossy_cells=0, time_period=(1, 1, 3))
    args.update(kw)
    return small_waveguide(**args)


def _box(case):
    L0, L1, L2 = case.L[:3]
    return (3, L0 - 3, 3, L1 - 3, L2 // 2 - 3, L2 // 2 + 3)


GRIDS = {
    "small": {},
    # several x/y tiles and z chunks of the fused transposed step, Bloch-periodic walls
    "tiles": dict(L0=40, L1=20, L2=48, outer=8, pml_xy=0, pml_z=4, iBc=(1, -1, 0), time_period=(1, 2, 4)),
    # z chunks [32, 64) and [64, 96) reach no PML plane: the split-free instantiation
    "long": dict(L0=40, L1=20, L2=120, outer=20, pml_xy=0, pml_z=4, iBc=(1, -1, 0), time_period=(1, 8, 10)),
}


@pytest.mark.parametrize("kernel,ckpt,grid", [("fused", 7, "small"), ("twopass", 5, "small"),
                                              ("fused", 1000, "small"), ("staged", 7, "small"),
                                              ("two-step", 10, "small"), ("two-step", 7, "small"),
                                              ("two-step", 1000, "small"), ("fused", 0, "small"),
                                              ("two-step", 0, "tiles"), ("fused", 9, "tiles"),
                                              ("two-step", 0, "long")])
def test_adjoint_matches_oracle(kernel, ckpt, grid, monkeypatch):
    """With the two-step kernel selected and an even checkpoint interval the forward pass
    runs two steps per pass and takes the objective's samples inside the pass; with an
    odd interval it runs one step at a time.  ckpt = 0: the design box's E stored after
    every step, no recomputation."""
    monkeypatch.setenv("SMX_F32_KERNEL", "twopass" if kernel == "twopass" else "fused")
    if kernel == "staged":
        monkeypatch.setenv("SMX_F32_TILE", "3,16")
    if kernel == "two-step":
        monkeypatch.setenv("SMX_F32_TWO", "2,8")
    case = _case(**GRIDS[grid])
    prep = casegen.prepare(case)
    zT = min(case.L[2] - 5, case.iI + 50)
    w = 0.9 * 2 * np.pi / prep.N
    K = case.time_period[2] * prep.N
    box = _box(case)
    with prep.plan(fdtd.F32) as p:
        J = p.adjoint_forward(K, ckpt, zT, w, box)
        g = p.adjoint_backward()
    N, iR, iI, _ = oracle.set_tau(case.omega, case.min_dt)
    st = oracle.setup_from_media(case.L, case.mu, case.eps, iR=iR, iI=iI)
    Jo, go = oracle.adjoint_run(case.L, st, case.iBc, (-1, -1, case.iI, 1), case.Hinc, case.Einc, N,
                                case.time_period, zT, w, iR, iI)
    assert abs(J - Jo) <= 1e-4 * Jo
    x0, x1, y0, y1, z0, z1 = box
    go3 = go.reshape(case.L[2], case.L[1], case.L[0], 3)[z0:z1, y0:y1, x0:x1, :].transpose(3, 0, 1, 2)
    err = np.linalg.norm(g - go3) / np.linalg.norm(go3)
    assert err <= 1e-3, err


def test_adjoint_vs_finite_differences():
    """Central differences of J on the FP64 GPU forward (delta eps = 1e-3), on cells
    where the gradient is large enough to resolve."""
    case = _case(seed=5)
    prep = casegen.prepare(case)
    zT = min(case.L[2] - 5, case.iI + 50)
    w = 0.9 * 2 * np.pi / prep.N
    K = case.time_period[2] * prep.N
    box = _box(case)
    with prep.plan(fdtd.F32) as p:
        p.adjoint_forward(K, 10, zT, w, box)
        g = p.adjoint_backward()

    def J64(eps):
        c2 = casegen.Case(**{**case.__dict__, "eps": eps})
        pr = casegen.prepare(c2)
        assert pr.N == prep.N
        with pr.plan(fdtd.F64) as q:
            q.set_dft([zT], [w], chunk=16)
            q.run_steps(0, K)
            D = q.read_dft()[0]
        return float(np.sum(np.abs(D[2:4, 0, :]) ** 2))

    x0, x1, y0, y1, z0, z1 = box
    L0, L1 = case.L[0], case.L[1]
    flat = np.abs(g).reshape(-1)
    order = np.argsort(flat)[::-1][:4]
    for idx in order:
        c, zz, yy, xx = np.unravel_index(idx, g.shape)
        cell = (x0 + xx) + L0 * ((y0 + yy) + L1 * (z0 + zz))
        h = 1e-3
        eps_p = case.eps.copy()
        eps_m = case.eps.copy()
        eps_p[3 * cell + c] += h
        eps_m[3 * cell + c] -= h
        fd = (J64(eps_p) - J64(eps_m)) / (2 * h)
        assert abs(fd - g[c, zz, yy, xx]) <= 1e-3 * abs(fd) + 1e-3 * flat.max(), (fd, g[c, zz, yy, xx])


def test_adjoint_forward_keeps_caller_dft(monkeypatch):
    """The paired forward pass runs its own DFT of the objective plane; a DFT the caller
    set stays set, with its accumulators untouched, and the forward's J is the one of
    the step-by-step forward."""
    monkeypatch.setenv("SMX_F32_TWO", "2,8")
    case = _case()
    prep = casegen.prepare(case)
    zT = min(case.L[2] - 5, case.iI + 50)
    w = 0.9 * 2 * np.pi / prep.N
    K = case.time_period[2] * prep.N
    box = _box(case)
    with prep.plan(fdtd.F32) as p:
        p.set_dft([zT - 1, zT], [w, 1.1 * w], chunk=8)
        J2 = p.adjoint_forward(K, 10, zT, w, box)
        D = p.read_dft()
        assert D.shape == (2, 4, 2, case.plane) and not np.any(D)
        J1 = p.adjoint_forward(K, 7, zT, w, box)  # odd interval: one step at a time
    assert abs(J2 - J1) <= 1e-5 * J1


@pytest.mark.parametrize("grid", list(GRIDS))
def test_fused_transposed_step_equals_gather_kernels(grid, monkeypatch):
    """The fused LDS-staged transposed step against the two gather kernels it replaces:
    the same gradient up to FP32 rounding order (1e-5 relative)."""
    case = _case(**GRIDS[grid])
    prep = casegen.prepare(case)
    zT = min(case.L[2] - 5, case.iI + 50)
    w = 0.9 * 2 * np.pi / prep.N
    K = case.time_period[2] * prep.N
    box = _box(case)
    out = []
    for gather in ("0", "1"):
        monkeypatch.setenv("SMX_ADJ_GATHER", gather)
        with prep.plan(fdtd.F32) as p:
            p.adjoint_forward(K, 0, zT, w, box)
            out.append(p.adjoint_backward())
    assert np.abs(out[1]).max() > 0
    err = np.linalg.norm(out[0] - out[1]) / np.linalg.norm(out[1])
    assert err <= 1e-5, err


TWO_GRIDS = dict(GRIDS, pml_tiles=dict(L0=40, L1=20, L2=60, outer=10, pml_xy=3, pml_z=4, iBc=(0, 0, 0),
                                       time_period=(1, 2, 4)))


@pytest.mark.parametrize("grid,ckpt", [(g, c) for g in TWO_GRIDS for c in (0, 7)])
def test_two_step_transposed_pass_equals_one_step(grid, ckpt, monkeypatch):
    """Two transposed steps per launch (adj_two_f32, with the injection into lambda^k and
    lambda_TE^k of the design box inside the pass) give bit for bit the gradient of one
    fused transposed step per launch, with whole-run box history and with checkpoint
    segments of 7 steps (odd: each segment ends with a single step)."""
    case = _case(**TWO_GRIDS[grid])
    prep = casegen.prepare(case)
    zT = min(case.L[2] - 5, case.iI + 50)
    w = 0.9 * 2 * np.pi / prep.N
    K = case.time_period[2] * prep.N
    box = _box(case)
    out = []
    for two in ("0", "1"):
        monkeypatch.setenv("SMX_ADJ_TWO", two)
        with prep.plan(fdtd.F32) as p:
            p.adjoint_forward(K, ckpt, zT, w, box)
            out.append(p.adjoint_backward())
    assert np.abs(out[0]).max() > 0
    assert np.array_equal(out[0], out[1]), np.abs(out[0] - out[1]).max() / np.abs(out[0]).max()


@pytest.mark.parametrize("ckpt", [0, 8, 9])
def test_box_history_two_step_forward(ckpt, monkeypatch):
    """The design box's E taken inside the two-step forward pass (both steps of each pass;
    the box-history forward, and the recomputation of checkpoint segments, even and odd)
    against one forward step per launch with a box copy after each: the same J and
    gradient up to FP32 rounding order."""
    monkeypatch.setenv("SMX_F32_TWO", "2,16")
    case = _case(**TWO_GRIDS["pml_tiles"])
    prep = casegen.prepare(case)
    zT = min(case.L[2] - 5, case.iI + 50)
    w = 0.9 * 2 * np.pi / prep.N
    K = case.time_period[2] * prep.N - 1  # odd: the last step runs alone
    box = _box(case)
    out = []
    for box2 in ("0", "1"):
        monkeypatch.setenv("SMX_ADJ_BOX2", box2)
        with prep.plan(fdtd.F32) as p:
            J = p.adjoint_forward(K, ckpt, zT, w, box)
            out.append((J, p.adjoint_backward()))
    (J0, g0), (J1, g1) = out
    assert J0 > 0 and abs(J1 - J0) <= 1e-6 * J0, (J0, J1)
    assert np.linalg.norm(g1 - g0) <= 1e-5 * np.linalg.norm(g0), np.linalg.norm(g1 - g0) / np.linalg.norm(g0)
