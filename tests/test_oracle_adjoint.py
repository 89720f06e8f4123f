"""The oracle's discrete adjoint: the dot-product test pins the transposed step, and
central finite differences of the FP64 forward pin the gradient."""
import numpy as np
import pytest

import oracle
from smatranem_amd import casegen

orc = oracle.load_oracle()


def _case(seed=0, pml_xy=2, **kw):
    args = dict(L0=10, L1=8, L2=20, outer=5, pml_xy=pml_xy, pml_z=3, omega=2 * np.pi / 10.0, core=(3, 2),
                seed=seed, iBc=(0, 0, 0), time_period=(1, 1, 3), metal_cells=0, lossy_cells=0)
    args.update(kw)
    return casegen.waveguide_case(**args)


def _setup(case):
    N, iR, iI, _ = oracle.set_tau(case.omega, case.min_dt)
    st = oracle.setup_from_media(case.L, case.mu, case.eps, iR=iR, iI=iI)
    return N, iR, iI, st


@pytest.mark.parametrize("seed,iBc", [(0, (0, 0, 0)), (1, (1, -1, 0)), (2, (-1, 1, 1))])
def test_step_adjoint_dot_product(seed, iBc):
    case = _case(seed=seed, iBc=iBc, lossy_cells=3)
    N, iR, iI, st = _setup(case)
    n, lnotn = case.n, st["Lnotn"]
    flen = 3 * (n + lnotn + 1)
    rng = np.random.default_rng(seed)
    x = [rng.normal(size=flen) for _ in range(2)]
    y = [rng.normal(size=flen) for _ in range(2)]
    for a in x + y:
        a[3 * (n + lnotn):] = 0.0
    plane = case.plane
    z = np.zeros(2 * plane)
    mons = [np.zeros(4 * plane) for _ in range(4)]
    Mx = [a.copy() for a in x]
    oracle.main_calc(orc.orc_main_calc, case.L, st["Ah"], st["Bh"], st["Ae"], st["Be"], iBc, Mx[0], Mx[1],
                     (-1, -1, case.iI, 1), z, z, np.zeros(4), *mons, st["Cont"], st["addr"])
    MTy = [a.copy() for a in y]
    assert orc.orc_step_adjoint(oracle.i32(case.L), st["Ah"], st["Bh"], st["Ae"], st["Be"], oracle.i32(iBc),
                                MTy[0], MTy[1], oracle.cont_bytes(st["Cont"]), oracle.i32(st["addr"]),
                                lnotn) == 0
    lhs = np.dot(Mx[0], y[0]) + np.dot(Mx[1], y[1])
    rhs = np.dot(x[0], MTy[0]) + np.dot(x[1], MTy[1])
    assert abs(lhs - rhs) <= 1e-12 * (abs(lhs) + 1.0)


def _J(case, eps_override=None):
    c = case
    N, iR, iI, _ = oracle.set_tau(c.omega, c.min_dt)
    eps = c.eps if eps_override is None else eps_override
    st = oracle.setup_from_media(c.L, c.mu, eps, iR=iR, iI=iI)
    return st, N, iR, iI


@pytest.mark.parametrize("pml_xy,lossy", [(2, 0), (0, 2)])
def test_adjoint_gradient_vs_central_differences(pml_xy, lossy):
    case = _case(seed=3, pml_xy=pml_xy, lossy_cells=lossy)
    st, N, iR, iI = _J(case)
    zT = case.L[2] - 5
    w = 2 * np.pi / N * 0.9
    iBTI = (-1, -1, case.iI, 1)
    J, grad = oracle.adjoint_run(case.L, st, case.iBc, iBTI, case.Hinc, case.Einc, N, case.time_period,
                                 zT, w, iR, iI)
    assert J > 0
    rng = np.random.default_rng(1)
    cand = [i for i in range(3 * case.n) if st["Cont"][i // 3] == ord("n")]
    picks = rng.choice(cand, size=6, replace=False)
    # plus the largest-gradient component, so the check is not all noise
    picks = list(picks) + [int(cand[np.argmax(np.abs(grad[cand]))])]
    for i in picks:
        h = 1e-5 * abs(case.eps[i].real)
        Js = []
        for sgn in (1, -1):
            eps = case.eps.copy()
            eps[i] += sgn * h
            stp, N2, iR2, iI2 = _J(case, eps)
            assert N2 == N
            Jp, _ = oracle.adjoint_run(case.L, stp, case.iBc, iBTI, case.Hinc, case.Einc, N, case.time_period,
                                       zT, w, iR, iI)
            Js.append(Jp)
        fd = (Js[0] - Js[1]) / (2 * h)
        scale = np.abs(grad[cand]).max()
        assert abs(fd - grad[i]) <= 1e-6 * scale + 1e-7 * abs(fd), (i, fd, grad[i])
