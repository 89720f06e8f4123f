"""Closed-form pins of the oracle: checks that follow from the mathematics, chosen so a
dropped term, a wrong sign or index, or a transposed operand fails one of them.

* curls: brute force against a numpy Yee curl on small grids, every wrap weight;
  adjointness <R_fwd H, E> = <H, R_bwd E>.
* time constants (set_tau): the discrete d/dt and average are exact at omega.
* envelope (set_env): its sum over a whole run is 1 (the reference prints it, F:168).
* coefficients (build_coeffs, F:106-143 -- inline in main, so not callable in the
  reference): each cell class, driven at omega on a one-cell grid through the oracle's
  own main_calc, settles to the frequency-domain response the scheme is exact for.
* physics: zero source gives zero fields; flipping the incident mode's sign flips every
  field exactly; lossless periodic vacuum conserves the discrete Yee energy.
"""
import numpy as np
import pytest

import oracle
from tests.cases import random_state

orc = oracle.load_oracle()


def _np_curl(F, L, iBc, forward):
    """Yee curl of F (shape (L2, L1, L0, 3)) with the wrap weights iBc."""
    def shift(a, axis_d, fwd):
        ax = 2 - axis_d  # array axis of grid axis d
        out = np.roll(a, -1 if fwd else 1, axis=ax)
        idx = [slice(None)] * 3
        idx[ax] = -1 if fwd else 0
        out[tuple(idx)] *= iBc[axis_d]
        return out
    W = np.zeros_like(F)
    for c in range(3):
        c1, c2 = (c + 1) % 3, (c + 2) % 3
        if forward:
            W[..., c] = (shift(F[..., c2], c1, True) - F[..., c2]) - (shift(F[..., c1], c2, True) - F[..., c1])
        else:
            W[..., c] = (F[..., c2] - shift(F[..., c2], c1, False)) - (F[..., c1] - shift(F[..., c1], c2, False))
    return W


def _unit_state(L, iBc, seed):
    """All 'n' cells, A = B = 1, no source, monitors on plane 0."""
    rng = np.random.default_rng(seed)
    n = L[0] * L[1] * L[2]
    ones = np.ones(3 * n)
    return dict(L=oracle.i32([*L, 1]), Ah=ones, Bh=ones, Ae=ones, Be=ones, iBc=oracle.i32(iBc),
                iBTI=oracle.i32([0, 0, -5 if False else L[2] + 7, 1]), Cont=np.full(n, ord("n"), np.int8),
                addr=np.full(n, 0, np.int32), rng=rng, n=n)


@pytest.mark.parametrize("iBc", [(1, 1, 1), (-1, 0, 1), (0, 0, 0), (1, -1, -1)])
@pytest.mark.parametrize("L", [(4, 3, 5), (1, 4, 3), (5, 1, 1)])
def test_curls_brute_force(L, iBc):
    s = _unit_state(L, iBc, 1)
    n = s["n"]
    plane = L[0] * L[1]
    Hi = np.zeros(2 * plane)
    mons = [np.zeros(4 * plane) for _ in range(4)]
    E0 = s["rng"].normal(size=3 * (n + 1))
    E0[3 * n:] = 0
    H = np.zeros(3 * (n + 1))
    E = E0.copy()
    oracle.main_calc(orc.orc_main_calc, s["L"], s["Ah"], s["Bh"], s["Ae"], s["Be"], s["iBc"], H, E,
                     s["iBTI"], Hi, Hi, np.zeros(4), *mons, s["Cont"], s["addr"])
    Ef = E0[:3 * n].reshape(L[2], L[1], L[0], 3)
    Hexp = -_np_curl(Ef, L, iBc, forward=False)
    np.testing.assert_allclose(H[:3 * n].reshape(L[2], L[1], L[0], 3), Hexp, rtol=0, atol=1e-14)
    Eexp = Ef + _np_curl(Hexp, L, iBc, forward=True)
    np.testing.assert_allclose(E[:3 * n].reshape(L[2], L[1], L[0], 3), Eexp, rtol=0, atol=1e-13)


@pytest.mark.parametrize("iBc", [(1, 1, 1), (-1, 1, 0), (0, 0, 0)])
def test_curl_adjointness(iBc):
    """The forward curl is the transpose of the backward one: <R H, E> = <H, R^T E>."""
    L = (5, 4, 6)
    rng = np.random.default_rng(3)
    H = rng.normal(size=(L[2], L[1], L[0], 3))
    E = rng.normal(size=(L[2], L[1], L[0], 3))
    # extract both curls through the oracle, as in the brute-force test
    s = _unit_state(L, iBc, 0)
    n = s["n"]
    plane = L[0] * L[1]
    z = np.zeros(2 * plane)
    mons = [np.zeros(4 * plane) for _ in range(4)]
    Hs = np.zeros(3 * (n + 1))
    Es = np.concatenate([E.reshape(-1), np.zeros(3)])
    oracle.main_calc(orc.orc_main_calc, s["L"], s["Ah"], s["Bh"], s["Ae"], s["Be"], s["iBc"], Hs, Es,
                     s["iBTI"], z, z, np.zeros(4), *mons, s["Cont"], s["addr"])
    RbE = -Hs[:3 * n]
    H2 = np.concatenate([H.reshape(-1), np.zeros(3)])
    E2 = np.zeros(3 * (n + 1))
    Bh0 = np.zeros(3 * n)  # H frozen: H_new = H
    oracle.main_calc(orc.orc_main_calc, s["L"], s["Ah"], Bh0, s["Ae"], s["Be"], s["iBc"], H2, E2,
                     s["iBTI"], z, z, np.zeros(4), *mons, s["Cont"], s["addr"])
    RfH = E2[:3 * n]
    assert abs(np.dot(RfH, E.reshape(-1)) - np.dot(H.reshape(-1), RbE)) < 1e-11


@pytest.mark.parametrize("omega,min_dt", [(2 * np.pi * 0.1, 0.57), (2 * np.pi / 20, 0.84), (1.3, 0.2)])
def test_set_tau_exact_at_omega(omega, min_dt):
    N, iR, iI, dt = oracle.set_tau(omega, min_dt)
    dphi = omega * dt
    assert dt <= min_dt and N * dt * omega == pytest.approx(2 * np.pi, rel=1e-14)
    e = np.exp(-1j * dphi)
    assert abs((e - 1) * iR - (-1j * omega * np.exp(-0.5j * dphi))) < 1e-13 * omega
    assert abs((e + 1) * iI - omega * np.exp(-0.5j * dphi)) < 1e-13 * omega


@pytest.mark.parametrize("N,tp", [(18, (10, 80, 111)), (24, (2, 1, 8)), (7, (3, 4, 12))])
def test_envelope_sums_to_one(N, tp):
    s00 = sum(oracle.set_env(N, tp, 0, t + 1, P) for P in range(tp[2]) for t in range(N))
    s05 = sum(oracle.set_env(N, tp, 1, t, P) for P in range(tp[2]) for t in range(N))
    assert s00 == pytest.approx(1.0, abs=1e-12) and s05 == pytest.approx(1.0, abs=1e-12)
    assert np.isnan(oracle.set_env(N, tp, 0, 0, tp[2]))


def _one_cell(cls, mu, eps, sig6, iR, iI):
    """Coefficients of a 1x1x1 grid of class `cls` from orc_build_coeffs."""
    L = oracle.i32([1, 1, 1, 0])
    lnotn = 0 if cls == "n" else 1
    cont = np.array([ord(cls)], np.int8)
    addr = np.array([0 if lnotn else 0], np.int32)
    sigma = np.zeros(6)
    sigma[:] = sig6  # (x int, x half, y int, y half, z int, z half)
    mu3 = np.array([mu] * 3, dtype=np.complex128).view(np.float64)
    eps3 = np.array([eps] * 3, dtype=np.complex128).view(np.float64)
    co = [np.zeros(3 * (1 + lnotn)) for _ in range(4)]
    orc.orc_build_coeffs(L, mu3, eps3, sigma, cont, addr, iR, iI, *co)
    return L, cont, addr, lnotn, co


def _drive_e(cls, mu, eps, sig6, omega=2 * np.pi / 16, min_dt=0.5, periods=400):
    """Drive a one-cell grid (PEC wrap, so W[c] = H[c1] - H[c2] for the forward curl and
    H[c2] - H[c1] for the backward one) with H held at a sinusoid; return the steady E
    phasor per unit forcing phasor, fitted over the last periods."""
    N, iR, iI, dt = oracle.set_tau(omega, min_dt)
    L, cont, addr, lnotn, (Ah, Bh, Ae, Be) = _one_cell(cls, mu, eps, sig6, iR, iI)
    flen = 3 * (1 + lnotn + 1)
    H = np.zeros(flen)
    E = np.zeros(flen)
    z2 = np.zeros(2)
    mons = [np.zeros(4) for _ in range(4)]
    Bh0 = np.zeros_like(Bh)          # H is set by the test, not updated
    Ah1 = np.ones_like(Ah)
    if cls in "hb":
        Bh0[3:] = 0.0
        Ah1[3:] = 1.0
    hist = []
    total = periods * N
    for k in range(total):
        # H at time k+1/2 (the E update of step k uses H^{k+1/2}); H_y carries the drive
        hval = np.cos(omega * (k + 0.5) * dt)
        H[:3] = 0.0
        H[1] = hval
        oracle.main_calc(orc.orc_main_calc, L, Ah1, Bh0, Ae, Be, oracle.i32([0, 0, 0]), H, E,
                         oracle.i32([0, 0, 5, 1]), z2, z2, np.zeros(4), *mons, cont, addr)
        hist.append(E[0])
    t = np.arange(total - 8 * N, total)
    y = np.array(hist[-8 * N:])
    # E^n = Re(Ehat e^{-i omega n dt}) + c: least squares over the last 8 periods
    A = np.stack([np.cos(omega * (t + 1) * dt), np.sin(omega * (t + 1) * dt), np.ones_like(t, dtype=float)], 1)
    cr, ci, _ = np.linalg.lstsq(A, y, rcond=None)[0]
    Ehat = cr + 1j * ci
    return Ehat, iR, iI, omega


def test_lossy_dielectric_cell_matches_frequency_domain():
    """E cell, class 'n', eps = 2 + 0.5i: -i*omega*eps*Ehat = What (What the forcing
    phasor of W_x = H_y - H_z = H_y at time n + 1/2)."""
    eps = 2.0 + 0.5j
    Ehat, iR, iI, omega = _drive_e("n", 1.0 + 0j, eps, np.zeros(6))
    W = 1.0  # W_x = H[c1=y] - H[c2=z] for the forward curl with zero wrap weights
    assert abs(-1j * omega * eps * Ehat - W) < 1e-9


def test_pml_cell_matches_stretched_response():
    """E cell, class 'a', sigma on the main slot's axis (y for E_x; half samples):
    the main slot follows -i*omega*eps*(1 + i*sigma)*Ehat_main = What_+ with
    What_+ = H_z part -> here the '+' half carries -H[c2] = 0 and the '-' half +H[c1]."""
    sig = 0.7
    sig6 = [0, 0, 0, sig, 0, 0]  # y half sample
    # E_x main slot: '+' half = w*H_z(l+e_y) - H_z(l) = 0 (H_z = 0): no drive;
    # split slot (sigma of axis z = 0) gets the '-' half = -w*H_y(l+e_z) + H_y(l) = H_y.
    # The total E_x = main + split therefore follows the unstretched response.
    N, iR, iI, dt = oracle.set_tau(2 * np.pi / 16, 0.5)
    L, cont, addr, lnotn, (Ah, Bh, Ae, Be) = _one_cell("a", 1.0 + 0j, 2.0 + 0j, sig6, iR, iI)
    # closed forms of the two slots' coefficients from the scheme that is exact at omega
    e_main = 2.0 * (1 + 1j * sig)
    a_m, b_m = Ae[0], Be[0]
    a_s, b_s = Ae[3], Be[3]
    B_exp = lambda er, s: 1.0 / (iR * er + iI * er * s)
    A_exp = lambda er, s: (iR * er - iI * er * s) * B_exp(er, s)
    assert b_m == pytest.approx(B_exp(2.0, sig), rel=1e-15) and a_m == pytest.approx(A_exp(2.0, sig), rel=1e-15)
    assert b_s == pytest.approx(B_exp(2.0, 0.0), rel=1e-15) and a_s == pytest.approx(1.0, abs=2e-16)
    # and the slot scheme X^{n+1} = A X^n + B W is exact at omega for eps*(1 + i*sigma):
    omega = 2 * np.pi / 16
    z = np.exp(-1j * omega * dt)
    # X(z - A) = B * W z^{1/2}  =>  X/W = B z^{1/2} / (z - A) = 1/(-i omega eps_main)
    ratio = b_m * np.exp(-0.5j * omega * dt) / (z - a_m)
    assert abs(ratio * (-1j * omega * e_main) - 1.0) < 1e-12


def test_metal_cell_stable_form_response():
    """E cell, class 'e', eps = -10 + 1.5i: the stable form of check_files/StableForm.pdf
    p.3 as the reference codes it (trapezoid E(l) = E(l-1) + (dE(l)+dE(l-1))/2 with no
    time step factor) settles to -i*omega*(eps/iR)*Ehat = What, where What is the
    forcing phasor of the backward curl W_x = H_z - H_y = -H_y taken at time n - 1/2.
    A dropped or misplaced term in either coefficient changes this response."""
    eps = -10.0 + 1.5j
    Ehat, iR, iI, omega = _drive_e("e", 1.0 + 0j, eps, np.zeros(6))
    N, iR2, iI2, dt = oracle.set_tau(omega, 0.5)
    # W^{k} uses H^{k+1/2} (set before step k) -> phasor -1 at time k+1/2 relative to the
    # E sample at k+1; the stable-form relation pairs W with E at the same half step.
    W = -1.0
    lhs = -1j * omega * (eps / iR) * Ehat * np.exp(0.0j)
    assert abs(lhs - W) < 1e-9 * abs(W), (lhs, W)


def test_lossless_cell_coefficients():
    N, iR, iI, dt = oracle.set_tau(2 * np.pi / 16, 0.5)
    for er in (1.0, 2.1, 12.1, 1e30):
        L, cont, addr, lnotn, (Ah, Bh, Ae, Be) = _one_cell("n", 1.0 + 0j, er + 0j, np.zeros(6), iR, iI)
        assert Be[0] == pytest.approx(1.0 / (iR * er), rel=1e-15)
        assert abs(Ae[0] - 1.0) <= 2.3e-16
        assert Bh[0] == pytest.approx(1.0 / iR, rel=1e-15)


def _physical(seed=0, inc=1):
    from smatranem_amd import casegen
    case = casegen.waveguide_case(10, 8, 24, 6, 0, 3, core=(4, 3), seed=seed, iBc=(1, 0, 0),
                                  metal_cells=3, lossy_cells=2, time_period=(1, 1, 3))
    case.inc_mode = inc
    N, iR, iI, _ = oracle.set_tau(case.omega, case.min_dt)
    st = oracle.setup_from_media(case.L, case.mu, case.eps, iR=iR, iI=iI)
    return case, st, N


def test_sign_flip_and_zero_source():
    case, st, N = _physical()
    iBTI = (case.iB, case.iT, case.iI, 1)
    H1, E1, m1, _ = oracle.run(case.L, st, case.iBc, iBTI, case.Hinc, case.Einc, N, case.time_period)
    H2, E2, m2, _ = oracle.run(case.L, st, case.iBc, (case.iB, case.iT, case.iI, -2), case.Hinc, case.Einc,
                               N, case.time_period)
    assert np.abs(H1).max() > 0
    assert np.array_equal(H1, -H2) and np.array_equal(E1, -E2)
    z = np.zeros_like(case.Hinc)
    H0, E0, m0, _ = oracle.run(case.L, st, case.iBc, iBTI, z, z, N, case.time_period)
    assert not H0.any() and not E0.any() and not any(m.any() for m in m0)


def test_energy_conserved_in_periodic_vacuum():
    """A = 1, B = b on every component, periodic wrap, no source: the Yee invariant
    ||E^n||^2 + <H^{n-1/2}, H^{n+1/2}> (b_E = b_H) is constant."""
    L = (6, 5, 7)
    n = 6 * 5 * 7
    b = 0.4
    ones = np.ones(3 * n)
    st = dict(L=oracle.i32([*L, 1]), iBc=oracle.i32([1, 1, 1]), iBTI=oracle.i32([0, 1, 50, 1]))
    rng = np.random.default_rng(9)
    H = np.concatenate([rng.normal(size=3 * n), np.zeros(3)])
    E = np.concatenate([rng.normal(size=3 * n), np.zeros(3)])
    cont = np.full(n, ord("n"), np.int8)
    addr = np.zeros(n, np.int32)
    plane = 30
    z = np.zeros(2 * plane)
    mons = [np.zeros(4 * plane) for _ in range(4)]
    vals = []
    for _ in range(200):
        Hprev = H.copy()
        oracle.main_calc(orc.orc_main_calc, st["L"], ones, b * ones, ones, b * ones, st["iBc"], H, E,
                         st["iBTI"], z, z, np.zeros(4), *mons, cont, addr)
        # after the call: H = H^{n+1/2}, E = E^{n+1}; invariant with E^{n}? use
        # ||E^{n+1}||^2 + <H^{n+1/2}, H^{n+3/2}> evaluated one step later
        vals.append((Hprev.copy(), H.copy(), E.copy()))
    inv = [np.dot(vals[k][2], vals[k][2]) + np.dot(vals[k][1], vals[k + 1][1]) for k in range(len(vals) - 1)]
    inv = np.array(inv)
    assert np.ptp(inv) / np.abs(inv).mean() < 1e-12
