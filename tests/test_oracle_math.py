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


# Distinct sigma samples per axis and per integer/half position, so that a coefficient
# taking another axis's sample, or the half sample instead of the integer one, changes
# the response.  Order: (x int, x half, y int, y half, z int, z half).
SIG6 = np.array([0.11, 0.23, 0.37, 0.41, 0.53, 0.67])
# Distinct drive phasors of the three components of the held field.
DRIVE = np.array([1.0 + 0.0j, 0.5j, -0.7 + 0.3j])


def _drive(cls, mu, eps, sig6, field, drive=DRIVE, omega=2 * np.pi / 16, min_dt=0.5, periods=400):
    """Drive a one-cell grid (PEC wrap, so every neighbour term of a curl vanishes) with
    the other field held at the sinusoids Re(drive[c] e^{-i omega t}), and return the
    steady phasors of the driven field's main slots [0:3] and split/aux slots [3:6],
    fitted over the last 8 periods, with (iR, iI, omega).

    field 'H': E is set to its values at t = k dt before step k, E's own update is
    discarded, and H after step k is H at (k + 1/2) dt.  field 'E': H is set to its
    values at (k + 1/2) dt before step k, and E after step k is E at (k + 1) dt.  The
    held field's split/aux slots are zero, so the curls see the drive as the total."""
    N, iR, iI, dt = oracle.set_tau(omega, min_dt)
    L, cont, addr, lnotn, (Ah, Bh, Ae, Be) = _one_cell(cls, mu, eps, sig6, iR, iI)
    held_h = field == "E"
    if held_h:
        # the H sweep runs before the E sweep of a step: A = 1, B = 0 keeps the held H
        Ah, Bh = np.ones_like(Ah), np.zeros_like(Bh)
    flen = 3 * (1 + lnotn + 1)
    H = np.zeros(flen)
    E = np.zeros(flen)
    z2 = np.zeros(2)
    mons = [np.zeros(4) for _ in range(4)]
    t_held = 0.5 if held_h else 0.0
    t_out = 1.0 if held_h else 0.5
    hist = []
    total = periods * N
    for k in range(total):
        th = omega * (k + t_held) * dt
        vals = drive.real * np.cos(th) + drive.imag * np.sin(th)
        held = H if held_h else E
        held[:] = 0.0
        held[:3] = vals
        oracle.main_calc(orc.orc_main_calc, L, Ah, Bh, Ae, Be, oracle.i32([0, 0, 0]), H, E,
                         oracle.i32([-1, -1, 5, 1]), z2, z2, np.zeros(4), *mons, cont, addr)
        out = E if held_h else H
        hist.append(out[:3 * (1 + lnotn)].copy())
    t = np.arange(total - 8 * N, total) + t_out
    y = np.array(hist[-8 * N:])
    A = np.stack([np.cos(omega * t * dt), np.sin(omega * t * dt), np.ones_like(t)], 1)
    coef = np.linalg.lstsq(A, y, rcond=None)[0]
    ph = coef[0] + 1j * coef[1]
    if lnotn == 0:
        ph = np.concatenate([ph, np.zeros(3)])
    return ph, iR, iI, omega


def _curl_halves(drive, backward):
    """The '+' and '-' halves (F:372-413, js = 0 and 1) of the one-cell curl of a field
    with component phasors `drive`: every neighbour term vanishes, so the forward curl
    is (-F[c2], +F[c1]) and the backward curl (+F[c2], -F[c1]), c1 = c+1, c2 = c+2."""
    plus = np.zeros(3, complex)
    minus = np.zeros(3, complex)
    for c in range(3):
        c1, c2 = (c + 1) % 3, (c + 2) % 3
        plus[c] = drive[c2] if backward else -drive[c2]
        minus[c] = -drive[c1] if backward else drive[c1]
    return plus, minus


def test_lossy_dielectric_cell_matches_frequency_domain():
    """E cell, class 'n', eps = 2 + 0.5i: -i omega eps Ehat = What, What the phasor of
    the forward curl of H (F:291-299)."""
    eps = 2.0 + 0.5j
    ph, iR, iI, omega = _drive("n", 1.0 + 0j, eps, np.zeros(6), "E")
    plus, minus = _curl_halves(DRIVE, backward=False)
    W = plus + minus
    np.testing.assert_allclose(-1j * omega * eps * ph[:3], W, rtol=0, atol=1e-9)


def test_lossy_magnetic_cell_matches_frequency_domain():
    """H cell, class 'n', mu = 2 + 0.5i: H <- A H + B (-W) with W the backward curl of E
    (F:259-267) is exact at omega for -i omega mu Hhat = -What.  A flipped sign of Im mu
    in A or B, or a dropped term, moves the response by O(Im mu / |mu|)."""
    mu = 2.0 + 0.5j
    ph, iR, iI, omega = _drive("n", mu, 1.0 + 0j, np.zeros(6), "H")
    plus, minus = _curl_halves(DRIVE, backward=True)
    W = plus + minus
    np.testing.assert_allclose(-1j * omega * mu * ph[:3], -W, rtol=0, atol=1e-9)


@pytest.mark.parametrize("family", ["E", "H"])
def test_pml_slots_match_stretched_response(family):
    """Class 'a' cell with lossy mu and eps and distinct sigma samples on every axis.
    The main slot of component c takes the '+' half of the curl, stretched along
    d1 = c+1; the split slot the '-' half, along d2 = c+2 (F:118-127, F:268-276,
    F:300-307).  H uses the integer sigma samples, E the half samples.  Each slot is
    exact at omega for the stretched medium m_eff = Re m (1 + i sigma) + i Im m:
    -i omega eps_eff Ehat = What (E) and -i omega mu_eff Hhat = -What (H)."""
    mu, eps = 1.5 + 0.3j, 2.5 + 0.4j
    ph, iR, iI, omega = _drive("a", mu, eps, SIG6, family)
    plus, minus = _curl_halves(DRIVE, backward=family == "H")
    m = mu if family == "H" else eps
    half = 0 if family == "H" else 1
    sgn = -1.0 if family == "H" else 1.0
    for c in range(3):
        d1, d2 = (c + 1) % 3, (c + 2) % 3
        m_main = m.real * (1 + 1j * SIG6[half + 2 * d1]) + 1j * m.imag
        m_split = m.real * (1 + 1j * SIG6[half + 2 * d2]) + 1j * m.imag
        assert abs(-1j * omega * m_main * ph[c] - sgn * plus[c]) < 1e-9, (family, c, "main")
        assert abs(-1j * omega * m_split * ph[3 + c] - sgn * minus[c]) < 1e-9, (family, c, "split")


@pytest.mark.parametrize("cls", ["e", "b"])
def test_metal_cell_stable_form_response(cls):
    """E in a class 'e' (or 'b') cell, eps = -10 + 1.5i: the stable form of
    check_files/StableForm.pdf p.3 as the reference codes it (trapezoid
    E(l) = E(l-1) + (dE(l)+dE(l-1))/2 with no time step factor; F:129-134, F:308-314)
    settles to -i omega (eps/iR) Ehat = What, What the phasor of the BACKWARD curl of H
    (Q1) at the same half step.  A dropped or misplaced term in either coefficient
    changes this response."""
    eps = -10.0 + 1.5j
    ph, iR, iI, omega = _drive(cls, -3.0 + 0.8j, eps, np.zeros(6), "E")
    plus, minus = _curl_halves(DRIVE, backward=True)
    W = plus + minus
    np.testing.assert_allclose(-1j * omega * (eps / iR) * ph[:3], W, rtol=0, atol=1e-9)


@pytest.mark.parametrize("cls", ["h", "b"])
def test_magnetic_metal_cell_stable_form_response(cls):
    """H in a class 'h' (or 'b') cell, mu = -3 + 0.8i: the aux slot X takes
    X <- A X + B (-W) with the stable-form coefficients (F:135-139) and the main slot
    H += (X_old + X_new)/2 (F:277-283).  With (z+1) iI = omega z^{1/2} and
    (z-1) iR = -i omega z^{1/2}, z = e^{-i omega dt}, this settles to
    -i omega (mu/iR) Hhat = -What.  A flipped sign of Im mu in the aux A or B moves it."""
    mu = -3.0 + 0.8j
    ph, iR, iI, omega = _drive(cls, mu, -10.0 + 1.5j, np.zeros(6), "H")
    plus, minus = _curl_halves(DRIVE, backward=True)
    W = plus + minus
    np.testing.assert_allclose(-1j * omega * (mu / iR) * ph[:3], -W, rtol=0, atol=1e-9)


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
