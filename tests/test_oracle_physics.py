"""Whole-grid physics pins of the oracle on a z column: periodic x/y wrap (weight 1), a
uniform E_x/H_y sheet source, z PML at both ends (tests/cases.py: column_case).  The
fields stay uniform in x and y, and the scheme reduces to the 1-D Yee recurrence in unit
cells (mu = 1)

    E_x[k+1] + E_x[k-1] = (2 - omega^2 eps_k) E_x[k],

exactly at the drive frequency, because set_tau (F:515-524) makes the time
discretisation exact at omega.  The monitors' running DFT (F:318-328) over a run whose
fields have left through the PML by the end is the source spectrum times the steady
transfer function, so ratios of monitor phasors are ratios of the recurrence's solution.

* dispersion: between two planes downstream of the source the phasor ratio is
  e^{i theta d} with 4 sin^2(theta/2) = omega^2 mu eps (the discrete dispersion of
  check_files/CheckMaxwell.pdf p.3-4), not the continuum e^{i omega d}: a dropped
  1/tau factor or a misplaced coefficient moves theta.
* slab: the reflection and transmission of a dielectric slab, lossless and lossy, equal
  the recurrence's exact solution in amplitude and phase (which pins the plane each
  eps_x sits on: a slab one plane off changes r by 0.1 and more); the lossless slabs
  have |r|^2 + |t|^2 = 1 (unitarity of S).

Tolerances: the residue is the discrete reflection of the graded PML and what is left in
the grid at the end of the run, together below 2.2e-6 here; the bounds are 2e-5.
"""
import numpy as np
import pytest

import oracle
from tests import cases as C


def _column(eps_slab=None, iB=C.COLUMN_UP):
    """The oracle's E_x phasors (at iB, at the top plane) of one column run."""
    case = C.column_case(eps_slab, iB=iB)
    N, iR, iI, _ = oracle.set_tau(case.omega, case.min_dt)
    st = oracle.setup_from_media(case.L, case.mu, case.eps, iR=iR, iI=iI)
    L0, L1, L2 = case.L[:3]
    cont = st["Cont"].reshape(L2, L1, L0)
    assert (cont[C.COLUMN_PML + 1:L2 - C.COLUMN_PML - 1] == ord("n")).all()
    H, E, (HB, EB, HT, ET), _ = oracle.run(case.L, st, case.iBc, (case.iB, case.iT, case.iI, 1),
                                           case.Hinc, case.Einc, N, case.time_period)
    # the grid must be (nearly) empty at the end for the DFT to be the transfer function;
    # the incident wave's amplitude is about 1
    assert np.abs(E).max() < 1e-6 and np.abs(H).max() < 1e-6
    return EB.view(np.complex128)[0], ET.view(np.complex128)[0]


def test_discrete_dispersion_between_planes():
    eb, et = _column(iB=C.COLUMN_DOWN)
    d = C.COLUMN_TOP - C.COLUMN_DOWN
    ratio = et / eb
    assert abs(ratio - np.exp(1j * C.column_theta() * d)) < 2e-5
    # the continuum phase is off by d * (theta - omega) = 0.16 rad: far outside the bound
    assert abs(ratio - np.exp(1j * C.COLUMN_OMEGA * d)) > 1e-1


@pytest.mark.parametrize("eps_slab", [2.0, 4.0, 4.0 + 0.8j])
def test_slab_reflection_transmission(eps_slab):
    r, t = C.column_r_t(_column(), _column(eps_slab))
    r_x, t_x = C.column_exact_slab(eps_slab)
    assert abs(r - r_x) < 2e-5, (r, r_x)
    assert abs(t - t_x) < 2e-5, (t, t_x)
    if np.imag(eps_slab) == 0:
        assert abs(abs(r) ** 2 + abs(t) ** 2 - 1.0) < 2e-5
    else:
        assert abs(r) ** 2 + abs(t) ** 2 < 0.6   # absorbed in the slab: 0.50 here


def test_exact_slab_solution_is_a_slab():
    """The closed-form side: a vacuum 'slab' reflects nothing and transmits 1; a slab
    shifted by one plane gives an r that the bounds above reject."""
    r, t = C.column_exact_slab(1.0)
    assert abs(r) < 1e-14 and abs(t - 1.0) < 1e-14
    r4, _ = C.column_exact_slab(4.0)
    a, b = C.COLUMN_SLAB
    saved = C.COLUMN_SLAB
    try:
        C.COLUMN_SLAB = (a + 1, b + 1)
        r4s, _ = C.column_exact_slab(4.0)
    finally:
        C.COLUMN_SLAB = saved
    assert abs(r4s - r4) > 0.1
