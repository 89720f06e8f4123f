"""The z-column physics pins of tests/test_oracle_physics.py through the HIP plans.

FP64: the monitor phasors are bit for bit the oracle's, so the oracle's pins carry over.
FP32: every kernel family (the two sweeps, the fused staged kernel with split slots, the
two-step pass with split slots, and that pass writing the multi-frequency DFT, at the drive
frequency and, through the discrete time derivative, at four others) reproduces
the exact discrete solution -- dispersion between two planes, the reflection and
transmission of lossless and lossy slabs, |r|^2 + |t|^2 = 1 -- within 5e-5.  Measured on
an MI355X, the FP32 residues are 1.1e-7 to 2.3e-6 for every family: the 2.2e-6 of the
PML and the run's end, as in FP64, with the FP32 rounding far below it (each test
prints its residues)."""
import numpy as np
import pytest

import oracle
from smatranem_amd import casegen, fdtd
from tests import cases as C

pytestmark = pytest.mark.gpu

L01 = 4
F32_BOUND = 5e-5

# environment of each FP32 kernel family
KERNELS = {
    "sweeps": {"SMX_F32_KERNEL": "twopass"},
    "fused_split": {"SMX_F32_KERNEL": "fused", "SMX_F32_TILE": "3,16"},
    "two_step_split": {"SMX_F32_KERNEL": "fused", "SMX_F32_TWO": "2,16"},
}


def _gpu_column(precision, eps_slab=None, iB=C.COLUMN_UP, dft=False):
    """E_x phasors (at iB, at the top plane) of one plan run; with dft, those of the
    plan's multi-frequency DFT at the drive frequency (and a second frequency) on the
    same two planes, which must agree with the monitors."""
    prep = casegen.prepare(C.column_case(eps_slab, iB=iB, L01=L01))
    with prep.plan(precision) as p:
        if dft:
            p.set_dft([iB, C.COLUMN_TOP], [2 * np.pi / prep.N, 0.9 * 2 * np.pi / prep.N], chunk=16)
        p.run(0, prep.case.time_period[2])
        HB, EB, HT, ET = p.read_monitors()
        kernels = p.kernels
        out = (EB[0], ET[0])
        if dft:
            D = p.read_dft()      # complex [plane][Hx, Hy, Ex, Ey][f][L0*L1]
            out_dft = (D[0, 2, 0, 0], D[1, 2, 0, 0])
    if dft:
        return out, out_dft, kernels
    return out, kernels


def test_f64_column_monitors_bitwise():
    prep = casegen.prepare(C.column_case(4.0, L01=L01))
    with prep.plan(fdtd.F64) as p:
        p.run(0, prep.case.time_period[2])
        mons = p.read_monitors()
    c = prep.case
    N, iR, iI, _ = oracle.set_tau(c.omega, c.min_dt)
    st = oracle.setup_from_media(c.L, c.mu, c.eps, iR=iR, iI=iI)
    _, _, mo, _ = oracle.run(c.L, st, c.iBc, (c.iB, c.iT, c.iI, 1), c.Hinc, c.Einc, N, c.time_period)
    for a, b in zip(mons, mo):
        assert np.array_equal(a.view(np.float64), b)
    r, t = C.column_r_t(_gpu_column(fdtd.F64)[0], (mons[1][0], mons[3][0]))
    r_x, t_x = C.column_exact_slab(4.0)
    assert abs(r - r_x) < 2e-5 and abs(t - t_x) < 2e-5


@pytest.mark.parametrize("family", list(KERNELS))
def test_f32_column_physics(family, monkeypatch):
    for k, v in KERNELS[family].items():
        monkeypatch.setenv(k, v)
    (eb, et), kernels = _gpu_column(fdtd.F32, iB=C.COLUMN_DOWN)
    print(family, kernels, flush=True)
    if family == "two_step_split":
        assert "two_step_f32<32,8,1,true,false,false,false>" in kernels
    d = C.COLUMN_TOP - C.COLUMN_DOWN
    err = [abs(et / eb - np.exp(1j * C.column_theta() * d))]
    assert err[0] < F32_BOUND
    vac = _gpu_column(fdtd.F32)[0]
    for eps_slab in (2.0, 4.0, 4.0 + 0.8j):
        r, t = C.column_r_t(vac, _gpu_column(fdtd.F32, eps_slab)[0])
        r_x, t_x = C.column_exact_slab(eps_slab)
        err += [abs(r - r_x), abs(t - t_x)]
        print(family, "eps", eps_slab, "|dr| %.1e |dt| %.1e" % tuple(err[-2:]), flush=True)
        assert abs(r - r_x) < F32_BOUND and abs(t - t_x) < F32_BOUND, (eps_slab, r, r_x, t, t_x)
        if np.imag(eps_slab) == 0:
            assert abs(abs(r) ** 2 + abs(t) ** 2 - 1.0) < F32_BOUND


def test_f32_column_dft_matches_monitors(monkeypatch):
    """The two-step pass writing the DFT samples: at the drive frequency the DFT equals
    the reference monitors' accumulation (the same sum in another order), and the slab's
    r and t from the DFT meet the same bounds."""
    for k, v in KERNELS["two_step_split"].items():
        monkeypatch.setenv(k, v)
    vac_m, vac_d, kernels = _gpu_column(fdtd.F32, dft=True)
    assert "two_step_f32<32,8,1,true,false,true,false>" in kernels, kernels
    slab_m, slab_d, _ = _gpu_column(fdtd.F32, 4.0, dft=True)
    for m, dd in zip(vac_m + slab_m, vac_d + slab_d):
        assert abs(m - dd) <= 1e-9 * abs(m), (m, dd)
    r, t = C.column_r_t(vac_d, slab_d)
    r_x, t_x = C.column_exact_slab(4.0)
    print("dft |dr| %.1e |dt| %.1e" % (abs(r - r_x), abs(t - t_x)), flush=True)
    assert abs(r - r_x) < F32_BOUND and abs(t - t_x) < F32_BOUND


def test_f32_column_dft_off_drive_frequencies(monkeypatch):
    """At a DFT phase step w other than the drive's 2 pi / N, the leapfrog in lossless cells
    is the same recurrence with omega replaced by omega_eff = 2 invtau_R sin(w/2) (the
    discrete time derivative at w; at w = 2 pi / N it is omega, set_tau's exactness).  The
    two-step pass's DFT of a lossless slab at five frequencies around the drive gives that
    exact solution's r and t, and |r|^2 + |t|^2 = 1, within 2e-5 (measured on an MI355X:
    2.2e-6 and below).  The continuum omega = w / Dt instead of omega_eff moves r by 2.9e-3
    to 7.2e-3 at these frequencies: far outside the bound."""
    for k, v in KERNELS["two_step_split"].items():
        monkeypatch.setenv(k, v)
    prep = casegen.prepare(C.column_case(L01=L01))
    w0 = 2 * np.pi / prep.N
    ws = [w0 * s for s in (0.92, 0.96, 1.0, 1.04, 1.08)]
    out = []
    for eps_slab in (None, 4.0):
        pr = casegen.prepare(C.column_case(eps_slab, L01=L01))
        with pr.plan(fdtd.F32) as p:
            p.set_dft([C.COLUMN_UP, C.COLUMN_TOP], ws, chunk=64)
            p.run(0, pr.case.time_period[2])
            D = p.read_dft()
        out.append(D[:, 2, :, 0])   # E_x phasors [plane][f]
    for f, w in enumerate(ws):
        om = 2 * prep.invtau_R * np.sin(w / 2)
        r, t = C.column_r_t((out[0][0, f], out[0][1, f]), (out[1][0, f], out[1][1, f]), omega=om)
        r_x, t_x = C.column_exact_slab(4.0, omega=om)
        print("w/w0 %.2f omega_eff/omega %.5f |dr| %.1e |dt| %.1e" % (w / w0, om / C.COLUMN_OMEGA,
              abs(r - r_x), abs(t - t_x)), flush=True)
        assert abs(r - r_x) < 2e-5 and abs(t - t_x) < 2e-5, (w / w0, r, r_x, t, t_x)
        assert abs(abs(r) ** 2 + abs(t) ** 2 - 1.0) < 2e-5
