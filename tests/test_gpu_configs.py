"""The benchmark configurations at their own sizes.

C1: 48 x 48 x 64 uniform grid, eps = 1 with a 16 x 16 x 24 block of eps = 12.1, 8-cell z
PML, omega = 2*pi*0.1, Dt factor 1/0.99 (N_2pi = 18), envelope (10, 80, 111): 1998 steps,
the reference monitors on planes 9 and 54 and a 4-frequency DFT on plane 54.  FP64 on
the GPU is bit-identical to the oracle; FP32 |S| within 1e-3 of FP64 and 1e-4 of each
|S|.
C5 (reduced): the 768 x 768 x 1024 workload's structure at 96 x 96 x 64.
C2: 160 x 160 x 320 curvilinear grid with PEC edges and 10-cell z PML (the bench
workload), FP32 against FP64 on the GPU over the ramp and 3000 CW steps: |S| of the
reference monitors and of the 16-frequency DFT, normalised by the envelope sum of the
steps run, within 1e-3 absolute and 1e-3 of each |S|.
C3: three runs drawn from the batch generator (sizes in [64, 192]^2 x [64, 256], 8-cell
PML on all faces), advanced together with graph replay over CW periods, equal to each
run alone; and one structure excited by 6 port modes in 2 polarisations, fused.
"""
import numpy as np
import pytest

import oracle
from smatranem_amd import casegen, fdtd
from tests.cases import rel_l2

pytestmark = pytest.mark.gpu

S_TOL = 1e-3  # |S| error of FP32 against FP64 on S normalised by the envelope sum run


def s_errors(pairs, env_sum):
    """(absolute errors, |S| values) of FP32 against FP64 over (S32, S64) overlap pairs,
    each S divided by the envelope sum of the steps run (sum00, F:159-168: the
    reference's monitors, normalised by 1/((ramp+cw) N), sum to O(|S|) only over the
    whole envelope).  Fails when an FP64 value is zero: a monitor the wave has not
    reached compares nothing."""
    errs, mags = [], []
    for s32, s64 in pairs:
        a, b = abs(s32) / env_sum, abs(s64) / env_sum
        assert b > 0, "a compared monitor is identically zero"
        errs.append(abs(a - b))
        mags.append(b)
    return np.array(errs), np.array(mags)


def overlap(case, H, E):
    """Projection of monitor phasors on the source profile (S_calc's overlap,
    FDTDcalc/FDTDsplitfieldMetal.c:437-445, with the incident profile as the mode)."""
    w = np.concatenate([case.Hinc, case.Einc])
    v = np.concatenate([H, E])
    return complex(np.vdot(w, v) / np.vdot(w, w).real)


def dft_phasors(D, m, f):
    """(H, E) complex[2*L0*L1], (x, y) interleaved, of DFT plane m and frequency f."""
    plane = D.shape[-1]
    H = np.empty(2 * plane, dtype=np.complex128)
    E = np.empty_like(H)
    H[0::2], H[1::2], E[0::2], E[1::2] = D[m, 0, f], D[m, 1, f], D[m, 2, f], D[m, 3, f]
    return H, E


def test_c1_fp64_bitwise_and_fp32_s_error():
    case = casegen.block_in_vacuum_case()
    prep = casegen.prepare(case)
    assert prep.N == 18 and prep.N * case.time_period[2] == 1998
    assert (case.iB, case.iT, case.iI) == (9, 54, 10)
    ws = 2 * np.pi / prep.N * np.array([0.9, 1.0, 1.1, 1.2])
    out = {}
    for prec in (fdtd.F64, fdtd.F32):
        with prep.plan(prec) as p:
            p.set_dft([case.iT], ws, chunk=32)
            sums = p.run(0, case.time_period[2])
            H, E = p.read_fields()
            out[prec] = (H, E, p.read_monitors(), p.read_dft(), sums)
    N, iR, iI, _ = oracle.set_tau(case.omega, case.min_dt / case.dt_factor)
    st = oracle.setup_from_media(case.L, case.mu, case.eps, iR=iR, iI=iI)
    Ho, Eo, mo, so = oracle.run(case.L, st, case.iBc, (case.iB, case.iT, case.iI, case.inc_mode), case.Hinc,
                                case.Einc, N, case.time_period, 0, case.time_period[2])
    H, E, mons, D64, sums = out[fdtd.F64]
    assert np.abs(Ho).max() > 0
    assert np.array_equal(H, Ho) and np.array_equal(E, Eo)
    for a, b in zip(mons, mo):
        assert np.array_equal(a.view(np.float64), b)
    assert np.array_equal(sums, so)
    # the DFT at the source frequency is the reference's top monitor
    HT, ET = mons[2], mons[3]
    ref = np.stack([HT[0::2], HT[1::2], ET[0::2], ET[1::2]])
    assert rel_l2(D64[0, :, 1, :], ref) < 1e-12
    # FP32 against FP64: |S| of both monitors and of every DFT frequency
    H32, E32, m32, D32, _ = out[fdtd.F32]
    # fields: FP32 rounding accumulated over 1998 steps (2.5e-4 measured); the bound that
    # matters is the one on |S| below
    assert rel_l2(H32, H) < 1e-3 and rel_l2(E32, E) < 1e-3
    pairs = [(overlap(case, m32[2 * k], m32[2 * k + 1]), overlap(case, mons[2 * k], mons[2 * k + 1]))
             for k in range(2)]
    pairs += [(overlap(case, *dft_phasors(D32, 0, f)), overlap(case, *dft_phasors(D64, 0, f)))
              for f in range(len(ws))]
    errs, mags = s_errors(pairs, sums[0])
    print(f"C1 |S| error: max abs {errs.max():.3e}, max rel {(errs / mags).max():.3e}, |S| {mags.min():.3e}..."
          f"{mags.max():.3e}", flush=True)
    assert errs.max() <= S_TOL, errs
    # relative to each |S|: the monitors' |S| is O(0.05), so the absolute bar alone
    # would allow a 2 % error
    assert (errs / mags).max() <= 1e-4, errs / mags


def test_c2_fp32_vs_fp64_s_error():
    case = casegen.curvilinear_waveguide_case()
    prep = casegen.prepare(case)
    ws = 2 * np.pi / prep.N * np.linspace(0.85, 1.15, 16)
    # the ramp, then 3000 steps of the CW part: the wave has crossed the grid
    steps = case.time_period[0] * prep.N + 3000
    out = {}
    for prec in (fdtd.F32, fdtd.F64):
        with prep.plan(prec) as p:
            p.set_dft([case.iB, case.iT], ws, chunk=32)
            if prec == fdtd.F32:
                p.autotune(10)
            sums = p.run_steps(0, steps)
            out[prec] = (p.read_monitors(), p.read_dft(), sums)
    m32, D32, _ = out[fdtd.F32]
    m64, D64, sums = out[fdtd.F64]
    pairs = []
    for k in range(2):
        pairs.append((overlap(case, m32[2 * k], m32[2 * k + 1]), overlap(case, m64[2 * k], m64[2 * k + 1])))
        for f in range(len(ws)):
            pairs.append((overlap(case, *dft_phasors(D32, k, f)), overlap(case, *dft_phasors(D64, k, f))))
    errs, mags = s_errors(pairs, sums[0])
    print(f"C2 |S| error: max abs {errs.max():.3e}, max rel {(errs / mags).max():.3e}, |S| {mags.min():.3e}..."
          f"{mags.max():.3e}", flush=True)
    assert errs.max() <= S_TOL, errs.max()
    # relative to each |S| (the DFT frequencies away from the drive have small |S|),
    # and to the largest
    assert (errs / mags).max() <= 1e-3, (errs / mags).max()
    assert errs.max() <= 1e-5 * mags.max(), (errs.max(), mags.max())


def test_c3_batch_members_graph_equals_alone(monkeypatch):
    """The two-step pass with split slots, chosen explicitly (the autotuner's timing-based
    choice could differ between the members alone and batched)."""
    monkeypatch.setenv("SMX_F32_TWO", "2,32,1")
    cases = casegen.batch_cases(3, seed=2025, time_period=(1, 12, 14))
    preps = [casegen.prepare(c) for c in cases]
    steps = min(pr.N * pr.case.time_period[2] for pr in preps)
    alone = []
    for pr in preps:
        with pr.plan(fdtd.F32) as p:
            assert p.kernels == "two_step_f32<32,8,1,true,false,false,false>"
            p.run_steps(0, steps)
            alone.append(p.read_fields())
    plans = [pr.plan(fdtd.F32) for pr in preps]
    try:
        info = {}
        fdtd.run_batch(plans, 0, steps, info=info)
        together = [p.read_fields() for p in plans]
    finally:
        for p in plans:
            p.close()
    assert info["graph_launches"] >= 1
    for a, b in zip(alone, together):
        assert np.abs(a[0]).max() > 0
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


def test_c3_port_modes_fused_equals_alone(monkeypatch):
    """The batch as the configuration states it, at a reduced size: one structure excited by
    6 Hermite-Gaussian port modes in 2 polarisations (12 runs sharing the media), all in
    the fused launch, each bit for bit the run alone; different modes give different
    fields."""
    import dataclasses
    monkeypatch.setenv("SMX_F32_TWO", "2,16")
    runs = casegen.port_batch_cases(1, 6, 2, seed=7, pml=4, time_period=(1, 12, 14),
                                    dims=((40, 48), (36, 44), (40, 48)))
    base = casegen.prepare(runs[0][3])
    preps = [dataclasses.replace(base, case=c) for _, _, _, c in runs]
    steps = 2 * base.N
    alone = []
    for pr in preps:
        with pr.plan(fdtd.F32) as p:
            p.run_steps(0, steps)
            alone.append(p.read_fields())
    plans = [pr.plan(fdtd.F32) for pr in preps]
    try:
        info = {}
        fdtd.run_batch(plans, 0, steps, mode=fdtd.BATCH_FUSED, info=info)
        together = [p.read_fields() for p in plans]
    finally:
        for p in plans:
            p.close()
    assert info["graph_launches"] == steps // 2
    for a, b in zip(alone, together):
        assert np.abs(a[1]).max() > 0
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    assert rel_l2(alone[0][1], alone[1][1]) > 0.5 and rel_l2(alone[0][1], alone[2][1]) > 0.5


@pytest.fixture(scope="module")
def c5_reduced():
    """The 96 x 96 x 64 layered stack, its oracle fields after 4 periods and the FP64
    plan's fields and DFT (computed once for both forms of the FP32 pass)."""
    case = casegen.layered_stack_case(L=(96, 96, 64), pml=12, layers=12, time_period=(1, 20, 24))
    prep = casegen.prepare(case)
    periods = 4
    planes = [case.iI + 2, int(case.L[2]) - int(case.L[3])]
    ws = 2 * np.pi / prep.N * np.linspace(0.9, 1.1, 8)
    with prep.plan(fdtd.F64) as p:
        p.set_dft(planes, ws, chunk=32)
        p.run(0, periods)
        H64, E64 = p.read_fields()
        D64 = p.read_dft()
    N, iR, iI, _ = oracle.set_tau(case.omega, case.min_dt / case.dt_factor)
    st = oracle.setup_from_media(case.L, case.mu, case.eps, iR=iR, iI=iI)
    Ho, Eo, _, _ = oracle.run(case.L, st, case.iBc, (case.iB, case.iT, case.iI, case.inc_mode), case.Hinc,
                              case.Einc, N, case.time_period, 0, periods)
    return dict(prep=prep, periods=periods, planes=planes, ws=ws, F64=(H64, E64, D64), oracle=(Ho, Eo))


@pytest.mark.parametrize("big", ["0", "1"])
def test_c5_layered_stack_reduced(big, c5_reduced, monkeypatch):
    """The large-grid workload's structure at 96 x 96 x 64: curvilinear scale factors, 12
    seeded dielectric layers, a PEC grating, 12-cell PML on every face, an 8-frequency DFT
    on two planes.  FP64 is bit-identical to the oracle; the FP32 two-step pass with the
    PML inside, in its usual and in its 64-bit form (the form grids of 179 M cells and
    more take), is within 2e-4 of it on the fields and of FP64 on the DFT."""
    c = c5_reduced
    prep = c["prep"]
    assert prep.case.meta["pec_edges"] > 0 and int(np.count_nonzero(prep.Cont == ord("a"))) > 0
    H64, E64, D64 = c["F64"]
    Ho, Eo = c["oracle"]
    assert np.abs(Ho).max() > 0 and np.abs(D64).max() > 0
    assert np.array_equal(H64, Ho) and np.array_equal(E64, Eo)
    monkeypatch.setenv("SMX_F32_TWO", "2,16")
    monkeypatch.setenv("SMX_F32_BIG", big)
    with prep.plan(fdtd.F32) as p:
        p.set_dft(c["planes"], c["ws"], chunk=32)
        assert p.kernels.startswith("two_step_f32<") and p.kernels.endswith(",true>" if big == "1" else ",false>")
        p.run(0, c["periods"])
        H32, E32 = p.read_fields()
        D32 = p.read_dft()
    assert rel_l2(H32, Ho) < 2e-4 and rel_l2(E32, Eo) < 2e-4, (rel_l2(H32, Ho), rel_l2(E32, Eo))
    assert rel_l2(D32, D64) < 2e-4, rel_l2(D32, D64)
