"""The broadband point-source workload on the GPU: a Gaussian-modulated E_z point current
source (smx_plan_set_point_source) in a 48 x 48 x 64 unit grid, vacuum with a centred
16 x 16 x 24 block of eps 12.1, 8-cell PML on every face, 2000 steps at 0.99 of the 3-D
CFL limit, a 4-frequency DFT on one z-plane.  FP64 against the oracle's point-source
run (orc_point_add after every orc_main_calc): fields bit for bit, DFT within 1e-12 of
its largest value (a sum in another order); FP32 within the bounds written below."""
import dataclasses

import numpy as np
import pytest

import oracle
from smatranem_amd import casegen, fdtd
from tests.cases import rel_l2

pytestmark = pytest.mark.gpu

SRC = (24, 24, 14)       # E_z source cell, in vacuum below the block
DFT_PLANE = 50           # above the block, below the top z PML
FREQS = (0.085, 0.095, 0.105, 0.115)


def _setup(L=(48, 48, 64), steps=2000):
    case = casegen.point_source_case(L)
    prep = casegen.prepare(case)
    iR, iI, dt, amp = casegen.point_source_drive(case, steps)
    prep = dataclasses.replace(prep, invtau_R=iR, invtau_I=iI, dt=dt)
    ws = [2 * np.pi * f * dt for f in FREQS]
    return case, prep, amp, ws


def _oracle(case, prep, amp, ws, plane, src=SRC, comp=2):
    """The oracle step by step with the point source (E component `comp` at cell `src`),
    and a DFT of the plane's total tangential fields with the library's phases (H of step
    k at e^{iw(k+1)}, E at e^{iw(k+1.5)}).  Returns (H, E, dft[4][nf][L0*L1], the largest
    |H| and |E| of the run)."""
    st = oracle.setup_from_media(case.L, case.mu, case.eps, iR=prep.invtau_R, iI=prep.invtau_I)
    L0, L1, L2 = (int(v) for v in case.L[:3])
    n, P = L0 * L1 * L2, L0 * L1
    lib = oracle.load_oracle()
    flen = 3 * (n + st["Lnotn"] + 1)
    H, E = np.zeros(flen), np.zeros(flen)
    mons = [np.zeros(4 * P) for _ in range(4)]
    zero = np.zeros(2 * P)
    ExpN = oracle.expn_table(prep.N)
    cells = np.arange(P) + P * plane
    a = st["Cont"][cells] == ord("a")
    sl = 3 * n + 3 * st["addr"][cells[a]]
    lp = src[0] + L0 * (src[1] + L1 * src[2])
    dft = np.zeros((4, len(ws), P), dtype=np.complex128)
    w = np.asarray(ws)[:, None]
    peak = np.zeros(2)
    for k in range(amp.size):
        tn = (k + 1) % prep.N
        oracle.main_calc(lib.orc_main_calc, case.L, st["Ah"], st["Bh"], st["Ae"], st["Be"], case.iBc, H, E,
                         (-1, -1, case.iI, case.inc_mode), zero, zero, ExpN[4 * tn:4 * tn + 4].copy(), *mons,
                         st["Cont"], st["addr"])
        lib.orc_point_add(E, st["Be"], comp, lp, float(amp[k]))
        peak = np.maximum(peak, (np.abs(H).max(), np.abs(E).max()))
        for v, (F, c) in enumerate(((H, 0), (H, 1), (E, 0), (E, 1))):
            x = F[c + 3 * cells].copy()
            x[a] += F[c + sl]
            dft[v] += np.exp(1j * w * (k + (1.0 if v < 2 else 1.5))) * x[None, :]
    return H, E, dft, peak


@pytest.fixture(scope="module")
def c1():
    case, prep, amp, ws = _setup()
    return case, prep, amp, ws, _oracle(case, prep, amp, ws, DFT_PLANE)


def _gpu(prep, amp, ws, precision, chunk=64):
    with prep.plan(precision) as p:
        p.set_dft([DFT_PLANE], ws, chunk=chunk)
        p.set_point_source(2, SRC, amp)
        if precision == fdtd.F32:
            assert p.autotune(4)[3] != 3  # one step per pass: the autotuner skips the two-step pass
        p.run_steps(0, amp.size)
        H, E = p.read_fields()
        D = p.read_dft()
    return H, E, D


def test_point_source_f64_bitwise(c1):
    case, prep, amp, ws, (Ho, Eo, Do, _) = c1
    H, E, D = _gpu(prep, amp, ws, fdtd.F64)
    assert np.abs(Eo).max() > 0
    assert np.array_equal(H, Ho) and np.array_equal(E, Eo)
    D = D.reshape(4, len(ws), -1)
    assert np.abs(D - Do).max() <= 1e-12 * np.abs(Do).max(), np.abs(D - Do).max() / np.abs(Do).max()


def test_point_source_f32(c1):
    """FP32 against the FP64 oracle: the DFT (the spectrum a broadband run exists for)
    within 1e-4 of its largest value at every frequency; the fields after 2000 steps --
    the late-time remainder of a pulse the PML has mostly absorbed -- within 1e-5 of the
    largest field of the run."""
    case, prep, amp, ws, (Ho, Eo, Do, peak) = c1
    H, E, D = _gpu(prep, amp, ws, fdtd.F32)
    D = D.reshape(4, len(ws), -1)
    for f in range(len(ws)):
        err = np.abs(D[:, f] - Do[:, f]).max() / np.abs(Do[:, f]).max()
        assert err < 1e-4, (f, err)
    eH, eE = np.abs(H - Ho).max() / peak[0], np.abs(E - Eo).max() / peak[1]
    assert eH < 1e-5 and eE < 1e-5, (eH, eE)


@pytest.mark.parametrize("planes", [[9], [9, 20], [9, 9]])
def test_point_source_on_a_dft_plane(planes):
    """An E_x source on a DFT plane: the sample kernel applies the step's source term in
    the thread that samples the source's cell, before its samples ([9], [9, 20]), or, with
    two ring planes on that cell, the term keeps a launch of its own ([9, 9]).  FP64
    fields bit for bit and the DFT within 1e-12 of its largest value against the oracle;
    FP32 DFT within 1e-4."""
    case, prep, amp, ws = _setup(L=(24, 24, 32), steps=300)
    src = (11, 13, 9)
    Ho, Eo, Do, _ = _oracle(case, prep, amp, ws, 9, src=src, comp=0)
    for precision in (fdtd.F64, fdtd.F32):
        with prep.plan(precision) as p:
            p.set_dft(planes, ws, chunk=32)
            p.set_point_source(0, src, amp)
            p.run_steps(0, amp.size)
            H, E = p.read_fields()
            D = p.read_dft()[0].reshape(4, len(ws), -1)
        assert np.abs(Do).max() > 0
        err = np.abs(D - Do).max() / np.abs(Do).max()
        if precision == fdtd.F64:
            assert np.array_equal(H, Ho) and np.array_equal(E, Eo)
            assert err <= 1e-12, err
        else:
            assert err < 1e-4, err


def test_point_source_refusals_and_removal():
    case, prep, amp, ws = _setup(L=(24, 24, 32), steps=40)
    with prep.plan(fdtd.F64) as p:
        for comp, cell in ((3, (12, 12, 12)), (2, (1, 12, 12)), (2, (12, 12, 40))):  # bad component, PML, outside
            with pytest.raises(fdtd.SmxError):
                p.set_point_source(comp, cell, amp)
        p.set_point_source(2, (12, 12, 12), amp)
        p.set_point_source(2, (12, 12, 12), None)  # removed: zero source, zero fields
        p.run_steps(0, amp.size)
        H, E = p.read_fields()
    assert not H.any() and not E.any()


def test_point_source_refused_by_the_adjoint():
    """The adjoint's transposed step has no point-source term: a plan with one is refused."""
    case, prep, amp, ws = _setup(L=(24, 24, 32), steps=40)
    with prep.plan(fdtd.F32) as p:
        p.set_point_source(2, (12, 12, 12), amp)
        with pytest.raises(fdtd.SmxError):
            p.adjoint_forward(40, 0, 20, ws[0], (10, 14, 10, 14, 10, 14))
        p.set_point_source(2, (12, 12, 12), None)
        J = p.adjoint_forward(40, 0, 20, ws[0], (10, 14, 10, 14, 10, 14))  # accepted once removed
    assert np.isfinite(J)


def test_point_source_plan_in_batch():
    """A plan with a point source runs eagerly inside a batch (no graph, no fused launch:
    its amplitude changes every step) and equals the same plan run alone, bit for bit,
    in FP32 and FP64, next to a plan without one."""
    case, prep, amp, ws = _setup(L=(24, 24, 32), steps=120)
    other = casegen.prepare(casegen.waveguide_case(20, 16, 30, outer=6, pml_xy=0, pml_z=3, seed=5,
                                                   time_period=(2, 100, 102)))
    for prec in (fdtd.F32, fdtd.F64):
        alone = []
        for pr, pt in ((prep, True), (other, False)):
            with pr.plan(prec) as p:
                if pt:
                    p.set_point_source(2, (12, 12, 10), amp)
                p.run_steps(0, amp.size)
                alone.append(p.read_fields())
        plans = [prep.plan(prec), other.plan(prec)]
        try:
            plans[0].set_point_source(2, (12, 12, 10), amp)
            for mode in (fdtd.BATCH_FUSED,):
                fdtd.run_batch(plans, 0, amp.size, mode=mode)
            together = [p.read_fields() for p in plans]
        finally:
            for p in plans:
                p.close()
        for a, b in zip(alone, together):
            assert np.abs(a[1]).max() > 0
            assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
