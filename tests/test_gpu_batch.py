"""Batched runs (smx_plan_run_batch): independent plans of different grids and
precisions advanced together give exactly what each plan gives when run alone, and the
FP64 members stay bit-identical to the oracle."""
import numpy as np
import pytest

import oracle
from smatranem_amd import casegen, fdtd
from tests.cases import small_waveguide

pytestmark = pytest.mark.gpu

MEMBERS = [
    (fdtd.F32, dict(seed=21, L0=12, L1=10, L2=28)),
    (fdtd.F32, dict(seed=22, L0=37, L1=19, L2=41, outer=8, pml_xy=2, pml_z=3, core=(3, 3))),
    (fdtd.F64, dict(seed=23, L0=9, L1=14, L2=30, outer=7, pml_xy=0, pml_z=3)),
    (fdtd.F32, dict(seed=24, L0=20, L1=16, L2=33, outer=8, pml_xy=0, pml_z=0, iBc=(1, -1, 1),
                    metal_cells=0, lossy_cells=0)),
]


def _preps():
    return [(prec, casegen.prepare(small_waveguide(**kw))) for prec, kw in MEMBERS]


def _state(p):
    H, E = p.read_fields()
    return H, E, tuple(p.read_monitors())


@pytest.mark.parametrize("two", [None, "0,16", "3,16"])
def test_batch_equals_separate_runs(two, monkeypatch):
    """With two = "v,fz" the PML-free member runs two steps per pass (and the batch
    loop has to keep the members' step counts in step)."""
    if two:
        monkeypatch.setenv("SMX_F32_TWO", two)
    preps = _preps()
    steps = min(pr.case.time_period[2] * pr.N for _, pr in preps)
    alone = []
    for prec, pr in preps:
        with pr.plan(prec) as p:
            sums = p.run_steps(0, steps)
            alone.append(_state(p) + (sums,))
    plans = [pr.plan(prec) for prec, pr in preps]
    try:
        sums = fdtd.run_batch(plans, 0, steps)
        together = [_state(p) for p in plans]
    finally:
        for p in plans:
            p.close()
    for i, (a, b) in enumerate(zip(alone, together)):
        assert np.abs(a[0]).max() > 0, i
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]), i
        for ma, mb in zip(a[2], b[2]):
            assert np.array_equal(ma, mb), i
        assert np.array_equal(a[3], sums[i]), i


def test_batch_fp64_member_matches_oracle():
    """Whole periods of the FP64 member, batched in two calls, against the oracle."""
    preps = _preps()
    N2 = preps[2][1].N
    total = min(pr.case.time_period[2] * pr.N for _, pr in preps)
    periods = total // N2
    assert periods >= 2
    plans = [pr.plan(prec) for prec, pr in preps]
    try:
        fdtd.run_batch(plans, 0, N2)
        fdtd.run_batch(plans, N2, (periods - 1) * N2)
        H, E = plans[2].read_fields()
        mons = plans[2].read_monitors()
    finally:
        for p in plans:
            p.close()
    c = preps[2][1].case
    N, iR, iI, _ = oracle.set_tau(c.omega, c.min_dt / c.dt_factor)
    st = oracle.setup_from_media(c.L, c.mu, c.eps, iR=iR, iI=iI)
    Ho, Eo, mo, _ = oracle.run(c.L, st, c.iBc, (c.iB, c.iT, c.iI, c.inc_mode), c.Hinc, c.Einc, N,
                               c.time_period, 0, periods)
    assert N == N2
    assert np.array_equal(H, Ho) and np.array_equal(E, Eo)
    for a, b in zip(mons, mo):
        assert np.array_equal(a.view(np.float64), b)


def test_batch_errors():
    preps = _preps()[:2]
    plans = [pr.plan(prec) for prec, pr in preps]
    try:
        with pytest.raises(fdtd.SmxError) as e:
            fdtd.run_batch([plans[0], plans[0]], 0, 1)
        assert e.value.status == 1
        with pytest.raises(fdtd.SmxError) as e:
            fdtd.run_batch(plans, 0, -1)
        assert e.value.status == 1
    finally:
        for p in plans:
            p.close()
    prec, pr = preps[0]
    g = pr.grid(prec)
    with fdtd.Plan(g, mu=pr.mu, eps=pr.eps, sigma=pr.sigma, Cont=pr.Cont, address_notn=pr.addr) as bare:
        with pytest.raises(fdtd.SmxError) as e:
            fdtd.run_batch([bare], 0, 1)
        assert e.value.status == 5


GRAPH_MEMBERS = [
    (fdtd.F32, dict(seed=31, L0=12, L1=10, L2=28, metal_cells=0, lossy_cells=0, time_period=(1, 24, 26))),
    # lossy cells put this member on the one-step kernels; no metal cells, whose stable
    # form as the reference codes it grows without bound over runs this long
    (fdtd.F32, dict(seed=32, L0=37, L1=19, L2=41, outer=8, pml_xy=2, pml_z=3, core=(3, 3),
                    metal_cells=0, lossy_cells=4, time_period=(2, 22, 25))),
    (fdtd.F64, dict(seed=33, L0=9, L1=14, L2=30, outer=7, pml_xy=0, pml_z=3, metal_cells=0, lossy_cells=4,
                    time_period=(1, 24, 26))),
    (fdtd.F32, dict(seed=34, L0=20, L1=16, L2=33, outer=8, pml_xy=0, pml_z=0, iBc=(1, -1, 1),
                    metal_cells=0, lossy_cells=0, time_period=(1, 24, 25))),
    (fdtd.F32, dict(seed=35, L0=24, L1=12, L2=30, outer=8, pml_xy=3, pml_z=3, metal_cells=0,
                    lossy_cells=0, omega=2 * np.pi / 9.0, time_period=(1, 30, 32))),
]


def _graph_preps():
    return [(prec, casegen.prepare(small_waveguide(**kw))) for prec, kw in GRAPH_MEMBERS]


@pytest.mark.parametrize("kernel", [None, "0,16", "2,16", "twopass"])
def test_batch_graph_replay_equals_streams(kernel, monkeypatch):
    """The graph mode (one captured segment of every plan, replayed over the CW periods)
    gives bit for bit what the streams mode and each plan alone give, over ramp, CW
    and ramp-down, from a step that is not a period start, and it does replay."""
    if kernel == "twopass":
        monkeypatch.setenv("SMX_F32_KERNEL", "twopass")
    elif kernel:
        monkeypatch.setenv("SMX_F32_TWO", kernel)
    preps = _graph_preps()
    assert len({pr.N for _, pr in preps}) > 1, "members should differ in N_2pi"
    steps = min(pr.case.time_period[2] * pr.N for _, pr in preps) - 3
    results = {}
    for mode in (fdtd.BATCH_STREAMS, fdtd.BATCH_GRAPH):
        plans = [pr.plan(prec) for prec, pr in preps]
        try:
            info = {}
            s1 = fdtd.run_batch(plans, 0, 3, mode=mode)
            s2 = fdtd.run_batch(plans, 3, steps - 3, mode=mode, info=info)
            results[mode] = ([_state(p) for p in plans], s1 + s2, info["graph_launches"])
        finally:
            for p in plans:
                p.close()
    alone = []
    for prec, pr in preps:
        with pr.plan(prec) as p:
            sums = p.run_steps(0, 3) + p.run_steps(3, steps - 3)
            alone.append(_state(p) + (sums,))
    st_streams, sums_streams, n_streams = results[fdtd.BATCH_STREAMS]
    st_graph, sums_graph, n_graph = results[fdtd.BATCH_GRAPH]
    assert n_streams == 0
    assert n_graph >= 3, n_graph
    for i in range(len(preps)):
        a, b, c = alone[i], st_streams[i], st_graph[i]
        assert np.abs(a[0]).max() > 0, i
        for x in (b, c):
            assert np.array_equal(a[0], x[0]) and np.array_equal(a[1], x[1]), i
            for ma, mx in zip(a[2], x[2]):
                assert np.array_equal(ma, mx), i
        assert np.array_equal(a[3], sums_streams[i]) and np.array_equal(a[3], sums_graph[i]), i


def test_batch_graph_off_with_dft():
    """A member with the multi-frequency DFT set keeps the batch on the streams path
    (its sample and fold launches depend on host counters), with the same results."""
    preps = _graph_preps()[:3]
    steps = min(pr.case.time_period[2] * pr.N for _, pr in preps)
    out = []
    for mode in (fdtd.BATCH_STREAMS, fdtd.BATCH_GRAPH):
        plans = [pr.plan(prec) for prec, pr in preps]
        try:
            plans[0].set_dft([preps[0][1].case.L[2] - 4], [0.3, 0.5], chunk=8)
            info = {}
            fdtd.run_batch(plans, 0, steps, mode=mode, info=info)
            out.append(([_state(p) for p in plans], plans[0].read_dft(), info["graph_launches"]))
        finally:
            for p in plans:
                p.close()
    assert out[1][2] == 0
    for a, b in zip(out[0][0], out[1][0]):
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    assert np.array_equal(out[0][1], out[1][1])


FUSED_MEMBERS = [
    (fdtd.F32, dict(seed=41, L0=12, L1=10, L2=28, metal_cells=0, lossy_cells=0, time_period=(1, 5, 7))),
    (fdtd.F32, dict(seed=42, L0=37, L1=19, L2=41, outer=8, pml_xy=2, pml_z=3, core=(3, 3), metal_cells=0,
                    lossy_cells=0, time_period=(1, 5, 7))),
    (fdtd.F32, dict(seed=43, L0=20, L1=16, L2=33, outer=8, pml_xy=0, pml_z=0, iBc=(1, -1, 1), metal_cells=0,
                    lossy_cells=0, time_period=(1, 5, 7))),
    # several x tiles, y tiles and z chunks (z chunks of 16 planes)
    (fdtd.F32, dict(seed=44, L0=70, L1=26, L2=75, outer=10, pml_xy=4, pml_z=4, core=(5, 4), metal_cells=0,
                    lossy_cells=0, omega=2 * np.pi / 9.0, time_period=(1, 5, 7))),
]


def _fused_preps(members):
    return [(prec, casegen.prepare(small_waveguide(**kw))) for prec, kw in members]


# multi-frequency DFT on members 2 (monitors on: the ring carries them) and 3: (planes,
# frequencies in units of 2 pi / N_2pi, chunk); chunk 7 is odd, so a pass that would not
# fit folds first
FUSED_DFT = {2: ([6, 20], [0.9, 1.0, 1.1], 8), 3: ([40], [1.0, 1.05], 7)}


@pytest.mark.parametrize("odd,fz,dft", [(False, None, False), (True, None, False), (True, "8", False),
                                        (False, "64", False), (True, None, True), (False, "8", True)])
def test_batch_fused_equals_separate_runs(odd, fz, dft, monkeypatch):
    """SMX_BATCH_FUSED (every plan's two-step pass in one launch) gives bit for bit the
    fields, monitors, DFT accumulators and envelope sums of each plan run alone on the
    split-slot two-step kernel, over two calls (the second starting on the other
    ping-pong buffer when the first ran an odd number of steps), with the batch's own z
    chunk or a forced one, with and without DFT members."""
    monkeypatch.setenv("SMX_F32_TWO", "2,16")
    if fz:
        monkeypatch.setenv("SMX_BATCH_FZ", fz)
    preps = _fused_preps(FUSED_MEMBERS)
    steps = min(pr.case.time_period[2] * pr.N for _, pr in preps) - (1 if odd else 2)
    first = 3 if odd else 4

    def setup(i, p, pr):
        if dft and i in FUSED_DFT:
            planes, ws, chunk = FUSED_DFT[i]
            p.set_dft(planes, [w * 2 * np.pi / pr.N for w in ws], chunk=chunk)

    def state(i, p):
        return _state(p) + ((p.read_dft(),) if dft and i in FUSED_DFT else (None,))
    alone = []
    for i, (prec, pr) in enumerate(preps):
        with pr.plan(prec) as p:
            setup(i, p, pr)
            assert p.kernels.startswith("two_step_f32<32,8,1,true,false,")
            sums = p.run_steps(0, first) + p.run_steps(first, steps - first)
            alone.append(state(i, p) + (sums,))
    plans = [pr.plan(prec) for prec, pr in preps]
    try:
        for i, (p, (_, pr)) in enumerate(zip(plans, preps)):
            setup(i, p, pr)
        info1, info2 = {}, {}
        s1 = fdtd.run_batch(plans, 0, first, mode=fdtd.BATCH_FUSED, info=info1)
        s2 = fdtd.run_batch(plans, first, steps - first, mode=fdtd.BATCH_FUSED, info=info2)
        together = [state(i, p) for i, p in enumerate(plans)]
    finally:
        for p in plans:
            p.close()
    assert info1["graph_launches"] == first // 2 and info2["graph_launches"] == (steps - first) // 2
    sums = s1 + s2
    for i, (a, b) in enumerate(zip(alone, together)):
        assert np.abs(a[0]).max() > 0, i
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]), i
        for ma, mb in zip(a[2], b[2]):
            assert np.array_equal(ma, mb), i
        if a[3] is not None:
            assert np.abs(a[3]).max() > 0 and np.array_equal(a[3], b[3]), i
        assert np.array_equal(a[4], sums[i]), i


def test_batch_fused_falls_back():
    """A batch the fused launch cannot take (an FP64 member, a member with lossy cells)
    runs as the graph mode, with each plan's results as alone."""
    # (the lossy member's own pass leaves its lossy tiles to region sweeps)
    members = [FUSED_MEMBERS[0],
               (fdtd.F64, dict(seed=46, L0=9, L1=14, L2=30, outer=7, pml_xy=0, pml_z=3, metal_cells=0,
                               lossy_cells=0, time_period=(1, 5, 7))),
               (fdtd.F32, dict(seed=47, L0=16, L1=12, L2=30, outer=7, pml_xy=2, pml_z=3, metal_cells=0,
                               lossy_cells=4, time_period=(1, 5, 7)))]
    for sel in ([0, 1], [0, 2]):
        preps = _fused_preps([members[i] for i in sel])
        steps = min(pr.case.time_period[2] * pr.N for _, pr in preps)
        alone = []
        for prec, pr in preps:
            with pr.plan(prec) as p:
                p.run_steps(0, steps)
                alone.append(p.read_fields())
        plans = [pr.plan(prec) for prec, pr in preps]
        try:
            fdtd.run_batch(plans, 0, steps, mode=fdtd.BATCH_FUSED)
            together = [p.read_fields() for p in plans]
        finally:
            for p in plans:
                p.close()
        for a, b in zip(alone, together):
            assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]), sel
