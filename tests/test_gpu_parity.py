"""Parity of the HIP library against the oracle on the GPU.

FP64 (plain path): bit-identical to the oracle, which is itself bit-identical to the
reference (tests/test_oracle_vs_ref.py).  FP32 (fused path): relative L2 against the
FP64 oracle within the tolerances written in each test."""
import numpy as np
import pytest

import oracle
from smatranem_amd import casegen, fdtd
from tests.cases import call_main_calc, random_state, rel_l2, small_waveguide

pytestmark = pytest.mark.gpu


def _eq(a, b):
    return np.array_equal(np.asarray(a), np.asarray(b))


@pytest.mark.parametrize("seed,L", [(0, (7, 6, 9)), (1, (65, 3, 4)), (2, (1, 5, 7)), (3, (6, 1, 5)),
                                    (4, (4, 5, 1)), (5, (20, 17, 6)), (6, (3, 3, 2))])
def test_main_calc_bitwise(seed, L):
    st = random_state(seed, L=L)
    Hg, Eg, mg, _ = call_main_calc(lambda *a: fdtd.main_calc(*a), st)
    Ho, Eo, mo, rc = call_main_calc(oracle.load_oracle().orc_main_calc, st)
    assert rc == 0
    assert _eq(Hg, Ho) and _eq(Eg, Eo)
    for a, b in zip(mg, mo):
        assert _eq(a, b)


def test_plan_steps_bitwise():
    st = random_state(7, L=(9, 8, 10), iBc=(1, -1, 0))
    g = fdtd.Grid(L=st["L"], iBc=st["iBc"], iBTIin=st["iBTI"], precision=fdtd.F64, Lnotn=st["Lnotn"])
    with fdtd.Plan(g, coeffs=(st["Ah"], st["Bh"], st["Ae"], st["Be"]), Cont=st["Cont"],
                   address_notn=st["addr"]) as p:
        p.write_fields(st["H"], st["E"])
        p.write_monitors(*[m.view(np.complex128) for m in st["mons"]])
        Ho, Eo, mo = st["H"].copy(), st["E"].copy(), [m.copy() for m in st["mons"]]
        for _ in range(6):
            p.step(st["Hi"], st["Ei"], st["ExpN2"])
            call_main_calc(oracle.load_oracle().orc_main_calc, st, Ho, Eo, mo)
        H, E = p.read_fields()
        assert _eq(H, Ho) and _eq(E, Eo)
        for a, b in zip(p.read_monitors(), mo):
            assert _eq(a.view(np.float64), b)


def _oracle_run(prep, P0, P1, fields=None, monitors=None):
    c = prep.case
    N, iR, iI, _ = oracle.set_tau(c.omega, c.min_dt / c.dt_factor)
    ref = oracle.setup_from_media(c.L, c.mu, c.eps, iR=iR, iI=iI)
    return oracle.run(c.L, ref, c.iBc, (c.iB, c.iT, c.iI, c.inc_mode), c.Hinc, c.Einc, N,
                      c.time_period, P0, P1, fields=fields, monitors=monitors)


CASES = {
    # z PML only, periodic/antiperiodic/PEC walls: the reference monitors are on
    "ports": dict(pml_xy=0, pml_z=3, iBc=(1, -1, 0), metal_cells=6, lossy_cells=4),
    # PML on every face: the reference monitors are off (plane -1)
    "box": dict(pml_xy=2, pml_z=3, iBc=(0, 0, 0), metal_cells=6, lossy_cells=4),
    "curvilinear": dict(pml_xy=0, pml_z=3, iBc=(1, 1, 0), scale_amp=0.3, pec_fraction=0.03,
                        metal_cells=0, lossy_cells=0),
    # lossy mu in 'n' cells next to lossy eps and metal: every H coefficient branch
    "lossy_mu": dict(pml_xy=2, pml_z=3, iBc=(0, 0, 0), metal_cells=3, lossy_cells=3, lossy_mu_cells=8),
}


@pytest.mark.parametrize("name", list(CASES))
def test_plan_run_f64_bitwise(name):
    prep = casegen.prepare(small_waveguide(seed=3, **CASES[name]))
    with prep.plan(fdtd.F64) as p:
        sums = p.run(0, 3)
        sums = p.run(3, prep.case.time_period[2], sums)
        H, E = p.read_fields()
        mons = p.read_monitors()
    Ho, Eo, mo, so = _oracle_run(prep, 0, prep.case.time_period[2])
    assert np.abs(Ho).max() > 0
    assert _eq(H, Ho) and _eq(E, Eo)
    for a, b in zip(mons, mo):
        assert _eq(a.view(np.float64), b)
    assert _eq(sums, so)


@pytest.fixture(params=["fused", "staged", "staged-deep", "twopass"])
def f32_kernel(request, monkeypatch):
    """The FP32 plan's step kernel: one fused launch staged through registers or by
    asynchronous copies to LDS (two planes ahead with split slots, or three planes
    ahead without them plus the region sweeps around the PML-free box), or the two
    in-place sweeps."""
    monkeypatch.setenv("SMX_F32_KERNEL", "twopass" if request.param == "twopass" else "fused")
    if request.param == "staged":
        monkeypatch.setenv("SMX_F32_TILE", "3,16")
    if request.param == "staged-deep":
        monkeypatch.setenv("SMX_F32_TILE", "5,16")
    return request.param


@pytest.mark.parametrize("name", list(CASES))
def test_plan_run_f32_close(name, f32_kernel):
    """FP32 fields vs the FP64 oracle: relative L2 <= 2e-4 on fields, monitors."""
    prep = casegen.prepare(small_waveguide(seed=3, **CASES[name]))
    with prep.plan(fdtd.F32) as p:
        p.run(0, prep.case.time_period[2])
        H, E = p.read_fields()
        mons = p.read_monitors()
    Ho, Eo, mo, _ = _oracle_run(prep, 0, prep.case.time_period[2])
    assert rel_l2(H, Ho) < 2e-4, rel_l2(H, Ho)
    assert rel_l2(E, Eo) < 2e-4, rel_l2(E, Eo)
    for a, b in zip(mons, mo):
        if np.linalg.norm(b) > 0:
            assert rel_l2(a.view(np.float64), b) < 2e-4


@pytest.mark.parametrize("dims", [(37, 19, 41), (33, 9, 17), (5, 3, 20)])
def test_fused_partial_tiles(dims, f32_kernel):
    """Grids that are not multiples of the fused tile (32 x 8 x 16): partial tiles,
    wrapped halos and chunk edges, against the FP64 oracle."""
    L0, L1, L2 = dims
    prep = casegen.prepare(small_waveguide(seed=8, L0=L0, L1=L1, L2=L2, outer=max(L2 // 4, 4),
                                           pml_xy=0, pml_z=2, iBc=(1, -1, 1), metal_cells=3,
                                           lossy_cells=2, core=(2, 2)))
    with prep.plan(fdtd.F32) as p:
        p.run(0, prep.case.time_period[2])
        H, E = p.read_fields()
        mons = p.read_monitors()
    Ho, Eo, mo, _ = _oracle_run(prep, 0, prep.case.time_period[2])
    assert rel_l2(H, Ho) < 2e-4 and rel_l2(E, Eo) < 2e-4, (rel_l2(H, Ho), rel_l2(E, Eo))
    for a, b in zip(mons, mo):
        if np.linalg.norm(b) > 0:
            assert rel_l2(a.view(np.float64), b) < 2e-4


def test_f32_round_trip_fields(f32_kernel):
    """write_fields/read_fields of the FP32 plan: total/split packing is lossless up to
    FP32 rounding of the main and split slots."""
    prep = casegen.prepare(small_waveguide(seed=5, **CASES["box"]))
    rng = np.random.default_rng(0)
    with prep.plan(fdtd.F32) as p:
        H = rng.normal(size=p.flen)
        E = rng.normal(size=p.flen)
        H[-3:] = 0
        E[-3:] = 0
        p.write_fields(H, E)
        H2, E2 = p.read_fields()
    assert rel_l2(H2, H) < 1e-6 and rel_l2(E2, E) < 1e-6


def test_monitor_plane_in_pml_is_refused():
    prep = casegen.prepare(small_waveguide(seed=1, **CASES["box"]))
    g = prep.grid(fdtd.F64)
    g.iBTIin = (5, 20, 6, 1)
    with pytest.raises(fdtd.SmxError) as e:
        fdtd.Plan(g, mu=prep.mu, eps=prep.eps, sigma=prep.sigma, Cont=prep.Cont,
                  address_notn=prep.addr)
    assert e.value.status == 100


def test_run_errors():
    prep = casegen.prepare(small_waveguide(seed=1, **CASES["ports"]))
    g = prep.grid(fdtd.F64)
    with fdtd.Plan(g, mu=prep.mu, eps=prep.eps, sigma=prep.sigma, Cont=prep.Cont,
                   address_notn=prep.addr) as p:
        with pytest.raises(fdtd.SmxError) as e:
            p.run(0, 1)
        assert e.value.status == 5
        p.set_source(prep.case.Hinc, prep.case.Einc, prep.N, prep.case.time_period)
        with pytest.raises(fdtd.SmxError) as e:
            p.run(prep.case.time_period[2] - 1, prep.case.time_period[2] + 1)
        assert e.value.status == 101


def test_zero_source_keeps_fields_zero(f32_kernel):
    prep = casegen.prepare(small_waveguide(seed=2, **CASES["ports"]))
    for prec in (fdtd.F64, fdtd.F32):
        with prep.plan(prec) as p:
            z = np.zeros(2 * prep.case.plane, dtype=np.complex128)
            p.set_source(z, z, prep.N, prep.case.time_period)
            p.run(0, 3)
            H, E = p.read_fields()
        assert not H.any() and not E.any()


SHAPES = ((32, 8), (64, 4), (16, 16), (32, 8), (16, 16), (32, 8), (16, 16), (32, 8), (32, 8), (32, 8))


@pytest.mark.parametrize("tile", ["0,16", "1,8", "2,32", "0,5", "3,16", "4,8", "3,5", "4,64", "3,41",
                                  "5,16", "6,8", "7,64", "5,5", "7,41", "8,16", "9,32", "9,3"])
def test_fused_tile_shapes(tile, monkeypatch):
    """Every fused tile shape and z chunk (SMX_F32_TILE: shapes 0-2 staged through
    registers, 3-4 by asynchronous copies to LDS with split slots, 5-7 without them
    plus the region sweeps) against the FP64 oracle."""
    monkeypatch.setenv("SMX_F32_KERNEL", "fused")
    monkeypatch.setenv("SMX_F32_TILE", tile)
    prep = casegen.prepare(small_waveguide(seed=9, L0=37, L1=19, L2=41, outer=8, pml_xy=2, pml_z=3,
                                           iBc=(1, 0, -1), metal_cells=3, lossy_cells=2, core=(3, 3)))
    with prep.plan(fdtd.F32) as p:
        shape, fz = (int(v) for v in tile.split(","))
        assert p.autotune(0) == SHAPES[shape] + (fz, 1 + (shape >= 3))
        p.run(0, prep.case.time_period[2])
        H, E = p.read_fields()
    Ho, Eo, _, _ = _oracle_run(prep, 0, prep.case.time_period[2])
    assert rel_l2(H, Ho) < 2e-4 and rel_l2(E, Eo) < 2e-4, (rel_l2(H, Ho), rel_l2(E, Eo))


@pytest.mark.parametrize("tile", ["5,16", "6,8", "7,64", "5,5", "7,41", "3,16", "8,8", "9,16"])
def test_staged_without_pml(tile, monkeypatch):
    """The staged variants without split slots (tiles 5-7, for grids without PML cells)
    against the FP64 oracle, on periodic / antiperiodic walls with metal and lossy cells."""
    monkeypatch.setenv("SMX_F32_KERNEL", "fused")
    monkeypatch.setenv("SMX_F32_TILE", tile)
    prep = casegen.prepare(small_waveguide(seed=10, L0=37, L1=19, L2=41, outer=8, pml_xy=0, pml_z=0,
                                           iBc=(1, -1, 1), metal_cells=3, lossy_cells=2, core=(3, 3)))
    assert not any(c == ord("a") for c in prep.Cont)
    with prep.plan(fdtd.F32) as p:
        shape, fz = (int(v) for v in tile.split(","))
        assert p.autotune(0) == SHAPES[shape] + (fz, 2)
        p.run(0, prep.case.time_period[2])
        H, E = p.read_fields()
    Ho, Eo, _, _ = _oracle_run(prep, 0, prep.case.time_period[2])
    assert rel_l2(H, Ho) < 2e-4 and rel_l2(E, Eo) < 2e-4, (rel_l2(H, Ho), rel_l2(E, Eo))


@pytest.mark.parametrize("tile", ["5,16", "7,8"])
def test_region_sweeps_carry_monitor_plane(tile, monkeypatch):
    """With the monitor planes right next to the z PML, the top one lies outside the
    staged kernel's PML-free box: the region sweeps accumulate it."""
    monkeypatch.setenv("SMX_F32_KERNEL", "fused")
    monkeypatch.setenv("SMX_F32_TILE", tile)
    prep = casegen.prepare(small_waveguide(seed=12, L0=21, L1=17, L2=36, outer=7, pml_xy=0, pml_z=3,
                                           monitor_gap=0, iBc=(1, -1, 0), metal_cells=3, lossy_cells=2))
    assert prep.case.iT == prep.case.L[2] - 4
    with prep.plan(fdtd.F32) as p:
        p.run(0, prep.case.time_period[2])
        H, E = p.read_fields()
        mons = p.read_monitors()
    Ho, Eo, mo, _ = _oracle_run(prep, 0, prep.case.time_period[2])
    assert rel_l2(H, Ho) < 2e-4 and rel_l2(E, Eo) < 2e-4
    for a, b in zip(mons, mo):
        assert np.linalg.norm(b) > 0
        assert rel_l2(a.view(np.float64), b) < 2e-4


# two-step variants: (copies in flight, split slots, 16-byte copies); tiles are 32 x 8
TWO_VARIANTS = ((1, False, False), (2, False, False), (1, True, False), (1, False, True))
TWO_NAMES = tuple(f"two_step_f32<32,8,{d},{str(sp).lower()},{str(wi).lower()},false,false>" for d, sp, wi in TWO_VARIANTS)
TWO_SPLIT = tuple(v[1] for v in TWO_VARIANTS)


@pytest.mark.parametrize("two,iBc", [("0,16", (1, -1, 1)), ("1,8", (0, 0, 0)), ("0,5", (1, 1, -1)),
                                     ("1,64", (-1, 1, 1)), ("0,41", (1, -1, 0)), ("2,16", (1, -1, 1)),
                                     ("3,16", (1, -1, 1)), ("3,5", (0, 0, 0)), ("3,64", (1, 1, -1)),
                                     ("3,41", (1, -1, 0)), ("3,8", (-1, 1, 1))])
def test_two_step_pass(two, iBc, monkeypatch):
    """Two steps per pass (grids without PML and lossy/metal cells) against the FP64
    oracle, fields and both monitor planes, over a run with an odd number of steps (the
    last one takes the one-step kernel)."""
    monkeypatch.setenv("SMX_F32_KERNEL", "fused")
    monkeypatch.setenv("SMX_F32_TWO", two)
    v, fz = (int(x) for x in two.split(","))
    L0 = 68 if TWO_VARIANTS[v][2] else 37  # 16-byte copies need L0 % 4 == 0
    prep = casegen.prepare(small_waveguide(seed=13, L0=L0, L1=19, L2=41, outer=8, pml_xy=0, pml_z=0,
                                           iBc=iBc, metal_cells=0, lossy_cells=0, core=(3, 3)))
    assert all(c == ord("n") for c in prep.Cont)
    steps = prep.case.time_period[2] * prep.N
    with prep.plan(fdtd.F32) as p:
        assert p.autotune(0) == (32, 8, fz, 3)
        assert p.kernels == TWO_NAMES[v]
        sums = p.run_steps(0, steps - 1)
        p.run_steps(steps - 1, 1, sums)
        H, E = p.read_fields()
        mons = p.read_monitors()
    Ho, Eo, mo, so = _oracle_run(prep, 0, prep.case.time_period[2])
    assert np.abs(Ho).max() > 0
    assert rel_l2(H, Ho) < 2e-4 and rel_l2(E, Eo) < 2e-4, (rel_l2(H, Ho), rel_l2(E, Eo))
    for a, b in zip(mons, mo):
        assert np.linalg.norm(b) > 0
        assert rel_l2(a.view(np.float64), b) < 2e-4
    assert np.allclose(sums, so, rtol=1e-14, atol=0)


@pytest.mark.parametrize("variant", ["0,16", "3,16"])
def test_two_step_equals_one_step_kernels(variant, monkeypatch):
    """The two-step pass against the one-step staged kernel on the same grid: both are
    FP32 restatements of the same map, so they agree far below the FP32-vs-FP64 gap."""
    prep = casegen.prepare(small_waveguide(seed=14, L0=40, L1=24, L2=36, outer=8, pml_xy=0, pml_z=0,
                                           iBc=(1, -1, 1), metal_cells=0, lossy_cells=0))
    out = []
    for two in (variant, None):
        monkeypatch.setenv("SMX_F32_KERNEL", "fused")
        monkeypatch.setenv("SMX_F32_TILE", "3,16")
        if two:
            monkeypatch.setenv("SMX_F32_TWO", two)
        else:
            monkeypatch.delenv("SMX_F32_TWO", raising=False)
        with prep.plan(fdtd.F32) as p:
            p.run(0, prep.case.time_period[2])
            out.append(p.read_fields())
    assert rel_l2(out[0][0], out[1][0]) < 1e-5 and rel_l2(out[0][1], out[1][1]) < 1e-5


@pytest.mark.parametrize("L0,L1", [(40, 20), (24, 36), (64, 13)])
def test_two_step_stage_rows(L0, L1, monkeypatch):
    """The 4-byte-copy variants stage each array component only on the rows of the stage
    region its values are read on; the 16-byte-copy variant stages every row.  On a
    PML-free grid all of them run the same per-cell arithmetic, so their fields agree bit
    for bit, with x tiles wrapping (L0 not a multiple of 32, or below one tile) and y
    tiles past the grid (L1 not a multiple of 8)."""
    prep = casegen.prepare(small_waveguide(seed=21, L0=L0, L1=L1, L2=30, outer=8, pml_xy=0, pml_z=0,
                                           iBc=(1, -1, 1), metal_cells=0, lossy_cells=0))
    out = {}
    for two in ("3,16", "0,16", "2,16", "2,5"):
        monkeypatch.setenv("SMX_F32_TWO", two)
        with prep.plan(fdtd.F32) as p:
            assert "two_step_f32<32,8,1," + ("false,true" if two[0] == "3" else
                                              "true,false" if two[0] == "2" else "false,false") in p.kernels
            p.run(0, prep.case.time_period[2])
            out[two] = p.read_fields()
    assert np.abs(out["3,16"][1]).max() > 0
    for two in ("0,16", "2,16", "2,5"):
        assert _eq(out[two][0], out["3,16"][0]) and _eq(out[two][1], out["3,16"][1]), two


@pytest.mark.parametrize("two,kw", [
    ("0,16", dict(pml_xy=0, pml_z=3, iBc=(1, -1, 0))),                 # z PML, both monitors in the box
    ("1,8", dict(pml_xy=0, pml_z=3, monitor_gap=0, iBc=(0, 0, 0))),    # monitor planes in the region sweeps
    ("0,5", dict(pml_xy=2, pml_z=3, iBc=(0, 0, 0))),                   # PML on every face
    ("1,64", dict(pml_xy=3, pml_z=4, iBc=(1, 1, 0), scale_amp=0.3)),
    # split slots: PML cells inside the two-step kernel
    ("2,16", dict(pml_xy=0, pml_z=3, iBc=(1, -1, 0))),
    ("2,8", dict(pml_xy=0, pml_z=3, monitor_gap=0, iBc=(0, 0, 0))),
    ("2,5", dict(pml_xy=2, pml_z=3, iBc=(0, 0, 0))),
    ("2,64", dict(pml_xy=3, pml_z=4, iBc=(1, 1, 0), scale_amp=0.3)),
    # z chunks balanced for the PML layers (third field: a z-PML stage counts 1 + k/2)
    ("2,16,2", dict(pml_xy=0, pml_z=3, iBc=(1, -1, 0))),
    ("2,8,1", dict(pml_xy=2, pml_z=3, monitor_gap=0, iBc=(0, 0, 0))),
    ("0,8,2", dict(pml_xy=0, pml_z=3, iBc=(1, -1, 0))),
    # 16-byte copies
    ("3,16", dict(pml_xy=0, pml_z=3, iBc=(1, -1, 0))),
    ("3,8", dict(pml_xy=0, pml_z=3, monitor_gap=0, iBc=(0, 0, 0))),
    ("3,5", dict(pml_xy=2, pml_z=3, iBc=(0, 0, 0))),
    ("3,64", dict(pml_xy=3, pml_z=4, iBc=(1, 1, 0), scale_amp=0.3)),
])
def test_two_step_pass_with_pml(two, kw, monkeypatch):
    """Two steps per pass on grids with PML, against the FP64 oracle: variants 0-1 on the
    box whose two-step cone is PML-free with the region sweeps (first step into a scratch
    set, second into the new buffer) on the rest; the split-slot variants with the PML
    cells inside."""
    monkeypatch.setenv("SMX_F32_KERNEL", "fused")
    monkeypatch.setenv("SMX_F32_TWO", two)
    v, fz = (int(x) for x in two.split(",")[:2])
    L0 = 44 if TWO_VARIANTS[v][2] else 37
    prep = casegen.prepare(small_waveguide(seed=16, L0=L0, L1=22, L2=44, outer=9, metal_cells=0,
                                           lossy_cells=0, core=(3, 3), **kw))
    assert any(c == ord("a") for c in prep.Cont)
    steps = prep.case.time_period[2] * prep.N
    with prep.plan(fdtd.F32) as p:
        assert p.autotune(0) == (32, 8, fz, 3)
        assert TWO_NAMES[v] in p.kernels and ("region_h_f32" in p.kernels) == (not TWO_SPLIT[v])
        p.run_steps(0, steps - 3)
        p.run_steps(steps - 3, 3)
        H, E = p.read_fields()
        mons = p.read_monitors()
    Ho, Eo, mo, _ = _oracle_run(prep, 0, prep.case.time_period[2])
    assert np.abs(Ho).max() > 0
    assert rel_l2(H, Ho) < 2e-4 and rel_l2(E, Eo) < 2e-4, (rel_l2(H, Ho), rel_l2(E, Eo))
    for a, b in zip(mons, mo):
        if np.linalg.norm(b) > 0:
            assert rel_l2(a.view(np.float64), b) < 2e-4


@pytest.mark.parametrize("kw", [
    dict(pml_xy=0, pml_z=3, metal_cells=6, lossy_cells=4, iBc=(1, -1, 0)),              # monitors on
    dict(pml_xy=2, pml_z=3, metal_cells=6, lossy_cells=4, lossy_mu_cells=3, iBc=(0, 0, 0)),
    dict(L0=40, L1=20, L2=48, outer=8, pml_xy=2, pml_z=3, metal_cells=5, lossy_cells=5, iBc=(1, 1, 0)),
    # a compact metal box: most tiles run the pass, the box's tiles the region sweeps
    dict(L0=70, L1=26, L2=75, outer=10, pml_xy=3, pml_z=4, core=(5, 4), metal_cells=0, lossy_cells=0,
         iBc=(0, 0, 0), metal_box=((33, 37, 12, 15, 36, 40), -20.0 + 2.0j)),
    dict(L0=70, L1=26, L2=75, outer=10, pml_xy=0, pml_z=4, core=(5, 4), metal_cells=0, lossy_cells=3,
         iBc=(1, -1, 0), metal_box=((2, 6, 1, 4, 36, 40), -20.0 + 2.0j)),               # near the x/y wrap
])
@pytest.mark.parametrize("dft", [False, True])
def test_two_step_pass_with_special_cells(kw, dft, monkeypatch):
    """Two steps per pass on grids with lossy and metal cells: the tiles whose stage
    region holds such a cell run the region sweeps and the special-cell passes, the rest
    the pass.  Against the FP64 oracle (fields, monitors, envelope sums) over a run with an
    odd number of steps, and against the one-step kernels; with a DFT set, the region
    sweeps' cells sampled from the scratch and new buffers, the pass's from inside it
    (and the monitors folded from the ring), the DFT against the one-step kernels'."""
    monkeypatch.setenv("SMX_F32_KERNEL", "fused")
    prep = casegen.prepare(small_waveguide(seed=17, **kw))
    assert int(np.count_nonzero(prep.Cont != ord("n")) + np.count_nonzero(prep.Cont == ord("a"))) > 0
    steps = prep.case.time_period[2] * prep.N
    L2 = int(prep.case.L[2])
    planes = [prep.case.iI + 3, L2 // 2, L2 - 9]
    ws = [2 * np.pi / prep.N * f for f in (0.9, 1.0)]
    out = []
    for two in ("2,16", None):
        if two:
            monkeypatch.setenv("SMX_F32_TWO", two)
        else:
            monkeypatch.delenv("SMX_F32_TWO", raising=False)
        with prep.plan(fdtd.F32) as p:
            if dft:
                p.set_dft(planes, ws, chunk=8)
            if two:
                assert p.autotune(0)[3] == 3 and p.kernels.startswith("two_step_f32<32,8,1,true")
            sums = p.run_steps(0, steps - 1)
            p.run_steps(steps - 1, 1, sums)
            out.append((p.read_fields(), p.read_monitors(), sums, p.read_dft() if dft else None))
    if dft:
        D2, D1 = out[0][3], out[1][3]
        assert np.abs(D1).max() > 0 and rel_l2(D2, D1) < 1e-5, rel_l2(D2, D1)
    (H, E), mons, sums, _ = out[0]
    Ho, Eo, mo, so = _oracle_run(prep, 0, prep.case.time_period[2])
    assert np.abs(Ho).max() > 0
    assert rel_l2(H, Ho) < 2e-4 and rel_l2(E, Eo) < 2e-4, (rel_l2(H, Ho), rel_l2(E, Eo))
    for a, b in zip(mons, mo):
        if np.linalg.norm(b) > 0:
            assert rel_l2(a.view(np.float64), b) < 2e-4
    assert np.allclose(sums, so, rtol=1e-14, atol=0)
    (H1, E1), _, _, _ = out[1]
    assert rel_l2(H, H1) < 1e-5 and rel_l2(E, E1) < 1e-5, (rel_l2(H, H1), rel_l2(E, E1))


def test_two_step_refused(monkeypatch):
    """The variants without split slots on a grid with lossy/metal cells, or a PML-free
    core too thin for the two-step box, keep the one-step kernels; so does an x extent
    that is not a multiple of 4 for the 16-byte copies."""
    monkeypatch.setenv("SMX_F32_KERNEL", "fused")
    for two, kw in (("0,16", dict(pml_xy=0, pml_z=0, metal_cells=2, lossy_cells=0)),
                    ("0,16", dict(L0=12, L1=12, pml_xy=4, pml_z=3, metal_cells=0, lossy_cells=0)),
                    ("3,16", dict(L0=37, pml_xy=0, pml_z=0, metal_cells=0, lossy_cells=0))):
        monkeypatch.setenv("SMX_F32_TWO", two)
        prep = casegen.prepare(small_waveguide(seed=15, **kw))
        with prep.plan(fdtd.F32) as p:
            assert p.autotune(0)[3] != 3
            assert "two_step" not in p.kernels


def test_autotune_keeps_state():
    """The autotuner restores fields and monitors, and the run it tunes stays correct."""
    prep = casegen.prepare(small_waveguide(seed=9, **CASES["ports"]))
    with prep.plan(fdtd.F32) as p:
        p.run(0, 1)
        H0, E0 = p.read_fields()
        m0 = p.read_monitors()
        choice = p.autotune(3)
        H1, E1 = p.read_fields()
        m1 = p.read_monitors()
        assert _eq(H0, H1) and _eq(E0, E1)
        assert all(_eq(a, b) for a, b in zip(m0, m1))
        assert choice == (0, 0, 0, 0) or (choice[0] * choice[1] in (256, 512) and choice[2] in (8, 16, 32, 64)
                                          and choice[3] in (1, 2, 3))
        p.run(1, prep.case.time_period[2])
        H, E = p.read_fields()
    Ho, Eo, _, _ = _oracle_run(prep, 0, prep.case.time_period[2])
    assert rel_l2(H, Ho) < 2e-4 and rel_l2(E, Eo) < 2e-4


@pytest.mark.parametrize("kw", [dict(pml_xy=3, pml_z=3, iBc=(0, 0, 0)), dict(pml_xy=0, pml_z=3, iBc=(1, -1, 0))])
def test_two_step_64bit_form(kw, monkeypatch):
    """The two-step pass's form for grids of 4 GiB component groups and more (a buffer
    resource per array and plane, 64-bit stores), forced on a small grid: fields, monitors
    and DFT are bit for bit those of the usual form."""
    prep = casegen.prepare(small_waveguide(seed=19, L0=40, L1=24, L2=36, outer=8, metal_cells=0, lossy_cells=0,
                                           **kw))
    ws = [0.3, 2 * np.pi / prep.N]
    out = []
    for big in ("0", "1"):
        monkeypatch.setenv("SMX_F32_TWO", "2,16")
        monkeypatch.setenv("SMX_F32_BIG", big)
        with prep.plan(fdtd.F32) as p:
            p.set_dft([10, 20], ws, chunk=8)
            assert p.kernels.endswith(",true>" if big == "1" else ",false>"), p.kernels
            p.run(0, prep.case.time_period[2])
            out.append(p.read_fields() + tuple(p.read_monitors()) + (p.read_dft(),))
    assert np.abs(out[0][0]).max() > 0
    for a, b in zip(out[0], out[1]):
        assert np.array_equal(a, b)


def _flagged_tiles(flag, f, s):
    """Tiles of f cells from origin s (wrapping) whose computed cells (1 before the tile, 2
    after it) hold a flagged cell."""
    Lv, k, t = len(flag), 0, 0
    while t * f < Lv:
        k += any(flag[u % Lv] for u in range(s + t * f - 1, s + t * f + f + 2))
        t += 1
    return k


def _split_tile_y_origin(prep, fy=8):
    """y origin of the split-slot two-step tiles that puts the y PML in the fewest tile rows
    (0 on ties), with the count of PML tile rows from 0 and from it."""
    L0, L1 = int(prep.case.L[0]), int(prep.case.L[1])
    tol = 10 * np.finfo(np.float64).eps
    sg = np.abs(prep.sigma)
    fyl = [bool(sg[2 * L0 + 2 * j] > tol or sg[2 * L0 + 2 * j + 1] > tol) for j in range(L1)]
    by, ky = 0, _flagged_tiles(fyl, fy, 0)
    for s in range(1, fy):
        if _flagged_tiles(fyl, fy, s) < ky:
            by, ky = s, _flagged_tiles(fyl, fy, s)
    return by, _flagged_tiles(fyl, fy, 0), ky


@pytest.mark.parametrize("kw", [
    dict(L0=96, L1=26, L2=40, outer=8, pml_xy=4, pml_z=3, iBc=(0, 0, 0), metal_cells=0, lossy_cells=0),
    dict(L0=96, L1=26, L2=40, outer=8, pml_xy=4, pml_z=3, iBc=(1, -1, 0), metal_cells=0, lossy_cells=3,
         metal_box=((86, 90, 5, 9, 18, 22), -20.0 + 2.0j)),
    dict(L0=70, L1=25, L2=36, outer=8, pml_xy=3, pml_z=3, iBc=(0, 0, 0), metal_cells=0, lossy_cells=0),
])
def test_two_step_pml_tile_origins(kw, monkeypatch):
    """The split-slot two-step pass starts its y tiles where the y-PML layers fall in the
    fewest tile rows (the row across the y wrap).  The tiling changes which block computes
    a cell, not its arithmetic: without lossy/metal cells the fields, monitors and DFT are
    bit for bit those of tiles from row 0 (SMX_F32_PML_ORIGIN=0).  With them, the box of
    tiles left to the region sweeps moves with the tiling, and the sweeps round some cells
    differently from the pass: there the two agree to 1e-5.  Both are close to the
    oracle."""
    prep = casegen.prepare(small_waveguide(seed=31, core=(5, 4), **kw))
    oy, ky0, ky = _split_tile_y_origin(prep)
    assert oy != 0 and ky < ky0, (oy, ky0, ky)
    L2 = int(prep.case.L[2])
    planes = [prep.case.iI + 3, L2 // 2]
    ws = [2 * np.pi / prep.N * f for f in (0.9, 1.0)]
    steps = prep.case.time_period[2] * prep.N
    out = []
    for origin in ("0", "1"):
        monkeypatch.setenv("SMX_F32_KERNEL", "fused")
        monkeypatch.setenv("SMX_F32_TWO", "2,16")
        monkeypatch.setenv("SMX_F32_PML_ORIGIN", origin)
        with prep.plan(fdtd.F32) as p:
            p.set_dft(planes, ws, chunk=8)
            assert p.autotune(0)[3] == 3 and p.kernels.startswith("two_step_f32<32,8,1,true"), p.kernels
            sums = p.run_steps(0, steps - 1)
            p.run_steps(steps - 1, 1, sums)
            out.append(p.read_fields() + tuple(p.read_monitors()) + (p.read_dft(),))
    assert np.abs(out[0][0]).max() > 0 and np.abs(out[0][-1]).max() > 0
    special = kw.get("metal_cells", 0) or kw.get("lossy_cells", 0) or kw.get("metal_box")
    for a, b in zip(out[0], out[1]):
        if special:
            assert rel_l2(a, b) < 1e-5
        else:
            assert np.array_equal(a, b)
    Ho, Eo, _, _ = _oracle_run(prep, 0, prep.case.time_period[2])
    assert rel_l2(out[1][0], Ho) < 2e-4 and rel_l2(out[1][1], Eo) < 2e-4


def test_batch_pml_tile_origins(monkeypatch):
    """The fused batch (every plan's pass in one launch) with the y tiles placed for the PML
    against each plan run alone with tiles from row 0: bit for bit."""
    kws = [dict(seed=32, L0=96, L1=26, L2=40, outer=8, pml_xy=4, pml_z=3, core=(5, 4), metal_cells=0,
                lossy_cells=0, iBc=(0, 0, 0)),
           dict(seed=33, L0=70, L1=25, L2=36, outer=8, pml_xy=3, pml_z=3, core=(5, 4), metal_cells=0,
                lossy_cells=0, iBc=(0, 0, 0))]
    preps = [casegen.prepare(small_waveguide(**kw)) for kw in kws]
    assert all(_split_tile_y_origin(pr)[0] != 0 for pr in preps)
    steps = 2 * (min(pr.case.time_period[2] * pr.N for pr in preps) // 2)
    monkeypatch.setenv("SMX_F32_KERNEL", "fused")
    monkeypatch.setenv("SMX_F32_TWO", "2,16")
    alone = []
    monkeypatch.setenv("SMX_F32_PML_ORIGIN", "0")
    for pr in preps:
        with pr.plan(fdtd.F32) as p:
            p.run_steps(0, steps)
            alone.append(p.read_fields() + tuple(p.read_monitors()))
    monkeypatch.delenv("SMX_F32_PML_ORIGIN")
    plans = [pr.plan(fdtd.F32) for pr in preps]
    try:
        info = {}
        fdtd.run_batch(plans, 0, steps, mode=fdtd.BATCH_FUSED, info=info)
        assert info["graph_launches"] == steps // 2
        together = [p.read_fields() + tuple(p.read_monitors()) for p in plans]
    finally:
        for p in plans:
            p.close()
    for a, b in zip(alone, together):
        assert np.abs(a[0]).max() > 0
        for x, y in zip(a, b):
            assert np.array_equal(x, y)
