"""Pin the oracle against the reference compiled from its own source (oracle/_ref).

Bitwise comparisons: the oracle restates FDTDcalc/FDTDsplitfieldMetal.c with the same
FP64 operation order and no contraction, so every value must be identical.  Skipped
where oracle/_ref was not built (it needs the reference sources)."""
import ctypes as C
import os

import numpy as np
import pytest

import oracle
from tests.cases import call_main_calc, random_state

ref = oracle.load_ref()
orc = oracle.load_oracle()
needs_ref = pytest.mark.skipif(ref is None, reason="oracle/_ref not built (no reference sources)")


def _eq(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return a.shape == b.shape and np.array_equal(a, b)


@needs_ref
@pytest.mark.parametrize("seed,L", [(0, (7, 6, 9)), (1, (5, 4, 6)), (2, (1, 5, 7)), (3, (6, 1, 5)),
                                    (4, (4, 5, 1)), (5, (8, 8, 8)), (6, (3, 3, 2))])
def test_main_calc_bitwise(seed, L):
    st = random_state(seed, L=L)
    Hr, Er, mr, _ = call_main_calc(ref.main_calc, st)
    Ho, Eo, mo, rc = call_main_calc(orc.orc_main_calc, st)
    assert rc == 0
    assert _eq(Hr, Ho) and _eq(Er, Eo)
    for a, b in zip(mr, mo):
        assert _eq(a, b)


@needs_ref
@pytest.mark.parametrize("iBc", [(1, 1, 1), (-1, 0, 1), (0, 0, 0), (-1, -1, -1)])
def test_main_calc_bitwise_multistep(iBc):
    st = random_state(11, L=(6, 5, 7), iBc=iBc)
    Hr, Er, mr = st["H"].copy(), st["E"].copy(), [m.copy() for m in st["mons"]]
    Ho, Eo, mo = st["H"].copy(), st["E"].copy(), [m.copy() for m in st["mons"]]
    for _ in range(5):
        call_main_calc(ref.main_calc, st, Hr, Er, mr)
        call_main_calc(orc.orc_main_calc, st, Ho, Eo, mo)
    assert _eq(Hr, Ho) and _eq(Er, Eo)
    assert all(_eq(a, b) for a, b in zip(mr, mo))


@needs_ref
def test_main_calc_only_normal_cells():
    st = random_state(21, L=(6, 6, 6), probs=(1.0, 0, 0, 0, 0))
    Hr, Er, _, _ = call_main_calc(ref.main_calc, st)
    Ho, Eo, _, _ = call_main_calc(orc.orc_main_calc, st)
    assert _eq(Hr, Ho) and _eq(Er, Eo)


@needs_ref
def test_set_env_bitwise():
    for N in (1, 7, 18, 24):
        for tp in ((2, 1, 8), (10, 80, 111), (3, 0, 6), (1, 5, 7)):
            tpa = oracle.i32(tp)
            for P in range(tp[2]):
                for t in range(N + 1):
                    for h in (0, 1):
                        a = ref.set_env(N, tpa, h, t, P)
                        b = orc.orc_set_env(N, tpa, h, t, P)
                        assert a == b or (np.isnan(a) and np.isnan(b)), (N, tp, P, t, h)


@needs_ref
@pytest.mark.parametrize("omega,min_dt", [(2 * np.pi * 0.1, 1 / np.sqrt(3) * 0.99),
                                          (2 * np.pi / 20, 0.8431), (0.3141592653589793, 0.5),
                                          (2.0, 0.1), (1e-3, 0.577)])
def test_set_tau_bitwise(omega, min_dt):
    N, r, i = C.c_int(), C.c_double(), C.c_double()
    ref.set_tau(omega, min_dt, C.byref(N), C.byref(r), C.byref(i))
    No, ro, io, _ = oracle.set_tau(omega, min_dt)
    assert (N.value, r.value, i.value) == (No, ro, io)


def _stretched_case(seed):
    from smatranem_amd import casegen
    return casegen.waveguide_case(10, 8, 20, 5, 2, 2, core=(4, 3), metal_cells=5, lossy_cells=3,
                                  seed=seed, scale_amp=0.2)


@needs_ref
@pytest.mark.parametrize("seed", [0, 1])
def test_sigma_address_bitwise(seed):
    case = _stretched_case(seed)
    L = oracle.i32(case.L)
    mu = case.mu.view(np.float64).copy()
    eps = case.eps.view(np.float64).copy()
    m = 2 * (int(L[0]) + int(L[1]) + int(L[2]))
    s_ref, s_orc = np.zeros(m), np.zeros(m)
    ref.set_sigma(L, mu, eps, s_ref)
    orc.orc_set_sigma(L, mu, eps, s_orc)
    assert _eq(s_ref, s_orc)
    assert np.count_nonzero(s_ref) > 0
    mu_r, eps_r, mu_o, eps_o = mu.copy(), eps.copy(), mu.copy(), eps.copy()
    ref.remove_sigma(L, mu_r, eps_r, s_ref)
    assert orc.orc_remove_sigma(L, mu_o, eps_o, s_orc) == 0
    assert _eq(mu_r, mu_o) and _eq(eps_r, eps_o)
    n = case.n
    cr, co = np.zeros(n, np.int8), np.zeros(n, np.int8)
    ar, ao = np.zeros(n, np.int32), np.zeros(n, np.int32)
    lr = C.c_int()
    ref.set_address(L, s_ref, mu_r, eps_r, cr, C.byref(lr), ar)
    lo = orc.orc_set_address(L, s_orc, mu_o, eps_o, co, ao)
    assert lr.value == lo and _eq(cr, co) and _eq(ar, ao)
    assert set(np.unique(co).tobytes()) >= set(b"naeh") or set(np.unique(co).tobytes()) >= set(b"nae")


def _write_mode_files(tmp, prefix, L, Xi, n_real, omega, which="b"):
    """Xi and Mode files in the layout ModeCalc/Mcalc.c:191-228 writes."""
    P = L[0] * L[1]
    NA = 2 * P
    with open(f"{prefix}_{which}Xi.dat", "w") as f:
        f.write("# Xi file written by the test\n# columns: Xi[j], Xi[j+NA]\n")
        for j in range(NA):
            f.write(f"# Xi[{j}], Xi[{j + NA}]\n")
            for r in range(2 * NA):
                a, b = Xi[r, j], Xi[r, j + NA]
                f.write(f"{a.real:.20e} {a.imag:.20e} {b.real:.20e} {b.imag:.20e}\n")
    with open(f"{prefix}_{which}Mode.dat", "w") as f:
        f.write(f"# Omega              : k/k_0 = {omega:.20e}\n")
        f.write(f"# Eigen-value info.  : real num./2 = {n_real}, comp. num./2 = {NA - n_real}\n")
        for j in range(NA):
            f.write(f"theta[{j}]=({0.1 * j:.6e}, {0.0:.6e}), theta[{j + NA}]=({-0.1 * j:.6e}, {0.0:.6e})\n")


def tilxi_columns(Xi, n_real, cols):
    """tilXi columns from Xi (F:617-623): rows rotated by NA; propagating columns keep
    j and negate j+NA; evanescent columns swap."""
    NA = Xi.shape[0] // 2
    out = []
    for col in cols:
        j = col % NA
        a, b = Xi[:, j], Xi[:, j + NA]
        if j < n_real:
            v = a if col < NA else -b
        else:
            v = b if col < NA else a
        out.append(np.roll(v, NA))
    return np.array(out)


@needs_ref
def test_s_calc_bitwise(tmp_path):
    rng = np.random.default_rng(5)
    L = oracle.i32([3, 2, 10, 4])
    P = 6
    NA = 2 * P
    Xi = rng.normal(size=(2 * NA, 2 * NA)) + 1j * rng.normal(size=(2 * NA, 2 * NA))
    n_real = 4
    prefix = str(tmp_path / "joint")
    _write_mode_files(tmp_path, prefix, L, Xi, n_real, 0.5)
    HB = rng.normal(size=4 * P)
    EB = rng.normal(size=4 * P)
    SB = np.zeros(8 * P)
    nr = ref.S_calc(L, 1, prefix.encode(), HB, EB, SB)
    assert nr == n_real
    cols = list(range(n_real)) + [NA + m for m in range(n_real)]
    tcols = tilxi_columns(Xi, n_real, cols)
    S_o = np.zeros(2 * len(cols))
    orc.orc_s_project(L, len(cols), np.ascontiguousarray(tcols).view(np.float64).reshape(-1), HB, EB, S_o)
    S_r = SB.view(np.complex128)[cols]
    assert _eq(S_r.view(np.float64), S_o)


@pytest.mark.parametrize("seed,L", [(31, (9, 7, 11)), (32, (5, 6, 17))])
def test_openmp_build_bitwise(seed, L):
    """The OpenMP build of the oracle (the all-cores CPU baseline) gives the serial
    oracle's fields and monitors bit for bit: each cell's arithmetic is unchanged."""
    omp = oracle.load_oracle_omp()
    if omp is None:
        pytest.skip("liboracle_omp.so not built")
    st = random_state(seed, L=L)
    Ha, Ea, ma, rca = call_main_calc(omp.orc_main_calc, st)
    Hb, Eb, mb, rcb = call_main_calc(oracle.load_oracle().orc_main_calc, st)
    assert rca == rcb == 0
    assert np.array_equal(Ha, Hb) and np.array_equal(Ea, Eb)
    for a, b in zip(ma, mb):
        assert np.array_equal(a, b)
