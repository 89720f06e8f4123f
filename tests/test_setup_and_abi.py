"""Host side of libsmx_fdtd on the CPU: the library loads and exports every symbol of
include/smx_fdtd.h, and its setup functions equal the oracle's bit for bit."""
import os
import re

import numpy as np
import pytest

import oracle
from smatranem_amd import casegen, fdtd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_symbols():
    text = open(os.path.join(ROOT, "include", "smx_fdtd.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(smx_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_header_symbols():
    lib = fdtd.lib()
    names = _declared_symbols()
    assert len(names) >= 20
    for name in names:
        assert hasattr(lib, name), name


def test_library_exports_nothing_unprefixed():
    """Every C symbol the library defines carries the smx_ prefix (C++ ones are mangled):
    an unprefixed C name such as `advance` can be bound, by the dynamic linker, to a
    same-named function of another library in the process."""
    import shutil
    import subprocess
    if shutil.which("nm") is None:
        pytest.skip("nm not available")
    out = subprocess.run(["nm", "-D", "--defined-only", fdtd.library_path], capture_output=True, text=True,
                         check=True).stdout
    bad = []
    for line in out.splitlines():
        parts = line.split()
        if len(parts) == 3 and parts[1] in "TWD":
            name = parts[2]
            if not (name.startswith("smx_") or name.startswith("_Z") or name.startswith("__hip")
                    or name.startswith("_fini") or name.startswith("_init")):
                bad.append(name)
    assert not bad, bad


def test_status_strings():
    lib = fdtd.lib()
    assert lib.smx_status_string(0) == b"ok"
    assert b"PML" in lib.smx_status_string(100)


@pytest.mark.parametrize("omega,min_dt", [(2 * np.pi * 0.1, 0.5715767664977296), (0.314159, 0.84),
                                          (2.0, 0.1)])
def test_set_tau_matches_oracle(omega, min_dt):
    assert fdtd.set_tau(omega, min_dt) == oracle.set_tau(omega, min_dt)


def test_set_env_and_expn_match_oracle():
    for N in (1, 9, 24):
        np.testing.assert_array_equal(fdtd.expn_table(N).view(np.float64), oracle.expn_table(N))
        for tp in ((2, 1, 8), (3, 0, 7)):
            for P in range(tp[2]):
                for t in range(N + 1):
                    for h in (0, 1):
                        assert fdtd.set_env(N, tp, h, t, P) == oracle.set_env(N, tp, h, t, P)
    with pytest.raises(fdtd.SmxError):
        fdtd.set_env(9, (2, 1, 8), 0, 0, 8)


@pytest.mark.parametrize("seed,scale", [(0, 0.0), (3, 0.3)])
def test_setup_chain_matches_oracle(seed, scale):
    case = casegen.waveguide_case(10, 8, 20, 5, 2, 2, core=(4, 3), metal_cells=5, lossy_cells=3,
                                  lossy_mu_cells=4, seed=seed, scale_amp=scale)
    prep = casegen.prepare(case)
    N, iR, iI, _ = oracle.set_tau(case.omega, case.min_dt)
    ref = oracle.setup_from_media(case.L, case.mu, case.eps, iR=iR, iI=iI)
    np.testing.assert_array_equal(prep.sigma, ref["sigma"])
    np.testing.assert_array_equal(prep.mu.view(np.float64), ref["mu"])
    np.testing.assert_array_equal(prep.eps.view(np.float64), ref["eps"])
    np.testing.assert_array_equal(prep.Cont, ref["Cont"])
    np.testing.assert_array_equal(prep.addr, ref["addr"])
    assert prep.Lnotn == ref["Lnotn"]
    assert (prep.N, prep.invtau_R, prep.invtau_I) == (N, iR, iI)
    Ah, Bh, Ae, Be = fdtd.build_coeffs(case.L, prep.mu, prep.eps, prep.sigma, prep.Cont, prep.addr,
                                       prep.Lnotn, iR, iI)
    for a, b in zip((Ah, Bh, Ae, Be), (ref["Ah"], ref["Bh"], ref["Ae"], ref["Be"])):
        np.testing.assert_array_equal(a, b)
    kinds = set(np.unique(prep.Cont).tobytes())
    assert {ord("n"), ord("a")} <= kinds and kinds & {ord("e"), ord("h"), ord("b")}
    mu_n = prep.mu.reshape(-1, 3)[prep.Cont == ord("n")]
    assert np.count_nonzero(mu_n.imag) >= 3, "no lossy-mu 'n' cell in the case"


def test_plan_without_device_fails_loudly():
    """On a host with no GPU the plan refuses to run; no CPU fallback exists."""
    import torch
    if torch.cuda.device_count() > 0:
        pytest.skip("a GPU is visible")
    case = casegen.waveguide_case(6, 6, 12, 3, 0, 2, core=(2, 2), seed=1)
    prep = casegen.prepare(case)
    with pytest.raises(fdtd.SmxError) as e:
        prep.plan(fdtd.F64)
    assert e.value.status == 3
