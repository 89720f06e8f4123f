"""The oracle: TEST INFRASTRUCTURE ONLY.

ctypes bindings to two CPU libraries built by oracle/Makefile:

* ``liboracle.so`` -- this repository's FP64 restatement of the split-field FDTD step of
  SMatrAnEM (``fdtd_oracle.c``; cites FDTDcalc/FDTDsplitfieldMetal.c line by line).
* ``_ref/libsmx_ref.so`` -- the reference's own FDTDsplitfieldMetal.c compiled from its
  source where that source is present (built in the development container only; the
  loader returns None elsewhere).

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg may
import this package, and only as the checker or the timed CPU baseline -- never as the
thing measured or shipped.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libsmx_ref.so")

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_cp = np.ctypeslib.ndpointer(dtype=np.int8, flags="C_CONTIGUOUS")

_MAIN_CALC_ARGS = [_ip, _dp, _dp, _dp, _dp, _ip, _dp, _dp, _ip, _dp, _dp, _dp,
                   _dp, _dp, _dp, _dp, _cp, _ip]

_oracle = None
_ref = None
_oracle_omp = None


def load_oracle():
    """liboracle.so (raises if it has not been built: run `make -C oracle`)."""
    global _oracle
    if _oracle is None:
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError(f"oracle not built: {ORACLE_SO} (run make -C oracle)")
        lib = C.CDLL(ORACLE_SO)
        lib.orc_main_calc.argtypes = _MAIN_CALC_ARGS
        lib.orc_main_calc.restype = C.c_int
        lib.orc_set_tau.argtypes = [C.c_double, C.c_double, C.POINTER(C.c_double),
                                    C.POINTER(C.c_double), C.POINTER(C.c_double)]
        lib.orc_set_tau.restype = C.c_int
        lib.orc_expn_table.argtypes = [C.c_int, _dp]
        lib.orc_set_env.argtypes = [C.c_int, _ip, C.c_int, C.c_int, C.c_int]
        lib.orc_set_env.restype = C.c_double
        lib.orc_set_sigma.argtypes = [_ip, _dp, _dp, _dp]
        lib.orc_remove_sigma.argtypes = [_ip, _dp, _dp, _dp]
        lib.orc_remove_sigma.restype = C.c_int
        lib.orc_set_address.argtypes = [_ip, _dp, _dp, _dp, _cp, _ip]
        lib.orc_set_address.restype = C.c_int
        lib.orc_build_coeffs.argtypes = [_ip, _dp, _dp, _dp, _cp, _ip, C.c_double, C.c_double,
                                         _dp, _dp, _dp, _dp]
        lib.orc_run.argtypes = [_ip, _dp, _dp, _dp, _dp, _ip, _dp, _dp, _ip, _dp, _dp, C.c_int,
                                _ip, C.c_int, C.c_int, _dp, _dp, _dp, _dp, _cp, _ip, _dp]
        lib.orc_run.restype = C.c_int
        lib.orc_run_point.argtypes = lib.orc_run.argtypes + [C.c_int, C.c_int64, _dp, C.c_int64]
        lib.orc_run_point.restype = C.c_int
        lib.orc_point_add.argtypes = [_dp, _dp, C.c_int, C.c_int64, C.c_double]
        lib.orc_s_project.argtypes = [_ip, C.c_int, _dp, _dp, _dp, _dp]
        lib.orc_step_adjoint.argtypes = [_ip, _dp, _dp, _dp, _dp, _ip, _dp, _dp, _cp, _ip, C.c_int]
        lib.orc_step_adjoint.restype = C.c_int
        lib.orc_adjoint_run.argtypes = [_ip, _dp, _dp, _dp, _dp, _ip, _ip, _dp, _dp, C.c_int, _ip, _cp, _ip,
                                        C.c_int, _dp, _dp, C.c_double, C.c_double, C.c_int, C.c_double, _dp]
        lib.orc_adjoint_run.restype = C.c_double
        _oracle = lib
    return _oracle


def load_oracle_omp():
    """liboracle_omp.so, the oracle's main_calc with its sweeps on OpenMP threads (the
    all-cores CPU baseline); None when it has not been built."""
    global _oracle_omp
    if _oracle_omp is None:
        path = os.path.join(os.path.dirname(ORACLE_SO), "liboracle_omp.so")
        if not os.path.exists(path):
            return None
        lib = C.CDLL(path)
        lib.orc_main_calc.argtypes = _MAIN_CALC_ARGS
        lib.orc_main_calc.restype = C.c_int
        _oracle_omp = lib
    return _oracle_omp


def load_ref():
    """The compiled reference (None when oracle/_ref was not built here)."""
    global _ref
    if _ref is None:
        if not os.path.exists(REF_SO):
            return None
        lib = C.CDLL(REF_SO)
        lib.main_calc.argtypes = _MAIN_CALC_ARGS
        lib.main_calc.restype = None
        lib.set_tau.argtypes = [C.c_double, C.c_double, C.POINTER(C.c_int),
                                C.POINTER(C.c_double), C.POINTER(C.c_double)]
        lib.set_tau.restype = None
        lib.set_env.argtypes = [C.c_int, _ip, C.c_int, C.c_int, C.c_int]
        lib.set_env.restype = C.c_double
        lib.set_sigma.argtypes = [_ip, _dp, _dp, _dp]
        lib.set_sigma.restype = None
        lib.remove_sigma.argtypes = [_ip, _dp, _dp, _dp]
        lib.remove_sigma.restype = None
        lib.set_address.argtypes = [_ip, _dp, _dp, _dp, _cp, C.POINTER(C.c_int), _ip]
        lib.set_address.restype = None
        lib.S_calc.argtypes = [_ip, C.c_int, C.c_char_p, _dp, _dp, _dp]
        lib.S_calc.restype = C.c_int
        lib.get_RealNum.argtypes = [C.c_char_p, _ip]
        lib.get_RealNum.restype = C.c_int
        lib.set_Xi.argtypes = [C.c_char_p, C.c_int, C.c_int, _dp]
        lib.set_Xi.restype = None
        lib.set_tilXi.argtypes = [C.c_char_p, C.c_int, C.c_int, _dp]
        lib.set_tilXi.restype = None
        lib.input_data.argtypes = [C.c_char_p, _ip, C.POINTER(C.c_double), _dp, _dp]
        lib.input_data.restype = None
        lib.input_Lf.argtypes = [C.c_char_p, _ip, C.POINTER(C.c_double)]
        lib.input_Lf.restype = None
        lib.num_u2.argtypes = [_ip, C.c_char_p]
        lib.num_u2.restype = C.c_int
        _ref = lib
    return _ref


def i32(x):
    return np.ascontiguousarray(np.asarray(x, dtype=np.int32))


def cont_bytes(cont):
    """Cell classes as an int8 array (from a str, bytes or array of chars)."""
    if isinstance(cont, (bytes, bytearray)):
        return np.frombuffer(bytes(cont), dtype=np.int8).copy()
    if isinstance(cont, str):
        return np.frombuffer(cont.encode(), dtype=np.int8).copy()
    return np.ascontiguousarray(np.asarray(cont, dtype=np.int8))


def main_calc(lib_fn, L, Ah, Bh, Ae, Be, iBc, H, E, iBTI, Hi, Ei, ExpN2, HB, EB, HT, ET, Cont, addr):
    """Call a main_calc-shaped function (oracle or reference) in place."""
    return lib_fn(i32(L), Ah, Bh, Ae, Be, i32(iBc), H, E, i32(iBTI), Hi, Ei,
                  np.ascontiguousarray(ExpN2, dtype=np.float64), HB, EB, HT, ET,
                  cont_bytes(Cont), i32(addr))


def set_tau(omega, min_dt):
    lib = load_oracle()
    r, i, dt = C.c_double(), C.c_double(), C.c_double()
    N = lib.orc_set_tau(omega, min_dt, C.byref(r), C.byref(i), C.byref(dt))
    return N, r.value, i.value, dt.value


def set_env(N, tp, j_half, j_time, j_period):
    return load_oracle().orc_set_env(N, i32(tp), j_half, j_time, j_period)


def expn_table(N):
    out = np.zeros(4 * N)
    load_oracle().orc_expn_table(N, out)
    return out


def setup_from_media(L, mu, eps, omega_dt=None, iR=None, iI=None):
    """sigma, classes, addresses and coefficients from stretched media (F:75-143).

    mu, eps: complex128[3n] as the Med files give them (PML stretch included); they are
    copied, then the stretch is removed as the reference does.  Returns a dict.
    """
    lib = load_oracle()
    L = i32(L)
    n = int(L[0]) * int(L[1]) * int(L[2])
    mu_r = np.ascontiguousarray(np.asarray(mu, dtype=np.complex128)).view(np.float64).copy()
    eps_r = np.ascontiguousarray(np.asarray(eps, dtype=np.complex128)).view(np.float64).copy()
    sigma = np.zeros(2 * (int(L[0]) + int(L[1]) + int(L[2])))
    lib.orc_set_sigma(L, mu_r, eps_r, sigma)
    rc = lib.orc_remove_sigma(L, mu_r, eps_r, sigma)
    if rc:
        raise ValueError(("mu", "epsi")[rc - 1] + " error in remove_sigma")
    cont = np.zeros(n, dtype=np.int8)
    addr = np.zeros(n, dtype=np.int32)
    lnotn = lib.orc_set_address(L, sigma, mu_r, eps_r, cont, addr)
    out = dict(sigma=sigma, mu=mu_r, eps=eps_r, Cont=cont, addr=addr, Lnotn=lnotn)
    if iR is not None:
        clen = 3 * (n + lnotn)
        Ah, Bh, Ae, Be = (np.zeros(clen) for _ in range(4))
        lib.orc_build_coeffs(L, mu_r, eps_r, sigma, cont, addr, iR, iI, Ah, Bh, Ae, Be)
        out.update(Ah=Ah, Bh=Bh, Ae=Ae, Be=Be)
    return out


def run(L, coeffs, iBc, iBTI, Hinc, Einc, N, tp, P0=0, P1=None, fields=None, monitors=None, point=None):
    """orc_run: the time loop of F:153-167 over periods [P0, P1).  point = (c, cell, amp)
    adds a point source on E component c of linear cell `cell` after every step k with
    amplitude amp[k] (orc_run_point).  Returns (H, E, (HB, EB, HT, ET), sums)."""
    lib = load_oracle()
    L = i32(L)
    n = int(L[0]) * int(L[1]) * int(L[2])
    plane = int(L[0]) * int(L[1])
    lnotn = coeffs["Lnotn"]
    flen = 3 * (n + lnotn + 1)
    if fields is None:
        H, E = np.zeros(flen), np.zeros(flen)
    else:
        H, E = (np.ascontiguousarray(f, dtype=np.float64).copy() for f in fields)
    if monitors is None:
        mons = [np.zeros(4 * plane) for _ in range(4)]
    else:
        mons = [np.ascontiguousarray(m, dtype=np.float64).copy() for m in monitors]
    sums = np.zeros(2)
    P1 = tp[2] if P1 is None else P1
    args = (L, coeffs["Ah"], coeffs["Bh"], coeffs["Ae"], coeffs["Be"], i32(iBc), H, E,
            i32(iBTI), np.ascontiguousarray(Hinc, dtype=np.complex128).view(np.float64),
            np.ascontiguousarray(Einc, dtype=np.complex128).view(np.float64), N, i32(tp),
            P0, P1, *mons, cont_bytes(coeffs["Cont"]), i32(coeffs["addr"]), sums)
    if point is None:
        rc = lib.orc_run(*args)
    else:
        c, cell, amp = point
        amp = np.ascontiguousarray(amp, dtype=np.float64)
        rc = lib.orc_run_point(*args, int(c), int(cell), amp, amp.size)
    if rc:
        raise RuntimeError(f"iBTIin[{rc - 1}] error!")
    return H, E, tuple(mons), sums


def adjoint_run(L, st, iBc, iBTI, Hinc, Einc, N, tp, zT, w, iR, iI):
    """orc_adjoint_run: (J, dJ/d eps_c [3n]) of the port-DFT objective on plane zT."""
    lib = load_oracle()
    n = int(L[0]) * int(L[1]) * int(L[2])
    grad = np.zeros(3 * n)
    J = lib.orc_adjoint_run(i32(L), st["Ah"], st["Bh"], st["Ae"], st["Be"], i32(iBc), i32(iBTI),
                            np.ascontiguousarray(Hinc, dtype=np.complex128).view(np.float64),
                            np.ascontiguousarray(Einc, dtype=np.complex128).view(np.float64), N, i32(tp),
                            cont_bytes(st["Cont"]), i32(st["addr"]), st["Lnotn"], st["eps"], st["sigma"],
                            iR, iI, int(zT), float(w), grad)
    return J, grad
