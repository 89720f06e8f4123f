"""ctypes mirror of include/smx_fdtd.h.

Names and argument meanings follow the reference's functions
(FDTDcalc/FDTDsplitfieldMetal.c): ``main_calc`` (F:242), ``set_tau`` (F:515),
``set_env`` (F:452), ``set_sigma`` (F:858), ``remove_sigma`` (F:822), ``set_address``
(F:188) and the coefficient build of ``main`` (F:106-143).  Where the reference calls
``exit(EXIT_FAILURE)`` these raise :class:`SmxError` carrying the reference's message.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
library_path = os.path.join(HERE, "libsmx_fdtd.so")

F64 = 0
F32 = 1

_STATUS = {0: "ok", 1: "invalid argument", 2: "HIP runtime error", 3: "no HIP device",
           4: "device out of memory", 5: "call out of order", 100: "iBTIin error",
           101: "f_env error!"}


class SmxError(RuntimeError):
    def __init__(self, status, detail=""):
        self.status = status
        super().__init__(f"{_STATUS.get(status, status)}: {detail}")


_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_cp = np.ctypeslib.ndpointer(dtype=np.int8, flags="C_CONTIGUOUS")


class _Grid(C.Structure):
    _fields_ = [("L", C.c_int * 4), ("iBc", C.c_int * 3), ("iBTIin", C.c_int * 4),
                ("precision", C.c_int), ("Lnotn", C.c_int), ("invtau_R", C.c_double),
                ("invtau_I", C.c_double)]


_lib = None


def lib():
    """Load libsmx_fdtd.so; raises if it has not been built (no fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(library_path):
        raise RuntimeError(f"libsmx_fdtd.so not built at {library_path}: run "
                           "`python -c 'import __graft_entry__ as g; g.build()'` or "
                           "`make -C smatranem_amd`")
    L = C.CDLL(library_path)
    L.smx_status_string.restype = C.c_char_p
    L.smx_status_string.argtypes = [C.c_int]
    L.smx_last_error.restype = C.c_char_p
    L.smx_last_error.argtypes = []
    L.smx_main_calc.argtypes = [_ip, _dp, _dp, _dp, _dp, _ip, _dp, _dp, _ip, _dp, _dp, _dp,
                                _dp, _dp, _dp, _dp, _cp, _ip]
    L.smx_main_calc.restype = C.c_int
    L.smx_main_calc_noalloc.argtypes = L.smx_main_calc.argtypes
    L.smx_main_calc_noalloc.restype = C.c_int
    L.smx_plan_create.argtypes = [C.POINTER(_Grid), _dp, _dp, _dp, _cp, _ip, C.POINTER(C.c_void_p)]
    L.smx_plan_create.restype = C.c_int
    L.smx_plan_create_coeffs.argtypes = [C.POINTER(_Grid), _dp, _dp, _dp, _dp, _cp, _ip,
                                         C.POINTER(C.c_void_p)]
    L.smx_plan_create_coeffs.restype = C.c_int
    L.smx_plan_destroy.argtypes = [C.c_void_p]
    L.smx_plan_field_len.argtypes = [C.c_void_p]
    L.smx_plan_field_len.restype = C.c_size_t
    for name in ("smx_plan_write_fields", "smx_plan_read_fields"):
        getattr(L, name).argtypes = [C.c_void_p, _dp, _dp]
        getattr(L, name).restype = C.c_int
    L.smx_plan_read_monitors.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp]
    L.smx_plan_read_monitors.restype = C.c_int
    L.smx_plan_write_monitors.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    L.smx_plan_write_monitors.restype = C.c_int
    L.smx_plan_step.argtypes = [C.c_void_p, _dp, _dp, _dp]
    L.smx_plan_step.restype = C.c_int
    L.smx_plan_set_source.argtypes = [C.c_void_p, _dp, _dp, C.c_int, _ip]
    L.smx_plan_set_source.restype = C.c_int
    L.smx_plan_run.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_void_p]
    L.smx_plan_run.restype = C.c_int
    L.smx_plan_run_steps.argtypes = [C.c_void_p, C.c_longlong, C.c_longlong, C.c_void_p]
    L.smx_plan_run_steps.restype = C.c_int
    L.smx_plan_stream.argtypes = [C.c_void_p]
    L.smx_plan_stream.restype = C.c_void_p
    L.smx_plan_set_dft.argtypes = [C.c_void_p, C.c_int, _ip, C.c_int, _dp, C.c_int]
    L.smx_plan_set_dft.restype = C.c_int
    L.smx_plan_read_dft.argtypes = [C.c_void_p, _dp]
    L.smx_plan_read_dft.restype = C.c_int
    L.smx_plan_set_point_source.argtypes = [C.c_void_p, C.c_int, _ip, C.c_longlong, C.c_void_p]
    L.smx_plan_set_point_source.restype = C.c_int
    L.smx_adjoint_forward.argtypes = [C.c_void_p, C.c_longlong, C.c_int, C.c_int, C.c_double, _ip,
                                      C.POINTER(C.c_double)]
    L.smx_adjoint_forward.restype = C.c_int
    L.smx_adjoint_backward.argtypes = [C.c_void_p, _dp]
    L.smx_adjoint_backward.restype = C.c_int
    L.smx_plan_kernels.argtypes = [C.c_void_p]
    L.smx_plan_kernels.restype = C.c_char_p
    L.smx_plan_run_batch.argtypes = [C.POINTER(C.c_void_p), C.c_int, C.c_longlong, C.c_longlong,
                                     C.POINTER(C.c_double)]
    L.smx_plan_run_batch.restype = C.c_int
    L.smx_plan_run_batch_mode.argtypes = [C.POINTER(C.c_void_p), C.c_int, C.c_longlong, C.c_longlong,
                                          C.POINTER(C.c_double), C.c_int, C.POINTER(C.c_int)]
    L.smx_plan_run_batch_mode.restype = C.c_int
    L.smx_plan_autotune.argtypes = [C.c_void_p, C.c_int, _ip]
    L.smx_plan_autotune.restype = C.c_int
    L.smx_plan_sync.argtypes = [C.c_void_p]
    L.smx_plan_sync.restype = C.c_int
    L.smx_plan_flush.argtypes = [C.c_void_p]
    L.smx_plan_flush.restype = C.c_int
    L.smx_plan_project.argtypes = [C.c_void_p, C.c_int, _dp, C.c_void_p, C.c_void_p]
    L.smx_plan_project.restype = C.c_int
    L.smx_set_tau.argtypes = [C.c_double, C.c_double, C.POINTER(C.c_double), C.POINTER(C.c_double),
                              C.POINTER(C.c_double)]
    L.smx_set_tau.restype = C.c_int
    L.smx_set_env.argtypes = [C.c_int, _ip, C.c_int, C.c_int, C.c_int]
    L.smx_set_env.restype = C.c_double
    L.smx_expn_table.argtypes = [C.c_int, _dp]
    L.smx_set_sigma.argtypes = [_ip, _dp, _dp, _dp]
    L.smx_remove_sigma.argtypes = [_ip, _dp, _dp, _dp]
    L.smx_remove_sigma.restype = C.c_int
    L.smx_set_address.argtypes = [_ip, _dp, _dp, _dp, _cp, _ip]
    L.smx_set_address.restype = C.c_int
    L.smx_build_coeffs.argtypes = [_ip, _dp, _dp, _dp, _cp, _ip, C.c_int, C.c_double, C.c_double,
                                   _dp, _dp, _dp, _dp]
    _lib = L
    return L


def _check(rc):
    if rc:
        detail = lib().smx_last_error().decode(errors="replace")
        raise SmxError(rc, detail)


def _i32(x):
    return np.ascontiguousarray(np.asarray(x, dtype=np.int32))


def _f64(x):
    return np.ascontiguousarray(np.asarray(x, dtype=np.float64))


def _cplx(x):
    """complex128 array -> interleaved float64 view (the C99 double complex layout)."""
    a = np.asarray(x)
    if np.iscomplexobj(a):
        return np.ascontiguousarray(a, dtype=np.complex128).view(np.float64)
    return _f64(a)


def _cont(c):
    if isinstance(c, (bytes, bytearray)):
        return np.frombuffer(bytes(c), dtype=np.int8).copy()
    if isinstance(c, str):
        return np.frombuffer(c.encode(), dtype=np.int8).copy()
    return np.ascontiguousarray(np.asarray(c, dtype=np.int8))


# ---------------------------------------------------------------------------------------
# The reference's setup functions (host, FP64, bit-identical)
# ---------------------------------------------------------------------------------------

def set_tau(omega, min_dt):
    """set_tau (F:515-524) -> (N_2pi, invtau_R, invtau_I, Dt)."""
    r, i, dt = C.c_double(), C.c_double(), C.c_double()
    n = lib().smx_set_tau(omega, min_dt, C.byref(r), C.byref(i), C.byref(dt))
    return n, r.value, i.value, dt.value


def set_env(N, time_period, j_half, j_time, j_period):
    """set_env (F:452-475); raises where the reference prints 'f_env error!' and exits."""
    v = lib().smx_set_env(N, _i32(time_period), j_half, j_time, j_period)
    if np.isnan(v):
        raise SmxError(101, "f_env error!")
    return v


def expn_table(N):
    """ExpN (F:93-99) as complex128[2N]."""
    out = np.zeros(4 * N)
    lib().smx_expn_table(N, out)
    return out.view(np.complex128)


def set_sigma(L, mu, eps):
    """set_sigma (F:858-926): mu, eps complex128[3n] -> sigma[2(L0+L1+L2)]."""
    sigma = np.zeros(2 * (int(L[0]) + int(L[1]) + int(L[2])))
    lib().smx_set_sigma(_i32(L), _cplx(mu), _cplx(eps), sigma)
    return sigma


def remove_sigma(L, mu, eps, sigma):
    """remove_sigma (F:822-853) in place on complex128 arrays."""
    m, e = _cplx(mu), _cplx(eps)
    rc = lib().smx_remove_sigma(_i32(L), m, e, _f64(sigma))
    if rc:
        raise SmxError(1, ("mu", "epsi")[rc - 1] + " error in remove_sigma")
    return m.view(np.complex128), e.view(np.complex128)


def set_address(L, sigma, mu, eps):
    """set_address (F:188-233) -> (Cont int8[n], address_notn int32[n], Lnotn)."""
    n = int(L[0]) * int(L[1]) * int(L[2])
    cont = np.zeros(n, dtype=np.int8)
    addr = np.zeros(n, dtype=np.int32)
    lnotn = lib().smx_set_address(_i32(L), _f64(sigma), _cplx(mu), _cplx(eps), cont, addr)
    return cont, addr, lnotn


def build_coeffs(L, mu, eps, sigma, cont, addr, lnotn, invtau_R, invtau_I):
    """Coefficient build of main (F:106-143) -> (Ah, Bh, Ae, Be)."""
    n = int(L[0]) * int(L[1]) * int(L[2])
    out = [np.zeros(3 * (n + lnotn)) for _ in range(4)]
    lib().smx_build_coeffs(_i32(L), _cplx(mu), _cplx(eps), _f64(sigma), _cont(cont), _i32(addr),
                           lnotn, invtau_R, invtau_I, *out)
    return tuple(out)


# ---------------------------------------------------------------------------------------
# Parity entry point and plans
# ---------------------------------------------------------------------------------------

def main_calc(L, Ah, Bh, Ae, Be, iBc, H_t, E_t, iBTIin, Hi, Ei, ExpN2, HB, EB, HT, ET, Cont,
              address_notn):
    """main_calc (F:242) on the GPU: arrays updated in place, as in the reference."""
    rc = lib().smx_main_calc(_i32(L), _f64(Ah), _f64(Bh), _f64(Ae), _f64(Be), _i32(iBc), H_t, E_t,
                             _i32(iBTIin), _f64(Hi), _f64(Ei), _cplx(ExpN2), HB, EB, HT, ET,
                             _cont(Cont), _i32(address_notn))
    _check(rc)


def main_calc_noalloc(L, Ah, Bh, Ae, Be, iBc, H_t, E_t, iBTIin, Hi, Ei, ExpN2, HB, EB, HT, ET,
                      Cont, address_notn):
    """main_calc (F:242) on the GPU through smx_main_calc_noalloc: the arrays are staged
    in the library's fixed device scratch (NOALLOC_MAX_BYTES) instead of buffers
    allocated per call; same results as :func:`main_calc`."""
    rc = lib().smx_main_calc_noalloc(_i32(L), _f64(Ah), _f64(Bh), _f64(Ae), _f64(Be), _i32(iBc), H_t,
                                     E_t, _i32(iBTIin), _f64(Hi), _f64(Ei), _cplx(ExpN2), HB, EB, HT,
                                     ET, _cont(Cont), _i32(address_notn))
    _check(rc)


NOALLOC_MAX_BYTES = 64 << 20


BATCH_STREAMS = 0
BATCH_GRAPH = 1
BATCH_FUSED = 2


def run_batch(plans, step0, nsteps, mode=None, info=None):
    """smx_plan_run_batch: steps [step0, step0 + nsteps) of independent plans, each on its
    own stream.  ``mode`` None is the library default (BATCH_GRAPH: repeating segments
    replayed from one captured HIP graph); BATCH_STREAMS launches every step from the
    host; BATCH_FUSED runs every two-step pass of all plans in one launch (FP32 plans
    without lossy/metal cells or DFT, each on the split-slot two-step kernel; others fall
    back to BATCH_GRAPH).  ``info`` (a dict, optional) receives ``graph_launches`` (for
    BATCH_FUSED the launches of the batch kernel).  Returns the envelope sums,
    float64[len(plans), 2]."""
    handles = (C.c_void_p * len(plans))(*[p._h.value for p in plans])
    sums = np.zeros((len(plans), 2))
    sp = sums.ctypes.data_as(C.POINTER(C.c_double))
    if mode is None and info is None:
        _check(lib().smx_plan_run_batch(handles, len(plans), int(step0), int(nsteps), sp))
    else:
        launches = C.c_int(0)
        _check(lib().smx_plan_run_batch_mode(handles, len(plans), int(step0), int(nsteps), sp,
                                             BATCH_GRAPH if mode is None else int(mode),
                                             C.byref(launches)))
        if info is not None:
            info["graph_launches"] = launches.value
    return sums


@dataclass
class Grid:
    """Scalars main() holds before its time loop (F:57-104)."""
    L: tuple
    iBc: tuple = (0, 0, 0)
    iBTIin: tuple = (0, 1, 0, 1)
    precision: int = F64
    Lnotn: int = 0
    invtau_R: float = 0.0
    invtau_I: float = 0.0

    def _c(self):
        g = _Grid()
        g.L[:] = [int(v) for v in self.L]
        g.iBc[:] = [int(v) for v in self.iBc]
        g.iBTIin[:] = [int(v) for v in self.iBTIin]
        g.precision = int(self.precision)
        g.Lnotn = int(self.Lnotn)
        g.invtau_R = float(self.invtau_R)
        g.invtau_I = float(self.invtau_I)
        return g

    @property
    def n(self):
        return int(self.L[0]) * int(self.L[1]) * int(self.L[2])

    @property
    def plane(self):
        return int(self.L[0]) * int(self.L[1])


class Plan:
    """Device-resident FDTD state of one grid (smx_plan_* in include/smx_fdtd.h)."""

    def __init__(self, grid: Grid, *, mu=None, eps=None, sigma=None, Cont=None, address_notn=None,
                 coeffs=None):
        self.grid = grid
        self._h = C.c_void_p()
        g = grid._c()
        if coeffs is not None:
            Ah, Bh, Ae, Be = (_f64(c) for c in coeffs)
            rc = lib().smx_plan_create_coeffs(C.byref(g), Ah, Bh, Ae, Be, _cont(Cont),
                                              _i32(address_notn), C.byref(self._h))
        else:
            rc = lib().smx_plan_create(C.byref(g), _cplx(mu), _cplx(eps), _f64(sigma), _cont(Cont),
                                       _i32(address_notn), C.byref(self._h))
        _check(rc)
        self.flen = lib().smx_plan_field_len(self._h)

    def close(self):
        if self._h:
            lib().smx_plan_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def write_fields(self, H_t, E_t):
        _check(lib().smx_plan_write_fields(self._h, _f64(H_t), _f64(E_t)))

    def read_fields(self):
        H, E = np.zeros(self.flen), np.zeros(self.flen)
        _check(lib().smx_plan_read_fields(self._h, H, E))
        return H, E

    def write_monitors(self, HB=None, EB=None, HT=None, ET=None):
        arrs = [None if m is None else _cplx(m) for m in (HB, EB, HT, ET)]
        ptrs = [None if a is None else a.ctypes.data_as(C.c_void_p) for a in arrs]
        _check(lib().smx_plan_write_monitors(self._h, *ptrs))

    def read_monitors(self):
        """(HB, EB, HT, ET) as complex128[2*L0*L1]."""
        out = [np.zeros(4 * self.grid.plane) for _ in range(4)]
        _check(lib().smx_plan_read_monitors(self._h, *out))
        return tuple(o.view(np.complex128) for o in out)

    def step(self, Hi, Ei, ExpN2):
        _check(lib().smx_plan_step(self._h, _f64(Hi), _f64(Ei), _cplx(ExpN2)))

    def set_source(self, Hinc, Einc, N_2pi, time_period):
        _check(lib().smx_plan_set_source(self._h, _cplx(Hinc), _cplx(Einc), int(N_2pi),
                                         _i32(time_period)))

    def run(self, P0, P1, sums=None):
        """Envelope periods [P0, P1) (F:153-167); returns the envelope sums added."""
        acc = np.zeros(2) if sums is None else sums
        _check(lib().smx_plan_run(self._h, int(P0), int(P1), acc.ctypes.data_as(C.c_void_p)))
        return acc

    def run_steps(self, step0, nsteps, sums=None):
        """Global steps [step0, step0 + nsteps) of the loop (period k // N, phase k % N)."""
        acc = np.zeros(2) if sums is None else sums
        _check(lib().smx_plan_run_steps(self._h, int(step0), int(nsteps),
                                        acc.ctypes.data_as(C.c_void_p)))
        return acc

    def set_dft(self, planes, w_per_step, chunk=32):
        """Port DFT on z-planes `planes` at phase advances w_per_step = omega_f*Dt."""
        self._dft = (len(planes), len(w_per_step))
        _check(lib().smx_plan_set_dft(self._h, len(planes), _i32(planes), len(w_per_step),
                                      _f64(w_per_step), int(chunk)))

    def set_point_source(self, comp, cell, amp):
        """smx_plan_set_point_source: after run step k, E_comp of cell (j0, j1, j2) gets
        B * amp[k] (amp[k] added to the curl of step k's E update); amp None or empty
        removes the source."""
        if amp is None or len(amp) == 0:
            _check(lib().smx_plan_set_point_source(self._h, 0, _i32([0, 0, 0]), 0, None))
            return
        a = np.ascontiguousarray(amp, dtype=np.float64)
        _check(lib().smx_plan_set_point_source(self._h, int(comp), _i32(cell), a.size, a.ctypes.data))

    def project(self, modes, monitors=True, dft=False):
        """smx_plan_project: the S_calc overlap (F:437-445) of each mode column (complex
        [nmodes, 4*L0*L1], tilXi's row order) with the monitors and/or every DFT plane
        and frequency, on the device.  Returns (S_mon [2, nmodes] or None,
        S_dft [nplanes, nfreq, nmodes] or None)."""
        modes = np.ascontiguousarray(np.asarray(modes, dtype=np.complex128))
        nm = modes.shape[0]
        smon = np.zeros((2, nm), dtype=np.complex128) if monitors else None
        sdft = None
        if dft:
            m, f = getattr(self, "_dft", (1, 1))  # no DFT set: the library refuses (status 5)
            sdft = np.zeros((m, f, nm), dtype=np.complex128)
        ptr = lambda a: None if a is None else a.ctypes.data_as(C.c_void_p)  # noqa: E731
        _check(lib().smx_plan_project(self._h, nm, modes.view(np.float64).reshape(-1), ptr(smon), ptr(sdft)))
        return smon, sdft

    def read_dft(self):
        """complex128[nplanes, 4 (Hx, Hy, Ex, Ey), nfreq, L0*L1]."""
        m, f = self._dft
        out = np.zeros(2 * m * 4 * f * self.grid.plane)
        _check(lib().smx_plan_read_dft(self._h, out))
        return out.view(np.complex128).reshape(m, 4, f, self.grid.plane)

    def adjoint_forward(self, nsteps, ckpt_every, zT, w, box):
        """Forward pass of the port-objective adjoint; returns J (smx_adjoint_forward)."""
        J = C.c_double()
        self._box = tuple(int(v) for v in box)
        _check(lib().smx_adjoint_forward(self._h, int(nsteps), int(ckpt_every), int(zT), float(w),
                                         _i32(box), C.byref(J)))
        return J.value

    def adjoint_backward(self):
        """dJ/d eps_c on the design box: float64[3, nz, ny, nx] (smx_adjoint_backward)."""
        x0, x1, y0, y1, z0, z1 = self._box
        out = np.zeros(3 * (x1 - x0) * (y1 - y0) * (z1 - z0))
        _check(lib().smx_adjoint_backward(self._h, out))
        return out.reshape(3, z1 - z0, y1 - y0, x1 - x0)

    def autotune(self, trial_steps=20):
        """Pick the fastest FP32 step kernel for this grid; returns (fx, fy, fz, variant):
        variant 0 = the two in-place sweeps (sizes 0), 1 = fused step staged through
        registers, 2 = fused step staged by asynchronous copies to LDS, 3 = two steps per
        pass.  `z_balance` then holds how the two-step kernel cuts z (0 equal chunks,
        k > 0 chunks balanced with a z-PML stage counted 1 + k/2 times)."""
        out = np.zeros(5, dtype=np.int32)
        _check(lib().smx_plan_autotune(self._h, int(trial_steps), out))
        self.z_balance = int(out[4])
        return tuple(int(v) for v in out[:4])

    @property
    def kernels(self):
        """Names of the kernels one step launches with the current kernel choice."""
        return lib().smx_plan_kernels(self._h).decode()

    @property
    def stream(self):
        """The plan's hipStream_t as an integer handle."""
        return lib().smx_plan_stream(self._h)

    def sync(self):
        _check(lib().smx_plan_sync(self._h))

    def flush(self):
        """Queue the folds of the DFT and monitor samples still in the ring (no wait)."""
        _check(lib().smx_plan_flush(self._h))
