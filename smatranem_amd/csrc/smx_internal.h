/*
 * smx_internal.h -- types shared by the HIP kernels and the plan code of libsmx_fdtd.
 * Not part of the C ABI.
 */
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "smx_fdtd.h"

/* Per-step scalars of the modal sheet source (F:155-158) and the monitor phases
 * (F:165, F:322-325), computed on the host for every step. */
struct smx_step_scalars {
    double srcE_f2, srcE_re, srcE_im; /* Ei = f2*(re*Re Einc + im*Im Einc) */
    double srcH_f2, srcH_re, srcH_im; /* Hi likewise from Hinc */
    double monH_re, monH_im;          /* ExpN2[0]: H monitors */
    double monE_re, monE_im;          /* ExpN2[1]: E monitors */
};

/* Geometry and boundary data common to every kernel. */
struct smx_geom {
    int L0, L1, L2;
    int iBc0, iBc1, iBc2; /* Bloch/PEC factors at the wrap, in {-1, 0, 1} */
    int iB, iT, iI;       /* monitor planes, source plane */
    double sgn;           /* sign of the incident mode (F:253) */
    int64_t n;            /* L0*L1*L2 */
    int Lnotn;
};

/* Plain FP64 step: the reference storage contract on the device. */
struct smx_f64_args {
    smx_geom g;
    const double *Ah, *Bh, *Ae, *Be; /* 3*(n+Lnotn) */
    double *H, *E;                   /* 3*(n+Lnotn+1) */
    const char *Cont;
    const int *addr;
    const double *srcH_plane, *srcE_plane; /* explicit Hi/Ei planes, or null */
    const double *Hinc, *Einc;             /* complex[2*L0*L1] for on-device synthesis */
    double *HB, *EB, *HT, *ET;             /* complex[2*L0*L1] */
    smx_step_scalars s;
};

extern "C" {
hipError_t smx_launch_f64_step(const smx_f64_args *a, hipStream_t stream);
}

/* Host-side checks and error text shared by the entry points (smx_plan.hip):
 * smx_check_grid validates a grid and its class map as smx_plan_create does (SMX_OK,
 * SMX_ERR_ARG or SMX_ERR_MONITOR_PML); smx_note_hip / smx_note_arg record the text
 * smx_last_error returns and give the status code. */
int smx_check_grid(const smx_grid *g, const char *Cont);
int smx_note_hip(hipError_t e, const char *what);
int smx_note_arg(const char *what);
