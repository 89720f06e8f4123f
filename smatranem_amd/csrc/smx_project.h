/*
 * smx_project.h -- device mode projection (kern_project.hip): kernel arguments.
 */
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

struct smx_project_args {
    int64_t P;             /* L0*L1 */
    int nmodes;            /* tilXi columns */
    int nmon;              /* monitor vectors first (0, 1 or 2) */
    int nfreq;             /* DFT frequencies (DFT vectors follow the monitors) */
    int nvec;              /* nmon + DFT planes * nfreq */
    int nchunks;           /* smx_project_chunks(P) */
    const double2 *modes;  /* [nmodes][4P] */
    const double2 *mon;    /* HB, EB, HT, ET: complex[2P] each, or null */
    const double2 *dft;    /* [plane][4][freq][P], or null */
    double2 *partial;      /* [nvec][nmodes][nchunks] */
    double2 *S;            /* [nvec][nmodes] */
};

extern "C" {
int smx_project_chunks(int64_t P);
hipError_t smx_project_launch(const smx_project_args *a, hipStream_t stream);
}
