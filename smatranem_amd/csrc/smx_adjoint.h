/*
 * smx_adjoint.h -- arguments of the adjoint kernels (kern_adjoint.hip).
 */
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

struct smx_adj_args {
    int L0, L1, L2;
    int64_t n, sz;
    float w[3];                /* Bloch/PEC factors at the wraps */
    int zT;                    /* objective plane */
    int box[6];                /* design box [x0,x1) x [y0,y1) x [z0,z1) */
    int bx, by;
    int64_t nbox;
    const float *b[6];         /* coefficient words of the FP32 plan */
    const float *pml;
    const unsigned char *pmlflag;
    float *lT[6], *lS[6];      /* adjoint of total and split fields */
    double *G;                 /* [3][nbox] sum_k lambda_TE (E^{k+1} - E^k) */
    double *D;                 /* [2][L0*L1] complex: the objective's DFT */
};

extern "C" {
hipError_t smx_adj_launch_a(const smx_adj_args *a, int finish_h, const float *cur, const float *next, hipStream_t s);
hipError_t smx_adj_launch_b(const smx_adj_args *a, int inject, double cs, double sn, hipStream_t s);
hipError_t smx_adj_launch_inject(const smx_adj_args *a, double cs, double sn, hipStream_t s);
hipError_t smx_adj_launch_sample(const smx_adj_args *a, const float *Ex, const float *Ey, double cs, double sn,
                                 hipStream_t s);
hipError_t smx_adj_launch_save_box(const smx_adj_args *a, const float *Ex, const float *Ey, const float *Ez,
                                   float *dst, hipStream_t s);
hipError_t smx_adj_launch_finish(const smx_adj_args *a, double iR, double *grad, hipStream_t s);
}
