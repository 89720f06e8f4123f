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

/* One fused, LDS-staged transposed step (adj_fused_f32): lambda^{k+1} -> lambda^k over
 * the whole grid.  lambda buffers are groups of 6 component arrays, ncomp floats apart
 * (24*ncomp < 2^32); T = total, S = split. */
struct smx_adj_fused_args {
    int L0, L1, L2;
    int64_t sz, ncomp;
    float w[3];
    int tile;                  /* tile shape: kAdjTiles[tile] */
    int gx, gy, fz;            /* tiles of FX x FY cells, z chunks of fz planes */
    const int *blocks;         /* device: block ids bx + gx*(by + gy*bz), bit 30 set when the
                                  block's reach holds a PML cell */
    int n_pml, n_free;         /* blocks with and without PML in reach */
    const float *lT, *lS;      /* lambda^{k+1} (T group, S group) */
    float *nT, *nS;            /* lambda^k */
    const float *b;            /* coefficient group: bH x y z, bE x y z */
    const float *pml;
    const unsigned char *pmlflag;
    int has_pml;
};

/* Two transposed steps per launch (adj_two_f32): lambda^{k+1} -> lambda^{k-1} over the
 * whole grid, tiles of 32 x 8 cells, z chunks of fz <= 48 planes (block ids in blocks,
 * bx + gx*(by + gy*bz)).  Inside: the objective's injection into lambda_TE^k on plane
 * zT with (cs, sn), and lambda_TE^k of the design box to lbox ([3][nbox]). */
struct smx_adj_two_args {
    int L0, L1, L2;
    int64_t sz, ncomp;
    float w[3];
    int gx, gy, fz, nblocks;
    const int *blocks;
    const float *lT, *lS;      /* lambda^{k+1} (T group, S group) */
    float *nT, *nS;            /* lambda^{k-1} */
    const float *b;
    const float *pml;
    const unsigned char *pmlflag;
    int has_pml;
    const double *D;           /* the objective's DFT [2][L0*L1] complex */
    int zT;
    double cs, sn;
    int box[6], bx, by;
    int64_t nbox;
    float *lbox;
};

constexpr int kAdjTwoMaxFZ = 48;

/* tile shapes of the fused transposed step (FX x FY cells, one thread each) */
constexpr int kAdjTiles[3][2] = {{32, 8}, {64, 8}, {64, 4}};
constexpr int kAdjFZ = 32;

extern "C" {
/* Launch one transposed step over the block list. */
hipError_t smx_adj_launch_fused(const smx_adj_fused_args *a, hipStream_t s);
/* G_c += lambda_TE^{k+1}(box) * (E^{k+1} - E^k)(box), the box arrays [3][nbox]. */
hipError_t smx_adj_launch_grad(const smx_adj_args *a, const float *lTE, const float *cur, const float *next,
                               hipStream_t s);
/* Two transposed steps (adj_two_f32). */
hipError_t smx_adj_launch_two(const smx_adj_two_args *a, hipStream_t s);
/* The gradient terms of steps k and k-1: lambda_TE^{k+1} (full group lTE) with
 * E^{k+1} - E^k, then lbox with E^k - E^{k-1} (box arrays [3][nbox]). */
hipError_t smx_adj_launch_grad2(const smx_adj_args *a, const float *lTE, const float *lbox, const float *kp1,
                                const float *k0, const float *km1, hipStream_t s);
hipError_t smx_adj_launch_a(const smx_adj_args *a, int finish_h, const float *cur, const float *next, hipStream_t s);
hipError_t smx_adj_launch_b(const smx_adj_args *a, int inject, double cs, double sn, hipStream_t s);
hipError_t smx_adj_launch_inject(const smx_adj_args *a, double cs, double sn, hipStream_t s);
hipError_t smx_adj_launch_sample(const smx_adj_args *a, const float *Ex, const float *Ey, double cs, double sn,
                                 hipStream_t s);
hipError_t smx_adj_launch_save_box(const smx_adj_args *a, const float *Ex, const float *Ey, const float *Ez,
                                   float *dst, hipStream_t s);
hipError_t smx_adj_launch_finish(const smx_adj_args *a, double iR, double *grad, hipStream_t s);
}
