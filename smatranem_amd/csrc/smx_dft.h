/*
 * smx_dft.h -- multi-frequency port-plane DFT (kern_dft.hip): kernel arguments.
 */
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define SMX_DFT_MAX_PLANES 8
#define SMX_DFT_MAX_FREQ 32
#define SMX_DFT_MAX_CHUNK 256

struct smx_dft_sample_args {
    int L0, L1, L2;
    int nplanes;
    int planes[SMX_DFT_MAX_PLANES];
    int64_t plane_cells;            /* L0*L1 */
    const float *T32[6], *S32[6];   /* FP32 plan: total and split (current buffer) */
    const unsigned char *pmlflag;   /* FP32 plan: per-axis PML flags */
    const float *bH0, *bH1;         /* FP32 plan: coefficient words bH_x, bH_y (NaN in bH_x marks a
                                       lossy/metal cell, whose bH_y bits index the special table) */
    const int *special;             /* FP32 plan: special table, the class char first in each entry */
    int special_stride;             /* ints per special-table entry */
    const double *H64, *E64;        /* FP64 plan: reference layout */
    const char *cont;               /* FP64 plan: cell classes ('a' cells add the split slot) */
    const int *addr;
    int64_t n;
    void *series;                   /* ring: [chunk][plane][4][L0*L1], float or double */
    /* FP32 plan, mask_on: only the cells in the ranges [mask[2d], mask[2d] + mask[2d+1]) of
     * each axis d, taken modulo its length */
    int mask_on, mask[6];
    /* a point source's term of this step, applied before the samples when pt_amp is set:
     * dst += coef * pt_amp[pt_k] (FP64: pt_dst64, pt_coef64; FP32: pt_dst32, pt_coef32, the
     * amplitude rounded to FP32), a separate multiply and add, in the thread that samples
     * cell pt_q of ring plane pt_m, or in thread 0 when pt_m < 0 (no ring plane holds the
     * source's cell) */
    const double *pt_amp;
    long long pt_k;
    int pt_m;
    int64_t pt_q;
    double *pt_dst64;
    const double *pt_coef64;
    float *pt_dst32;
    const float *pt_coef32;
};

struct smx_dft_fold_args {
    int nplanes, nfreq, count;      /* folded planes: the first nplanes of each ring slot */
    int ring;                       /* planes per ring slot (>= nplanes) */
    int64_t plane_cells;
    int table_nt;                   /* chunk length the table was built for (>= count) */
    const double2 *table;           /* [2 (H, E)][table_nt][row]: e^{i*w_f*(t + 1 + h/2)}, zero for f >= nfreq */
    double rot[SMX_DFT_MAX_FREQ][2]; /* e^{i*w_f*k0}, k0 = run step index of ring slot 0 */
    const void *series;
    double *acc;                    /* [plane][4][freq][L0*L1] complex */
};

/* The reference monitors (F:318-328) taken from the ring instead of per step: for the
 * ring slots whose bit is set in mask, HB/HT += e^{i*dphi*j}*(Hx, Hy) and EB/ET +=
 * e^{i*dphi*(j + 1/2)}*(Ex, Ey) of the monitor planes' samples, j = (run step + 1) mod N,
 * with the phases of the reference's table (expn, F:93-99). */
struct smx_dft_mon_args {
    int ring;                       /* planes per ring slot */
    int mon[2];                     /* ring plane of the bottom / top monitor; -1: none */
    int count;                      /* slots to fold */
    int N, j0;                      /* period of the table; j of slot 0 */
    int64_t plane_cells;
    uint64_t mask[SMX_DFT_MAX_CHUNK / 64];
    const double *expn;             /* [N][4]: e^{i*dphi*j}, e^{i*dphi*(j + 1/2)} */
    const void *series;
    double *monitors;               /* HB, EB, HT, ET: complex[2*L0*L1] each, in this order */
};

/* Entries per row of the phase table: nfreq rounded up to a multiple of 16. */
__host__ __device__ inline int smx_dft_table_row(int nfreq) { return (nfreq + 15) / 16 * 16; }

/* The phase table of a plan's DFT (host): tab[h][t][f] = e^{i*w_f*(t + 1 + h/2)} for
 * t < chunk, f < nfreq, and 0 for nfreq <= f < smx_dft_table_row(nfreq);
 * 4 * chunk * smx_dft_table_row(nfreq) doubles. */
void smx_dft_table(int nfreq, const double *w, int chunk, double *tab);

extern "C" {
hipError_t smx_dft_sample(const smx_dft_sample_args *a, int slot, hipStream_t stream);
hipError_t smx_dft_fold(const smx_dft_fold_args *a, int f64, hipStream_t stream);
hipError_t smx_dft_fold_monitors(const smx_dft_mon_args *a, int f64, hipStream_t stream);
}
