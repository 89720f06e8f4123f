/*
 * smx_f32.h -- the FP32 plan: packed device layout and its kernels (kern_f32.hip).
 *
 * Layout (all planar, j0 fastest, n = L0*L1*L2 floats each):
 *   T[6]  total fields Hx Hy Hz Ex Ey Ez.  Total = reference main + split slot, the
 *         field every neighbour's curl sees (main_calc adds the split curl through
 *         address_notn, FDTDcalc/FDTDsplitfieldMetal.c:266, :298).
 *   S[6]  split ('a' cells: the '-' part) or stable-form aux (metal cells); zero in
 *         every other cell.  Reference main slot = T - S.
 *   b[6]  per-cell coefficient words bH_c = 1/(iR*mu_c), bE_c = 1/(iR*eps_c).  A
 *         lossy or metal cell is flagged by bH_x = bE_x = NaN, with bH_y/bE_y holding
 *         (as raw bits) its index in the special table.
 *   pml   per-axis tables for 'a' cells: aH, gH (integer samples), aE, gE (half
 *         samples) with A = (iR - iI*s)/(iR + iI*s), B = b*iR/(iR + iI*s), plus the
 *         per-axis 'a' flag.
 *   special  full per-component coefficients and class of lossy/metal cells.
 * Fields are double-buffered: the fused step reads buffer cur and writes 1 - cur, so a
 * block's halo reads never see a neighbour block's new values.
 */
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "smx_internal.h"

struct smx_special_f32 {
    int cls; /* reference class char: 'n', 'a', 'e', 'h', 'b' */
    int pad[3];
    float hA[3], hB[3], hAs[3], hBs[3]; /* H: main A, B; split/aux A, B */
    float eA[3], eB[3], eAs[3], eBs[3]; /* E likewise */
};

/* Up to 6 boxes [x0, x1) x [y0, y1) x [z0, z1) covering the grid outside one inner
 * box, with the running cell count before each (the region sweeps' thread index). */
struct smx_region_set {
    int n;
    int box[8][6];
    int64_t start[9];
};

struct smx_f32_state {
    int L[3];
    int64_t n;
    float *T[2][6];         /* ping-pong buffers; [cur] holds the current state */
    float *S[2][6];
    float *b[6];
    float *Tmem[2], *Smem[2], *bmem; /* one allocation per group: X[i] = Xmem + i*n */
    float *pml;             /* 4*(L0+L1+L2) floats: per axis aH[L], gH[L], aE[L], gE[L] */
    unsigned char *pmlflag; /* L0+L1+L2 */
    smx_special_f32 *special;
    int64_t *special_cells; /* linear index of each special cell, table order */
    int64_t *special_host;  /* the same on the host */
    /* two-step pass on a grid with lossy/metal cells (smx_f32_step2): the box R of the
     * tiles whose stage region holds such a cell (left to region sweeps), R grown by one
     * and two cells, the monitor ranges of R, and the other tiles' ids (device);
     * recomputed when the tiling (sp_key) changes */
    int sp_valid, sp_key[6], sp_nclean, sp_monin[6];
    int *sp_blist;
    smx_region_set sp_R, sp_R1, sp_R2;
    int nspecial;
    int has_pml; /* any PML cell on the grid */
    /* split of a grid with PML for the staged kernel without split slots: fbox, the
     * cells whose +x/+y/+z neighbours hold no PML cell either ([x0, x1) x [y0, y1) x
     * [z0, z1)), and up to 6 boxes covering the rest (region sweeps); split_ok = 0
     * when the PML-free cells do not form one box */
    int split_ok;
    int fbox[6];
    int nreg;
    int reg[6][6];
    int64_t reg_start[7];
    /* the same for the two-step kernel on a grid with PML: its box (cells whose
     * two-step cone stays PML-free), the rest (P) and the two first-step regions
     * around it (P and both one-cell shells of the box; P and its +x/+y/+z shell);
     * tmpT/tmpS hold the first step there (allocated on first use) */
    int split2_ok;
    int fbox2[6];
    smx_region_set regP2, regH1, regE1;
    float *tmpT, *tmpS;
    int cur;
    int fused; /* 1: one fused H+E launch per step (+ the special-cell E pass);
                  0: two in-place sweeps */
    int bx;    /* x width of a block of the in-place sweeps (32 or 64) */
    int tile;  /* fused tile variant (kern_f32.hip kTiles): 0 = 32x8, 1 = 64x4, 2 = 16x16
                  (x by y) staged through registers; 3-7 staged by asynchronous copies
                  to LDS (5-7 only on grids without PML cells) */
    int fz;    /* fused z chunk */
    int two;   /* two-step kernel in use for pairs of steps: -1 off, else its variant
                  (kern_f32.hip kTwo); single steps still take the choice above */
    int two_fz;
    int two_bal; /* z chunks of the two-step kernel: 0 uniform (two_fz planes), k > 0
                    balanced with a stage of a z-PML plane counted as 1 + k/2 */
    unsigned char *zpml_host; /* z-PML flag of every plane (host copy) */
    unsigned char *xypml_host; /* x-PML flag of every column, then y-PML flag of every row */
    int force_big; /* the two-step pass in its 64-bit form on any grid (tests) */
    int pml_origin; /* split-slot two-step pass: y tiles placed to put the y PML in the
                       fewest tile rows (1, the default), or from row 0 (0) */
    /* host copies for the reference-layout field import/export */
    char *cont_host;
    int *addr_host;
    int Lnotn;
};

struct smx_f32_launch_args {
    smx_geom g;
    smx_step_scalars s;
    smx_step_scalars s2; /* the second step's scalars (smx_f32_step2) */
    const double *srcH_plane, *srcE_plane;
    const double *Hinc, *Einc;
    double *HB, *EB, *HT, *ET;
    /* smx_f32_step2 only: multi-frequency DFT ring (float [slot][plane][4][L0*L1]); the
     * two steps' samples go to slots dft_slot and dft_slot + 1; dft_n = 0: none */
    float *dft;
    int dft_n, dft_slot;
    int dft_planes[8];
    /* smx_f32_step2 only: E of the design box box[6] (non-PML cells) after each of the two
     * steps, to box_out and box_out + 3*nbox ([3][nbox], x fastest); null: none.  Runs
     * the split-slot variant (smx_f32_box_fz(st) != 0). */
    float *box_out;
    int box[6];
};

int smx_f32_create(smx_f32_state *st, const smx_geom *g, double iR, double iI, const double *mu,
                   const double *eps, const double *sigma, const char *Cont, const int *addr,
                   hipStream_t stream);
void smx_f32_free(smx_f32_state *st);
hipError_t smx_f32_step(smx_f32_state *st, const smx_f32_launch_args *a, hipStream_t stream);
/* Two steps in one pass (scalars a->s, then a->s2), when smx_f32_two_ok(st, variant, fz). */
hipError_t smx_f32_step2(smx_f32_state *st, const smx_f32_launch_args *a, hipStream_t stream);
/* Allocate every buffer the current kernel choice allocates on first use (the scratch
 * level of the two-step pass with region sweeps), so that the steps that follow make
 * no allocation and can be captured into a graph. */
hipError_t smx_f32_reserve(smx_f32_state *st, hipStream_t stream);
int smx_f32_num_two(void);
/* Whether two-step variant `variant` updates every cell of this grid in its one launch
 * (no region sweeps), so that it can also take the DFT samples of both steps. */
int smx_f32_two_full(const smx_f32_state *st, int variant);
int smx_f32_two_ok(const smx_f32_state *st, int variant, int fz);
/* z chunk of the design-box two-step pass, or 0 where the grid cannot take it */
int smx_f32_box_fz(const smx_f32_state *st);
/* tile of two-step variant `variant`: fx x fy cells, one thread each */
void smx_f32_two_shape(int variant, int *fx, int *fy);
int smx_f32_write_fields(smx_f32_state *st, const double *H_t, const double *E_t, hipStream_t stream);
int smx_f32_read_fields(const smx_f32_state *st, double *H_t, double *E_t, hipStream_t stream);
const char *smx_f32_last_error(void);
/* Number of fused tile variants, and whether (tile, z chunk) can run on this grid. */
int smx_f32_num_tiles(void);
void smx_f32_tile_shape(int tile, int *fx, int *fy);
/* Kernels of one step (one pass of the two-step kernel) in the current choice,
 * " + "-separated (the lossy/metal passes of the fused variants not included); dft != 0:
 * the two-step kernel that also samples the DFT planes. */
const char *smx_f32_kernel_name(const smx_f32_state *st, int dft);
int smx_f32_tile_ok(const smx_f32_state *st, int tile, int fz);

/* Several plans per launch (two_step_multi, the split-slot two-step variant 2): whether
 * a plan can take part with z chunks of fz planes, the blocks of the kernel resident
 * on the device at once, a plan's two argument blocks (smx_f32_multi_args_size() bytes
 * each: the pass from buffer cur, then the next pass) and block count; the launch takes
 * the device copy of the plans' argument blocks, the block list (plan << 20 | block id)
 * and SMX_MULTI_SCALARS doubles per plan (the 20 step scalars, the second step's after
 * the first's, in the order of fused_args; then the DFT ring slot of the pass's first
 * step); dft selects the instantiation that takes DFT samples. */
#define SMX_MULTI_SCALARS 21 /* per plan and pass: 20 step scalars, then the DFT ring slot */
size_t smx_f32_multi_args_size(void);
int smx_f32_multi_ok(const smx_f32_state *st, int fz);
int smx_f32_multi_slots(void);
int smx_f32_multi_prepare(const smx_f32_state *st, const smx_f32_launch_args *la, int fz, void *args2, int *nblocks);
hipError_t smx_f32_multi_launch(const void *args, const int *blk, const double *scal, int nblocks, int dft,
                                 hipStream_t stream);
const char *smx_f32_multi_kernel_name(void);
