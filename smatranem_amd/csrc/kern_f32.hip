/*
 * kern_f32.hip -- FP32 split-field step on the packed layout of smx_f32.h.
 *
 * Physics of FDTDcalc/FDTDsplitfieldMetal.c:242-329 (main_calc) restated on total
 * fields T = main + split, which every neighbour's curl uses (the reference adds the
 * split curl through address_notn, :266, :298, :274, :305).  Per cell kind:
 *   regular 'n' (lossless):  X += -/+ b*(curl T)                (A = 1)
 *   regular 'a' (PML):       main = T - S; main' = aM*main -/+ b*gM*T+;
 *                            S' = aS*S -/+ b*gS*T-; T' = main' + S'
 *   special (lossy/metal):   table coefficients, the reference's branch per class,
 *                            including its backward curl of H main in metal E cells
 *                            (:310) and frozen fields (:259-314).
 * Monitors accumulate in FP64 on the monitor planes (:318-328).
 *
 * Launch: 256-thread blocks of BX x (256/BX) cells over (j0, j1), one z plane per
 * block, blocks renumbered so that each XCD works through a contiguous run of z
 * planes (its L2 then holds the z-neighbour plane of the next block it runs).
 */
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <vector>

#include "smx_dft.h"
#include "smx_f32.h"
#include "smx_fdtd.h"
#include "smx_setup.h"
#include "smx_stage.h"

namespace {

thread_local char f32_err[512] = "";

#define F32_HIP(call)                                                                     \
    do {                                                                                  \
        hipError_t e_ = (call);                                                           \
        if (e_ != hipSuccess) {                                                           \
            snprintf(f32_err, sizeof f32_err, "%s: %s", #call, hipGetErrorString(e_));    \
            return e_ == hipErrorOutOfMemory ? SMX_ERR_OOM : SMX_ERR_HIP;                 \
        }                                                                                 \
    } while (0)

constexpr int kThreads = 256;

struct f32_args {
    int L0, L1, L2;
    int gx, gy, nblocks;
    int64_t sy, sz;
    float w[3];
    int iB, iT, iI;
    float sgn;
    float *T[6];
    float *S[6];
    const float *b[6];
    const float *pml;
    const unsigned char *pmlflag;
    const smx_special_f32 *special;
    /* source: explicit plane (double[2*L0*L1]) or synthesis from the profile */
    const double *src_plane;
    const double *inc;
    double src_f2, src_re, src_im;
    /* monitors */
    double *HB, *EB, *HT, *ET;
    double monH_re, monH_im, monE_re, monE_im;
};


struct cellpos {
    int j0, j1, j2;
    int64_t l, q;
};

template <int BX>
__device__ inline bool decode(const f32_args &a, cellpos &p)
{
    constexpr int BY = kThreads / BX;
    const int id = xcd_logical(blockIdx.x, a.nblocks);
    const int bx = id % a.gx, rest = id / a.gx;
    const int by = rest % a.gy;
    p.j2 = rest / a.gy;
    p.j0 = bx * BX + (int)(threadIdx.x % BX);
    p.j1 = by * BY + (int)(threadIdx.x / BX);
    if (p.j0 >= a.L0 || p.j1 >= a.L1) return false;
    p.q = p.j0 + (int64_t)a.L0 * p.j1;
    p.l = p.q + a.sz * p.j2;
    return true;
}

/* PML tables of axis d: aH, gH, aE, gE, each L_d long. */
__device__ inline const float *pml_axis(const f32_args &a, int d)
{
    return d == 0 ? a.pml : d == 1 ? a.pml + 4 * a.L0 : a.pml + 4 * (a.L0 + a.L1);
}

__device__ inline double source_value(const f32_args &a, int64_t q, int v)
{
    if (a.src_plane) return a.src_plane[v + 2 * q];
    const double *z = a.inc + 2 * (v + 2 * q);
    return a.src_f2 * (a.src_re * z[0] + a.src_im * z[1]);
}

/* Offsets and weights of the backward (j-1) or forward (j+1) neighbours. */
struct nb3 {
    int64_t o[3];
    float w[3];
};

__device__ inline nb3 nb_backward(const f32_args &a, const cellpos &p)
{
    nb3 n;
    n.o[0] = p.j0 == 0 ? (int64_t)(a.L0 - 1) : -1;
    n.o[1] = p.j1 == 0 ? (int64_t)(a.L1 - 1) * a.sy : -a.sy;
    n.o[2] = p.j2 == 0 ? (int64_t)(a.L2 - 1) * a.sz : -a.sz;
    n.w[0] = p.j0 == 0 ? a.w[0] : 1.f;
    n.w[1] = p.j1 == 0 ? a.w[1] : 1.f;
    n.w[2] = p.j2 == 0 ? a.w[2] : 1.f;
    return n;
}

__device__ inline nb3 nb_forward(const f32_args &a, const cellpos &p)
{
    nb3 n;
    n.o[0] = p.j0 + 1 == a.L0 ? -(int64_t)(a.L0 - 1) : 1;
    n.o[1] = p.j1 + 1 == a.L1 ? -(int64_t)(a.L1 - 1) * a.sy : a.sy;
    n.o[2] = p.j2 + 1 == a.L2 ? -(int64_t)(a.L2 - 1) * a.sz : a.sz;
    n.w[0] = p.j0 + 1 == a.L0 ? a.w[0] : 1.f;
    n.w[1] = p.j1 + 1 == a.L1 ? a.w[1] : 1.f;
    n.w[2] = p.j2 + 1 == a.L2 ? a.w[2] : 1.f;
    return n;
}

/*
 * Curl halves of field F (F[0..2] = x, y, z arrays, value = F - G when G is given,
 * i.e. a main slot rebuilt from total and split) at cell l.
 * Backward: Pc = F[c2](l) - w[c1] F[c2](l-e_c1); Mc = -F[c1](l) + w[c2] F[c1](l-e_c2).
 * Forward:  Pc = w[c1] F[c2](l+e_c1) - F[c2](l); Mc = -w[c2] F[c1](l+e_c2) + F[c1](l).
 */
template <bool FWD>
__device__ inline void curl_halves(float *const *F, float *const *G, const nb3 &n, int64_t l, float P[3],
                                   float M[3])
{
    auto v = [&](int c, int64_t i) { return G ? F[c][i] - G[c][i] : F[c][i]; };
#pragma unroll
    for (int c = 0; c < 3; c++) {
        const int c1 = c == 2 ? 0 : c + 1, c2 = c == 0 ? 2 : c - 1;
        if (FWD) {
            P[c] = n.w[c1] * v(c2, l + n.o[c1]) - v(c2, l);
            M[c] = -n.w[c2] * v(c1, l + n.o[c2]) + v(c1, l);
        } else {
            P[c] = v(c2, l) - n.w[c1] * v(c2, l + n.o[c1]);
            M[c] = -v(c1, l) + n.w[c2] * v(c1, l + n.o[c2]);
        }
    }
}

/* ---------------------------------------------------------------------------------- */
/* Special (lossy / metal) cells: the reference branch of each class, FP32             */
/* ---------------------------------------------------------------------------------- */

/* H update of a special cell; sign = -1 (H takes -curl). */
__device__ void special_h(const f32_args &a, const cellpos &p, const smx_special_f32 &sp)
{
    float *const *H = a.T;
    float *const *Hs = a.S;
    float *const *E = a.T + 3;
    float *const *Es = a.S + 3;
    const int k = sp.cls;
    if (k == 'e') return;
    const nb3 n = nb_backward(a, p);
    float P[3], M[3];
    if (k == 'n' || k == 'a') {
        curl_halves<false>(E, nullptr, n, p.l, P, M);
        if (k == 'n') {
            float W[3] = {P[0] + M[0], P[1] + M[1], P[2] + M[2]};
            if (p.j2 == a.iI) {
                W[0] += a.sgn * (float)source_value(a, p.q, 1);
                W[1] -= a.sgn * (float)source_value(a, p.q, 0);
            }
            for (int c = 0; c < 3; c++) H[c][p.l] = sp.hA[c] * H[c][p.l] - sp.hB[c] * W[c];
        } else {
            for (int c = 0; c < 3; c++) {
                const float s = Hs[c][p.l], m = H[c][p.l] - s;
                const float m2 = sp.hA[c] * m - sp.hB[c] * P[c];
                const float s2 = sp.hAs[c] * s - sp.hBs[c] * M[c];
                H[c][p.l] = m2 + s2;
                Hs[c][p.l] = s2;
            }
        }
        return;
    }
    /* 'h', 'b': curl of the E main slots (F:279), aux in the split slot */
    curl_halves<false>(E, Es, n, p.l, P, M);
    for (int c = 0; c < 3; c++) {
        const float s = Hs[c][p.l];
        float m = H[c][p.l] - s;
        m += 0.5f * s;
        const float s2 = sp.hAs[c] * s - sp.hBs[c] * (P[c] + M[c]);
        m += 0.5f * s2;
        H[c][p.l] = m + s2;
        Hs[c][p.l] = s2;
    }
}

__device__ void special_e(const f32_args &a, const cellpos &p, const smx_special_f32 &sp)
{
    float *const *H = a.T;
    float *const *Hs = a.S;
    float *const *E = a.T + 3;
    float *const *Es = a.S + 3;
    const int k = sp.cls;
    if (k == 'h') return;
    float P[3], M[3];
    if (k == 'n' || k == 'a') {
        const nb3 n = nb_forward(a, p);
        curl_halves<true>(H, nullptr, n, p.l, P, M);
        if (k == 'n') {
            float W[3] = {P[0] + M[0], P[1] + M[1], P[2] + M[2]};
            if (p.j2 == a.iI) {
                W[0] += a.sgn * (float)source_value(a, p.q, 1);
                W[1] -= a.sgn * (float)source_value(a, p.q, 0);
            }
            for (int c = 0; c < 3; c++) E[c][p.l] = sp.eA[c] * E[c][p.l] + sp.eB[c] * W[c];
        } else {
            for (int c = 0; c < 3; c++) {
                const float s = Es[c][p.l], m = E[c][p.l] - s;
                const float m2 = sp.eA[c] * m + sp.eB[c] * P[c];
                const float s2 = sp.eAs[c] * s + sp.eBs[c] * M[c];
                E[c][p.l] = m2 + s2;
                Es[c][p.l] = s2;
            }
        }
        return;
    }
    /* 'e', 'b': BACKWARD curl of the H main slots (F:310) */
    const nb3 n = nb_backward(a, p);
    curl_halves<false>(H, Hs, n, p.l, P, M);
    for (int c = 0; c < 3; c++) {
        const float s = Es[c][p.l];
        float m = E[c][p.l] - s;
        m += 0.5f * s;
        const float s2 = sp.eAs[c] * s + sp.eBs[c] * (P[c] + M[c]);
        m += 0.5f * s2;
        E[c][p.l] = m + s2;
        Es[c][p.l] = s2;
    }
}

/* ---------------------------------------------------------------------------------- */
/* Sweeps                                                                             */
/* ---------------------------------------------------------------------------------- */

template <int BX>
__global__ __launch_bounds__(kThreads) void h_sweep_f32(f32_args a)
{
    cellpos p;
    if (!decode<BX>(a, p)) return;
    const float b0 = a.b[0][p.l];
    if (__builtin_isnan(b0)) {
        special_h(a, p, a.special[__float_as_uint(a.b[1][p.l])]);
        return;
    }
    const float bc[3] = {b0, a.b[1][p.l], a.b[2][p.l]};
    const nb3 n = nb_backward(a, p);
    float P[3], M[3];
    curl_halves<false>(a.T + 3, nullptr, n, p.l, P, M);
    const bool is_pml = a.pmlflag[p.j0] | a.pmlflag[a.L0 + p.j1] | a.pmlflag[a.L0 + a.L1 + p.j2];
    if (!is_pml) {
        if (p.j2 == a.iI) { /* sheet source, 'n' cells only */
            P[0] += a.sgn * (float)source_value(a, p.q, 1);
            P[1] -= a.sgn * (float)source_value(a, p.q, 0);
        }
#pragma unroll
        for (int c = 0; c < 3; c++) a.T[c][p.l] -= bc[c] * (P[c] + M[c]);
        return;
    }
    const int jj[3] = {p.j0, p.j1, p.j2};
#pragma unroll
    for (int c = 0; c < 3; c++) {
        const int dm = c == 2 ? 0 : c + 1, ds = c == 0 ? 2 : c - 1; /* main axis c+1, split c+2 */
        const float *tm = pml_axis(a, dm), *ts = pml_axis(a, ds);
        const int Lm = dm == 0 ? a.L0 : dm == 1 ? a.L1 : a.L2;
        const int Ls = ds == 0 ? a.L0 : ds == 1 ? a.L1 : a.L2;
        const float aM = tm[jj[dm]], gM = tm[Lm + jj[dm]];
        const float aS = ts[jj[ds]], gS = ts[Ls + jj[ds]];
        const float s = a.S[c][p.l], m = a.T[c][p.l] - s;
        const float m2 = aM * m - bc[c] * gM * P[c];
        const float s2 = aS * s - bc[c] * gS * M[c];
        a.T[c][p.l] = m2 + s2;
        a.S[c][p.l] = s2;
    }
}

template <int BX>
__global__ __launch_bounds__(kThreads) void e_sweep_f32(f32_args a)
{
    cellpos p;
    if (!decode<BX>(a, p)) return;
    float *const *E = a.T + 3;
    float *const *Es = a.S + 3;
    const float b0 = a.b[3][p.l];
    if (__builtin_isnan(b0)) {
        special_e(a, p, a.special[__float_as_uint(a.b[4][p.l])]);
    } else {
        const float bc[3] = {b0, a.b[4][p.l], a.b[5][p.l]};
        const nb3 n = nb_forward(a, p);
        float P[3], M[3];
        curl_halves<true>(a.T, nullptr, n, p.l, P, M);
        const bool is_pml = a.pmlflag[p.j0] | a.pmlflag[a.L0 + p.j1] | a.pmlflag[a.L0 + a.L1 + p.j2];
        if (!is_pml) {
            if (p.j2 == a.iI) {
                P[0] += a.sgn * (float)source_value(a, p.q, 1);
                P[1] -= a.sgn * (float)source_value(a, p.q, 0);
            }
#pragma unroll
            for (int c = 0; c < 3; c++) E[c][p.l] += bc[c] * (P[c] + M[c]);
        } else {
            const int jj[3] = {p.j0, p.j1, p.j2};
#pragma unroll
            for (int c = 0; c < 3; c++) {
                const int dm = c == 2 ? 0 : c + 1, ds = c == 0 ? 2 : c - 1;
                const float *tm = pml_axis(a, dm), *ts = pml_axis(a, ds);
                const int Lm = dm == 0 ? a.L0 : dm == 1 ? a.L1 : a.L2;
                const int Ls = ds == 0 ? a.L0 : ds == 1 ? a.L1 : a.L2;
                const float aM = tm[2 * Lm + jj[dm]], gM = tm[3 * Lm + jj[dm]];
                const float aS = ts[2 * Ls + jj[ds]], gS = ts[3 * Ls + jj[ds]];
                const float s = Es[c][p.l], m = E[c][p.l] - s;
                const float m2 = aM * m + bc[c] * gM * P[c];
                const float s2 = aS * s + bc[c] * gS * M[c];
                E[c][p.l] = m2 + s2;
                Es[c][p.l] = s2;
            }
        }
    }
    /* monitors: reference main slots, main = T - S (F:318-328) */
    if (p.j2 == a.iB || p.j2 == a.iT) {
        double hv[2], ev[2];
        for (int v = 0; v < 2; v++) {
            hv[v] = (double)a.T[v][p.l] - (double)a.S[v][p.l];
            ev[v] = (double)E[v][p.l] - (double)Es[v][p.l];
        }
        for (int m = 0; m < 2; m++) {
            if (p.j2 != (m == 0 ? a.iB : a.iT)) continue;
            double *HM = m == 0 ? a.HB : a.HT, *EM = m == 0 ? a.EB : a.ET;
            for (int v = 0; v < 2; v++) {
                const int64_t o = 2 * (v + 2 * p.q);
                HM[o] += a.monH_re * hv[v];
                HM[o + 1] += a.monH_im * hv[v];
                EM[o] += a.monE_re * ev[v];
                EM[o + 1] += a.monE_im * ev[v];
            }
        }
    }
}

/* ---------------------------------------------------------------------------------- */
/* Host side                                                                          */
/* ---------------------------------------------------------------------------------- */

const double kTol = 10. * 2.220446049250313e-16; /* constant.h:7 */

f32_args base_args(const smx_f32_state *st, const smx_geom &g)
{
    f32_args a;
    std::memset(&a, 0, sizeof a);
    a.L0 = g.L0;
    a.L1 = g.L1;
    a.L2 = g.L2;
    const int BY = kThreads / st->bx;
    a.gx = (g.L0 + st->bx - 1) / st->bx;
    a.gy = (g.L1 + BY - 1) / BY;
    a.nblocks = a.gx * a.gy * g.L2;
    a.sy = g.L0;
    a.sz = (int64_t)g.L0 * g.L1;
    a.w[0] = (float)g.iBc0;
    a.w[1] = (float)g.iBc1;
    a.w[2] = (float)g.iBc2;
    a.iB = g.iB;
    a.iT = g.iT;
    a.iI = g.iI;
    a.sgn = (float)g.sgn;
    for (int i = 0; i < 6; i++) {
        a.T[i] = st->T[st->cur][i];
        a.S[i] = st->S[st->cur][i];
        a.b[i] = st->b[i];
    }
    a.pml = st->pml;
    a.pmlflag = st->pmlflag;
    a.special = st->special;
    return a;
}

/* ---------------------------------------------------------------------------------- */
/* Fused step: H and E of one step in a single pass over HBM                          */
/* ---------------------------------------------------------------------------------- */
/*
 * A block owns an FX x FY tile of (j0, j1) and marches an fz-plane chunk of j2.  At
 * plane k it holds in LDS the old E of plane k+1 on the tile grown by one cell on
 * every side (for the backward curl of H) and the new H of plane k on the tile grown
 * by one cell on the +x/+y sides (for the forward curl of E); the z neighbours travel
 * in registers.  Each iteration computes H(k+1) on the grown tile, then E(k) on the
 * tile, so every field word is read and written once per step; halo H cells are
 * recomputed, not exchanged.  Reads come from buffer cur, writes go to 1 - cur.
 * Lossy/metal cells run their class branch in two small passes around this kernel:
 * special_h_pass before it (this kernel then reads their new H) and special_e_pass
 * after it (the metal E update takes the backward curl of the new H, F:310, which
 * reaches outside the tile).
 */
/* Tile shapes (FX x FY = 256 threads) the autotuner chooses among; FZ is a runtime
 * chunk length. */
struct tile_shape {
    int fx, fy;
};
/* 0-2: staged through registers; 3-4: staged by copies to LDS two planes ahead, split
 * slots included; 5-9: without split slots (on a grid with PML: the PML-free box, the
 * region sweeps doing the rest), 5-6 three planes ahead, 7 four, 8 two and 9 one (the
 * shallower rings leave room for three or four blocks per CU) */
constexpr tile_shape kTiles[10] = {{32, 8}, {64, 4}, {16, 16}, {32, 8}, {16, 16},
                                   {32, 8}, {16, 16}, {32, 8}, {32, 8}, {32, 8}};
constexpr int kNumTiles = 10, kFirstDmaTile = 3, kFirstNoSplitTile = 5;
constexpr int kDefaultTile = 0, kDefaultFZ = 16;

/* most z chunks of one two-step launch */
constexpr int kMaxZChunks = 256;
constexpr int kMaxRegions = 8; /* boxes of a region sweep */
constexpr int kDftMaxPlanes = 8; /* as SMX_DFT_MAX_PLANES */

struct fused_args {
    int L0, L1, L2;
    int gx, gy, nblocks;
    int fz;
    int64_t sz;
    float w[3];
    int iB, iT, iI;
    float sgn;
    const float *To[6];
    float *Tn[6];
    const float *So[6];
    float *Sn[6];
    const float *b[6];
    int64_t ncomp; /* component stride of To, So, b (their arrays are one allocation each) */
    const float *pml;
    const unsigned char *pmlflag;
    const smx_special_f32 *special;
    const int64_t *special_cells;
    int nspecial;
    const double *srcE_plane, *srcH_plane, *Einc, *Hinc;
    double srcE_f2, srcE_re, srcE_im, srcH_f2, srcH_re, srcH_im;
    double *HB, *EB, *HT, *ET;
    double monH_re, monH_im, monE_re, monE_im;
    /* the second step's scalars (two-step kernel) */
    double srcE2_f2, srcE2_re, srcE2_im, srcH2_f2, srcH2_re, srcH2_im;
    double monH2_re, monH2_im, monE2_re, monE2_im;
    /* box of the staged kernel's tiles: [x0, x1) x [y0, y1) x [z0, z1) (the whole grid
     * unless the cells outside it are left to the region sweeps) */
    int box[6];
    /* two-step kernel: x of its first tile column, and the (unwrapped) x range it stores;
     * its z chunks [zc[i], zc[i+1]) for i < nzc */
    int xorg, xlo, xhi;
    int nzc;
    int zc[kMaxZChunks + 1];
    /* boxes of the region sweeps and their first thread index */
    int nreg;
    int reg_specials; /* region sweeps: the special cells' update after the boxes' cells */
    int reg[kMaxRegions][6];
    int64_t reg_start[kMaxRegions + 1];
    /* region_e_f32 accumulates no monitor for cells inside this box */
    int monskip[6];
    /* region_e_f32 with monin_on: monitors only for cells in the ranges
     * [monin[2d], monin[2d] + monin[2d+1]) of each axis d, taken modulo its length */
    int monin[6], monin_on;
    /* two-step kernel: ring of the multi-frequency DFT ([slot][plane][Hx, Hy, Ex, Ey][q],
     * floats); the first step's samples go to slot dft_slot, the second's to the next;
     * dft_n = 0: no DFT */
    float *dft;
    int dft_n, dft_slot;
    int dft_planes[kDftMaxPlanes];
    /* two-step kernel (BOX instantiations): E of both steps on the design box
     * [dbox[0], dbox[1]) x [dbox[2], dbox[3]) x [dbox[4], dbox[5]), non-PML cells, to
     * dbox_out (after the first step) and dbox_out + 3*dnbox (after the second), each
     * [3][dnbox] with x fastest */
    float *dbox_out;
    int dbox[6], dbx, dby;
    int64_t dnbox;
    /* two-step kernel: the block ids it runs (device), null = all of them */
    const int *blist;
};


__device__ inline double plane_source(const double *plane, const double *inc, double f2, double re, double im,
                                      int64_t q, int v)
{
    if (plane) return plane[v + 2 * q];
    const double *z = inc + 2 * (v + 2 * q);
    return f2 * (re * z[0] + im * z[1]);
}

__device__ inline int axis_len(const fused_args &a, int d) { return d == 0 ? a.L0 : d == 1 ? a.L1 : a.L2; }

__device__ inline const float *axis_tab(const fused_args &a, int d)
{
    return d == 0 ? a.pml : d == 1 ? a.pml + 4 * a.L0 : a.pml + 4 * (a.L0 + a.L1);
}

/* Neighbour offsets/weights of cell (x, y, z): backward (fwd = false) or forward. */
__device__ inline nb3 nbrs_at(const fused_args &a, int x, int y, int z, bool fwd)
{
    const int64_t sy = a.L0;
    nb3 n;
    if (fwd) {
        n.o[0] = x + 1 == a.L0 ? -(int64_t)(a.L0 - 1) : 1;
        n.o[1] = y + 1 == a.L1 ? -(int64_t)(a.L1 - 1) * sy : sy;
        n.o[2] = z + 1 == a.L2 ? -(int64_t)(a.L2 - 1) * a.sz : a.sz;
        n.w[0] = x + 1 == a.L0 ? a.w[0] : 1.f;
        n.w[1] = y + 1 == a.L1 ? a.w[1] : 1.f;
        n.w[2] = z + 1 == a.L2 ? a.w[2] : 1.f;
    } else {
        n.o[0] = x == 0 ? (int64_t)(a.L0 - 1) : -1;
        n.o[1] = y == 0 ? (int64_t)(a.L1 - 1) * sy : -sy;
        n.o[2] = z == 0 ? (int64_t)(a.L2 - 1) * a.sz : -a.sz;
        n.w[0] = x == 0 ? a.w[0] : 1.f;
        n.w[1] = y == 0 ? a.w[1] : 1.f;
        n.w[2] = z == 0 ? a.w[2] : 1.f;
    }
    return n;
}

/* H of a lossy/metal cell from the old buffers (the class branches of F:259-283). */
__device__ void fused_special_h(const fused_args &a, int x, int y, int z, int64_t l, const smx_special_f32 &sp,
                                float Hn[3], float Sn[3])
{
    float *const *E = const_cast<float *const *>(a.To + 3);
    float *const *Es = const_cast<float *const *>(a.So + 3);
    const int k = sp.cls;
    for (int c = 0; c < 3; c++) {
        Hn[c] = a.To[c][l];
        Sn[c] = a.So[c][l];
    }
    if (k == 'e') return; /* frozen */
    const nb3 n = nbrs_at(a, x, y, z, false);
    float P[3], M[3];
    if (k == 'n' || k == 'a') {
        curl_halves<false>(E, nullptr, n, l, P, M);
        if (k == 'n') {
            float W[3] = {P[0] + M[0], P[1] + M[1], P[2] + M[2]};
            if (z == a.iI) {
                const int64_t q = x + (int64_t)a.L0 * y;
                W[0] += a.sgn * (float)plane_source(a.srcE_plane, a.Einc, a.srcE_f2, a.srcE_re, a.srcE_im, q, 1);
                W[1] -= a.sgn * (float)plane_source(a.srcE_plane, a.Einc, a.srcE_f2, a.srcE_re, a.srcE_im, q, 0);
            }
            for (int c = 0; c < 3; c++) Hn[c] = sp.hA[c] * Hn[c] - sp.hB[c] * W[c];
        } else {
            for (int c = 0; c < 3; c++) {
                const float s = Sn[c], m = Hn[c] - s;
                const float s2 = sp.hAs[c] * s - sp.hBs[c] * M[c];
                Hn[c] = sp.hA[c] * m - sp.hB[c] * P[c] + s2;
                Sn[c] = s2;
            }
        }
        return;
    }
    curl_halves<false>(E, Es, n, l, P, M); /* 'h', 'b': curl of the E main slots */
    for (int c = 0; c < 3; c++) {
        const float s = Sn[c];
        float m = Hn[c] - s + 0.5f * s;
        const float s2 = sp.hAs[c] * s - sp.hBs[c] * (P[c] + M[c]);
        m += 0.5f * s2;
        Hn[c] = m + s2;
        Sn[c] = s2;
    }
}

/* Kinds of a cell in the fused step. */
enum { K_NORMAL = 0, K_PML = 1, K_SPECIAL = 2 };

__device__ inline void store3(float *const *dst, int64_t l, const float v[3])
{
    dst[0][l] = v[0];
    dst[1][l] = v[1];
    dst[2][l] = v[2];
}

/* Components c0 .. c0+2 of cell l through a group's buffer resource (component stride
 * cb bytes; the group spans less than 4 GiB). */
__device__ inline void bstore3(__amdgpu_buffer_rsrc_t r, unsigned cb, int c0, int64_t l, const float v[3])
{
    const unsigned off = 4u * (unsigned)l + (unsigned)c0 * cb;
    __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(v[0]), r, off, 0, 0);
    __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(v[1]), r, off + cb, 0, 0);
    __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(v[2]), r, off + 2 * cb, 0, 0);
}

/* New H of one cell from preloaded old values: E at the cell (ec) and its -x, -y, -z
 * neighbours, old H (h), coefficient words (b), old split (sp, PML only). */
__device__ inline int fused_h_pre(const fused_args &a, int x, int y, int z, int64_t l, const float ec[3],
                                  const float exm[3], const float eym[3], const float ezm[3], const float h[3],
                                  const float b[3], const float sp[3], bool pml, float Hn[3], float Sn[3])
{
    if (__builtin_isnan(b[0])) { /* computed by special_h_pass before this kernel */
#pragma unroll
        for (int c = 0; c < 3; c++) {
            Hn[c] = a.Tn[c][l];
            Sn[c] = a.Sn[c][l];
        }
        return K_SPECIAL;
    }
    const float wx = x == 0 ? a.w[0] : 1.f, wy = y == 0 ? a.w[1] : 1.f, wz = z == 0 ? a.w[2] : 1.f;
    float P[3], M[3];
    P[0] = ec[2] - wy * eym[2];
    M[0] = -ec[1] + wz * ezm[1];
    P[1] = ec[0] - wz * ezm[0];
    M[1] = -ec[2] + wx * exm[2];
    P[2] = ec[1] - wx * exm[1];
    M[2] = -ec[0] + wy * eym[0];
    if (!pml) {
        if (z == a.iI) {
            const int64_t q = x + (int64_t)a.L0 * y;
            P[0] += a.sgn * (float)plane_source(a.srcE_plane, a.Einc, a.srcE_f2, a.srcE_re, a.srcE_im, q, 1);
            P[1] -= a.sgn * (float)plane_source(a.srcE_plane, a.Einc, a.srcE_f2, a.srcE_re, a.srcE_im, q, 0);
        }
#pragma unroll
        for (int c = 0; c < 3; c++) {
            Hn[c] = h[c] - b[c] * (P[c] + M[c]);
            Sn[c] = 0.f; /* defined on every path, so the array stays in registers */
        }
        return K_NORMAL;
    }
    const int jj[3] = {x, y, z};
#pragma unroll
    for (int c = 0; c < 3; c++) {
        const int dm = c == 2 ? 0 : c + 1, ds = c == 0 ? 2 : c - 1;
        const float *tm = axis_tab(a, dm), *ts = axis_tab(a, ds);
        const int Lm = axis_len(a, dm), Ls = axis_len(a, ds);
        const float aM = tm[jj[dm]], gM = tm[Lm + jj[dm]];
        const float aS = ts[jj[ds]], gS = ts[Ls + jj[ds]];
        const float s2 = aS * sp[c] - b[c] * gS * M[c];
        Hn[c] = aM * (h[c] - sp[c]) - b[c] * gM * P[c] + s2;
        Sn[c] = s2;
    }
    return K_PML;
}

/* Everything plane p contributes from global memory, loaded one plane ahead. */
struct plane_bundle {
    float e[2][3];                   /* old E of the two staged cells of this thread */
    float hA[3], bA[3], sA[3];       /* slot A: old H, bH, old H split (PML) */
    float hB[3], bB[3], sB[3];       /* slot B likewise */
    float be[3], se[3];              /* slot A: bE, old E split (PML), for E at plane p */
    bool pA, pB;                     /* slot A / B in PML at plane p */
};

template <int FX, int FY>
__global__ __launch_bounds__(kThreads) void fused_step_f32(fused_args a)
{
    constexpr int EXW = FX + 2, EYW = FY + 2;  /* E stage: x0-1 .. x0+FX, y0-1 .. y0+FY */
    constexpr int HXW = FX + 1, HYW = FY + 1;  /* H plane: x0 .. x0+FX, y0 .. y0+FY */
    constexpr int NHALO = HXW * HYW - FX * FY; /* grown-tile H cells outside the tile */
    constexpr int NESLOT = EXW * EYW;
    static_assert(NHALO <= kThreads && NESLOT <= 2 * kThreads && FX * FY == kThreads, "tile shape");
    __shared__ float sE[3][EYW][EXW];
    __shared__ float sH[3][HYW][HXW];

    const int id = xcd_logical(blockIdx.x, a.nblocks);
    const int bxi = id % a.gx, rest = id / a.gx;
    const int byi = rest % a.gy, bzi = rest / a.gy;
    const int x0 = bxi * FX, y0 = byi * FY, z0 = bzi * a.fz;
    const int z1 = min(z0 + a.fz, a.L2);
    const int t = threadIdx.x;

    /* tile cell of this thread (slot A) and, for t < NHALO, one grown-tile cell (slot B) */
    const int ai = t % FX, aj = t / FX;
    const bool owned = x0 + ai < a.L0 && y0 + aj < a.L1;
    const int ax = wrap_index(x0 + ai, a.L0), ay = wrap_index(y0 + aj, a.L1);
    const int64_t aq = ax + (int64_t)a.L0 * ay;
    const bool hasB = t < NHALO;
    const int bi = t < HXW ? t : FX, bj = t < HXW ? FY : t - HXW;
    const int bx = wrap_index(x0 + bi, a.L0), by = wrap_index(y0 + bj, a.L1);
    const int64_t bq = bx + (int64_t)a.L0 * by;
    const bool pxyA = a.pmlflag[ax] | a.pmlflag[a.L0 + ay];
    const bool pxyB = a.pmlflag[bx] | a.pmlflag[a.L0 + by];

    /* E stage cells loaded by this thread: t and t + 256 (< NESLOT) */
    int64_t eq[2];
    int es[2];
#pragma unroll
    for (int r = 0; r < 2; r++) {
        const int s = t + r * kThreads;
        es[r] = s < NESLOT ? s : -1;
        const int ex = s % EXW, ey = s / EXW;
        eq[r] = wrap_index(x0 - 1 + ex, a.L0) + (int64_t)a.L0 * wrap_index(y0 - 1 + ey, a.L1);
    }
    auto load_bundle = [&](int p, plane_bundle &B) {
        const int64_t off = (int64_t)p * a.sz;
        const bool pz = a.pmlflag[a.L0 + a.L1 + p];
        B.pA = pxyA | pz;
        B.pB = pxyB | pz;
#pragma unroll
        for (int r = 0; r < 2; r++)
#pragma unroll
            for (int c = 0; c < 3; c++) B.e[r][c] = es[r] >= 0 ? a.To[3 + c][off + eq[r]] : 0.f;
#pragma unroll
        for (int c = 0; c < 3; c++) {
            B.hA[c] = a.To[c][off + aq];
            B.bA[c] = a.b[c][off + aq];
            B.sA[c] = B.pA ? a.So[c][off + aq] : 0.f;
            B.be[c] = a.b[3 + c][off + aq];
            B.se[c] = B.pA ? a.So[3 + c][off + aq] : 0.f;
            B.hB[c] = hasB ? a.To[c][off + bq] : 0.f;
            B.bB[c] = hasB ? a.b[c][off + bq] : 0.f;
            B.sB[c] = hasB && B.pB ? a.So[c][off + bq] : 0.f;
        }
    };
    auto stage_to_lds = [&](const plane_bundle &B) {
#pragma unroll
        for (int r = 0; r < 2; r++) {
            if (es[r] < 0) continue;
            const int ex = es[r] % EXW, ey = es[r] / EXW;
#pragma unroll
            for (int c = 0; c < 3; c++) sE[c][ey][ex] = B.e[r][c];
        }
    };
    /* H at a grown-tile cell (i, j) of plane p from the staged E */
    auto evalH = [&](int i, int j, int x, int y, int p, int64_t l, const float ezm[3], const float h[3],
                     const float b[3], const float sp[3], bool pml, float Hn[3], float Sn[3], float ecenter[3]) {
        float exm[3], eym[3];
#pragma unroll
        for (int c = 0; c < 3; c++) {
            ecenter[c] = sE[c][j + 1][i + 1];
            exm[c] = sE[c][j + 1][i];
            eym[c] = sE[c][j][i + 1];
        }
        return fused_h_pre(a, x, y, p, l, ecenter, exm, eym, ezm, h, b, sp, pml, Hn, Sn);
    };

    float eA[3], eB[3]; /* old E of the previous plane at slots A, B (the -z neighbours) */
    float beK[3], seK[3];
    bool pK;
    plane_bundle cur, nxt;

    /* prologue: H(z0) on the grown tile */
    {
        const int zm = z0 == 0 ? a.L2 - 1 : z0 - 1;
        const int64_t offm = (int64_t)zm * a.sz;
#pragma unroll
        for (int c = 0; c < 3; c++) {
            eA[c] = a.To[3 + c][offm + aq];
            eB[c] = hasB ? a.To[3 + c][offm + bq] : 0.f;
        }
        load_bundle(z0, cur);
        load_bundle(z0 + 1 == a.L2 ? 0 : z0 + 1, nxt);
        stage_to_lds(cur);
        __syncthreads();
        const int64_t off = (int64_t)z0 * a.sz;
        float hA[3], sA[3], hB[3], sB[3], ec[3];
        const int kind = evalH(ai, aj, ax, ay, z0, off + aq, eA, cur.hA, cur.bA, cur.sA, cur.pA, hA, sA, ec);
        if (owned && kind != K_SPECIAL) {
            store3(a.Tn, off + aq, hA);
            if (kind == K_PML) store3(a.Sn, off + aq, sA);
        }
#pragma unroll
        for (int c = 0; c < 3; c++) eA[c] = ec[c];
        if (hasB) {
            evalH(bi, bj, bx, by, z0, off + bq, eB, cur.hB, cur.bB, cur.sB, cur.pB, hB, sB, ec);
#pragma unroll
            for (int c = 0; c < 3; c++) eB[c] = ec[c];
        }
#pragma unroll
        for (int c = 0; c < 3; c++) {
            beK[c] = cur.be[c];
            seK[c] = cur.se[c];
        }
        pK = cur.pA;
#pragma unroll
        for (int c = 0; c < 3; c++) {
            sH[c][aj][ai] = hA[c];
            if (hasB) sH[c][bj][bi] = hB[c];
        }
        __syncthreads();
    }

    for (int k = z0; k < z1; k++) {
        const int p = k + 1 == a.L2 ? 0 : k + 1;
        const bool p_owned = k + 1 < z1;
        const int64_t offk = (int64_t)k * a.sz, offp = (int64_t)p * a.sz;
        cur = nxt;                                   /* bundle of plane p */
        if (k + 1 < z1) load_bundle(p + 1 == a.L2 ? 0 : p + 1, nxt); /* prefetch plane p+1 */
        stage_to_lds(cur);
        __syncthreads();
        /* H(k+1) on the grown tile; eA/eB hold old E(k), the -z neighbours */
        float hnA[3], hnB[3], snA[3], snB[3], ecA[3], ecB[3];
        const int kindA = evalH(ai, aj, ax, ay, p, offp + aq, eA, cur.hA, cur.bA, cur.sA, cur.pA, hnA, snA, ecA);
        if (owned && p_owned && kindA != K_SPECIAL) {
            store3(a.Tn, offp + aq, hnA);
            if (kindA == K_PML) store3(a.Sn, offp + aq, snA);
        }
        if (hasB) evalH(bi, bj, bx, by, p, offp + bq, eB, cur.hB, cur.bB, cur.sB, cur.pB, hnB, snB, ecB);

        /* E(k) on the tile cell: forward curl of the new H */
        const int64_t l = offk + aq;
        if (owned && !__builtin_isnan(beK[0])) {
            float hc[3], hxp[3], hyp[3];
#pragma unroll
            for (int c = 0; c < 3; c++) {
                hc[c] = sH[c][aj][ai];
                hxp[c] = sH[c][aj][ai + 1];
                hyp[c] = sH[c][aj + 1][ai];
            }
            const float wx = ax + 1 == a.L0 ? a.w[0] : 1.f, wy = ay + 1 == a.L1 ? a.w[1] : 1.f;
            const float wz = k + 1 == a.L2 ? a.w[2] : 1.f;
            float P[3], M[3];
            P[0] = wy * hyp[2] - hc[2];
            M[0] = -wz * hnA[1] + hc[1];
            P[1] = wz * hnA[0] - hc[0];
            M[1] = -wx * hxp[2] + hc[2];
            P[2] = wx * hxp[1] - hc[1];
            M[2] = -wy * hyp[0] + hc[0];
            float En[3];
            if (!pK) {
                if (k == a.iI) {
                    P[0] += a.sgn * (float)plane_source(a.srcH_plane, a.Hinc, a.srcH_f2, a.srcH_re, a.srcH_im, aq, 1);
                    P[1] -= a.sgn * (float)plane_source(a.srcH_plane, a.Hinc, a.srcH_f2, a.srcH_re, a.srcH_im, aq, 0);
                }
#pragma unroll
                for (int c = 0; c < 3; c++) En[c] = eA[c] + beK[c] * (P[c] + M[c]);
                store3(a.Tn + 3, l, En);
                for (int m = 0; m < 2; m++) { /* monitors: main = total here (S = 0) */
                    if (k != (m == 0 ? a.iB : a.iT)) continue;
                    double *HM = m == 0 ? a.HB : a.HT, *EM = m == 0 ? a.EB : a.ET;
                    for (int v = 0; v < 2; v++) {
                        const int64_t o = 2 * (v + 2 * aq);
                        HM[o] += a.monH_re * (double)hc[v];
                        HM[o + 1] += a.monH_im * (double)hc[v];
                        EM[o] += a.monE_re * (double)En[v];
                        EM[o + 1] += a.monE_im * (double)En[v];
                    }
                }
            } else {
                const int jj[3] = {ax, ay, k};
                float Sn[3];
#pragma unroll
                for (int c = 0; c < 3; c++) {
                    const int dm = c == 2 ? 0 : c + 1, ds = c == 0 ? 2 : c - 1;
                    const float *tm = axis_tab(a, dm), *ts = axis_tab(a, ds);
                    const int Lm = axis_len(a, dm), Ls = axis_len(a, ds);
                    const float aM = tm[2 * Lm + jj[dm]], gM = tm[3 * Lm + jj[dm]];
                    const float aS = ts[2 * Ls + jj[ds]], gS = ts[3 * Ls + jj[ds]];
                    const float s2 = aS * seK[c] + beK[c] * gS * M[c];
                    En[c] = aM * (eA[c] - seK[c]) + beK[c] * gM * P[c] + s2;
                    Sn[c] = s2;
                }
                store3(a.Tn + 3, l, En);
                store3(a.Sn + 3, l, Sn);
            }
        }
#pragma unroll
        for (int c = 0; c < 3; c++) {
            eA[c] = ecA[c];
            if (hasB) eB[c] = ecB[c];
            beK[c] = cur.be[c];
            seK[c] = cur.se[c];
        }
        pK = cur.pA;
        __syncthreads();
#pragma unroll
        for (int c = 0; c < 3; c++) {
            sH[c][aj][ai] = hnA[c];
            if (hasB) sH[c][bj][bi] = hnB[c];
        }
    }
}

/* ---------------------------------------------------------------------------------- */
/* Fused step with asynchronous LDS staging                                           */
/* ---------------------------------------------------------------------------------- */
/*
 * The dataflow of fused_step_f32 (H(p) on the grown tile, then E(p-1) on the tile,
 * marching p over a z chunk), with every per-plane input copied from HBM straight into
 * LDS by buffer_load_dword ... lds two planes ahead of the plane being computed, into a
 * ring of three stage slots:
 *   E    old E on the E halo region (x0-1 .. x0+FX, y0-1 .. y0+FY)
 *   Ho   old H, bH (and the H split in PML stages) on the grown H tile
 *   bE   bE (and the E split in PML stages) on the tile
 * The copies occupy no registers, so a block keeps two planes of loads in flight while it
 * computes, where fused_step_f32 could hold one plane in registers.
 *
 * Ordering: every copy instruction writes 64 consecutive LDS words (lane i -> word i)
 * and counts in the issuing wave's vmcnt.  Vector memory operations of one wave complete
 * in issue order, and each wave issues at least kMinOps copies per stage (the list of a
 * stage is dealt round-robin to the 4 waves), so after issuing stage p+2 a wave that
 * waits for vmcnt <= 2*kMinOps holds its share of stage p in LDS; the s_barrier that
 * follows makes the whole stage visible.  Waiting for fewer outstanding operations than
 * are actually in flight (PML stages carry more copies, stores add more) only waits
 * longer.  Global stores and the few direct loads (source, monitors, lossy/metal cells)
 * count in the same vmcnt and the compiler's own waits account for the copies.
 */
template <int FX, int FY, bool SPLIT = true>
struct dma_layout {
    static constexpr int EXW = FX + 2, EYW = FY + 2, NE = EXW * EYW, NEC = (NE + 63) / 64;
    static constexpr int HXW = FX + 1, HYW = FY + 1, NH = HXW * HYW, NHC = (NH + 63) / 64;
    static constexpr int NTC = FX * FY / 64;
    /* word offsets inside a stage slot; each region is [3 components][chunks * 64] */
    static constexpr int E_OFF = 0;
    static constexpr int HO_OFF = E_OFF + 3 * NEC * 64;
    static constexpr int BH_OFF = HO_OFF + 3 * NHC * 64;
    static constexpr int BE_OFF = BH_OFF + 3 * NHC * 64;
    static constexpr int SH_OFF = BE_OFF + 3 * NTC * 64;
    static constexpr int SE_OFF = SH_OFF + 3 * NHC * 64;
    /* a slot carries the split regions only when the launch region may hold PML cells */
    static constexpr int SLOT = SPLIT ? SE_OFF + 3 * NTC * 64 : SH_OFF;
    static constexpr int NBASE = 3 * NEC + 6 * NHC + 3 * NTC; /* copies of every stage */
    static constexpr int NSPLIT = 3 * NHC + 3 * NTC;          /* extra copies of PML stages */
    static constexpr int kMinOps = NBASE / 4;                 /* per wave, at least */
    static constexpr int kMaxFZ = 64;
};


/* Copies of stage slot `slot` (plane byte offset soff) dealt to wave W.  Component c of
 * a group starts c*cb bytes into the group's buffer. */
template <int FX, int FY, bool SPLIT, int W>
__device__ inline void dma_issue(float *slot, __amdgpu_buffer_rsrc_t rT, __amdgpu_buffer_rsrc_t rB,
                                 __amdgpu_buffer_rsrc_t rS, unsigned cb, unsigned soff, bool split, const unsigned *vE,
                                 const unsigned *vH, const unsigned *vT)
{
    using Ly = dma_layout<FX, FY, SPLIT>;
#pragma unroll
    for (int j = W; j < Ly::NBASE; j += 4) {
        if (j < 3 * Ly::NEC) {
            const int c = j / Ly::NEC, ch = j % Ly::NEC;
            __builtin_amdgcn_raw_ptr_buffer_load_lds(rT, SMX_LDS(slot + Ly::E_OFF + c * Ly::NEC * 64 + ch * 64), 4,
                                                     vE[ch] + (3 + c) * cb, soff, 0, 0);
        } else if (j < 3 * Ly::NEC + 3 * Ly::NHC) {
            const int r = j - 3 * Ly::NEC, c = r / Ly::NHC, ch = r % Ly::NHC;
            __builtin_amdgcn_raw_ptr_buffer_load_lds(rT, SMX_LDS(slot + Ly::HO_OFF + c * Ly::NHC * 64 + ch * 64), 4,
                                                     vH[ch] + c * cb, soff, 0, 0);
        } else if (j < 3 * Ly::NEC + 6 * Ly::NHC) {
            const int r = j - 3 * Ly::NEC - 3 * Ly::NHC, c = r / Ly::NHC, ch = r % Ly::NHC;
            __builtin_amdgcn_raw_ptr_buffer_load_lds(rB, SMX_LDS(slot + Ly::BH_OFF + c * Ly::NHC * 64 + ch * 64), 4,
                                                     vH[ch] + c * cb, soff, 0, 0);
        } else {
            const int r = j - 3 * Ly::NEC - 6 * Ly::NHC, c = r / Ly::NTC, ch = r % Ly::NTC;
            __builtin_amdgcn_raw_ptr_buffer_load_lds(rB, SMX_LDS(slot + Ly::BE_OFF + c * Ly::NTC * 64 + ch * 64), 4,
                                                     vT[ch] + (3 + c) * cb, soff, 0, 0);
        }
    }
    if (SPLIT && split) {
#pragma unroll
        for (int j = W; j < Ly::NSPLIT; j += 4) {
            if (j < 3 * Ly::NHC) {
                const int c = j / Ly::NHC, ch = j % Ly::NHC;
                __builtin_amdgcn_raw_ptr_buffer_load_lds(rS, SMX_LDS(slot + Ly::SH_OFF + c * Ly::NHC * 64 + ch * 64),
                                                         4, vH[ch] + c * cb, soff, 0, 0);
            } else {
                const int r = j - 3 * Ly::NHC, c = r / Ly::NTC, ch = r % Ly::NTC;
                __builtin_amdgcn_raw_ptr_buffer_load_lds(rS, SMX_LDS(slot + Ly::SE_OFF + c * Ly::NTC * 64 + ch * 64),
                                                         4, vT[ch] + (3 + c) * cb, soff, 0, 0);
            }
        }
    }
}


/* New H of one regular cell from its curl halves: lossless X -= b*(P + M), or the PML
 * main/split pair with the axis factors pa (A) and pg (B/b) of the cell. */
__device__ inline void h_regular(const float h[3], const float b[3], const float sp[3], const float P[3],
                                 const float M[3], bool pml, const float pa[3], const float pg[3], float Hn[3],
                                 float Sn[3])
{
    if (!pml) {
#pragma unroll
        for (int c = 0; c < 3; c++) {
            Hn[c] = h[c] - b[c] * (P[c] + M[c]);
            Sn[c] = 0.f; /* defined on every path, so the array stays in registers */
        }
        return;
    }
#pragma unroll
    for (int c = 0; c < 3; c++) {
        const int dm = c == 2 ? 0 : c + 1, ds = c == 0 ? 2 : c - 1;
        const float s2 = pa[ds] * sp[c] - b[c] * pg[ds] * M[c];
        Hn[c] = pa[dm] * (h[c] - sp[c]) - b[c] * pg[dm] * P[c] + s2;
        Sn[c] = s2;
    }
}

template <int FX, int FY, int D, bool SPLIT>
__global__ __launch_bounds__(kThreads) void fused_dma_f32(fused_args a)
{
    using Ly = dma_layout<FX, FY, SPLIT>;
    constexpr int NSLOT = D + 1;
    static_assert(D >= 1 && D * Ly::kMinOps < 64, "copies in flight must fit the vmcnt field");
    constexpr int EXW = Ly::EXW, HXW = Ly::HXW, HYW = Ly::HYW;
    constexpr int NHALO = HXW * HYW - FX * FY;
    static_assert(FX * FY == kThreads && NHALO <= kThreads, "tile shape");
    /* One LDS array holds the three stage slots, the new H of the previous plane on the
     * grown tile (sH) and the chunk's z tables.  It must stay a single variable: with
     * several LDS variables the compiler starts tracking the asynchronous copies itself
     * and waits for all of them (vmcnt(0)) before every read of a slot. */
    constexpr int SH_BASE = NSLOT * Ly::SLOT, NHG = HYW * HXW;
    constexpr int ZT_BASE = SH_BASE + 3 * NHG;
    constexpr int FLAG = ZT_BASE + 5 * (Ly::kMaxFZ + 1);
    __shared__ float lds[FLAG + 1];
    auto sH = [&](int c, int j, int i) -> float & { return lds[SH_BASE + c * NHG + j * HXW + i]; };
    auto zt = [&](int p, int k) -> float & { return lds[ZT_BASE + 5 * p + k]; }; /* aH, gH, aE, gE, flag */

    const int id = xcd_logical(blockIdx.x, a.nblocks);
    const int bxi = id % a.gx, rest = id / a.gx;
    const int byi = rest % a.gy, bzi = rest / a.gy;
    const int x0 = a.box[0] + bxi * FX, y0 = a.box[2] + byi * FY, z0 = a.box[4] + bzi * a.fz;
    const int z1 = min(z0 + a.fz, a.box[5]);
    const int t = threadIdx.x, lane = t & 63;
    const int wv = __builtin_amdgcn_readfirstlane(t >> 6);

    /* per-lane source offsets (bytes within a plane) of every chunk of the three regions;
     * lanes past a region's end copy a valid cell into the slot's padding */
    unsigned vE[Ly::NEC], vH[Ly::NHC], vT[Ly::NTC];
#pragma unroll
    for (int ch = 0; ch < Ly::NEC; ch++) {
        const int e = min(ch * 64 + lane, Ly::NE - 1);
        vE[ch] = 4u * (unsigned)(wrap_index(x0 - 1 + e % EXW, a.L0) + a.L0 * wrap_index(y0 - 1 + e / EXW, a.L1));
    }
#pragma unroll
    for (int ch = 0; ch < Ly::NHC; ch++) {
        const int e = min(ch * 64 + lane, Ly::NH - 1);
        vH[ch] = 4u * (unsigned)(wrap_index(x0 + e % HXW, a.L0) + a.L0 * wrap_index(y0 + e / HXW, a.L1));
    }
#pragma unroll
    for (int ch = 0; ch < Ly::NTC; ch++) {
        const int e = ch * 64 + lane;
        vT[ch] = 4u * (unsigned)(wrap_index(x0 + e % FX, a.L0) + a.L0 * wrap_index(y0 + e / FX, a.L1));
    }

    /* this thread's tile cell (slot A) and, for t < NHALO, one grown-tile cell (slot B) */
    const int ai = t % FX, aj = t / FX;
    const bool owned = x0 + ai < a.box[1] && y0 + aj < a.box[3];
    const int ax = wrap_index(x0 + ai, a.L0), ay = wrap_index(y0 + aj, a.L1);
    const int64_t aq = ax + (int64_t)a.L0 * ay;
    const bool hasB = t < NHALO;
    const int bi = t < HXW ? t : FX, bj = t < HXW ? FY : t - HXW;
    const int bx = wrap_index(x0 + bi, a.L0), by = wrap_index(y0 + bj, a.L1);
    const int64_t bq = bx + (int64_t)a.L0 * by;

    /* x/y PML factors of the two cells (H: integer samples, E: half samples) */
    const float *tx = a.pml, *ty = a.pml + 4 * a.L0;
    /* without split slots the launch region holds no PML cell (checked on the host) */
    const bool pxyA = SPLIT && (a.pmlflag[ax] | a.pmlflag[a.L0 + ay]);
    const bool pxyB = SPLIT && hasB && (a.pmlflag[bx] | a.pmlflag[a.L0 + by]);
    const float hAx = tx[ax], hGx = tx[a.L0 + ax], eAx = tx[2 * a.L0 + ax], eGx = tx[3 * a.L0 + ax];
    const float hAy = ty[ay], hGy = ty[a.L1 + ay], eAy = ty[2 * a.L1 + ay], eGy = ty[3 * a.L1 + ay];
    const float hBx = hasB ? tx[bx] : 1.f, hBgx = hasB ? tx[a.L0 + bx] : 1.f;
    const float hBy = hasB ? ty[by] : 1.f, hBgy = hasB ? ty[a.L1 + by] : 1.f;
    /* z tables of the chunk's planes z0 .. z1 (z1 wraps to 0 at the top) */
    for (int i = t; i <= z1 - z0; i += kThreads) {
        const int z = z0 + i == a.L2 ? 0 : z0 + i;
        const float *tz = a.pml + 4 * (a.L0 + a.L1);
        zt(i, 0) = tz[z];
        zt(i, 1) = tz[a.L2 + z];
        zt(i, 2) = tz[2 * a.L2 + z];
        zt(i, 3) = tz[3 * a.L2 + z];
        zt(i, 4) = SPLIT && a.pmlflag[a.L0 + a.L1 + z] ? 1.f : 0.f;
    }
    /* does any cell of the grown tile lie in x/y PML?  (a vote through the same LDS
     * array: __syncthreads_or would bring an LDS variable of its own) */
    if (t == 0) lds[FLAG] = 0.f;
    __syncthreads();
    if (pxyA | pxyB) lds[FLAG] = 1.f;
    __syncthreads();
    const bool tile_pml = lds[FLAG] != 0.f;

    /* old E of plane z0-1 at the two cells: the -z neighbours of H(z0) */
    float eA[3], eB[3];
    {
        const int zm = z0 == 0 ? a.L2 - 1 : z0 - 1;
        const int64_t offm = (int64_t)zm * a.sz;
#pragma unroll
        for (int c = 0; c < 3; c++) {
            eA[c] = a.To[3 + c][offm + aq];
            eB[c] = hasB ? a.To[3 + c][offm + bq] : 0.f;
        }
        /* retire these loads before the copies start, so that no wait for them lands
         * inside the loop */
        retire_loads();
    }

    const __amdgpu_buffer_rsrc_t rT = group_rsrc(a.To[0], a.ncomp), rB = group_rsrc(a.b[0], a.ncomp),
                                 rS = group_rsrc(a.So[0], a.ncomp);
    const unsigned cb = 4u * (unsigned)a.ncomp;
    auto issue = [&](int p) { /* stage of chunk plane index p (0 .. z1-z0) into slot p % NSLOT */
        const int z = z0 + p == a.L2 ? 0 : z0 + p;
        const unsigned soff = 4u * (unsigned)((int64_t)z * a.sz);
        const bool split = tile_pml || zt(p, 4) != 0.f;
        float *slot = lds + (p % NSLOT) * Ly::SLOT;
        switch (wv) {
        case 0: dma_issue<FX, FY, SPLIT, 0>(slot, rT, rB, rS, cb, soff, split, vE, vH, vT); break;
        case 1: dma_issue<FX, FY, SPLIT, 1>(slot, rT, rB, rS, cb, soff, split, vE, vH, vT); break;
        case 2: dma_issue<FX, FY, SPLIT, 2>(slot, rT, rB, rS, cb, soff, split, vE, vH, vT); break;
        default: dma_issue<FX, FY, SPLIT, 3>(slot, rT, rB, rS, cb, soff, split, vE, vH, vT); break;
        }
    };

    const int np = z1 - z0; /* planes p = 0 .. np: H at z0+p, E at z0+p-1 */
    for (int p = 0; p < D && p <= np; p++) issue(p);

    float beK[3] = {0.f, 0.f, 0.f}, seK[3] = {0.f, 0.f, 0.f}; /* bE, E split of plane p-1 */
    bool pK = false;
    for (int p = 0; p <= np; p++) {
        /* stage p is complete once at most (stages issued after it) * kMinOps copies of
         * this wave are outstanding */
        if (p + D <= np) {
            issue(p + D);
            wait_vm_barrier<D * Ly::kMinOps>();
        } else {
            switch (np - p) { /* stages already issued after p: np - p < D */
            case 0: wait_vm_barrier<0>(); break;
            case 1: wait_vm_barrier<Ly::kMinOps>(); break;
            case 2: wait_vm_barrier<(D > 2 ? 2 : 0) * Ly::kMinOps>(); break;
            default: wait_vm_barrier<(D > 3 ? 3 : 0) * Ly::kMinOps>(); break;
            }
        }
        const float *S = lds + (p % NSLOT) * Ly::SLOT;
        const int z = z0 + p == a.L2 ? 0 : z0 + p;
        const int64_t offp = (int64_t)z * a.sz;
        const bool zp = zt(p, 4) != 0.f;
        const float hz[2] = {zt(p, 0), zt(p, 1)};

        /* H(z) on the grown tile */
        auto evalH = [&](int i, int j, int x, int y, int64_t l, const float ezm[3], bool pml, const float pa[3],
                         const float pg[3], float Hn[3], float Sn[3], float ecen[3]) {
            float h[3], b[3], sp[3], exm[3], eym[3];
#pragma unroll
            for (int c = 0; c < 3; c++) {
                const float *Ec = S + Ly::E_OFF + c * Ly::NEC * 64;
                ecen[c] = Ec[(j + 1) * EXW + i + 1];
                exm[c] = Ec[(j + 1) * EXW + i];
                eym[c] = Ec[j * EXW + i + 1];
                h[c] = S[Ly::HO_OFF + c * Ly::NHC * 64 + j * HXW + i];
                b[c] = S[Ly::BH_OFF + c * Ly::NHC * 64 + j * HXW + i];
                sp[c] = SPLIT && pml ? S[Ly::SH_OFF + c * Ly::NHC * 64 + j * HXW + i] : 0.f;
            }
            if (__builtin_isnan(b[0])) { /* lossy/metal: computed by special_h_pass */
#pragma unroll
                for (int c = 0; c < 3; c++) {
                    Hn[c] = a.Tn[c][l];
                    Sn[c] = a.Sn[c][l];
                }
                retire_loads();
                return K_SPECIAL;
            }
            const float wx = x == 0 ? a.w[0] : 1.f, wy = y == 0 ? a.w[1] : 1.f, wz = z == 0 ? a.w[2] : 1.f;
            float P[3], M[3];
            P[0] = ecen[2] - wy * eym[2];
            M[0] = -ecen[1] + wz * ezm[1];
            P[1] = ecen[0] - wz * ezm[0];
            M[1] = -ecen[2] + wx * exm[2];
            P[2] = ecen[1] - wx * exm[1];
            M[2] = -ecen[0] + wy * eym[0];
            if (!pml && z == a.iI) {
                const int64_t q = x + (int64_t)a.L0 * y;
                P[0] += a.sgn * (float)plane_source(a.srcE_plane, a.Einc, a.srcE_f2, a.srcE_re, a.srcE_im, q, 1);
                P[1] -= a.sgn * (float)plane_source(a.srcE_plane, a.Einc, a.srcE_f2, a.srcE_re, a.srcE_im, q, 0);
                retire_loads();
            }
            h_regular(h, b, sp, P, M, pml, pa, pg, Hn, Sn);
            return pml ? K_PML : K_NORMAL;
        };
        const bool pmlA = pxyA | zp, pmlB = pxyB | zp;
        float hnA[3], snA[3], hnB[3], snB[3], ecA[3], ecB[3];
        {
            const float pa[3] = {hAx, hAy, hz[0]}, pg[3] = {hGx, hGy, hz[1]};
            const int kind = evalH(ai, aj, ax, ay, offp + aq, eA, pmlA, pa, pg, hnA, snA, ecA);
            if (owned && p < np && kind != K_SPECIAL) {
                store3(a.Tn, offp + aq, hnA);
                if (kind == K_PML) store3(a.Sn, offp + aq, snA);
            }
        }
        if (hasB) {
            const float pa[3] = {hBx, hBy, hz[0]}, pg[3] = {hBgx, hBgy, hz[1]};
            evalH(bi, bj, bx, by, offp + bq, eB, pmlB, pa, pg, hnB, snB, ecB);
        }

        /* E(z-1) on the tile cell: forward curl of the new H */
        if (p > 0 && owned && !__builtin_isnan(beK[0])) {
            const int k = z0 + p - 1;
            const int64_t l = (int64_t)k * a.sz + aq;
            float hc[3], hxp[3], hyp[3];
#pragma unroll
            for (int c = 0; c < 3; c++) {
                hc[c] = sH(c, aj, ai);
                hxp[c] = sH(c, aj, ai + 1);
                hyp[c] = sH(c, aj + 1, ai);
            }
            const float wx = ax + 1 == a.L0 ? a.w[0] : 1.f, wy = ay + 1 == a.L1 ? a.w[1] : 1.f;
            const float wz = k + 1 == a.L2 ? a.w[2] : 1.f;
            float P[3], M[3];
            P[0] = wy * hyp[2] - hc[2];
            M[0] = -wz * hnA[1] + hc[1];
            P[1] = wz * hnA[0] - hc[0];
            M[1] = -wx * hxp[2] + hc[2];
            P[2] = wx * hxp[1] - hc[1];
            M[2] = -wy * hyp[0] + hc[0];
            float En[3];
            if (!pK) {
                if (k == a.iI) {
                    P[0] += a.sgn * (float)plane_source(a.srcH_plane, a.Hinc, a.srcH_f2, a.srcH_re, a.srcH_im, aq, 1);
                    P[1] -= a.sgn * (float)plane_source(a.srcH_plane, a.Hinc, a.srcH_f2, a.srcH_re, a.srcH_im, aq, 0);
                    retire_loads();
                }
#pragma unroll
                for (int c = 0; c < 3; c++) En[c] = eA[c] + beK[c] * (P[c] + M[c]);
                store3(a.Tn + 3, l, En);
                if (k == a.iB || k == a.iT) { /* monitors: main = total here (S = 0) */
                    for (int m = 0; m < 2; m++) {
                        if (k != (m == 0 ? a.iB : a.iT)) continue;
                        double *HM = m == 0 ? a.HB : a.HT, *EM = m == 0 ? a.EB : a.ET;
                        for (int v = 0; v < 2; v++) {
                            const int64_t o = 2 * (v + 2 * aq);
                            HM[o] += a.monH_re * (double)hc[v];
                            HM[o + 1] += a.monH_im * (double)hc[v];
                            EM[o] += a.monE_re * (double)En[v];
                            EM[o + 1] += a.monE_im * (double)En[v];
                        }
                    }
                    retire_loads();
                }
            } else {
                const float pa[3] = {eAx, eAy, zt(p - 1, 2)}, pg[3] = {eGx, eGy, zt(p - 1, 3)};
                float Sn[3];
#pragma unroll
                for (int c = 0; c < 3; c++) {
                    const int dm = c == 2 ? 0 : c + 1, ds = c == 0 ? 2 : c - 1;
                    const float s2 = pa[ds] * seK[c] + beK[c] * pg[ds] * M[c];
                    En[c] = pa[dm] * (eA[c] - seK[c]) + beK[c] * pg[dm] * P[c] + s2;
                    Sn[c] = s2;
                }
                store3(a.Tn + 3, l, En);
                store3(a.Sn + 3, l, Sn);
            }
        }
        /* carry plane z to the next iteration */
#pragma unroll
        for (int c = 0; c < 3; c++) {
            eA[c] = ecA[c];
            if (hasB) eB[c] = ecB[c];
            beK[c] = S[Ly::BE_OFF + c * Ly::NTC * 64 + t];
            seK[c] = SPLIT && pmlA ? S[Ly::SE_OFF + c * Ly::NTC * 64 + t] : 0.f;
        }
        pK = pmlA;
        /* every wave is done with stage p and with sH before either is overwritten */
        wait_vm_barrier<63>();
#pragma unroll
        for (int c = 0; c < 3; c++) {
            sH(c, aj, ai) = hnA[c];
            if (hasB) sH(c, bj, bi) = hnB[c];
        }
    }
}

/* ---------------------------------------------------------------------------------- */
/* Two steps per pass: temporal blocking of the staged z-march                        */
/* ---------------------------------------------------------------------------------- */
/*
 * The step is bound by HBM (one step reads 12 words per cell and writes 6), so the way
 * past the streaming ceiling is to read and write the fields once per two steps.  A
 * block owns an FX x FY tile and a z chunk [z0, z1), and marches planes p = z0-1 ..
 * z1+1 carrying two time levels, skewed by one plane:
 *   a. H1(p)   on G3 from E0(p), E0(p-1)            (first step, H)
 *   b. E1(p-1) on G2 from H1(p-1), H1(p)            (first step, E)
 *   c. H2(p-1) on G1 from E1(p-1), E1(p-2)          (second step, H)
 *   d. E2(p-2) on T  from H2(p-2), H2(p-1)          (second step, E)
 * with T the tile, G1 = T grown by one cell on +x/+y, G2 = G1 grown by one on -x/-y,
 * G3 = G2 grown by one on +x/+y, G4 = G3 grown by one on -x/-y (the stage region).
 * Only H2 and E2 of the tile reach HBM; the halo levels are recomputed by every tile
 * that needs them, and the chunk ends recompute two planes each side.  Per plane the
 * old E, H, bH, bE of G4 come in by asynchronous copies to LDS (as in fused_dma_f32);
 * H1, E1, H2 of the neighbours go through LDS, the -z / +z neighbours of a thread's own
 * cells stay in its registers.
 *
 * On grids with lossy/metal cells the tiles near them are left to region sweeps and the
 * special-cell passes (smx_f32_step2) and the kernel runs the rest.  With SPLIT
 * the stages of PML planes and of tiles touching x/y PML also carry the old splits, and
 * PML cells take the main/split update of the one-step kernels, their splits carried
 * like the fields; without it the kernel runs only where no PML cell is reached (the
 * whole grid, or the box the region sweeps surround).  Source terms and the reference
 * monitors of both steps are applied as in the one-step kernels.
 */
template <int FX, int FY, bool SPLIT, bool WIDE>
struct two_layout {
    static constexpr int NT = FX * FY, NW = NT / 64;        /* one thread per tile cell */
    /* G4 starts OX columns left of the tile.  WIDE: every copy moves 16 bytes per lane (4
     * consecutive x, lane i -> words 4i..4i+3 of a 256-word run), the region starts at a
     * multiple of 4 and is 4 columns wider, a quarter of the copy instructions */
    static constexpr int OX = WIDE ? 4 : 2;
    static constexpr int W4 = FX + 2 * OX, H4 = FY + 4, N4 = W4 * H4;
    static constexpr int CH = WIDE ? 256 : 64;             /* words per copy instruction */
    static constexpr int NC = (N4 + CH - 1) / CH;          /* copies per component */
    static constexpr int ARR = 3 * NC * CH;                /* one array (3 components) of a stage */
    static constexpr int SLOT = (SPLIT ? 6 : 4) * ARR;     /* E0, H0, bH, bE [, SH0, SE0] */
    /* rows of the G4 region each array component k = 3 arr + c is read on: E0 on all of
     * G4, but its y component (read one row up only for Ex and Ez) from the second row;
     * H0, bH, SH0 on G3, from the second row, their y components (whose H1 the last row
     * does not need: that row is read only for Hx and Hz) a row short; bE, SE0 on G2, the
     * second row to the second-last.  A component's copies start at its first row, and
     * its LDS words count from there */
    static constexpr int row0(int k) { return WIDE || k == 0 || k == 2 ? 0 : 1; }
    static constexpr int rows(int k)
    {
        return WIDE || k == 0 || k == 2 ? H4
               : k == 1                 ? H4 - 1
               : k >= 9 && k < 12       ? H4 - 2 /* bE */
               : k >= 15                ? H4 - 2 /* SE0 */
               : k % 3 == 1             ? H4 - 2 /* H0, bH, SH0: y */
                                        : H4 - 1;
    }
    static constexpr int nca(int k) { return WIDE ? NC : (rows(k) * W4 + CH - 1) / CH; }
    static constexpr int cum(int k) { return k == 0 ? 0 : cum(k - 1) + nca(k - 1); }
    /* the component copy j of a stage belongs to */
    static constexpr int kof(int j) { return j < cum(1) ? 0 : 1 + kof_from(j, 1); }
    static constexpr int kof_from(int j, int k) { return k >= 17 || j < cum(k + 1) ? 0 : 1 + kof_from(j, k + 1); }
    static constexpr int NBASE = cum(12);                  /* copies of every stage */
    static constexpr int NSPLIT = cum(18) - cum(12);       /* extra copies of PML stages */
    static constexpr int kMinOps = NBASE / NW;             /* per wave, at least */
    static constexpr int CELLS = 2;                        /* a tile cell and a ring cell */
    static constexpr int kMaxFZ = 80;                      /* z tables of a chunk in LDS */
};

/* Copies of one stage dealt to wave W: arrays E0, H0, bH, bE (and SH0, SE0 when `split`),
 * each component k in nca(k) runs of CH words of the rows of the G4 region it is read on
 * (vq: from the first row; vr: from the second; vlast: the last run of the components that
 * also end a row before the region's last). */
/* Where a stage's copies read: one buffer resource per group of 6 component arrays
 * (component c at c*cb bytes, the plane at soff), or -- with BIG, for groups of 4 GiB and
 * more -- a resource built per copy on one array's plane (base + comp*ncomp + zoff
 * floats, plane_bytes long), so that every offset stays below 2^32. */
struct two_src {
    __amdgpu_buffer_rsrc_t rT, rB, rS;
    unsigned cb, soff;
    const float *pT, *pB, *pS;
    int64_t ncomp, zoff;
    unsigned plane_bytes;
};

template <bool BIG>
__device__ inline __amdgpu_buffer_rsrc_t two_res(const two_src &q, int group, unsigned comp, unsigned &voff)
{
    if constexpr (BIG) {
        const float *base = group == 0 ? q.pT : group == 1 ? q.pB : q.pS;
        return __builtin_amdgcn_make_buffer_rsrc((void *)(base + comp * q.ncomp + q.zoff), (short)0,
                                                 (int)q.plane_bytes, (int)kBufWord3);
    } else {
        voff += comp * q.cb;
        return group == 0 ? q.rT : group == 1 ? q.rB : q.rS;
    }
}

template <int FX, int FY, bool SPLIT, bool WIDE, bool BIG, int W>
__device__ inline void two_issue(float *slot, const two_src &q, bool split, const unsigned *vq, const unsigned *vr,
                                 unsigned vlast)
{
    using Ly = two_layout<FX, FY, SPLIT, WIDE>;
    /* copy j of a stage: component k = 3 arr + c, run ch (all known at compile time) */
    auto one = [&](auto jc) {
        const unsigned so = BIG ? 0u : q.soff;
        constexpr int j = decltype(jc)::value, k = Ly::kof(j), arr = k / 3, c = k % 3, ch = j - Ly::cum(k);
        /* E0 = T comp 3+c, H0 = T comp c, bH = b comp c, bE = b comp 3+c, SH0 = S comp c,
         * SE0 = S comp 3+c */
        constexpr unsigned comp = arr >= 4 ? (arr - 4) * 3 + c : (arr == 0 || arr == 3) ? 3 + c : c;
        constexpr int group = arr < 2 ? 0 : arr < 4 ? 1 : 2;
        float *dst = slot + arr * Ly::ARR + (c * Ly::NC + ch) * Ly::CH;
        unsigned voff = Ly::row0(k) == 0                                      ? vq[ch]
                        : Ly::rows(k) == Ly::H4 - 2 && ch == Ly::nca(k) - 1 ? vlast
                                                                               : vr[ch];
        const __amdgpu_buffer_rsrc_t res = two_res<BIG>(q, group, comp, voff);
        /* the copy's size must be a literal */
        if constexpr (WIDE) __builtin_amdgcn_raw_ptr_buffer_load_lds(res, SMX_LDS(dst), 16, voff, so, 0, 0);
        else __builtin_amdgcn_raw_ptr_buffer_load_lds(res, SMX_LDS(dst), 4, voff, so, 0, 0);
    };
    static_for<W, Ly::NBASE, Ly::NW>(one);
    if (SPLIT && split) static_for<Ly::NBASE + W, Ly::NBASE + Ly::NSPLIT, Ly::NW>(one);
}

/* New E of one regular cell: lossless X += b*(P + M), or the PML main/split pair. */
__device__ inline void e_regular(const float e[3], const float b[3], const float se[3], const float P[3],
                                 const float M[3], bool pml, const float pa[3], const float pg[3], float En[3],
                                 float Sn[3])
{
    if (!pml) {
#pragma unroll
        for (int c = 0; c < 3; c++) {
            En[c] = e[c] + b[c] * (P[c] + M[c]);
            Sn[c] = 0.f;
        }
        return;
    }
#pragma unroll
    for (int c = 0; c < 3; c++) {
        const int dm = c == 2 ? 0 : c + 1, ds = c == 0 ? 2 : c - 1;
        const float s2 = pa[ds] * se[c] + b[c] * pg[ds] * M[c];
        En[c] = pa[dm] * (e[c] - se[c]) + b[c] * pg[dm] * P[c] + s2;
        Sn[c] = s2;
    }
}

/* The pass of one block `id` of a plan whose arguments `a` are the kernel's own (one
 * plan per launch) or an entry of a table in constant memory (several plans per launch,
 * two_step_multi); scal (may be null) holds the 20 per-step scalars instead of a. */
template <int FX, int FY, int D, bool SPLIT, bool WIDE, bool DFT, bool BIG, bool BOX, typename A>
__device__ __forceinline__ void two_step_body(const A &a, int id, const double *scal)
{
    using Ly = two_layout<FX, FY, SPLIT, WIDE>;
    constexpr int W4 = Ly::W4, N4 = Ly::N4, NSLOT = D + 1, CELLS = Ly::CELLS;
    constexpr int NT = Ly::NT, OX = Ly::OX;
    static_assert((NT == 256 || NT == 512) && FX % 32 == 0, "tile shape");
    /* ring cells: G1\T, G2\G1, G3\G2 */
    constexpr int RC = FY + 1 + FX, RB = RC + FY + 2 + FX + 1, RA = RB + FY + 3 + FX + 2;
    static_assert(RA <= NT, "rings fit one cell per thread");
    static_assert(D >= 1 && D * Ly::kMinOps < 64, "copies in flight must fit the vmcnt field");
    /* one LDS variable (see fused_dma_f32): stage slots, H1, E1, H2 of the previous plane
     * in G4 layout, the chunk's z tables, a flag word */
    constexpr int L_H1 = NSLOT * Ly::SLOT, L_E1 = L_H1 + 3 * N4, L_H2 = L_E1 + 3 * N4;
    constexpr int L_ZT = L_H2 + 3 * N4, L_FLAG = L_ZT + (SPLIT ? 5 * (Ly::kMaxFZ + 4) : 0);
    constexpr int L_SC = (L_FLAG + 2) & ~1; /* 20 doubles: the two steps' source and monitor scalars */
    /* DFT: ring plane of each stage's plane, or -1 (with SPLIT, folded into the z table's
     * flag word, which keeps the LDS of two blocks within a CU's 160 KiB) */
    constexpr int L_DZ = L_SC + 2 * 20;
    __shared__ __attribute__((aligned(16))) float lds[L_DZ + (DFT && !SPLIT ? Ly::kMaxFZ + 4 : 0)];

    const int bxi = id % a.gx, rest = id / a.gx;
    const int byi = rest % a.gy, bzi = rest / a.gy;
    const int x0 = a.xorg + bxi * FX, y0 = a.box[2] + byi * FY, z0 = a.zc[bzi], z1 = a.zc[bzi + 1];
    const int t = threadIdx.x, lane = t & 63;
    const int wv = __builtin_amdgcn_readfirstlane(t >> 6);
    /* The per-step scalars are read only on the source and monitor planes: kept in LDS
     * rather than in scalar registers across the z-march (the first barrier of the
     * march publishes them). */
    if (t == 0 && scal) {
        double *d = reinterpret_cast<double *>(lds + L_SC);
        for (int i = 0; i < 20; i++) d[i] = scal[i];
    } else if (t == 0) {
        double *d = reinterpret_cast<double *>(lds + L_SC);
        d[0] = a.srcE_f2, d[1] = a.srcE_re, d[2] = a.srcE_im, d[3] = a.srcH_f2, d[4] = a.srcH_re;
        d[5] = a.srcH_im, d[6] = a.monH_re, d[7] = a.monH_im, d[8] = a.monE_re, d[9] = a.monE_im;
        d[10] = a.srcE2_f2, d[11] = a.srcE2_re, d[12] = a.srcE2_im, d[13] = a.srcH2_f2, d[14] = a.srcH2_re;
        d[15] = a.srcH2_im, d[16] = a.monH2_re, d[17] = a.monH2_im, d[18] = a.monE2_re, d[19] = a.monE2_im;
    }
    const double *sc = reinterpret_cast<const double *>(lds + L_SC);
    /* source profiles Hinc, Einc and the monitors HB, EB, HT, ET are consecutive arrays
     * of 4*L0*L1 doubles (the host checks) */
    const double *Hinc = a.Hinc, *Einc = a.Hinc + 4 * a.sz;
    auto mon = [&](int k) { return a.HB + 4 * k * a.sz; }; /* HB, EB, HT, ET */
    const int ns = z1 - z0 + 4; /* stages: planes z0-2 .. z1+1 */
    auto plane_of = [&](int s) { return wrap_index(z0 - 2 + s, a.L2); };
    /* DFT planes among this block's own planes: the first two, and whether there are
     * more (then every plane is checked against the whole list) */
    /* DFT: for each stage, the ring plane its plane is sampled into (only the block's own
     * planes [z0, z1), once), published by the first barrier of the march; the ring of
     * the first step as a buffer resource, the second step's slot dstep bytes further */
    __amdgpu_buffer_rsrc_t rD = __builtin_amdgcn_make_buffer_rsrc(nullptr, (short)0, 0, 0);
    unsigned dstep = 0;
    auto dft_plane_of = [&](int i) { /* ring plane of stage i's plane, -1 if none or not the block's */
        const int z = z0 - 2 + i;
        int m = -1;
        for (int k = 0; k < a.dft_n && z >= z0 && z < z1; k++)
            if (a.dft_planes[k] == z) m = k;
        return m;
    };
    if constexpr (DFT) {
        if constexpr (!SPLIT)
            for (int i = t; i < ns; i += NT) reinterpret_cast<int *>(lds + L_DZ)[i] = dft_plane_of(i);
        const int64_t slot_words = 4 * (int64_t)a.dft_n * a.sz;
        const int dslot = scal ? (int)scal[20] : a.dft_slot; /* several plans: per pass, after the scalars */
        rD = __builtin_amdgcn_make_buffer_rsrc((void *)(a.dft + dslot * slot_words), (short)0,
                                               (int)(unsigned)(8 * slot_words), (int)kBufWord3);
        dstep = 4u * (unsigned)slot_words;
    }
    /* sample (Hx, Hy, Ex, Ey) of cell q at stage s into the ring (second step: soff = dstep):
     * the total field, which is the main slot on the lossless grids this kernel takes
     * outside PML, and the sample taken in PML cells */
    /* ring plane of stage s_ (-1: none); read one stage ahead of its use, so that the
     * next barrier's LDS wait covers the read */
    auto dft_m = [&](int s_) {
        if constexpr (!DFT) return -1;
        return SPLIT ? ((int)lds[L_ZT + 5 * s_ + 4] >> 1) - 1 : reinterpret_cast<const int *>(lds + L_DZ)[s_];
    };
    /* E of a tile cell after step `step` (0, 1) into the design-box output */
    auto box_save = [&](int x, int y, int z, int step, const float v[3]) {
        if constexpr (BOX) {
            if (x < a.dbox[0] || x >= a.dbox[1] || y < a.dbox[2] || y >= a.dbox[3] || z < a.dbox[4] || z >= a.dbox[5])
                return;
            const int64_t i = (x - a.dbox[0]) + (int64_t)a.dbx * ((y - a.dbox[2]) + (int64_t)a.dby * (z - a.dbox[4]));
            float *o = a.dbox_out + step * 3 * a.dnbox;
            o[i] = v[0];
            o[a.dnbox + i] = v[1];
            o[2 * a.dnbox + i] = v[2];
        }
    };
    auto dft_sample = [&](int m, int64_t q, unsigned soff, float hx, float hy, float ex, float ey) {
        if constexpr (DFT) {
            if (m < 0) return;
            const unsigned o = 4u * (unsigned)(4 * m * a.sz + q), pl = 4u * (unsigned)a.sz;
            __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(hx), rD, o, soff, 0);
            __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(hy), rD, o + pl, soff, 0);
            __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(ex), rD, o + 2 * pl, soff, 0);
            __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(ey), rD, o + 3 * pl, soff, 0);
        }
    };

    /* byte offset in a plane of the first word lane `lane` copies in run ch (WIDE: a group
     * of 4 x; the groups past the region repeat the last one, into the run's padding) */
    unsigned vq[Ly::NC], vr[Ly::NC], vlast = 0;
    auto region_off = [&](int e) {
        return 4u * (unsigned)(wrap_index(x0 - OX + e % W4, a.L0) + a.L0 * wrap_index(y0 - 2 + e / W4, a.L1));
    };
#pragma unroll
    for (int ch = 0; ch < Ly::NC; ch++) {
        vq[ch] = region_off(WIDE ? 4 * min(ch * 64 + lane, N4 / 4 - 1) : min(ch * 64 + lane, N4 - 1));
        /* the arrays read from the second row on (from the third-last: to the last row) */
        vr[ch] = WIDE ? vq[ch] : region_off(min(W4 + ch * 64 + lane, N4 - 1));
    }
    if constexpr (!WIDE) vlast = region_off(min(W4 + (Ly::nca(9) - 1) * 64 + lane, N4 - W4 - 1));

    /* this thread's cells: r = 0 the tile cell t (every step computes it); r = 1 one cell
     * of the rings the first three updates add around the tile -- G1\T (used by c, b,
     * a), G2\G1 (b, a), G3\G2 (a), 129 cells for 32x8 -- ring cell h on thread
     * (h % 4)*64 + h/4, so the four waves share the extra work evenly; G4\G3 is only
     * read.  Local coordinates (gx, gy) count from the G4 corner (x0-OX, y0-2). */
    int ce[CELLS], cx[CELLS], cy[CELLS];
    bool inG3[CELLS], inG2[CELLS], inG1[CELLS], inT[CELLS], pxy[CELLS];
    float tH[CELLS][4], tE[CELLS][4]; /* x/y PML factors: aH_x, gH_x, aH_y, gH_y (E likewise) */
    int64_t cq[CELLS];
#pragma unroll
    for (int r = 0; r < CELLS; r++) {
        int gx, gy, ring = 0; /* ring: 0 tile, 1 G1\T, 2 G2\G1, 3 G3\G2, 4 none */
        if (r == 0) {
            gx = OX + t % FX;
            gy = 2 + t / FX;
        } else {
            /* 256 threads (two blocks per CU): ring cells spread over the 4 waves; 512
             * threads (one block per CU): on the first waves, so the SIMDs that run
             * them run a ring-free wave beside each */
            const int h = NT == 256 ? (t & 63) * 4 + (t >> 6) : t;
            if (h < RC) { /* x = x0+FX column, then y = y0+FY row */
                ring = 1;
                gx = h < FY + 1 ? OX + FX : OX + (h - (FY + 1));
                gy = h < FY + 1 ? 2 + h : FY + 2;
            } else if (h < RB) { /* x = x0-1 column, then y = y0-1 row */
                const int k = h - RC;
                ring = 2;
                gx = k < FY + 2 ? OX - 1 : OX + (k - (FY + 2));
                gy = k < FY + 2 ? 1 + k : 1;
            } else if (h < RA) { /* x = x0+FX+1 column, then y = y0+FY+1 row */
                const int k = h - RB;
                ring = 3;
                gx = k < FY + 3 ? OX + FX + 1 : OX - 1 + (k - (FY + 3));
                gy = k < FY + 3 ? 1 + k : FY + 3;
            } else {
                ring = 4;
                gx = gy = 0;
            }
        }
        const bool valid = ring < 4;
        const int e = gy * W4 + gx;
        ce[r] = valid ? e : W4 + 1; /* idle lanes compute on a cell whose neighbours exist */
        cx[r] = wrap_index(x0 - OX + gx, a.L0);
        cy[r] = wrap_index(y0 - 2 + gy, a.L1);
        cq[r] = cx[r] + (int64_t)a.L0 * cy[r];
        inG3[r] = ring <= 3;
        inG2[r] = ring <= 2;
        inG1[r] = ring <= 1;
        inT[r] = r == 0 && x0 - OX + gx >= a.xlo && x0 - OX + gx < a.xhi && y0 - 2 + gy < a.box[3];
        pxy[r] = false;
#pragma unroll
        for (int k = 0; k < 4; k++) tH[r][k] = tE[r][k] = 1.f;
        if constexpr (SPLIT) {
            const float *tx = a.pml, *ty = a.pml + 4 * a.L0;
            pxy[r] = valid && (a.pmlflag[cx[r]] | a.pmlflag[a.L0 + cy[r]]);
            tH[r][0] = tx[cx[r]], tH[r][1] = tx[a.L0 + cx[r]], tH[r][2] = ty[cy[r]], tH[r][3] = ty[a.L1 + cy[r]];
            tE[r][0] = tx[2 * a.L0 + cx[r]], tE[r][1] = tx[3 * a.L0 + cx[r]];
            tE[r][2] = ty[2 * a.L1 + cy[r]], tE[r][3] = ty[3 * a.L1 + cy[r]];
        }
    }
    bool tile_pml = false;
    if constexpr (SPLIT) {
        for (int i = t; i < ns; i += NT) {
            const int z = plane_of(i);
            const float *tz = a.pml + 4 * (a.L0 + a.L1);
            lds[L_ZT + 5 * i + 0] = tz[z];
            lds[L_ZT + 5 * i + 1] = tz[a.L2 + z];
            lds[L_ZT + 5 * i + 2] = tz[2 * a.L2 + z];
            lds[L_ZT + 5 * i + 3] = tz[3 * a.L2 + z];
            /* bit 0: z PML; bits 1..: 1 + the DFT ring plane (DFT variants) */
            lds[L_ZT + 5 * i + 4] = (float)((a.pmlflag[a.L0 + a.L1 + z] ? 1 : 0) + (DFT ? 2 * (dft_plane_of(i) + 1) : 0));
        }
        /* does any cell of the G4 region lie in x/y PML? (vote through the LDS array) */
        if (t == 0) lds[L_FLAG] = 0.f;
        __syncthreads();
        if (pxy[0] | pxy[1]) lds[L_FLAG] = 1.f;
        __syncthreads();
        tile_pml = lds[L_FLAG] != 0.f;
        retire_loads(); /* the table loads, before the copies start */
    }
    auto zpml = [&](int s) { return SPLIT && ((int)lds[L_ZT + 5 * s + 4] & 1); };

    const __amdgpu_buffer_rsrc_t rT = group_rsrc(a.To[0], a.ncomp), rB = group_rsrc(a.b[0], a.ncomp),
                                 rS = group_rsrc(a.So[0], a.ncomp);
    const unsigned cb = 4u * (unsigned)a.ncomp;
    /* the new fields and splits: one resource per group, as the copies (BIG: 64-bit
     * addresses from the groups' first arrays) */
    const __amdgpu_buffer_rsrc_t rTn = group_rsrc(a.Tn[0], a.ncomp), rSn = group_rsrc(a.Sn[0], a.ncomp);
    auto store = [&](bool split_group, int c0, int64_t l, const float v[3]) {
        if constexpr (BIG) {
            float *base = (split_group ? a.Sn[0] : a.Tn[0]) + (int64_t)c0 * a.ncomp + l;
            base[0] = v[0];
            base[a.ncomp] = v[1];
            base[2 * a.ncomp] = v[2];
        } else {
            bstore3(split_group ? rSn : rTn, cb, c0, l, v);
        }
    };
    two_src src;
    src.rT = rT, src.rB = rB, src.rS = rS, src.cb = cb, src.soff = 0;
    src.pT = a.To[0], src.pB = a.b[0], src.pS = a.So[0], src.ncomp = a.ncomp, src.zoff = 0;
    src.plane_bytes = 4u * (unsigned)a.sz;
    auto issue = [&](int s, int z) { /* stage s, plane z = plane_of(s) */
        src.soff = 4u * (unsigned)((int64_t)z * a.sz);
        src.zoff = (int64_t)z * a.sz;
        const bool split = SPLIT && (tile_pml || zpml(s));
        float *slot = lds + (s % NSLOT) * Ly::SLOT;
        if constexpr (Ly::NW == 4) {
            switch (wv) {
            case 0: two_issue<FX, FY, SPLIT, WIDE, BIG, 0>(slot, src, split, vq, vr, vlast); break;
            case 1: two_issue<FX, FY, SPLIT, WIDE, BIG, 1>(slot, src, split, vq, vr, vlast); break;
            case 2: two_issue<FX, FY, SPLIT, WIDE, BIG, 2>(slot, src, split, vq, vr, vlast); break;
            default: two_issue<FX, FY, SPLIT, WIDE, BIG, 3>(slot, src, split, vq, vr, vlast); break;
            }
        } else {
            switch (wv) {
            case 0: two_issue<FX, FY, SPLIT, WIDE, BIG, 0>(slot, src, split, vq, vr, vlast); break;
            case 1: two_issue<FX, FY, SPLIT, WIDE, BIG, 1>(slot, src, split, vq, vr, vlast); break;
            case 2: two_issue<FX, FY, SPLIT, WIDE, BIG, 2>(slot, src, split, vq, vr, vlast); break;
            case 3: two_issue<FX, FY, SPLIT, WIDE, BIG, 3>(slot, src, split, vq, vr, vlast); break;
            case 4: two_issue<FX, FY, SPLIT, WIDE, BIG, 4>(slot, src, split, vq, vr, vlast); break;
            case 5: two_issue<FX, FY, SPLIT, WIDE, BIG, 5>(slot, src, split, vq, vr, vlast); break;
            case 6: two_issue<FX, FY, SPLIT, WIDE, BIG, 6>(slot, src, split, vq, vr, vlast); break;
            default: two_issue<FX, FY, SPLIT, WIDE, BIG, 7>(slot, src, split, vq, vr, vlast); break;
            }
        }
    };
    /* stage arrays: 0 = E0, 1 = H0, 2 = bH, 3 = bE, 4 = SH0, 5 = SE0 */
    auto st = [&](const float *S, int arr, int c, int e) {
        return S[arr * Ly::ARR + c * Ly::NC * Ly::CH + e - Ly::row0(3 * arr + c) * W4];
    };
    /* PML factors of cell r on the plane of stage s: H (integer samples) or E (half) */
    auto factors = [&](int r, int s, bool e_field, float pa[3], float pg[3]) {
        const float *tt = e_field ? tE[r] : tH[r];
        pa[0] = tt[0], pg[0] = tt[1], pa[1] = tt[2], pg[1] = tt[3];
        pa[2] = SPLIT ? lds[L_ZT + 5 * s + (e_field ? 2 : 0)] : 1.f;
        pg[2] = SPLIT ? lds[L_ZT + 5 * s + (e_field ? 3 : 1)] : 1.f;
    };

    for (int s = 0; s < D && s < ns; s++) issue(s, plane_of(s));
    int zi = plane_of(D); /* plane of the next stage to issue */
    /* carried per cell: E0, SE0, bE of the previous plane; bH, H1, SH1 of the previous
     * plane; E1, SE1, bE two planes back */
    float e0p[CELLS][3], se0p[CELLS][3], bep[CELLS][3], bhp[CELLS][3], h1p[CELLS][3], sh1p[CELLS][3];
    float e1pp[CELLS][3], se1pp[CELLS][3], bepp[CELLS][3];
#pragma unroll
    for (int r = 0; r < CELLS; r++)
#pragma unroll
        for (int c = 0; c < 3; c++)
            e0p[r][c] = se0p[r][c] = bep[r][c] = bhp[r][c] = h1p[r][c] = sh1p[r][c] = e1pp[r][c] = se1pp[r][c] =
                bepp[r][c] = 0.f;

    int zq0 = plane_of(0), zq1 = plane_of(-1), zq2 = plane_of(-2); /* planes of stages s, s-1, s-2 */
    int dm_s1 = -1, dm_s2 = -1; /* DFT ring planes of stages s-1, s-2 (read in earlier iterations) */
    for (int s = 0; s < ns; s++) {
        const int p = z0 - 2 + s;     /* unwrapped plane of this stage */
        const int zp = zq0, zp1 = zq1, zp2 = zq2; /* wrapped planes of stages s, s-1, s-2 */
        const bool doA = s >= 1, doB = p >= z0, doC = p >= z0 + 1, doD = p >= z0 + 2;
        if (s + D < ns) {
            issue(s + D, zi);
            zi = zi + 1 == a.L2 ? 0 : zi + 1;
            wait_vm_barrier<D * Ly::kMinOps>();
        } else {
            switch (ns - 1 - s) { /* stages already issued after s */
            case 0: wait_vm_barrier<0>(); break;
            case 1: wait_vm_barrier<Ly::kMinOps>(); break;
            default: wait_vm_barrier<(D > 2 ? 2 : 0) * Ly::kMinOps>(); break;
            }
        }
        const float *S = lds + (s % NSLOT) * Ly::SLOT;
        const bool zs = zpml(s), zs1 = s >= 1 && zpml(s - 1), zs2 = s >= 2 && zpml(s - 2);
        const int dm1 = __builtin_amdgcn_readfirstlane(dm_s1), dm2 = __builtin_amdgcn_readfirstlane(dm_s2);
        const int dm0 = dft_m(s); /* this stage's, used in the next two iterations */

        float h1n[CELLS][3], sh1n[CELLS][3], e1n[CELLS][3], se1n[CELLS][3], h2n[CELLS][3];
#pragma unroll
        for (int r = 0; r < CELLS; r++)
#pragma unroll
            for (int c = 0; c < 3; c++) h1n[r][c] = sh1n[r][c] = e1n[r][c] = se1n[r][c] = h2n[r][c] = 0.f;
        /* a. H1(p) on G3 */
#pragma unroll
        for (int r = 0; r < CELLS; r++) {
            /* the ring cell runs on every lane (no divergent branch, so the two cells'
             * updates interleave); a lane outside the update's region writes nothing */
            if (!doA) continue;
            const int e = ce[r];
            const bool pml = pxy[r] || zs;
            float P[3], M[3], ec[3], exm[3], eym[3], h[3], b[3], sp[3];
#pragma unroll
            for (int c = 0; c < 3; c++) {
                ec[c] = st(S, 0, c, e);
                exm[c] = st(S, 0, c, e - 1);
                eym[c] = st(S, 0, c, e - W4);
                h[c] = st(S, 1, c, e);
                b[c] = st(S, 2, c, e);
                sp[c] = SPLIT && pml ? st(S, 4, c, e) : 0.f;
            }
            const float wx = cx[r] == 0 ? a.w[0] : 1.f, wy = cy[r] == 0 ? a.w[1] : 1.f, wz = zp == 0 ? a.w[2] : 1.f;
            P[0] = ec[2] - wy * eym[2];
            M[0] = -ec[1] + wz * e0p[r][1];
            P[1] = ec[0] - wz * e0p[r][0];
            M[1] = -ec[2] + wx * exm[2];
            P[2] = ec[1] - wx * exm[1];
            M[2] = -ec[0] + wy * eym[0];
            if (!pml && zp == a.iI) {
                P[0] += a.sgn * (float)plane_source(nullptr, Einc, sc[0], sc[1], sc[2], cq[r], 1);
                P[1] -= a.sgn * (float)plane_source(nullptr, Einc, sc[0], sc[1], sc[2], cq[r], 0);
                retire_loads();
            }
            float pa[3], pg[3];
            factors(r, s, false, pa, pg);
            h_regular(h, b, sp, P, M, pml, pa, pg, h1n[r], sh1n[r]);
        }
        /* b. E1(p-1) on G2: forward curl of H1 (previous plane from LDS, +z from a.) */
#pragma unroll
        for (int r = 0; r < CELLS; r++) {
            if (!doB) continue;
            const int e = ce[r];
            const bool pml = pxy[r] || zs1;
            float P[3], M[3], hc[3], hxp[3], hyp[3];
#pragma unroll
            for (int c = 0; c < 3; c++) {
                hc[c] = lds[L_H1 + c * N4 + e];
                hxp[c] = lds[L_H1 + c * N4 + e + 1];
                hyp[c] = lds[L_H1 + c * N4 + e + W4];
            }
            const float wx = cx[r] + 1 == a.L0 ? a.w[0] : 1.f, wy = cy[r] + 1 == a.L1 ? a.w[1] : 1.f;
            const float wz = zp1 + 1 == a.L2 ? a.w[2] : 1.f;
            P[0] = wy * hyp[2] - hc[2];
            M[0] = -wz * h1n[r][1] + hc[1];
            P[1] = wz * h1n[r][0] - hc[0];
            M[1] = -wx * hxp[2] + hc[2];
            P[2] = wx * hxp[1] - hc[1];
            M[2] = -wy * hyp[0] + hc[0];
            if (!pml && zp1 == a.iI) {
                P[0] += a.sgn * (float)plane_source(nullptr, Hinc, sc[3], sc[4], sc[5], cq[r], 1);
                P[1] -= a.sgn * (float)plane_source(nullptr, Hinc, sc[3], sc[4], sc[5], cq[r], 0);
                retire_loads();
            }
            float pa[3], pg[3];
            factors(r, s - 1, true, pa, pg);
            e_regular(e0p[r], bep[r], se0p[r], P, M, pml, pa, pg, e1n[r], se1n[r]);
            if (!pml && inT[r] && p - 1 >= z0 && p - 1 < z1 && (zp1 == a.iB || zp1 == a.iT)) {
                /* first step's monitors (main = total, S = 0): atomic adds that return
                 * nothing, so no load has to be retired behind the copies in flight; one
                 * lane per cell, in step order */
                for (int m = 0; m < 2; m++) {
                    if (zp1 != (m == 0 ? a.iB : a.iT)) continue;
                    double *HM = mon(2 * m), *EM = mon(2 * m + 1);
                    for (int v = 0; v < 2; v++) {
                        const int64_t o = 2 * (v + 2 * cq[r]);
                        unsafeAtomicAdd(HM + o, sc[6] * (double)hc[v]);
                        unsafeAtomicAdd(HM + o + 1, sc[7] * (double)hc[v]);
                        unsafeAtomicAdd(EM + o, sc[8] * (double)e1n[r][v]);
                        unsafeAtomicAdd(EM + o + 1, sc[9] * (double)e1n[r][v]);
                    }
                }
            }
            if (DFT && inT[r]) dft_sample(dm1, cq[r], 0u, hc[0], hc[1], e1n[r][0], e1n[r][1]);
            if (BOX && inT[r] && p - 1 >= z0 && p - 1 < z1) box_save(cx[r], cy[r], zp1, 0, e1n[r]);
            if (inG2[r]) {
                lds[L_E1 + 0 * N4 + e] = e1n[r][0];
                lds[L_E1 + 1 * N4 + e] = e1n[r][1];
                lds[L_E1 + 2 * N4 + e] = e1n[r][2];
            }
        }
        wait_vm_barrier<63>(); /* E1(p-1) of the neighbours in LDS */
        /* c. H2(p-1) on G1 */
#pragma unroll
        for (int r = 0; r < CELLS; r++) {
            if (!doC) continue;
            const int e = ce[r];
            const bool pml = pxy[r] || zs1;
            float P[3], M[3], ec[3], exm[3], eym[3], sh2[3];
#pragma unroll
            for (int c = 0; c < 3; c++) {
                ec[c] = lds[L_E1 + c * N4 + e];
                exm[c] = lds[L_E1 + c * N4 + e - 1];
                eym[c] = lds[L_E1 + c * N4 + e - W4];
            }
            const float wx = cx[r] == 0 ? a.w[0] : 1.f, wy = cy[r] == 0 ? a.w[1] : 1.f, wz = zp1 == 0 ? a.w[2] : 1.f;
            P[0] = ec[2] - wy * eym[2];
            M[0] = -ec[1] + wz * e1pp[r][1];
            P[1] = ec[0] - wz * e1pp[r][0];
            M[1] = -ec[2] + wx * exm[2];
            P[2] = ec[1] - wx * exm[1];
            M[2] = -ec[0] + wy * eym[0];
            if (!pml && zp1 == a.iI) {
                P[0] += a.sgn * (float)plane_source(nullptr, Einc, sc[10], sc[11], sc[12], cq[r], 1);
                P[1] -= a.sgn * (float)plane_source(nullptr, Einc, sc[10], sc[11], sc[12], cq[r], 0);
                retire_loads();
            }
            float pa[3], pg[3];
            factors(r, s - 1, false, pa, pg);
            h_regular(h1p[r], bhp[r], sh1p[r], P, M, pml, pa, pg, h2n[r], sh2);
            if (inT[r] && p - 1 < z1) {
                store(false, 0, cq[r] + a.sz * zp1, h2n[r]);
                if (SPLIT && pml) store(true, 0, cq[r] + a.sz * zp1, sh2);
            }
        }
        /* d. E2(p-2) on T: forward curl of H2 (previous plane from LDS, +z from c.) */
#pragma unroll
        for (int r = 0; r < CELLS; r++) {
            if (!(doD && inT[r])) continue;
            const int e = ce[r];
            const bool pml = pxy[r] || zs2;
            float P[3], M[3], hc[3], hxp[3], hyp[3], en[3], sn[3];
#pragma unroll
            for (int c = 0; c < 3; c++) {
                hc[c] = lds[L_H2 + c * N4 + e];
                hxp[c] = lds[L_H2 + c * N4 + e + 1];
                hyp[c] = lds[L_H2 + c * N4 + e + W4];
            }
            const float wx = cx[r] + 1 == a.L0 ? a.w[0] : 1.f, wy = cy[r] + 1 == a.L1 ? a.w[1] : 1.f;
            const float wz = zp2 + 1 == a.L2 ? a.w[2] : 1.f;
            P[0] = wy * hyp[2] - hc[2];
            M[0] = -wz * h2n[r][1] + hc[1];
            P[1] = wz * h2n[r][0] - hc[0];
            M[1] = -wx * hxp[2] + hc[2];
            P[2] = wx * hxp[1] - hc[1];
            M[2] = -wy * hyp[0] + hc[0];
            if (!pml && zp2 == a.iI) {
                P[0] += a.sgn * (float)plane_source(nullptr, Hinc, sc[13], sc[14], sc[15], cq[r], 1);
                P[1] -= a.sgn * (float)plane_source(nullptr, Hinc, sc[13], sc[14], sc[15], cq[r], 0);
                retire_loads();
            }
            float pa[3], pg[3];
            factors(r, s - 2, true, pa, pg);
            e_regular(e1pp[r], bepp[r], se1pp[r], P, M, pml, pa, pg, en, sn);
            store(false, 3, cq[r] + a.sz * zp2, en);
            if (SPLIT && pml) store(true, 3, cq[r] + a.sz * zp2, sn);
            if (DFT) dft_sample(dm2, cq[r], dstep, hc[0], hc[1], en[0], en[1]);
            if (BOX) box_save(cx[r], cy[r], zp2, 1, en);
            if (!pml && (zp2 == a.iB || zp2 == a.iT)) { /* second step's monitors */
                for (int m = 0; m < 2; m++) {
                    if (zp2 != (m == 0 ? a.iB : a.iT)) continue;
                    double *HM = mon(2 * m), *EM = mon(2 * m + 1);
                    for (int v = 0; v < 2; v++) {
                        const int64_t o = 2 * (v + 2 * cq[r]);
                        unsafeAtomicAdd(HM + o, sc[16] * (double)hc[v]);
                        unsafeAtomicAdd(HM + o + 1, sc[17] * (double)hc[v]);
                        unsafeAtomicAdd(EM + o, sc[18] * (double)en[v]);
                        unsafeAtomicAdd(EM + o + 1, sc[19] * (double)en[v]);
                    }
                }
            }
        }
        /* carry to the next plane */
        dm_s2 = dm_s1;
        dm_s1 = dm0;
        const bool split_s = SPLIT && (tile_pml || zs);
#pragma unroll
        for (int r = 0; r < CELLS; r++)
#pragma unroll
            for (int c = 0; c < 3; c++) {
                e1pp[r][c] = e1n[r][c];
                se1pp[r][c] = se1n[r][c];
                bepp[r][c] = bep[r][c];
                h1p[r][c] = h1n[r][c];
                sh1p[r][c] = sh1n[r][c];
                bhp[r][c] = st(S, 2, c, ce[r]);
                bep[r][c] = st(S, 3, c, ce[r]);
                e0p[r][c] = st(S, 0, c, ce[r]);
                se0p[r][c] = split_s ? st(S, 5, c, ce[r]) : 0.f;
            }
        zq2 = zq1;
        zq1 = zq0;
        zq0 = zq0 + 1 == a.L2 ? 0 : zq0 + 1;
        /* every wave is done with this stage and with H1, H2 of the previous plane */
        wait_vm_barrier<63>();
#pragma unroll
        for (int r = 0; r < CELLS; r++) {
            if (doA && inG3[r])
#pragma unroll
                for (int c = 0; c < 3; c++) lds[L_H1 + c * N4 + ce[r]] = h1n[r][c];
            if (doC && inG1[r])
#pragma unroll
                for (int c = 0; c < 3; c++) lds[L_H2 + c * N4 + ce[r]] = h2n[r][c];
        }
    }
}

template <int FX, int FY, int D, bool SPLIT, bool WIDE, bool DFT, bool BIG = false, bool BOX = false>
__global__ __launch_bounds__(FX * FY) void two_step_f32(fused_args a)
{
    const int h = xcd_logical(blockIdx.x, gridDim.x);
    two_step_body<FX, FY, D, SPLIT, WIDE, DFT, BIG, BOX>(a, a.blist ? a.blist[h] : h, nullptr);
}

/* A two-step pass of several plans in one launch (a batch of independent runs): block i
 * (XCD-aware order) takes entry blk[i] = plan << 20 | the block's id within its plan, the
 * plan's arguments from args[plan] (constant memory, read through the scalar cache) and
 * its 20 step scalars and (DFT) the ring slot of the pass's first step from
 * scal[kMultiScalars * plan].  The split-slot variant, which updates every cell of a grid
 * with or without PML; with DFT, plans whose dft_n is 0 take no samples. */
typedef const fused_args __attribute__((address_space(4))) const_fused_args;

template <bool DFT>
__global__ __launch_bounds__(256) void two_step_multi(const fused_args *args, const int *blk, const double *scal)
{
    const int e = blk[xcd_logical(blockIdx.x, gridDim.x)];
    const int plan = e >> 20, id = e & ((1 << 20) - 1);
    const const_fused_args &a = ((const_fused_args *)args)[plan];
    two_step_body<32, 8, 1, true, false, DFT, false, false>(a, id, scal + SMX_MULTI_SCALARS * plan);
}

/* ---------------------------------------------------------------------------------- */
/* Region sweeps: the cells outside the staged kernel's box                           */
/* ---------------------------------------------------------------------------------- */
/* H of the lossy/metal cells, before the fused pass (which reads the result for its
 * E update and leaves these cells' H alone). */
__device__ void special_h_one(const fused_args &a, int i)
{
    const int64_t l = a.special_cells[i];
    const int x = (int)(l % a.L0), y = (int)((l / a.L0) % a.L1), z = (int)(l / a.sz);
    float Hn[3], Sn[3];
    fused_special_h(a, x, y, z, l, a.special[i], Hn, Sn);
    for (int c = 0; c < 3; c++) {
        a.Tn[c][l] = Hn[c];
        a.Sn[c][l] = Sn[c];
    }
}

/* E of the lossy/metal cells after the fused pass (the class branches of F:291-314),
 * from the new H in buffer 1 - cur; also their monitor samples. */
__device__ void special_e_one(const fused_args &a, int i)
{
    const int64_t l = a.special_cells[i];
    const smx_special_f32 &sp = a.special[i];
    const int x = (int)(l % a.L0), y = (int)((l / a.L0) % a.L1), z = (int)(l / a.sz);
    float *const *H = a.Tn;
    float *const *Hs = a.Sn;
    float En[3], Sn[3];
    for (int c = 0; c < 3; c++) {
        En[c] = a.To[3 + c][l];
        Sn[c] = a.So[3 + c][l];
    }
    const int k = sp.cls;
    float P[3], M[3];
    if (k == 'n' || k == 'a') {
        curl_halves<true>(H, nullptr, nbrs_at(a, x, y, z, true), l, P, M);
        if (k == 'n') {
            float W[3] = {P[0] + M[0], P[1] + M[1], P[2] + M[2]};
            if (z == a.iI) {
                const int64_t q = x + (int64_t)a.L0 * y;
                W[0] += a.sgn * (float)plane_source(a.srcH_plane, a.Hinc, a.srcH_f2, a.srcH_re, a.srcH_im, q, 1);
                W[1] -= a.sgn * (float)plane_source(a.srcH_plane, a.Hinc, a.srcH_f2, a.srcH_re, a.srcH_im, q, 0);
            }
            for (int c = 0; c < 3; c++) En[c] = sp.eA[c] * En[c] + sp.eB[c] * W[c];
        } else {
            for (int c = 0; c < 3; c++) {
                const float s = Sn[c], m = En[c] - s;
                const float s2 = sp.eAs[c] * s + sp.eBs[c] * M[c];
                En[c] = sp.eA[c] * m + sp.eB[c] * P[c] + s2;
                Sn[c] = s2;
            }
        }
    } else if (k != 'h') { /* 'e', 'b': backward curl of the new H main slots (F:310) */
        curl_halves<false>(H, Hs, nbrs_at(a, x, y, z, false), l, P, M);
        for (int c = 0; c < 3; c++) {
            const float s = Sn[c];
            float m = En[c] - s + 0.5f * s;
            const float s2 = sp.eAs[c] * s + sp.eBs[c] * (P[c] + M[c]);
            m += 0.5f * s2;
            En[c] = m + s2;
            Sn[c] = s2;
        }
    }
    for (int c = 0; c < 3; c++) {
        a.Tn[3 + c][l] = En[c];
        a.Sn[3 + c][l] = Sn[c];
    }
    if (z == a.iB || z == a.iT) {
        const int64_t q = x + (int64_t)a.L0 * y;
        for (int m = 0; m < 2; m++) {
            if (z != (m == 0 ? a.iB : a.iT)) continue;
            double *HM = m == 0 ? a.HB : a.HT, *EM = m == 0 ? a.EB : a.ET;
            for (int v = 0; v < 2; v++) {
                const double hv = (double)H[v][l] - (double)Hs[v][l];
                const double ev = (double)En[v] - (double)Sn[v];
                const int64_t o = 2 * (v + 2 * q);
                HM[o] += a.monH_re * hv;
                HM[o + 1] += a.monH_im * hv;
                EM[o] += a.monE_re * ev;
                EM[o + 1] += a.monE_im * ev;
            }
        }
    }
}

__global__ __launch_bounds__(kThreads) void special_h_pass(fused_args a)
{
    const int i = blockIdx.x * kThreads + threadIdx.x;
    if (i < a.nspecial) special_h_one(a, i);
}

__global__ __launch_bounds__(kThreads) void special_e_pass(fused_args a)
{
    const int i = blockIdx.x * kThreads + threadIdx.x;
    if (i < a.nspecial) special_e_one(a, i);
}

/*
 * On a grid with PML the staged kernel without split slots covers only the box of
 * cells whose +x, +y, +z neighbours hold no PML cell either (its recomputed halo then
 * needs no split value).  These two kernels update every other cell from the old
 * buffer into the new one: region_h_f32 their H (independent of the staged kernel),
 * then, after both, region_e_f32 their E from the new H of all cells.  Lossy/metal
 * cells are left to the special passes.  The thread index runs x fastest through each
 * of up to 6 boxes in turn.
 */
__device__ inline bool region_cell(const fused_args &a, int64_t g, int &x, int &y, int &z)
{
    if (g >= a.reg_start[a.nreg]) return false;
    int r = 0;
    while (r + 1 < a.nreg && g >= a.reg_start[r + 1]) r++;
    const int *b = a.reg[r];
    /* 32-bit arithmetic: a box holds far fewer than 2^32 cells */
    const unsigned o = (unsigned)(g - a.reg_start[r]);
    const unsigned wx = (unsigned)(b[1] - b[0]), wy = (unsigned)(b[3] - b[2]);
    const unsigned row = o / wx;
    x = b[0] + (int)(o - row * wx);
    y = b[2] + (int)(row % wy);
    z = b[4] + (int)(row / wy);
    return true;
}

/* PML factors of the cell for H (half = 0: integer samples) or E (half = 1). */
__device__ inline void cell_pml(const fused_args &a, int x, int y, int z, int half, float pa[3], float pg[3])
{
    const int jj[3] = {x, y, z};
#pragma unroll
    for (int d = 0; d < 3; d++) {
        const float *t = axis_tab(a, d);
        const int L = axis_len(a, d);
        pa[d] = t[2 * half * L + jj[d]];
        pg[d] = t[(2 * half + 1) * L + jj[d]];
    }
}

__global__ __launch_bounds__(kThreads) void region_h_f32(fused_args a)
{
    int x, y, z;
    const int64_t g = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    if (!region_cell(a, g, x, y, z)) { /* with reg_specials: the special cells follow the boxes */
        const int64_t i = g - a.reg_start[a.nreg];
        if (a.reg_specials && i < a.nspecial) special_h_one(a, (int)i);
        return;
    }
    const int64_t q = x + (int64_t)a.L0 * y, l = q + a.sz * z;
    const float b0 = a.b[0][l];
    if (__builtin_isnan(b0)) return; /* special_h_pass */
    const float b[3] = {b0, a.b[1][l], a.b[2][l]};
    const nb3 n = nbrs_at(a, x, y, z, false);
    float ec[3], exm[3], eym[3], ezm[3], h[3], sp[3];
    const bool pml = a.pmlflag[x] | a.pmlflag[a.L0 + y] | a.pmlflag[a.L0 + a.L1 + z];
#pragma unroll
    for (int c = 0; c < 3; c++) {
        ec[c] = a.To[3 + c][l];
        exm[c] = a.To[3 + c][l + n.o[0]];
        eym[c] = a.To[3 + c][l + n.o[1]];
        ezm[c] = a.To[3 + c][l + n.o[2]];
        h[c] = a.To[c][l];
        sp[c] = pml ? a.So[c][l] : 0.f;
    }
    const float wx = n.w[0], wy = n.w[1], wz = n.w[2];
    float P[3], M[3];
    P[0] = ec[2] - wy * eym[2];
    M[0] = -ec[1] + wz * ezm[1];
    P[1] = ec[0] - wz * ezm[0];
    M[1] = -ec[2] + wx * exm[2];
    P[2] = ec[1] - wx * exm[1];
    M[2] = -ec[0] + wy * eym[0];
    if (!pml && z == a.iI) {
        P[0] += a.sgn * (float)plane_source(a.srcE_plane, a.Einc, a.srcE_f2, a.srcE_re, a.srcE_im, q, 1);
        P[1] -= a.sgn * (float)plane_source(a.srcE_plane, a.Einc, a.srcE_f2, a.srcE_re, a.srcE_im, q, 0);
    }
    float pa[3], pg[3], Hn[3], Sn[3];
    cell_pml(a, x, y, z, 0, pa, pg);
    h_regular(h, b, sp, P, M, pml, pa, pg, Hn, Sn);
    store3(a.Tn, l, Hn);
    if (pml) store3(a.Sn, l, Sn);
}

__global__ __launch_bounds__(kThreads) void region_e_f32(fused_args a)
{
    int x, y, z;
    const int64_t g = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    if (!region_cell(a, g, x, y, z)) {
        const int64_t i = g - a.reg_start[a.nreg];
        if (a.reg_specials && i < a.nspecial) special_e_one(a, (int)i);
        return;
    }
    const int64_t q = x + (int64_t)a.L0 * y, l = q + a.sz * z;
    const float b0 = a.b[3][l];
    if (__builtin_isnan(b0)) return; /* special_e_pass */
    const float b[3] = {b0, a.b[4][l], a.b[5][l]};
    const nb3 n = nbrs_at(a, x, y, z, true);
    float hc[3], hxp[3], hyp[3], hzp[3], e[3], se[3];
    const bool pml = a.pmlflag[x] | a.pmlflag[a.L0 + y] | a.pmlflag[a.L0 + a.L1 + z];
#pragma unroll
    for (int c = 0; c < 3; c++) {
        hc[c] = a.Tn[c][l];
        hxp[c] = a.Tn[c][l + n.o[0]];
        hyp[c] = a.Tn[c][l + n.o[1]];
        hzp[c] = a.Tn[c][l + n.o[2]];
        e[c] = a.To[3 + c][l];
        se[c] = pml ? a.So[3 + c][l] : 0.f;
    }
    const float wx = n.w[0], wy = n.w[1], wz = n.w[2];
    float P[3], M[3];
    P[0] = wy * hyp[2] - hc[2];
    M[0] = -wz * hzp[1] + hc[1];
    P[1] = wz * hzp[0] - hc[0];
    M[1] = -wx * hxp[2] + hc[2];
    P[2] = wx * hxp[1] - hc[1];
    M[2] = -wy * hyp[0] + hc[0];
    float En[3];
    if (!pml) {
        if (z == a.iI) {
            P[0] += a.sgn * (float)plane_source(a.srcH_plane, a.Hinc, a.srcH_f2, a.srcH_re, a.srcH_im, q, 1);
            P[1] -= a.sgn * (float)plane_source(a.srcH_plane, a.Hinc, a.srcH_f2, a.srcH_re, a.srcH_im, q, 0);
        }
#pragma unroll
        for (int c = 0; c < 3; c++) En[c] = e[c] + b[c] * (P[c] + M[c]);
        store3(a.Tn + 3, l, En);
        auto in_mod = [](int v, int lo, int len, int L) { return ((v - lo) % L + L) % L < len; };
        const bool skip = a.monin_on ? !(in_mod(x, a.monin[0], a.monin[1], a.L0) && in_mod(y, a.monin[2], a.monin[3], a.L1) &&
                                         in_mod(z, a.monin[4], a.monin[5], a.L2))
                                     : x >= a.monskip[0] && x < a.monskip[1] && y >= a.monskip[2] && y < a.monskip[3] &&
                                           z >= a.monskip[4] && z < a.monskip[5];
        for (int m = 0; m < 2 && !skip; m++) { /* monitors: main = total here (S = 0) */
            if (z != (m == 0 ? a.iB : a.iT)) continue;
            double *HM = m == 0 ? a.HB : a.HT, *EM = m == 0 ? a.EB : a.ET;
            for (int v = 0; v < 2; v++) {
                const int64_t o = 2 * (v + 2 * q);
                HM[o] += a.monH_re * (double)hc[v];
                HM[o + 1] += a.monH_im * (double)hc[v];
                EM[o] += a.monE_re * (double)En[v];
                EM[o + 1] += a.monE_im * (double)En[v];
            }
        }
        return;
    }
    float pa[3], pg[3], Sn[3];
    cell_pml(a, x, y, z, 1, pa, pg);
#pragma unroll
    for (int c = 0; c < 3; c++) {
        const int dm = c == 2 ? 0 : c + 1, ds = c == 0 ? 2 : c - 1;
        const float s2 = pa[ds] * se[c] + b[c] * pg[ds] * M[c];
        En[c] = pa[dm] * (e[c] - se[c]) + b[c] * pg[dm] * P[c] + s2;
        Sn[c] = s2;
    }
    store3(a.Tn + 3, l, En);
    store3(a.Sn + 3, l, Sn);
}

} // namespace

/* The box of cells whose +x/+y/+z neighbours are PML-free too, and boxes covering the
 * rest of the grid (see smx_f32_state). */
/* Boxes covering the grid outside `inner`; false when a box holds 2^32 cells or more
 * (region_cell decodes with 32-bit offsets). */
static bool complement(const int L[3], const int in[6], smx_region_set *rs)
{
    rs->n = 0;
    auto add = [&](int x0, int x1, int y0, int y1, int z0, int z1) {
        if (x1 <= x0 || y1 <= y0 || z1 <= z0) return;
        int *b = rs->box[rs->n++];
        b[0] = x0, b[1] = x1, b[2] = y0, b[3] = y1, b[4] = z0, b[5] = z1;
    };
    add(0, L[0], 0, L[1], 0, in[4]);
    add(0, L[0], 0, L[1], in[5], L[2]);
    add(0, L[0], 0, in[2], in[4], in[5]);
    add(0, L[0], in[3], L[1], in[4], in[5]);
    add(0, in[0], in[2], in[3], in[4], in[5]);
    add(in[1], L[0], in[2], in[3], in[4], in[5]);
    rs->start[0] = 0;
    for (int r = 0; r < rs->n; r++) {
        const int *b = rs->box[r];
        const int64_t cells = (int64_t)(b[1] - b[0]) * (b[3] - b[2]) * (b[5] - b[4]);
        if (cells >= ((int64_t)1 << 32)) return false;
        rs->start[r + 1] = rs->start[r] + cells;
    }
    return true;
}

/* The PML-free run [lo, hi) of every axis (lo = 0, hi = L and wrapping halos where an
 * axis has no PML); false when the PML-free cells of an axis are not one run. */
static bool pml_free_runs(const smx_f32_state *st, const unsigned char *flag, int lo[3], int hi[3], bool has[3])
{
    const unsigned char *f = flag;
    for (int d = 0; d < 3; d++) {
        const int L = st->L[d];
        int first = -1, last = -1, count = 0;
        for (int j = 0; j < L; j++)
            if (!f[j]) {
                if (first < 0) first = j;
                last = j;
                count++;
            }
        f += L;
        has[d] = count != L;
        if (!has[d]) {
            lo[d] = 0;
            hi[d] = L;
            continue;
        }
        if (count == 0 || last - first + 1 != count) return false;
        lo[d] = first;
        hi[d] = last + 1;
    }
    return true;
}

/* The boxes of the staged kernel without split slots (one step) and of the two-step
 * kernel on a grid with PML (see smx_f32_state). */
static void plan_split(smx_f32_state *st, const unsigned char *flag)
{
    const int L[3] = {st->L[0], st->L[1], st->L[2]};
    int lo[3], hi[3];
    bool has[3];
    st->split_ok = st->split2_ok = 0;
    st->nreg = 0;
    if (!pml_free_runs(st, flag, lo, hi, has)) return;
    /* one step: cells whose +1 neighbours are PML-free, i.e. [lo, hi-1) on a PML axis */
    int in1[6];
    for (int d = 0; d < 3; d++) {
        in1[2 * d] = lo[d];
        in1[2 * d + 1] = has[d] ? hi[d] - 1 : hi[d];
    }
    bool ok = true;
    for (int d = 0; d < 3; d++) ok &= in1[2 * d + 1] > in1[2 * d];
    smx_region_set rs;
    if (ok && complement(L, in1, &rs)) {
        for (int i = 0; i < 6; i++) st->fbox[i] = in1[i];
        st->nreg = rs.n;
        std::memcpy(st->reg, rs.box, sizeof st->reg);
        std::memcpy(st->reg_start, rs.start, sizeof st->reg_start);
        st->split_ok = 1;
    }
    /* two steps: outputs on [a, b) need H1 on [a-1, b+1] PML-free: a = lo+1, b = hi-2;
     * the x start rounded up to a multiple of 4 (the 16-byte copies read 4-aligned x) */
    int f2[6], h1[6], e1[6];
    ok = true;
    for (int d = 0; d < 3; d++) {
        int a = has[d] ? lo[d] + 1 : 0;
        const int b = has[d] ? hi[d] - 2 : L[d];
        if (d == 0) a = (a + 3) / 4 * 4;
        f2[2 * d] = a;
        f2[2 * d + 1] = b;
        /* first step outside the box shrunk by one on both sides (H1), on + sides (E1) */
        h1[2 * d] = has[d] ? a + 1 : 0;
        h1[2 * d + 1] = has[d] ? b - 1 : L[d];
        e1[2 * d] = has[d] ? a : 0;
        e1[2 * d + 1] = has[d] ? b - 1 : L[d];
        ok &= b - a >= 3;
    }
    if (ok && complement(L, f2, &st->regP2) && complement(L, h1, &st->regH1) && complement(L, e1, &st->regE1)) {
        for (int i = 0; i < 6; i++) st->fbox2[i] = f2[i];
        st->split2_ok = 1;
    }
}

const char *smx_f32_last_error(void) { return f32_err; }

int smx_f32_num_tiles(void) { return kNumTiles; }

/* two-step variants: 32x8 tiles (one thread per cell), d planes of copies in flight,
 * split slots (PML cells inside the kernel) or not (PML only outside the kernel's box,
 * region sweeps for the rest), 4- or 16-byte copies (wide: x extent and box start
 * multiples of 4) */
struct two_variant {
    int fx, fy, d;
    bool split, wide;
    int ox; /* columns of the stage region left of a tile */
    const char *name, *with_regions, *name_dft; /* as profilers list them; the last with the DFT */
};
#define SMX_TWO(FX, FY, D, SP, WI)                                                                         \
    {FX, FY, D, SP, WI, two_layout<FX, FY, SP, WI>::OX,                                                   \
     "two_step_f32<" #FX "," #FY "," #D "," #SP "," #WI ",false,false>",                                 \
     "region_h_f32 + region_e_f32 + two_step_f32<" #FX "," #FY "," #D "," #SP "," #WI                     \
     ",false,false> + region_h_f32 + region_e_f32",                                                       \
     "two_step_f32<" #FX "," #FY "," #D "," #SP "," #WI ",true,false>"}
constexpr two_variant kTwo[] = {SMX_TWO(32, 8, 1, false, false), SMX_TWO(32, 8, 2, false, false),
                                SMX_TWO(32, 8, 1, true, false), SMX_TWO(32, 8, 1, false, true)};
constexpr int kMultiVariant = 2; /* the variant two_step_multi runs: <32, 8, 1, split, narrow> */
#undef SMX_TWO
constexpr int kNumTwo = sizeof kTwo / sizeof kTwo[0];

int smx_f32_num_two(void) { return kNumTwo; }

/* Whether the two-step pass runs in its 64-bit form (groups of 4 GiB and more, or forced). */
static bool two_big(const smx_f32_state *st) { return st->force_big || 24 * st->n >= ((int64_t)1 << 32); }

int smx_f32_two_full(const smx_f32_state *st, int variant)
{
    return variant >= 0 && variant < kNumTwo && (!st->has_pml || kTwo[variant].split);
}

/* the kernels of kTwo, in its order */
using two_kernel = void (*)(fused_args);
constexpr two_kernel kTwoKernel[] = {
    two_step_f32<32, 8, 1, false, false, false>, two_step_f32<32, 8, 2, false, false, false>,
    two_step_f32<32, 8, 1, true, false, false>, two_step_f32<32, 8, 1, false, true, false>};
static_assert(sizeof kTwoKernel / sizeof kTwoKernel[0] == kNumTwo, "one kernel per variant");
/* the same with the DFT samples of both steps (run while a DFT is set) */
constexpr two_kernel kTwoKernelDft[] = {
    two_step_f32<32, 8, 1, false, false, true>, two_step_f32<32, 8, 2, false, false, true>,
    two_step_f32<32, 8, 1, true, false, true>, two_step_f32<32, 8, 1, false, true, true>};
static_assert(sizeof kTwoKernelDft / sizeof kTwoKernelDft[0] == kNumTwo, "one kernel per variant");
/* the split-slot variant for groups of 4 GiB and more (24 n >= 2^32), without and with the DFT */
constexpr int kTwoBigVariant = 2;
/* the split-slot variant writing E of both steps on the design box (adjoint forward and
 * recomputation), without and with the DFT samples */
constexpr two_kernel kTwoKernelBox[2] = {two_step_f32<32, 8, 1, true, false, false, false, true>,
                                         two_step_f32<32, 8, 1, true, false, true, false, true>};
constexpr two_kernel kTwoKernelBig[2] = {two_step_f32<32, 8, 1, true, false, false, true>,
                                         two_step_f32<32, 8, 1, true, false, true, true>};

void smx_f32_two_shape(int variant, int *fx, int *fy)
{
    const bool ok = variant >= 0 && variant < kNumTwo;
    *fx = ok ? kTwo[variant].fx : 0;
    *fy = ok ? kTwo[variant].fy : 0;
}

int smx_f32_two_ok(const smx_f32_state *st, int variant, int fz)
{
    /* no lossy/metal cell; 32-bit offsets as the staged kernel; the variants without
     * split slots need the PML outside the kernel's box, those with them keep the
     * chunk's z tables in LDS; 16-byte copies need x groups of 4 that never straddle
     * the x wrap (the two-step box starts at a multiple of 4) */
    if (variant < 0 || variant >= kNumTwo || fz < 1) return 0;
    /* lossy/metal cells: the split-slot variant in its 32-bit form, the tiles near them
     * left to region sweeps (smx_f32_step2) */
    if (st->nspecial != 0 && (variant != kMultiVariant || two_big(st))) return 0;
    /* groups of 4 GiB and more: the split-slot variant's 64-bit form, one array below 4 GiB */
    if ((24 * st->n >= ((int64_t)1 << 32) || st->force_big) && (variant != kTwoBigVariant || 4 * st->n >= ((int64_t)1 << 32)))
        return 0;
    if (kTwo[variant].wide && st->L[0] % 4 != 0) return 0;
    if ((st->L[2] + fz - 1) / fz > kMaxZChunks) return 0;
    if (kTwo[variant].split) return fz <= two_layout<32, 8, true, false>::kMaxFZ;
    return !st->has_pml || st->split2_ok;
}

const char *smx_f32_kernel_name(const smx_f32_state *st, int dft)
{
    static const char *const fused[kNumTiles] = {
        "fused_step_f32<32,8>", "fused_step_f32<64,4>", "fused_step_f32<16,16>",
        "fused_dma_f32<32,8,2,true>", "fused_dma_f32<16,16,2,true>", "fused_dma_f32<32,8,3,false>",
        "fused_dma_f32<16,16,3,false>", "fused_dma_f32<32,8,4,false>", "fused_dma_f32<32,8,2,false>",
        "fused_dma_f32<32,8,1,false>"};
    static const char *const split[kNumTiles] = {
        nullptr, nullptr, nullptr, nullptr, nullptr,
        "region_h_f32 + fused_dma_f32<32,8,3,false> + region_e_f32",
        "region_h_f32 + fused_dma_f32<16,16,3,false> + region_e_f32",
        "region_h_f32 + fused_dma_f32<32,8,4,false> + region_e_f32",
        "region_h_f32 + fused_dma_f32<32,8,2,false> + region_e_f32",
        "region_h_f32 + fused_dma_f32<32,8,1,false> + region_e_f32"};
    if (st->two >= 0 && st->two < kNumTwo) {
        if (two_big(st))
            return dft ? "two_step_f32<32,8,1,true,false,true,true>" : "two_step_f32<32,8,1,true,false,false,true>";
        if (st->has_pml && !kTwo[st->two].split) return kTwo[st->two].with_regions;
        if (st->nspecial > 0)
            return dft ? "two_step_f32<32,8,1,true,false,true,false> + region_h_f32 + region_e_f32 + sample_planes_f32 "
                         "(the tiles around lossy/metal cells, and those cells)"
                       : "two_step_f32<32,8,1,true,false,false,false> + region_h_f32 + region_e_f32 (the tiles "
                         "around lossy/metal cells, and those cells)";
        return dft ? kTwo[st->two].name_dft : kTwo[st->two].name;
    }
    if (!st->fused) return st->bx == 64 ? "h_sweep_f32<64> + e_sweep_f32<64>" : "h_sweep_f32<32> + e_sweep_f32<32>";
    if (st->tile < 0 || st->tile >= kNumTiles) return "";
    return st->tile >= kFirstNoSplitTile && st->has_pml ? split[st->tile] : fused[st->tile];
}

void smx_f32_tile_shape(int tile, int *fx, int *fy)
{
    const bool ok = tile >= 0 && tile < kNumTiles;
    *fx = ok ? kTiles[tile].fx : 0;
    *fy = ok ? kTiles[tile].fy : 0;
}

int smx_f32_tile_ok(const smx_f32_state *st, int tile, int fz)
{
    if (tile < 0 || tile >= kNumTiles || fz < 1) return 0;
    if (tile < kFirstDmaTile) return 1;
    if (tile >= kFirstNoSplitTile && st->has_pml && !st->split_ok) return 0;
    /* the staged kernel keeps the chunk's z tables in LDS and addresses each group of
     * six component arrays with 32-bit byte offsets */
    return fz <= dma_layout<32, 8>::kMaxFZ && 24 * st->n < ((int64_t)1 << 32);
}

void smx_f32_free(smx_f32_state *st)
{
    for (int k = 0; k < 2; k++) {
        if (st->Tmem[k]) (void)hipFree(st->Tmem[k]);
        if (st->Smem[k]) (void)hipFree(st->Smem[k]);
        st->Tmem[k] = st->Smem[k] = nullptr;
    }
    if (st->bmem) (void)hipFree(st->bmem);
    st->bmem = nullptr;
    if (st->tmpT) (void)hipFree(st->tmpT);
    if (st->tmpS) (void)hipFree(st->tmpS);
    st->tmpT = st->tmpS = nullptr;
    for (int i = 0; i < 6; i++) {
        st->T[0][i] = st->T[1][i] = st->S[0][i] = st->S[1][i] = st->b[i] = nullptr;
    }
    if (st->pml) (void)hipFree(st->pml);
    if (st->pmlflag) (void)hipFree(st->pmlflag);
    delete[] st->zpml_host;
    st->zpml_host = nullptr;
    delete[] st->xypml_host;
    st->xypml_host = nullptr;
    if (st->special) (void)hipFree(st->special);
    if (st->special_cells) (void)hipFree(st->special_cells);
    delete[] st->special_host;
    st->special_host = nullptr;
    if (st->sp_blist) (void)hipFree(st->sp_blist);
    st->sp_blist = nullptr;
    st->sp_valid = 0;
    st->pml = nullptr;
    st->pmlflag = nullptr;
    st->special = nullptr;
    st->special_cells = nullptr;
    delete[] st->cont_host;
    delete[] st->addr_host;
    st->cont_host = nullptr;
    st->addr_host = nullptr;
}

int smx_f32_create(smx_f32_state *st, const smx_geom *g, double iR, double iI, const double *mu,
                   const double *eps, const double *sigma, const char *Cont, const int *addr,
                   hipStream_t stream)
{
    const int L[4] = {g->L0, g->L1, g->L2, 0};
    const int64_t L0 = g->L0, L1 = g->L1, L2 = g->L2, n = L0 * L1 * L2;
    std::memset(st, 0, sizeof *st);
    st->L[0] = g->L0;
    st->L[1] = g->L1;
    st->L[2] = g->L2;
    st->n = n;
    st->Lnotn = g->Lnotn;
    st->bx = (g->L0 % 64 == 0 || (g->L0 > 32 && g->L0 % 32 != 0)) ? 64 : 32;

    /* per-axis PML tables and flags */
    const int64_t off[3] = {0, 2 * L0, 2 * L0 + 2 * L1};
    const int Ld[3] = {g->L0, g->L1, g->L2};
    std::vector<float> pml(4 * (size_t)(L0 + L1 + L2));
    std::vector<unsigned char> flag((size_t)(L0 + L1 + L2));
    {
        size_t base = 0, fbase = 0;
        for (int d = 0; d < 3; d++) {
            for (int j = 0; j < Ld[d]; j++) {
                const double sh = sigma[off[d] + 2 * j], se = sigma[off[d] + 2 * j + 1];
                pml[base + j] = (float)((iR - iI * sh) / (iR + iI * sh));
                pml[base + Ld[d] + j] = (float)(iR / (iR + iI * sh));
                pml[base + 2 * Ld[d] + j] = (float)((iR - iI * se) / (iR + iI * se));
                pml[base + 3 * Ld[d] + j] = (float)(iR / (iR + iI * se));
                flag[fbase + j] = std::fabs(sh) > kTol || std::fabs(se) > kTol;
            }
            base += 4 * (size_t)Ld[d];
            fbase += Ld[d];
        }
    }

    /* coefficient words; lossy and metal cells go to the special table */
    std::vector<float> bw(6 * (size_t)n);
    std::vector<smx_special_f32> special;
    std::vector<int64_t> special_cells;
    for (int64_t l = 0; l < n; l++) {
        const int64_t x = l % L0, y = (l / L0) % L1, z = l / (L0 * L1);
        const bool is_a = flag[x] || flag[L0 + y] || flag[L0 + L1 + z];
        const char k = Cont[l];
        bool lossy = false;
        for (int c = 0; c < 3; c++) lossy |= mu[6 * l + 2 * c + 1] != 0. || eps[6 * l + 2 * c + 1] != 0.;
        const bool regular = !lossy && (k == 'n' || k == 'a');
        if (regular && (k == 'a') != is_a) {
            snprintf(f32_err, sizeof f32_err,
                     "cell %lld: class '%c' disagrees with the separable PML profile", (long long)l, k);
            return SMX_ERR_ARG;
        }
        if (regular) {
            for (int c = 0; c < 3; c++) {
                bw[(size_t)c * n + l] = (float)(1. / (iR * mu[6 * l + 2 * c]));
                bw[(size_t)(3 + c) * n + l] = (float)(1. / (iR * eps[6 * l + 2 * c]));
            }
            continue;
        }
        double sg[6], w[24] = {0};
        smx_cell_sigma(L, sigma, x, y, z, sg);
        smx_cell_coeffs(k, mu + 6 * l, eps + 6 * l, sg, iR, iI, w);
        smx_special_f32 sp;
        std::memset(&sp, 0, sizeof sp);
        sp.cls = k;
        for (int c = 0; c < 3; c++) {
            sp.hA[c] = (float)w[c];
            sp.hB[c] = (float)w[3 + c];
            sp.eA[c] = (float)w[6 + c];
            sp.eB[c] = (float)w[9 + c];
            sp.hAs[c] = (float)w[12 + c];
            sp.hBs[c] = (float)w[15 + c];
            sp.eAs[c] = (float)w[18 + c];
            sp.eBs[c] = (float)w[21 + c];
        }
        const uint32_t idx = (uint32_t)special.size();
        special.push_back(sp);
        special_cells.push_back(l);
        float fidx;
        std::memcpy(&fidx, &idx, sizeof fidx);
        const float nan = std::numeric_limits<float>::quiet_NaN();
        bw[0 * (size_t)n + l] = nan;
        bw[3 * (size_t)n + l] = nan;
        bw[1 * (size_t)n + l] = fidx;
        bw[4 * (size_t)n + l] = fidx;
        bw[2 * (size_t)n + l] = 0.f;
        bw[5 * (size_t)n + l] = 0.f;
    }
    st->nspecial = (int)special.size();
    if (!special_cells.empty()) {
        st->special_host = new int64_t[special_cells.size()];
        std::memcpy(st->special_host, special_cells.data(), sizeof(int64_t) * special_cells.size());
    }
    st->has_pml = 0;
    for (unsigned char f : flag) st->has_pml |= f;
    plan_split(st, flag.data());

    /* both buffers always exist, so the autotuner may switch between the fused step
     * (buffer cur -> 1 - cur) and the in-place sweeps (on buffer cur) at any time */
    const char *mode = getenv("SMX_F32_KERNEL");
    st->fused = !(mode && !strcmp(mode, "twopass"));
    st->cur = 0;
    st->tile = kDefaultTile;
    st->fz = kDefaultFZ;
    st->two = -1;
    st->two_fz = kDefaultFZ;
    st->two_bal = 0;
    /* the two-step pass's 64-bit form on any grid ("1"), for tests of it on small grids */
    st->force_big = getenv("SMX_F32_BIG") && !strcmp(getenv("SMX_F32_BIG"), "1");
    st->pml_origin = !(getenv("SMX_F32_PML_ORIGIN") && !strcmp(getenv("SMX_F32_PML_ORIGIN"), "0"));
    if (const char *t = getenv("SMX_F32_TILE")) { /* "<variant>,<z chunk>", e.g. "1,32" */
        int ti = kDefaultTile, zc = kDefaultFZ;
        if (sscanf(t, "%d,%d", &ti, &zc) >= 1 && smx_f32_tile_ok(st, ti, zc)) {
            st->tile = ti;
            st->fz = zc;
        }
    }
    if (const char *t = getenv("SMX_F32_TWO")) { /* two-step kernel "<variant>,<z chunk>[,<balance>]" */
        int v = -1, zc = kDefaultFZ, bal = 0;
        if (sscanf(t, "%d,%d,%d", &v, &zc, &bal) >= 1 && smx_f32_two_ok(st, v, zc) && bal >= 0 && bal <= 4) {
            st->two = v;
            st->two_fz = zc;
            st->two_bal = bal;
        }
    }
    for (int k = 0; k < 2; k++) {
        F32_HIP(hipMalloc(&st->Tmem[k], sizeof(float) * 6 * n));
        F32_HIP(hipMalloc(&st->Smem[k], sizeof(float) * 6 * n));
        F32_HIP(hipMemsetAsync(st->Tmem[k], 0, sizeof(float) * 6 * n, stream));
        F32_HIP(hipMemsetAsync(st->Smem[k], 0, sizeof(float) * 6 * n, stream));
    }
    F32_HIP(hipMalloc(&st->bmem, sizeof(float) * 6 * n));
    F32_HIP(hipMemcpyAsync(st->bmem, bw.data(), sizeof(float) * 6 * n, hipMemcpyHostToDevice, stream));
    for (int i = 0; i < 6; i++) {
        for (int k = 0; k < 2; k++) {
            st->T[k][i] = st->Tmem[k] + (size_t)i * n;
            st->S[k][i] = st->Smem[k] + (size_t)i * n;
        }
        st->b[i] = st->bmem + (size_t)i * n;
    }
    F32_HIP(hipMalloc(&st->pml, sizeof(float) * pml.size()));
    F32_HIP(hipMemcpyAsync(st->pml, pml.data(), sizeof(float) * pml.size(), hipMemcpyHostToDevice, stream));
    F32_HIP(hipMalloc(&st->pmlflag, flag.size()));
    F32_HIP(hipMemcpyAsync(st->pmlflag, flag.data(), flag.size(), hipMemcpyHostToDevice, stream));
    st->zpml_host = new unsigned char[g->L2];
    for (int z = 0; z < g->L2; z++) st->zpml_host[z] = flag[g->L0 + g->L1 + z];
    st->xypml_host = new unsigned char[g->L0 + g->L1];
    for (int i = 0; i < g->L0 + g->L1; i++) st->xypml_host[i] = flag[i];
    if (!special.empty()) {
        F32_HIP(hipMalloc(&st->special, sizeof(smx_special_f32) * special.size()));
        F32_HIP(hipMemcpyAsync(st->special, special.data(), sizeof(smx_special_f32) * special.size(),
                               hipMemcpyHostToDevice, stream));
        F32_HIP(hipMalloc(&st->special_cells, sizeof(int64_t) * special_cells.size()));
        F32_HIP(hipMemcpyAsync(st->special_cells, special_cells.data(), sizeof(int64_t) * special_cells.size(),
                               hipMemcpyHostToDevice, stream));
    }
    st->cont_host = new char[n];
    st->addr_host = new int[n];
    std::memcpy(st->cont_host, Cont, n);
    std::memcpy(st->addr_host, addr, sizeof(int) * n);
    F32_HIP(hipStreamSynchronize(stream));
    return SMX_OK;
}

hipError_t smx_f32_step(smx_f32_state *st, const smx_f32_launch_args *la, hipStream_t stream)
{
    if (st->fused) {
        fused_args a;
        std::memset(&a, 0, sizeof a);
        const smx_geom &g = la->g;
        a.L0 = g.L0;
        a.L1 = g.L1;
        a.L2 = g.L2;
        if (!smx_f32_tile_ok(st, st->tile, st->fz)) return hipErrorInvalidValue;
        const tile_shape ts = kTiles[st->tile];
        /* the staged kernel without split slots covers only the PML-free box on a grid
         * with PML; the region sweeps take the rest */
        const bool split = st->tile >= kFirstNoSplitTile && st->has_pml;
        const int full[6] = {0, g.L0, 0, g.L1, 0, g.L2};
        for (int i = 0; i < 6; i++) a.box[i] = split ? st->fbox[i] : full[i];
        if (split) {
            a.nreg = st->nreg;
            std::memcpy(a.reg, st->reg, sizeof a.reg);
            std::memcpy(a.reg_start, st->reg_start, sizeof a.reg_start);
        }
        a.fz = st->fz;
        a.gx = (a.box[1] - a.box[0] + ts.fx - 1) / ts.fx;
        a.gy = (a.box[3] - a.box[2] + ts.fy - 1) / ts.fy;
        a.nblocks = a.gx * a.gy * ((a.box[5] - a.box[4] + a.fz - 1) / a.fz);
        a.sz = (int64_t)g.L0 * g.L1;
        a.w[0] = (float)g.iBc0;
        a.w[1] = (float)g.iBc1;
        a.w[2] = (float)g.iBc2;
        a.iB = g.iB;
        a.iT = g.iT;
        a.iI = g.iI;
        a.sgn = (float)g.sgn;
        for (int i = 0; i < 6; i++) {
            a.To[i] = st->T[st->cur][i];
            a.Tn[i] = st->T[1 - st->cur][i];
            a.So[i] = st->S[st->cur][i];
            a.Sn[i] = st->S[1 - st->cur][i];
            a.b[i] = st->b[i];
        }
        a.ncomp = st->n;
        a.pml = st->pml;
        a.pmlflag = st->pmlflag;
        a.special = st->special;
        a.special_cells = st->special_cells;
        a.nspecial = st->nspecial;
        a.srcE_plane = la->srcE_plane;
        a.srcH_plane = la->srcH_plane;
        a.Einc = la->Einc;
        a.Hinc = la->Hinc;
        a.srcE_f2 = la->s.srcE_f2;
        a.srcE_re = la->s.srcE_re;
        a.srcE_im = la->s.srcE_im;
        a.srcH_f2 = la->s.srcH_f2;
        a.srcH_re = la->s.srcH_re;
        a.srcH_im = la->s.srcH_im;
        a.HB = la->HB;
        a.EB = la->EB;
        a.HT = la->HT;
        a.ET = la->ET;
        a.monH_re = la->s.monH_re;
        a.monH_im = la->s.monH_im;
        a.monE_re = la->s.monE_re;
        a.monE_im = la->s.monE_im;
        hipError_t e;
        if (a.nspecial > 0) {
            hipLaunchKernelGGL(special_h_pass, dim3((a.nspecial + kThreads - 1) / kThreads), dim3(kThreads), 0,
                               stream, a);
            if ((e = hipGetLastError()) != hipSuccess) return e;
        }
        const unsigned rblocks = (unsigned)((a.reg_start[a.nreg] + kThreads - 1) / kThreads);
        if (a.nreg > 0) {
            hipLaunchKernelGGL(region_h_f32, dim3(rblocks), dim3(kThreads), 0, stream, a);
            if ((e = hipGetLastError()) != hipSuccess) return e;
        }
        switch (st->tile) {
        case 9: hipLaunchKernelGGL((fused_dma_f32<32, 8, 1, false>), dim3(a.nblocks), dim3(kThreads), 0, stream, a); break;
        case 8: hipLaunchKernelGGL((fused_dma_f32<32, 8, 2, false>), dim3(a.nblocks), dim3(kThreads), 0, stream, a); break;
        case 7: hipLaunchKernelGGL((fused_dma_f32<32, 8, 4, false>), dim3(a.nblocks), dim3(kThreads), 0, stream, a); break;
        case 6: hipLaunchKernelGGL((fused_dma_f32<16, 16, 3, false>), dim3(a.nblocks), dim3(kThreads), 0, stream, a); break;
        case 5: hipLaunchKernelGGL((fused_dma_f32<32, 8, 3, false>), dim3(a.nblocks), dim3(kThreads), 0, stream, a); break;
        case 4: hipLaunchKernelGGL((fused_dma_f32<16, 16, 2, true>), dim3(a.nblocks), dim3(kThreads), 0, stream, a); break;
        case 3: hipLaunchKernelGGL((fused_dma_f32<32, 8, 2, true>), dim3(a.nblocks), dim3(kThreads), 0, stream, a); break;
        case 1: hipLaunchKernelGGL((fused_step_f32<64, 4>), dim3(a.nblocks), dim3(kThreads), 0, stream, a); break;
        case 2: hipLaunchKernelGGL((fused_step_f32<16, 16>), dim3(a.nblocks), dim3(kThreads), 0, stream, a); break;
        default: hipLaunchKernelGGL((fused_step_f32<32, 8>), dim3(a.nblocks), dim3(kThreads), 0, stream, a);
        }
        if ((e = hipGetLastError()) != hipSuccess) return e;
        if (a.nreg > 0) {
            hipLaunchKernelGGL(region_e_f32, dim3(rblocks), dim3(kThreads), 0, stream, a);
            if ((e = hipGetLastError()) != hipSuccess) return e;
        }
        if (a.nspecial > 0) {
            hipLaunchKernelGGL(special_e_pass, dim3((a.nspecial + kThreads - 1) / kThreads), dim3(kThreads), 0,
                               stream, a);
            if ((e = hipGetLastError()) != hipSuccess) return e;
        }
        st->cur = 1 - st->cur;
        return hipSuccess;
    }
    f32_args a = base_args(st, la->g);
    a.HB = la->HB;
    a.EB = la->EB;
    a.HT = la->HT;
    a.ET = la->ET;
    a.monH_re = la->s.monH_re;
    a.monH_im = la->s.monH_im;
    a.monE_re = la->s.monE_re;
    a.monE_im = la->s.monE_im;
    const dim3 grid(a.nblocks), block(kThreads);
    /* H sweep: source from Ei */
    a.src_plane = la->srcE_plane;
    a.inc = la->Einc;
    a.src_f2 = la->s.srcE_f2;
    a.src_re = la->s.srcE_re;
    a.src_im = la->s.srcE_im;
    if (st->bx == 64) hipLaunchKernelGGL(h_sweep_f32<64>, grid, block, 0, stream, a);
    else hipLaunchKernelGGL(h_sweep_f32<32>, grid, block, 0, stream, a);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    /* E sweep: source from Hi */
    a.src_plane = la->srcH_plane;
    a.inc = la->Hinc;
    a.src_f2 = la->s.srcH_f2;
    a.src_re = la->s.srcH_re;
    a.src_im = la->s.srcH_im;
    if (st->bx == 64) hipLaunchKernelGGL(e_sweep_f32<64>, grid, block, 0, stream, a);
    else hipLaunchKernelGGL(e_sweep_f32<32>, grid, block, 0, stream, a);
    return hipGetLastError();
}

/* z chunks of the two-step kernel over [a.box[4], a.box[5]): as many as chunks of fz
 * planes would make, each of at most `maxlen` planes.  bal = 0 cuts them uniformly; bal > 0
 * balances their cost, a stage (a chunk's planes and the two halo planes on each side)
 * counting 2, a stage of a z-PML plane 2 + bal (split slots, the PML update) and one of a
 * monitor or source plane 2 + 2*bal (the accumulators, the profile loads), so the chunks
 * holding them get fewer planes: the launch lasts as long as its slowest block.  False
 * when more than kMaxZChunks chunks would be needed. */
static bool two_chunks_z(fused_args &a, const smx_f32_state *st, int fz, int bal, int maxlen)
{
    const int zb = a.box[4], ze = a.box[5], len = ze - zb;
    const int n = (len + fz - 1) / fz;
    if (n > kMaxZChunks || (n > 0 && (len + n - 1) / n > maxlen)) return false;
    a.nzc = n;
    if (bal == 0 || n <= 1) {
        for (int i = 0; i <= n; i++) a.zc[i] = min(zb + i * fz, ze);
        return true;
    }
    auto cost = [&](int p) { /* stage of plane p (unwrapped) */
        const int z = ((p % st->L[2]) + st->L[2]) % st->L[2];
        return 2 + (st->zpml_host[z] ? bal : 0) + (z == a.iB || z == a.iT || z == a.iI ? 2 * bal : 0);
    };
    /* greedy cut under a cost bound T: the number of chunks it needs */
    auto cuts = [&](int64_t T, int *zc) {
        int k = 0, start = zb;
        if (zc) zc[0] = zb;
        while (start < ze) {
            int64_t c = cost(start - 2) + cost(start - 1) + cost(start) + cost(start + 1) + cost(start + 2);
            int end = start + 1;
            while (end < ze && end - start < maxlen && c + cost(end + 2) <= T) c += cost(end + 2), end++;
            start = end;
            k++;
            if (zc && k <= kMaxZChunks) zc[k] = end;
        }
        return k;
    };
    int64_t lo = 1, hi = 0;
    for (int p = zb - 2; p < ze + 2; p++) hi += cost(p);
    while (lo < hi) { /* least bound that needs at most n chunks */
        const int64_t mid = (lo + hi) / 2;
        if (cuts(mid, nullptr) <= n) hi = mid;
        else lo = mid + 1;
    }
    a.nzc = cuts(lo, a.zc);
    return a.nzc <= kMaxZChunks;
}

/* x tiles of the two-step kernel over a.box, placed so that every stage row (ox columns
 * left of its tile) starts on a 128-byte line: a row of FX + 2*ox words then spans 2
 * lines where it would span 3, and the copies ask the L2 for a third fewer lines.  A box
 * over the whole x axis shifts its tiles by ox (the last tile wraps around); a box
 * inside it starts its tiles at the aligned column at or below its start, and the
 * kernel stores only inside it. */
static void two_tiles_x(fused_args &a, int ox, int fx, int L0)
{
    if (a.box[0] == 0 && a.box[1] == L0) {
        a.xorg = a.xlo = ox;
        a.xhi = L0 + ox;
    } else {
        a.xorg = a.box[0] - ((a.box[0] - ox) % 32 + 32) % 32;
        a.xlo = a.box[0];
        a.xhi = a.box[1];
    }
    a.gx = (a.xhi - a.xorg + fx - 1) / fx;
}

/* Tiles of `f` cells along an axis of Lv cells from origin s (the last wraps around):
 * how many hold a flagged cell in the cells they compute, which reach 1 cell before the
 * tile and 2 after it (the region G3 of the two-step pass). */
static int flagged_tiles(const unsigned char *flag, int Lv, int f, int s)
{
    int k = 0;
    for (int t = 0; t * f < Lv; t++) {
        bool hit = false;
        for (int u = s + t * f - 1; u < s + t * f + f + 2 && !hit; u++) hit = flag[((u % Lv) + Lv) % Lv] != 0;
        k += hit;
    }
    return k;
}

/* x and y tiles of the two-step pass over the whole (x, y) plane (a.box[0..3] set to it).
 * A tile whose computed cells touch x/y PML copies the split slots of its stage region on
 * every plane, so the split-slot variant starts its y tiles where the y-PML layers of both
 * faces (which meet across the wrap) fall in the fewest tile rows, 0 on ties: on a 160 x 160
 * plane with 10-cell PML, 3 tile rows instead of 4 (0.9 % less time per pass).  The x tiles
 * keep the line-aligned origin: the origin that puts both x-PML layers in the one tile
 * across the wrap (11 there, 32 PML tiles instead of 52) starts the stage rows 9 words into
 * a line and measured 4 % slower. */
static void two_tiles(fused_args &a, const smx_f32_state *st, const two_variant &tv, int L0, int L1)
{
    two_tiles_x(a, tv.ox, tv.fx, L0);
    a.gy = (L1 + tv.fy - 1) / tv.fy;
    if (!tv.split || !st->pml_origin || !st->has_pml || a.box[2] != 0 || a.box[3] != L1) return;
    const unsigned char *fy = st->xypml_host + L0;
    int by = 0, ky = flagged_tiles(fy, L1, tv.fy, 0);
    for (int s = 1; s < tv.fy && ky > 0; s++) {
        const int k = flagged_tiles(fy, L1, tv.fy, s);
        if (k < ky) ky = k, by = s;
    }
    a.box[2] = by;
    a.box[3] = L1 + by;
}

hipError_t smx_f32_reserve(smx_f32_state *st, hipStream_t stream)
{
    const bool need = st->two >= 0 && ((st->has_pml && !kTwo[st->two].split) || st->nspecial > 0);
    if (!need || st->tmpT) return hipSuccess;
    hipError_t e;
    if ((e = hipMalloc(&st->tmpT, sizeof(float) * 6 * st->n)) != hipSuccess) return e;
    if ((e = hipMalloc(&st->tmpS, sizeof(float) * 6 * st->n)) != hipSuccess) return e;
    return hipMemsetAsync(st->tmpS, 0, sizeof(float) * 6 * st->n, stream);
}

/* ---------------------------------------------------------------------------------- */
/* Several plans per launch (batches of independent runs)                              */
/* ---------------------------------------------------------------------------------- */

size_t smx_f32_multi_args_size(void) { return sizeof(fused_args); }

int smx_f32_multi_ok(const smx_f32_state *st, int fz)
{
    return st->nspecial == 0 && !two_big(st) && smx_f32_two_ok(st, kMultiVariant, fz);
}

int smx_f32_multi_slots(void)
{
    int dev = 0, ncu = 0, per_cu = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, two_step_multi<true>, 256, 0) != hipSuccess)
        return 0;
    return ncu * per_cu;
}

int smx_f32_multi_prepare(const smx_f32_state *st, const smx_f32_launch_args *la, int fz, void *out, int *nblocks)
{
    if (two_big(st) || !smx_f32_two_ok(st, kMultiVariant, fz)) return SMX_ERR_ARG;
    const smx_geom &g = la->g;
    const two_variant tv = kTwo[kMultiVariant];
    const int64_t pl4 = 4 * (int64_t)g.L0 * g.L1;
    if (la->Einc != la->Hinc + pl4 || la->EB != la->HB + pl4 || la->HT != la->HB + 2 * pl4 ||
        la->ET != la->HB + 3 * pl4)
        return SMX_ERR_ARG;
    fused_args *args = static_cast<fused_args *>(out);
    for (int k = 0; k < 2; k++) { /* k = 0: buffers cur -> 1 - cur; k = 1: the next pass */
        fused_args &a = args[k];
        std::memset(&a, 0, sizeof a);
        a.L0 = g.L0, a.L1 = g.L1, a.L2 = g.L2;
        const int box[6] = {0, g.L0, 0, g.L1, 0, g.L2};
        for (int i = 0; i < 6; i++) a.box[i] = box[i];
        a.fz = fz;
        a.iB = g.iB, a.iT = g.iT, a.iI = g.iI;
        two_tiles(a, st, tv, g.L0, g.L1);
        if (!two_chunks_z(a, st, fz, 0, two_layout<32, 8, true, false>::kMaxFZ)) return SMX_ERR_ARG;
        a.nblocks = a.gx * a.gy * a.nzc;
        a.sz = (int64_t)g.L0 * g.L1;
        a.w[0] = (float)g.iBc0, a.w[1] = (float)g.iBc1, a.w[2] = (float)g.iBc2;
        a.sgn = (float)g.sgn;
        const int c = k == 0 ? st->cur : 1 - st->cur;
        for (int i = 0; i < 6; i++) {
            a.To[i] = st->T[c][i];
            a.Tn[i] = st->T[1 - c][i];
            a.So[i] = st->S[c][i];
            a.Sn[i] = st->S[1 - c][i];
            a.b[i] = st->b[i];
        }
        a.ncomp = st->n;
        a.pml = st->pml;
        a.pmlflag = st->pmlflag;
        a.Einc = la->Einc;
        a.Hinc = la->Hinc;
        a.HB = la->HB, a.EB = la->EB, a.HT = la->HT, a.ET = la->ET;
        if (la->dft_n > 0) { /* the ring; the slot of each pass comes with its scalars */
            if (la->dft_n > kDftMaxPlanes) return SMX_ERR_ARG;
            a.dft = la->dft;
            a.dft_n = la->dft_n;
            for (int m = 0; m < la->dft_n; m++) a.dft_planes[m] = la->dft_planes[m];
        }
    }
    *nblocks = args[0].nblocks;
    return args[0].nblocks < (1 << 20) ? SMX_OK : SMX_ERR_ARG;
}

hipError_t smx_f32_multi_launch(const void *args, const int *blk, const double *scal, int nblocks, int dft,
                                 hipStream_t stream)
{
    if (dft)
        hipLaunchKernelGGL(two_step_multi<true>, dim3(nblocks), dim3(256), 0, stream,
                           static_cast<const fused_args *>(args), blk, scal);
    else
        hipLaunchKernelGGL(two_step_multi<false>, dim3(nblocks), dim3(256), 0, stream,
                           static_cast<const fused_args *>(args), blk, scal);
    return hipGetLastError();
}

const char *smx_f32_multi_kernel_name(void) { return "two_step_multi"; }

/* Boxes in [0, L) covering the unwrapped ranges [lo[d], lo[d] + len[d]) of each axis
 * (a range as long as its axis covers it once). */
static void wrap_boxes(const int L[3], const int lo[3], const int len[3], smx_region_set *rs)
{
    int iv[3][2][2], ni[3];
    for (int d = 0; d < 3; d++) {
        if (len[d] >= L[d]) {
            ni[d] = 1, iv[d][0][0] = 0, iv[d][0][1] = L[d];
            continue;
        }
        const int a = ((lo[d] % L[d]) + L[d]) % L[d];
        if (a + len[d] <= L[d]) {
            ni[d] = 1, iv[d][0][0] = a, iv[d][0][1] = a + len[d];
        } else {
            ni[d] = 2, iv[d][0][0] = a, iv[d][0][1] = L[d], iv[d][1][0] = 0, iv[d][1][1] = a + len[d] - L[d];
        }
    }
    rs->n = 0;
    rs->start[0] = 0;
    for (int i = 0; i < ni[0]; i++)
        for (int j = 0; j < ni[1]; j++)
            for (int k = 0; k < ni[2]; k++) {
                int *b = rs->box[rs->n];
                b[0] = iv[0][i][0], b[1] = iv[0][i][1], b[2] = iv[1][j][0], b[3] = iv[1][j][1];
                b[4] = iv[2][k][0], b[5] = iv[2][k][1];
                rs->start[rs->n + 1] = rs->start[rs->n] + (int64_t)(b[1] - b[0]) * (b[3] - b[2]) * (b[5] - b[4]);
                rs->n++;
            }
}

/*
 * The tiles of the two-step pass `a` (its x tiles, y tiles and z chunks) on a grid with
 * lossy/metal cells: a tile is dirty when a special cell lies in its stage region G4
 * (x0-2 .. x0+FX+1, y0-2 .. y0+FY+1) on its stage planes (z0-2 .. z1+1).  R is the box of
 * the dirty tiles (in tile indices, so a tile inside it that is clean belongs to it too);
 * the tiles outside R run the pass (st->sp_blist), R runs the region sweeps.
 */
static hipError_t special_tiles(smx_f32_state *st, const fused_args &a, int fx, int fy)
{
    int key[6] = {a.gx, a.gy, a.nzc, a.xorg, a.box[2], 0};
    for (int i = 0; i <= a.nzc; i++) key[5] = key[5] * 31 + a.zc[i];
    if (st->sp_valid && std::memcmp(key, st->sp_key, sizeof key) == 0) return hipSuccess;
    const int L[3] = {st->L[0], st->L[1], st->L[2]};
    std::vector<char> dirty((size_t)a.gx * a.gy * a.nzc, 0);
    auto tiles = [](int v, int Lv, int ntiles, auto lo_of, auto hi_of, std::vector<int> &out) {
        out.clear();
        for (int u = v - Lv; u <= v + 2 * Lv; u += Lv) /* tiles reach past Lv by their origin */
            for (int t = 0; t < ntiles; t++)
                if (u >= lo_of(t) && u < hi_of(t)) out.push_back(t);
    };
    std::vector<int> tx, ty, tz;
    for (int i = 0; i < st->nspecial; i++) {
        const int64_t l = st->special_host[i];
        const int x = (int)(l % L[0]), y = (int)((l / L[0]) % L[1]), z = (int)(l / ((int64_t)L[0] * L[1]));
        tiles(x, L[0], a.gx, [&](int t) { return a.xorg + t * fx - 2; }, [&](int t) { return a.xorg + t * fx + fx + 2; },
              tx);
        tiles(y, L[1], a.gy, [&](int t) { return a.box[2] + t * fy - 2; }, [&](int t) { return a.box[2] + t * fy + fy + 2; },
              ty);
        tiles(z, L[2], a.nzc, [&](int t) { return a.zc[t] - 2; }, [&](int t) { return a.zc[t + 1] + 2; }, tz);
        for (int bz : tz)
            for (int by : ty)
                for (int bx : tx) dirty[(size_t)bx + (size_t)a.gx * (by + (size_t)a.gy * bz)] = 1;
    }
    int lo[3] = {a.gx, a.gy, a.nzc}, hi[3] = {-1, -1, -1};
    for (int bz = 0; bz < a.nzc; bz++)
        for (int by = 0; by < a.gy; by++)
            for (int bx = 0; bx < a.gx; bx++)
                if (dirty[(size_t)bx + (size_t)a.gx * (by + (size_t)a.gy * bz)]) {
                    const int t[3] = {bx, by, bz};
                    for (int d = 0; d < 3; d++) lo[d] = std::min(lo[d], t[d]), hi[d] = std::max(hi[d], t[d]);
                }
    std::vector<int> clean;
    for (int bz = 0; bz < a.nzc; bz++)
        for (int by = 0; by < a.gy; by++)
            for (int bx = 0; bx < a.gx; bx++)
                if (bx < lo[0] || bx > hi[0] || by < lo[1] || by > hi[1] || bz < lo[2] || bz > hi[2])
                    clean.push_back(bx + a.gx * (by + a.gy * bz));
    /* R in cells (unwrapped: the last x tile reaches past L0 by xorg) */
    const int rlo[3] = {a.xorg + lo[0] * fx, a.box[2] + lo[1] * fy, a.zc[lo[2]]};
    const int rhi[3] = {std::min(a.xorg + (hi[0] + 1) * fx, a.xhi), std::min(a.box[2] + (hi[1] + 1) * fy, a.box[3]),
                        a.zc[hi[2] + 1]};
    for (int k = 0; k <= 2; k++) {
        int glo[3], glen[3];
        for (int d = 0; d < 3; d++) {
            glo[d] = rlo[d] - k;
            glen[d] = rhi[d] - rlo[d] + 2 * k;
        }
        wrap_boxes(L, glo, glen, k == 0 ? &st->sp_R : k == 1 ? &st->sp_R1 : &st->sp_R2);
        if (k == 0)
            for (int d = 0; d < 3; d++) {
                st->sp_monin[2 * d] = ((glo[d] % L[d]) + L[d]) % L[d];
                st->sp_monin[2 * d + 1] = glen[d];
            }
    }
    if (st->sp_blist) (void)hipFree(st->sp_blist);
    st->sp_blist = nullptr;
    st->sp_nclean = (int)clean.size();
    hipError_t e;
    if (!clean.empty() &&
        ((e = hipMalloc(&st->sp_blist, sizeof(int) * clean.size())) != hipSuccess ||
         (e = hipMemcpy(st->sp_blist, clean.data(), sizeof(int) * clean.size(), hipMemcpyHostToDevice)) != hipSuccess))
        return e;
    std::memcpy(st->sp_key, key, sizeof key);
    st->sp_valid = 1;
    return hipSuccess;
}

int smx_f32_box_fz(const smx_f32_state *st)
{
    const int fz = st->two == kMultiVariant ? st->two_fz : 64;
    return !two_big(st) && smx_f32_two_ok(st, kMultiVariant, fz) ? fz : 0;
}

hipError_t smx_f32_step2(smx_f32_state *st, const smx_f32_launch_args *la, hipStream_t stream)
{
    /* the design-box output runs the split-slot variant whatever the plan's choice */
    const bool boxed = la->box_out != nullptr;
    const int vr = boxed ? kMultiVariant : st->two, vfz = boxed ? smx_f32_box_fz(st) : st->two_fz;
    if (boxed ? vfz == 0 : !smx_f32_two_ok(st, vr, vfz)) return hipErrorInvalidValue;
    fused_args a;
    std::memset(&a, 0, sizeof a);
    const smx_geom &g = la->g;
    a.L0 = g.L0;
    a.L1 = g.L1;
    a.L2 = g.L2;
    const int box[6] = {0, g.L0, 0, g.L1, 0, g.L2};
    for (int i = 0; i < 6; i++) a.box[i] = box[i];
    a.fz = vfz;
    a.iB = g.iB; /* also read by the z chunking */
    a.iT = g.iT;
    a.iI = g.iI;
    const two_variant tv = kTwo[vr];
    if (boxed) {
        for (int i = 0; i < 6; i++) a.dbox[i] = la->box[i];
        a.dbx = la->box[1] - la->box[0];
        a.dby = la->box[3] - la->box[2];
        a.dnbox = (int64_t)a.dbx * a.dby * (la->box[5] - la->box[4]);
        a.dbox_out = la->box_out;
    }
    if (la->dft_n > 0) {
        /* the kernel samples only the cells it updates */
        if (!smx_f32_two_full(st, vr) || la->dft_n > kDftMaxPlanes) return hipErrorInvalidValue;
        a.dft = la->dft;
        a.dft_n = la->dft_n;
        a.dft_slot = la->dft_slot;
        for (int m = 0; m < la->dft_n; m++) a.dft_planes[m] = la->dft_planes[m];
    }
    two_tiles(a, st, tv, g.L0, g.L1);
    const int maxlen = tv.split ? two_layout<32, 8, true, false>::kMaxFZ : g.L2;
    if (!two_chunks_z(a, st, a.fz, st->two_bal, maxlen)) {
        snprintf(f32_err, sizeof f32_err, "two-step kernel: more than %d z chunks", kMaxZChunks);
        return hipErrorInvalidValue;
    }
    a.nblocks = a.gx * a.gy * a.nzc;
    auto launch_two = [&](const fused_args &k) {
        const dim3 grid(k.nblocks), block(tv.fx * tv.fy);
        const two_kernel kern = boxed ? kTwoKernelBox[k.dft_n ? 1 : 0]
                                : two_big(st) ? kTwoKernelBig[k.dft_n ? 1 : 0]
                                : k.dft_n ? kTwoKernelDft[st->two] : kTwoKernel[st->two];
        hipLaunchKernelGGL(kern, grid, block, 0, stream, k);
        return hipGetLastError();
    };
    a.sz = (int64_t)g.L0 * g.L1;
    a.w[0] = (float)g.iBc0;
    a.w[1] = (float)g.iBc1;
    a.w[2] = (float)g.iBc2;
    a.iB = g.iB;
    a.iT = g.iT;
    a.iI = g.iI;
    a.sgn = (float)g.sgn;
    for (int i = 0; i < 6; i++) {
        a.To[i] = st->T[st->cur][i];
        a.Tn[i] = st->T[1 - st->cur][i];
        a.So[i] = st->S[st->cur][i];
        a.Sn[i] = st->S[1 - st->cur][i];
        a.b[i] = st->b[i];
    }
    a.ncomp = st->n;
    a.pml = st->pml;
    a.pmlflag = st->pmlflag;
    a.Einc = la->Einc;
    a.Hinc = la->Hinc;
    a.srcE_f2 = la->s.srcE_f2;
    a.srcE_re = la->s.srcE_re;
    a.srcE_im = la->s.srcE_im;
    a.srcH_f2 = la->s.srcH_f2;
    a.srcH_re = la->s.srcH_re;
    a.srcH_im = la->s.srcH_im;
    a.monH_re = la->s.monH_re;
    a.monH_im = la->s.monH_im;
    a.monE_re = la->s.monE_re;
    a.monE_im = la->s.monE_im;
    a.srcE2_f2 = la->s2.srcE_f2;
    a.srcE2_re = la->s2.srcE_re;
    a.srcE2_im = la->s2.srcE_im;
    a.srcH2_f2 = la->s2.srcH_f2;
    a.srcH2_re = la->s2.srcH_re;
    a.srcH2_im = la->s2.srcH_im;
    a.monH2_re = la->s2.monH_re;
    a.monH2_im = la->s2.monH_im;
    a.monE2_re = la->s2.monE_re;
    a.monE2_im = la->s2.monE_im;
    a.HB = la->HB;
    a.EB = la->EB;
    a.HT = la->HT;
    a.ET = la->ET;
    /* the kernel reaches Einc, EB, HT, ET and the field components from one base each */
    const int64_t pl4 = 4 * (int64_t)g.L0 * g.L1;
    if (la->Einc != la->Hinc + pl4 || la->EB != la->HB + pl4 || la->HT != la->HB + 2 * pl4 ||
        la->ET != la->HB + 3 * pl4)
        return hipErrorInvalidValue;
    for (int i = 1; i < 6; i++)
        if (st->T[1 - st->cur][i] != st->T[1 - st->cur][0] + (size_t)i * st->n ||
            st->S[1 - st->cur][i] != st->S[1 - st->cur][0] + (size_t)i * st->n)
            return hipErrorInvalidValue;
    if (la->dft_n > 0)
        /* the kernel keeps a table of the stages of its z chunk (at most kMaxFZ planes) */
        for (int i = 0; i < a.nzc; i++)
            if (a.zc[i + 1] - a.zc[i] > two_layout<32, 8, true, false>::kMaxFZ) return hipErrorInvalidValue;
    if (st->nspecial > 0) {
        /*
         * Lossy/metal cells.  A tile's rings recompute the first step on its stage region
         * with the regular update, and the metal E update's backward curl (F:310) reaches
         * one cell further than the rings hold, so the box R of the tiles whose stage
         * region holds a special cell goes to the region sweeps and the special passes:
         *   1. H1 on R grown by 2, the specials' H1         old -> tmp
         *   2. E1 on R grown by 1, the specials' E1         old, tmp H1 -> tmp
         *   3. two steps on the tiles outside R             old -> new
         *   4. H2 on R, the specials' H2                    tmp -> new
         *   5. E2 on R, the specials' E2                    tmp E1, new H2 -> new
         * Every special cell lies in R, so each region step reads only cells the steps
         * before it wrote.  The monitors of R's first step come from step 2 (R only), of
         * its second from step 5; the pass takes those of its own tiles.
         */
        if (!tv.split || boxed) return hipErrorInvalidValue;
        hipError_t e;
        if (!st->tmpT && (e = smx_f32_reserve(st, stream)) != hipSuccess) return e;
        if (!st->tmpT) return hipErrorInvalidValue;
        if ((e = special_tiles(st, a, tv.fx, tv.fy)) != hipSuccess) return e;
        float *tT[6], *tS[6], *oT[6], *oS[6], *nT[6], *nS[6];
        for (int i = 0; i < 6; i++) {
            tT[i] = st->tmpT + (size_t)i * st->n;
            tS[i] = st->tmpS + (size_t)i * st->n;
            oT[i] = st->T[st->cur][i], oS[i] = st->S[st->cur][i];
            nT[i] = st->T[1 - st->cur][i], nS[i] = st->S[1 - st->cur][i];
        }
        auto with = [&](float *const *inT, float *const *inS, float *const *outT, float *const *outS, int step) {
            fused_args r = a;
            for (int i = 0; i < 6; i++) {
                r.To[i] = inT[i];
                r.So[i] = inS[i];
                r.Tn[i] = outT[i];
                r.Sn[i] = outS[i];
            }
            r.special = st->special;
            r.special_cells = st->special_cells;
            r.nspecial = st->nspecial;
            if (step == 2) {
                r.srcE_f2 = a.srcE2_f2, r.srcE_re = a.srcE2_re, r.srcE_im = a.srcE2_im;
                r.srcH_f2 = a.srcH2_f2, r.srcH_re = a.srcH2_re, r.srcH_im = a.srcH2_im;
                r.monH_re = a.monH2_re, r.monH_im = a.monH2_im, r.monE_re = a.monE2_re, r.monE_im = a.monE2_im;
            }
            r.monin_on = 1;
            std::memcpy(r.monin, st->sp_monin, sizeof r.monin);
            return r;
        };
        /* one launch per half step: the boxes' cells, then the special cells (every special
         * cell lies in R, so its update reads what the same half step of R reads) */
        auto sweep = [&](const smx_region_set &rs, bool e_sweep, const fused_args &r0) -> hipError_t {
            fused_args r = r0;
            r.nreg = rs.n;
            r.reg_specials = 1;
            std::memcpy(r.reg, rs.box, sizeof r.reg);
            std::memcpy(r.reg_start, rs.start, sizeof r.reg_start);
            const unsigned blocks = (unsigned)((rs.start[rs.n] + r.nspecial + kThreads - 1) / kThreads);
            if (e_sweep) hipLaunchKernelGGL(region_e_f32, dim3(blocks), dim3(kThreads), 0, stream, r);
            else hipLaunchKernelGGL(region_h_f32, dim3(blocks), dim3(kThreads), 0, stream, r);
            return hipGetLastError();
        };
        /* DFT samples of R's cells (the pass samples its own tiles): after the first step
         * from the scratch buffer into the pass's first ring slot, after the second from
         * the new buffer into the next */
        auto sample = [&](float *const *T, float *const *S, int slot) -> hipError_t {
            if (la->dft_n <= 0) return hipSuccess;
            smx_dft_sample_args d;
            std::memset(&d, 0, sizeof d);
            d.L0 = g.L0, d.L1 = g.L1, d.L2 = g.L2;
            d.nplanes = la->dft_n;
            std::memcpy(d.planes, la->dft_planes, sizeof(int) * (size_t)la->dft_n);
            d.plane_cells = (int64_t)g.L0 * g.L1;
            for (int i = 0; i < 6; i++) d.T32[i] = T[i], d.S32[i] = S[i];
            d.pmlflag = st->pmlflag;
            d.bH0 = st->b[0];
            d.bH1 = st->b[1];
            d.special = reinterpret_cast<const int *>(st->special);
            d.special_stride = (int)(sizeof(smx_special_f32) / sizeof(int));
            d.series = la->dft;
            d.mask_on = 1;
            std::memcpy(d.mask, st->sp_monin, sizeof d.mask);
            return smx_dft_sample(&d, slot, stream);
        };
        const fused_args s1 = with(oT, oS, tT, tS, 1), s2 = with(tT, tS, nT, nS, 2);
        if ((e = sweep(st->sp_R2, false, s1)) != hipSuccess || (e = sweep(st->sp_R1, true, s1)) != hipSuccess ||
            (e = sample(tT, tS, la->dft_slot)) != hipSuccess)
            return e;
        if (st->sp_nclean > 0) {
            fused_args k = a;
            k.blist = st->sp_blist;
            k.nblocks = st->sp_nclean;
            if ((e = launch_two(k)) != hipSuccess) return e;
        }
        if ((e = sweep(st->sp_R, false, s2)) != hipSuccess || (e = sweep(st->sp_R, true, s2)) != hipSuccess ||
            (e = sample(nT, nS, la->dft_slot + 1)) != hipSuccess)
            return e;
        st->cur = 1 - st->cur;
        return hipSuccess;
    }
    if (!st->has_pml || tv.split) {
        const hipError_t e = launch_two(a);
        if (e != hipSuccess) return e;
        st->cur = 1 - st->cur;
        return hipSuccess;
    }
    /*
     * With PML: the two-step kernel on its box F, the region sweeps on the rest P.
     *   1. H1 on P and both one-cell shells of F        old -> tmp
     *   2. E1 on P and the +x/+y/+z shell of F          old, tmp H1 -> tmp
     *   3. two steps on F                               old -> new
     *   4. H2 on P                                      tmp -> new
     *   5. E2 on P                                      tmp E1, new H2 -> new
     * Each region step reads only cells the steps before it wrote (the shells supply
     * the first step next to F, where P's second step needs it; F's own outputs supply
     * the second step's +x/+y/+z neighbours of P).
     */
    if (!st->tmpT) {
        const hipError_t e = smx_f32_reserve(st, stream);
        if (e != hipSuccess) return e;
    }
    float *tT[6], *tS[6];
    for (int i = 0; i < 6; i++) {
        tT[i] = st->tmpT + (size_t)i * st->n;
        tS[i] = st->tmpS + (size_t)i * st->n;
    }
    float *const oT[6] = {st->T[st->cur][0], st->T[st->cur][1], st->T[st->cur][2],
                          st->T[st->cur][3], st->T[st->cur][4], st->T[st->cur][5]};
    float *const oS[6] = {st->S[st->cur][0], st->S[st->cur][1], st->S[st->cur][2],
                          st->S[st->cur][3], st->S[st->cur][4], st->S[st->cur][5]};
    float *const nT[6] = {st->T[1 - st->cur][0], st->T[1 - st->cur][1], st->T[1 - st->cur][2],
                          st->T[1 - st->cur][3], st->T[1 - st->cur][4], st->T[1 - st->cur][5]};
    float *const nS[6] = {st->S[1 - st->cur][0], st->S[1 - st->cur][1], st->S[1 - st->cur][2],
                          st->S[1 - st->cur][3], st->S[1 - st->cur][4], st->S[1 - st->cur][5]};
    auto region = [&](const smx_region_set &rs, bool e_sweep, const float *const *inT, const float *const *inS,
                      float *const *outT, float *const *outS, int step) -> hipError_t {
        fused_args r = a;
        r.nreg = rs.n;
        std::memcpy(r.reg, rs.box, sizeof r.reg);
        std::memcpy(r.reg_start, rs.start, sizeof r.reg_start);
        for (int i = 0; i < 6; i++) {
            r.To[i] = inT[i];
            r.So[i] = inS[i];
            r.Tn[i] = outT[i];
            r.Sn[i] = outS[i];
            r.monskip[i] = st->fbox2[i];
        }
        if (step == 2) {
            r.srcE_f2 = a.srcE2_f2, r.srcE_re = a.srcE2_re, r.srcE_im = a.srcE2_im;
            r.srcH_f2 = a.srcH2_f2, r.srcH_re = a.srcH2_re, r.srcH_im = a.srcH2_im;
            r.monH_re = a.monH2_re, r.monH_im = a.monH2_im, r.monE_re = a.monE2_re, r.monE_im = a.monE2_im;
        }
        if (rs.n == 0) return hipSuccess;
        const unsigned blocks = (unsigned)((rs.start[rs.n] + kThreads - 1) / kThreads);
        if (e_sweep) hipLaunchKernelGGL(region_e_f32, dim3(blocks), dim3(kThreads), 0, stream, r);
        else hipLaunchKernelGGL(region_h_f32, dim3(blocks), dim3(kThreads), 0, stream, r);
        return hipGetLastError();
    };
    hipError_t e;
    /* 1: H1 reads old H, E, S; writes tmp */
    if ((e = region(st->regH1, false, oT, oS, tT, tS, 1)) != hipSuccess) return e;
    /* 2: E1 reads old E, S and the new H in its output set: tmp */
    if ((e = region(st->regE1, true, oT, oS, tT, tS, 1)) != hipSuccess) return e;
    /* 3: the box */
    for (int i = 0; i < 6; i++) a.box[i] = st->fbox2[i];
    two_tiles_x(a, tv.ox, tv.fx, g.L0);
    a.gy = (a.box[3] - a.box[2] + tv.fy - 1) / tv.fy;
    if (!two_chunks_z(a, st, a.fz, st->two_bal, maxlen)) {
        snprintf(f32_err, sizeof f32_err, "two-step kernel: more than %d z chunks", kMaxZChunks);
        return hipErrorInvalidValue;
    }
    a.nblocks = a.gx * a.gy * a.nzc;
    if ((e = launch_two(a)) != hipSuccess) return e;
    /* 4: H2 from tmp (H1, E1, H1 split) into new */
    if ((e = region(st->regP2, false, tT, tS, nT, nS, 2)) != hipSuccess) return e;
    /* 5: E2 from tmp (E1, E1 split) and the new H into new */
    if ((e = region(st->regP2, true, tT, tS, nT, nS, 2)) != hipSuccess) return e;
    st->cur = 1 - st->cur;
    return hipSuccess;
}

int smx_f32_write_fields(smx_f32_state *st, const double *H_t, const double *E_t, hipStream_t stream)
{
    const int64_t n = st->n;
    std::vector<float> T(6 * (size_t)n), S(6 * (size_t)n);
    for (int64_t l = 0; l < n; l++) {
        const bool special = st->cont_host[l] != 'n';
        const int64_t s = 3 * n + 3 * (int64_t)st->addr_host[l];
        for (int c = 0; c < 3; c++) {
            const double hs = special ? H_t[s + c] : 0., es = special ? E_t[s + c] : 0.;
            T[(size_t)c * n + l] = (float)(H_t[3 * l + c] + hs);
            T[(size_t)(3 + c) * n + l] = (float)(E_t[3 * l + c] + es);
            S[(size_t)c * n + l] = (float)hs;
            S[(size_t)(3 + c) * n + l] = (float)es;
        }
    }
    /* both buffers: cells that never update (frozen metal fields) must read the same
     * value from either */
    for (int i = 0; i < 6; i++)
        for (int k = 0; k < 2; k++) {
            F32_HIP(hipMemcpyAsync(st->T[k][i], T.data() + (size_t)i * n, sizeof(float) * n, hipMemcpyHostToDevice,
                                   stream));
            F32_HIP(hipMemcpyAsync(st->S[k][i], S.data() + (size_t)i * n, sizeof(float) * n, hipMemcpyHostToDevice,
                                   stream));
        }
    F32_HIP(hipStreamSynchronize(stream));
    return SMX_OK;
}

int smx_f32_read_fields(const smx_f32_state *st, double *H_t, double *E_t, hipStream_t stream)
{
    const int64_t n = st->n;
    std::vector<float> T(6 * (size_t)n), S(6 * (size_t)n);
    for (int i = 0; i < 6; i++) {
        F32_HIP(hipMemcpyAsync(T.data() + (size_t)i * n, st->T[st->cur][i], sizeof(float) * n, hipMemcpyDeviceToHost,
                               stream));
        F32_HIP(hipMemcpyAsync(S.data() + (size_t)i * n, st->S[st->cur][i], sizeof(float) * n, hipMemcpyDeviceToHost,
                               stream));
    }
    F32_HIP(hipStreamSynchronize(stream));
    const size_t flen = 3 * (size_t)(n + st->Lnotn + 1);
    std::memset(H_t, 0, sizeof(double) * flen);
    std::memset(E_t, 0, sizeof(double) * flen);
    for (int64_t l = 0; l < n; l++) {
        const bool special = st->cont_host[l] != 'n';
        const int64_t s = 3 * n + 3 * (int64_t)st->addr_host[l];
        for (int c = 0; c < 3; c++) {
            const double th = T[(size_t)c * n + l], sh = S[(size_t)c * n + l];
            const double te = T[(size_t)(3 + c) * n + l], se = S[(size_t)(3 + c) * n + l];
            H_t[3 * l + c] = th - sh;
            E_t[3 * l + c] = te - se;
            if (special) {
                H_t[s + c] = sh;
                E_t[s + c] = se;
            }
        }
    }
    return SMX_OK;
}
