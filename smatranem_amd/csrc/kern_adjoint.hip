/*
 * kern_adjoint.hip -- reverse-time (discrete adjoint) kernels of the FP32 step.
 *
 * No counterpart in the reference.  Objective: the transmitted port DFT
 *   J = sum_{v = x, y} sum_q |D_v(q)|^2,  D_v(q) = sum_k e^{i w (k + 1.5)} E_v(q, zT)^{k+1}
 * (the E monitor convention of FDTDcalc/FDTDsplitfieldMetal.c:323-325 at frequency w).
 *
 * In the packed variables (T = total, S = split, smx_f32.h) one step of a lossless
 * cell is affine:
 *   H half:  T_H' = aM T_H + (aS - aM) S_H - b gM P(T_E) - b gS M(T_E),  S_H' = aS S_H - b gS M(T_E)
 *   E half:  T_E' = aM T_E + (aS - aM) S_E + b gM P'(T_H') + b gS M'(T_H'), S_E' = aS S_E + b gS M'(T_H')
 * (P, M: the two halves of the backward curl; P', M': of the forward curl; a = g = 1,
 * S = 0 outside PML).  Its transpose, step k from lambda^{k+1} to lambda^k:
 *   kernel A: lambda_H += C_b-like gather of (b gM lambda_TE, b gS (lambda_TE + lambda_SE))
 *             (E half transposed; backward differences), after finishing the own-cell
 *             H transform left over from step k+1; gradient term of step k.
 *   kernel B: lambda_E = own-cell E transform - C_f-like gather of
 *             (b gM lambda_TH, b gS (lambda_TH + lambda_SH)) (H half transposed), plus the
 *             objective sample of state k.
 * A reads lambda_E and writes its own lambda_H; B reads lambda_H and writes its own
 * lambda_E: no races.  Gradient: dE'/d eps_c = -iR*bE_c*(E^{k+1} - E^k) in design
 * cells ('n', non-PML), so dJ/d eps_c = -iR*bE_c * sum_k lambda_TE^{k+1} (E^{k+1} - E^k).
 */
#include <hip/hip_runtime.h>

#include <stdint.h>

#include "smx_adjoint.h"
#include "smx_stage.h"

namespace {

constexpr int kThreads = 256;

struct cell3 {
    int x, y, z;
};

__device__ inline cell3 cell_of(const smx_adj_args &a, int64_t l)
{
    cell3 c;
    c.x = (int)(l % a.L0);
    c.y = (int)((l / a.L0) % a.L1);
    c.z = (int)(l / a.sz);
    return c;
}

__device__ inline int axis_len(const smx_adj_args &a, int d) { return d == 0 ? a.L0 : d == 1 ? a.L1 : a.L2; }

__device__ inline const float *axis_tab(const smx_adj_args &a, int d)
{
    return d == 0 ? a.pml : d == 1 ? a.pml + 4 * a.L0 : a.pml + 4 * (a.L0 + a.L1);
}

__device__ inline bool in_pml(const smx_adj_args &a, const cell3 &p)
{
    return a.pmlflag[p.x] | a.pmlflag[a.L0 + p.y] | a.pmlflag[a.L0 + a.L1 + p.z];
}

/* PML table values of component c of cell p: fam 0 = H (integer samples), 1 = E
 * (half samples); returns aM, aS, gM, gS (1 outside PML). */
__device__ inline void coef4(const smx_adj_args &a, const cell3 &p, int fam, int c, float &aM, float &aS, float &gM,
                             float &gS)
{
    if (!in_pml(a, p)) {
        aM = aS = gM = gS = 1.f;
        return;
    }
    const int jj[3] = {p.x, p.y, p.z};
    const int dm = c == 2 ? 0 : c + 1, ds = c == 0 ? 2 : c - 1;
    const float *tm = axis_tab(a, dm), *ts = axis_tab(a, ds);
    const int Lm = axis_len(a, dm), Ls = axis_len(a, ds);
    aM = tm[2 * fam * Lm + jj[dm]];
    gM = tm[(2 * fam + 1) * Lm + jj[dm]];
    aS = ts[2 * fam * Ls + jj[ds]];
    gS = ts[(2 * fam + 1) * Ls + jj[ds]];
}

/* Offset of the neighbour of p along axis d, backward (dir = -1) or forward (+1). */
__device__ inline int64_t nb_off(const smx_adj_args &a, const cell3 &p, int d, int dir)
{
    const int j = d == 0 ? p.x : d == 1 ? p.y : p.z;
    const int L = axis_len(a, d);
    const int64_t s = d == 0 ? 1 : d == 1 ? (int64_t)a.L0 : a.sz;
    if (dir < 0) return j == 0 ? (int64_t)(L - 1) * s : -s;
    return j + 1 == L ? -(int64_t)(L - 1) * s : s;
}

__device__ inline cell3 shifted(const smx_adj_args &a, cell3 p, int d, int dir)
{
    int *j = d == 0 ? &p.x : d == 1 ? &p.y : &p.z;
    const int L = axis_len(a, d);
    *j = dir < 0 ? (*j == 0 ? L - 1 : *j - 1) : (*j + 1 == L ? 0 : *j + 1);
    return p;
}

/* u_c = b gM lambda_T (P part) and v_c = b gS (lambda_T + lambda_S) (M part) of
 * component c at cell p for family fam (0: H, coefficients bH; 1: E, bE). */
__device__ inline void uv(const smx_adj_args &a, const cell3 &p, int64_t l, int fam, int c, float &u, float &v)
{
    float aM, aS, gM, gS;
    coef4(a, p, fam, c, aM, aS, gM, gS);
    const float b = a.b[3 * fam + c][l];
    const float lt = a.lT[3 * fam + c][l];
    const float ls = a.lS[3 * fam + c][l];
    u = b * gM * lt;
    v = b * gS * (lt + ls);
}

__global__ __launch_bounds__(kThreads) void adj_kernel_a(smx_adj_args a, int finish_h, const float *Eseg_cur,
                                                       const float *Eseg_next)
{
    const int64_t l = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    if (l >= a.n) return;
    const cell3 p = cell_of(a, l);
    float lt[3], ls[3];
#pragma unroll
    for (int d = 0; d < 3; d++) {
        lt[d] = a.lT[d][l];
        ls[d] = a.lS[d][l];
    }
    if (finish_h && in_pml(a, p)) { /* own-cell transpose of the H half of the later step */
#pragma unroll
        for (int d = 0; d < 3; d++) {
            float aM, aS, gM, gS;
            coef4(a, p, 0, d, aM, aS, gM, gS);
            const float t = lt[d];
            lt[d] = aM * t;
            ls[d] = (aS - aM) * t + aS * ls[d];
        }
    }
    /* E half of this step, transposed: backward-difference gather */
#pragma unroll
    for (int d = 0; d < 3; d++) {
        const int cP = d == 2 ? 0 : d + 1, cM = d == 0 ? 2 : d - 1; /* c = d+1 (P term), d+2 (M term) */
        const int aP = cM, aMx = cP;                                  /* axes c1 = d+2, c2 = d+1 */
        float u0, v0, u1, v1, dummy;
        uv(a, p, l, 1, cP, u0, dummy);
        const cell3 q1 = shifted(a, p, aP, -1);
        uv(a, q1, l + nb_off(a, p, aP, -1), 1, cP, u1, dummy);
        uv(a, p, l, 1, cM, dummy, v0);
        const cell3 q2 = shifted(a, p, aMx, -1);
        uv(a, q2, l + nb_off(a, p, aMx, -1), 1, cM, dummy, v1);
        const int jP = aP == 0 ? p.x : aP == 1 ? p.y : p.z;
        const int jM = aMx == 0 ? p.x : aMx == 1 ? p.y : p.z;
        const float wP = jP == 0 ? a.w[aP] : 1.f, wM = jM == 0 ? a.w[aMx] : 1.f;
        lt[d] += (-u0 + wP * u1) + (v0 - wM * v1);
    }
#pragma unroll
    for (int d = 0; d < 3; d++) {
        a.lT[d][l] = lt[d];
        a.lS[d][l] = ls[d];
    }
    /* gradient of this step in the design box */
    if (p.x >= a.box[0] && p.x < a.box[1] && p.y >= a.box[2] && p.y < a.box[3] && p.z >= a.box[4] &&
        p.z < a.box[5]) {
        const int64_t bi = (p.x - a.box[0]) + (int64_t)a.bx * ((p.y - a.box[2]) + (int64_t)a.by * (p.z - a.box[4]));
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const float dE = Eseg_next[c * a.nbox + bi] - Eseg_cur[c * a.nbox + bi];
            a.G[c * a.nbox + bi] += (double)a.lT[3 + c][l] * (double)dE;
        }
    }
}

__global__ __launch_bounds__(kThreads) void adj_kernel_b(smx_adj_args a, int inject, double cs, double sn)
{
    const int64_t l = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    if (l >= a.n) return;
    const cell3 p = cell_of(a, l);
    const bool pml = in_pml(a, p);
    float lt[3], ls[3];
#pragma unroll
    for (int d = 0; d < 3; d++) {
        const float t = a.lT[3 + d][l];
        if (pml) {
            float aM, aS, gM, gS;
            coef4(a, p, 1, d, aM, aS, gM, gS);
            lt[d] = aM * t;
            ls[d] = (aS - aM) * t + aS * a.lS[3 + d][l];
        } else {
            lt[d] = t;
            ls[d] = 0.f;
        }
    }
    /* H half transposed: minus a forward-difference gather */
#pragma unroll
    for (int d = 0; d < 3; d++) {
        const int cP = d == 2 ? 0 : d + 1, cM = d == 0 ? 2 : d - 1;
        const int aP = cM, aMx = cP; /* c1 = d+2 for the P term, c2 = d+1 for the M term */
        float U0, U1, V0, V1, dummy;
        uv(a, p, l, 0, cP, U0, dummy);
        uv(a, shifted(a, p, aP, 1), l + nb_off(a, p, aP, 1), 0, cP, U1, dummy);
        uv(a, p, l, 0, cM, dummy, V0);
        uv(a, shifted(a, p, aMx, 1), l + nb_off(a, p, aMx, 1), 0, cM, dummy, V1);
        const int jP = aP == 0 ? p.x : aP == 1 ? p.y : p.z;
        const int jM = aMx == 0 ? p.x : aMx == 1 ? p.y : p.z;
        const float wP = jP + 1 == axis_len(a, aP) ? a.w[aP] : 1.f;
        const float wM = jM + 1 == axis_len(a, aMx) ? a.w[aMx] : 1.f;
        lt[d] -= (U0 - wP * U1) + (-V0 + wM * V1);
    }
    /* objective sample of the state this step starts from */
    if (inject && p.z == a.zT) {
        const int64_t q = p.x + (int64_t)a.L0 * p.y;
#pragma unroll
        for (int v = 0; v < 2; v++) {
            const double *D = a.D + 2 * (v * (int64_t)a.L0 * a.L1 + q);
            lt[v] += (float)(2. * (D[0] * cs + D[1] * sn));
        }
    }
#pragma unroll
    for (int d = 0; d < 3; d++) {
        a.lT[3 + d][l] = lt[d];
        a.lS[3 + d][l] = ls[d];
    }
}

/* lambda_TE += 2 Re(conj(D) e^{i theta}) on the objective plane (the last state). */
__global__ __launch_bounds__(kThreads) void adj_inject(smx_adj_args a, double cs, double sn)
{
    const int64_t q = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    const int64_t plane = (int64_t)a.L0 * a.L1;
    if (q >= plane) return;
    const int64_t l = q + plane * a.zT;
#pragma unroll
    for (int v = 0; v < 2; v++) {
        const double *D = a.D + 2 * (v * plane + q);
        a.lT[3 + v][l] += (float)(2. * (D[0] * cs + D[1] * sn));
    }
}

/* D_v(q) += e^{i theta} E_v(q, zT) (the objective's DFT, FP64). */
__global__ __launch_bounds__(kThreads) void adj_sample(smx_adj_args a, const float *Ex, const float *Ey, double cs,
                                                       double sn)
{
    const int64_t q = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    const int64_t plane = (int64_t)a.L0 * a.L1;
    if (q >= plane) return;
    const int64_t l = q + plane * a.zT;
    const double ev[2] = {(double)Ex[l], (double)Ey[l]};
#pragma unroll
    for (int v = 0; v < 2; v++) {
        double *D = a.D + 2 * (v * plane + q);
        D[0] += cs * ev[v];
        D[1] += sn * ev[v];
    }
}

/* design-box copy of E: dst[c][box cell] = E_c at the cell */
__global__ __launch_bounds__(kThreads) void adj_save_box(smx_adj_args a, const float *Ex, const float *Ey,
                                                         const float *Ez, float *dst)
{
    const int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    if (i >= a.nbox) return;
    const int x = (int)(i % a.bx) + a.box[0];
    const int y = (int)((i / a.bx) % a.by) + a.box[2];
    const int z = (int)(i / ((int64_t)a.bx * a.by)) + a.box[4];
    const int64_t l = x + (int64_t)a.L0 * y + a.sz * z;
    dst[i] = Ex[l];
    dst[a.nbox + i] = Ey[l];
    dst[2 * a.nbox + i] = Ez[l];
}

/* grad_c = -iR * bE_c * G_c */
__global__ __launch_bounds__(kThreads) void adj_finish(smx_adj_args a, double iR, double *grad)
{
    const int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    if (i >= a.nbox) return;
    const int x = (int)(i % a.bx) + a.box[0];
    const int y = (int)((i / a.bx) % a.by) + a.box[2];
    const int z = (int)(i / ((int64_t)a.bx * a.by)) + a.box[4];
    const int64_t l = x + (int64_t)a.L0 * y + a.sz * z;
    for (int c = 0; c < 3; c++) grad[c * a.nbox + i] = -iR * (double)a.b[3 + c][l] * a.G[c * a.nbox + i];
}

/* ---------------------------------------------------------------------------------- */
/* The transposed step fused into one LDS-staged z-march                               */
/* ---------------------------------------------------------------------------------- */
/*
 * One launch takes lambda^{k+1} to lambda^k over the whole grid.  The transposed step
 * has the stencil footprint of the forward step with its halves swapped: first a
 * backward-difference gather over the E cells (the E half transposed) giving
 * lambda_yH on the H cells, then a forward-difference gather over those (the H half
 * transposed) giving lambda_E.  So the march is the forward fused kernel's: a block owns
 * an FX x FY tile and a z chunk; at plane p it forms lambda_yH(p) on the tile grown by
 * one cell on +x/+y (G1) from lambda_E^{k+1} of planes p and p-1, and lambda_E^k(p-1) on
 * the tile from lambda_yH of planes p-1 and p.  Per plane the stage copies HBM -> LDS by
 * buffer_load_dword ... lds, D planes ahead:
 *   E region [x0-1, x0+FX] x [y0-1, y0+FY]:  lambda_TE, bE (and lambda_SE in PML stages)
 *   H region G1 = [x0, x0+FX] x [y0, y0+FY]:  lambda_TH, bH (and lambda_SH)
 * The gathers need, at a neighbour, u_c = bE_c gM_c lambda_TE_c and
 * v_c = bE_c gS_c (lambda_TE_c + lambda_SE_c) (gM_c: the PML factor of axis c+1, gS_c:
 * of axis c+2), and every neighbour a gather reads lies along the axis of the factor it
 * needs there, so the factors come from the per-axis tables at the cell and one cell
 * back (forward: one cell on).  lambda_yH enters the second gather as U_d = bH_d gM_d
 * lambda_yTH_d and V_d = bH_d gS_d (lambda_yTH_d + lambda_ySH_d), kept in LDS for the
 * previous plane and in registers for the current one.  Own-cell transforms: E,
 * lambda_yTE = aM lambda_TE, lambda_ySE = (aS - aM) lambda_TE + aS lambda_SE, before the
 * second gather; H, lambda_TH = aM lambda_yTH, lambda_SH = (aS - aM) lambda_yTH +
 * aS lambda_ySH, after the first.  Copy ordering and waits: as fused_dma_f32
 * (kern_f32.hip).
 */
template <int FX, int FY, bool SPLIT>
struct adj_layout {
    static constexpr int EXW = FX + 2, EYW = FY + 2, NE = EXW * EYW, NEC = (NE + 63) / 64;
    static constexpr int HXW = FX + 1, HYW = FY + 1, NH = HXW * HYW, NHC = (NH + 63) / 64;
    /* word offsets in a stage slot, each region [3 components][chunks * 64], in the order
     * the copies are numbered */
    static constexpr int LTE = 0, BE = LTE + 3 * NEC * 64, LTH = BE + 3 * NEC * 64, BH = LTH + 3 * NHC * 64;
    /* a slot carries the split regions only in the instantiation for blocks near PML */
    static constexpr int LSE = BH + 3 * NHC * 64, LSH = LSE + 3 * NEC * 64;
    static constexpr int SLOT = SPLIT ? LSH + 3 * NHC * 64 : LSE;
    static constexpr int NBASE = 6 * NEC + 6 * NHC;  /* copies of every stage */
    static constexpr int NSPLIT = 3 * NEC + 3 * NHC; /* extra copies of PML stages */
    static constexpr int kMinOps = NBASE / 4;        /* per wave, at least */
    static constexpr int kMaxFZ = 64;
};

template <int FX, int FY, bool SPLIT, int W>
__device__ inline void adj_issue(float *slot, __amdgpu_buffer_rsrc_t rT, __amdgpu_buffer_rsrc_t rB,
                                 __amdgpu_buffer_rsrc_t rS, unsigned cb, unsigned soff, bool split, const unsigned *vE,
                                 const unsigned *vH)
{
    using Ly = adj_layout<FX, FY, SPLIT>;
#pragma unroll
    for (int j = W; j < Ly::NBASE; j += 4) {
        float *dst = slot + j * 64;
        if (j < 6 * Ly::NEC) { /* lambda_TE, then bE: components 3 + c */
            const int r = j % (3 * Ly::NEC), c = r / Ly::NEC, ch = r % Ly::NEC;
            __builtin_amdgcn_raw_ptr_buffer_load_lds(j < 3 * Ly::NEC ? rT : rB, SMX_LDS(dst), 4, vE[ch] + (3 + c) * cb,
                                                     soff, 0, 0);
        } else { /* lambda_TH, then bH: components c */
            const int r = (j - 6 * Ly::NEC) % (3 * Ly::NHC), c = r / Ly::NHC, ch = r % Ly::NHC;
            __builtin_amdgcn_raw_ptr_buffer_load_lds(j < 6 * Ly::NEC + 3 * Ly::NHC ? rT : rB, SMX_LDS(dst), 4,
                                                     vH[ch] + c * cb, soff, 0, 0);
        }
    }
    if (SPLIT && split) {
#pragma unroll
        for (int j = W; j < Ly::NSPLIT; j += 4) {
            float *dst = slot + Ly::LSE + j * 64;
            if (j < 3 * Ly::NEC) {
                const int c = j / Ly::NEC, ch = j % Ly::NEC;
                __builtin_amdgcn_raw_ptr_buffer_load_lds(rS, SMX_LDS(dst), 4, vE[ch] + (3 + c) * cb, soff, 0, 0);
            } else {
                const int r = j - 3 * Ly::NEC, c = r / Ly::NHC, ch = r % Ly::NHC;
                __builtin_amdgcn_raw_ptr_buffer_load_lds(rS, SMX_LDS(dst), 4, vH[ch] + c * cb, soff, 0, 0);
            }
        }
    }
}

/* Per-axis PML factors a cell of the first gather needs: gE at the cell and one cell
 * back along x and y; aH, gH at the cell (for its own transform and U, V). */
struct adj_cellf {
    float gEx, gExm, gEy, gEym, aHx, gHx, aHy, gHy;
};

/* Words of the kernel's LDS array for a march with D stages in flight: stage slots, U
 * and V of the previous plane on G1, the chunk's z tables, a flag word. */
template <int FX, int FY, int D, bool SPLIT>
constexpr int adj_lds_words()
{
    using Ly = adj_layout<FX, FY, SPLIT>;
    return (D + 1) * Ly::SLOT + 6 * Ly::NH + 5 * (Ly::kMaxFZ + 1) + 1;
}

template <int FX, int FY, int D, bool SPLIT>
__device__ __forceinline__ void adj_march(const smx_adj_fused_args &a, int id, float *lds)
{
    using Ly = adj_layout<FX, FY, SPLIT>;
    constexpr int NSLOT = D + 1;
    static_assert(D >= 1 && D * Ly::kMinOps < 64, "copies in flight must fit the vmcnt field");
    constexpr int EXW = Ly::EXW, HXW = Ly::HXW, HYW = Ly::HYW, NHG = HXW * HYW;
    constexpr int NHALO = NHG - FX * FY;
    static_assert(FX * FY == kThreads && NHALO <= kThreads, "tile shape");
    constexpr int UV_BASE = NSLOT * Ly::SLOT, ZT_BASE = UV_BASE + 6 * NHG, FLAG = ZT_BASE + 5 * (Ly::kMaxFZ + 1);
    auto sUV = [&](int k, int j, int i) -> float & { return lds[UV_BASE + k * NHG + j * HXW + i]; }; /* k: U0-2, V0-2 */
    auto zt = [&](int p, int k) -> float & { return lds[ZT_BASE + 5 * p + k]; }; /* aH, gH, aE, gE, flag */

    const int bxi = id % a.gx, rest = id / a.gx;
    const int byi = rest % a.gy, bzi = rest / a.gy;
    const int x0 = bxi * FX, y0 = byi * FY, z0 = bzi * a.fz;
    const int z1 = min(z0 + a.fz, a.L2);
    const int t = threadIdx.x, lane = t & 63;
    const int wv = __builtin_amdgcn_readfirstlane(t >> 6);

    unsigned vE[Ly::NEC], vH[Ly::NHC];
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

    /* this thread's cells: A the tile cell, B (t < NHALO) one cell of G1 outside it */
    const int ai = t % FX, aj = t / FX;
    const bool owned = x0 + ai < a.L0 && y0 + aj < a.L1;
    const int ax = wrap_index(x0 + ai, a.L0), ay = wrap_index(y0 + aj, a.L1);
    const int64_t aq = ax + (int64_t)a.L0 * ay;
    const bool hasB = t < NHALO;
    const int bi = t < HXW ? t : FX, bj = t < HXW ? FY : t - HXW;
    const int bx = wrap_index(x0 + bi, a.L0), by = wrap_index(y0 + bj, a.L1);
    const int64_t bq = bx + (int64_t)a.L0 * by;
    const float *tx = a.pml, *ty = a.pml + 4 * a.L0, *tz = a.pml + 4 * (a.L0 + a.L1);
    auto factors = [&](int x, int y) {
        adj_cellf f;
        const int xm = x == 0 ? a.L0 - 1 : x - 1, ym = y == 0 ? a.L1 - 1 : y - 1;
        f.gEx = tx[3 * a.L0 + x], f.gExm = tx[3 * a.L0 + xm], f.gEy = ty[3 * a.L1 + y], f.gEym = ty[3 * a.L1 + ym];
        f.aHx = tx[x], f.gHx = tx[a.L0 + x], f.aHy = ty[y], f.gHy = ty[a.L1 + y];
        return f;
    };
    const adj_cellf fA = factors(ax, ay), fB = factors(bx, by);
    const float aEx = tx[2 * a.L0 + ax], aEy = ty[2 * a.L1 + ay];
    /* the split-free instantiation runs only on blocks whose reach holds no PML cell
     * (the host's lists): every factor is 1 there */
    const bool pxyA = SPLIT && a.has_pml && (a.pmlflag[ax] | a.pmlflag[a.L0 + ay]);
    const bool pxyB = SPLIT && a.has_pml && (a.pmlflag[bx] | a.pmlflag[a.L0 + by]);
    for (int i = t; i <= z1 - z0; i += kThreads) {
        const int z = z0 + i == a.L2 ? 0 : z0 + i;
        zt(i, 0) = tz[z];
        zt(i, 1) = tz[a.L2 + z];
        zt(i, 2) = tz[2 * a.L2 + z];
        zt(i, 3) = tz[3 * a.L2 + z];
        zt(i, 4) = SPLIT && a.has_pml && a.pmlflag[a.L0 + a.L1 + z] ? 1.f : 0.f;
    }
    /* does any cell of the E region lie in x/y PML?  (its split slots are then loaded) */
    if (t == 0) lds[FLAG] = 0.f;
    __syncthreads();
    if (SPLIT && a.has_pml) {
        if (t < Ly::EXW && a.pmlflag[wrap_index(x0 - 1 + t, a.L0)]) lds[FLAG] = 1.f;
        if (t < Ly::EYW && a.pmlflag[a.L0 + wrap_index(y0 - 1 + t, a.L1)]) lds[FLAG] = 1.f;
    }
    __syncthreads();
    const bool tile_pml = lds[FLAG] != 0.f;

    /* u_1 and v_0 of plane z0-1 at the two cells: the -z neighbours of the first gather */
    float u1z[2], v0z[2];
    {
        const int zm = z0 == 0 ? a.L2 - 1 : z0 - 1;
        const float gz = tz[3 * a.L2 + zm];
        const int64_t off = (int64_t)zm * a.sz;
        const int64_t qs[2] = {aq, bq};
#pragma unroll
        for (int r = 0; r < 2; r++) {
            const int64_t l = off + qs[r];
            const float lt0 = a.lT[3 * a.ncomp + l], lt1 = a.lT[4 * a.ncomp + l];
            const float ls0 = SPLIT ? a.lS[3 * a.ncomp + l] : 0.f;
            const float b0 = a.b[3 * a.ncomp + l], b1 = a.b[4 * a.ncomp + l];
            u1z[r] = b1 * gz * lt1;      /* u_1: gM of E_y is the z factor */
            v0z[r] = b0 * gz * (lt0 + ls0); /* v_0: gS of E_x is the z factor */
        }
        retire_loads();
    }

    const __amdgpu_buffer_rsrc_t rT = group_rsrc(a.lT, a.ncomp), rB = group_rsrc(a.b, a.ncomp),
                                 rS = group_rsrc(a.lS, a.ncomp);
    const unsigned cb = 4u * (unsigned)a.ncomp;
    auto issue = [&](int p) {
        const int z = z0 + p == a.L2 ? 0 : z0 + p;
        const unsigned soff = 4u * (unsigned)((int64_t)z * a.sz);
        const bool split = tile_pml || zt(p, 4) != 0.f;
        float *slot = lds + (p % NSLOT) * Ly::SLOT;
        switch (wv) {
        case 0: adj_issue<FX, FY, SPLIT, 0>(slot, rT, rB, rS, cb, soff, split, vE, vH); break;
        case 1: adj_issue<FX, FY, SPLIT, 1>(slot, rT, rB, rS, cb, soff, split, vE, vH); break;
        case 2: adj_issue<FX, FY, SPLIT, 2>(slot, rT, rB, rS, cb, soff, split, vE, vH); break;
        default: adj_issue<FX, FY, SPLIT, 3>(slot, rT, rB, rS, cb, soff, split, vE, vH); break;
        }
    };

    const int np = z1 - z0; /* p = 0 .. np: first gather at z0+p, second at z0+p-1 */
    for (int p = 0; p < D && p <= np; p++) issue(p);
    float lteK[3] = {0.f, 0.f, 0.f}, lseK[3] = {0.f, 0.f, 0.f}; /* lambda_E^{k+1} of plane p-1 at A */
    bool pK = false;
    for (int p = 0; p <= np; p++) {
        if (p + D <= np) {
            issue(p + D);
            wait_vm_barrier<D * Ly::kMinOps>();
        } else {
            switch (np - p) {
            case 0: wait_vm_barrier<0>(); break;
            case 1: wait_vm_barrier<Ly::kMinOps>(); break;
            case 2: wait_vm_barrier<(D > 2 ? 2 : 0) * Ly::kMinOps>(); break;
            default: wait_vm_barrier<(D > 3 ? 3 : 0) * Ly::kMinOps>(); break;
            }
        }
        const float *S = lds + (p % NSLOT) * Ly::SLOT;
        const int z = z0 + p == a.L2 ? 0 : z0 + p;
        const int64_t offp = (int64_t)z * a.sz;
        const bool zp = zt(p, 4) != 0.f, split = SPLIT && (tile_pml || zp);
        const float gEz = zt(p, 3), aHz = zt(p, 0), gHz = zt(p, 1);

        /* lambda_yH(z) on G1, its U and V; the own-cell H transform for the store */
        auto gather1 = [&](int i, int j, int x, int y, const adj_cellf &f, float &u1c, float &v0c, float UV[6],
                           float nth[3], float nsh[3], bool pml) {
            auto E = [&](int off, int c, int e) { return S[off + c * Ly::NEC * 64 + e]; };
            const int eo = (j + 1) * EXW + i + 1, exm = eo - 1, eym = eo - EXW;
            float lt[3], ls[3], be[3];
#pragma unroll
            for (int c = 0; c < 3; c++) {
                lt[c] = E(Ly::LTE, c, eo);
                be[c] = E(Ly::BE, c, eo);
                ls[c] = split ? E(Ly::LSE, c, eo) : 0.f;
            }
            /* own u_c (gM: axes y, z, x) and v_c (gS: axes z, x, y) */
            const float u0 = be[0] * f.gEy * lt[0], u1 = be[1] * gEz * lt[1], u2 = be[2] * f.gEx * lt[2];
            const float v0 = be[0] * gEz * (lt[0] + ls[0]), v1 = be[1] * f.gEx * (lt[1] + ls[1]);
            const float v2 = be[2] * f.gEy * (lt[2] + ls[2]);
            /* one cell back: u_2, v_1 along x; u_0, v_2 along y */
            const float u2xm = E(Ly::BE, 2, exm) * f.gExm * E(Ly::LTE, 2, exm);
            const float v1xm = E(Ly::BE, 1, exm) * f.gExm * (E(Ly::LTE, 1, exm) + (split ? E(Ly::LSE, 1, exm) : 0.f));
            const float u0ym = E(Ly::BE, 0, eym) * f.gEym * E(Ly::LTE, 0, eym);
            const float v2ym = E(Ly::BE, 2, eym) * f.gEym * (E(Ly::LTE, 2, eym) + (split ? E(Ly::LSE, 2, eym) : 0.f));
            const float wx = x == 0 ? a.w[0] : 1.f, wy = y == 0 ? a.w[1] : 1.f, wz = z == 0 ? a.w[2] : 1.f;
            const int ho = j * HXW + i;
            float ly[3], lsh[3], bh[3];
#pragma unroll
            for (int c = 0; c < 3; c++) {
                ly[c] = S[Ly::LTH + c * Ly::NHC * 64 + ho];
                bh[c] = S[Ly::BH + c * Ly::NHC * 64 + ho];
                lsh[c] = split ? S[Ly::LSH + c * Ly::NHC * 64 + ho] : 0.f;
            }
            ly[0] += (-u1 + wz * u1c) + (v2 - wy * v2ym);
            ly[1] += (-u2 + wx * u2xm) + (v0 - wz * v0c);
            ly[2] += (-u0 + wy * u0ym) + (v1 - wx * v1xm);
            u1c = u1;
            v0c = v0;
            /* U_d = bH gM lambda_yTH, V_d = bH gS (lambda_yTH + lambda_ySH); gM: y, z, x; gS: z, x, y */
            UV[0] = bh[0] * f.gHy * ly[0];
            UV[1] = bh[1] * gHz * ly[1];
            UV[2] = bh[2] * f.gHx * ly[2];
            UV[3] = bh[0] * gHz * (ly[0] + lsh[0]);
            UV[4] = bh[1] * f.gHx * (ly[1] + lsh[1]);
            UV[5] = bh[2] * f.gHy * (ly[2] + lsh[2]);
            if (pml) {
                const float aM[3] = {f.aHy, aHz, f.aHx}, aS[3] = {aHz, f.aHx, f.aHy};
#pragma unroll
                for (int c = 0; c < 3; c++) {
                    nth[c] = aM[c] * ly[c];
                    nsh[c] = (aS[c] - aM[c]) * ly[c] + aS[c] * lsh[c];
                }
            } else {
#pragma unroll
                for (int c = 0; c < 3; c++) {
                    nth[c] = ly[c];
                    nsh[c] = 0.f;
                }
            }
        };
        float uvA[6], uvB[6], nthA[3], nshA[3], nthB[3], nshB[3];
        const bool pmlA = pxyA | zp, pmlB = pxyB | zp;
        gather1(ai, aj, ax, ay, fA, u1z[0], v0z[0], uvA, nthA, nshA, pmlA);
        if (owned && p < np) {
#pragma unroll
            for (int c = 0; c < 3; c++) a.nT[c * a.ncomp + offp + aq] = nthA[c];
            if (pmlA)
#pragma unroll
                for (int c = 0; c < 3; c++) a.nS[c * a.ncomp + offp + aq] = nshA[c];
        }
        if (hasB) gather1(bi, bj, bx, by, fB, u1z[1], v0z[1], uvB, nthB, nshB, pmlB);

        /* lambda_E^k(z-1) on the tile cell: own transform minus the forward gather of U, V */
        if (p > 0 && owned) {
            const int k = z0 + p - 1;
            const int64_t l = (int64_t)k * a.sz + aq;
            const float wx = ax + 1 == a.L0 ? a.w[0] : 1.f, wy = ay + 1 == a.L1 ? a.w[1] : 1.f;
            const float wz = k + 1 == a.L2 ? a.w[2] : 1.f;
            float lt[3], ls[3];
            if (pK) {
                const float aEz = zt(p - 1, 2);
                const float aM[3] = {aEy, aEz, aEx}, aS[3] = {aEz, aEx, aEy};
#pragma unroll
                for (int c = 0; c < 3; c++) {
                    lt[c] = aM[c] * lteK[c];
                    ls[c] = (aS[c] - aM[c]) * lteK[c] + aS[c] * lseK[c];
                }
            } else {
#pragma unroll
                for (int c = 0; c < 3; c++) {
                    lt[c] = lteK[c];
                    ls[c] = 0.f;
                }
            }
            /* d: (U_{d+1} - wP U_{d+1}(+e_{d+2})) + (-V_{d+2} + wM V_{d+2}(+e_{d+1})) */
            lt[0] -= (sUV(1, aj, ai) - wz * uvA[1]) + (-sUV(5, aj, ai) + wy * sUV(5, aj + 1, ai));
            lt[1] -= (sUV(2, aj, ai) - wx * sUV(2, aj, ai + 1)) + (-sUV(3, aj, ai) + wz * uvA[3]);
            lt[2] -= (sUV(0, aj, ai) - wy * sUV(0, aj + 1, ai)) + (-sUV(4, aj, ai) + wx * sUV(4, aj, ai + 1));
#pragma unroll
            for (int c = 0; c < 3; c++) a.nT[(3 + c) * a.ncomp + l] = lt[c];
            if (pK)
#pragma unroll
                for (int c = 0; c < 3; c++) a.nS[(3 + c) * a.ncomp + l] = ls[c];
        }
        /* carry plane z to the next iteration */
        {
            const int eo = (aj + 1) * EXW + ai + 1;
#pragma unroll
            for (int c = 0; c < 3; c++) {
                lteK[c] = S[Ly::LTE + c * Ly::NEC * 64 + eo];
                lseK[c] = split ? S[Ly::LSE + c * Ly::NEC * 64 + eo] : 0.f;
            }
        }
        pK = pmlA;
        wait_vm_barrier<63>(); /* every wave is done with stage p and with U, V of plane p-1 */
#pragma unroll
        for (int k = 0; k < 6; k++) {
            sUV(k, aj, ai) = uvA[k];
            if (hasB) sUV(k, bj, bi) = uvB[k];
        }
    }
}

/* One launch over the grid's blocks (ids bx + gx*(by + gy*bz), XCD-aware order); the
 * host's list marks those whose reach holds a PML cell (bit 30): they march with split
 * slots and one stage in flight, the others without split slots and two stages in
 * flight.  Both marches fit the same LDS array, so the two kinds share the launch. */
template <int FX, int FY>
__global__ __launch_bounds__(kThreads) void adj_fused_f32(smx_adj_fused_args a)
{
    constexpr int W1 = adj_lds_words<FX, FY, 1, true>(), W2 = adj_lds_words<FX, FY, 2, false>();
    __shared__ float lds[W1 > W2 ? W1 : W2];
    const int e = a.blocks[xcd_logical(blockIdx.x, gridDim.x)];
    const int id = e & ~(1 << 30);
    if (e & (1 << 30)) adj_march<FX, FY, 1, true>(a, id, lds);
    else adj_march<FX, FY, 2, false>(a, id, lds);
}

/* G_c += lambda_TE^{k+1} (E^{k+1} - E^k) in the design box (FP64 accumulators). */
__global__ __launch_bounds__(kThreads) void adj_grad(smx_adj_args a, const float *lTE, const float *cur,
                                                     const float *next)
{
    const int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    if (i >= a.nbox) return;
    const int x = (int)(i % a.bx) + a.box[0];
    const int y = (int)((i / a.bx) % a.by) + a.box[2];
    const int z = (int)(i / ((int64_t)a.bx * a.by)) + a.box[4];
    const int64_t l = x + (int64_t)a.L0 * y + a.sz * z;
#pragma unroll
    for (int c = 0; c < 3; c++) {
        const float dE = next[c * a.nbox + i] - cur[c * a.nbox + i];
        a.G[c * a.nbox + i] += (double)lTE[c * a.n + l] * (double)dE;
    }
}

dim3 grid_for(int64_t count) { return dim3((unsigned)((count + kThreads - 1) / kThreads)); }

} // namespace

extern "C" {

hipError_t smx_adj_launch_fused(const smx_adj_fused_args *a, hipStream_t s)
{
    hipLaunchKernelGGL((adj_fused_f32<kAdjFX, kAdjFY>), dim3(a->n_pml + a->n_free), dim3(kThreads), 0, s, *a);
    return hipGetLastError();
}

hipError_t smx_adj_launch_grad(const smx_adj_args *a, const float *lTE, const float *cur, const float *next,
                               hipStream_t s)
{
    hipLaunchKernelGGL(adj_grad, grid_for(a->nbox), dim3(kThreads), 0, s, *a, lTE, cur, next);
    return hipGetLastError();
}

hipError_t smx_adj_launch_a(const smx_adj_args *a, int finish_h, const float *cur, const float *next, hipStream_t s)
{
    hipLaunchKernelGGL(adj_kernel_a, grid_for(a->n), dim3(kThreads), 0, s, *a, finish_h, cur, next);
    return hipGetLastError();
}

hipError_t smx_adj_launch_b(const smx_adj_args *a, int inject, double cs, double sn, hipStream_t s)
{
    hipLaunchKernelGGL(adj_kernel_b, grid_for(a->n), dim3(kThreads), 0, s, *a, inject, cs, sn);
    return hipGetLastError();
}

hipError_t smx_adj_launch_inject(const smx_adj_args *a, double cs, double sn, hipStream_t s)
{
    hipLaunchKernelGGL(adj_inject, grid_for((int64_t)a->L0 * a->L1), dim3(kThreads), 0, s, *a, cs, sn);
    return hipGetLastError();
}

hipError_t smx_adj_launch_sample(const smx_adj_args *a, const float *Ex, const float *Ey, double cs, double sn,
                                 hipStream_t s)
{
    hipLaunchKernelGGL(adj_sample, grid_for((int64_t)a->L0 * a->L1), dim3(kThreads), 0, s, *a, Ex, Ey, cs, sn);
    return hipGetLastError();
}

hipError_t smx_adj_launch_save_box(const smx_adj_args *a, const float *Ex, const float *Ey, const float *Ez,
                                   float *dst, hipStream_t s)
{
    hipLaunchKernelGGL(adj_save_box, grid_for(a->nbox), dim3(kThreads), 0, s, *a, Ex, Ey, Ez, dst);
    return hipGetLastError();
}

hipError_t smx_adj_launch_finish(const smx_adj_args *a, double iR, double *grad, hipStream_t s)
{
    hipLaunchKernelGGL(adj_finish, grid_for(a->nbox), dim3(kThreads), 0, s, *a, iR, grad);
    return hipGetLastError();
}

} // extern "C"
