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

/* No fused multiply-adds: the one-step and the two-step transposed kernels (and the
 * injection and gradient kernels) then round every product and sum alike, so the
 * two-step pass gives bit for bit what one step per launch gives. */
#pragma clang fp contract(off)

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
    static constexpr int NT = FX * FY, NW = NT / 64;  /* one thread per tile cell */
    static constexpr int NBASE = 6 * NEC + 6 * NHC;  /* copies of every stage */
    static constexpr int NSPLIT = 3 * NEC + 3 * NHC; /* extra copies of PML stages */
    static constexpr int kMinOps = NBASE / NW;       /* per wave, at least */
    static constexpr int kMaxFZ = 64;
};

template <int FX, int FY, bool SPLIT, int W>
__device__ inline void adj_issue(float *slot, __amdgpu_buffer_rsrc_t rT, __amdgpu_buffer_rsrc_t rB,
                                 __amdgpu_buffer_rsrc_t rS, unsigned cb, unsigned soff, bool split, const unsigned *vE,
                                 const unsigned *vH)
{
    using Ly = adj_layout<FX, FY, SPLIT>;
#pragma unroll
    for (int j = W; j < Ly::NBASE; j += Ly::NW) {
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
        for (int j = W; j < Ly::NSPLIT; j += Ly::NW) {
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
    constexpr int NT = Ly::NT, NHALO = NHG - NT;
    static_assert((NT == 256 || NT == 512) && NHALO <= NT, "tile shape");
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
    for (int i = t; i <= z1 - z0; i += NT) {
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
        case 3: adj_issue<FX, FY, SPLIT, 3>(slot, rT, rB, rS, cb, soff, split, vE, vH); break;
        case 4: if constexpr (Ly::NW > 4) adj_issue<FX, FY, SPLIT, 4>(slot, rT, rB, rS, cb, soff, split, vE, vH); break;
        case 5: if constexpr (Ly::NW > 4) adj_issue<FX, FY, SPLIT, 5>(slot, rT, rB, rS, cb, soff, split, vE, vH); break;
        case 6: if constexpr (Ly::NW > 4) adj_issue<FX, FY, SPLIT, 6>(slot, rT, rB, rS, cb, soff, split, vE, vH); break;
        default: if constexpr (Ly::NW > 4) adj_issue<FX, FY, SPLIT, 7>(slot, rT, rB, rS, cb, soff, split, vE, vH); break;
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
__global__ __launch_bounds__(FX * FY) void adj_fused_f32(smx_adj_fused_args a)
{
    constexpr int W1 = adj_lds_words<FX, FY, 1, true>(), W2 = adj_lds_words<FX, FY, 2, false>();
    __shared__ float lds[W1 > W2 ? W1 : W2];
    const int e = a.blocks[xcd_logical(blockIdx.x, gridDim.x)];
    const int id = e & ~(1 << 30);
    if (e & (1 << 30)) adj_march<FX, FY, 1, true>(a, id, lds);
    else adj_march<FX, FY, 2, false>(a, id, lds);
}

/* ---------------------------------------------------------------------------------- */
/* Two transposed steps per pass                                                       */
/* ---------------------------------------------------------------------------------- */
/*
 * One launch takes lambda^{k+1} to lambda^{k-1}: the transposed step is bound by HBM as
 * the forward step is, so it gets the same temporal blocking (two_step_f32,
 * kern_f32.hip).  The transposed step has the forward step's footprint -- its H half
 * gathers E-like values at the cell and one cell back, its E half H-like values at the
 * cell and one cell on -- so the march is the forward pass's: a block owns a 32 x 8 tile
 * T and a z chunk [z0, z1), and marches stages s (plane p = z0-2+s) carrying two levels,
 * skewed by one plane:
 *   a. lambda_yH1(p)   on G3 from lambda_E0 of planes p, p-1   (step k, H half)
 *   b. lambda_E1(p-1)  on G2 from U1, V1 of planes p-1, p      (step k, E half)
 *   c. lambda_yH2(p-1) on G1 from u, v of lambda_E1, p-1, p-2 (step k-1, H half)
 *   d. lambda_E2(p-2)  on T  from U2, V2 of planes p-2, p-1    (step k-1, E half)
 * (G1 = T grown by one cell on +x/+y, G2 = G1 by one on -x/-y, G3 = G2 by one on +x/+y,
 * G4 = G3 by one on -x/-y, the stage region.)  The per-cell arithmetic is adj_march's,
 * operation for operation, so the pass gives bit for bit what two launches of
 * adj_fused_f32 give.  Per stage the old lambda_TE, lambda_TH, bH, bE (and in PML stages
 * lambda_SH, lambda_SE) of G4 come in by asynchronous copies to LDS, one stage ahead.
 * The neighbour values go through LDS in compact arrays, four per level (the ones read
 * one cell away: U0, U2, V1, V2 of the H halves on +x/+y; u0, u2, v1, v2 of lambda_E1 on
 * -x/-y); own-cell and -z/+z values stay in registers.  Inside the pass: the objective's
 * injection into lambda_TE^k on plane zT (every cell of level b, as adj_inject does
 * after a one-step launch), and lambda_TE^k of the tile cells in the design box written
 * to lbox for the gradient.  lambda^{k-1} gets its injection from adj_inject afterwards.
 */
namespace adj2 {
constexpr int FX = 32, FY = 8, NT = FX * FY, NW = NT / 64, OX = 2;
constexpr int W4 = FX + 4, H4 = FY + 4, N4 = W4 * H4; /* G4, the stage region */
constexpr int W3 = FX + 3, N3 = W3 * (FY + 3);        /* G3 = [1, W4) x [1, H4) of G4 */
constexpr int W2 = FX + 2, N2 = W2 * (FY + 2);        /* G2 = [1, W4-1) x [1, H4-1) */
constexpr int W1 = FX + 1, N1 = W1 * (FY + 1);        /* G1 = [2, W4-1) x [2, H4-1) */
constexpr int NC = (N4 + 63) / 64, ARR = 3 * NC * 64, SLOT = 6 * ARR;
/* rows of G4 each array component k = 3 arr + c is read on (arrays 0 lambda_TE, 1
 * lambda_TH, 2 bH, 3 bE, 4 lambda_SH, 5 lambda_SE): the E side on all of G4, but its y
 * components (a cell's -y neighbour is read only for its x and z components) from the
 * second row; the H
 * side on G3, from the second row, its components that the last row's U0 and V2 do not
 * need (lambda_TH and bH y, lambda_SH x and y) a row short.  A component's copies start
 * at its first row, and its LDS words count from there. */
constexpr int row0(int k) { return k / 3 == 0 || k / 3 == 3 || k / 3 == 5 ? (k % 3 == 1 ? 1 : 0) : 1; }
constexpr int rows(int k)
{
    return k / 3 == 0 || k / 3 == 3 || k / 3 == 5 ? (k % 3 == 1 ? H4 - 1 : H4)
           : k == 4 || k == 7 || k == 12 || k == 13 ? H4 - 2
                                                    : H4 - 1;
}
constexpr int nca(int k) { return (rows(k) * W4 + 63) / 64; }
constexpr int cum(int k) { return k == 0 ? 0 : cum(k - 1) + nca(k - 1); }
constexpr int kof_from(int j, int k) { return k >= 17 || j < cum(k + 1) ? 0 : 1 + kof_from(j, k + 1); }
constexpr int kof(int j) { return j < cum(1) ? 0 : 1 + kof_from(j, 1); } /* the component of copy j */
constexpr int NBASE = cum(12), NSPLIT = cum(18) - cum(12), kMinOps = NBASE / NW;
constexpr int kMaxFZ = 48;
constexpr int L_UV1 = 2 * SLOT, L_UE1 = L_UV1 + 4 * N3, L_UV2 = L_UE1 + 4 * N2, L_ZT = L_UV2 + 4 * N1;
constexpr int L_FLAG = L_ZT + 5 * (kMaxFZ + 4), L_END = L_FLAG + 1;
static_assert(4 * L_END <= 81920, "two blocks per CU");
static_assert(2 * kMinOps < 64, "copies in flight must fit the vmcnt field");
/* compact arrays: the four values read one cell away */
enum { A0 = 0, A2 = 1, B1 = 2, B2 = 3 }; /* U0|u0, U2|u2, V1|v1, V2|v2 */
} // namespace adj2

/* Copies of one stage dealt to wave W: lambda_TE (T comps 3..5), lambda_TH (T 0..2), bH
 * (b 0..2), bE (b 3..5), and in PML stages lambda_SH (S 0..2), lambda_SE (S 3..5), each
 * component in nca(k) runs of 64 words of its rows of G4 (vq: from the first row; vr:
 * from the second; vlast: the last run of the components that stop a row short). */
template <int W>
__device__ inline void adj2_issue(float *slot, __amdgpu_buffer_rsrc_t rT, __amdgpu_buffer_rsrc_t rB,
                                  __amdgpu_buffer_rsrc_t rS, unsigned cb, unsigned soff, bool split, const unsigned *vq,
                                  const unsigned *vr, unsigned vlast)
{
    auto one = [&](auto jc) {
        using namespace adj2;
        constexpr int j = decltype(jc)::value, k = kof(j), arr = k / 3, c = k % 3, ch = j - cum(k);
        constexpr unsigned comp = arr >= 4 ? (arr - 4) * 3 + c : (arr == 0 || arr == 3) ? 3 + c : c;
        float *dst = slot + arr * ARR + (c * NC + ch) * 64;
        const unsigned voff = row0(k) == 0 ? vq[ch] : rows(k) == H4 - 2 && ch == nca(k) - 1 ? vlast : vr[ch];
        __builtin_amdgcn_raw_ptr_buffer_load_lds(arr < 2 ? rT : arr < 4 ? rB : rS, SMX_LDS(dst), 4, voff + comp * cb,
                                                 soff, 0, 0);
    };
    static_for<W, adj2::NBASE, adj2::NW>(one);
    if (split) static_for<adj2::NBASE + W, adj2::NBASE + adj2::NSPLIT, adj2::NW>(one);
}

/* Own-cell transform of a PML cell: main aM x, split (aS - aM) x + aS s. */
__device__ inline void adj_own(const float x[3], const float s[3], bool pml, const float aM[3], const float aS[3],
                               float t[3], float u[3])
{
    if (pml) {
#pragma unroll
        for (int c = 0; c < 3; c++) {
            t[c] = aM[c] * x[c];
            u[c] = (aS[c] - aM[c]) * x[c] + aS[c] * s[c];
        }
    } else {
#pragma unroll
        for (int c = 0; c < 3; c++) {
            t[c] = x[c];
            u[c] = 0.f;
        }
    }
}

__global__ __launch_bounds__(adj2::NT) void adj_two_f32(smx_adj_two_args a)
{
    using namespace adj2;
    __shared__ __attribute__((aligned(16))) float lds[L_END];
    const int id = a.blocks[xcd_logical(blockIdx.x, gridDim.x)];
    const int bxi = id % a.gx, rest = id / a.gx;
    const int byi = rest % a.gy, bzi = rest / a.gy;
    /* x tiles start OX columns right of a multiple of 32, so that every row of the stage
     * region spans two 128-byte lines; the last one wraps around (its cells from L0 + OX
     * on are computed, not stored) */
    const int x0 = OX + bxi * FX, y0 = byi * FY, z0 = bzi * a.fz, z1 = min(z0 + a.fz, a.L2);
    const int t = threadIdx.x, lane = t & 63;
    const int wv = __builtin_amdgcn_readfirstlane(t >> 6);
    const int ns = z1 - z0 + 4; /* stages: planes z0-2 .. z1+1 */
    auto plane_of = [&](int s) { return wrap_index(z0 - 2 + s, a.L2); };

    unsigned vq[NC], vr[NC], vlast;
    auto region_off = [&](int e) {
        return 4u * (unsigned)(wrap_index(x0 - OX + e % W4, a.L0) + a.L0 * wrap_index(y0 - 2 + e / W4, a.L1));
    };
#pragma unroll
    for (int ch = 0; ch < NC; ch++) {
        vq[ch] = region_off(min(ch * 64 + lane, N4 - 1));
        vr[ch] = region_off(min(W4 + ch * 64 + lane, N4 - 1)); /* from the second row */
    }
    vlast = region_off(min(W4 + (nca(4) - 1) * 64 + lane, N4 - W4 - 1)); /* ... to the second-last */

    /* this thread's cells: r = 0 the tile cell, r = 1 one cell of the rings G1\T, G2\G1,
     * G3\G2 (ring cell h on thread (h % 4)*64 + h/4).  Local coordinates count from the
     * G4 corner (x0-OX, y0-2).  Compact-array indices: i3, i2, i1 where the cell writes
     * (-1: not in that region); rb, rc, rd where its level b, c, d reads (a cell of the
     * level's region stands in for cells outside it, whose results are dropped). */
    constexpr int RC = FY + 1 + FX, RB = RC + FY + 2 + FX + 1, RA = RB + FY + 3 + FX + 2;
    static_assert(RA <= NT, "rings fit one cell per thread");
    int ce[2], i3[2], i2[2], i1[2], rb[2], rc[2], rd[2], cx[2], cy[2];
    int64_t cq[2];
    bool own[2], pxy[2];
    float tH[2][4], tE[2][4], gm[2][2]; /* aH_x gH_x aH_y gH_y; aE_x gE_x aE_y gE_y; gE_x(x-1) gE_y(y-1) */
    const float *tx = a.pml, *ty = a.pml + 4 * a.L0, *tz = a.pml + 4 * (a.L0 + a.L1);
#pragma unroll
    for (int r = 0; r < 2; r++) {
        int gx, gy, ring = 0;
        if (r == 0) {
            gx = OX + t % FX;
            gy = 2 + t / FX;
        } else {
            const int h = (t & 63) * 4 + (t >> 6);
            if (h < RC) {
                ring = 1;
                gx = h < FY + 1 ? OX + FX : OX + (h - (FY + 1));
                gy = h < FY + 1 ? 2 + h : FY + 2;
            } else if (h < RB) {
                const int k = h - RC;
                ring = 2;
                gx = k < FY + 2 ? OX - 1 : OX + (k - (FY + 2));
                gy = k < FY + 2 ? 1 + k : 1;
            } else if (h < RA) {
                const int k = h - RB;
                ring = 3;
                gx = k < FY + 3 ? OX + FX + 1 : OX - 1 + (k - (FY + 3));
                gy = k < FY + 3 ? 1 + k : FY + 3;
            } else {
                ring = 4;
                gx = gy = 1;
            }
        }
        ce[r] = gy * W4 + gx;
        i3[r] = ring <= 3 ? (gy - 1) * W3 + (gx - 1) : -1;
        i2[r] = ring <= 2 ? (gy - 1) * W2 + (gx - 1) : -1;
        i1[r] = ring <= 1 ? (gy - 2) * W1 + (gx - 2) : -1;
        rb[r] = ring <= 2 ? i3[r] : 0;
        rc[r] = ring <= 1 ? i2[r] : W2 + 1;
        rd[r] = r == 0 ? i1[r] : 0;
        cx[r] = wrap_index(x0 - OX + gx, a.L0);
        cy[r] = wrap_index(y0 - 2 + gy, a.L1);
        cq[r] = cx[r] + (int64_t)a.L0 * cy[r];
        own[r] = r == 0 && x0 - OX + gx < a.L0 + OX && y0 - 2 + gy < a.L1;
        pxy[r] = ring < 4 && a.has_pml && (a.pmlflag[cx[r]] | a.pmlflag[a.L0 + cy[r]]);
        const int xm = cx[r] == 0 ? a.L0 - 1 : cx[r] - 1, ym = cy[r] == 0 ? a.L1 - 1 : cy[r] - 1;
        tH[r][0] = tx[cx[r]], tH[r][1] = tx[a.L0 + cx[r]], tH[r][2] = ty[cy[r]], tH[r][3] = ty[a.L1 + cy[r]];
        tE[r][0] = tx[2 * a.L0 + cx[r]], tE[r][1] = tx[3 * a.L0 + cx[r]];
        tE[r][2] = ty[2 * a.L1 + cy[r]], tE[r][3] = ty[3 * a.L1 + cy[r]];
        gm[r][0] = tx[3 * a.L0 + xm], gm[r][1] = ty[3 * a.L1 + ym];
    }
    /* z factors of every stage's plane (aH, gH, aE, gE) and its z-PML flag */
    for (int i = t; i < ns; i += NT) {
        const int z = plane_of(i);
        lds[L_ZT + 5 * i + 0] = tz[z];
        lds[L_ZT + 5 * i + 1] = tz[a.L2 + z];
        lds[L_ZT + 5 * i + 2] = tz[2 * a.L2 + z];
        lds[L_ZT + 5 * i + 3] = tz[3 * a.L2 + z];
        lds[L_ZT + 5 * i + 4] = a.has_pml && a.pmlflag[a.L0 + a.L1 + z] ? 1.f : 0.f;
    }
    /* does any cell of G4 lie in x/y PML? (its split arrays are then copied) */
    if (t == 0) lds[L_FLAG] = 0.f;
    __syncthreads();
    if (pxy[0] | pxy[1]) lds[L_FLAG] = 1.f;
    __syncthreads();
    const bool tile_pml = lds[L_FLAG] != 0.f;
    retire_loads(); /* the table loads, before the copies start */
    auto zt = [&](int s, int k) { return lds[L_ZT + 5 * s + k]; };
    auto zpml = [&](int s) { return zt(s, 4) != 0.f; };

    const __amdgpu_buffer_rsrc_t rT = group_rsrc(a.lT, a.ncomp), rB = group_rsrc(a.b, a.ncomp),
                                 rS = group_rsrc(a.lS, a.ncomp);
    const __amdgpu_buffer_rsrc_t rTn = group_rsrc(a.nT, a.ncomp), rSn = group_rsrc(a.nS, a.ncomp);
    const unsigned cb = 4u * (unsigned)a.ncomp;
    auto store = [&](__amdgpu_buffer_rsrc_t r, int c0, int64_t l, const float v[3]) {
        const unsigned off = 4u * (unsigned)l + (unsigned)c0 * cb;
        __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(v[0]), r, off, 0, 0);
        __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(v[1]), r, off + cb, 0, 0);
        __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(v[2]), r, off + 2 * cb, 0, 0);
    };
    auto issue = [&](int s, int z) {
        const unsigned soff = 4u * (unsigned)((int64_t)z * a.sz);
        const bool split = tile_pml || zpml(s);
        float *slot = lds + (s & 1) * SLOT;
        switch (wv) {
        case 0: adj2_issue<0>(slot, rT, rB, rS, cb, soff, split, vq, vr, vlast); break;
        case 1: adj2_issue<1>(slot, rT, rB, rS, cb, soff, split, vq, vr, vlast); break;
        case 2: adj2_issue<2>(slot, rT, rB, rS, cb, soff, split, vq, vr, vlast); break;
        default: adj2_issue<3>(slot, rT, rB, rS, cb, soff, split, vq, vr, vlast); break;
        }
    };
    /* stage arrays: 0 lambda_TE, 1 lambda_TH, 2 bH, 3 bE, 4 lambda_SH, 5 lambda_SE */
    auto st = [&](const float *S, int arr, int c, int e) { return S[arr * ARR + c * NC * 64 + e - row0(3 * arr + c) * W4]; };
    float *const UV1 = lds + L_UV1, *const UE1 = lds + L_UE1, *const UV2 = lds + L_UV2;

    issue(0, plane_of(0));
    int zi = plane_of(1);
    /* carried per cell: lambda_TE0, lambda_SE0, bE, bH of the previous plane; lambda_TH1,
     * lambda_SH1 of the previous plane; lambda_TE1, lambda_SE1 two planes back; the -z
     * values U1, V0 of the previous plane's two H halves and u1, v0 of its lambda_E1 */
    float e0p[2][3], se0p[2][3], bep[2][3], bhp[2][3], h1p[2][3], sh1p[2][3], e1pp[2][3], se1pp[2][3];
    float uv1p[2][2], ue1p[2][2], uv2p[2][2];
#pragma unroll
    for (int r = 0; r < 2; r++) {
#pragma unroll
        for (int c = 0; c < 3; c++)
            e0p[r][c] = se0p[r][c] = bep[r][c] = bhp[r][c] = h1p[r][c] = sh1p[r][c] = e1pp[r][c] = se1pp[r][c] = 0.f;
#pragma unroll
        for (int c = 0; c < 2; c++) uv1p[r][c] = ue1p[r][c] = uv2p[r][c] = 0.f;
    }
    int zq0 = plane_of(0), zq1 = plane_of(-1), zq2 = plane_of(-2);
    for (int s = 0; s < ns; s++) {
        const int p = z0 - 2 + s;
        const int zp = zq0, zp1 = zq1, zp2 = zq2;
        const bool doA = s >= 1, doB = p >= z0, doC = p >= z0 + 1, doD = p >= z0 + 2;
        if (s + 1 < ns) {
            issue(s + 1, zi);
            zi = zi + 1 == a.L2 ? 0 : zi + 1;
            wait_vm_barrier<kMinOps>();
        } else {
            wait_vm_barrier<0>();
        }
        const float *S = lds + (s & 1) * SLOT;
        const int s1 = s >= 1 ? s - 1 : 0, s2 = s >= 2 ? s - 2 : 0;
        const bool zs = zpml(s), zs1 = s >= 1 && zpml(s1), zs2 = s >= 2 && zpml(s2);
        const bool split = tile_pml || zs;
        const float aHz = zt(s, 0), gHz = zt(s, 1), gEz = zt(s, 3);
        const float aHz1 = zt(s1, 0), gHz1 = zt(s1, 1), aEz1 = zt(s1, 2), gEz1 = zt(s1, 3), aEz2 = zt(s2, 2);

        float uv1n[2][6], h1n[2][3], sh1n[2][3], e1n[2][3], se1n[2][3], ue1n[2][6], uv2n[2][6];
#pragma unroll
        for (int r = 0; r < 2; r++)
#pragma unroll
            for (int c = 0; c < 6; c++) {
                uv1n[r][c] = ue1n[r][c] = uv2n[r][c] = 0.f;
                if (c < 3) h1n[r][c] = sh1n[r][c] = e1n[r][c] = se1n[r][c] = 0.f;
            }
        /* a. lambda_yH1(p) on G3: gather of u, v of lambda_E0 (planes p, p-1) */
#pragma unroll
        for (int r = 0; r < 2; r++) {
            if (!doA) continue;
            const int e = ce[r], exm = e - 1, eym = e - W4;
            const bool pml = pxy[r] || zs;
            const float gEx = tE[r][1], gEy = tE[r][3];
            float lt[3], ls[3], be[3];
#pragma unroll
            for (int c = 0; c < 3; c++) {
                lt[c] = st(S, 0, c, e);
                be[c] = st(S, 3, c, e);
                ls[c] = split ? st(S, 5, c, e) : 0.f;
            }
            const float u0 = be[0] * gEy * lt[0], u1 = be[1] * gEz * lt[1], u2 = be[2] * gEx * lt[2];
            const float v0 = be[0] * gEz * (lt[0] + ls[0]), v1 = be[1] * gEx * (lt[1] + ls[1]);
            const float v2 = be[2] * gEy * (lt[2] + ls[2]);
            const float u2xm = st(S, 3, 2, exm) * gm[r][0] * st(S, 0, 2, exm);
            const float v1xm = st(S, 3, 1, exm) * gm[r][0] * (st(S, 0, 1, exm) + (split ? st(S, 5, 1, exm) : 0.f));
            const float u0ym = st(S, 3, 0, eym) * gm[r][1] * st(S, 0, 0, eym);
            const float v2ym = st(S, 3, 2, eym) * gm[r][1] * (st(S, 0, 2, eym) + (split ? st(S, 5, 2, eym) : 0.f));
            /* the -z neighbour: u_1, v_0 of plane p-1 */
            const float u1c = bep[r][1] * gEz1 * e0p[r][1];
            const float v0c = bep[r][0] * gEz1 * (e0p[r][0] + se0p[r][0]);
            const float wx = cx[r] == 0 ? a.w[0] : 1.f, wy = cy[r] == 0 ? a.w[1] : 1.f, wz = zp == 0 ? a.w[2] : 1.f;
            float ly[3], lsh[3], bh[3];
#pragma unroll
            for (int c = 0; c < 3; c++) {
                ly[c] = st(S, 1, c, e);
                bh[c] = st(S, 2, c, e);
                lsh[c] = split ? st(S, 4, c, e) : 0.f;
            }
            ly[0] += (-u1 + wz * u1c) + (v2 - wy * v2ym);
            ly[1] += (-u2 + wx * u2xm) + (v0 - wz * v0c);
            ly[2] += (-u0 + wy * u0ym) + (v1 - wx * v1xm);
            const float gHx = tH[r][1], gHy = tH[r][3];
            uv1n[r][0] = bh[0] * gHy * ly[0];
            uv1n[r][1] = bh[1] * gHz * ly[1];
            uv1n[r][2] = bh[2] * gHx * ly[2];
            uv1n[r][3] = bh[0] * gHz * (ly[0] + lsh[0]);
            uv1n[r][4] = bh[1] * gHx * (ly[1] + lsh[1]);
            uv1n[r][5] = bh[2] * gHy * (ly[2] + lsh[2]);
            const float aM[3] = {tH[r][2], aHz, tH[r][0]}, aS[3] = {aHz, tH[r][0], tH[r][2]};
            adj_own(ly, lsh, pml, aM, aS, h1n[r], sh1n[r]);
        }
        /* b. lambda_E1(p-1) on G2: own transform minus the gather of U1, V1 (planes p-1,
         * p); the injection on plane zT; u, v of lambda_E1 for level c */
#pragma unroll
        for (int r = 0; r < 2; r++) {
            if (!doB) continue;
            const bool pml = pxy[r] || zs1;
            const float aM[3] = {tE[r][2], aEz1, tE[r][0]}, aS[3] = {aEz1, tE[r][0], tE[r][2]};
            float lt[3], ls[3];
            adj_own(e0p[r], se0p[r], pml, aM, aS, lt, ls);
            const int k = rb[r];
            const float wx = cx[r] + 1 == a.L0 ? a.w[0] : 1.f, wy = cy[r] + 1 == a.L1 ? a.w[1] : 1.f;
            const float wz = zp1 + 1 == a.L2 ? a.w[2] : 1.f;
            lt[0] -= (uv1p[r][0] - wz * uv1n[r][1]) + (-UV1[B2 * N3 + k] + wy * UV1[B2 * N3 + k + W3]);
            lt[1] -= (UV1[A2 * N3 + k] - wx * UV1[A2 * N3 + k + 1]) + (-uv1p[r][1] + wz * uv1n[r][3]);
            lt[2] -= (UV1[A0 * N3 + k] - wy * UV1[A0 * N3 + k + W3]) + (-UV1[B1 * N3 + k] + wx * UV1[B1 * N3 + k + 1]);
            if (zp1 == a.zT) {
                const int64_t P = (int64_t)a.L0 * a.L1;
#pragma unroll
                for (int v = 0; v < 2; v++) {
                    const double *D = a.D + 2 * (v * P + cq[r]);
                    lt[v] += (float)(2. * (D[0] * a.cs + D[1] * a.sn));
                }
                retire_loads();
            }
            if (own[r] && p - 1 >= z0 && p - 1 < z1 && cx[r] >= a.box[0] && cx[r] < a.box[1] && cy[r] >= a.box[2] &&
                cy[r] < a.box[3] && zp1 >= a.box[4] && zp1 < a.box[5]) {
                const int64_t i = (cx[r] - a.box[0]) + (int64_t)a.bx * ((cy[r] - a.box[2]) + (int64_t)a.by * (zp1 - a.box[4]));
#pragma unroll
                for (int c = 0; c < 3; c++) a.lbox[c * a.nbox + i] = lt[c];
            }
            const float gEx = tE[r][1], gEy = tE[r][3];
            ue1n[r][0] = bep[r][0] * gEy * lt[0];
            ue1n[r][1] = bep[r][1] * gEz1 * lt[1];
            ue1n[r][2] = bep[r][2] * gEx * lt[2];
            ue1n[r][3] = bep[r][0] * gEz1 * (lt[0] + ls[0]);
            ue1n[r][4] = bep[r][1] * gEx * (lt[1] + ls[1]);
            ue1n[r][5] = bep[r][2] * gEy * (lt[2] + ls[2]);
            if (i2[r] >= 0) {
                UE1[A0 * N2 + i2[r]] = ue1n[r][0];
                UE1[A2 * N2 + i2[r]] = ue1n[r][2];
                UE1[B1 * N2 + i2[r]] = ue1n[r][4];
                UE1[B2 * N2 + i2[r]] = ue1n[r][5];
            }
#pragma unroll
            for (int c = 0; c < 3; c++) {
                e1n[r][c] = lt[c];
                se1n[r][c] = ls[c];
            }
        }
        wait_vm_barrier<63>(); /* u, v of lambda_E1(p-1) of the neighbours in LDS */
        /* c. lambda_yH2(p-1) on G1: gather of u, v of lambda_E1 (planes p-1, p-2) */
#pragma unroll
        for (int r = 0; r < 2; r++) {
            if (!doC) continue;
            const bool pml = pxy[r] || zs1;
            const int k = rc[r];
            const float u2xm = UE1[A2 * N2 + k - 1], v1xm = UE1[B1 * N2 + k - 1];
            const float u0ym = UE1[A0 * N2 + k - W2], v2ym = UE1[B2 * N2 + k - W2];
            const float wx = cx[r] == 0 ? a.w[0] : 1.f, wy = cy[r] == 0 ? a.w[1] : 1.f, wz = zp1 == 0 ? a.w[2] : 1.f;
            float ly[3];
#pragma unroll
            for (int c = 0; c < 3; c++) ly[c] = h1p[r][c];
            const float *u = ue1n[r]; /* u0 u1 u2 v0 v1 v2 */
            ly[0] += (-u[1] + wz * ue1p[r][0]) + (u[5] - wy * v2ym);
            ly[1] += (-u[2] + wx * u2xm) + (u[3] - wz * ue1p[r][1]);
            ly[2] += (-u[0] + wy * u0ym) + (u[4] - wx * v1xm);
            const float gHx = tH[r][1], gHy = tH[r][3];
            uv2n[r][0] = bhp[r][0] * gHy * ly[0];
            uv2n[r][1] = bhp[r][1] * gHz1 * ly[1];
            uv2n[r][2] = bhp[r][2] * gHx * ly[2];
            uv2n[r][3] = bhp[r][0] * gHz1 * (ly[0] + sh1p[r][0]);
            uv2n[r][4] = bhp[r][1] * gHx * (ly[1] + sh1p[r][1]);
            uv2n[r][5] = bhp[r][2] * gHy * (ly[2] + sh1p[r][2]);
            const float aM[3] = {tH[r][2], aHz1, tH[r][0]}, aS[3] = {aHz1, tH[r][0], tH[r][2]};
            float nth[3], nsh[3];
            adj_own(ly, sh1p[r], pml, aM, aS, nth, nsh);
            if (own[r] && p - 1 < z1) {
                store(rTn, 0, cq[r] + a.sz * zp1, nth);
                if (pml) store(rSn, 0, cq[r] + a.sz * zp1, nsh);
            }
        }
        /* d. lambda_E2(p-2) on T: own transform minus the gather of U2, V2 (planes p-2, p-1) */
        if (doD && own[0]) {
            const bool pml = pxy[0] || zs2;
            const float aM[3] = {tE[0][2], aEz2, tE[0][0]}, aS[3] = {aEz2, tE[0][0], tE[0][2]};
            float lt[3], ls[3];
            adj_own(e1pp[0], se1pp[0], pml, aM, aS, lt, ls);
            const int k = rd[0];
            const float wx = cx[0] + 1 == a.L0 ? a.w[0] : 1.f, wy = cy[0] + 1 == a.L1 ? a.w[1] : 1.f;
            const float wz = zp2 + 1 == a.L2 ? a.w[2] : 1.f;
            lt[0] -= (uv2p[0][0] - wz * uv2n[0][1]) + (-UV2[B2 * N1 + k] + wy * UV2[B2 * N1 + k + W1]);
            lt[1] -= (UV2[A2 * N1 + k] - wx * UV2[A2 * N1 + k + 1]) + (-uv2p[0][1] + wz * uv2n[0][3]);
            lt[2] -= (UV2[A0 * N1 + k] - wy * UV2[A0 * N1 + k + W1]) + (-UV2[B1 * N1 + k] + wx * UV2[B1 * N1 + k + 1]);
            store(rTn, 3, cq[0] + a.sz * zp2, lt);
            if (pml) store(rSn, 3, cq[0] + a.sz * zp2, ls);
        }
        /* carry to the next plane */
#pragma unroll
        for (int r = 0; r < 2; r++) {
#pragma unroll
            for (int c = 0; c < 3; c++) {
                e1pp[r][c] = e1n[r][c];
                se1pp[r][c] = se1n[r][c];
                h1p[r][c] = h1n[r][c];
                sh1p[r][c] = sh1n[r][c];
                bhp[r][c] = st(S, 2, c, ce[r]);
                bep[r][c] = st(S, 3, c, ce[r]);
                e0p[r][c] = st(S, 0, c, ce[r]);
                se0p[r][c] = split ? st(S, 5, c, ce[r]) : 0.f;
            }
            uv1p[r][0] = uv1n[r][1], uv1p[r][1] = uv1n[r][3];
            ue1p[r][0] = ue1n[r][1], ue1p[r][1] = ue1n[r][3];
            uv2p[r][0] = uv2n[r][1], uv2p[r][1] = uv2n[r][3];
        }
        zq2 = zq1;
        zq1 = zq0;
        zq0 = zq0 + 1 == a.L2 ? 0 : zq0 + 1;
        wait_vm_barrier<63>(); /* every wave is done with this stage and the compact arrays */
#pragma unroll
        for (int r = 0; r < 2; r++) {
            if (doA && i3[r] >= 0) {
                UV1[A0 * N3 + i3[r]] = uv1n[r][0];
                UV1[A2 * N3 + i3[r]] = uv1n[r][2];
                UV1[B1 * N3 + i3[r]] = uv1n[r][4];
                UV1[B2 * N3 + i3[r]] = uv1n[r][5];
            }
            if (doC && i1[r] >= 0) {
                UV2[A0 * N1 + i1[r]] = uv2n[r][0];
                UV2[A2 * N1 + i1[r]] = uv2n[r][2];
                UV2[B1 * N1 + i1[r]] = uv2n[r][4];
                UV2[B2 * N1 + i1[r]] = uv2n[r][5];
            }
        }
    }
}

/* G_c += lambda_TE^{k+1} (E^{k+1} - E^k), then += lambda_TE^k (E^k - E^{k-1}) (the order
 * of two adj_grad launches): lambda^{k+1} from the full group, lambda_TE^k from lbox. */
__global__ __launch_bounds__(kThreads) void adj_grad2(smx_adj_args a, const float *lTE, const float *lbox,
                                                      const float *kp1, const float *k0, const float *km1)
{
    const int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    if (i >= a.nbox) return;
    const int x = (int)(i % a.bx) + a.box[0];
    const int y = (int)((i / a.bx) % a.by) + a.box[2];
    const int z = (int)(i / ((int64_t)a.bx * a.by)) + a.box[4];
    const int64_t l = x + (int64_t)a.L0 * y + a.sz * z;
#pragma unroll
    for (int c = 0; c < 3; c++) {
        const int64_t j = c * a.nbox + i;
        const float d1 = kp1[j] - k0[j], d2 = k0[j] - km1[j];
        double g = a.G[j];
        g += (double)lTE[c * a.n + l] * (double)d1;
        g += (double)lbox[j] * (double)d2;
        a.G[j] = g;
    }
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
    const dim3 grid(a->n_pml + a->n_free);
    switch (a->tile) {
    case 1: hipLaunchKernelGGL((adj_fused_f32<64, 8>), grid, dim3(512), 0, s, *a); break;
    case 2: hipLaunchKernelGGL((adj_fused_f32<64, 4>), grid, dim3(256), 0, s, *a); break;
    default: hipLaunchKernelGGL((adj_fused_f32<32, 8>), grid, dim3(256), 0, s, *a); break;
    }
    return hipGetLastError();
}

hipError_t smx_adj_launch_two(const smx_adj_two_args *a, hipStream_t s)
{
    hipLaunchKernelGGL(adj_two_f32, dim3(a->nblocks), dim3(adj2::NT), 0, s, *a);
    return hipGetLastError();
}

hipError_t smx_adj_launch_grad2(const smx_adj_args *a, const float *lTE, const float *lbox, const float *kp1,
                                const float *k0, const float *km1, hipStream_t s)
{
    hipLaunchKernelGGL(adj_grad2, grid_for(a->nbox), dim3(kThreads), 0, s, *a, lTE, lbox, kp1, k0, km1);
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
