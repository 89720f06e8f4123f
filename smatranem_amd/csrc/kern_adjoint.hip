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

dim3 grid_for(int64_t count) { return dim3((unsigned)((count + kThreads - 1) / kThreads)); }

} // namespace

extern "C" {

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
