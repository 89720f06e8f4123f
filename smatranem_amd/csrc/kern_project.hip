/*
 * kern_project.hip -- the mode projection of S_calc (FDTDcalc/FDTDsplitfieldMetal.c:437-445)
 * on the device, for the reference monitors and for every plane and frequency of the DFT.
 *
 *   S[v][m] = sum_q conj(t_m[q])     Hx_v[q] + conj(t_m[q + P])      Hy_v[q]
 *           + conj(t_m[q + 2P]) (-Ey_v[q]) + conj(t_m[q + 3P])      Ex_v[q]
 *
 * with t_m a column of tilXi (2NA = 4P complex rows) and v a set of phasors (the
 * monitors HB/EB or HT/ET, or one DFT plane at one frequency).  FP64 throughout.
 *
 * Stage 1 (project_partial): block (chunk, v) owns 256 consecutive q of vector v; each
 * thread loads its four vector entries once and, for every mode, forms its term and
 * joins a fixed-order reduction (xor butterfly in the wave, then the four waves in
 * order), writing partial[v][m][chunk].  The mode rows a wave reads are 64 consecutive
 * q: coalesced.  Stage 2 (project_sum): one thread per (v, m) adds the chunks in order.
 * The result does not depend on timing; it differs from the reference's sequential
 * sum only by rounding.
 */
#include <hip/hip_runtime.h>

#include <stdint.h>

#include "smx_project.h"

namespace {

constexpr int kThreads = 256;

__device__ inline double2 cmul_conj(double2 t, double2 x) /* conj(t) * x */
{
    return make_double2(t.x * x.x + t.y * x.y, t.x * x.y - t.y * x.x);
}

/* Phasor of component c (0 Hx, 1 Hy, 2 Ex, 3 Ey) of vector v at in-plane index q. */
__device__ inline double2 vec_at(const smx_project_args &a, int v, int c, int64_t q)
{
    if (v < a.nmon) {
        /* monitors: HB, EB (v = 0) or HT, ET (v = 1), complex[2P], (x, y) interleaved */
        const double2 *H = a.mon + (int64_t)(2 * v) * 2 * a.P, *E = a.mon + (int64_t)(2 * v + 1) * 2 * a.P;
        return c < 2 ? H[2 * q + c] : E[2 * q + (c - 2)];
    }
    /* DFT accumulators [plane][comp][freq][q] */
    const int d = v - a.nmon, m = d / a.nfreq, f = d % a.nfreq;
    return a.dft[(((int64_t)m * 4 + c) * a.nfreq + f) * a.P + q];
}

__device__ inline double wave_sum(double x)
{
#pragma unroll
    for (int s = 1; s < 64; s <<= 1) x += __shfl_xor(x, s);
    return x;
}

__global__ __launch_bounds__(kThreads) void project_partial(smx_project_args a)
{
    __shared__ double2 red[kThreads / 64];
    const int v = blockIdx.y, chunk = blockIdx.x;
    const int64_t q = (int64_t)chunk * kThreads + threadIdx.x;
    const bool in = q < a.P;
    double2 hx = make_double2(0., 0.), hy = hx, ex = hx, mey = hx;
    if (in) {
        hx = vec_at(a, v, 0, q);
        hy = vec_at(a, v, 1, q);
        ex = vec_at(a, v, 2, q);
        const double2 ey = vec_at(a, v, 3, q);
        mey = make_double2(-ey.x, -ey.y);
    }
    const int lane = threadIdx.x % 64, wave = threadIdx.x / 64;
    const int64_t rows = 4 * a.P;
    for (int m = 0; m < a.nmodes; m++) {
        double2 acc = make_double2(0., 0.);
        if (in) {
            const double2 *t = a.modes + (int64_t)m * rows;
            double2 p = cmul_conj(t[q], hx);
            acc.x += p.x;
            acc.y += p.y;
            p = cmul_conj(t[q + a.P], hy);
            acc.x += p.x;
            acc.y += p.y;
            p = cmul_conj(t[q + 2 * a.P], mey);
            acc.x += p.x;
            acc.y += p.y;
            p = cmul_conj(t[q + 3 * a.P], ex);
            acc.x += p.x;
            acc.y += p.y;
        }
        acc.x = wave_sum(acc.x);
        acc.y = wave_sum(acc.y);
        if (lane == 0) red[wave] = acc;
        __syncthreads();
        if (threadIdx.x == 0) {
            double2 s = red[0];
            for (int w = 1; w < kThreads / 64; w++) {
                s.x += red[w].x;
                s.y += red[w].y;
            }
            a.partial[((int64_t)v * a.nmodes + m) * a.nchunks + chunk] = s;
        }
        __syncthreads();
    }
}

__global__ __launch_bounds__(kThreads) void project_sum(smx_project_args a)
{
    const int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    if (i >= (int64_t)a.nvec * a.nmodes) return;
    const double2 *p = a.partial + i * a.nchunks;
    double2 s = make_double2(0., 0.);
    for (int c = 0; c < a.nchunks; c++) {
        s.x += p[c].x;
        s.y += p[c].y;
    }
    a.S[i] = s;
}

} // namespace

extern "C" int smx_project_chunks(int64_t P) { return (int)((P + kThreads - 1) / kThreads); }

extern "C" hipError_t smx_project_launch(const smx_project_args *a, hipStream_t stream)
{
    if (a->nvec <= 0 || a->nmodes <= 0) return hipSuccess;
    hipLaunchKernelGGL(project_partial, dim3(a->nchunks, a->nvec), dim3(kThreads), 0, stream, *a);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const int64_t total = (int64_t)a->nvec * a->nmodes;
    hipLaunchKernelGGL(project_sum, dim3((unsigned)((total + kThreads - 1) / kThreads)), dim3(kThreads), 0, stream,
                       *a);
    return hipGetLastError();
}
