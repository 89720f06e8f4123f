/*
 * kern_dft.hip -- running DFT of the tangential fields on port planes at many
 * frequencies (the extension of the reference's single-frequency monitors,
 * FDTDcalc/FDTDsplitfieldMetal.c:318-328).
 *
 * The reference adds e^{i*dphi*(t+1)}*H and e^{i*dphi*(t+1.5)}*E to its monitors after
 * every step.  Done per step at F frequencies that is a read-modify-write of
 * 2 planes x 4 components x F complex FP64 accumulators per step.  Here each step only
 * appends the plane samples to a ring buffer (sample_planes: 4 words per plane cell),
 * and every `chunk` steps fold_series multiplies the chunk by the phase table of every
 * frequency and adds it to the FP64 accumulators: the accumulator traffic drops by the
 * chunk length.  Step k of the run (k from 0) is sampled with phase
 * e^{i*w*(k+1)} for H and e^{i*w*(k+1.5)} for E, w = omega_f*Dt; at w = 2*pi/N_2pi this
 * is the reference's monitor.
 *
 * Sample of a cell = the reference's main slot (total minus split/aux); in class 'a'
 * (PML) cells, which the reference refuses on monitor planes, the total field.
 */
#include <hip/hip_runtime.h>

#include <stdint.h>

#include <cmath>
#include <cstddef>

#include "smx_dft.h"

namespace {

constexpr int kThreads = 256;

/* The point source's term of the step (smx_dft_sample_args::pt_amp), rounded as the
 * oracle rounds it: no contraction into a fused multiply-add. */
__device__ inline void point_term(const smx_dft_sample_args &a)
{
#pragma clang fp contract(off)
    if (a.pt_dst64) {
        const double t = a.pt_coef64[0] * a.pt_amp[a.pt_k];
        a.pt_dst64[0] = a.pt_dst64[0] + t;
    } else {
        const float t = a.pt_coef32[0] * (float)a.pt_amp[a.pt_k];
        a.pt_dst32[0] = a.pt_dst32[0] + t;
    }
}

__device__ inline bool point_thread(const smx_dft_sample_args &a, int64_t i, int m, int64_t q)
{
    return a.pt_amp && (a.pt_m < 0 ? i == 0 : m == a.pt_m && q == a.pt_q);
}

/* series layout: [t][plane][comp][q], comp = Hx, Hy, Ex, Ey */
template <typename T>
__global__ __launch_bounds__(kThreads) void sample_planes_f32(smx_dft_sample_args a, int slot)
{
    const int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    const int64_t per_plane = a.plane_cells;
    if (i >= per_plane * a.nplanes) return;
    const int m = (int)(i / per_plane);
    const int64_t q = i - m * per_plane;
    const int64_t l = q + per_plane * a.planes[m];
    const int x = (int)(q % a.L0), y = (int)(q / a.L0), z = a.planes[m];
    if (point_thread(a, i, m, q)) point_term(a); /* before this thread's own samples */
    if (a.mask_on) {
        auto in_mod = [](int v, int lo, int len, int L) { return ((v - lo) % L + L) % L < len; };
        if (!(in_mod(x, a.mask[0], a.mask[1], a.L0) && in_mod(y, a.mask[2], a.mask[3], a.L1) &&
              in_mod(z, a.mask[4], a.mask[5], a.L2)))
            return;
    }
    /* the total field in class 'a' cells, as sample_planes_f64: the separable PML flag
     * for regular cells, the table's class for lossy and metal cells (a metal cell
     * inside a PML layer is class 'e'/'h'/'b', F:188-233, and samples its main slot) */
    bool total = a.pmlflag && (a.pmlflag[x] | a.pmlflag[a.L0 + y] | a.pmlflag[a.L0 + a.L1 + z]);
    const float b0 = a.bH0[l];
    if (__builtin_isnan(b0)) total = a.special[(int64_t)__float_as_uint(a.bH1[l]) * a.special_stride] == 'a';
    T *dst = (T *)a.series + ((int64_t)slot * a.nplanes + m) * 4 * per_plane + q;
    const int comp[4] = {0, 1, 3, 4};
#pragma unroll
    for (int v = 0; v < 4; v++) {
        const float t = a.T32[comp[v]][l];
        dst[v * per_plane] = total ? t : t - a.S32[comp[v]][l];
    }
}

__global__ __launch_bounds__(kThreads) void sample_planes_f64(smx_dft_sample_args a, int slot)
{
    const int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    const int64_t per_plane = a.plane_cells;
    if (i >= per_plane * a.nplanes) return;
    const int m = (int)(i / per_plane);
    const int64_t q = i - m * per_plane;
    const int64_t l = q + per_plane * a.planes[m];
    if (point_thread(a, i, m, q)) point_term(a); /* before this thread's own samples */
    double *dst = (double *)a.series + ((int64_t)slot * a.nplanes + m) * 4 * per_plane + q;
    double v[4] = {a.H64[0 + 3 * l], a.H64[1 + 3 * l], a.E64[0 + 3 * l], a.E64[1 + 3 * l]};
    if (a.cont[l] == 'a') { /* PML: main + split = total field */
        const int64_t s = 3 * a.n + 3 * (int64_t)a.addr[l];
        v[0] += a.H64[s + 0];
        v[1] += a.H64[s + 1];
        v[2] += a.E64[s + 0];
        v[3] += a.E64[s + 1];
    }
#pragma unroll
    for (int k = 0; k < 4; k++) dst[k * per_plane] = v[k];
}

/*
 * acc[m][v][f][q] += sum_t e^{i*w_f*(k0 + t + 1 + h_v)} * series[t][m][v][q],
 * h_v = 0 for H, 1/2 for E, computed as rot_f * sum_t tab[h_v][t][f] * series[t] with
 * tab = e^{i*w_f*(t + 1 + h)} (the plan's table, the same for every chunk) and
 * rot_f = e^{i*w_f*k0}.  One thread per (m, v, q), all frequencies in one pass over the
 * chunk (16 at a time).  Grid: x over q, y over (m, v), so the table row a wave reads
 * is the same in every lane: it is read through the scalar cache (constant address
 * space), and its rows are padded with zeros to a multiple of 16 frequencies, so the
 * inner loop has no per-frequency test.  The samples are loaded 8 slots ahead of use.
 */
typedef const double2 __attribute__((address_space(4))) *const_table_ptr;

template <typename T>
__device__ inline void fold_slot(double (&re)[16], double (&im)[16], const_table_ptr p, double x)
{
#pragma unroll
    for (int j = 0; j < 16; j++) {
        re[j] += p[j].x * x;
        im[j] += p[j].y * x;
    }
}

template <typename T>
__global__ __launch_bounds__(kThreads) void fold_series(smx_dft_fold_args a)
{
    const int F = a.nfreq, nt = a.count, FR = smx_dft_table_row(a.nfreq);
    const int64_t per = a.plane_cells;
    const int64_t q = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    if (q >= per) return;
    const int mv = blockIdx.y; /* m*4 + v */
    const int h = (mv % 4) >= 2;
    const T *src = (const T *)a.series + mv * per + q;
    const int64_t tstride = (int64_t)a.ring * 4 * per;
    double *acc = a.acc + ((int64_t)mv * F) * 2 * per + 2 * q;
    const const_table_ptr tab = (const_table_ptr)a.table + h * a.table_nt * FR;
    constexpr int kAhead = 8;
    const int nt8 = nt - nt % kAhead;
    for (int f0 = 0; f0 < F; f0 += 16) {
        double re[16], im[16];
#pragma unroll
        for (int j = 0; j < 16; j++) re[j] = im[j] = 0.;
        for (int t0 = 0; t0 < nt8; t0 += kAhead) {
            T xs[kAhead];
#pragma unroll
            for (int u = 0; u < kAhead; u++) xs[u] = src[(int64_t)(t0 + u) * tstride];
#pragma unroll
            for (int u = 0; u < kAhead; u++) fold_slot<T>(re, im, tab + (t0 + u) * FR + f0, (double)xs[u]);
        }
        for (int t = nt8; t < nt; t++) fold_slot<T>(re, im, tab + t * FR + f0, (double)src[(int64_t)t * tstride]);
#pragma unroll
        for (int j = 0; j < 16; j++) {
            if (f0 + j < F) {
                const double cr = a.rot[f0 + j][0], ci = a.rot[f0 + j][1];
                double *o = acc + (int64_t)(f0 + j) * 2 * per;
                o[0] += cr * re[j] - ci * im[j];
                o[1] += cr * im[j] + ci * re[j];
            }
        }
    }
}

/* One thread per (monitor, component, q); grid y = 2 monitors x 4 components. */
template <typename T>
__global__ __launch_bounds__(kThreads) void fold_monitors(smx_dft_mon_args a)
{
    const int64_t per = a.plane_cells;
    const int64_t q = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    const int m = blockIdx.y / 4, c = blockIdx.y % 4;
    if (q >= per || a.mon[m] < 0) return;
    const int h = c >= 2; /* E: the half-step phase */
    const T *src = (const T *)a.series + ((int64_t)a.mon[m] * 4 + c) * per + q;
    const int64_t tstride = (int64_t)a.ring * 4 * per;
    const const_table_ptr ph = (const_table_ptr)a.expn + h;
    double re = 0., im = 0.;
    int j = a.j0;
    auto add = [&](int t, double x) {
        if ((a.mask[t / 64] >> (t % 64)) & 1u) {
            re += ph[2 * j].x * x;
            im += ph[2 * j].y * x;
        }
        if (++j == a.N) j = 0;
    };
    /* 16 slots loaded before any is used: one memory latency per 16 slots */
    constexpr int kAhead = 16;
    const int n16 = a.count - a.count % kAhead;
    for (int t0 = 0; t0 < n16; t0 += kAhead) {
        T xs[kAhead];
#pragma unroll
        for (int u = 0; u < kAhead; u++) xs[u] = src[(int64_t)(t0 + u) * tstride];
#pragma unroll
        for (int u = 0; u < kAhead; u++) add(t0 + u, (double)xs[u]);
    }
    for (int t = n16; t < a.count; t++) add(t, (double)src[(int64_t)t * tstride]);
    double *dst = a.monitors + (int64_t)(2 * m + h) * 4 * per + 2 * ((c % 2) + 2 * q);
    dst[0] += re;
    dst[1] += im;
}

} // namespace

void smx_dft_table(int nfreq, const double *w, int chunk, double *tab)
{
    const int row = smx_dft_table_row(nfreq);
    for (int h = 0; h < 2; h++)
        for (int t = 0; t < chunk; t++)
            for (int f = 0; f < row; f++) {
                const double arg = f < nfreq ? w[f] * ((double)(t + 1) + 0.5 * h) : 0.;
                double *o = tab + 2 * (((size_t)h * chunk + t) * row + f);
                o[0] = f < nfreq ? std::cos(arg) : 0.;
                o[1] = f < nfreq ? std::sin(arg) : 0.;
            }
}

extern "C" hipError_t smx_dft_sample(const smx_dft_sample_args *a, int slot, hipStream_t stream)
{
    const int64_t total = a->plane_cells * a->nplanes;
    const dim3 grid((unsigned)((total + kThreads - 1) / kThreads)), block(kThreads);
    if (a->H64) hipLaunchKernelGGL(sample_planes_f64, grid, block, 0, stream, *a, slot);
    else hipLaunchKernelGGL(sample_planes_f32<float>, grid, block, 0, stream, *a, slot);
    return hipGetLastError();
}

extern "C" hipError_t smx_dft_fold_monitors(const smx_dft_mon_args *a, int f64, hipStream_t stream)
{
    if (a->count <= 0 || (a->mon[0] < 0 && a->mon[1] < 0)) return hipSuccess;
    const dim3 grid((unsigned)((a->plane_cells + kThreads - 1) / kThreads), 8), block(kThreads);
    if (f64) hipLaunchKernelGGL(fold_monitors<double>, grid, block, 0, stream, *a);
    else hipLaunchKernelGGL(fold_monitors<float>, grid, block, 0, stream, *a);
    return hipGetLastError();
}

extern "C" hipError_t smx_dft_fold(const smx_dft_fold_args *a, int f64, hipStream_t stream)
{
    if (a->count <= 0) return hipSuccess;
    const dim3 grid((unsigned)((a->plane_cells + kThreads - 1) / kThreads), (unsigned)(4 * a->nplanes)),
        block(kThreads);
    if (f64) hipLaunchKernelGGL(fold_series<double>, grid, block, 0, stream, *a);
    else hipLaunchKernelGGL(fold_series<float>, grid, block, 0, stream, *a);
    return hipGetLastError();
}
