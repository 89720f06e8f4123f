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
 * Sample of a cell = the reference's main slot (total minus split/aux); in PML cells,
 * which the reference refuses on monitor planes, the total field.
 */
#include <hip/hip_runtime.h>

#include <stdint.h>

#include "smx_dft.h"

namespace {

constexpr int kThreads = 256;

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
    const bool pml = a.pmlflag && (a.pmlflag[x] | a.pmlflag[a.L0 + y] | a.pmlflag[a.L0 + a.L1 + z]);
    T *dst = (T *)a.series + ((int64_t)slot * a.nplanes + m) * 4 * per_plane + q;
    const int comp[4] = {0, 1, 3, 4};
#pragma unroll
    for (int v = 0; v < 4; v++) {
        const float t = a.T32[comp[v]][l];
        dst[v * per_plane] = pml ? t : t - a.S32[comp[v]][l];
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
 * h_v = 0 for H, 1/2 for E.  One thread per (m, v, q); the phase table of the chunk is
 * built once per block in LDS.
 */
template <typename T>
__global__ __launch_bounds__(kThreads) void fold_series(smx_dft_fold_args a)
{
    __shared__ double ph[2][SMX_DFT_MAX_CHUNK * SMX_DFT_MAX_FREQ][2]; /* [H/E][t*F + f][re, im] */
    const int F = a.nfreq, nt = a.count;
    for (int i = threadIdx.x; i < 2 * nt * F; i += kThreads) {
        const int e = i / (nt * F), r = i % (nt * F);
        const int t = r / F, f = r % F;
        const double arg = a.w[f] * ((double)(a.k0 + t + 1) + 0.5 * e);
        double s, c;
        sincos(arg, &s, &c);
        ph[e][r][0] = c;
        ph[e][r][1] = s;
    }
    __syncthreads();
    const int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    const int64_t per = a.plane_cells;
    if (i >= per * 4 * a.nplanes) return;
    const int64_t mv = i / per, q = i - mv * per; /* mv = m*4 + v */
    const int v = (int)(mv % 4);
    const int e = v >= 2;
    const T *src = (const T *)a.series + mv * per + q;
    const int64_t tstride = (int64_t)a.nplanes * 4 * per;
    double *acc = a.acc + (mv * F) * 2 * per + 2 * q;
    for (int f0 = 0; f0 < F; f0 += 8) {
        double re[8], im[8];
#pragma unroll
        for (int j = 0; j < 8; j++) re[j] = im[j] = 0.;
        for (int t = 0; t < nt; t++) {
            const double x = (double)src[t * tstride];
            const double(*p)[2] = &ph[e][t * F + f0];
#pragma unroll
            for (int j = 0; j < 8; j++) {
                if (f0 + j < F) {
                    re[j] += p[j][0] * x;
                    im[j] += p[j][1] * x;
                }
            }
        }
#pragma unroll
        for (int j = 0; j < 8; j++) {
            if (f0 + j < F) {
                double *o = acc + (int64_t)(f0 + j) * 2 * per;
                o[0] += re[j];
                o[1] += im[j];
            }
        }
    }
}

} // namespace

extern "C" hipError_t smx_dft_sample(const smx_dft_sample_args *a, int slot, hipStream_t stream)
{
    const int64_t total = a->plane_cells * a->nplanes;
    const dim3 grid((unsigned)((total + kThreads - 1) / kThreads)), block(kThreads);
    if (a->H64) hipLaunchKernelGGL(sample_planes_f64, grid, block, 0, stream, *a, slot);
    else hipLaunchKernelGGL(sample_planes_f32<float>, grid, block, 0, stream, *a, slot);
    return hipGetLastError();
}

extern "C" hipError_t smx_dft_fold(const smx_dft_fold_args *a, int f64, hipStream_t stream)
{
    if (a->count <= 0) return hipSuccess;
    const int64_t total = a->plane_cells * 4 * a->nplanes;
    const dim3 grid((unsigned)((total + kThreads - 1) / kThreads)), block(kThreads);
    if (f64) hipLaunchKernelGGL(fold_series<double>, grid, block, 0, stream, *a);
    else hipLaunchKernelGGL(fold_series<float>, grid, block, 0, stream, *a);
    return hipGetLastError();
}
