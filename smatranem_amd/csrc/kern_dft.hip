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

#include <cmath>
#include <cstddef>

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
 * h_v = 0 for H, 1/2 for E, computed as rot_f * sum_t tab[h_v][t][f] * series[t] with
 * tab = e^{i*w_f*(t + 1 + h)} (the plan's table, the same for every chunk) and
 * rot_f = e^{i*w_f*k0}.  One thread per (m, v, q), all frequencies in one pass over the
 * chunk (16 at a time); the table is copied to LDS once per block.
 */
template <typename T>
__global__ __launch_bounds__(kThreads) void fold_series(smx_dft_fold_args a)
{
    extern __shared__ double2 ph[]; /* [h][t][f], 2 * count * nfreq entries */
    const int F = a.nfreq, nt = a.count;
    for (int i = threadIdx.x; i < 2 * nt * F; i += kThreads) {
        const int h = i / (nt * F), r = i % (nt * F);
        ph[i] = a.table[h * a.table_nt * F + r];
    }
    __syncthreads();
    const int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    const int64_t per = a.plane_cells;
    if (i >= per * 4 * a.nplanes) return;
    const int64_t mv = i / per, q = i - mv * per; /* mv = m*4 + v */
    const int h = (int)(mv % 4) >= 2;
    const T *src = (const T *)a.series + mv * per + q;
    const int64_t tstride = (int64_t)a.nplanes * 4 * per;
    double *acc = a.acc + (mv * F) * 2 * per + 2 * q;
    const double2 *tab = ph + h * nt * F;
    for (int f0 = 0; f0 < F; f0 += 16) {
        double re[16], im[16];
#pragma unroll
        for (int j = 0; j < 16; j++) re[j] = im[j] = 0.;
        for (int t = 0; t < nt; t++) {
            const double x = (double)src[t * tstride];
            const double2 *p = tab + t * F + f0;
#pragma unroll
            for (int j = 0; j < 16; j++) {
                if (f0 + j < F) {
                    re[j] += p[j].x * x;
                    im[j] += p[j].y * x;
                }
            }
        }
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

} // namespace

void smx_dft_table(int nfreq, const double *w, int chunk, double *tab)
{
    for (int h = 0; h < 2; h++)
        for (int t = 0; t < chunk; t++)
            for (int f = 0; f < nfreq; f++) {
                const double arg = w[f] * ((double)(t + 1) + 0.5 * h);
                double *o = tab + 2 * (((size_t)h * chunk + t) * nfreq + f);
                o[0] = std::cos(arg);
                o[1] = std::sin(arg);
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

extern "C" hipError_t smx_dft_fold(const smx_dft_fold_args *a, int f64, hipStream_t stream)
{
    if (a->count <= 0) return hipSuccess;
    const int64_t total = a->plane_cells * 4 * a->nplanes;
    const dim3 grid((unsigned)((total + kThreads - 1) / kThreads)), block(kThreads);
    const size_t lds = sizeof(double2) * 2 * (size_t)a->count * (size_t)a->nfreq; /* <= 64 KiB */
    if (f64) hipLaunchKernelGGL(fold_series<double>, grid, block, lds, stream, *a);
    else hipLaunchKernelGGL(fold_series<float>, grid, block, lds, stream, *a);
    return hipGetLastError();
}
