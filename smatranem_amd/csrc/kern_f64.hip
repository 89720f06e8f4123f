/*
 * kern_f64.hip -- plain FP64 split-field step on the reference storage contract.
 *
 * One thread per cell, two sweeps (H then E) as two launches, the E launch also
 * accumulating the monitor planes.  The per-cell arithmetic keeps the reference's
 * operation order (main_calc and ope_*, FDTDcalc/FDTDsplitfieldMetal.c:242-413) with
 * floating-point contraction off, so results are bit-identical to the reference
 * compiled with gcc -O2 on x86-64.  This is the parity path; the FP32 kernels in
 * kern_f32.hip are the fast path.
 *
 * Launch shape: block (64, 4) over (j0, j1), grid z = j2.  Each wavefront covers 64
 * consecutive cells of one x-row, so the component-interleaved loads of a wavefront
 * fall in one contiguous 1.5 KiB span.
 */
#include "smx_internal.h"

#pragma clang fp contract(off)

namespace {

constexpr int BX = 64, BY = 4;

struct nbrs {
    int64_t off[3];
    double w[3];
};

/* j-1 with wrap; weight iBc at j == 0 (F:396-402). */
__device__ inline nbrs backward_nbrs(const smx_geom &g, int j0, int j1, int j2)
{
    const int64_t sy = g.L0, sz = (int64_t)g.L0 * g.L1;
    nbrs b;
    b.off[0] = j0 == 0 ? (int64_t)(g.L0 - 1) : -1;
    b.off[1] = j1 == 0 ? (int64_t)(g.L1 - 1) * sy : -sy;
    b.off[2] = j2 == 0 ? (int64_t)(g.L2 - 1) * sz : -sz;
    b.w[0] = j0 == 0 ? (double)g.iBc0 : 1.;
    b.w[1] = j1 == 0 ? (double)g.iBc1 : 1.;
    b.w[2] = j2 == 0 ? (double)g.iBc2 : 1.;
    return b;
}

/* j+1 with wrap; weight iBc at j == L-1 (F:375-381). */
__device__ inline nbrs forward_nbrs(const smx_geom &g, int j0, int j1, int j2)
{
    const int64_t sy = g.L0, sz = (int64_t)g.L0 * g.L1;
    nbrs f;
    f.off[0] = j0 + 1 == g.L0 ? -(int64_t)(g.L0 - 1) : 1;
    f.off[1] = j1 + 1 == g.L1 ? -(int64_t)(g.L1 - 1) * sy : sy;
    f.off[2] = j2 + 1 == g.L2 ? -(int64_t)(g.L2 - 1) * sz : sz;
    f.w[0] = j0 + 1 == g.L0 ? (double)g.iBc0 : 1.;
    f.w[1] = j1 + 1 == g.L1 ? (double)g.iBc1 : 1.;
    f.w[2] = j2 + 1 == g.L2 ? (double)g.iBc2 : 1.;
    return f;
}

/* Main-slot value of component c at cell q, or the split slot through addr (map). */
template <bool MAP>
__device__ inline double val(const double *P, const int *addr, int c, int64_t q)
{
    if (MAP) return P[c + 3 * (int64_t)addr[q]];
    return P[c + 3 * q];
}

/* Backward curl (ope_Rd / ope_Rdmap).  HALF: 0 '+', 1 '-', 2 both. */
template <bool MAP>
__device__ inline void curl_bwd(int half, const nbrs &b, const double *P, const int *addr,
                                int64_t l, double W[3])
{
#pragma unroll
    for (int c = 0; c < 3; c++) {
        const int c1 = c == 2 ? 0 : c + 1, c2 = c == 0 ? 2 : c - 1;
        double acc = 0.;
        if (half != 1) acc += val<MAP>(P, addr, c2, l) - b.w[c1] * val<MAP>(P, addr, c2, l + b.off[c1]);
        if (half != 0) acc += -val<MAP>(P, addr, c1, l) + b.w[c2] * val<MAP>(P, addr, c1, l + b.off[c2]);
        W[c] = acc;
    }
}

/* Forward curl (ope_R / ope_Rmap). */
template <bool MAP>
__device__ inline void curl_fwd(int half, const nbrs &f, const double *P, const int *addr,
                                int64_t l, double W[3])
{
#pragma unroll
    for (int c = 0; c < 3; c++) {
        const int c1 = c == 2 ? 0 : c + 1, c2 = c == 0 ? 2 : c - 1;
        double acc = 0.;
        if (half != 1) acc += f.w[c1] * val<MAP>(P, addr, c2, l + f.off[c1]) - val<MAP>(P, addr, c2, l);
        if (half != 0) acc += -f.w[c2] * val<MAP>(P, addr, c1, l + f.off[c2]) + val<MAP>(P, addr, c1, l);
        W[c] = acc;
    }
}

/* Source plane value (x: v=0, y: v=1) at in-plane index q: explicit, or
 * f2*Re(conj(e)*inc) = f2*(e_re*inc_re + e_im*inc_im) (F:156-157). */
__device__ inline double source_at(const double *plane, const double *inc, double f2, double er,
                                   double ei, int64_t q, int v)
{
    if (plane) return plane[v + 2 * q];
    const double *z = inc + 2 * (v + 2 * q);
    return f2 * (er * z[0] + ei * z[1]);
}

__global__ __launch_bounds__(BX *BY) void h_sweep_f64(smx_f64_args a)
{
    const smx_geom &g = a.g;
    const int j0 = blockIdx.x * BX + threadIdx.x, j1 = blockIdx.y * BY + threadIdx.y, j2 = blockIdx.z;
    if (j0 >= g.L0 || j1 >= g.L1) return;
    const int64_t q = j0 + (int64_t)g.L0 * j1, l = q + (int64_t)g.L0 * g.L1 * j2;
    const char k = a.Cont[l];
    if (k == 'e') return; /* H frozen in epsilon-metal cells */
    const nbrs b = backward_nbrs(g, j0, j1, j2);
    double *H = a.H, *Hs = a.H + 3 * g.n;
    const double *E = a.E, *Es = a.E + 3 * g.n;
    double W[3];
    if (k == 'n') {
        curl_bwd<false>(2, b, E, nullptr, l, W);
        if (j2 == g.iI) { /* F:261-264 */
            const double sx = source_at(a.srcE_plane, a.Einc, a.s.srcE_f2, a.s.srcE_re, a.s.srcE_im, q, 0);
            const double sy = source_at(a.srcE_plane, a.Einc, a.s.srcE_f2, a.s.srcE_re, a.s.srcE_im, q, 1);
            W[0] += g.sgn * sy;
            W[1] -= g.sgn * sx;
        }
#pragma unroll
        for (int c = 0; c < 3; c++) H[c + 3 * l] = a.Ah[c + 3 * l] * H[c + 3 * l] + a.Bh[c + 3 * l] * (-W[c]);
        curl_bwd<true>(2, b, Es, a.addr, l, W);
#pragma unroll
        for (int c = 0; c < 3; c++) H[c + 3 * l] += a.Bh[c + 3 * l] * (-W[c]);
    } else if (k == 'a') { /* F:268-276 */
        for (int half = 0; half < 2; half++) {
            const int64_t s = half == 0 ? 3 * l : 3 * g.n + 3 * (int64_t)a.addr[l];
            curl_bwd<false>(half, b, E, nullptr, l, W);
#pragma unroll
            for (int c = 0; c < 3; c++) H[s + c] = a.Ah[s + c] * H[s + c] + a.Bh[s + c] * (-W[c]);
            curl_bwd<true>(half, b, Es, a.addr, l, W);
#pragma unroll
            for (int c = 0; c < 3; c++) H[s + c] += a.Bh[s + c] * (-W[c]);
        }
    } else { /* 'h', 'b': stable form (F:277-283) */
        const int64_t s = 3 * g.n + 3 * (int64_t)a.addr[l];
        curl_bwd<false>(2, b, E, nullptr, l, W);
#pragma unroll
        for (int c = 0; c < 3; c++) {
            double h = H[c + 3 * l];
            h += 0.5 * H[s + c];
            const double aux = a.Ah[s + c] * H[s + c] + a.Bh[s + c] * (-W[c]);
            H[s + c] = aux;
            H[c + 3 * l] = h + 0.5 * aux;
        }
    }
    (void)Hs;
}

__global__ __launch_bounds__(BX *BY) void e_sweep_f64(smx_f64_args a)
{
    const smx_geom &g = a.g;
    const int j0 = blockIdx.x * BX + threadIdx.x, j1 = blockIdx.y * BY + threadIdx.y, j2 = blockIdx.z;
    if (j0 >= g.L0 || j1 >= g.L1) return;
    const int64_t q = j0 + (int64_t)g.L0 * j1, l = q + (int64_t)g.L0 * g.L1 * j2;
    const char k = a.Cont[l];
    double *E = a.E;
    const double *H = a.H, *Hs = a.H + 3 * g.n;
    double W[3];
    if (k == 'n') {
        const nbrs f = forward_nbrs(g, j0, j1, j2);
        curl_fwd<false>(2, f, H, nullptr, l, W);
        if (j2 == g.iI) { /* F:293-296 */
            const double sx = source_at(a.srcH_plane, a.Hinc, a.s.srcH_f2, a.s.srcH_re, a.s.srcH_im, q, 0);
            const double sy = source_at(a.srcH_plane, a.Hinc, a.s.srcH_f2, a.s.srcH_re, a.s.srcH_im, q, 1);
            W[0] += g.sgn * sy;
            W[1] -= g.sgn * sx;
        }
#pragma unroll
        for (int c = 0; c < 3; c++) E[c + 3 * l] = a.Ae[c + 3 * l] * E[c + 3 * l] + a.Be[c + 3 * l] * W[c];
        curl_fwd<true>(2, f, Hs, a.addr, l, W);
#pragma unroll
        for (int c = 0; c < 3; c++) E[c + 3 * l] += a.Be[c + 3 * l] * W[c];
    } else if (k == 'a') { /* F:300-307 */
        const nbrs f = forward_nbrs(g, j0, j1, j2);
        for (int half = 0; half < 2; half++) {
            const int64_t s = half == 0 ? 3 * l : 3 * g.n + 3 * (int64_t)a.addr[l];
            curl_fwd<false>(half, f, H, nullptr, l, W);
#pragma unroll
            for (int c = 0; c < 3; c++) E[s + c] = a.Ae[s + c] * E[s + c] + a.Be[s + c] * W[c];
            curl_fwd<true>(half, f, Hs, a.addr, l, W);
#pragma unroll
            for (int c = 0; c < 3; c++) E[s + c] += a.Be[s + c] * W[c];
        }
    } else if (k != 'h') { /* 'e', 'b': stable form with the BACKWARD curl of H (F:308-314) */
        const nbrs b = backward_nbrs(g, j0, j1, j2);
        const int64_t s = 3 * g.n + 3 * (int64_t)a.addr[l];
        curl_bwd<false>(2, b, H, nullptr, l, W);
#pragma unroll
        for (int c = 0; c < 3; c++) {
            double e = E[c + 3 * l];
            e += 0.5 * E[s + c];
            const double aux = a.Ae[s + c] * E[s + c] + a.Be[s + c] * W[c];
            E[s + c] = aux;
            E[c + 3 * l] = e + 0.5 * aux;
        }
    }
    /* Monitors: main slots after both sweeps (F:318-328).  This thread alone writes
     * E at l and H at l is final, so no other synchronisation is needed. */
    if (j2 == g.iB || j2 == g.iT) {
        double hv[2], ev[2];
#pragma unroll
        for (int v = 0; v < 2; v++) {
            hv[v] = H[v + 3 * l];
            ev[v] = E[v + 3 * l];
        }
#pragma unroll
        for (int m = 0; m < 2; m++) {
            if (j2 != (m == 0 ? g.iB : g.iT)) continue;
            double *HM = m == 0 ? a.HB : a.HT, *EM = m == 0 ? a.EB : a.ET;
#pragma unroll
            for (int v = 0; v < 2; v++) {
                const int64_t o = 2 * (v + 2 * q);
                HM[o] += a.s.monH_re * hv[v];
                HM[o + 1] += a.s.monH_im * hv[v];
                EM[o] += a.s.monE_re * ev[v];
                EM[o + 1] += a.s.monE_im * ev[v];
            }
        }
    }
}

} // namespace

extern "C" hipError_t smx_launch_f64_step(const smx_f64_args *a, hipStream_t stream)
{
    const dim3 block(BX, BY, 1);
    const dim3 grid((a->g.L0 + BX - 1) / BX, (a->g.L1 + BY - 1) / BY, a->g.L2);
    hipLaunchKernelGGL(h_sweep_f64, grid, block, 0, stream, *a);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(e_sweep_f64, grid, block, 0, stream, *a);
    return hipGetLastError();
}
