/*
 * smx_plan.hip -- plan lifetime, device memory, the per-step driver and the C ABI
 * entry points of libsmx_fdtd.so (declared in include/smx_fdtd.h).
 *
 * A plan owns one grid's device state.  The host computes the per-step scalars of
 * the reference's time loop (envelope, phases; FDTDcalc/FDTDsplitfieldMetal.c:153-167)
 * and launches the sweeps on the plan's stream; no host/device copy happens inside
 * smx_plan_run.
 */
#include <hip/hip_runtime.h>

#include <climits>
#include <cmath>
#include <cstddef>
#include <cstdio>
#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <new>
#include <vector>

#include "smx_fdtd.h"
#include "smx_internal.h"
#include "smx_adjoint.h"
#include "smx_dft.h"
#include "smx_f32.h"
#include "smx_project.h"

namespace {

thread_local char g_last_error[512] = "";

int hip_fail(hipError_t e, const char *what)
{
    snprintf(g_last_error, sizeof g_last_error, "%s: %s", what, hipGetErrorString(e));
    return e == hipErrorOutOfMemory ? SMX_ERR_OOM : SMX_ERR_HIP;
}

#define SMX_HIP(call)                                                                    \
    do {                                                                                 \
        hipError_t e_ = (call);                                                          \
        if (e_ != hipSuccess) return hip_fail(e_, #call);                                \
    } while (0)

int arg_fail(const char *what)
{
    snprintf(g_last_error, sizeof g_last_error, "invalid argument: %s", what);
    return SMX_ERR_ARG;
}

} // namespace

struct smx_plan {
    smx_grid g{};
    int64_t n = 0;        /* cells */
    int64_t plane = 0;    /* L0*L1 */
    int64_t flen = 0;     /* 3*(n+Lnotn+1), one reference field array */
    int device = 0;
    hipStream_t stream = nullptr;

    /* F64: reference storage contract */
    double *Ah = nullptr, *Bh = nullptr, *Ae = nullptr, *Be = nullptr;
    double *H = nullptr, *E = nullptr;
    char *Cont = nullptr;
    int *addr = nullptr;

    /* F32: packed layout (kern_f32.hip) */
    smx_f32_state f32{};

    /* shared: monitors (4 x complex[2*plane]), source profiles and explicit planes */
    double *mon = nullptr;
    double *inc = nullptr; /* Hinc then Einc, complex[2*plane] each */
    double *src = nullptr; /* Hi then Ei, double[2*plane] each */

    bool have_source = false;
    int N = 0;
    int tp[3] = {0, 0, 0};
    std::vector<double> expn; /* complex[2N] */
    long long next_step = 0;  /* run step index of the next smx_plan_step */

    /* multi-frequency port DFT (kern_dft.hip) */
    int dft_nplanes = 0, dft_nfreq = 0, dft_chunk = 0, dft_fill = 0;
    int dft_planes[SMX_DFT_MAX_PLANES] = {0};
    double dft_w[SMX_DFT_MAX_FREQ] = {0};
    long long dft_k0 = 0;
    void *dft_series = nullptr;
    double *dft_acc = nullptr;
    double *dft_table = nullptr; /* [2][chunk][row] complex phases (smx_dft_table) */
    /* FP32 plans with the DFT on take the reference monitors from the same ring: its slots
     * hold the DFT planes, then the monitor planes not among them (dft_ring planes in
     * all); dft_mon = ring plane of each monitor (-1: the kernels accumulate the
     * monitors per step, as FP64 plans always do); a slot's monitors are folded at
     * flush time when its bit in dft_mon_mask is set (steps run by smx_plan_step keep
     * the caller's phases and accumulate in the kernel). */
    int dft_ring = 0;
    int dft_mon[2] = {-1, -1};
    uint64_t dft_mon_mask[SMX_DFT_MAX_CHUNK / 64] = {0};
    double *expn_dev = nullptr; /* the phase table on the device */

    /* adjoint (kern_adjoint.hip): checkpoints of the forward run, adjoint fields */
    long long adj_K = 0;
    int adj_C = 0, adj_zT = 0, adj_box[6] = {0};
    double adj_w = 0.;
    bool adj_ready = false;
    float *adj_ckpt = nullptr; /* [ceil(K/C)][12][n]: T[6], S[6] at steps 0, C, 2C, ... */
    float *adj_hist = nullptr; /* C = 0: [K+1][3][nbox], E of the design box after every step */
    double *adj_D = nullptr;   /* [2][L0*L1] complex */

    /* point current source (smx_plan_set_point_source): E component (-1: none), cell,
     * amplitudes of steps [0, pt_n) on the device */
    int pt_comp = -1;
    int64_t pt_cell = 0;
    long long pt_n = 0;
    double *pt_amp = nullptr;
};

namespace {

smx_geom make_geom(const smx_plan *p)
{
    smx_geom g;
    g.L0 = p->g.L[0];
    g.L1 = p->g.L[1];
    g.L2 = p->g.L[2];
    g.iBc0 = p->g.iBc[0];
    g.iBc1 = p->g.iBc[1];
    g.iBc2 = p->g.iBc[2];
    g.iB = p->g.iBTIin[0];
    g.iT = p->g.iBTIin[1];
    g.iI = p->g.iBTIin[2];
    g.sgn = (double)(p->g.iBTIin[3] / std::abs(p->g.iBTIin[3])); /* F:253 */
    g.n = p->n;
    g.Lnotn = p->g.Lnotn;
    return g;
}

int validate_grid(const smx_grid *g, const char *Cont)
{
    for (int d = 0; d < 3; d++)
        if (g->L[d] < 1) return arg_fail("L[0..2] must be >= 1");
    for (int d = 0; d < 3; d++)
        if (g->iBc[d] < -1 || g->iBc[d] > 1) return arg_fail("iBc must be -1, 0 or 1");
    if (g->iBTIin[3] == 0) return arg_fail("incident mode must be nonzero");
    /* a monitor plane of -1 switches that monitor off (an extension: with PML on the
     * x/y faces every plane holds PML cells, which the reference refuses, F:320) */
    for (int m = 0; m < 2; m++)
        if (g->iBTIin[m] < -1 || g->iBTIin[m] >= g->L[2]) return arg_fail("monitor plane outside the grid");
    if (g->precision != SMX_F64 && g->precision != SMX_F32) return arg_fail("precision");
    const int64_t n = (int64_t)g->L[0] * g->L[1] * g->L[2];
    int64_t count = 0;
    for (int64_t l = 0; l < n; l++) {
        const char k = Cont[l];
        if (k != 'n' && k != 'a' && k != 'e' && k != 'h' && k != 'b') return arg_fail("Cont holds an unknown class");
        count += k != 'n';
    }
    if (count != g->Lnotn) return arg_fail("Lnotn differs from the count of non-'n' cells");
    /* The reference exits at the first step when a monitor plane holds a PML cell
     * (F:320); the class map never changes, so the plan refuses it up front. */
    const int64_t plane = (int64_t)g->L[0] * g->L[1];
    for (int m = 0; m < 2; m++)
        for (int64_t q = 0; q < plane && g->iBTIin[m] >= 0; q++)
            if (Cont[q + plane * g->iBTIin[m]] == 'a') {
                snprintf(g_last_error, sizeof g_last_error, "iBTIin[%d] error!", m);
                return SMX_ERR_MONITOR_PML;
            }
    return SMX_OK;
}

int alloc_common(smx_plan *p)
{
    SMX_HIP(hipGetDevice(&p->device));
    SMX_HIP(hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking));
    const size_t plane = (size_t)p->plane;
    SMX_HIP(hipMalloc(&p->mon, sizeof(double) * 16 * plane));
    SMX_HIP(hipMemsetAsync(p->mon, 0, sizeof(double) * 16 * plane, p->stream));
    SMX_HIP(hipMalloc(&p->inc, sizeof(double) * 8 * plane));
    SMX_HIP(hipMemsetAsync(p->inc, 0, sizeof(double) * 8 * plane, p->stream));
    SMX_HIP(hipMalloc(&p->src, sizeof(double) * 4 * plane));
    return SMX_OK;
}

int alloc_f64(smx_plan *p, const double *Ah, const double *Bh, const double *Ae, const double *Be,
              const char *Cont, const int *addr)
{
    const size_t clen = 3 * (size_t)(p->n + p->g.Lnotn), flen = (size_t)p->flen, n = (size_t)p->n;
    double **coef[4] = {&p->Ah, &p->Bh, &p->Ae, &p->Be};
    const double *host[4] = {Ah, Bh, Ae, Be};
    for (int i = 0; i < 4; i++) {
        SMX_HIP(hipMalloc(coef[i], sizeof(double) * clen));
        SMX_HIP(hipMemcpyAsync(*coef[i], host[i], sizeof(double) * clen, hipMemcpyHostToDevice, p->stream));
    }
    SMX_HIP(hipMalloc(&p->H, sizeof(double) * flen));
    SMX_HIP(hipMalloc(&p->E, sizeof(double) * flen));
    SMX_HIP(hipMemsetAsync(p->H, 0, sizeof(double) * flen, p->stream));
    SMX_HIP(hipMemsetAsync(p->E, 0, sizeof(double) * flen, p->stream));
    SMX_HIP(hipMalloc(&p->Cont, n));
    SMX_HIP(hipMemcpyAsync(p->Cont, Cont, n, hipMemcpyHostToDevice, p->stream));
    SMX_HIP(hipMalloc(&p->addr, sizeof(int) * n));
    SMX_HIP(hipMemcpyAsync(p->addr, addr, sizeof(int) * n, hipMemcpyHostToDevice, p->stream));
    SMX_HIP(hipStreamSynchronize(p->stream));
    return SMX_OK;
}

smx_plan *new_plan(const smx_grid *g)
{
    smx_plan *p = new (std::nothrow) smx_plan();
    if (!p) return nullptr;
    p->g = *g;
    p->n = (int64_t)g->L[0] * g->L[1] * g->L[2];
    p->plane = (int64_t)g->L[0] * g->L[1];
    p->flen = 3 * (p->n + g->Lnotn + 1);
    return p;
}

void free_plan(smx_plan *p)
{
    if (!p) return;
    if (p->stream) (void)hipStreamSynchronize(p->stream);
    void *bufs[] = {p->Ah, p->Bh, p->Ae, p->Be, p->H, p->E, p->Cont, p->addr, p->mon, p->inc, p->src,
                    p->dft_series, p->dft_acc, p->dft_table, p->adj_ckpt, p->adj_D, p->expn_dev, p->adj_hist,
                    p->pt_amp};
    for (void *b : bufs)
        if (b) (void)hipFree(b);
    smx_f32_free(&p->f32);
    if (p->stream) (void)hipStreamDestroy(p->stream);
    delete p;
}

int check_device()
{
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count < 1) {
        snprintf(g_last_error, sizeof g_last_error, "no HIP device: %s", hipGetErrorString(e));
        return SMX_ERR_NO_DEVICE;
    }
    return SMX_OK;
}

/* Scalars of step (P, t) of the loop at F:153-167. */
int step_scalars(const smx_plan *p, int P, int t, smx_step_scalars *s, double *envH, double *envE)
{
    const int N = p->N;
    const double fE = smx_set_env(N, p->tp, 1, t, P);
    const double fH = smx_set_env(N, p->tp, 0, t + 1, P);
    if (std::isnan(fE) || std::isnan(fH)) {
        snprintf(g_last_error, sizeof g_last_error, "f_env error!");
        return SMX_ERR_ENVELOPE;
    }
    const double *e = p->expn.data();
    const int tn = (t + 1) % N;
    s->srcE_f2 = (2.) * fE;
    s->srcE_re = e[4 * t + 2];
    s->srcE_im = e[4 * t + 3];
    s->srcH_f2 = (2.) * fH;
    s->srcH_re = e[4 * tn + 0];
    s->srcH_im = e[4 * tn + 1];
    s->monH_re = e[4 * tn + 0];
    s->monH_im = e[4 * tn + 1];
    s->monE_re = e[4 * tn + 2];
    s->monE_im = e[4 * tn + 3];
    *envH = fH;
    *envE = fE;
    return SMX_OK;
}

smx_f64_args f64_args(smx_plan *p)
{
    smx_f64_args a;
    a.g = make_geom(p);
    a.Ah = p->Ah;
    a.Bh = p->Bh;
    a.Ae = p->Ae;
    a.Be = p->Be;
    a.H = p->H;
    a.E = p->E;
    a.Cont = p->Cont;
    a.addr = p->addr;
    a.srcH_plane = nullptr;
    a.srcE_plane = nullptr;
    a.Hinc = p->inc;
    a.Einc = p->inc + 4 * p->plane;
    a.HB = p->mon;
    a.EB = p->mon + 4 * p->plane;
    a.HT = p->mon + 8 * p->plane;
    a.ET = p->mon + 12 * p->plane;
    std::memset(&a.s, 0, sizeof a.s);
    return a;
}

/* Fold the DFT samples held in the ring into the accumulators. */
int dft_flush(smx_plan *p)
{
    if (!p->dft_nplanes || p->dft_fill == 0) return SMX_OK;
    smx_dft_fold_args f;
    std::memset(&f, 0, sizeof f);
    f.nplanes = p->dft_nplanes;
    f.nfreq = p->dft_nfreq;
    f.count = p->dft_fill;
    f.plane_cells = p->plane;
    f.table_nt = p->dft_chunk;
    f.table = reinterpret_cast<const double2 *>(p->dft_table);
    for (int i = 0; i < p->dft_nfreq; i++) {
        const double arg = p->dft_w[i] * (double)p->dft_k0;
        f.rot[i][0] = std::cos(arg);
        f.rot[i][1] = std::sin(arg);
    }
    f.ring = p->dft_ring;
    f.series = p->dft_series;
    f.acc = p->dft_acc;
    SMX_HIP(smx_dft_fold(&f, p->g.precision == SMX_F64, p->stream));
    bool any = false;
    for (uint64_t w : p->dft_mon_mask) any |= w != 0;
    if (any) {
        smx_dft_mon_args m;
        std::memset(&m, 0, sizeof m);
        m.ring = p->dft_ring;
        m.mon[0] = p->dft_mon[0];
        m.mon[1] = p->dft_mon[1];
        m.count = p->dft_fill;
        m.N = p->N;
        m.j0 = (int)((p->dft_k0 + 1) % p->N);
        m.plane_cells = p->plane;
        std::memcpy(m.mask, p->dft_mon_mask, sizeof m.mask);
        m.expn = p->expn_dev;
        m.series = p->dft_series;
        m.monitors = p->mon;
        SMX_HIP(smx_dft_fold_monitors(&m, p->g.precision == SMX_F64, p->stream));
        std::memset(p->dft_mon_mask, 0, sizeof p->dft_mon_mask);
    }
    p->dft_fill = 0;
    return SMX_OK;
}

bool mon_ring(const smx_plan *p) { return p->dft_mon[0] >= 0 || p->dft_mon[1] >= 0; }

void mark_mon_slot(smx_plan *p, int slot) { p->dft_mon_mask[slot / 64] |= (uint64_t)1 << (slot % 64); }

/* Sample the ring's planes after run step k (steps must come in order within a chunk);
 * mon: the ring carries this step's monitors. */
/* The ring plane that holds the point source's cell (-1: none), or -2 when several do
 * (the sample kernel then cannot apply the source's term before every sample of it). */
int point_ring_plane(const smx_plan *p)
{
    const int z = (int)(p->pt_cell / p->plane);
    int m = -1;
    for (int i = 0; i < p->dft_ring; i++)
        if (p->dft_planes[i] == z) m = m == -1 ? i : -2;
    return m;
}

/* Whether the DFT sample kernel applies the point source's term of each step (one launch
 * less per step): with a DFT set and the source's cell on at most one ring plane. */
bool point_in_sample(const smx_plan *p)
{
    return p->pt_comp >= 0 && p->dft_nplanes > 0 && point_ring_plane(p) != -2;
}

int dft_after_step(smx_plan *p, long long k, bool mon)
{
    if (!p->dft_nplanes) return SMX_OK;
    int rc;
    if (p->dft_fill > 0 && k != p->dft_k0 + p->dft_fill && (rc = dft_flush(p))) return rc;
    if (p->dft_fill == 0) p->dft_k0 = k;
    smx_dft_sample_args a;
    std::memset(&a, 0, sizeof a);
    a.L0 = p->g.L[0];
    a.L1 = p->g.L[1];
    a.L2 = p->g.L[2];
    a.nplanes = p->dft_ring;
    std::memcpy(a.planes, p->dft_planes, sizeof a.planes);
    a.plane_cells = p->plane;
    if (p->g.precision == SMX_F64) {
        a.H64 = p->H;
        a.E64 = p->E;
        a.cont = p->Cont;
        a.addr = p->addr;
        a.n = p->n;
    } else {
        for (int i = 0; i < 6; i++) {
            a.T32[i] = p->f32.T[p->f32.cur][i];
            a.S32[i] = p->f32.S[p->f32.cur][i];
        }
        a.pmlflag = p->f32.pmlflag;
        a.bH0 = p->f32.b[0];
        a.bH1 = p->f32.b[1];
        static_assert(offsetof(smx_special_f32, cls) == 0 && sizeof(smx_special_f32) % sizeof(int) == 0,
                      "the DFT sample reads the class as the first int of a special-table entry");
        a.special = reinterpret_cast<const int *>(p->f32.special);
        a.special_stride = (int)(sizeof(smx_special_f32) / sizeof(int));
    }
    a.series = p->dft_series;
    if (point_in_sample(p) && k >= 0 && k < p->pt_n) {
        a.pt_amp = (const double *)p->pt_amp;
        a.pt_k = k;
        a.pt_m = point_ring_plane(p);
        a.pt_q = p->pt_cell % p->plane;
        if (p->g.precision == SMX_F64) {
            const int64_t i = p->pt_comp + 3 * p->pt_cell;
            a.pt_dst64 = p->E + i;
            a.pt_coef64 = p->Be + i;
        } else {
            const int c = 3 + p->pt_comp;
            a.pt_dst32 = p->f32.T[p->f32.cur][c] + p->pt_cell;
            a.pt_coef32 = p->f32.b[c] + p->pt_cell;
        }
    }
    SMX_HIP(smx_dft_sample(&a, p->dft_fill, p->stream));
    if (mon) mark_mon_slot(p, p->dft_fill);
    if (++p->dft_fill == p->dft_chunk) return dft_flush(p);
    return SMX_OK;
}

/* The point source's term after run step k: E += B * amp[k] on the source's component,
 * one thread; a separate multiply and add (no contraction into a fused multiply-add),
 * as the oracle rounds it. */
__global__ void point_add_f64(double *E, const double *B, const double *amp, long long k)
{
#pragma clang fp contract(off)
    const double t = B[0] * amp[k];
    E[0] = E[0] + t;
}

__global__ void point_add_f32(float *T, const float *b, const double *amp, long long k)
{
#pragma clang fp contract(off)
    const float t = b[0] * (float)amp[k];
    T[0] = T[0] + t;
}

int point_after_step(smx_plan *p, long long k)
{
    if (p->pt_comp < 0 || k < 0 || k >= p->pt_n || point_in_sample(p)) return SMX_OK;
    if (p->g.precision == SMX_F64) {
        const int64_t i = p->pt_comp + 3 * p->pt_cell;
        hipLaunchKernelGGL(point_add_f64, dim3(1), dim3(1), 0, p->stream, p->E + i, p->Be + i,
                           (const double *)p->pt_amp, k);
    } else {
        const int c = 3 + p->pt_comp;
        hipLaunchKernelGGL(point_add_f32, dim3(1), dim3(1), 0, p->stream, p->f32.T[p->f32.cur][c] + p->pt_cell,
                           (const float *)p->f32.b[c] + p->pt_cell, (const double *)p->pt_amp, k);
    }
    SMX_HIP(hipGetLastError());
    return SMX_OK;
}

/* One step; mon_kernel = false leaves the monitors to the ring (FP32 only). */
int launch_step(smx_plan *p, const smx_step_scalars &s, bool explicit_src, bool mon_kernel = true)
{
    if (p->g.precision == SMX_F64) {
        smx_f64_args a = f64_args(p);
        a.s = s;
        if (explicit_src) {
            a.srcH_plane = p->src;
            a.srcE_plane = p->src + 2 * p->plane;
        }
        SMX_HIP(smx_launch_f64_step(&a, p->stream));
        return SMX_OK;
    }
    smx_f32_launch_args a{};
    a.g = make_geom(p);
    if (!mon_kernel) a.g.iB = a.g.iT = -1;
    a.s = s;
    a.srcH_plane = explicit_src ? p->src : nullptr;
    a.srcE_plane = explicit_src ? p->src + 2 * p->plane : nullptr;
    a.Hinc = p->inc;
    a.Einc = p->inc + 4 * p->plane;
    a.HB = p->mon;
    a.EB = p->mon + 4 * p->plane;
    a.HT = p->mon + 8 * p->plane;
    a.ET = p->mon + 12 * p->plane;
    SMX_HIP(smx_f32_step(&p->f32, &a, p->stream));
    return SMX_OK;
}

} // namespace

int smx_check_grid(const smx_grid *g, const char *Cont) { return validate_grid(g, Cont); }
int smx_note_hip(hipError_t e, const char *what) { return hip_fail(e, what); }
int smx_note_arg(const char *what) { return arg_fail(what); }

/* ==================================================================================== */
/* C ABI                                                                                */
/* ==================================================================================== */

extern "C" {

const char *smx_status_string(int status)
{
    switch (status) {
    case SMX_OK: return "ok";
    case SMX_ERR_ARG: return "invalid argument";
    case SMX_ERR_HIP: return "HIP runtime error";
    case SMX_ERR_NO_DEVICE: return "no HIP device";
    case SMX_ERR_OOM: return "device out of memory";
    case SMX_ERR_STATE: return "call out of order";
    case SMX_ERR_MONITOR_PML: return "monitor plane holds a PML cell";
    case SMX_ERR_ENVELOPE: return "step beyond the stop period";
    default: return "unknown status";
    }
}

const char *smx_last_error(void) { return g_last_error; }

int smx_plan_create_coeffs(const smx_grid *grid, const double *Ah, const double *Bh,
                           const double *Ae, const double *Be, const char *Cont,
                           const int *address_notn, smx_plan **out)
{
    if (!grid || !Ah || !Bh || !Ae || !Be || !Cont || !address_notn || !out) return arg_fail("null pointer");
    *out = nullptr;
    if (grid->precision != SMX_F64) return arg_fail("coefficient plans are FP64 only");
    int rc = validate_grid(grid, Cont);
    if (rc) return rc;
    if ((rc = check_device())) return rc;
    smx_plan *p = new_plan(grid);
    if (!p) return SMX_ERR_OOM;
    if ((rc = alloc_common(p)) || (rc = alloc_f64(p, Ah, Bh, Ae, Be, Cont, address_notn))) {
        free_plan(p);
        return rc;
    }
    *out = p;
    return SMX_OK;
}

int smx_plan_create(const smx_grid *grid, const double *mu, const double *eps, const double *sigma,
                    const char *Cont, const int *address_notn, smx_plan **out)
{
    if (!grid || !mu || !eps || !sigma || !Cont || !address_notn || !out) return arg_fail("null pointer");
    *out = nullptr;
    int rc = validate_grid(grid, Cont);
    if (rc) return rc;
    if ((rc = check_device())) return rc;
    smx_plan *p = new_plan(grid);
    if (!p) return SMX_ERR_OOM;
    if ((rc = alloc_common(p))) {
        free_plan(p);
        return rc;
    }
    if (grid->precision == SMX_F64) {
        const size_t clen = 3 * (size_t)(p->n + grid->Lnotn);
        std::vector<double> A_h(clen, 0.), B_h(clen, 0.), A_e(clen, 0.), B_e(clen, 0.);
        smx_build_coeffs(grid->L, mu, eps, sigma, Cont, address_notn, grid->Lnotn, grid->invtau_R,
                         grid->invtau_I, A_h.data(), B_h.data(), A_e.data(), B_e.data());
        rc = alloc_f64(p, A_h.data(), B_h.data(), A_e.data(), B_e.data(), Cont, address_notn);
    } else {
        smx_geom g = make_geom(p);
        rc = smx_f32_create(&p->f32, &g, grid->invtau_R, grid->invtau_I, mu, eps, sigma, Cont,
                            address_notn, p->stream);
        if (rc == SMX_ERR_HIP || rc == SMX_ERR_OOM)
            snprintf(g_last_error, sizeof g_last_error, "%s", smx_f32_last_error());
    }
    if (rc) {
        free_plan(p);
        return rc;
    }
    *out = p;
    return SMX_OK;
}

int smx_plan_destroy(smx_plan *plan)
{
    free_plan(plan);
    return SMX_OK;
}

size_t smx_plan_field_len(const smx_plan *plan) { return plan ? (size_t)plan->flen : 0; }

int smx_plan_write_fields(smx_plan *p, const double *H_t, const double *E_t)
{
    if (!p || !H_t || !E_t) return arg_fail("null pointer");
    if (p->g.precision == SMX_F64) {
        SMX_HIP(hipMemcpyAsync(p->H, H_t, sizeof(double) * p->flen, hipMemcpyHostToDevice, p->stream));
        SMX_HIP(hipMemcpyAsync(p->E, E_t, sizeof(double) * p->flen, hipMemcpyHostToDevice, p->stream));
        SMX_HIP(hipStreamSynchronize(p->stream));
        return SMX_OK;
    }
    return smx_f32_write_fields(&p->f32, H_t, E_t, p->stream);
}

int smx_plan_read_fields(const smx_plan *p, double *H_t, double *E_t)
{
    if (!p || !H_t || !E_t) return arg_fail("null pointer");
    if (p->g.precision == SMX_F64) {
        SMX_HIP(hipMemcpyAsync(H_t, p->H, sizeof(double) * p->flen, hipMemcpyDeviceToHost, p->stream));
        SMX_HIP(hipMemcpyAsync(E_t, p->E, sizeof(double) * p->flen, hipMemcpyDeviceToHost, p->stream));
        SMX_HIP(hipStreamSynchronize(p->stream));
        return SMX_OK;
    }
    return smx_f32_read_fields(&p->f32, H_t, E_t, p->stream);
}

int smx_plan_read_monitors(const smx_plan *p, double *HB, double *EB, double *HT, double *ET)
{
    if (!p || !HB || !EB || !HT || !ET) return arg_fail("null pointer");
    /* fold the ring's pending samples first: the monitors they carry belong to steps
     * already run (folding early changes only the order of the sums) */
    int rc = dft_flush(const_cast<smx_plan *>(p));
    if (rc) return rc;
    double *dst[4] = {HB, EB, HT, ET};
    for (int i = 0; i < 4; i++)
        SMX_HIP(hipMemcpyAsync(dst[i], p->mon + 4 * p->plane * i, sizeof(double) * 4 * p->plane,
                               hipMemcpyDeviceToHost, p->stream));
    SMX_HIP(hipStreamSynchronize(p->stream));
    return SMX_OK;
}

int smx_plan_write_monitors(smx_plan *p, const double *HB, const double *EB, const double *HT,
                            const double *ET)
{
    if (!p) return arg_fail("null pointer");
    int rc = dft_flush(p); /* pending monitor samples belong to the values being replaced */
    if (rc) return rc;
    const double *src[4] = {HB, EB, HT, ET};
    for (int i = 0; i < 4; i++) {
        double *d = p->mon + 4 * p->plane * i;
        if (src[i]) SMX_HIP(hipMemcpyAsync(d, src[i], sizeof(double) * 4 * p->plane, hipMemcpyHostToDevice, p->stream));
        else SMX_HIP(hipMemsetAsync(d, 0, sizeof(double) * 4 * p->plane, p->stream));
    }
    SMX_HIP(hipStreamSynchronize(p->stream));
    return SMX_OK;
}

int smx_plan_step(smx_plan *p, const double *Hi, const double *Ei, const double *ExpN2)
{
    if (!p || !Hi || !Ei || !ExpN2) return arg_fail("null pointer");
    SMX_HIP(hipMemcpyAsync(p->src, Hi, sizeof(double) * 2 * p->plane, hipMemcpyHostToDevice, p->stream));
    SMX_HIP(hipMemcpyAsync(p->src + 2 * p->plane, Ei, sizeof(double) * 2 * p->plane, hipMemcpyHostToDevice,
                           p->stream));
    smx_step_scalars s;
    std::memset(&s, 0, sizeof s);
    s.monH_re = ExpN2[0];
    s.monH_im = ExpN2[1];
    s.monE_re = ExpN2[2];
    s.monE_im = ExpN2[3];
    int rc = launch_step(p, s, true);
    if (rc || (rc = point_after_step(p, p->next_step)) || (rc = dft_after_step(p, p->next_step, false))) return rc;
    p->next_step++;
    SMX_HIP(hipStreamSynchronize(p->stream));
    return SMX_OK;
}

int smx_plan_set_source(smx_plan *p, const double *Hinc, const double *Einc, int N_2pi,
                        const int time_period[3])
{
    if (!p || !Hinc || !Einc || !time_period) return arg_fail("null pointer");
    if (N_2pi < 1) return arg_fail("N_2pi must be >= 1");
    SMX_HIP(hipMemcpyAsync(p->inc, Hinc, sizeof(double) * 4 * p->plane, hipMemcpyHostToDevice, p->stream));
    SMX_HIP(hipMemcpyAsync(p->inc + 4 * p->plane, Einc, sizeof(double) * 4 * p->plane, hipMemcpyHostToDevice,
                           p->stream));
    int rc = dft_flush(p); /* pending monitor samples use the old phase table */
    if (rc) return rc;
    SMX_HIP(hipStreamSynchronize(p->stream));
    p->N = N_2pi;
    std::memcpy(p->tp, time_period, sizeof p->tp);
    p->expn.assign(4 * (size_t)N_2pi, 0.);
    smx_expn_table(N_2pi, p->expn.data());
    if (p->expn_dev) SMX_HIP(hipFree(p->expn_dev));
    p->expn_dev = nullptr;
    SMX_HIP(hipMalloc(&p->expn_dev, sizeof(double) * p->expn.size()));
    SMX_HIP(hipMemcpy(p->expn_dev, p->expn.data(), sizeof(double) * p->expn.size(), hipMemcpyHostToDevice));
    p->have_source = true;
    return SMX_OK;
}

/* Launch arguments of a two-step pass without its step scalars. */
static smx_f32_launch_args step2_args(smx_plan *p, bool mon_kernel)
{
    smx_f32_launch_args a{};
    a.g = make_geom(p);
    if (!mon_kernel) a.g.iB = a.g.iT = -1;
    a.srcH_plane = nullptr;
    a.srcE_plane = nullptr;
    a.Hinc = p->inc;
    a.Einc = p->inc + 4 * p->plane;
    a.HB = p->mon;
    a.EB = p->mon + 4 * p->plane;
    a.HT = p->mon + 8 * p->plane;
    a.ET = p->mon + 12 * p->plane;
    return a;
}

static int launch_step2(smx_plan *p, const smx_step_scalars &s, const smx_step_scalars &s2, int dft_slot,
                        bool mon_kernel, float *box_out = nullptr)
{
    smx_f32_launch_args a = step2_args(p, mon_kernel);
    a.s = s;
    a.s2 = s2;
    if (box_out) {
        a.box_out = box_out;
        std::memcpy(a.box, p->adj_box, sizeof a.box);
    }
    if (dft_slot >= 0) {
        a.dft = (float *)p->dft_series;
        a.dft_n = p->dft_ring;
        a.dft_slot = dft_slot;
        std::memcpy(a.dft_planes, p->dft_planes, sizeof(int) * (size_t)p->dft_ring);
    }
    SMX_HIP(smx_f32_step2(&p->f32, &a, p->stream));
    return SMX_OK;
}

/* Whether steps k and k+1 may run as one two-step pass: an FP32 plan with the two-step
 * kernel selected, and with the DFT on only where that kernel samples every cell and
 * the ring holds two slots. */
static bool pair_ok(const smx_plan *p)
{
    if (p->g.precision != SMX_F32 || p->f32.two < 0 || p->pt_comp >= 0) return false;
    return p->dft_nplanes == 0 || (smx_f32_two_full(&p->f32, p->f32.two) && p->dft_chunk >= 2);
}

/* The same for the pass that also writes E of the design box (the split-slot variant,
 * whatever kernel the plan selected). */
static bool pair_box_ok(const smx_plan *p)
{
    return p->g.precision == SMX_F32 && p->pt_comp < 0 && smx_f32_box_fz(&p->f32) > 0 && (p->dft_nplanes == 0 || p->dft_chunk >= 2);
}

/* Ring slot of the pair (k, k+1) of DFT samples, flushing the ring first where the
 * pair does not continue its chunk or does not fit. */
static int dft_pair_slot(smx_plan *p, long long k, int *slot)
{
    int rc;
    if (p->dft_fill > 0 && (k != p->dft_k0 + p->dft_fill || p->dft_fill + 2 > p->dft_chunk) &&
        (rc = dft_flush(p)))
        return rc;
    if (p->dft_fill == 0) p->dft_k0 = k;
    *slot = p->dft_fill;
    return SMX_OK;
}

static int save_box(smx_plan *p, float *dst);

/* Step k of the source loop, or steps k and k+1 in one pass when the plan's two-step
 * kernel is in use (pair_ok) and both steps are to run; *done = steps taken.  sums get
 * the envelope values in the same order as step-by-step.  box_out (may be null): E of
 * the design box after step k to box_out, after step k+1 to box_out + 3*nbox. */
static int advance(smx_plan *p, long long k, long long nleft, double *sums, int *done, float *box_out = nullptr)
{
    smx_step_scalars s;
    double fH, fE;
    int rc = step_scalars(p, (int)(k / p->N), (int)(k % p->N), &s, &fH, &fE);
    if (rc) return rc;
    if ((box_out ? pair_box_ok(p) : pair_ok(p)) && nleft >= 2) {
        smx_step_scalars s2;
        double fH2, fE2;
        if (step_scalars(p, (int)((k + 1) / p->N), (int)((k + 1) % p->N), &s2, &fH2, &fE2) == SMX_OK) {
            if (sums) {
                sums[0] += fH;
                sums[1] += fE;
                sums[0] += fH2;
                sums[1] += fE2;
            }
            *done = 2;
            int slot = -1;
            if (p->dft_nplanes && (rc = dft_pair_slot(p, k, &slot))) return rc;
            if ((rc = launch_step2(p, s, s2, slot, !mon_ring(p), box_out))) return rc;
            if (p->dft_nplanes) {
                if (mon_ring(p)) {
                    mark_mon_slot(p, slot);
                    mark_mon_slot(p, slot + 1);
                }
                p->dft_fill += 2;
                if (p->dft_fill == p->dft_chunk) return dft_flush(p);
            }
            return SMX_OK;
        }
    }
    if (sums) {
        sums[0] += fH;
        sums[1] += fE;
    }
    *done = 1;
    if ((rc = launch_step(p, s, false, !mon_ring(p))) || (rc = point_after_step(p, k))) return rc;
    if (box_out && (rc = save_box(p, box_out))) return rc;
    return dft_after_step(p, k, mon_ring(p));
}

int smx_plan_run_steps(smx_plan *p, long long step0, long long nsteps, double sums[2])
{
    if (!p) return arg_fail("null pointer");
    if (!p->have_source) {
        snprintf(g_last_error, sizeof g_last_error, "smx_plan_run before smx_plan_set_source");
        return SMX_ERR_STATE;
    }
    if (step0 < 0 || nsteps < 0) return arg_fail("negative step");
    for (long long k = step0; k < step0 + nsteps;) {
        int done = 0;
        const int rc = advance(p, k, step0 + nsteps - k, sums, &done);
        if (rc) return rc;
        k += done;
        p->next_step = k;
    }
    p->next_step = step0 + nsteps;
    return SMX_OK;
}

} // extern "C"

/* ---------------------------------------------------------------------------------- */
/* Batch replay through a HIP graph                                                    */
/* ---------------------------------------------------------------------------------- */

namespace {

/* Scalars and envelope values (fH, fE per step) of steps [k0, k0 + G) of plan p;
 * false where the envelope is undefined (past the stop period). */
bool segment_scalars(const smx_plan *p, long long k0, int G, std::vector<smx_step_scalars> &s,
                     std::vector<double> &env)
{
    s.resize((size_t)G);
    env.resize(2 * (size_t)G);
    for (int j = 0; j < G; j++) {
        const long long k = k0 + j;
        if (step_scalars(p, (int)(k / p->N), (int)(k % p->N), &s[(size_t)j], &env[2 * (size_t)j],
                         &env[2 * (size_t)j + 1]))
            return false;
    }
    return true;
}

bool same_scalars(const std::vector<smx_step_scalars> &a, const std::vector<smx_step_scalars> &b)
{
    return a.size() == b.size() && std::memcmp(a.data(), b.data(), sizeof(smx_step_scalars) * a.size()) == 0;
}

/* Host-side state that advance() changes: restored when a capture fails. */
struct plan_host_state {
    int cur;
    long long next_step;
};

/*
 * One graph holding, for every plan of the batch, one segment of its source loop: G_p
 * steps, G_p a multiple of N_2pi and of 4, captured from the plan's own stream.  The
 * kernels of a step take the step's scalars (envelope, phases) by value, and those
 * repeat with period N_2pi while the envelope is constant (the CW periods), so the
 * same graph advances every plan by another G_p steps wherever the scalars of those
 * steps equal the captured ones bit for bit.  A segment flips the FP32 ping-pong
 * buffers an even number of times (G_p one-step passes or G_p/2 two-step passes), so
 * every replay starts from the buffers the capture started from.
 */
struct batch_graph {
    hipStream_t ms = nullptr;              /* capture and launch stream */
    hipEvent_t fork = nullptr, join_ev = nullptr;
    std::vector<hipEvent_t> ev;            /* one per plan, for launches */
    std::vector<hipEvent_t> cap_ev;        /* one per plan, inside the capture */
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<int> G;                    /* steps per segment, per plan */
    std::vector<int> cur0;                 /* FP32 buffer index at the segment start */
    std::vector<std::vector<smx_step_scalars>> tpl; /* captured scalars per plan */

    ~batch_graph()
    {
        /* the graph's last launch may still run: destroy nothing under it */
        if (ms) (void)hipStreamSynchronize(ms);
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        for (hipEvent_t e : ev)
            if (e) (void)hipEventDestroy(e);
        for (hipEvent_t e : cap_ev)
            if (e) (void)hipEventDestroy(e);
        if (fork) (void)hipEventDestroy(fork);
        if (join_ev) (void)hipEventDestroy(join_ev);
        if (ms) (void)hipStreamDestroy(ms);
    }
};

long long gcd_ll(long long a, long long b) { return b ? gcd_ll(b, a % b) : a; }

constexpr long long kBatchMaxSegment = 1024; /* longest segment captured, in steps */
constexpr int kBatchCheckEvery = 8;          /* eager rounds between replay checks */

/* Work already queued on the plans' streams precedes what is launched on ms. */
int join_into(batch_graph &bg, smx_plan *const *plans, int nplans)
{
    for (int i = 0; i < nplans; i++) {
        SMX_HIP(hipEventRecord(bg.ev[(size_t)i], plans[i]->stream));
        SMX_HIP(hipStreamWaitEvent(bg.ms, bg.ev[(size_t)i], 0));
    }
    return SMX_OK;
}

/* Work launched on ms precedes what is queued on the plans' streams afterwards. */
int fork_out(batch_graph &bg, smx_plan *const *plans, int nplans)
{
    SMX_HIP(hipEventRecord(bg.join_ev, bg.ms));
    for (int i = 0; i < nplans; i++) SMX_HIP(hipStreamWaitEvent(plans[i]->stream, bg.join_ev, 0));
    return SMX_OK;
}

int graph_launch(batch_graph &bg, smx_plan *const *plans, int nplans)
{
    int rc = join_into(bg, plans, nplans);
    if (rc) return rc;
    SMX_HIP(hipGraphLaunch(bg.exec, bg.ms));
    return fork_out(bg, plans, nplans);
}

/*
 * Capture one segment of every plan (plan i from step next[i]) into bg and launch it
 * once.  The host state advances as if the steps had run, which the launch makes
 * true.  Returns SMX_OK with bg.exec == nullptr when the segment cannot be replayed
 * (host state restored, nothing launched), an error code on a HIP error.
 */
int capture_segment(batch_graph &bg, smx_plan *const *plans, int nplans, std::vector<long long> &next,
                    double *sums)
{
    std::vector<plan_host_state> saved((size_t)nplans);
    std::vector<double> saved_sums(sums ? 2 * (size_t)nplans : 0);
    for (int i = 0; i < nplans; i++) {
        smx_plan *p = plans[i];
        if (p->g.precision == SMX_F32) SMX_HIP(smx_f32_reserve(&p->f32, p->stream));
        saved[(size_t)i] = {p->f32.cur, p->next_step};
    }
    if (sums) std::memcpy(saved_sums.data(), sums, sizeof(double) * saved_sums.size());
    bg.cur0.resize((size_t)nplans);
    bg.tpl.resize((size_t)nplans);
    std::vector<double> env;
    for (int i = 0; i < nplans; i++) {
        bg.cur0[(size_t)i] = plans[i]->f32.cur;
        if (!segment_scalars(plans[i], next[(size_t)i], bg.G[(size_t)i], bg.tpl[(size_t)i], env))
            return arg_fail("batch segment beyond the stop period");
    }

    SMX_HIP(hipStreamBeginCapture(bg.ms, hipStreamCaptureModeThreadLocal));
    int rc = SMX_OK;
    hipError_t he = hipEventRecord(bg.fork, bg.ms);
    for (int i = 0; i < nplans && he == hipSuccess; i++) he = hipStreamWaitEvent(plans[i]->stream, bg.fork, 0);
    if (he != hipSuccess) rc = hip_fail(he, "batch capture fork");
    /* step-major within the segment, as the eager loop issues them */
    std::vector<long long> k((size_t)nplans);
    for (int i = 0; i < nplans; i++) k[(size_t)i] = next[(size_t)i];
    for (bool more = true; more && rc == SMX_OK;) {
        more = false;
        for (int i = 0; i < nplans && rc == SMX_OK; i++) {
            const long long end = next[(size_t)i] + bg.G[(size_t)i];
            if (k[(size_t)i] >= end) continue;
            int done = 0;
            rc = advance(plans[i], k[(size_t)i], end - k[(size_t)i], sums ? sums + 2 * i : nullptr, &done);
            k[(size_t)i] += done;
            more = more || k[(size_t)i] < end;
        }
    }
    for (int i = 0; i < nplans && rc == SMX_OK; i++) {
        if ((he = hipEventRecord(bg.cap_ev[(size_t)i], plans[i]->stream)) != hipSuccess ||
            (he = hipStreamWaitEvent(bg.ms, bg.cap_ev[(size_t)i], 0)) != hipSuccess)
            rc = hip_fail(he, "batch capture join");
    }
    hipGraph_t graph = nullptr;
    he = hipStreamEndCapture(bg.ms, &graph);
    if (rc == SMX_OK && he != hipSuccess) rc = hip_fail(he, "hipStreamEndCapture");
    bool replayable = rc == SMX_OK;
    for (int i = 0; i < nplans && replayable; i++)
        replayable = plans[i]->f32.cur == bg.cur0[(size_t)i];
    if (rc == SMX_OK && replayable) {
        he = hipGraphInstantiate(&bg.exec, graph, nullptr, nullptr, 0);
        if (he != hipSuccess) {
            bg.exec = nullptr;
            rc = hip_fail(he, "hipGraphInstantiate");
        }
    }
    bg.graph = graph;
    if (rc != SMX_OK || !replayable) {
        /* nothing of the segment ran: undo the host-side advance */
        for (int i = 0; i < nplans; i++) {
            plans[i]->f32.cur = saved[(size_t)i].cur;
            plans[i]->next_step = saved[(size_t)i].next_step;
        }
        if (sums) std::memcpy(sums, saved_sums.data(), sizeof(double) * saved_sums.size());
        if (bg.exec) (void)hipGraphExecDestroy(bg.exec);
        bg.exec = nullptr;
        return rc;
    }
    rc = graph_launch(bg, plans, nplans);
    if (rc) return rc;
    for (int i = 0; i < nplans; i++) {
        next[(size_t)i] += bg.G[(size_t)i];
        plans[i]->next_step = next[(size_t)i];
    }
    return SMX_OK;
}

/* Replay bg for every plan from next[i] if all their scalars match; *ran = 1 if so. */
int try_replay(batch_graph &bg, smx_plan *const *plans, int nplans, long long end, std::vector<long long> &next,
               double *sums, int *ran)
{
    *ran = 0;
    for (int i = 0; i < nplans; i++)
        if (next[(size_t)i] + bg.G[(size_t)i] > end || plans[i]->f32.cur != bg.cur0[(size_t)i]) return SMX_OK;
    std::vector<std::vector<double>> env((size_t)nplans);
    std::vector<smx_step_scalars> s;
    for (int i = 0; i < nplans; i++) {
        if (!segment_scalars(plans[i], next[(size_t)i], bg.G[(size_t)i], s, env[(size_t)i]) ||
            !same_scalars(s, bg.tpl[(size_t)i]))
            return SMX_OK;
    }
    int rc = graph_launch(bg, plans, nplans);
    if (rc) return rc;
    for (int i = 0; i < nplans; i++) {
        if (sums)
            for (int j = 0; j < bg.G[(size_t)i]; j++) {
                sums[2 * i] += env[(size_t)i][2 * (size_t)j];
                sums[2 * i + 1] += env[(size_t)i][2 * (size_t)j + 1];
            }
        next[(size_t)i] += bg.G[(size_t)i];
        plans[i]->next_step = next[(size_t)i];
    }
    *ran = 1;
    return SMX_OK;
}

/* Whether plan i's segment at next[i] repeats in the G_i steps after it. */
bool segment_repeats(const smx_plan *p, long long k0, int G, long long end)
{
    if (k0 + 2 * (long long)G > end) return false;
    std::vector<smx_step_scalars> a, b;
    std::vector<double> env;
    return segment_scalars(p, k0, G, a, env) && segment_scalars(p, k0 + G, G, b, env) && same_scalars(a, b);
}

/*
 * Mode SMX_BATCH_FUSED: the two-step passes of all plans, one launch of two_step_multi
 * per pass.  Its grid is a block list over every plan's tiles (plan << 20 | tile), so
 * small plans fill the device together and the host issues one launch per pass instead
 * of one per plan and pass.  Each plan's kernel arguments sit in a device table, two
 * entries per plan (the passes from either ping-pong buffer); the 20 step scalars per
 * plan and pass go up a chunk of passes at a time, into one of two device buffers,
 * while the passes of the chunk before run from the other.
 */
constexpr int kFusedChunk = 256; /* passes per upload of step scalars */
constexpr int kFusedFZ[] = {8, 12, 16, 24, 32, 48, 64}; /* z chunks the batch chooses from */
constexpr int kFusedDefaultFZ = 16;
constexpr int kScalarsPerPass = SMX_MULTI_SCALARS; /* 20 step scalars, then the DFT ring slot */
static_assert(sizeof(smx_step_scalars) == 10 * sizeof(double), "step scalars: 10 doubles in kernel order");

struct fused_batch {
    hipStream_t ms = nullptr;
    hipEvent_t join_ev = nullptr;
    std::vector<hipEvent_t> ev; /* one per plan */
    hipEvent_t done[2] = {nullptr, nullptr};
    void *args = nullptr;
    int *blk = nullptr;
    double *scal[2] = {nullptr, nullptr};
    std::vector<double> host[2]; /* the chunks' scalars until their upload has run */

    ~fused_batch()
    {
        if (ms) (void)hipStreamSynchronize(ms);
        for (double *d : scal)
            if (d) (void)hipFree(d);
        if (blk) (void)hipFree(blk);
        if (args) (void)hipFree(args);
        for (hipEvent_t e : done)
            if (e) (void)hipEventDestroy(e);
        for (hipEvent_t e : ev)
            if (e) (void)hipEventDestroy(e);
        if (join_ev) (void)hipEventDestroy(join_ev);
        if (ms) (void)hipStreamDestroy(ms);
    }
};

bool fused_ok(smx_plan *const *plans, int nplans)
{
    if (nplans >= (1 << 11)) return false;
    for (int i = 0; i < nplans; i++) {
        const smx_plan *p = plans[i];
        /* a DFT: its ring must hold the samples of a pass (chunk >= 2) */
        if (p->g.precision != SMX_F32 || (p->dft_nplanes != 0 && p->dft_chunk < 2) || p->pt_comp >= 0 ||
            !smx_f32_multi_ok(&p->f32, kFusedDefaultFZ))
            return false;
    }
    return true;
}

/*
 * The z chunk of a fused batch.  A pass lasts about (rounds of resident blocks) x (stages
 * a block marches, fz + 4): where the batch's blocks fill the device a few times at
 * most (small grids), the chunk is chosen to make that product least (the measured pass
 * times of the 24 small grids, 8 <= fz <= 64, follow it within 10%); on batches many
 * times the device, chunks of 16 planes (measured best there: the chunks then balance
 * the blocks of one round, and short marches keep the halo planes in the L2).
 * SMX_BATCH_FZ overrides the choice.
 */
int fused_fz(smx_plan *const *plans, int nplans)
{
    if (const char *e = getenv("SMX_BATCH_FZ")) return atoi(e);
    const long long slots = smx_f32_multi_slots();
    if (slots <= 0) return kFusedDefaultFZ;
    auto blocks = [&](int fz) {
        long long B = 0;
        for (int i = 0; i < nplans; i++) {
            const smx_f32_state &st = plans[i]->f32;
            if (!smx_f32_multi_ok(&st, fz)) return -1LL;
            B += (long long)((st.L[0] + 31) / 32) * ((st.L[1] + 7) / 8) * ((st.L[2] + fz - 1) / fz);
        }
        return B;
    };
    if (blocks(kFusedDefaultFZ) > 4 * slots) return kFusedDefaultFZ;
    int best = kFusedDefaultFZ;
    long long best_cost = -1;
    for (int fz : kFusedFZ) {
        const long long B = blocks(fz);
        if (B <= 0) continue;
        const long long cost = (B + slots - 1) / slots * (fz + 4);
        if (best_cost < 0 || cost < best_cost) best = fz, best_cost = cost;
    }
    return best;
}

/*
 * Passes [0, npass) of every plan from step k0 (steps k0 .. k0 + 2 npass - 1), in
 * order; *ran = passes launched, fewer where a step's envelope is undefined (the
 * caller's step loop then meets that step and reports it).
 */
int run_fused(smx_plan *const *plans, int nplans, long long k0, long long npass, double *sums, long long *ran)
{
    *ran = 0;
    const size_t asz = smx_f32_multi_args_size();
    std::vector<unsigned char> args(2 * (size_t)nplans * asz), two(2 * asz);
    std::vector<int> blk;
    const int fz = fused_fz(plans, nplans);
    bool any_dft = false;
    for (int i = 0; i < nplans; i++) {
        smx_plan *p = plans[i];
        /* the monitors in the pass unless the DFT ring carries them; the ring's slot of
         * each pass goes with its scalars */
        smx_f32_launch_args la = step2_args(p, !mon_ring(p));
        if (p->dft_nplanes) {
            any_dft = true;
            la.dft = (float *)p->dft_series;
            la.dft_n = p->dft_ring;
            std::memcpy(la.dft_planes, p->dft_planes, sizeof(int) * (size_t)p->dft_ring);
        }
        int nb = 0;
        if (smx_f32_multi_prepare(&p->f32, &la, fz, two.data(), &nb) != SMX_OK)
            return arg_fail("batch: two-step tiling of a plan");
        std::memcpy(args.data() + (size_t)i * asz, two.data(), asz);
        std::memcpy(args.data() + ((size_t)nplans + i) * asz, two.data() + asz, asz);
        for (int b = 0; b < nb; b++) blk.push_back(i << 20 | b);
    }
    if (blk.size() >= ((size_t)1 << 31)) return arg_fail("batch: too many blocks");
    fused_batch fb;
    SMX_HIP(hipStreamCreateWithFlags(&fb.ms, hipStreamNonBlocking));
    SMX_HIP(hipEventCreateWithFlags(&fb.join_ev, hipEventDisableTiming));
    for (hipEvent_t &e : fb.done) SMX_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    fb.ev.assign((size_t)nplans, nullptr);
    for (hipEvent_t &e : fb.ev) SMX_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    const long long chunk = std::min<long long>(npass, kFusedChunk);
    const size_t per_pass = (size_t)nplans * kScalarsPerPass;
    SMX_HIP(hipMalloc(&fb.args, args.size()));
    SMX_HIP(hipMalloc(&fb.blk, sizeof(int) * blk.size()));
    for (double *&d : fb.scal) SMX_HIP(hipMalloc(&d, sizeof(double) * per_pass * (size_t)chunk));
    SMX_HIP(hipMemcpy(fb.args, args.data(), args.size(), hipMemcpyHostToDevice));
    SMX_HIP(hipMemcpy(fb.blk, blk.data(), sizeof(int) * blk.size(), hipMemcpyHostToDevice));
    /* work queued on the plans' streams runs first */
    for (int i = 0; i < nplans; i++) {
        SMX_HIP(hipEventRecord(fb.ev[(size_t)i], plans[i]->stream));
        SMX_HIP(hipStreamWaitEvent(fb.ms, fb.ev[(size_t)i], 0));
    }
    /* the plans' own launches (DFT folds) go to the batch stream meanwhile */
    std::vector<hipStream_t> own((size_t)nplans);
    for (int i = 0; i < nplans; i++) own[(size_t)i] = plans[i]->stream, plans[i]->stream = fb.ms;
    struct restore_streams {
        smx_plan *const *pl;
        std::vector<hipStream_t> &s;
        ~restore_streams()
        {
            for (size_t i = 0; i < s.size(); i++) pl[i]->stream = s[i];
        }
    } restore{plans, own};
    std::vector<double> env(4 * (size_t)nplans);
    bool stop = false;
    long long done = 0;
    for (long long c = 0; done < npass && !stop; c++) {
        const int b = (int)(c & 1);
        if (c >= 2) SMX_HIP(hipEventSynchronize(fb.done[b])); /* chunk c - 2 has run */
        std::vector<double> &h = fb.host[b];
        h.resize(per_pass * (size_t)chunk);
        /* the DFT rings' fill as the passes of this chunk will leave it (dft_pair_slot) */
        std::vector<int> fill((size_t)nplans);
        std::vector<long long> kfirst((size_t)nplans);
        for (int i = 0; i < nplans; i++) fill[(size_t)i] = plans[i]->dft_fill, kfirst[(size_t)i] = plans[i]->dft_k0;
        long long n = 0;
        for (; n < chunk && done + n < npass; n++) {
            const long long k = k0 + 2 * (done + n);
            bool ok = true;
            for (int i = 0; i < nplans && ok; i++) {
                const smx_plan *p = plans[i];
                smx_step_scalars st[2];
                for (int j = 0; j < 2 && ok; j++)
                    ok = step_scalars(p, (int)((k + j) / p->N), (int)((k + j) % p->N), &st[j],
                                      &env[4 * (size_t)i + 2 * j], &env[4 * (size_t)i + 2 * j + 1]) == SMX_OK;
                double *row = &h[per_pass * (size_t)n + kScalarsPerPass * (size_t)i];
                std::memcpy(row, st, sizeof st);
                int slot = 0;
                if (p->dft_nplanes) {
                    int &f = fill[(size_t)i];
                    if (f > 0 && (k != kfirst[(size_t)i] + f || f + 2 > p->dft_chunk)) f = 0;
                    if (f == 0) kfirst[(size_t)i] = k;
                    slot = f;
                    f += 2;
                    if (f == p->dft_chunk) f = 0;
                }
                row[20] = (double)slot;
            }
            if (!ok) {
                stop = true;
                break;
            }
            if (sums) /* per plan in step order, as the step loop adds them */
                for (int i = 0; i < nplans; i++)
                    for (int j = 0; j < 2; j++) {
                        sums[2 * i] += env[4 * (size_t)i + 2 * j];
                        sums[2 * i + 1] += env[4 * (size_t)i + 2 * j + 1];
                    }
        }
        if (n == 0) break;
        SMX_HIP(hipMemcpyAsync(fb.scal[b], h.data(), sizeof(double) * per_pass * (size_t)n, hipMemcpyHostToDevice,
                               fb.ms));
        for (long long j = 0; j < n; j++) {
            const long long k = k0 + 2 * (done + j);
            /* the rings' bookkeeping as advance() keeps it: a fold before a pass that does
             * not continue the chunk or does not fit, after one that fills it */
            for (int i = 0; i < nplans; i++) {
                smx_plan *p = plans[i];
                if (!p->dft_nplanes) continue;
                int slot = 0, rc;
                if ((rc = dft_pair_slot(p, k, &slot))) return rc;
                if (slot != (int)h[per_pass * (size_t)j + kScalarsPerPass * (size_t)i + 20])
                    return arg_fail("batch: DFT ring slot out of step");
            }
            const void *a = static_cast<const unsigned char *>(fb.args) + (size_t)((done + j) & 1) * nplans * asz;
            SMX_HIP(smx_f32_multi_launch(a, fb.blk, fb.scal[b] + per_pass * (size_t)j, (int)blk.size(), any_dft ? 1 : 0,
                                         fb.ms));
            for (int i = 0; i < nplans; i++) {
                smx_plan *p = plans[i];
                if (!p->dft_nplanes) continue;
                const int slot = p->dft_fill;
                if (mon_ring(p)) {
                    mark_mon_slot(p, slot);
                    mark_mon_slot(p, slot + 1);
                }
                p->dft_fill += 2;
                int rc;
                if (p->dft_fill == p->dft_chunk && (rc = dft_flush(p))) return rc;
            }
        }
        SMX_HIP(hipEventRecord(fb.done[b], fb.ms));
        done += n;
    }
    /* and what is queued on them afterwards runs after the passes */
    SMX_HIP(hipEventRecord(fb.join_ev, fb.ms));
    for (int i = 0; i < nplans; i++) {
        SMX_HIP(hipStreamWaitEvent(own[(size_t)i], fb.join_ev, 0));
        if (done & 1) plans[i]->f32.cur = 1 - plans[i]->f32.cur;
        plans[i]->next_step = k0 + 2 * done;
    }
    *ran = done;
    return SMX_OK;
}

} // namespace

extern "C" {

int smx_plan_run_batch_mode(smx_plan *const *plans, int nplans, long long step0, long long nsteps, double *sums,
                            int mode, int *graph_launches)
{
    if (graph_launches) *graph_launches = 0;
    if (mode != SMX_BATCH_STREAMS && mode != SMX_BATCH_GRAPH && mode != SMX_BATCH_FUSED) return arg_fail("batch mode");
    if (!plans || nplans <= 0) return arg_fail("empty batch");
    for (int i = 0; i < nplans; i++) {
        if (!plans[i]) return arg_fail("null plan in batch");
        for (int j = 0; j < i; j++)
            if (plans[j] == plans[i]) return arg_fail("plan listed twice in batch");
        if (!plans[i]->have_source) {
            snprintf(g_last_error, sizeof g_last_error, "batch plan %d: run before smx_plan_set_source", i);
            return SMX_ERR_STATE;
        }
    }
    if (step0 < 0 || nsteps < 0) return arg_fail("negative step");
    long long end = step0 + nsteps;

    std::vector<long long> next((size_t)nplans, step0);
    if (mode == SMX_BATCH_FUSED) {
        if (!fused_ok(plans, nplans)) mode = SMX_BATCH_GRAPH;
        else if (nsteps >= 2) {
            long long ran = 0;
            const int rc = run_fused(plans, nplans, step0, nsteps / 2, sums, &ran);
            if (rc) return rc;
            if (graph_launches) *graph_launches = (int)std::min<long long>(ran, INT_MAX);
            for (long long &k : next) k = step0 + 2 * ran;
            /* an odd last step, or the step whose envelope is undefined, as one step */
            mode = SMX_BATCH_STREAMS;
        }
    }

    /* graph replay needs every plan's steps to be plain kernel launches (no DFT
     * samples, whose fold points depend on host counters, and no point source, whose
     * amplitude changes every step) */
    bool graph = mode == SMX_BATCH_GRAPH;
    batch_graph bg;
    if (graph) {
        bg.G.resize((size_t)nplans);
        for (int i = 0; i < nplans && graph; i++) {
            const long long N = plans[i]->N;
            const long long G = N / gcd_ll(N, 4) * 4;
            graph = plans[i]->dft_nplanes == 0 && plans[i]->pt_comp < 0 && G <= kBatchMaxSegment && 2 * G <= nsteps;
            bg.G[(size_t)i] = (int)G;
        }
    }
    if (graph) {
        SMX_HIP(hipStreamCreateWithFlags(&bg.ms, hipStreamNonBlocking));
        SMX_HIP(hipEventCreateWithFlags(&bg.fork, hipEventDisableTiming));
        SMX_HIP(hipEventCreateWithFlags(&bg.join_ev, hipEventDisableTiming));
        bg.ev.assign((size_t)nplans, nullptr);
        bg.cap_ev.assign((size_t)nplans, nullptr);
        for (int i = 0; i < nplans; i++) {
            SMX_HIP(hipEventCreateWithFlags(&bg.ev[(size_t)i], hipEventDisableTiming));
            SMX_HIP(hipEventCreateWithFlags(&bg.cap_ev[(size_t)i], hipEventDisableTiming));
        }
    }

    bool captured = false; /* a capture was tried (one per call) */
    int wait = 0;          /* eager rounds until the next replay or capture check */
    for (;;) {
        if (graph && wait-- <= 0) {
            wait = kBatchCheckEvery;
            if (bg.exec) {
                int ran = 0;
                const int rc = try_replay(bg, plans, nplans, end, next, sums, &ran);
                if (rc) return rc;
                if (ran) {
                    if (graph_launches) ++*graph_launches;
                    wait = 0;
                    continue;
                }
            } else if (!captured) {
                /* capture once every plan has entered a stretch of repeating scalars */
                bool ready = true;
                for (int i = 0; i < nplans && ready; i++)
                    ready = segment_repeats(plans[i], next[(size_t)i], bg.G[(size_t)i], end);
                if (ready) {
                    captured = true;
                    const int rc = capture_segment(bg, plans, nplans, next, sums);
                    if (rc) return rc;
                    if (bg.exec) {
                        if (graph_launches) ++*graph_launches;
                        wait = 0;
                        continue;
                    }
                }
            }
        }
        /* one eager round: every plan that is behind the front advances by one step
         * or one two-step pass, in step order */
        long long kmin = end;
        for (int i = 0; i < nplans; i++) kmin = std::min(kmin, next[(size_t)i]);
        if (kmin >= end) break;
        for (int i = 0; i < nplans; i++) {
            if (next[(size_t)i] != kmin) continue;
            smx_plan *p = plans[i];
            int done = 0;
            const int rc = advance(p, kmin, end - kmin, sums ? sums + 2 * i : nullptr, &done);
            if (rc) return rc;
            next[(size_t)i] = kmin + done;
            p->next_step = kmin + done;
        }
    }
    return SMX_OK;
}

int smx_plan_run_batch(smx_plan *const *plans, int nplans, long long step0, long long nsteps, double *sums)
{
    return smx_plan_run_batch_mode(plans, nplans, step0, nsteps, sums, SMX_BATCH_GRAPH, nullptr);
}

int smx_plan_run(smx_plan *p, int P0, int P1, double sums[2])
{
    if (!p) return arg_fail("null pointer");
    if (!p->have_source) {
        snprintf(g_last_error, sizeof g_last_error, "smx_plan_run before smx_plan_set_source");
        return SMX_ERR_STATE;
    }
    if (P1 < P0) return arg_fail("P1 < P0");
    return smx_plan_run_steps(p, (long long)P0 * p->N, (long long)(P1 - P0) * p->N, sums);
}

void *smx_plan_stream(const smx_plan *p) { return p ? (void *)p->stream : nullptr; }

const char *smx_plan_kernels(const smx_plan *p)
{
    if (!p) return "";
    if (p->g.precision == SMX_F64) return "h_sweep_f64 + e_sweep_f64";
    return smx_f32_kernel_name(&p->f32, p->dft_nplanes > 0 && pair_ok(p));
}

int smx_plan_set_dft(smx_plan *p, int nplanes, const int *planes, int nfreq, const double *w_per_step,
                     int chunk)
{
    if (!p) return arg_fail("null pointer");
    if (nplanes < 0 || nplanes > SMX_DFT_MAX_PLANES) return arg_fail("DFT planes: 0..8");
    if (nplanes > 0 && (!planes || !w_per_step)) return arg_fail("null pointer");
    if (nplanes > 0 && (nfreq < 1 || nfreq > SMX_DFT_MAX_FREQ)) return arg_fail("DFT frequencies: 1..32");
    if (nplanes > 0 && (chunk < 1 || chunk > SMX_DFT_MAX_CHUNK)) return arg_fail("DFT chunk: 1..256");
    for (int m = 0; m < nplanes; m++)
        if (planes[m] < 0 || planes[m] >= p->g.L[2]) return arg_fail("DFT plane outside the grid");
    /* samples still in the ring: their monitor contributions are kept */
    int rc = dft_flush(p);
    if (rc) return rc;
    SMX_HIP(hipStreamSynchronize(p->stream));
    if (p->dft_series) SMX_HIP(hipFree(p->dft_series));
    if (p->dft_acc) SMX_HIP(hipFree(p->dft_acc));
    if (p->dft_table) SMX_HIP(hipFree(p->dft_table));
    p->dft_series = nullptr;
    p->dft_acc = nullptr;
    p->dft_table = nullptr;
    p->dft_nplanes = 0;
    p->dft_ring = 0;
    p->dft_fill = 0;
    p->dft_mon[0] = p->dft_mon[1] = -1;
    std::memset(p->dft_mon_mask, 0, sizeof p->dft_mon_mask);
    if (nplanes == 0) return SMX_OK;
    /* ring planes: the DFT planes, then (FP32) the monitor planes not among them */
    int ring[SMX_DFT_MAX_PLANES];
    std::memcpy(ring, planes, sizeof(int) * (size_t)nplanes);
    int nring = nplanes, mon[2] = {-1, -1};
    bool fits = true;
    for (int m = 0; m < 2 && p->g.precision == SMX_F32; m++) {
        const int z = p->g.iBTIin[m];
        if (z < 0) continue;
        int r = 0;
        while (r < nring && ring[r] != z) r++;
        if (r == nring) {
            if (nring == SMX_DFT_MAX_PLANES) {
                fits = false;
                break;
            }
            ring[nring++] = z;
        }
        mon[m] = r;
    }
    if (!fits) { /* no room for the monitor planes: the kernels keep accumulating them */
        nring = nplanes;
        mon[0] = mon[1] = -1;
    }
    const size_t word = p->g.precision == SMX_F64 ? sizeof(double) : sizeof(float);
    const size_t cells = (size_t)nplanes * 4 * (size_t)p->plane;
    SMX_HIP(hipMalloc(&p->dft_series, word * (size_t)nring * 4 * (size_t)p->plane * (size_t)chunk));
    SMX_HIP(hipMalloc(&p->dft_acc, sizeof(double) * 2 * cells * (size_t)nfreq));
    SMX_HIP(hipMemsetAsync(p->dft_acc, 0, sizeof(double) * 2 * cells * (size_t)nfreq, p->stream));
    p->dft_nplanes = nplanes;
    p->dft_nfreq = nfreq;
    p->dft_chunk = chunk;
    std::memset(p->dft_planes, 0, sizeof p->dft_planes);
    std::memset(p->dft_w, 0, sizeof p->dft_w);
    std::memcpy(p->dft_planes, ring, sizeof(int) * nring);
    std::memcpy(p->dft_w, w_per_step, sizeof(double) * nfreq);
    p->dft_ring = nring;
    p->dft_mon[0] = mon[0];
    p->dft_mon[1] = mon[1];
    std::vector<double> tab(4 * (size_t)chunk * (size_t)smx_dft_table_row(nfreq));
    smx_dft_table(nfreq, w_per_step, chunk, tab.data());
    SMX_HIP(hipMalloc(&p->dft_table, sizeof(double) * tab.size()));
    SMX_HIP(hipMemcpyAsync(p->dft_table, tab.data(), sizeof(double) * tab.size(), hipMemcpyHostToDevice,
                           p->stream));
    SMX_HIP(hipStreamSynchronize(p->stream));
    return SMX_OK;
}

int smx_plan_set_point_source(smx_plan *p, int comp, const int cell[3], long long nsteps, const double *amp)
{
    if (!p) return arg_fail("null pointer");
    if (nsteps <= 0 || !amp) {
        SMX_HIP(hipStreamSynchronize(p->stream));
        if (p->pt_amp) SMX_HIP(hipFree(p->pt_amp));
        p->pt_amp = nullptr;
        p->pt_comp = -1;
        p->pt_n = 0;
        return SMX_OK;
    }
    if (!cell) return arg_fail("null pointer");
    if (comp < 0 || comp > 2) return arg_fail("point source: component must be 0, 1 or 2 (E x, y, z)");
    for (int d = 0; d < 3; d++)
        if (cell[d] < 0 || cell[d] >= p->g.L[d]) return arg_fail("point source: cell outside the grid");
    const int64_t l = cell[0] + (int64_t)p->g.L[0] * (cell[1] + (int64_t)p->g.L[1] * cell[2]);
    /* the class of the cell, and in FP32 plans the lossy/metal flag (NaN in bE_x) */
    char cls = 'n';
    float flag = 0.f;
    SMX_HIP(hipStreamSynchronize(p->stream));
    if (p->g.precision == SMX_F64) {
        SMX_HIP(hipMemcpy(&cls, p->Cont + l, 1, hipMemcpyDeviceToHost));
    } else {
        cls = p->f32.cont_host ? p->f32.cont_host[l] : 'n';
        SMX_HIP(hipMemcpy(&flag, p->f32.b[3] + l, sizeof(float), hipMemcpyDeviceToHost));
    }
    if (cls != 'n' || std::isnan(flag)) return arg_fail("point source: the cell must be a lossless class-'n' cell");
    double *dev = nullptr;
    SMX_HIP(hipMalloc(&dev, sizeof(double) * (size_t)nsteps));
    const hipError_t e = hipMemcpy(dev, amp, sizeof(double) * (size_t)nsteps, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        (void)hipFree(dev);
        return hip_fail(e, "point source amplitudes");
    }
    if (p->pt_amp) SMX_HIP(hipFree(p->pt_amp));
    p->pt_amp = dev;
    p->pt_comp = comp;
    p->pt_cell = l;
    p->pt_n = nsteps;
    return SMX_OK;
}

int smx_plan_read_dft(smx_plan *p, double *out)
{
    if (!p || !out) return arg_fail("null pointer");
    if (!p->dft_nplanes) {
        snprintf(g_last_error, sizeof g_last_error, "no DFT planes set");
        return SMX_ERR_STATE;
    }
    int rc = dft_flush(p);
    if (rc) return rc;
    const size_t len = 2 * (size_t)p->dft_nplanes * 4 * (size_t)p->plane * (size_t)p->dft_nfreq;
    SMX_HIP(hipMemcpyAsync(out, p->dft_acc, sizeof(double) * len, hipMemcpyDeviceToHost, p->stream));
    SMX_HIP(hipStreamSynchronize(p->stream));
    return SMX_OK;
}

int smx_plan_sync(smx_plan *p)
{
    if (!p) return arg_fail("null pointer");
    SMX_HIP(hipStreamSynchronize(p->stream));
    return SMX_OK;
}

int smx_plan_flush(smx_plan *p)
{
    if (!p) return arg_fail("null pointer");
    return dft_flush(p);
}

int smx_plan_project(smx_plan *p, int nmodes, const double *modes, double *S_mon, double *S_dft)
{
    if (!p || !modes || (!S_mon && !S_dft)) return arg_fail("null pointer");
    if (nmodes < 1) return arg_fail("nmodes must be >= 1");
    if (S_dft && !p->dft_nplanes) {
        snprintf(g_last_error, sizeof g_last_error, "smx_plan_project: no DFT set");
        return SMX_ERR_STATE;
    }
    int rc = dft_flush(p); /* pending samples into the monitors and the DFT accumulators */
    if (rc) return rc;
    smx_project_args a;
    std::memset(&a, 0, sizeof a);
    a.P = p->plane;
    a.nmodes = nmodes;
    a.nmon = S_mon ? 2 : 0;
    a.nfreq = S_dft ? p->dft_nfreq : 0;
    a.nvec = a.nmon + (S_dft ? p->dft_nplanes * p->dft_nfreq : 0);
    a.nchunks = smx_project_chunks(a.P);
    a.mon = reinterpret_cast<const double2 *>(p->mon);
    a.dft = reinterpret_cast<const double2 *>(p->dft_acc);
    const size_t mbytes = sizeof(double2) * (size_t)nmodes * 4 * (size_t)a.P;
    const size_t pbytes = sizeof(double2) * (size_t)a.nvec * nmodes * (size_t)a.nchunks;
    const size_t sbytes = sizeof(double2) * (size_t)a.nvec * nmodes;
    void *dm = nullptr, *dp = nullptr, *ds = nullptr;
    hipError_t e = hipMalloc(&dm, mbytes);
    if (e == hipSuccess) e = hipMalloc(&dp, pbytes);
    if (e == hipSuccess) e = hipMalloc(&ds, sbytes);
    if (e == hipSuccess) e = hipMemcpyAsync(dm, modes, mbytes, hipMemcpyHostToDevice, p->stream);
    if (e == hipSuccess) {
        a.modes = static_cast<const double2 *>(dm);
        a.partial = static_cast<double2 *>(dp);
        a.S = static_cast<double2 *>(ds);
        e = smx_project_launch(&a, p->stream);
    }
    std::vector<double> S(2 * (size_t)a.nvec * nmodes);
    if (e == hipSuccess) e = hipMemcpyAsync(S.data(), ds, sbytes, hipMemcpyDeviceToHost, p->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(p->stream);
    for (void *b : {dm, dp, ds})
        if (b) (void)hipFree(b);
    if (e != hipSuccess) return hip_fail(e, "smx_plan_project");
    const size_t per = 2 * (size_t)nmodes;
    if (S_mon) std::memcpy(S_mon, S.data(), sizeof(double) * 2 * per);
    if (S_dft) std::memcpy(S_dft, S.data() + a.nmon * per, sizeof(double) * (a.nvec - a.nmon) * per);
    return SMX_OK;
}

/* ---------------------------------------------------------------------------------- */
/* Adjoint gradient (extension)                                                        */
/* ---------------------------------------------------------------------------------- */

static smx_adj_args adj_args(smx_plan *p)
{
    smx_adj_args a;
    std::memset(&a, 0, sizeof a);
    a.L0 = p->g.L[0];
    a.L1 = p->g.L[1];
    a.L2 = p->g.L[2];
    a.n = p->n;
    a.sz = p->plane;
    for (int d = 0; d < 3; d++) a.w[d] = (float)p->g.iBc[d];
    a.zT = p->adj_zT;
    std::memcpy(a.box, p->adj_box, sizeof a.box);
    a.bx = a.box[1] - a.box[0];
    a.by = a.box[3] - a.box[2];
    a.nbox = (int64_t)a.bx * a.by * (a.box[5] - a.box[4]);
    for (int i = 0; i < 6; i++) a.b[i] = p->f32.b[i];
    a.pml = p->f32.pml;
    a.pmlflag = p->f32.pmlflag;
    a.D = p->adj_D;
    return a;
}

} // extern "C"

/* E of the design box (the adjoint's box) from the current fields to dst, [3][nbox]. */
static int save_box(smx_plan *p, float *dst)
{
    const smx_adj_args a = adj_args(p);
    float **T = p->f32.T[p->f32.cur];
    SMX_HIP(smx_adj_launch_save_box(&a, T[3], T[4], T[5], dst, p->stream));
    return SMX_OK;
}

namespace {

/* The DFT state of a plan (set by smx_plan_set_dft), parked while the adjoint forward
 * runs its own. */
struct dft_state {
    int nplanes, nfreq, chunk, fill, ring;
    int mon[2];
    uint64_t mon_mask[SMX_DFT_MAX_CHUNK / 64];
    int planes[SMX_DFT_MAX_PLANES];
    double w[SMX_DFT_MAX_FREQ];
    long long k0;
    void *series;
    double *acc, *table;
};

dft_state park_dft(smx_plan *p)
{
    dft_state d;
    d.nplanes = p->dft_nplanes, d.nfreq = p->dft_nfreq, d.chunk = p->dft_chunk, d.fill = p->dft_fill;
    d.ring = p->dft_ring, d.mon[0] = p->dft_mon[0], d.mon[1] = p->dft_mon[1];
    std::memcpy(d.mon_mask, p->dft_mon_mask, sizeof d.mon_mask);
    std::memcpy(d.planes, p->dft_planes, sizeof d.planes);
    std::memcpy(d.w, p->dft_w, sizeof d.w);
    d.k0 = p->dft_k0;
    d.series = p->dft_series, d.acc = p->dft_acc, d.table = p->dft_table;
    p->dft_nplanes = 0, p->dft_fill = 0, p->dft_ring = 0;
    p->dft_mon[0] = p->dft_mon[1] = -1;
    std::memset(p->dft_mon_mask, 0, sizeof p->dft_mon_mask);
    p->dft_series = nullptr, p->dft_acc = nullptr, p->dft_table = nullptr;
    return d;
}

void restore_dft(smx_plan *p, const dft_state &d)
{
    p->dft_nplanes = d.nplanes, p->dft_nfreq = d.nfreq, p->dft_chunk = d.chunk, p->dft_fill = d.fill;
    p->dft_ring = d.ring, p->dft_mon[0] = d.mon[0], p->dft_mon[1] = d.mon[1];
    std::memcpy(p->dft_mon_mask, d.mon_mask, sizeof d.mon_mask);
    std::memcpy(p->dft_planes, d.planes, sizeof d.planes);
    std::memcpy(p->dft_w, d.w, sizeof d.w);
    p->dft_k0 = d.k0;
    p->dft_series = d.series, p->dft_acc = d.acc, p->dft_table = d.table;
}

/* Whether the forward pass can take two steps per pass: the plan's two-step kernel
 * samples the objective plane itself (split slots, or no PML) and checkpoints fall on
 * pass boundaries. */
bool adjoint_forward_paired(const smx_plan *p, int ckpt_every)
{
    return p->f32.two >= 0 && smx_f32_two_full(&p->f32, p->f32.two) && ckpt_every % 2 == 0;
}

/*
 * The forward pass with two steps per pass: D_v(q) is the plan's DFT of E_v on plane zT
 * at w, whose E phase e^{i*w*(k+1.5)} is the objective's, taken by the two-step kernel
 * inside each pass; the caller's DFT, if any, is parked meanwhile and left as it was.
 */
int adjoint_forward_two_step(smx_plan *p, long long nsteps, int ckpt_every, int zT, double w, double *J)
{
    const int64_t n = p->n;
    const bool hist = ckpt_every == 0; /* E of the design box after every step, no checkpoints */
    const smx_adj_args ab = adj_args(p);
    const dft_state saved = park_dft(p);
    int rc = smx_plan_set_dft(p, 1, &zT, 1, &w, 64);
    if (rc == SMX_OK && hist) rc = save_box(p, p->adj_hist);
    for (long long k = 0; k < nsteps && rc == SMX_OK;) {
        if (hist) {
            int done = 0;
            rc = advance(p, k, nsteps - k, nullptr, &done, p->adj_hist + (size_t)(k + 1) * 3 * (size_t)ab.nbox);
            k += done;
            continue;
        }
        if (k % ckpt_every == 0) {
            float *dst = p->adj_ckpt + (size_t)(k / ckpt_every) * 12 * (size_t)n;
            hipError_t e = hipSuccess;
            for (int i = 0; i < 6 && e == hipSuccess; i++) {
                e = hipMemcpyAsync(dst + (size_t)i * n, p->f32.T[p->f32.cur][i], sizeof(float) * n,
                                   hipMemcpyDeviceToDevice, p->stream);
                if (e == hipSuccess)
                    e = hipMemcpyAsync(dst + (size_t)(6 + i) * n, p->f32.S[p->f32.cur][i], sizeof(float) * n,
                                       hipMemcpyDeviceToDevice, p->stream);
            }
            if (e != hipSuccess) {
                rc = hip_fail(e, "adjoint checkpoint");
                break;
            }
        }
        /* up to the next checkpoint, so that a pair never straddles one */
        const long long stop = std::min(nsteps, (k / ckpt_every + 1) * (long long)ckpt_every);
        int done = 0;
        rc = advance(p, k, stop - k, nullptr, &done);
        k += done;
    }
    if (rc == SMX_OK) rc = dft_flush(p);
    if (rc == SMX_OK) {
        /* acc[plane 0][Ex, Ey][freq 0][q]: the objective's D, laid out as adj_D */
        const size_t len = 4 * (size_t)p->plane;
        hipError_t e = hipMemcpyAsync(p->adj_D, p->dft_acc + len, sizeof(double) * len, hipMemcpyDeviceToDevice,
                                      p->stream);
        std::vector<double> D(len);
        if (e == hipSuccess)
            e = hipMemcpyAsync(D.data(), p->dft_acc + len, sizeof(double) * len, hipMemcpyDeviceToHost, p->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(p->stream);
        if (e != hipSuccess) rc = hip_fail(e, "adjoint objective");
        double acc = 0.;
        for (size_t i = 0; i < D.size(); i++) acc += D[i] * D[i];
        *J = acc;
    }
    (void)hipStreamSynchronize(p->stream);
    for (void *b : {(void *)p->dft_series, (void *)p->dft_acc, (void *)p->dft_table})
        if (b) (void)hipFree(b);
    restore_dft(p, saved);
    return rc;
}

} // namespace

extern "C" {

int smx_adjoint_forward(smx_plan *p, long long nsteps, int ckpt_every, int zT, double w, const int box[6],
                        double *J)
{
    if (!p || !box || !J) return arg_fail("null pointer");
    if (p->g.precision != SMX_F32) return arg_fail("the adjoint runs on FP32 plans");
    if (p->f32.nspecial > 0) return arg_fail("the adjoint does not support lossy or metal cells");
    if (p->pt_comp >= 0) return arg_fail("the adjoint does not support a point source");
    if (!p->have_source) {
        snprintf(g_last_error, sizeof g_last_error, "smx_adjoint_forward before smx_plan_set_source");
        return SMX_ERR_STATE;
    }
    if (nsteps < 1 || ckpt_every < 0) return arg_fail("nsteps must be >= 1 and ckpt_every >= 0");
    if (zT < 0 || zT >= p->g.L[2]) return arg_fail("objective plane outside the grid");
    for (int d = 0; d < 3; d++)
        if (box[2 * d] < 0 || box[2 * d + 1] > p->g.L[d] || box[2 * d] >= box[2 * d + 1])
            return arg_fail("design box outside the grid or empty");
    for (int z = box[4]; z < box[5]; z++)
        for (int y = box[2]; y < box[3]; y++)
            for (int x = box[0]; x < box[1]; x++)
                if (p->f32.cont_host[x + (int64_t)p->g.L[0] * (y + (int64_t)p->g.L[1] * z)] != 'n')
                    return arg_fail("design box holds a PML or special cell");
    const int64_t n = p->n;
    const bool hist = ckpt_every == 0;
    const long long nck = hist ? 0 : (nsteps + ckpt_every - 1) / ckpt_every;
    SMX_HIP(hipStreamSynchronize(p->stream));
    if (p->adj_ckpt) SMX_HIP(hipFree(p->adj_ckpt));
    if (p->adj_D) SMX_HIP(hipFree(p->adj_D));
    if (p->adj_hist) SMX_HIP(hipFree(p->adj_hist));
    p->adj_ckpt = nullptr;
    p->adj_D = nullptr;
    p->adj_hist = nullptr;
    p->adj_ready = false;
    const int64_t nbox = (int64_t)(box[1] - box[0]) * (box[3] - box[2]) * (box[5] - box[4]);
    if (hist) SMX_HIP(hipMalloc(&p->adj_hist, sizeof(float) * 3 * (size_t)nbox * (size_t)(nsteps + 1)));
    else SMX_HIP(hipMalloc(&p->adj_ckpt, sizeof(float) * 12 * (size_t)n * (size_t)nck));
    SMX_HIP(hipMalloc(&p->adj_D, sizeof(double) * 4 * (size_t)p->plane));
    SMX_HIP(hipMemsetAsync(p->adj_D, 0, sizeof(double) * 4 * (size_t)p->plane, p->stream));
    p->adj_K = nsteps;
    p->adj_C = ckpt_every;
    p->adj_zT = zT;
    p->adj_w = w;
    std::memcpy(p->adj_box, box, sizeof p->adj_box);
    /* the run starts from zero fields */
    for (int k = 0; k < 2; k++)
        for (int i = 0; i < 6; i++) {
            if (p->f32.T[k][i]) SMX_HIP(hipMemsetAsync(p->f32.T[k][i], 0, sizeof(float) * n, p->stream));
            if (p->f32.S[k][i]) SMX_HIP(hipMemsetAsync(p->f32.S[k][i], 0, sizeof(float) * n, p->stream));
        }
    const char *box2 = getenv("SMX_ADJ_BOX2"); /* "0": the history forward one step per launch */
    if (hist ? pair_box_ok(p) && !(box2 && !strcmp(box2, "0")) : adjoint_forward_paired(p, ckpt_every)) {
        const int rc = adjoint_forward_two_step(p, nsteps, ckpt_every, zT, w, J);
        if (rc == SMX_OK) p->adj_ready = true;
        return rc;
    }
    smx_adj_args a = adj_args(p);
    auto save_hist = [&](long long slot) {
        float **T = p->f32.T[p->f32.cur];
        return smx_adj_launch_save_box(&a, T[3], T[4], T[5], p->adj_hist + (size_t)slot * 3 * (size_t)nbox,
                                       p->stream);
    };
    if (hist) SMX_HIP(save_hist(0));
    for (long long k = 0; k < nsteps; k++) {
        if (!hist && k % ckpt_every == 0) {
            float *dst = p->adj_ckpt + (size_t)(k / ckpt_every) * 12 * (size_t)n;
            for (int i = 0; i < 6; i++) {
                SMX_HIP(hipMemcpyAsync(dst + (size_t)i * n, p->f32.T[p->f32.cur][i], sizeof(float) * n,
                                       hipMemcpyDeviceToDevice, p->stream));
                SMX_HIP(hipMemcpyAsync(dst + (size_t)(6 + i) * n, p->f32.S[p->f32.cur][i], sizeof(float) * n,
                                       hipMemcpyDeviceToDevice, p->stream));
            }
        }
        smx_step_scalars s;
        double fH, fE;
        int rc = step_scalars(p, (int)(k / p->N), (int)(k % p->N), &s, &fH, &fE);
        if (rc || (rc = launch_step(p, s, false))) return rc;
        const double th = w * ((double)k + 1.5);
        SMX_HIP(smx_adj_launch_sample(&a, p->f32.T[p->f32.cur][3], p->f32.T[p->f32.cur][4], std::cos(th),
                                      std::sin(th), p->stream));
        if (hist) SMX_HIP(save_hist(k + 1));
    }
    std::vector<double> D(4 * (size_t)p->plane);
    SMX_HIP(hipMemcpyAsync(D.data(), p->adj_D, sizeof(double) * D.size(), hipMemcpyDeviceToHost, p->stream));
    SMX_HIP(hipStreamSynchronize(p->stream));
    double acc = 0.;
    for (size_t i = 0; i < D.size(); i++) acc += D[i] * D[i];
    *J = acc;
    p->adj_ready = true;
    return SMX_OK;
}

int smx_adjoint_backward(smx_plan *p, double *grad)
{
    if (!p || !grad) return arg_fail("null pointer");
    if (!p->adj_ready) {
        snprintf(g_last_error, sizeof g_last_error, "smx_adjoint_backward before smx_adjoint_forward");
        return SMX_ERR_STATE;
    }
    const int64_t n = p->n;
    smx_adj_args a = adj_args(p);
    const long long K = p->adj_K, C = p->adj_C;
    const bool hist = C == 0;
    const long long seglen = hist ? K : C, nck = (K + seglen - 1) / seglen;
    /* the fused transposed step addresses a group of 6 arrays with 32-bit offsets; larger
     * grids take the two gather kernels (in place) */
    const char *gather = getenv("SMX_ADJ_GATHER"); /* "1": the two gather kernels on any grid (tests) */
    const bool fused = 24 * n < ((int64_t)1 << 32) && !(gather && !strcmp(gather, "1"));
    float *lam = nullptr, *seg = nullptr;
    double *G = nullptr, *gdev = nullptr;
    int rc = SMX_OK;
    auto fail = [&](hipError_t e, const char *what) {
        rc = hip_fail(e, what);
        return rc;
    };
    hipError_t e;
    const size_t lam_words = (fused ? 24 : 12) * (size_t)n;
    if ((e = hipMalloc(&lam, sizeof(float) * lam_words)) != hipSuccess) return fail(e, "hipMalloc(lambda)");
    if ((!hist && (e = hipMalloc(&seg, sizeof(float) * 3 * (size_t)a.nbox * (size_t)(C + 1))) != hipSuccess) ||
        (e = hipMalloc(&G, sizeof(double) * 3 * (size_t)a.nbox)) != hipSuccess ||
        (e = hipMalloc(&gdev, sizeof(double) * 3 * (size_t)a.nbox)) != hipSuccess) {
        fail(e, "hipMalloc(adjoint)");
    }
    float *cur = lam, *nxt = lam + 12 * (size_t)n; /* lambda^{k+1}, lambda^k (fused) */
    auto point = [&](float *buf) {
        for (int i = 0; i < 6; i++) {
            a.lT[i] = buf + (size_t)i * n;
            a.lS[i] = buf + (size_t)(6 + i) * n;
        }
    };
    smx_adj_fused_args f;
    std::memset(&f, 0, sizeof f);
    f.L0 = a.L0, f.L1 = a.L1, f.L2 = a.L2, f.sz = a.sz, f.ncomp = n;
    for (int d = 0; d < 3; d++) f.w[d] = a.w[d];
    f.b = p->f32.b[0];
    f.pml = p->f32.pml;
    f.pmlflag = p->f32.pmlflag;
    f.has_pml = p->f32.has_pml;
    int *blocks_dev = nullptr;
    if (fused && rc == SMX_OK) {
        /* blocks whose reach -- x in [x0-1, x0+FX], y in [y0-1, y0+FY], z in [z0-1, z1] --
         * holds a PML cell take the split instantiation; the list puts them first */
        f.tile = 1;
        if (const char *t = getenv("SMX_ADJ_TILE")) f.tile = std::min(std::max(atoi(t), 0), 2);
        const int FX = kAdjTiles[f.tile][0], FY = kAdjTiles[f.tile][1];
        f.gx = (a.L0 + FX - 1) / FX;
        f.gy = (a.L1 + FY - 1) / FY;
        f.fz = kAdjFZ;
        const int gz = (a.L2 + kAdjFZ - 1) / kAdjFZ;
        std::vector<unsigned char> flag(a.L0 + a.L1 + a.L2);
        if ((e = hipMemcpy(flag.data(), p->f32.pmlflag, flag.size(), hipMemcpyDeviceToHost)) != hipSuccess)
            fail(e, "read PML flags");
        auto any = [&](int off, int L, int lo, int hi) {
            for (int v = lo; v <= hi; v++)
                if (flag[off + ((v % L) + L) % L]) return true;
            return false;
        };
        std::vector<int> ids;
        for (int bz = 0; bz < gz; bz++)
            for (int by = 0; by < f.gy; by++)
                for (int bx = 0; bx < f.gx; bx++) {
                    const int x0 = bx * FX, y0 = by * FY, z0 = bz * kAdjFZ;
                    const int z1 = std::min(z0 + kAdjFZ, a.L2);
                    const bool near = f.has_pml && (any(0, a.L0, x0 - 1, x0 + FX) ||
                                                    any(a.L0, a.L1, y0 - 1, y0 + FY) ||
                                                    any(a.L0 + a.L1, a.L2, z0 - 1, z1));
                    ids.push_back((bx + f.gx * (by + f.gy * bz)) | (near ? 1 << 30 : 0));
                    (near ? f.n_pml : f.n_free)++;
                }
        if (rc == SMX_OK && ((e = hipMalloc(&blocks_dev, sizeof(int) * ids.size())) != hipSuccess ||
                             (e = hipMemcpy(blocks_dev, ids.data(), sizeof(int) * ids.size(),
                                            hipMemcpyHostToDevice)) != hipSuccess))
            fail(e, "adjoint block list");
        f.blocks = blocks_dev;
    }
    /* two transposed steps per launch where both steps lie in one segment (SMX_ADJ_TWO=0:
     * one step per launch throughout) */
    const char *two_env = getenv("SMX_ADJ_TWO");
    const bool two = fused && !(two_env && !strcmp(two_env, "0"));
    smx_adj_two_args t2;
    std::memset(&t2, 0, sizeof t2);
    int *blocks2 = nullptr;
    float *lbox = nullptr;
    if (two && rc == SMX_OK) {
        t2.L0 = a.L0, t2.L1 = a.L1, t2.L2 = a.L2, t2.sz = a.sz, t2.ncomp = n;
        for (int d = 0; d < 3; d++) t2.w[d] = a.w[d];
        t2.b = p->f32.b[0], t2.pml = p->f32.pml, t2.pmlflag = p->f32.pmlflag, t2.has_pml = p->f32.has_pml;
        t2.D = p->adj_D, t2.zT = a.zT;
        std::memcpy(t2.box, a.box, sizeof t2.box);
        t2.bx = a.bx, t2.by = a.by, t2.nbox = a.nbox;
        t2.gx = (a.L0 + 31) / 32, t2.gy = (a.L1 + 7) / 8;
        /* z chunks: the fewest rounds of resident blocks (2 per CU) times stages per block */
        int dev = 0, ncu = 0;
        if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
            ncu = 256;
        const long long slots = 2LL * std::max(ncu, 1), layer = (long long)t2.gx * t2.gy;
        long long best = -1;
        for (int nz = (a.L2 + kAdjTwoMaxFZ - 1) / kAdjTwoMaxFZ; nz <= a.L2; nz++) {
            const int fz = (a.L2 + nz - 1) / nz, nzc = (a.L2 + fz - 1) / fz;
            const long long cost = (layer * nzc + slots - 1) / slots * (fz + 4);
            if (best < 0 || cost < best) best = cost, t2.fz = fz, t2.nblocks = (int)(layer * nzc);
            if (fz <= 8) break;
        }
        std::vector<int> ids((size_t)t2.nblocks);
        for (int i = 0; i < t2.nblocks; i++) ids[(size_t)i] = i;
        if ((e = hipMalloc(&blocks2, sizeof(int) * ids.size())) != hipSuccess ||
            (e = hipMemcpy(blocks2, ids.data(), sizeof(int) * ids.size(), hipMemcpyHostToDevice)) != hipSuccess ||
            (e = hipMalloc(&lbox, sizeof(float) * 3 * (size_t)a.nbox)) != hipSuccess)
            fail(e, "adjoint two-step tables");
        t2.blocks = blocks2;
        t2.lbox = lbox;
    }
    if (rc == SMX_OK) {
        point(cur);
        a.G = G;
        (void)hipMemsetAsync(lam, 0, sizeof(float) * lam_words, p->stream);
        (void)hipMemsetAsync(G, 0, sizeof(double) * 3 * (size_t)a.nbox, p->stream);
        const double th_last = p->adj_w * ((double)(K - 1) + 1.5);
        if ((e = smx_adj_launch_inject(&a, std::cos(th_last), std::sin(th_last), p->stream)) != hipSuccess)
            fail(e, "adjoint inject");
    }
    bool first = true;
    for (long long s = nck - 1; s >= 0 && rc == SMX_OK; s--) {
        const long long ks = s * seglen, ke = std::min(ks + seglen, K);
        /* E of the design box after step k - 1 (state k) */
        auto box_state = [&](long long k) -> const float * {
            return hist ? p->adj_hist + (size_t)k * 3 * a.nbox : seg + (size_t)(k - ks) * 3 * a.nbox;
        };
        if (!hist) { /* recompute the segment from its checkpoint, saving the box after every step */
            const float *src = p->adj_ckpt + (size_t)s * 12 * (size_t)n;
            for (int i = 0; i < 6 && rc == SMX_OK; i++) {
                if ((e = hipMemcpyAsync(p->f32.T[p->f32.cur][i], src + (size_t)i * n, sizeof(float) * n,
                                        hipMemcpyDeviceToDevice, p->stream)) != hipSuccess ||
                    (e = hipMemcpyAsync(p->f32.S[p->f32.cur][i], src + (size_t)(6 + i) * n, sizeof(float) * n,
                                        hipMemcpyDeviceToDevice, p->stream)) != hipSuccess)
                    fail(e, "restore checkpoint");
            }
            /* two steps per pass writing the box after both (the split-slot two-step
             * kernel, SMX_ADJ_BOX2=0: one step per launch), monitors left alone */
            const char *box2 = getenv("SMX_ADJ_BOX2");
            const bool pairs = pair_box_ok(p) && !(box2 && !strcmp(box2, "0"));
            if (rc == SMX_OK) rc = save_box(p, seg);
            for (long long k = ks; k < ke && rc == SMX_OK;) {
                smx_step_scalars sc, sc2;
                double fH, fE;
                if ((rc = step_scalars(p, (int)(k / p->N), (int)(k % p->N), &sc, &fH, &fE))) break;
                float *slot = seg + (size_t)(k - ks + 1) * 3 * a.nbox;
                if (pairs && ke - k >= 2) {
                    if ((rc = step_scalars(p, (int)((k + 1) / p->N), (int)((k + 1) % p->N), &sc2, &fH, &fE)) ||
                        (rc = launch_step2(p, sc, sc2, -1, false, slot)))
                        break;
                    k += 2;
                    continue;
                }
                if ((rc = launch_step(p, sc, false, false)) || (rc = save_box(p, slot))) break;
                k++;
            }
        }
        for (long long k = ke - 1; k >= ks && rc == SMX_OK; k--) {
            const float *bcur = box_state(k), *bnext = box_state(k + 1);
            const double th = p->adj_w * ((double)(k - 1) + 1.5);
            if (two && k - 1 >= ks) {
                /* steps k and k-1: lambda^{k+1} -> lambda^{k-1}, injecting into lambda^k inside;
                 * then both gradient terms, and the injection into lambda^{k-1} */
                t2.lT = cur, t2.lS = cur + 6 * (size_t)n, t2.nT = nxt, t2.nS = nxt + 6 * (size_t)n;
                t2.cs = std::cos(th), t2.sn = std::sin(th);
                if ((e = smx_adj_launch_two(&t2, p->stream)) != hipSuccess ||
                    (e = smx_adj_launch_grad2(&a, cur + 3 * (size_t)n, lbox, bnext, bcur, box_state(k - 1),
                                              p->stream)) != hipSuccess) {
                    fail(e, "adjoint two steps");
                    break;
                }
                std::swap(cur, nxt);
                point(cur);
                const double th2 = p->adj_w * ((double)(k - 2) + 1.5);
                if (k - 1 >= 1 &&
                    (e = smx_adj_launch_inject(&a, std::cos(th2), std::sin(th2), p->stream)) != hipSuccess) {
                    fail(e, "adjoint inject");
                    break;
                }
                k--; /* and the loop's own decrement passes step k-1 */
                continue;
            }
            if (fused) {
                /* gradient with lambda_TE^{k+1}, then lambda^k, then the objective's sample of state k */
                if ((e = smx_adj_launch_grad(&a, cur + 3 * (size_t)n, bcur, bnext, p->stream)) != hipSuccess) {
                    fail(e, "adjoint gradient");
                    break;
                }
                f.lT = cur, f.lS = cur + 6 * (size_t)n, f.nT = nxt, f.nS = nxt + 6 * (size_t)n;
                if ((e = smx_adj_launch_fused(&f, p->stream)) != hipSuccess) {
                    fail(e, "adjoint step");
                    break;
                }
                std::swap(cur, nxt);
                point(cur);
                if (k >= 1 && (e = smx_adj_launch_inject(&a, std::cos(th), std::sin(th), p->stream)) != hipSuccess) {
                    fail(e, "adjoint inject");
                    break;
                }
                continue;
            }
            if ((e = smx_adj_launch_a(&a, first ? 0 : 1, bcur, bnext, p->stream)) != hipSuccess) {
                fail(e, "adjoint kernel A");
                break;
            }
            first = false;
            if ((e = smx_adj_launch_b(&a, k >= 1, std::cos(th), std::sin(th), p->stream)) != hipSuccess) {
                fail(e, "adjoint kernel B");
                break;
            }
        }
    }
    if (rc == SMX_OK) {
        if ((e = smx_adj_launch_finish(&a, p->g.invtau_R, gdev, p->stream)) != hipSuccess ||
            (e = hipMemcpyAsync(grad, gdev, sizeof(double) * 3 * (size_t)a.nbox, hipMemcpyDeviceToHost,
                                p->stream)) != hipSuccess ||
            (e = hipStreamSynchronize(p->stream)) != hipSuccess)
            fail(e, "adjoint finish");
    }
    (void)hipStreamSynchronize(p->stream);
    if (lam) (void)hipFree(lam);
    if (seg) (void)hipFree(seg);
    if (G) (void)hipFree(G);
    if (gdev) (void)hipFree(gdev);
    if (blocks_dev) (void)hipFree(blocks_dev);
    if (blocks2) (void)hipFree(blocks2);
    if (lbox) (void)hipFree(lbox);
    return rc;
}

int smx_plan_autotune(smx_plan *p, int trial_steps, int choice[5])
{
    static const int chunks[4] = {8, 16, 32, 64};
    if (!p) return arg_fail("null pointer");
    smx_f32_state &st = p->f32;
    if (p->g.precision == SMX_F32 && trial_steps > 0) {
        const int64_t n = p->n;
        const size_t field = sizeof(float) * (size_t)n;
        float *save = nullptr;
        double *mon = nullptr;
        SMX_HIP(hipMalloc(&save, 12 * field));
        SMX_HIP(hipMalloc(&mon, sizeof(double) * 16 * (size_t)p->plane));
        for (int i = 0; i < 6; i++) {
            SMX_HIP(hipMemcpyAsync((char *)save + i * field, st.T[st.cur][i], field, hipMemcpyDeviceToDevice, p->stream));
            SMX_HIP(hipMemcpyAsync((char *)save + (6 + i) * field, st.S[st.cur][i], field, hipMemcpyDeviceToDevice,
                                   p->stream));
        }
        SMX_HIP(hipMemcpyAsync(mon, p->mon, sizeof(double) * 16 * (size_t)p->plane, hipMemcpyDeviceToDevice, p->stream));
        smx_step_scalars s;
        std::memset(&s, 0, sizeof s);
        hipEvent_t e0, e1;
        SMX_HIP(hipEventCreate(&e0));
        SMX_HIP(hipEventCreate(&e1));
        float best = 1e30f;
        int best_f = st.fused, best_t = st.tile, best_z = st.fz;
        st.two = -1; /* single-step candidates first */
        /* candidate 0: the in-place sweeps; then every fused shape and z chunk */
        for (int cand = 0; cand < 1 + smx_f32_num_tiles() * 4; cand++) {
            st.fused = cand > 0;
            if (cand > 0) {
                st.tile = (cand - 1) / 4;
                st.fz = chunks[(cand - 1) % 4];
                if (!smx_f32_tile_ok(&st, st.tile, st.fz)) continue;
            }
            int rc;
            for (int k = 0; k < 2; k++)
                if ((rc = launch_step(p, s, false, !mon_ring(p)))) return rc;
            SMX_HIP(hipEventRecord(e0, p->stream));
            for (int k = 0; k < trial_steps; k++)
                if ((rc = launch_step(p, s, false, !mon_ring(p)))) return rc;
            SMX_HIP(hipEventRecord(e1, p->stream));
            SMX_HIP(hipEventSynchronize(e1));
            float ms = 0.f;
            SMX_HIP(hipEventElapsedTime(&ms, e0, e1));
            if (ms < best) {
                best = ms;
                best_f = st.fused;
                best_t = st.tile;
                best_z = st.fz;
            }
        }
        st.fused = best_f;
        st.tile = best_t;
        st.fz = best_z;
        /* two steps per pass: timed per step over the same number of steps (even) */
        float best_step = best / (float)trial_steps;
        int best_two = -1, best_two_fz = st.two_fz, best_two_bal = 0;
        const int pairs = (trial_steps + 1) / 2;
        bool zpml = false; /* balanced z chunks differ from uniform ones only with z PML */
        for (int z = 0; z < st.L[2]; z++) zpml |= st.zpml_host[z] != 0;
        const char *two_env = getenv("SMX_F32_TWO"); /* "off": one step per pass only */
        const int ntwo = (two_env && !strcmp(two_env, "off")) || p->pt_comp >= 0 ? 0 : smx_f32_num_two();
        for (int v = 0; v < ntwo; v++)
            for (int ci = 0; ci < 4 * 3; ci++) {
                const int fz = chunks[ci % 4], bal = ci / 4;
                if (!smx_f32_two_ok(&st, v, fz) || (bal > 0 && !zpml)) continue;
                /* with the DFT on, only a variant that samples both steps itself pairs steps */
                if (p->dft_nplanes && (!smx_f32_two_full(&st, v) || p->dft_chunk < 2)) continue;
                st.two = v;
                st.two_fz = fz;
                st.two_bal = bal;
                int rc;
                if ((rc = launch_step2(p, s, s, -1, !mon_ring(p)))) return rc;
                SMX_HIP(hipEventRecord(e0, p->stream));
                for (int k = 0; k < pairs; k++)
                    if ((rc = launch_step2(p, s, s, -1, !mon_ring(p)))) return rc;
                SMX_HIP(hipEventRecord(e1, p->stream));
                SMX_HIP(hipEventSynchronize(e1));
                float ms = 0.f;
                SMX_HIP(hipEventElapsedTime(&ms, e0, e1));
                if (ms / (2.f * pairs) < best_step) {
                    best_step = ms / (2.f * pairs);
                    best_two = v;
                    best_two_fz = fz;
                    best_two_bal = bal;
                }
            }
        st.two = best_two;
        st.two_fz = best_two_fz;
        st.two_bal = best_two_bal;
        for (int k = 0; k < 2; k++)
            for (int i = 0; i < 6; i++) {
                SMX_HIP(hipMemcpyAsync(st.T[k][i], (char *)save + i * field, field, hipMemcpyDeviceToDevice, p->stream));
                SMX_HIP(hipMemcpyAsync(st.S[k][i], (char *)save + (6 + i) * field, field, hipMemcpyDeviceToDevice,
                                       p->stream));
            }
        SMX_HIP(hipMemcpyAsync(p->mon, mon, sizeof(double) * 16 * (size_t)p->plane, hipMemcpyDeviceToDevice, p->stream));
        SMX_HIP(hipStreamSynchronize(p->stream));
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        (void)hipFree(save);
        (void)hipFree(mon);
    }
    if (choice) {
        const bool f = p->g.precision == SMX_F32 && st.fused;
        int fx = 0, fy = 0;
        smx_f32_tile_shape(st.tile, &fx, &fy);
        choice[0] = f ? fx : 0;
        choice[1] = f ? fy : 0;
        choice[2] = f ? st.fz : 0;
        choice[3] = f ? (st.tile >= 3 ? 2 : 1) : 0;
        choice[4] = 0;
        if (p->g.precision == SMX_F32 && st.two >= 0) {
            smx_f32_two_shape(st.two, &choice[0], &choice[1]);
            choice[2] = st.two_fz;
            choice[3] = 3;
            choice[4] = st.two_bal;
        }
    }
    return SMX_OK;
}

int smx_main_calc(const int L[4], const double *Ah, const double *Bh, const double *Ae,
                  const double *Be, const int iBc[3], double *H_t, double *E_t,
                  const int iBTIin[4], const double *Hi, const double *Ei, const double *ExpN2,
                  double *HB, double *EB, double *HT, double *ET, const char *Cont,
                  const int *address_notn)
{
    if (!L || !iBc || !iBTIin || !Cont) return arg_fail("null pointer");
    smx_grid g;
    std::memset(&g, 0, sizeof g);
    std::memcpy(g.L, L, sizeof g.L);
    std::memcpy(g.iBc, iBc, sizeof g.iBc);
    std::memcpy(g.iBTIin, iBTIin, sizeof g.iBTIin);
    g.precision = SMX_F64;
    const int64_t n = (int64_t)L[0] * L[1] * L[2];
    for (int64_t l = 0; l < n; l++) g.Lnotn += Cont[l] != 'n';
    smx_plan *p = nullptr;
    int rc = smx_plan_create_coeffs(&g, Ah, Bh, Ae, Be, Cont, address_notn, &p);
    if (rc) return rc;
    if (!(rc = smx_plan_write_fields(p, H_t, E_t)) && !(rc = smx_plan_write_monitors(p, HB, EB, HT, ET)) &&
        !(rc = smx_plan_step(p, Hi, Ei, ExpN2)) && !(rc = smx_plan_read_fields(p, H_t, E_t)))
        rc = smx_plan_read_monitors(p, HB, EB, HT, ET);
    free_plan(p);
    return rc;
}

} // extern "C"
