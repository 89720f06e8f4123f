/*
 * smx_noalloc.hip -- smx_main_calc_noalloc: main_calc (FDTDcalc/FDTDsplitfieldMetal.c:242)
 * for callers that cannot allocate device memory.
 *
 * smx_main_calc stages the reference's arrays in device buffers it allocates for the
 * call.  This entry point stages them instead in `scratch`, a file-scope device array of
 * SMX_NOALLOC_MAX_BYTES, and runs the same FP64 kernels (kern_f64.hip) on it, so the two
 * give bit-identical results.  It calls no hipMalloc, hipFree or allocating library
 * routine: only hipGetSymbolAddress (once per device), hipMemcpy and the kernel launches.
 *
 * The device array is a symbol of this code object; host code never passes its name to
 * the device.  On its first call on a device the entry point asks the runtime for that
 * device's address of the array and keeps it; every later use goes through that pointer.
 * Every byte the kernels read is copied in from the caller's arrays before the launch,
 * so nothing left by an earlier call is read.  Calls on one device share its array: the
 * entry point holds a per-device lock for the whole call, so calls from several host
 * threads of one process run one after another; work the caller queued on other streams
 * is not ordered against it.
 */
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <mutex>

#include "smx_fdtd.h"
#include "smx_internal.h"

namespace {

static __device__ __align__(256) unsigned char scratch[SMX_NOALLOC_MAX_BYTES];

constexpr int kMaxDevices = 64;
std::mutex g_device_lock[kMaxDevices];
unsigned char *g_scratch_ptr[kMaxDevices]; /* device address of `scratch`, per device */

#define NA_HIP(call)                                                                     \
    do {                                                                                 \
        hipError_t e_ = (call);                                                          \
        if (e_ != hipSuccess) return smx_note_hip(e_, #call);                            \
    } while (0)

/* Device address of this device's `scratch` (the caller holds the device's lock). */
int scratch_of(int dev, unsigned char **out)
{
    if (!g_scratch_ptr[dev]) {
        void *ptr = nullptr;
        NA_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        g_scratch_ptr[dev] = static_cast<unsigned char *>(ptr);
    }
    *out = g_scratch_ptr[dev];
    return SMX_OK;
}

/* Offsets of the staged arrays in the scratch array, each 256-byte aligned. */
struct staging {
    size_t coef[4], H, E, Cont, addr, Hi, Ei, mon[4], total;
};

size_t up256(size_t x) { return (x + 255) & ~(size_t)255; }

staging lay_out(int64_t n, int64_t Lnotn, int64_t plane)
{
    staging s;
    size_t at = 0;
    const size_t clen = sizeof(double) * 3 * (size_t)(n + Lnotn);
    const size_t flen = sizeof(double) * 3 * (size_t)(n + Lnotn + 1);
    for (int i = 0; i < 4; i++) {
        s.coef[i] = at;
        at = up256(at + clen);
    }
    s.H = at;
    at = up256(at + flen);
    s.E = at;
    at = up256(at + flen);
    s.Cont = at;
    at = up256(at + (size_t)n);
    s.addr = at;
    at = up256(at + sizeof(int) * (size_t)n);
    s.Hi = at;
    at = up256(at + sizeof(double) * 2 * (size_t)plane);
    s.Ei = at;
    at = up256(at + sizeof(double) * 2 * (size_t)plane);
    for (int i = 0; i < 4; i++) {
        s.mon[i] = at;
        at = up256(at + sizeof(double) * 4 * (size_t)plane);
    }
    s.total = at;
    return s;
}

} // namespace

extern "C" int smx_main_calc_noalloc(const int L[4], const double *Ah, const double *Bh, const double *Ae,
                                     const double *Be, const int iBc[3], double *H_t, double *E_t,
                                     const int iBTIin[4], const double *Hi, const double *Ei,
                                     const double *ExpN2, double *HB, double *EB, double *HT, double *ET,
                                     const char *Cont, const int *address_notn)
{
    if (!L || !Ah || !Bh || !Ae || !Be || !iBc || !H_t || !E_t || !iBTIin || !Hi || !Ei || !ExpN2 || !HB ||
        !EB || !HT || !ET || !Cont || !address_notn)
        return smx_note_arg("null pointer");
    smx_grid g;
    std::memset(&g, 0, sizeof g);
    std::memcpy(g.L, L, sizeof g.L);
    std::memcpy(g.iBc, iBc, sizeof g.iBc);
    std::memcpy(g.iBTIin, iBTIin, sizeof g.iBTIin);
    g.precision = SMX_F64;
    for (int d = 0; d < 3; d++)
        if (L[d] < 1) return smx_note_arg("L[0..2] must be >= 1");
    const int64_t n = (int64_t)L[0] * L[1] * L[2], plane = (int64_t)L[0] * L[1];
    if (n > (int64_t)SMX_NOALLOC_MAX_BYTES) return smx_note_arg("grid larger than the no-allocation scratch");
    for (int64_t l = 0; l < n; l++) g.Lnotn += Cont[l] != 'n';
    int rc = smx_check_grid(&g, Cont);
    if (rc) return rc;
    const staging st = lay_out(n, g.Lnotn, plane);
    if (st.total > (size_t)SMX_NOALLOC_MAX_BYTES) {
        char msg[200];
        snprintf(msg, sizeof msg, "grid needs %zu bytes of device scratch; smx_main_calc_noalloc holds %zu",
                 st.total, (size_t)SMX_NOALLOC_MAX_BYTES);
        return smx_note_arg(msg);
    }

    int count = 0, dev = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count < 1) {
        smx_note_hip(e == hipSuccess ? hipErrorNoDevice : e, "no HIP device");
        return SMX_ERR_NO_DEVICE;
    }
    NA_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return smx_note_arg("device index beyond the no-allocation table");
    std::lock_guard<std::mutex> hold(g_device_lock[dev]);
    unsigned char *base = nullptr;
    if ((rc = scratch_of(dev, &base))) return rc;

    const size_t clen = sizeof(double) * 3 * (size_t)(n + g.Lnotn);
    const size_t flen = sizeof(double) * 3 * (size_t)(n + g.Lnotn + 1);
    const size_t mlen = sizeof(double) * 4 * (size_t)plane;
    const double *coef[4] = {Ah, Bh, Ae, Be};
    double *mon[4] = {HB, EB, HT, ET};
    for (int i = 0; i < 4; i++) NA_HIP(hipMemcpy(base + st.coef[i], coef[i], clen, hipMemcpyHostToDevice));
    NA_HIP(hipMemcpy(base + st.H, H_t, flen, hipMemcpyHostToDevice));
    NA_HIP(hipMemcpy(base + st.E, E_t, flen, hipMemcpyHostToDevice));
    NA_HIP(hipMemcpy(base + st.Cont, Cont, (size_t)n, hipMemcpyHostToDevice));
    NA_HIP(hipMemcpy(base + st.addr, address_notn, sizeof(int) * (size_t)n, hipMemcpyHostToDevice));
    NA_HIP(hipMemcpy(base + st.Hi, Hi, sizeof(double) * 2 * (size_t)plane, hipMemcpyHostToDevice));
    NA_HIP(hipMemcpy(base + st.Ei, Ei, sizeof(double) * 2 * (size_t)plane, hipMemcpyHostToDevice));
    for (int i = 0; i < 4; i++) NA_HIP(hipMemcpy(base + st.mon[i], mon[i], mlen, hipMemcpyHostToDevice));

    smx_f64_args a;
    std::memset(&a, 0, sizeof a);
    a.g.L0 = L[0];
    a.g.L1 = L[1];
    a.g.L2 = L[2];
    a.g.iBc0 = iBc[0];
    a.g.iBc1 = iBc[1];
    a.g.iBc2 = iBc[2];
    a.g.iB = iBTIin[0];
    a.g.iT = iBTIin[1];
    a.g.iI = iBTIin[2];
    a.g.sgn = (double)(iBTIin[3] / (iBTIin[3] < 0 ? -iBTIin[3] : iBTIin[3])); /* F:253 */
    a.g.n = n;
    a.g.Lnotn = g.Lnotn;
    a.Ah = reinterpret_cast<const double *>(base + st.coef[0]);
    a.Bh = reinterpret_cast<const double *>(base + st.coef[1]);
    a.Ae = reinterpret_cast<const double *>(base + st.coef[2]);
    a.Be = reinterpret_cast<const double *>(base + st.coef[3]);
    a.H = reinterpret_cast<double *>(base + st.H);
    a.E = reinterpret_cast<double *>(base + st.E);
    a.Cont = reinterpret_cast<const char *>(base + st.Cont);
    a.addr = reinterpret_cast<const int *>(base + st.addr);
    a.srcH_plane = reinterpret_cast<const double *>(base + st.Hi);
    a.srcE_plane = reinterpret_cast<const double *>(base + st.Ei);
    a.Hinc = a.Einc = nullptr;
    a.HB = reinterpret_cast<double *>(base + st.mon[0]);
    a.EB = reinterpret_cast<double *>(base + st.mon[1]);
    a.HT = reinterpret_cast<double *>(base + st.mon[2]);
    a.ET = reinterpret_cast<double *>(base + st.mon[3]);
    a.s.monH_re = ExpN2[0];
    a.s.monH_im = ExpN2[1];
    a.s.monE_re = ExpN2[2];
    a.s.monE_im = ExpN2[3];
    NA_HIP(smx_launch_f64_step(&a, nullptr));

    NA_HIP(hipMemcpy(H_t, base + st.H, flen, hipMemcpyDeviceToHost));
    NA_HIP(hipMemcpy(E_t, base + st.E, flen, hipMemcpyDeviceToHost));
    for (int i = 0; i < 4; i++) NA_HIP(hipMemcpy(mon[i], base + st.mon[i], mlen, hipMemcpyDeviceToHost));
    return SMX_OK;
}
