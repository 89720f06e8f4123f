/*
 * smx_stage.h -- device helpers shared by the LDS-staged z-march kernels (kern_f32.hip,
 * kern_adjoint.hip): XCD-aware block numbering, index wrap, buffer resources, the
 * asynchronous HBM -> LDS copies, a compile-time loop over them and the counted waits
 * that order them.
 */
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

namespace {

constexpr int kXCD = 8; /* MI355X: 8 accelerator complex dies, blocks dealt round-robin */

/* Hardware block h runs on XCD h % 8; give each XCD a contiguous run of logical ids. */
__device__ inline int xcd_logical(int h, int nblocks)
{
    const int q = nblocks / kXCD, r = nblocks % kXCD;
    const int x = h % kXCD, s = h / kXCD;
    return x * q + (x < r ? x : r) + s;
}

__device__ inline int wrap_index(int v, int L)
{
    const int r = v % L;
    return r < 0 ? r + L : r;
}

constexpr unsigned kBufWord3 = 0x00020000u; /* raw 32-bit buffer resource (gfx9 family) */

/* Buffer resource over one group of 6 component arrays (6*ncomp floats). */
__device__ inline __amdgpu_buffer_rsrc_t group_rsrc(const float *base, int64_t ncomp)
{
    return __builtin_amdgcn_make_buffer_rsrc((void *)base, (short)0, (int)(unsigned)(24 * ncomp), (int)kBufWord3);
}

/* LDS address of a generic pointer into the kernel's LDS array: its low 32 bits (a
 * plain cast would add a null test, which the code generator cannot emit when the
 * pointer folds to the array's base) */
#define SMX_LDS(p) ((__attribute__((address_space(3))) void *)(size_t)(unsigned)(size_t)(p))

/* s_waitcnt vmcnt(VM) lgkmcnt(0), then s_barrier; the empty asm statements keep the
 * compiler from moving LDS accesses across it. */
template <int VM>
__device__ inline void wait_vm_barrier()
{
    static_assert(VM >= 0 && VM < 64, "vmcnt field");
    asm volatile("" ::: "memory");
    __builtin_amdgcn_s_waitcnt((VM & 0xF) | ((VM >> 4) << 14) | (0x7 << 4));
    __builtin_amdgcn_s_barrier();
    asm volatile("" ::: "memory");
}

/* f(integral_constant<int, J>) for J = J0, J0 + STEP, ... below END, unrolled at compile
 * time (the copies of a stage: which array, component and run a copy takes is then a
 * constant, and the run offsets stay in registers) */
template <int J, int END, int STEP, typename F>
__device__ __forceinline__ void static_for(F &f)
{
    if constexpr (J < END) {
        f(std::integral_constant<int, J>{});
        static_for<J + STEP, END, STEP>(f);
    }
}

/* Wait for every outstanding vector memory operation of this wave.  Used right after
 * the rare direct loads inside the staged loop (source planes, monitors, lossy/metal
 * cells): the compiler does not count the asynchronous copies in its own vmcnt
 * bookkeeping, so a load whose result it still thinks pending at a join would make it
 * wait for vmcnt(0) -- every copy in flight -- on the common path too. */
__device__ inline void retire_loads()
{
    asm volatile("" ::: "memory");
    __builtin_amdgcn_s_waitcnt((0x7 << 4) | (0xF << 8)); /* vmcnt(0) */
    asm volatile("" ::: "memory");
}

} // namespace
