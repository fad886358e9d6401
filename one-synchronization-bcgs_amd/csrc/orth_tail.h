// The synchronization at the end of a fused pass, inside the pass launch:
// the fixed-order sum of the workgroups' projection partials and the small
// step that follows it (the reference's MPI_Allreduce and the s x s work
// after it, SRC/orth.c:114, :133, :188, :197).  Workgroups never wait for
// one another: each takes a ticket after storing its partial, and only the
// last workgroup of a group (then the last group) goes on.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>

#include "orth_kernels.h"
#include "orth_small.h"

namespace orth {

// LDS of the tail: the small step's region and the ticket flag after it.
constexpr size_t kTailLdsBytes = (size_t)(kSmallLds + 1) * sizeof(double);

// Sum of the partials [k0, k1) of element e in workgroup order, starting
// from the first (no added zero, so the sign of an exact zero is kept).
// The loads of W partials are issued together (indices past the range are
// clamped to its last partial and their values not added), then added in
// order: one memory round trip per W partials.
template <int W>
static __device__ __forceinline__ double part_range_sum_w(const double *p, int64_t E, int n)
{
    double s = 0.0;
    for (int b = 0; b < n; b += W) {
        double v[W];
#pragma unroll
        for (int i = 0; i < W; i++) v[i] = p[(int64_t)min(b + i, n - 1) * E];
#pragma unroll
        for (int i = 0; i < W; i++) {
            const int k = b + i;
            s = k == 0 ? v[i] : k < n ? s + v[i] : s;
        }
    }
    return s;
}

static __device__ __forceinline__ double part_range_sum(const double *part, int64_t E, int64_t e,
                                                        int k0, int k1)
{
    const double *p = part + (int64_t)k0 * E + e;
    const int n = k1 - k0;
    if (n <= 8) return part_range_sum_w<8>(p, E, n);
    if (n <= 16) return part_range_sum_w<16>(p, E, n);
    return part_range_sum_w<32>(p, E, n);
}

// Publish what this workgroup stored and take a ticket of `counter`
// (release at device scope, no invalidation: the workgroups still streaming
// keep what they cached); returns, in every thread, whether the workgroup
// drew ticket `last`.  Every wave waits for its own stores before the
// barrier, so thread 0's release covers the whole workgroup's stores.
static __device__ __forceinline__ bool tail_ticket(int *counter, int last, int *flag)
{
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    __syncthreads();
    if (threadIdx.x == 0) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        flag[0] = atomicAdd(counter, 1) == last ? 1 : 0;
    }
    __syncthreads();
    const bool won = flag[0] != 0;
    if (won) {
        // Acquire every other workgroup's published stores.
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    }
    return won;
}

// The address of the kernel's argument segment, taken at the kernel entry
// and kept in a vector register until the tail: a scalar register pair would
// be held through the whole pass, where scalar registers are scarce.
static __device__ __forceinline__ uint64_t kernarg_stash()
{
    uint64_t p = (uint64_t)__builtin_amdgcn_kernarg_segment_ptr();
    asm volatile("" : "+v"(p));
    return p;
}

// The kernel's argument segment (PassArgs is each pass kernel's only
// argument, at offset 0) from the stashed address; reads through it stay at
// the tail, so that none of the tail's arguments occupies a register
// during the pass.
typedef const char __attribute__((address_space(4))) *kernarg_ptr;
static __device__ __forceinline__ kernarg_ptr kernarg_base(uint64_t stash)
{
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)stash);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(stash >> 32));
    return (kernarg_ptr)((uint64_t)hi << 32 | lo);
}

static __device__ __forceinline__ TailArgs tail_args(kernarg_ptr kp)
{
    static_assert(sizeof(TailArgs) % 8 == 0 && offsetof(PassArgs, tail) % 8 == 0,
                  "tail arguments are copied as 8-byte words");
    typedef const uint64_t __attribute__((address_space(4))) *wptr;
    const wptr src = (wptr)(kp + offsetof(PassArgs, tail));
    uint64_t w[sizeof(TailArgs) / 8];
#pragma unroll
    for (int i = 0; i < (int)(sizeof(TailArgs) / 8); i++) w[i] = src[i];
    TailArgs t;
    __builtin_memcpy(&t, w, sizeof(t));
    return t;
}

// Every thread of every workgroup of the pass calls this after its
// workgroup's partial is stored; it returns at once unless the pass
// synchronizes itself (PassArgs::tail_on).  stash: kernarg_stash() from the
// kernel entry; nt: the workgroup's thread count (block size and grid size
// come from the arguments, so that nothing of the dispatch stays live
// through the pass).  sm: the pass's dynamic LDS, at least pass_tail_lds()
// bytes; the pass no longer uses it.
static __device__ __forceinline__ void pass_tail(const PassArgs &a, double *sm, uint64_t stash,
                                                 int nt)
{
    const kernarg_ptr kp = kernarg_base(stash);
    typedef const int __attribute__((address_space(4))) *iptr;
    if (*(iptr)(kp + offsetof(PassArgs, tail_on)) == 0) return;
    const TailArgs t = tail_args(kp);
    const int nwg = t.nwg;
    const int64_t E = (int64_t)a.na * a.nb;
    const int ngrp = (nwg + t.grp - 1) / t.grp;
    const int g = (int)blockIdx.x / t.grp;
    const int k0 = g * t.grp, k1 = min(nwg, k0 + t.grp);
    int *flag = (int *)(sm + kSmallLds);

    // The pass's own LDS and memory traffic is finished: tail_ticket waits
    // for every wave's loads and stores and synchronizes the workgroup.
    if (!tail_ticket(&t.tickets[g], k1 - k0 - 1, flag)) return;
    // Last of its group: reset the counter, sum the group's partials.
    if (threadIdx.x == 0) t.tickets[g] = 0;
    double *dst = ngrp == 1 ? t.out : t.lvl + (int64_t)g * E;
    for (int64_t e = threadIdx.x; e < E; e += nt) dst[e] = part_range_sum(a.part, E, e, k0, k1);
    if (ngrp > 1) {
        if (!tail_ticket(&t.tickets[kTailTop], ngrp - 1, flag)) return;
        if (threadIdx.x == 0) t.tickets[kTailTop] = 0;
        for (int64_t e = threadIdx.x; e < E; e += nt)
            t.out[e] = part_range_sum(t.lvl, E, e, 0, ngrp);
    }
    if (!t.small) return;
    __syncthreads();
    small_step(nt, t.sm, sm);
}

}  // namespace orth
