// Allocation-free entry point of liborth_hip.so: orth_bcgs_device_noalloc.
//
// Its scratch is the file-scope device array `scratch` below (48 MiB of
// doubles per device, held by every device the library runs on until the
// process ends).  On a device's first call, with that device current, the
// array's device address is looked up once with hipGetSymbolAddress and
// kept per device; host code uses only that address (the array's name is a
// host-side stand-in, not a device address).  The address is carved into
// the buffers of one call (factor slots, status, the synchronization
// kernel's ticket counter, reduced projections, R, the small steps' work
// matrix, the next pass's coefficients, and the projection partials, which
// take the rest).  Nothing is allocated: no hipMalloc, no event, no pinned
// staging.  Every call on a device shares that device's array, and a call
// clears what must start at zero with kernels (R, the factor slots, the
// status, the ticket counter), so that it can be captured in a graph.  Calls
// that overlap on one device race on the array.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>

#include <mutex>

#include "../../include/orth_ext.h"
#include "orth_kernels.h"
#include "orth_scratch.h"

namespace orth {

// Widest slab (columns) the fixed buffers are carved for.
constexpr int64_t kNoAllocMaxN = 512;
// Doubles of the fixed region: 48 MiB.
constexpr size_t kNoAllocDoubles = (size_t)6 << 20;

static __device__ __align__(256) double scratch[kNoAllocDoubles];

namespace {

constexpr int kMaxDev = 64;

struct FixedDev {
    std::mutex mu;
    bool init = false;
    int ncu = 0;
    Scratch sc;
};

FixedDev g_fixed[kMaxDev];

// One-time set-up on device dev (current): the region's device address,
// carved into the call's buffers (each start 256-byte aligned).
bool fixed_setup(FixedDev &f, int dev)
{
    if (f.init) return true;
    void *base = nullptr;
    if (hipGetSymbolAddress(&base, HIP_SYMBOL(scratch)) != hipSuccess || base == nullptr) {
        (void)hipGetLastError();
        fprintf(stderr, "orth_bcgs_device_noalloc: no device address for the fixed scratch on "
                        "device %d\n", dev);
        return false;
    }
    if (hipDeviceGetAttribute(&f.ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    double *q = (double *)base;
    size_t left = kNoAllocDoubles;
    auto take = [&](size_t n) {
        n = (n + 31) & ~(size_t)31;
        double *r = q;
        q += n;
        left -= n;
        return r;
    };
    Scratch &sc = f.sc;
    sc.stash = take(scratch_stash_need());
    sc.status_d = (int *)take(32);
    sc.tickets_n = kTailTickets;
    sc.tickets = (int *)take(kTailTickets / 2);
    sc.P_n = scratch_p_need(kNoAllocMaxN);
    sc.P[0] = take(sc.P_n);
    sc.P[1] = take(sc.P_n);
    sc.Rd_n = scratch_r_need(kNoAllocMaxN);
    sc.Rd = take(sc.Rd_n);
    sc.Tbig_n = scratch_tbig_need(kNoAllocMaxN);
    sc.Tbig = take(sc.Tbig_n);
    sc.Cbuf_n = scratch_cbuf_need(kNoAllocMaxN);
    sc.Cbuf = take(sc.Cbuf_n);
    sc.part = q;
    sc.part_n = left;
    sc.fixed = true;
    f.init = true;
    return true;
}

}  // namespace
}  // namespace orth

extern "C" int orth_bcgs_device_noalloc(int variant, int64_t m, int64_t n, int64_t s, double *dX,
                                        int64_t ldX, double *dR, int64_t ldR, int *dInfo,
                                        void *stream)
{
    using namespace orth;
    if (n > kNoAllocMaxN) {
        fprintf(stderr, "orth_bcgs_device_noalloc: n = %lld columns exceeds the fixed scratch "
                        "limit of %lld\n", (long long)n, (long long)kNoAllocMaxN);
        return 2;
    }
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDev) return 2;
    FixedDev &f = g_fixed[dev];
    std::lock_guard<std::mutex> lk(f.mu);
    if (!fixed_setup(f, dev)) return 2;
    return run_on_fixed_scratch(variant, m, n, s, dX, ldX, dR, ldR, dInfo, f.sc, f.ncu, stream);
}
