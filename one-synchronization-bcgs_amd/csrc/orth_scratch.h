// Device scratch of one orthogonalization call, shared by the host
// schedules (orth_host.cpp) and the allocation-free entry point
// (orth_noalloc.hip).
#pragma once

#include <stddef.h>
#include <stdint.h>

namespace orth {

// Projection partials, the reduced projections (ping-pong), the factor slots,
// the small steps' work matrix, the next pass's coefficients, R and the
// breakdown status.  `fixed` scratch (a region the library does not own the
// allocation of, or buffers frozen while a graph is captured) is never
// reallocated: the passes then use fewer workgroups so that their partials
// fit, and a call that needs more of anything else fails.
struct Scratch {
    double *part = nullptr;
    size_t part_n = 0;
    double *P[2] = {nullptr, nullptr};
    size_t P_n = 0;
    double *stash = nullptr;  // Sig | Om | Cq[3] | Ctmp (6 factor slots)
    double *Tbig = nullptr;   // (n + kMaxS) x kMaxS scratch of the small steps
    size_t Tbig_n = 0;
    double *Cbuf = nullptr;   // coefficients of the next fused pass
    size_t Cbuf_n = 0;
    double *Rd = nullptr;
    size_t Rd_n = 0;
    int *status_d = nullptr;  // 4 ints
    int *tickets = nullptr;   // ticket counter of the synchronization kernel (zero between launches)
    size_t tickets_n = 0;
    bool fixed = false;
};

// Ints reserved for the ticket counter (one used; padded to 256 bytes).
constexpr size_t kTailTickets = 64;

// Doubles of each buffer a call with n columns needs (any block width).
size_t scratch_p_need(int64_t n);
size_t scratch_r_need(int64_t n);
size_t scratch_tbig_need(int64_t n);
size_t scratch_cbuf_need(int64_t n);
size_t scratch_stash_need();

// Enqueue one orth_bcgs_device call on `stream` using only the fixed scratch
// sc (no allocation, no event, no host staging); the first-breakdown
// (block, pivot) pair goes to dInfo (device, 2 ints) when it is not NULL.
// ncu = compute units of the current device.  Returns 0, or 2 for an
// invalid shape, a shape whose buffers exceed sc, or a HIP error.
int run_on_fixed_scratch(int variant, int64_t m, int64_t n, int64_t s, double *dX, int64_t ldX,
                         double *dR, int64_t ldR, int *dInfo, Scratch &sc, int ncu,
                         void *stream);

}  // namespace orth
