// Host side of liborth_hip.so: the C-ABI entry points of include/orth.h and
// include/orth_ext.h, and the per-variant schedules of passes over the row
// slab.
//
// Every MPI_Allreduce of the reference becomes one projection pass whose
// per-workgroup partials are summed in a fixed order on the device; the s x s
// work between passes runs in a one-workgroup kernel.  Nothing returns to
// the host between blocks: a call enqueues its whole schedule on one stream.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <math.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <climits>
#include <fstream>
#include <iomanip>
#include <mutex>
#include <sstream>
#include <string>
#include <vector>

#include <unistd.h>

#include "../../include/orth.h"
#include "../../include/orth_ext.h"
#include "orth_kernels.h"
#include "orth_scratch.h"

namespace orth {
namespace {

constexpr int kMaxDevices = 64;
constexpr int kProjTargetWG = 1024;
// Rows per CU from which a fused pass synchronizes inside its own launch
// (2^23 rows on 256 CUs); shorter slabs end each pass with one
// k_reduce_small launch.
constexpr int64_t kTailMinRowsPerCU = 32768;

struct DevCtx {
    std::mutex mu;
    bool init = false;
    hipEvent_t done = nullptr;  // end of the last enqueued call
    bool done_valid = false;
    Scratch sc;
    int ncu = 0;
    double *Xd = nullptr;
    size_t Xd_n = 0;
    int *status_h = nullptr;
    orth_status_t last = {0, 0, 0, 0, 0};
    int last_passes = 0, last_syncs = 0;
    // Optional per-launch event timing, by kernel class (ORTH_K_*).
    bool ktime = false;
    std::vector<hipEvent_t> ev_pool;
    size_t ev_used = 0;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> kt[ORTH_K_COUNT];
    double kt_flops[ORTH_K_COUNT] = {0};  // algorithmic flops of the timed launches
    double kt_bytes[ORTH_K_COUNT] = {0};  // algorithmic HBM bytes of the timed launches
};

DevCtx g_ctx[kMaxDevices];

// Shift constant of the Pythagorean Cholesky steps (orth_set_pythagorean_shift).
std::atomic<double> g_pshift{0.0};

thread_local int tl_hip_error = 0;

bool ok(hipError_t e)
{
    if (e != hipSuccess) {
        tl_hip_error = (int)e;
        return false;
    }
    return true;
}

bool ensure(double *&p, size_t &cap, size_t n)
{
    if (cap >= n && p) return true;
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
    if (!ok(hipMalloc(&p, n * sizeof(double)))) return false;
    cap = n;
    return true;
}

// Grow one scratch buffer of sc to n doubles (fixed scratch: fail instead).
bool grow(Scratch &sc, double *&p, size_t &cap, size_t n)
{
    if (cap >= n && p) return true;
    if (sc.fixed) {
        tl_hip_error = (int)hipErrorOutOfMemory;
        return false;
    }
    return ensure(p, cap, n);
}

int current_device()
{
    int d = 0;
    (void)hipGetDevice(&d);
    return d;
}

bool ctx_init(DevCtx &c)
{
    if (c.init) return true;
    if (!ok(hipEventCreateWithFlags(&c.done, hipEventDisableTiming))) return false;
    if (!ok(hipMalloc(&c.sc.stash, scratch_stash_need() * sizeof(double)))) return false;
    if (!ok(hipMalloc(&c.sc.status_d, 4 * sizeof(int)))) return false;
    if (!ok(hipMalloc(&c.sc.tickets, kTailTickets * sizeof(int)))) return false;
    c.sc.tickets_n = kTailTickets;
    if (!ok(hipHostMalloc(&c.status_h, 4 * sizeof(int), hipHostMallocDefault))) return false;
    int dev = 0;
    (void)hipGetDevice(&dev);
    if (!ok(hipDeviceGetAttribute(&c.ncu, hipDeviceAttributeMultiprocessorCount, dev)))
        return false;
    c.init = true;
    return true;
}

// Environment switches that select kernels (alternate and fallback paths,
// read at each call): part of the key of captured batch graphs.
int64_t env_bits()
{
    static const char *const names[] = {"ORTH_UNFUSED", "ORTH_UPDATE_ROWWISE", "ORTH_PASS256",
                                        "ORTH_GENERIC_PASS", "ORTH_MIXED_PASS32",
                                        "ORTH_PASS_TILE32", "ORTH_PASS_LOAD8",
                                        "ORTH_SEPARATE_REDUCE", "ORTH_PASS_REGPF", "ORTH_PASS16D_ONEWG",
                                        "ORTH_PASS16D_MINCOLS", "ORTH_REDUCE_LAUNCH",
                                        "ORTH_IN_PASS_SYNC"};
    int64_t b = 0;
    for (int i = 0; i < (int)(sizeof(names) / sizeof(names[0])); i++)
        if (getenv(names[i])) b |= (int64_t)1 << i;
    return b;
}

int block_template(int64_t s)
{
    if (s <= 4) return 4;
    if (s <= 8) return 8;
    if (s <= 16) return 16;
    if (s <= 32) return 32;
    return 0;
}

// Records start/stop events around one launch when timing is enabled.
template <class F>
bool timed(DevCtx &c, int kind, hipStream_t st, double flops, double bytes, F launch)
{
    if (!c.ktime) return ok(launch());
    c.kt_flops[kind] += flops;
    c.kt_bytes[kind] += bytes;
    while (c.ev_pool.size() < c.ev_used + 2) {
        hipEvent_t e;
        if (!ok(hipEventCreate(&e))) return false;
        c.ev_pool.push_back(e);
    }
    hipEvent_t a = c.ev_pool[c.ev_used], b = c.ev_pool[c.ev_used + 1];
    c.ev_used += 2;
    if (!ok(hipEventRecord(a, st))) return false;
    if (!ok(launch())) return false;
    if (!ok(hipEventRecord(b, st))) return false;
    c.kt[kind].push_back({a, b});
    return true;
}

// One orthogonalization call: the schedule of passes for one variant.
class Run {
  public:
    Run(DevCtx &c, Scratch &sc, int64_t m, int64_t n, int64_t s, double *X, int64_t ldX,
        hipStream_t st)
        : c_(c), sc_(sc), m_(m), n_(n), s_(s), X_(X), ldX_(ldX), st_(st)
    {
        S_ = block_template(s);
        N_ = 1 + (n - s + s - 1) / s;
        n1_ = std::max(16, 2 * S_);
        ldR_ = n;
        Sig_ = sc.stash;
        Om_ = Sig_ + kSlot;
        Cq_ = Om_ + kSlot;
        Ctmp_ = Cq_ + 3 * kSlot;
    }

    bool prepare()
    {
        // Largest projection: (n + 3 s) x (3 s) covers every schedule; the
        // extra kMaxS columns keep coefficient reads past a ragged block in
        // bounds.
        const size_t pmax = scratch_p_need(n_);
        if (sc_.P_n < pmax || !sc_.P[0] || !sc_.P[1]) {
            if (sc_.fixed) {
                tl_hip_error = (int)hipErrorOutOfMemory;
                return false;
            }
            for (int k = 0; k < 2; k++) {
                if (sc_.P[k]) (void)hipFree(sc_.P[k]);
                sc_.P[k] = nullptr;
            }
            sc_.P_n = 0;
            if (!ok(hipMalloc(&sc_.P[0], pmax * sizeof(double)))) return false;
            if (!ok(hipMalloc(&sc_.P[1], pmax * sizeof(double)))) return false;
            sc_.P_n = pmax;
        }
        // Growable scratch reserves the partials of the widest plan up front;
        // fixed scratch is checked pass by pass.
        if (!sc_.fixed) {
            const size_t partmax = (size_t)kProjTargetWG * (n_ + 3 * kMaxS) * (3 * kMaxS);
            if (!grow(sc_, sc_.part, sc_.part_n, partmax)) return false;
        }
        const size_t rmax = scratch_r_need(n_);
        if (!grow(sc_, sc_.Rd, sc_.Rd_n, rmax)) return false;
        if (!grow(sc_, sc_.Tbig, sc_.Tbig_n, scratch_tbig_need(n_))) return false;
        if (!grow(sc_, sc_.Cbuf, sc_.Cbuf_n, scratch_cbuf_need(n_))) return false;
        if (!ok(launch_fill(sc_.Rd, rmax, st_))) return false;
        if (!ok(launch_fill(sc_.stash, scratch_stash_need(), st_))) return false;
        if (!ok(launch_fill_int(sc_.status_d, 4, st_))) return false;
        if (sc_.tickets && !ok(launch_fill_int(sc_.tickets, sc_.tickets_n, st_))) return false;
        Rd_ = sc_.Rd;
        return true;
    }

    int64_t bw(int64_t b) const { return b == 0 ? s_ : std::min<int64_t>(s_, n_ - b * s_); }

    // P <- [X(:, a0:a0+na)]^T [X(:, b0:b0+nb)], then the small step `after`
    // on it when given; returns the buffer (ld na).
    double *proj(int64_t a0, int64_t na, int64_t b0, int64_t nb, const SmallArgs *after = nullptr)
    {
        ProjPlan pl = plan_proj(m_, na, nb, std::max(1, (int)(kProjTargetWG * wg_scale_)));
        if (sc_.fixed && (size_t)pl.nwg * na * nb > sc_.part_n) {
            // Fixed scratch: fewer row chunks, so that the partials fit.
            const int64_t cap = (int64_t)(sc_.part_n / ((size_t)na * nb));
            if (cap < 1) return nullptr;
            pl = plan_proj(m_, na, nb, (int)std::min<int64_t>(INT_MAX, cap * pl.ngy));
        }
        const size_t need = (size_t)pl.nwg * na * nb;
        if (!grow(sc_, sc_.part, sc_.part_n, need)) return nullptr;
        if (pl.lds_bytes > 64 * 1024) return nullptr;
        // Algorithmic work: 2 flops per row and output entry; each column of
        // A union B read once.
        const int64_t lo = std::max(a0, b0), hi = std::min(a0 + na, b0 + nb);
        const int64_t ucols = na + nb - std::max<int64_t>(0, hi - lo);
        if (!timed(c_, ORTH_K_PROJ, st_, 2.0 * m_ * na * nb, 8.0 * m_ * ucols, [&] {
                return launch_proj(pl, X_, ldX_, m_, a0, na, b0, nb, sc_.part, st_);
            }))
            return nullptr;
        pcur_ ^= 1;
        double *P = sc_.P[pcur_];
        if (!sync(pl.nwg, na, nb, P, after)) return nullptr;
        passes_++;
        syncs_++;
        return P;
    }

    // The library's pointers and constants of a small step.
    void finish_small(SmallArgs &a) const
    {
        a.m = m_;
        a.R = Rd_;
        a.ldR = ldR_;
        a.Sig = Sig_;
        a.Om = Om_;
        a.Cq = Cq_;
        a.Ctmp = Ctmp_;
        a.status = sc_.status_d;
        a.Tbig = sc_.Tbig;
        a.pshift = g_pshift.load();
        if (fused_) {
            a.Cout = sc_.Cbuf;
            a.n1 = n1_;
            a.S = S_;
        }
    }

    bool small(SmallArgs a)
    {
        finish_small(a);
        if (!timed(c_, ORTH_K_SMALL, st_, 0.0, 0.0, [&] { return launch_small(a, st_); }))
            return false;
        if (debug_) dump(a);
        return true;
    }

    static SmallArgs sa(int op, int flags, int64_t blk, int64_t c, int64_t wb, int64_t wn,
                        double *P, int64_t na)
    {
        SmallArgs a;
        memset(&a, 0, sizeof(a));
        a.op = op;
        a.flags = flags;
        a.blk = (int)blk;
        a.c = (int)c;
        a.wb = (int)wb;
        a.wn = (int)wn;
        a.P = P;
        a.na = (int)na;
        return a;
    }

    static Upd upd(int64_t col0, int64_t w, int64_t c, const double *W, int64_t ldW,
                   const double *T)
    {
        Upd u;
        u.col0 = col0;
        u.w = (int)w;
        u.c = (int)c;
        u.W = W;
        u.ldW = ldW;
        u.T = T;
        return u;
    }

    // Algorithmic work of an update: 2 c w + w^2 flops per row (product and
    // triangular solve); prefix and target read once, target written once.
    double upd_flops(const Upd &u) const
    {
        return (double)m_ * (2.0 * u.c * u.w + (u.T ? (double)u.w * u.w : 0.0));
    }
    bool update(const Upd &u0)
    {
        passes_++;
        return timed(c_, ORTH_K_UPDATE, st_, upd_flops(u0), 8.0 * m_ * (u0.c + 2.0 * u0.w),
                     [&] { return launch_update(S_, X_, ldX_, m_, &u0, 1, st_); });
    }
    bool update2(const Upd &u0, const Upd &u1)
    {
        Upd u[2] = {u0, u1};
        passes_++;
        const double bytes = 8.0 * m_ * (u1.c + u1.w + u0.w + u1.w);
        return timed(c_, ORTH_K_UPDATE, st_, upd_flops(u0) + upd_flops(u1), bytes,
                     [&] { return launch_update(S_, X_, ldX_, m_, u, 2, st_); });
    }

    // Intra-block QR of columns [col0, col0+w): shifted CholQR3; the factor
    // R3 R2 R1 goes to Rdst (ld ldRdst).  Replaces Trilinos TSQR
    // factorExplicit (SRC/orth.c:106-107 and the other TSQR call sites).
    bool intra_qr(int64_t blk, int64_t col0, int64_t w, double *Rdst, int64_t ldRdst)
    {
        const int ops[3] = {OP_CQR1, OP_CQR2, OP_CQR3};
        for (int k = 0; k < 3; k++) {
            SmallArgs a = sa(ops[k], 0, blk, 0, w, 0, nullptr, w);
            a.Rdst = Rdst;
            a.ldRdst = ldRdst;
            if (!proj(col0, w, col0, w, &a)) return false;
            if (!update(upd(col0, w, 0, nullptr, 0, Cq_ + k * kSlot))) return false;
        }
        return true;
    }

    // ---------------- variant schedules ----------------

    // BCGSI+P-1S (SRC/orth.c:48-222).
    bool p1s()
    {
        if (!intra_qr(0, 0, s_, Rd_, ldR_)) return false;  // :105-107
        if (N_ == 1) return true;
        const int64_t w1 = bw(1);
        {
            // [Q1 X2]^T X2 (:111-114); Sigma, R12 = S (:115-119).
            SmallArgs q = sa(OP_OPEN_PYTH, 0, 1, s_, w1, 0, nullptr, s_ + w1);
            double *P = proj(0, s_ + w1, s_, w1, &q);
            if (!P) return false;
            // X2 = (X2 - Q1 S) Sigma^{-1} (:121-124).
            if (!update(upd(s_, w1, s_, P, s_ + w1, Sig_))) return false;
        }
        int64_t c = s_;
        for (int64_t b = 1; b < N_; b++) {
            const int64_t wb = bw(b), wn = (b + 1 < N_) ? bw(b + 1) : 0;
            const int64_t na = c + wb + wn;
            // The one sync: [Q Q^_b]^T [Q^_b X_n] and X_n^T X_n (:127-133, :182-197).
            SmallArgs q = sa(OP_CLOSE_OPEN, F_PYTH_NEXT, b, c, wb, wn, nullptr, na);
            double *P = proj(0, na, c, wb + wn, &q);
            if (!P) return false;
            // Q_b = (Q^_b - Q Y) Omega^{-1} (:139-142, :206-209) and the next
            // block's first pass (:173-176) in one pass.
            Upd u0 = upd(c, wb, c, P, na, Om_);
            if (wn > 0) {
                Upd u1 = upd(c + wb, wn, c + wb, P + wb * na, na, Sig_);
                if (!update2(u0, u1)) return false;
            } else if (!update(u0)) {
                return false;
            }
            c += wb;
        }
        return true;
    }

    // BCGSI+A-1S (SRC/orth.c:224-367).
    bool a1s()
    {
        if (!intra_qr(0, 0, s_, Rd_, ldR_)) return false;  // :269-271
        if (N_ == 1) return true;
        const int64_t w1 = bw(1);
        {
            // S = Q1^T X2 (:275-278); V2 = X2 - Q1 S (:280-281).
            SmallArgs q = sa(OP_PLAIN_S, 0, 1, s_, w1, 0, nullptr, s_);
            double *P = proj(0, s_, s_, w1, &q);
            if (!P) return false;
            if (!update(upd(s_, w1, s_, P, s_, nullptr))) return false;
        }
        int64_t c = s_;
        for (int64_t b = 1; b < N_; b++) {
            const int64_t wb = bw(b), wn = (b + 1 < N_) ? bw(b + 1) : 0;
            const int64_t na = c + wb;
            // [Q V_b]^T [V_b X_n] (:284-287, :330-342).
            SmallArgs q = sa(OP_CLOSE_OPEN, F_1S, b, c, wb, wn, nullptr, na);
            double *P = proj(0, na, c, wb + wn, &q);
            if (!P) return false;
            Upd u0 = upd(c, wb, c, P, na, Om_);
            if (wn > 0) {
                Upd u1 = upd(c + wb, wn, c + wb, P + wb * na, na, nullptr);
                if (!update2(u0, u1)) return false;
            } else if (!update(u0)) {
                return false;
            }
            c += wb;
        }
        return true;
    }

    // BCGS-PIPI+ (SRC/orth.c:369-462).
    bool pipi()
    {
        if (!intra_qr(0, 0, s_, Rd_, ldR_)) return false;  // :411-413
        int64_t c = s_;
        for (int64_t b = 1; b < N_; b++) {
            const int64_t wb = bw(b), na = c + wb;
            SmallArgs q = sa(OP_PIPI1, 0, b, c, wb, 0, nullptr, na);  // :421-424
            double *P = proj(0, na, c, wb, &q);
            if (!P) return false;
            if (!update(upd(c, wb, c, P, na, Sig_))) return false;  // :431-434
            q = sa(OP_PIPI2, 0, b, c, wb, 0, nullptr, na);  // :435-438
            P = proj(0, na, c, wb, &q);
            if (!P) return false;
            if (!update(upd(c, wb, c, P, na, Om_))) return false;  // :446-449
            c += wb;
        }
        return true;
    }

    // BCGSI+ (SRC/orth.c:464-548).
    bool i2()
    {
        if (!intra_qr(0, 0, s_, Rd_, ldR_)) return false;  // :503-505
        int64_t c = s_;
        for (int64_t b = 1; b < N_; b++) {
            const int64_t wb = bw(b);
            SmallArgs q = sa(OP_PLAIN_S, 0, b, c, wb, 0, nullptr, c);  // S = Q^T X (:512-515)
            double *P = proj(0, c, c, wb, &q);
            if (!P) return false;
            if (!update(upd(c, wb, c, P, c, nullptr))) return false;  // :518
            if (!intra_qr(b, c, wb, Sig_, kLdT)) return false;         // :520-523
            q = sa(OP_I2_W, 0, b, c, wb, 0, nullptr, c);  // W = Q^T Q^ (:524-527)
            P = proj(0, c, c, wb, &q);
            if (!P) return false;
            if (!update(upd(c, wb, c, P, c, nullptr))) return false;  // :530
            if (!intra_qr(b, c, wb, Om_, kLdT)) return false;          // :532-535
            if (!small(sa(OP_RBB, 0, b, c, wb, 0, nullptr, 1))) return false;  // :537-541
            c += wb;
        }
        return true;
    }

    // BCGSI+P-2S (SRC/orth.c:550-697).
    bool p2s()
    {
        if (!intra_qr(0, 0, s_, Rd_, ldR_)) return false;  // :599-601
        if (N_ == 1) return true;
        const int64_t w1 = bw(1);
        {
            // S = Q1^T X2, R12 = S (:605-609); V = X2 - Q1 S (:610-611); TSQR (:612-614).
            SmallArgs q = sa(OP_PLAIN_S, 0, 1, s_, w1, 0, nullptr, s_);
            double *P = proj(0, s_, s_, w1, &q);
            if (!P) return false;
            if (!update(upd(s_, w1, s_, P, s_, nullptr))) return false;
            if (!intra_qr(1, s_, w1, Sig_, kLdT)) return false;
        }
        int64_t c = s_;
        for (int64_t b = 1; b < N_; b++) {
            const int64_t wb = bw(b), wn = (b + 1 < N_) ? bw(b + 1) : 0;
            const int64_t na = c + wb;
            // [Q Q^_b]^T [Q^_b X_n] (:617-620, :658-675); the unused
            // X_n^T X_n Gram of :660-662 is not formed.
            SmallArgs q = sa(OP_CLOSE_OPEN, 0, b, c, wb, wn, nullptr, na);
            double *P = proj(0, na, c, wb + wn, &q);
            if (!P) return false;
            Upd u0 = upd(c, wb, c, P, na, Om_);
            if (wn > 0) {
                Upd u1 = upd(c + wb, wn, c + wb, P + wb * na, na, nullptr);
                if (!update2(u0, u1)) return false;
                if (!intra_qr(b + 1, c + wb, wn, Sig_, kLdT)) return false;  // :649-652
            } else if (!update(u0)) {
                return false;
            }
            c += wb;
        }
        return true;
    }

    // ---------------- fused schedules ----------------
    //
    // Each small step leaves a pending update (coefficients in Cbuf, target
    // groups); the next streaming pass applies it and accumulates the next
    // projection in the same sweep, so every block costs one pass per sync.

    struct Pending {
        int k1 = 0;
        int ngrp = 0;
        int seq = 0;
        PassGroup g[2];
    } pend_;

    void pend_group(int64_t t0, int64_t w, int zoff, int64_t k, const double *T, bool lp = false)
    {
        PassGroup &g = pend_.g[pend_.ngrp++];
        g.t0 = (int)t0;
        g.w = (int)w;
        g.zoff = zoff;
        g.k = (int)k;
        g.T = T;
        g.lp = lp ? 1 : 0;
    }

    // Mixed form: the input comes from FP32 storage and the marked
    // first-pass products run in the reduced precision mx (kMx*).
    // mx may carry kMxProjOnly: then only the first-pass projection and
    // Gram run in the reduced precision.
    void set_mixed(int mx, const float *x32, int64_t ldx32)
    {
        mx_ = mx & ~kMxProjOnly;
        proj_only_ = (mx & kMxProjOnly) != 0;
        x32_ = x32;
        ldx32_ = ldx32;
    }
    // The first-pass projection tiles run in the reduced precision.
    bool lp_first_pass() const { return mx_ != 0 && mx_ != kMxFP64; }
    // So does the first-pass update.
    bool lp_update() const { return lp_first_pass() && !proj_only_; }

    // Fraction of the device's workgroups one pass uses (the batched form
    // runs several calls side by side).
    void set_wg_scale(double f) { wg_scale_ = std::min(1.0, std::max(f, 1e-3)); }
    // Column bound of the double-buffered s = 16 pass (0: the default).
    void set_pass16d_min_cols(int c) { minc_ = std::max(0, c); }
    // Synchronize inside the passes (default) or in a launch after each.
    void set_in_pass_sync(bool on) { in_pass_sync_ = on; }

    // One fused pass: the pending update, then P <- A^T B (na == 0: none),
    // then the small step `after` on P when given.
    double *pass(int64_t a0, int64_t na, int64_t b0, int64_t nb, const SmallArgs *after = nullptr)
    {
        // The tile starts at the first column any part of the pass touches
        // (column 0 whenever a prefix product is applied).
        int64_t cb = pend_.k1 > 0 ? 0 : INT64_MAX;
        for (int g = 0; g < pend_.ngrp; g++) cb = std::min<int64_t>(cb, pend_.g[g].t0);
        if (na > 0) cb = std::min(cb, std::min(a0, b0));
        if (cb == INT64_MAX) cb = 0;
        PassArgs pa;
        memset(&pa, 0, sizeof(pa));
        pa.X = X_ + cb * ldX_;
        pa.ldX = ldX_;
        pa.m = m_;
        pa.k1 = pend_.k1;
        pa.n1 = n1_;
        pa.C = sc_.Cbuf;
        pa.ngrp = pend_.ngrp;
        pa.seq = pend_.seq;
        for (int g = 0; g < pend_.ngrp; g++) {
            pa.g[g] = pend_.g[g];
            pa.g[g].t0 -= (int)cb;
        }
        int64_t ncols = pend_.k1;
        for (int g = 0; g < pend_.ngrp; g++)
            ncols = std::max<int64_t>(ncols, pa.g[g].t0 + pa.g[g].w);
        if (na > 0) ncols = std::max(ncols, std::max(a0 + na, b0 + nb) - cb);
        int64_t wl = ncols;
        if (na > 0) {
            wl = std::max(wl, a0 - cb + ((na + 15) / 16) * 16);
            wl = std::max(wl, b0 - cb + ((nb + 15) / 16) * 16);
        }
        wl = std::max<int64_t>(wl, (pend_.k1 + 3) & ~3);
        wl = (wl + 15) / 16 * 16;
        pa.ncols = (int)ncols;
        pa.wl = (int)wl;
        pa.a0 = (int)(a0 - cb);
        pa.na = (int)na;
        pa.b0 = (int)(b0 - cb);
        pa.nb = (int)nb;
        // Kernel choice: the 512-thread s = 16 pass when it applies, else the
        // 256-thread passes (whose s = 16 form needs the tile padded to
        // ceil(Ma / 4) projection tiles per wave).
        if (mx_) {
            // Columns at or past the written frontier still live only in X32.
            pa.X32 = x32_ + cb * ldx32_;
            pa.ldX32 = ldx32_;
            pa.raw0 = (int)std::max<int64_t>(0, frontier_ - cb);
            if (pa.raw0 < pa.ncols && pa.raw0 % 16 != 0) return nullptr;
            pa.lpb = lp_from_ >= 0 ? (int)(lp_from_ - cb) : (1 << 30);
        }
        int mt = 0, nbt = 0;
        const int td = (S_ == 16 && !getenv("ORTH_PASS256") && !getenv("ORTH_MIXED_PASS32"))
                           ? pass16d_plan(pa, &mt, &nbt, lp_first_pass(), minc_)
                           : 0;
        const int ts16 = td > 0 ? 32
                         : (S_ == 16 && !mx_ && !getenv("ORTH_PASS256")) ? pass16b_plan(pa, &mt, &nbt)
                                                                          : 0;
        int64_t unit = kPassRows;
        int tsm = 0;
        if (td > 0) {
            unit = 32;
        } else if (mx_) {
            // The 512-thread pass with super-tiles when the slab allows it,
            // else the 32-row mixed pass.
            tsm = getenv("ORTH_MIXED_PASS32") ? 0 : pass16bm_plan(pa);
            if (tsm > 0)
                unit = tsm;
            else if (pass16m_lds_bytes(pa) == 0)
                return nullptr;
        } else if (ts16 > 0) {
            unit = ts16;
        } else if (S_ == 16 && na > 0) {
            const int64_t mtp = ((na + 15) / 16 + 3) / 4;
            pa.wl = (int)std::max<int64_t>(pa.wl, a0 - cb + 64 * mtp);
        }
        if (!mx_ && td == 0 && ts16 == 0) unit = pass_rows(S_, pa);
        // Two 512-thread workgroups per CU when the s = 16 pass applies (its
        // LDS is under half a CU's), so one hides the other's barriers.
        // The mixed pass with reduced-precision products runs one 512-thread
        // workgroup per CU (its operand conversions need two waves' worth of
        // registers); with FP64 products it runs two, like the FP64 pass.
        // The 256-thread passes keep as many workgroups per CU as their
        // registers and LDS allow.
        // The double-buffered pass with 16 waves holds a whole CU (one
        // workgroup).
        // The pass synchronizes itself (reduction and small step in its last
        // workgroups) when the scratch has the ticket counters and the slab
        // is long: the tail puts device-scope cache maintenance (each
        // workgroup's release, the last workgroups' acquires at two levels)
        // on the critical path, 4-15 us per synchronization more than a
        // launch boundary, which only passes of kTailMinRowsPerCU rows per CU
        // and more hide.  Its group sums follow the partials in `part`.
        bool tail = na > 0 && sc_.tickets && in_pass_sync_ && !getenv("ORTH_SEPARATE_REDUCE") &&
                    !getenv("ORTH_REDUCE_LAUNCH") &&
                    (m_ >= kTailMinRowsPerCU * (int64_t)std::max(1, c_.ncu) ||
                     getenv("ORTH_IN_PASS_SYNC"));
        const int per_cu = td == 16                          ? 1
                           : (ts16 > 0 || mx_ == kMxFP64) ? 2
                           : mx_                           ? 1
                                                           : pass_wg_per_cu(S_, pa, tail);
        int nwg = std::max(1, (int)(std::max(1, c_.ncu) * per_cu * wg_scale_));
        if (na > 0 && sc_.fixed) {
            // Fixed scratch: fewer row chunks, so that the partials fit.
            int64_t cap = (int64_t)(sc_.part_n / ((size_t)na * nb));
            if (tail && cap > 2 * kTailMaxGroups)
                cap -= kTailMaxGroups;
            else
                tail = false;
            if (cap < 1) return nullptr;
            nwg = (int)std::min<int64_t>(nwg, cap);
        }
        int64_t rows = (m_ + nwg - 1) / nwg;
        rows = (rows + unit - 1) / unit * unit;
        pa.rows_per = rows;
        nwg = (int)((m_ + rows - 1) / rows);
        const int ngrp = tail ? (nwg + tail_group(nwg) - 1) / tail_group(nwg) : 0;
        if (na > 0 && !grow(sc_, sc_.part, sc_.part_n, (size_t)(nwg + ngrp) * na * nb))
            return nullptr;
        pa.part = sc_.part;
        double *Pn = na > 0 ? sc_.P[pcur_ ^ 1] : nullptr;
        if (tail) {
            pa.tail_on = 1;
            pa.tail.nwg = nwg;
            pa.tail.grp = tail_group(nwg);
            pa.tail.tickets = sc_.tickets;
            pa.tail.lvl = sc_.part + (size_t)nwg * na * nb;
            pa.tail.out = Pn;
            if (after) {
                pa.tail.small = 1;
                pa.tail.sm = *after;
                pa.tail.sm.P = Pn;
                finish_small(pa.tail.sm);
            }
        }
        double flops = 2.0 * m_ * na * nb, wcols = 0.0;
        for (int g = 0; g < pend_.ngrp; g++) {
            const double kk = pend_.seq ? pend_.g[g].k : pend_.k1;
            flops += (double)m_ * pend_.g[g].w * (2.0 * kk + (pend_.g[g].T ? pend_.g[g].w : 0));
            wcols += pend_.g[g].w;
        }
        if (!timed(c_, ORTH_K_PASS, st_, flops, 8.0 * m_ * (ncols + wcols), [&] {
                if (td > 0) return mx_ ? launch_pass16d_mixed(pa, mx_, td, nbt, nwg, st_)
                                       : launch_pass16d(pa, td, nbt, nwg, st_);
                if (mx_)
                    return tsm > 0 ? launch_pass16bm(pa, mx_, nwg, st_)
                                   : launch_pass16m(pa, mx_, nwg, st_);
                return ts16 > 0 ? launch_pass16b(pa, nwg, st_) : launch_pass(S_, pa, nwg, st_);
            }))
            return nullptr;
        for (int g = 0; g < pend_.ngrp; g++)
            frontier_ = std::max<int64_t>(frontier_, pend_.g[g].t0 + pend_.g[g].w);
        lp_from_ = -1;
        pend_ = Pending();
        passes_++;
        if (na == 0) return sc_.P[pcur_];
        pcur_ ^= 1;
        syncs_++;
        if (tail) {
            if (debug_ && after) dump(pa.tail.sm);
            return Pn;
        }
        return sync(nwg, na, nb, Pn, after) ? Pn : nullptr;
    }

    // The synchronization after a projection: P <- the fixed-order sum of the
    // nwg partials, then the small step `after` (if any) on P -- in one
    // launch when the scratch has the ticket counter, else as two.
    bool sync(int nwg, int64_t na, int64_t nb, double *P, const SmallArgs *after)
    {
        const double rflops = (double)nwg * na * nb, rbytes = 8.0 * ((double)nwg + 1) * na * nb;
        if (after && sc_.tickets && !getenv("ORTH_SEPARATE_REDUCE")) {
            SmallArgs a = *after;
            a.P = P;
            finish_small(a);
            if (!timed(c_, ORTH_K_REDUCE, st_, rflops, rbytes, [&] {
                    return launch_reduce_small(sc_.part, nwg, na, nb, P, sc_.tickets, a, st_);
                }))
                return false;
            if (debug_) dump(a);
            return true;
        }
        if (!timed(c_, ORTH_K_REDUCE, st_, rflops, rbytes,
                   [&] { return launch_reduce(sc_.part, nwg, na, nb, P, na, st_); }))
            return false;
        if (!after) return true;
        SmallArgs a = *after;
        a.P = P;
        return small(a);
    }

    // A pass followed by the small step s on its reduced projection.
    bool pass_small(int64_t a0, int64_t na, int64_t b0, int64_t nb, const SmallArgs &s)
    {
        return pass(a0, na, b0, nb, &s) != nullptr;
    }

    bool flush()
    {
        if (pend_.ngrp == 0) return true;
        return pass(0, 0, 0, 0) != nullptr;
    }

    // Can every pass of this call run fused?  Widest tile: all n columns;
    // largest coefficient block: (n + s) x n1.
    bool fused_fits() const
    {
        // The generic fused pass addresses up to four columns with 32-bit
        // lane offsets.
        if (ldX_ >= ((int64_t)1 << 27)) return false;
        PassArgs pa;
        memset(&pa, 0, sizeof(pa));
        pa.ncols = (int)n_;
        pa.wl = (int)((n_ + 15) / 16 * 16);
        pa.k1 = (int)std::min<int64_t>(n_, pa.wl);
        pa.n1 = n1_;
        const int64_t Mt = (n_ + 15) / 16;
        pa.na = (int)std::min<int64_t>(n_, Mt * 16);
        pa.nb = (int)std::min<int64_t>(n_, 2 * s_);
        pa.a0 = 0;
        pa.b0 = (int)(pa.wl - ((pa.nb + 15) / 16) * 16);
        return pass_lds_bytes(pa) != 0;
    }

    bool intra_qr_f(int64_t blk, int64_t col0, int64_t w, double *Rdst, int64_t ldRdst)
    {
        const int ops[3] = {OP_CQR1, OP_CQR2, OP_CQR3};
        for (int k = 0; k < 3; k++) {
            SmallArgs a = sa(ops[k], 0, blk, 0, w, 0, nullptr, w);
            a.Rdst = Rdst;
            a.ldRdst = ldRdst;
            if (!pass(col0, w, col0, w, &a)) return false;
            pend_group(col0, w, 0, 0, Cq_ + k * kSlot);
        }
        return true;
    }

    void pend_close_open(int64_t c, int64_t wb, int64_t wn, bool next_T, bool seq)
    {
        pend_.k1 = (int)(c + wb);
        pend_.seq = seq ? 1 : 0;
        pend_group(c, wb, 0, c, Om_);
        if (wn > 0) pend_group(c + wb, wn, S_, c + wb, next_T ? Sig_ : nullptr, lp_update());
    }

    bool p1s_f()
    {
        if (!intra_qr_f(0, 0, s_, Rd_, ldR_)) return false;
        if (N_ == 1) return flush();
        const int64_t w1 = bw(1);
        // Mixed form: the projection of a block's raw columns (first pass)
        // and its first-pass update run in the reduced precision.
        if (lp_first_pass()) lp_from_ = s_;
        if (!pass_small(0, s_ + w1, s_, w1, sa(OP_OPEN_PYTH, 0, 1, s_, w1, 0, nullptr, s_ + w1))) return false;
        pend_.k1 = (int)s_;
        pend_group(s_, w1, 0, s_, Sig_, lp_update());
        int64_t c = s_;
        for (int64_t b = 1; b < N_; b++) {
            const int64_t wb = bw(b), wn = (b + 1 < N_) ? bw(b + 1) : 0;
            const int64_t na = c + wb + wn;
            if (lp_first_pass() && wn > 0) lp_from_ = c + wb;
            if (!pass_small(0, na, c, wb + wn, sa(OP_CLOSE_OPEN, F_PYTH_NEXT, b, c, wb, wn, nullptr, na))) return false;
            pend_close_open(c, wb, wn, true, false);
            c += wb;
        }
        return flush();
    }

    bool a1s_f()
    {
        if (!intra_qr_f(0, 0, s_, Rd_, ldR_)) return false;
        if (N_ == 1) return flush();
        const int64_t w1 = bw(1);
        if (!pass_small(0, s_, s_, w1, sa(OP_PLAIN_S, 0, 1, s_, w1, 0, nullptr, s_))) return false;
        pend_.k1 = (int)s_;
        pend_group(s_, w1, 0, s_, nullptr);
        int64_t c = s_;
        for (int64_t b = 1; b < N_; b++) {
            const int64_t wb = bw(b), wn = (b + 1 < N_) ? bw(b + 1) : 0;
            const int64_t na = c + wb;
            if (!pass_small(0, na, c, wb + wn, sa(OP_CLOSE_OPEN, F_1S | F_SEQ, b, c, wb, wn, nullptr, na))) return false;
            pend_close_open(c, wb, wn, false, true);
            c += wb;
        }
        return flush();
    }

    bool pipi_f()
    {
        if (!intra_qr_f(0, 0, s_, Rd_, ldR_)) return false;
        int64_t c = s_;
        for (int64_t b = 1; b < N_; b++) {
            const int64_t wb = bw(b), na = c + wb;
            if (!pass_small(0, na, c, wb, sa(OP_PIPI1, 0, b, c, wb, 0, nullptr, na))) return false;
            pend_.k1 = (int)c;
            pend_group(c, wb, 0, c, Sig_);
            if (!pass_small(0, na, c, wb, sa(OP_PIPI2, 0, b, c, wb, 0, nullptr, na))) return false;
            pend_.k1 = (int)c;
            pend_group(c, wb, 0, c, Om_);
            c += wb;
        }
        return flush();
    }

    bool i2_f()
    {
        if (!intra_qr_f(0, 0, s_, Rd_, ldR_)) return false;
        int64_t c = s_;
        for (int64_t b = 1; b < N_; b++) {
            const int64_t wb = bw(b);
            if (!pass_small(0, c, c, wb, sa(OP_PLAIN_S, 0, b, c, wb, 0, nullptr, c))) return false;
            pend_.k1 = (int)c;
            pend_group(c, wb, 0, c, nullptr);
            if (!intra_qr_f(b, c, wb, Sig_, kLdT)) return false;
            if (!pass_small(0, c, c, wb, sa(OP_I2_W, 0, b, c, wb, 0, nullptr, c))) return false;
            pend_.k1 = (int)c;
            pend_group(c, wb, 0, c, nullptr);
            if (!intra_qr_f(b, c, wb, Om_, kLdT)) return false;
            if (!small(sa(OP_RBB, 0, b, c, wb, 0, nullptr, 1))) return false;
            c += wb;
        }
        return flush();
    }

    bool p2s_f()
    {
        if (!intra_qr_f(0, 0, s_, Rd_, ldR_)) return false;
        if (N_ == 1) return flush();
        const int64_t w1 = bw(1);
        if (!pass_small(0, s_, s_, w1, sa(OP_PLAIN_S, 0, 1, s_, w1, 0, nullptr, s_))) return false;
        pend_.k1 = (int)s_;
        pend_group(s_, w1, 0, s_, nullptr);
        if (!intra_qr_f(1, s_, w1, Sig_, kLdT)) return false;
        int64_t c = s_;
        for (int64_t b = 1; b < N_; b++) {
            const int64_t wb = bw(b), wn = (b + 1 < N_) ? bw(b + 1) : 0;
            const int64_t na = c + wb;
            if (!pass_small(0, na, c, wb + wn, sa(OP_CLOSE_OPEN, 0, b, c, wb, wn, nullptr, na))) return false;
            pend_close_open(c, wb, wn, false, false);
            if (wn > 0 && !intra_qr_f(b + 1, c + wb, wn, Sig_, kLdT)) return false;
            c += wb;
        }
        return flush();
    }

    bool run(int variant)
    {
        fused_ = !getenv("ORTH_UNFUSED") && fused_fits();
        if (mx_) {
            if (!fused_ || S_ != 16 || variant != ORTH_P1S) return false;
            return p1s_f();
        }
        if (fused_) {
            switch (variant) {
            case ORTH_P1S: return p1s_f();
            case ORTH_1S: return a1s_f();
            case ORTH_PIPI: return pipi_f();
            case ORTH_I2: return i2_f();
            case ORTH_P2S: return p2s_f();
            default: return false;
            }
        }
        switch (variant) {
        case ORTH_P1S: return p1s();
        case ORTH_1S: return a1s();
        case ORTH_PIPI: return pipi();
        case ORTH_I2: return i2();
        case ORTH_P2S: return p2s();
        default: return false;
        }
    }

    // Copy R to the caller: for block column k only rows 0 .. end of block k.
    bool copy_R_out(double *Rdst, int64_t ldRdst, hipMemcpyKind kind)
    {
        int64_t c = 0;
        for (int64_t b = 0; b < N_; b++) {
            const int64_t w = bw(b);
            const int64_t rows = c + w;
            if (!ok(hipMemcpy2DAsync(Rdst + c * ldRdst, ldRdst * sizeof(double),
                                     Rd_ + c * ldR_, ldR_ * sizeof(double),
                                     rows * sizeof(double), w, kind, st_)))
                return false;
            c += w;
        }
        return true;
    }

    int passes_ = 0, syncs_ = 0;
    bool debug_ = getenv("ORTH_DEBUG_DUMP") != nullptr;
    bool fused_ = false;
    int n1_ = 32;
    int mx_ = 0;                // mixed form: reduced precision (kMx*), 0 = plain FP64
    bool proj_only_ = false;    // mixed form: the first-pass update stays FP64
    const float *x32_ = nullptr;
    int64_t ldx32_ = 0;
    int64_t frontier_ = 0;      // columns [0, frontier_) have been written to X
    int64_t lp_from_ = -1;      // next pass: B columns from here on are reduced precision
    double wg_scale_ = 1.0;
    int minc_ = 0;
    bool in_pass_sync_ = true;

    // Debug aid (ORTH_DEBUG_DUMP set): print the reduced projection and the
    // factors after each small step.
    void dump(const SmallArgs &a)
    {
        (void)hipStreamSynchronize(st_);
        std::vector<double> h(6 * kSlot);
        (void)hipMemcpy(h.data(), sc_.stash, h.size() * sizeof(double), hipMemcpyDeviceToHost);
        int stat[4];
        (void)hipMemcpy(stat, sc_.status_d, sizeof(stat), hipMemcpyDeviceToHost);
        fprintf(stderr, "[orth] op=%d blk=%d c=%d wb=%d wn=%d na=%d status=%d/%d\n", a.op,
                a.blk, a.c, a.wb, a.wn, a.na, stat[0], stat[1]);
        if (a.P) {
            std::vector<double> p((size_t)a.na * (a.wb + a.wn));
            (void)hipMemcpy(p.data(), a.P, p.size() * sizeof(double), hipMemcpyDeviceToHost);
            fprintf(stderr, "  P[0..3][0..3]:");
            for (int i = 0; i < 4 && i < a.na; i++)
                for (int j = 0; j < 4 && j < a.wb; j++) fprintf(stderr, " %.6e", p[i + j * a.na]);
            fprintf(stderr, "\n");
        }
        const char *names[6] = {"Sig", "Om", "Cq0", "Cq1", "Cq2", "Ctmp"};
        for (int k = 0; k < 6; k++) {
            fprintf(stderr, "  %s diag:", names[k]);
            for (int i = 0; i < 4; i++) fprintf(stderr, " %.6e", h[k * kSlot + i * (kLdT + 1)]);
            fprintf(stderr, "\n");
        }
    }

  private:
    DevCtx &c_;
    Scratch &sc_;
    int64_t m_, n_, s_;
    double *X_;
    int64_t ldX_;
    hipStream_t st_;
    int S_ = 16;
    int64_t N_ = 1;
    double *Rd_ = nullptr;
    int64_t ldR_ = 0;
    double *Sig_, *Om_, *Cq_, *Ctmp_;
    int pcur_ = 0;
};

bool is_device_ptr(const void *p)
{
    if (!p) return false;
    hipPointerAttribute_t at;
    memset(&at, 0, sizeof(at));
    hipError_t e = hipPointerGetAttributes(&at, p);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    return at.type == hipMemoryTypeDevice || at.type == hipMemoryTypeManaged;
}

bool shape_ok(int64_t m, int64_t n, int64_t s, int64_t ldX, int64_t ldR)
{
    return s >= 1 && n >= s && m >= n && ldX >= m && ldR >= n && block_template(s) != 0 &&
           n <= 4096;
}

// Enqueue one variant on device data.  Rdev_out receives R (device or host
// memory, per kind).  Returns 0 or 2.
int run_device(DevCtx &c, int variant, int64_t m, int64_t n, int64_t s, double *dX,
               int64_t ldX, double *Rout, int64_t ldR, hipMemcpyKind rkind, hipStream_t st,
               int mx = 0, const float *x32 = nullptr, int64_t ldx32 = 0)
{
    if (c.done_valid && !ok(hipStreamWaitEvent(st, c.done, 0))) return 2;
    Run r(c, c.sc, m, n, s, dX, ldX, st);
    if (mx) r.set_mixed(mx, x32, ldx32);
    if (!r.prepare()) return 2;
    if (!r.run(variant)) return 2;
    if (!r.copy_R_out(Rout, ldR, rkind)) return 2;
    if (!ok(hipMemcpyAsync(c.status_h, c.sc.status_d, 4 * sizeof(int), hipMemcpyDeviceToHost,
                            st)))
        return 2;
    if (!ok(hipEventRecord(c.done, st))) return 2;
    c.done_valid = true;
    c.last_passes = r.passes_;
    c.last_syncs = r.syncs_;
    return 0;
}

void fill_status(DevCtx &c)
{
    if (c.done_valid) (void)hipEventSynchronize(c.done);
    c.last.info_block = c.status_h ? c.status_h[0] : 0;
    c.last.info_pivot = c.status_h ? c.status_h[1] : 0;
    c.last.hip_error = tl_hip_error;
    c.last.passes = c.last_passes;
    c.last.syncs = c.last_syncs;
}

// Drop-in entry: host or device X / R, synchronous.
int dropin(int variant, int m, int n, int s, double *X, int ldX, double *R, int ldR,
           double *work, int lwork)
{
    const int64_t need = 5 * (int64_t)n * s + 9 * (int64_t)s * s +
                         (int64_t)m * (variant == ORTH_I2 ? s : n);
    if (work != nullptr && (int64_t)lwork < need) {
        printf("%% Work space is too small. It should be larger than 6*n*s + 8*s*s.\n");
        return 1;
    }
    if (!X || !R || !shape_ok(m, n, s, ldX, ldR)) return 2;
    tl_hip_error = 0;
    const int dev = current_device();
    if (dev < 0 || dev >= kMaxDevices) return 2;
    DevCtx &c = g_ctx[dev];
    std::lock_guard<std::mutex> lk(c.mu);
    if (!ctx_init(c)) return 2;
    hipStream_t st = nullptr;
    const bool xdev = is_device_ptr(X), rdev = is_device_ptr(R);
    double *dX = X;
    int64_t ldd = ldX;
    if (!xdev) {
        if (!ensure(c.Xd, c.Xd_n, (size_t)m * n)) return 2;
        dX = c.Xd;
        ldd = m;
        if (!ok(hipMemcpy2DAsync(dX, (size_t)ldd * sizeof(double), X, (size_t)ldX * sizeof(double),
                                 (size_t)m * sizeof(double), n, hipMemcpyHostToDevice, st)))
            return 2;
    }
    int rc = run_device(c, variant, m, n, s, dX, ldd, R, ldR,
                        rdev ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, st);
    if (rc != 0) return rc;
    if (!xdev) {
        if (!ok(hipMemcpy2DAsync(X, (size_t)ldX * sizeof(double), dX, (size_t)ldd * sizeof(double),
                                 (size_t)m * sizeof(double), n, hipMemcpyDeviceToHost, st)))
            return 2;
    }
    if (!ok(hipStreamSynchronize(st))) return 2;
    fill_status(c);
    return 0;
}

// ---------------------------------------------------------------------------
// Batched form: many independent calls side by side.
//
// Instances are assigned to up to kMaxSlots streams (longest first, by
// algorithmic bytes); each stream owns a private scratch set (a DevCtx) and
// runs its instances one after the other.  The whole fork / per-stream
// sequences / join is captured once into a hipGraph and replayed while the
// same batch (shapes and pointers) comes back.  The stream count and the
// share of the device's workgroups each pass uses are tunable; the tuner
// times the candidates on the batch itself and keeps the best per shape
// signature.
// ---------------------------------------------------------------------------
constexpr int kMaxSlots = 8;

struct BatchCfg {
    int streams = 4;
    double wg_scale = 0.5;
    bool graph = true;  // capture into hipGraphs (else launch directly on the streams)
    int minc = 0;       // column bound of the double-buffered s = 16 pass (0: default)
};

// Graph nodes per captured graph at most: batches beyond it are captured as
// several graphs launched back to back.
constexpr int64_t kMaxGraphNodes = 12000;

struct BatchCtx {
    bool init = false;
    bool broken = false;  // a stream stayed in capture mode: refuse further batches
    DevCtx slot[kMaxSlots];
    hipStream_t st[kMaxSlots] = {};
    hipStream_t main = nullptr;
    hipEvent_t fork = nullptr;
    hipEvent_t join[kMaxSlots] = {};
    int *info_d = nullptr;
    int *info_h = nullptr;
    size_t info_n = 0;
    std::vector<hipGraphExec_t> execs;  // the captured batch, in launch order
    std::vector<int64_t> key;           // batch the graphs were captured for
    std::vector<int64_t> nocap;         // batch whose capture failed (runs directly)
    std::vector<std::pair<std::vector<int64_t>, BatchCfg>> tuned;  // shape signature -> config
    bool cache_read = false;  // the tuning cache file was read for this device
};

BatchCtx g_batch[kMaxDevices];

struct BatchArgs {
    int variant = 0;
    int mx = 0;
    int count = 0;
    const int64_t *m = nullptr, *n = nullptr, *s = nullptr;
    double *const *X = nullptr;
    const int64_t *ldX = nullptr;
    double *const *R = nullptr;
    const int64_t *ldR = nullptr;
    const float *const *X32 = nullptr;
    const int64_t *ldX32 = nullptr;
};

bool batch_init(BatchCtx &b)
{
    if (b.init) return true;
    for (int k = 0; k < kMaxSlots; k++) {
        if (!ctx_init(b.slot[k])) return false;
        if (!ok(hipStreamCreateWithFlags(&b.st[k], hipStreamNonBlocking))) return false;
        if (!ok(hipEventCreateWithFlags(&b.join[k], hipEventDisableTiming))) return false;
    }
    if (!ok(hipStreamCreateWithFlags(&b.main, hipStreamNonBlocking))) return false;
    if (!ok(hipEventCreateWithFlags(&b.fork, hipEventDisableTiming))) return false;
    b.init = true;
    return true;
}

double inst_bytes(int64_t m, int64_t n, int64_t s)
{
    const int64_t N = 1 + (n - s + s - 1) / s;
    return 8.0 * m * (2.0 * n + (double)s * N * (N - 1) / 2);
}

std::vector<int64_t> shape_sig(const BatchArgs &a)
{
    std::vector<int64_t> k = {a.variant, a.mx, a.count};
    for (int i = 0; i < a.count; i++) {
        k.push_back(a.m[i]);
        k.push_back(a.n[i]);
        k.push_back(a.s[i]);
    }
    return k;
}

// Graph nodes of one instance: per synchronization a pass, a reduction and
// a small step; the fills, the R copies and the status copy.
int64_t inst_nodes(int variant, int64_t n, int64_t s)
{
    const int64_t N = 1 + (n - s + s - 1) / s;
    int64_t syncs = 3 + N;
    if (variant == ORTH_PIPI) syncs = 3 + 2 * (N - 1);
    if (variant == ORTH_I2) syncs = 3 + 8 * (N - 1);
    if (variant == ORTH_P2S) syncs = 7 + (N - 1) + 3 * std::max<int64_t>(0, N - 2);
    return 3 * syncs + N + 8;
}

std::vector<int64_t> full_key(const BatchArgs &a, const BatchCfg &cfg)
{
    std::vector<int64_t> k = shape_sig(a);
    k.push_back(cfg.graph ? 1 : 0);
    k.push_back(cfg.streams);
    k.push_back((int64_t)(cfg.wg_scale * 1e6));
    k.push_back(cfg.minc);
    // Captured graphs hold the small steps' arguments, the shift among them,
    // and the kernels the environment switches selected.
    const double ps = g_pshift.load();
    int64_t psb = 0;
    memcpy(&psb, &ps, sizeof(psb));
    k.push_back(psb);
    k.push_back(env_bits());
    for (int i = 0; i < a.count; i++) {
        k.push_back((int64_t)(intptr_t)a.X[i]);
        k.push_back(a.ldX[i]);
        k.push_back((int64_t)(intptr_t)a.R[i]);
        k.push_back(a.ldR[i]);
        if (a.mx) {
            k.push_back((int64_t)(intptr_t)a.X32[i]);
            k.push_back(a.ldX32[i]);
        }
    }
    return k;
}

// Longest-first assignment of instances (indices idx) to streams.
std::vector<std::vector<int>> assign(const BatchArgs &a, const std::vector<int> &idx, int streams)
{
    std::vector<int> order = idx;
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
        return inst_bytes(a.m[x], a.n[x], a.s[x]) > inst_bytes(a.m[y], a.n[y], a.s[y]);
    });
    std::vector<std::vector<int>> lists(streams);
    std::vector<double> load(streams, 0.0);
    for (int i : order) {
        const int k = (int)(std::min_element(load.begin(), load.end()) - load.begin());
        lists[k].push_back(i);
        load[k] += inst_bytes(a.m[i], a.n[i], a.s[i]);
    }
    return lists;
}

bool ensure_info(BatchCtx &b, int count)
{
    if (b.info_n >= (size_t)count * 2) return true;
    if (b.info_d) (void)hipFree(b.info_d);
    if (b.info_h) (void)hipHostFree(b.info_h);
    b.info_d = nullptr;
    b.info_h = nullptr;
    b.info_n = 0;
    if (!ok(hipMalloc(&b.info_d, (size_t)count * 2 * sizeof(int)))) return false;
    if (!ok(hipHostMalloc(&b.info_h, (size_t)count * 2 * sizeof(int), hipHostMallocDefault)))
        return false;
    b.info_n = (size_t)count * 2;
    return true;
}

void drop_graphs(BatchCtx &b)
{
    for (hipGraphExec_t e : b.execs) (void)hipGraphExecDestroy(e);
    b.execs.clear();
    b.key.clear();
}

// Size every slot's scratch for the widest instance of the batch.  Runs only
// while no graph of the batch exists (a captured graph keeps the buffer
// addresses); the batch's own runs then never reallocate (the scratch is
// frozen while they are enqueued).
bool batch_reserve(BatchCtx &b, const BatchArgs &a, int streams)
{
    int64_t nmax = 0;
    for (int i = 0; i < a.count; i++) nmax = std::max(nmax, a.n[i]);
    for (int k = 0; k < streams; k++) {
        Scratch &sc = b.slot[k].sc;
        const size_t pmax = scratch_p_need(nmax);
        if (sc.P_n < pmax || !sc.P[0] || !sc.P[1]) {
            for (int j = 0; j < 2; j++) {
                if (sc.P[j]) (void)hipFree(sc.P[j]);
                sc.P[j] = nullptr;
            }
            sc.P_n = 0;
            if (!ok(hipMalloc(&sc.P[0], pmax * sizeof(double)))) return false;
            if (!ok(hipMalloc(&sc.P[1], pmax * sizeof(double)))) return false;
            sc.P_n = pmax;
        }
        const size_t partmax = (size_t)kProjTargetWG * (nmax + 3 * kMaxS) * (3 * kMaxS);
        if (!grow(sc, sc.part, sc.part_n, partmax) || !grow(sc, sc.Rd, sc.Rd_n, scratch_r_need(nmax)) ||
            !grow(sc, sc.Tbig, sc.Tbig_n, scratch_tbig_need(nmax)) ||
            !grow(sc, sc.Cbuf, sc.Cbuf_n, scratch_cbuf_need(nmax)))
            return false;
    }
    return ensure_info(b, a.count);
}

void freeze(BatchCtx &b, bool f)
{
    for (int k = 0; k < kMaxSlots; k++) b.slot[k].sc.fixed = f;
}

// Enqueue instances on b.main: fork, per-stream sequences, join; `last` adds
// the copy of every instance's status to the host.  Every forked stream is
// joined back into b.main, also when enqueueing fails half way, so that a
// capture can always be ended.
bool batch_enqueue(BatchCtx &b, const BatchArgs &a, const BatchCfg &cfg,
                   const std::vector<std::vector<int>> &lists, bool last)
{
    if (!ok(hipEventRecord(b.fork, b.main))) return false;
    const int ns = (int)lists.size();
    int forked = 0;
    bool good = true;
    for (; forked < ns; forked++)
        if (!ok(hipStreamWaitEvent(b.st[forked], b.fork, 0))) {
            good = false;
            break;
        }
    for (int k = 0; k < forked && good; k++) {
        hipStream_t st = b.st[k];
        for (int i : lists[k]) {
            Run r(b.slot[k], b.slot[k].sc, a.m[i], a.n[i], a.s[i], a.X[i], a.ldX[i], st);
            r.set_wg_scale(cfg.wg_scale);
            r.set_pass16d_min_cols(cfg.minc);
            // Concurrent instances synchronize in a launch after each pass:
            // inside the passes, every workgroup's device-scope release and
            // the last workgroups' acquires reach into the L2 the other
            // streams' instances are using (the 512-instance batch ran 8 %
            // (P-1S) and 21 % (BCGSI+) slower that way).
            r.set_in_pass_sync(false);
            if (a.mx) r.set_mixed(a.mx, a.X32[i], a.ldX32[i]);
            good = r.prepare() && r.run(a.variant) &&
                   r.copy_R_out(a.R[i], a.ldR[i], hipMemcpyDeviceToDevice) &&
                   ok(hipMemcpyAsync(b.info_d + 2 * i, b.slot[k].sc.status_d, 2 * sizeof(int),
                                     hipMemcpyDeviceToDevice, st));
            if (!good) break;
        }
    }
    for (int k = 0; k < forked; k++) {
        if (!ok(hipEventRecord(b.join[k], b.st[k]))) good = false;
        if (!ok(hipStreamWaitEvent(b.main, b.join[k], 0))) good = false;
    }
    if (good && last)
        good = ok(hipMemcpyAsync(b.info_h, b.info_d, (size_t)a.count * 2 * sizeof(int),
                                 hipMemcpyDeviceToHost, b.main));
    return good;
}

// Instance chunks of at most kMaxGraphNodes graph nodes each.
std::vector<std::vector<int>> graph_chunks(const BatchArgs &a)
{
    std::vector<std::vector<int>> chunks(1);
    int64_t nodes = 0;
    for (int i = 0; i < a.count; i++) {
        const int64_t c = inst_nodes(a.variant, a.n[i], a.s[i]);
        if (nodes + c > kMaxGraphNodes && !chunks.back().empty()) {
            chunks.push_back({});
            nodes = 0;
        }
        chunks.back().push_back(i);
        nodes += c;
    }
    return chunks;
}

bool capturing(hipStream_t st)
{
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(st, &cs) != hipSuccess) {
        (void)hipGetLastError();
        return true;  // unknown state: treat as unusable
    }
    return cs != hipStreamCaptureStatusNone;
}

// Capture the batch into graphs for later replays; nothing runs.  On failure
// no graph is kept and every stream has left capture mode (else the batch
// context is marked broken).
bool batch_capture(BatchCtx &b, const BatchArgs &a, const BatchCfg &cfg)
{
    const auto chunks = graph_chunks(a);
    bool good = true;
    freeze(b, true);
    for (size_t ci = 0; ci < chunks.size() && good; ci++) {
        const auto lists = assign(a, chunks[ci], cfg.streams);
        if (!ok(hipStreamBeginCapture(b.main, hipStreamCaptureModeThreadLocal))) {
            good = false;
            break;
        }
        const bool enq = batch_enqueue(b, a, cfg, lists, ci + 1 == chunks.size());
        hipGraph_t g = nullptr;
        const bool ended = ok(hipStreamEndCapture(b.main, &g));
        good = enq && ended && g;
        hipGraphExec_t e = nullptr;
        if (good && ok(hipGraphInstantiate(&e, g, nullptr, nullptr, 0)))
            b.execs.push_back(e);
        else
            good = false;
        if (g) (void)hipGraphDestroy(g);
    }
    freeze(b, false);
    if (good) return true;
    drop_graphs(b);
    bool stuck = capturing(b.main);
    for (int k = 0; k < kMaxSlots; k++) stuck = capturing(b.st[k]) || stuck;
    if (stuck) b.broken = true;
    (void)hipGetLastError();
    tl_hip_error = 0;
    return false;
}

// Run the batch with cfg on the current contents of its inputs.  Graph mode
// replays the graphs captured for this exact batch; a batch met for the
// first time is launched directly on the streams (which also makes every
// kernel's first launch, and its launch attributes, happen outside capture)
// and then captured, without running, for the next call.
bool batch_run(BatchCtx &b, const BatchArgs &a, const BatchCfg &cfg)
{
    const std::vector<int64_t> key = full_key(a, cfg);
    if (cfg.graph && !b.execs.empty() && key == b.key) {
        for (hipGraphExec_t e : b.execs)
            if (!ok(hipGraphLaunch(e, b.main))) return false;
        return ok(hipStreamSynchronize(b.main));
    }
    drop_graphs(b);
    if (!batch_reserve(b, a, cfg.streams)) return false;
    std::vector<int> all(a.count);
    for (int i = 0; i < a.count; i++) all[i] = i;
    freeze(b, true);
    const bool enq = batch_enqueue(b, a, cfg, assign(a, all, cfg.streams), true);
    freeze(b, false);
    const bool done = ok(hipStreamSynchronize(b.main));
    if (!enq || !done) return false;
    if (cfg.graph && key != b.nocap) {
        if (batch_capture(b, a, cfg))
            b.key = key;
        else if (b.broken)
            return false;
        else
            b.nocap = key;  // launched directly from now on
    }
    return true;
}

// Device copies of the batch's FP64 inputs, so that every timing run of the
// tuner and the final run start from the caller's X (the FP64 form works in
// place; the mixed form never writes its FP32 input).
struct InputBackup {
    double *buf = nullptr;
    std::vector<size_t> off;
    ~InputBackup()
    {
        if (buf) (void)hipFree(buf);
    }
    bool save(const BatchArgs &a, hipStream_t st)
    {
        size_t tot = 0;
        for (int i = 0; i < a.count; i++) {
            off.push_back(tot);
            tot += (size_t)a.m[i] * a.n[i];
        }
        if (hipMalloc(&buf, tot * sizeof(double)) != hipSuccess) {
            (void)hipGetLastError();
            buf = nullptr;
            return false;
        }
        for (int i = 0; i < a.count; i++)
            if (!ok(hipMemcpy2DAsync(buf + off[i], a.m[i] * sizeof(double), a.X[i],
                                     a.ldX[i] * sizeof(double), a.m[i] * sizeof(double), a.n[i],
                                     hipMemcpyDeviceToDevice, st)))
                return false;
        return ok(hipStreamSynchronize(st));
    }
    bool restore(const BatchArgs &a, hipStream_t st) const
    {
        if (!buf) return true;
        for (int i = 0; i < a.count; i++)
            if (!ok(hipMemcpy2DAsync(a.X[i], a.ldX[i] * sizeof(double), buf + off[i],
                                     a.m[i] * sizeof(double), a.m[i] * sizeof(double), a.n[i],
                                     hipMemcpyDeviceToDevice, st)))
                return false;
        return true;
    }
};

// Persisted tuning: when ORTH_TUNE_CACHE names a file, the tuner's results
// are kept there across processes, one line per device and batch signature:
//   orth-tune-1 <device> <signature ...> : <streams> <wg_scale> <graph> <minc>
// (<device>: architecture and CU count; <signature>: variant, mixed
// precision, count, then m n s of every instance).  The file is read on the
// first batched call on a device and rewritten whole (a temporary file
// renamed over it) after each tuning; other devices' lines are kept, and
// lines that do not parse are dropped.
const char *const kTuneTag = "orth-tune-1";

std::string tune_device_id(int dev)
{
    hipDeviceProp_t p;
    if (hipGetDeviceProperties(&p, dev) != hipSuccess) {
        (void)hipGetLastError();
        return "unknown";
    }
    std::string arch = p.gcnArchName;
    const size_t colon = arch.find(':');  // target features
    if (colon != std::string::npos) arch.resize(colon);
    for (char &ch : arch)
        if (ch == ' ') ch = '_';
    return arch + "/" + std::to_string(p.multiProcessorCount);
}

bool parse_tune_line(const std::string &line, std::string *dev, std::vector<int64_t> *sig,
                     BatchCfg *cfg)
{
    std::istringstream in(line);
    std::string tag, tok;
    if (!(in >> tag) || tag != kTuneTag || !(in >> *dev)) return false;
    sig->clear();
    while (in >> tok && tok != ":") {
        char *end = nullptr;
        const long long v = strtoll(tok.c_str(), &end, 10);
        if (!end || *end) return false;
        sig->push_back(v);
    }
    int graph = -1;
    if (tok != ":" || !(in >> cfg->streams >> cfg->wg_scale >> graph >> cfg->minc)) return false;
    cfg->graph = graph == 1;
    return sig->size() >= 3 && (*sig)[2] >= 1 && sig->size() == 3 + 3 * (size_t)(*sig)[2] &&
           cfg->streams >= 1 && cfg->streams <= kMaxSlots && cfg->wg_scale > 0.0 &&
           cfg->wg_scale <= 1.0 && (graph == 0 || graph == 1) && cfg->minc >= 0;
}

std::string tune_line(const std::string &dev, const std::vector<int64_t> &sig, const BatchCfg &c)
{
    std::ostringstream o;
    o << kTuneTag << ' ' << dev;
    for (int64_t v : sig) o << ' ' << v;
    o << " : " << c.streams << ' ' << std::setprecision(17) << c.wg_scale << ' '
      << (c.graph ? 1 : 0) << ' ' << c.minc;
    return o.str();
}

void tune_cache_read(BatchCtx &b, int dev)
{
    if (b.cache_read) return;
    b.cache_read = true;
    const char *path = getenv("ORTH_TUNE_CACHE");
    if (!path || !*path) return;
    std::ifstream f(path);
    const std::string me = tune_device_id(dev);
    std::string line, d;
    std::vector<int64_t> sig;
    BatchCfg c;
    while (f && std::getline(f, line)) {
        if (!parse_tune_line(line, &d, &sig, &c) || d != me) continue;
        bool known = false;
        for (const auto &t : b.tuned) known = known || t.first == sig;
        if (!known) b.tuned.push_back({sig, c});
    }
}

void tune_cache_write(const BatchCtx &b, int dev)
{
    const char *path = getenv("ORTH_TUNE_CACHE");
    if (!path || !*path) return;
    const std::string me = tune_device_id(dev);
    std::vector<std::string> lines;
    {
        std::ifstream f(path);
        std::string line, d;
        std::vector<int64_t> sig;
        BatchCfg c;
        while (f && std::getline(f, line)) {
            if (!parse_tune_line(line, &d, &sig, &c)) continue;
            bool mine = false;
            if (d == me)
                for (const auto &t : b.tuned) mine = mine || t.first == sig;
            if (!mine) lines.push_back(line);
        }
    }
    for (const auto &t : b.tuned) lines.push_back(tune_line(me, t.first, t.second));
    const std::string tmp = std::string(path) + ".tmp" + std::to_string((long long)getpid());
    bool good = false;
    {
        std::ofstream o(tmp, std::ios::trunc);
        for (const auto &l : lines) o << l << '\n';
        o.flush();
        good = o.good();
    }
    if (!good || rename(tmp.c_str(), path) != 0) {
        fprintf(stderr, "orth: could not write the tuning cache %s\n", path);
        (void)remove(tmp.c_str());
    }
}

BatchCfg batch_cfg(const BatchCtx &b, const BatchArgs &a)
{
    const std::vector<int64_t> sig = shape_sig(a);
    for (const auto &t : b.tuned)
        if (t.first == sig) return t.second;
    BatchCfg c;
    c.streams = std::min(kMaxSlots, std::max(1, a.count));
    c.streams = std::min(c.streams, 4);
    return c;
}

bool batch_args_ok(const BatchArgs &a)
{
    if (a.count < 1 || !a.m || !a.n || !a.s || !a.X || !a.ldX || !a.R || !a.ldR) return false;
    // The batched mixed form is switched off: three runs of its tuning inside
    // the batched benchmark ended the process (segmentation fault, heap
    // corruption) with the cause not identified.
    if (a.mx) return false;
    for (int i = 0; i < a.count; i++) {
        if (!a.X[i] || !a.R[i] || !shape_ok(a.m[i], a.n[i], a.s[i], a.ldX[i], a.ldR[i])) return false;
        if (a.mx && (!a.X32[i] || a.ldX32[i] < a.m[i] || a.variant != ORTH_P1S || a.s[i] != 16 ||
                     a.n[i] > 256))
            return false;
    }
    return a.variant >= 0 && a.variant <= 4;
}

// Time the tuner's candidates on the batch itself; every run starts from the
// caller's input (restored from bk).  Returns false on a HIP failure.
bool batch_tune(BatchCtx &b, const BatchArgs &a, const InputBackup &bk, BatchCfg *best_out)
{
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (!ok(hipEventCreate(&e0)) || !ok(hipEventCreate(&e1))) return false;
    float best = 1e30f;
    BatchCfg bestc = *best_out;
    bool good = true;
    // Candidates: (streams, workgroup share) from two streams sharing the
    // device to eight; each with captured graphs and with direct launches.
    const std::pair<int, double> cands[] = {{2, 0.5}, {4, 0.5}, {8, 0.25}, {8, 0.5}, {4, 1.0}};
    for (const bool graph : {true, false}) {
        for (const auto &cd : cands) {
            if (cd.first > a.count || !good) continue;
            BatchCfg cand;
            cand.streams = cd.first;
            cand.wg_scale = cd.second;
            cand.graph = graph;
            // First run: reserves scratch, launches directly and captures;
            // the second (a replay in graph mode) is the one timed.
            good = bk.restore(a, b.main) && batch_run(b, a, cand) && bk.restore(a, b.main) &&
                   ok(hipEventRecord(e0, b.main)) && batch_run(b, a, cand) &&
                   ok(hipEventRecord(e1, b.main)) && ok(hipEventSynchronize(e1));
            float ms = 0.f;
            if (good && ok(hipEventElapsedTime(&ms, e0, e1)) && ms < best) {
                best = ms;
                bestc = cand;
            }
        }
    }
    // Tile shape of the s = 16 passes, with the best placement: the
    // double-buffered 32-row pass on every slab of 16 columns or more, or
    // never (the 512-thread super-tile pass throughout), against the
    // library's column bound.
    for (const int mc : {16, 1 << 20}) {
        if (!good) break;
        BatchCfg cand = bestc;
        cand.minc = mc;
        good = bk.restore(a, b.main) && batch_run(b, a, cand) && bk.restore(a, b.main) &&
               ok(hipEventRecord(e0, b.main)) && batch_run(b, a, cand) &&
               ok(hipEventRecord(e1, b.main)) && ok(hipEventSynchronize(e1));
        float ms = 0.f;
        if (good && ok(hipEventElapsedTime(&ms, e0, e1)) && ms < best) {
            best = ms;
            bestc = cand;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *best_out = bestc;
    return good;
}

int batched_entry(const BatchArgs &a, int *info, void *stream, int tune, int *cfg_out)
{
    if (!batch_args_ok(a)) return 2;
    tl_hip_error = 0;
    const int dev = current_device();
    DevCtx &c = g_ctx[dev];
    std::lock_guard<std::mutex> lk(c.mu);
    if (!ctx_init(c)) return 2;
    BatchCtx &b = g_batch[dev];
    if (b.broken || !batch_init(b)) return 2;
    tune_cache_read(b, dev);
    // Order after the caller's stream and the library's earlier calls.
    if (!ok(hipStreamSynchronize((hipStream_t)stream))) return 2;
    if (c.done_valid && !ok(hipEventSynchronize(c.done))) return 2;
    BatchCfg cfg = batch_cfg(b, a);
    if (tune) {
        // Tuning needs a device copy of the FP64 inputs (sum of 8 m n bytes);
        // without room for it the call runs untuned.
        InputBackup bk;
        if (a.mx != 0 || bk.save(a, b.main)) {
            BatchCfg best = cfg;
            if (!batch_tune(b, a, bk, &best)) return 2;
            const std::vector<int64_t> sig = shape_sig(a);
            bool found = false;
            for (auto &t : b.tuned)
                if (t.first == sig) {
                    t.second = best;
                    found = true;
                }
            if (!found) b.tuned.push_back({sig, best});
            tune_cache_write(b, dev);
            cfg = best;
            if (!bk.restore(a, b.main)) return 2;
        } else if (tl_hip_error) {
            return 2;
        }
    }
    if (!batch_run(b, a, cfg)) return 2;
    int worst = 0;
    for (int i = 0; i < a.count; i++) {
        if (info) {
            info[2 * i] = b.info_h[2 * i];
            info[2 * i + 1] = b.info_h[2 * i + 1];
        }
        if (b.info_h[2 * i] && !worst) worst = i + 1;
    }
    // orth_last_status reports the first instance that broke down.
    c.status_h[0] = worst ? b.info_h[2 * (worst - 1)] : 0;
    c.status_h[1] = worst ? b.info_h[2 * (worst - 1) + 1] : 0;
    c.last_passes = 0;
    c.last_syncs = 0;
    c.done_valid = false;
    if (cfg_out) {
        cfg_out[0] = cfg.streams * (cfg.graph ? 1 : -1);
        cfg_out[1] = (int)(cfg.wg_scale * 1000 + 0.5);
    }
    return 0;
}

}  // namespace

size_t scratch_p_need(int64_t n) { return (size_t)(n + 3 * kMaxS) * (4 * kMaxS) + 64; }
size_t scratch_r_need(int64_t n) { return (size_t)n * (n + kMaxS) + 64; }
size_t scratch_tbig_need(int64_t n) { return (size_t)(n + kMaxS) * kMaxS; }
size_t scratch_cbuf_need(int64_t n) { return (size_t)(n + 2 * kMaxS) * 2 * kMaxS; }
size_t scratch_stash_need() { return 6 * (size_t)kSlot; }

int run_on_fixed_scratch(int variant, int64_t m, int64_t n, int64_t s, double *dX, int64_t ldX,
                         double *dR, int64_t ldR, int *dInfo, Scratch &sc, int ncu, void *stream)
{
    if (!dX || !dR || !shape_ok(m, n, s, ldX, ldR) || variant < ORTH_P1S || variant > ORTH_P2S)
        return 2;
    tl_hip_error = 0;
    // A context without events or staging: per-kernel timing stays off.
    DevCtx c;
    c.ncu = ncu;
    hipStream_t st = (hipStream_t)stream;
    Run r(c, sc, m, n, s, dX, ldX, st);
    if (!r.prepare() || !r.run(variant) || !r.copy_R_out(dR, ldR, hipMemcpyDeviceToDevice))
        return 2;
    if (dInfo && !ok(hipMemcpyAsync(dInfo, sc.status_d, 2 * sizeof(int), hipMemcpyDeviceToDevice,
                                    st)))
        return 2;
    return 0;
}

}  // namespace orth

using namespace orth;

extern "C" {

int bcgsi2P1s(int m, int n, int s, double *Xsub, int ldXsub, double *R, int ldR, double *work,
              int lwork)
{
    return dropin(ORTH_P1S, m, n, s, Xsub, ldXsub, R, ldR, work, lwork);
}

int bcgsi21s(int m, int n, int s, double *Xsub, int ldXsub, double *R, int ldR, double *work,
             int lwork)
{
    return dropin(ORTH_1S, m, n, s, Xsub, ldXsub, R, ldR, work, lwork);
}

int bcgspipi2(int m, int n, int s, double *Xsub, int ldXsub, double *R, int ldR, double *work,
              int lwork)
{
    return dropin(ORTH_PIPI, m, n, s, Xsub, ldXsub, R, ldR, work, lwork);
}

int bcgsi2(int m, int n, int s, double *Xsub, int ldXsub, double *R, int ldR, double *work,
           int lwork)
{
    return dropin(ORTH_I2, m, n, s, Xsub, ldXsub, R, ldR, work, lwork);
}

int bcgsi2P2s(int m, int n, int s, double *Xsub, int ldXsub, double *R, int ldR, double *work,
              int lwork)
{
    return dropin(ORTH_P2S, m, n, s, Xsub, ldXsub, R, ldR, work, lwork);
}

int orth_last_status(orth_status_t *st)
{
    if (!st) return 2;
    const int dev = current_device();
    DevCtx &c = g_ctx[dev];
    std::lock_guard<std::mutex> lk(c.mu);
    fill_status(c);
    *st = c.last;
    return 0;
}

int orth_bcgs_mixed_device(int variant, int precision, int64_t m, int64_t n, int64_t s,
                           const float *dX32, int64_t ldX32, double *dQ, int64_t ldQ, double *dR,
                           int64_t ldR, void *stream)
{
    if (!dX32 || !dQ || !dR || !shape_ok(m, n, s, ldQ, ldR) || ldX32 < m) return 2;
    if (variant < ORTH_P1S || variant > ORTH_P2S) return 2;
    // The fused mixed pass covers BCGSI+P-1S with s = 16 and n <= 256; FP64
    // products on FP32 storage are also accepted for every other variant
    // and shape (the input is widened into Q, then the FP64 form runs).
    const bool fused_shape = variant == ORTH_P1S && s == 16 && n <= 256;
    if (!fused_shape && precision != ORTH_PREC_FP64) return 2;
    int mx = 0;
    switch (precision) {
    case ORTH_PREC_FP64: mx = kMxFP64; break;
    case ORTH_PREC_BF16: mx = kMxBF16; break;
    case ORTH_PREC_FP16: mx = kMxFP16; break;
    case ORTH_PREC_BF16X3: mx = kMxBF16X3; break;
    case ORTH_PREC_BF16_PROJ: mx = kMxBF16 | kMxProjOnly; break;
    case ORTH_PREC_FP16_PROJ: mx = kMxFP16 | kMxProjOnly; break;
    case ORTH_PREC_BF16X3_PROJ: mx = kMxBF16X3 | kMxProjOnly; break;
    default: return 2;
    }
    tl_hip_error = 0;
    const int dev = current_device();
    DevCtx &c = g_ctx[dev];
    std::lock_guard<std::mutex> lk(c.mu);
    if (!ctx_init(c)) return 2;
    if (!fused_shape) {
        // The widened copy overwrites Q: order it after the library's last
        // call, which may still be writing there.
        if (c.done_valid && !ok(hipStreamWaitEvent((hipStream_t)stream, c.done, 0))) return 2;
        if (!ok(launch_widen(dX32, ldX32, dQ, ldQ, m, n, (hipStream_t)stream))) return 2;
        return run_device(c, variant, m, n, s, dQ, ldQ, dR, ldR, hipMemcpyDeviceToDevice,
                          (hipStream_t)stream);
    }
    return run_device(c, variant, m, n, s, dQ, ldQ, dR, ldR, hipMemcpyDeviceToDevice,
                      (hipStream_t)stream, mx, dX32, ldX32);
}

int orth_bcgs_batched_device(int variant, int count, const int64_t *m, const int64_t *n,
                             const int64_t *s, double *const *dX, const int64_t *ldX,
                             double *const *dR, const int64_t *ldR, int *info, int tune,
                             int *config, void *stream)
{
    BatchArgs a;
    a.variant = variant;
    a.count = count;
    a.m = m;
    a.n = n;
    a.s = s;
    a.X = dX;
    a.ldX = ldX;
    a.R = dR;
    a.ldR = ldR;
    return batched_entry(a, info, stream, tune, config);
}

int orth_bcgs_batched_mixed_device(int variant, int precision, int count, const int64_t *m,
                                   const int64_t *n, const int64_t *s, const float *const *dX32,
                                   const int64_t *ldX32, double *const *dQ, const int64_t *ldQ,
                                   double *const *dR, const int64_t *ldR, int *info, int tune,
                                   int *config, void *stream)
{
    BatchArgs a;
    a.variant = variant;
    a.count = count;
    a.m = m;
    a.n = n;
    a.s = s;
    a.X = dQ;
    a.ldX = ldQ;
    a.R = dR;
    a.ldR = ldR;
    a.X32 = dX32;
    a.ldX32 = ldX32;
    switch (precision) {
    case ORTH_PREC_FP64: a.mx = kMxFP64; break;
    case ORTH_PREC_BF16: a.mx = kMxBF16; break;
    case ORTH_PREC_FP16: a.mx = kMxFP16; break;
    case ORTH_PREC_BF16X3: a.mx = kMxBF16X3; break;
    case ORTH_PREC_BF16_PROJ: a.mx = kMxBF16 | kMxProjOnly; break;
    case ORTH_PREC_FP16_PROJ: a.mx = kMxFP16 | kMxProjOnly; break;
    case ORTH_PREC_BF16X3_PROJ: a.mx = kMxBF16X3 | kMxProjOnly; break;
    default: return 2;
    }
    return batched_entry(a, info, stream, tune, config);
}

int orth_bcgs_device(int variant, int64_t m, int64_t n, int64_t s, double *dX, int64_t ldX,
                     double *dR, int64_t ldR, void *stream)
{
    if (!dX || !dR || !shape_ok(m, n, s, ldX, ldR)) return 2;
    tl_hip_error = 0;
    const int dev = current_device();
    DevCtx &c = g_ctx[dev];
    std::lock_guard<std::mutex> lk(c.mu);
    if (!ctx_init(c)) return 2;
    return run_device(c, variant, m, n, s, dX, ldX, dR, ldR, hipMemcpyDeviceToDevice,
                      (hipStream_t)stream);
}

int orth_gram_device(int64_t m, const double *dX, int64_t ldX, int64_t a0, int64_t na,
                     int64_t b0, int64_t nb, double *dC, int64_t ldc, void *stream)
{
    if (!dX || !dC || m < 1 || na < 1 || nb < 1 || ldc < na) return 2;
    tl_hip_error = 0;
    const int dev = current_device();
    DevCtx &c = g_ctx[dev];
    std::lock_guard<std::mutex> lk(c.mu);
    if (!ctx_init(c)) return 2;
    hipStream_t st = (hipStream_t)stream;
    if (c.done_valid && !ok(hipStreamWaitEvent(st, c.done, 0))) return 2;
    ProjPlan pl = plan_proj(m, na, nb, kProjTargetWG);
    if (pl.lds_bytes > 64 * 1024) return 2;
    if (!grow(c.sc, c.sc.part, c.sc.part_n, (size_t)pl.nwg * na * nb)) return 2;
    if (!ok(launch_proj(pl, dX, ldX, m, a0, na, b0, nb, c.sc.part, st))) return 2;
    if (!ok(launch_reduce(c.sc.part, pl.nwg, na, nb, dC, ldc, st))) return 2;
    if (!ok(hipEventRecord(c.done, st))) return 2;
    c.done_valid = true;
    return 0;
}

int orth_accuracy_device(int64_t m, int64_t n, const double *dX, int64_t ldX, const double *dQ,
                         int64_t ldQ, const double *dR, int64_t ldR, double *out, void *stream)
{
    if (!out || m < 1 || n < 1) return 2;
    hipStream_t st = (hipStream_t)stream;
    // Q^T Q with compensated accumulation (per row chunk and 64 x 64 block
    // of the upper block triangle), summed over the chunks in double-double
    // on the host in a fixed order.
    const int nch = gram_dd_chunks(m);
    const int64_t nb = (n + 63) / 64, nblk = nb * (nb + 1) / 2;
    const size_t npart = (size_t)nch * nblk * 256 * 32;
    double *part = nullptr;
    if (!ok(hipMalloc(&part, npart * sizeof(double)))) return 2;
    const int nwg = resid_wg_count(m);
    double *part2 = nullptr;
    if (!ok(hipMalloc(&part2, (size_t)2 * nwg * sizeof(double)))) {
        (void)hipFree(part);
        return 2;
    }
    std::vector<double> gp(npart), ph((size_t)2 * nwg);
    bool good = ok(launch_gram_dd(dQ, ldQ, m, n, part, st)) &&
                ok(launch_resid(m, n, dX, ldX, dQ, ldQ, dR, ldR, part2, st)) &&
                ok(hipMemcpyAsync(gp.data(), part, npart * sizeof(double),
                                  hipMemcpyDeviceToHost, st)) &&
                ok(hipMemcpyAsync(ph.data(), part2, ph.size() * sizeof(double),
                                  hipMemcpyDeviceToHost, st)) &&
                ok(hipStreamSynchronize(st));
    (void)hipFree(part);
    (void)hipFree(part2);
    if (!good) return 2;
    // ||Q^T Q - I||_F^2 over the pairs i <= j (off-diagonal ones twice).
    double loo = 0.0;
    int64_t blk = 0;
    for (int64_t I = 0; I < nb; I++)
        for (int64_t J = I; J < nb; J++, blk++)
            for (int t = 0; t < 256; t++)
                for (int k = 0; k < 16; k++) {
                    const int64_t i = 64 * I + 4 * (t % 16) + k / 4;
                    const int64_t j = 64 * J + 4 * (t / 16) + k % 4;
                    if (i >= n || j >= n || i > j) continue;
                    double H = 0.0, L = 0.0;
                    for (int c = 0; c < nch; c++) {
                        const double *g = &gp[(((size_t)c * nblk + blk) * 256 + t) * 32 + 2 * k];
                        const double s = H + g[0], bp = s - H;
                        L += (H - (s - bp)) + (g[0] - bp) + g[1];
                        H = s;
                    }
                    const double v = (i == j ? H - 1.0 : H) + L;
                    loo += (i == j ? 1.0 : 2.0) * v * v;
                }
    double sr = 0.0, sx = 0.0;
    for (int k = 0; k < nwg; k++) {
        sr += ph[2 * k];
        sx += ph[2 * k + 1];
    }
    out[0] = sqrt(loo);
    out[1] = sqrt(sr);
    out[2] = sqrt(sx);
    return 0;
}

int orth_set_pythagorean_shift(double c)
{
    if (!(c >= 0.0) || !std::isfinite(c)) return 2;
    g_pshift.store(c);
    return 0;
}

int orth_kernel_timing(int enable)
{
    const int dev = current_device();
    DevCtx &c = g_ctx[dev];
    std::lock_guard<std::mutex> lk(c.mu);
    if (c.done_valid) (void)hipEventSynchronize(c.done);
    c.ktime = enable != 0;
    c.ev_used = 0;
    for (auto &v : c.kt) v.clear();
    for (int k = 0; k < ORTH_K_COUNT; k++) c.kt_flops[k] = c.kt_bytes[k] = 0.0;
    return 0;
}

int orth_kernel_times(int kind, double *total_ms, int *count, double *flops, double *bytes)
{
    if (kind < 0 || kind >= ORTH_K_COUNT || !total_ms || !count) return 2;
    const int dev = current_device();
    DevCtx &c = g_ctx[dev];
    std::lock_guard<std::mutex> lk(c.mu);
    if (c.done_valid && !ok(hipEventSynchronize(c.done))) return 2;
    double t = 0.0;
    for (auto &pr : c.kt[kind]) {
        float ms = 0.0f;
        if (!ok(hipEventElapsedTime(&ms, pr.first, pr.second))) return 2;
        t += ms;
    }
    *total_ms = t;
    *count = (int)c.kt[kind].size();
    if (flops) *flops = c.kt_flops[kind];
    if (bytes) *bytes = c.kt_bytes[kind];
    return 0;
}

int orth_update_device(int64_t m, double *dX, int64_t ldX, int64_t col0, int64_t w,
                       int64_t c, const double *dW, int64_t ldW, const double *dT,
                       void *stream)
{
    const int S = block_template(w);
    if (!dX || m < 1 || w < 1 || S == 0 || c < 0 || col0 < c || (c > 0 && !dW) ||
        (c > 0 && ldW < c))
        return 2;
    tl_hip_error = 0;
    const int dev = current_device();
    DevCtx &ctx = g_ctx[dev];
    std::lock_guard<std::mutex> lk(ctx.mu);
    if (!ctx_init(ctx)) return 2;
    hipStream_t st = (hipStream_t)stream;
    if (ctx.done_valid && !ok(hipStreamWaitEvent(st, ctx.done, 0))) return 2;
    // Coefficients are staged with S zero-padded columns and the factor with
    // leading dimension kLdT, the layout the update kernel reads.  The
    // staging buffers are released on every return (the call is synchronous).
    struct Staging {
        double *p = nullptr;
        ~Staging()
        {
            if (p) (void)hipFree(p);
        }
    } Wp, Tp;
    if (c > 0) {
        if (!ok(hipMalloc(&Wp.p, (size_t)c * S * sizeof(double)))) return 2;
        if (!ok(launch_fill(Wp.p, (size_t)c * S, st)) ||
            !ok(hipMemcpy2DAsync(Wp.p, c * sizeof(double), dW, ldW * sizeof(double),
                                 c * sizeof(double), w, hipMemcpyDeviceToDevice, st))) {
            (void)hipStreamSynchronize(st);
            return 2;
        }
    }
    if (dT) {
        if (!ok(hipMalloc(&Tp.p, (size_t)kLdT * kLdT * sizeof(double)))) {
            (void)hipStreamSynchronize(st);
            return 2;
        }
        if (!ok(launch_fill(Tp.p, (size_t)kLdT * kLdT, st)) ||
            !ok(hipMemcpy2DAsync(Tp.p, kLdT * sizeof(double), dT, w * sizeof(double),
                                 w * sizeof(double), w, hipMemcpyDeviceToDevice, st))) {
            (void)hipStreamSynchronize(st);
            return 2;
        }
    }
    Upd u;
    u.col0 = col0;
    u.w = (int)w;
    u.c = (int)c;
    u.W = Wp.p;
    u.ldW = c;
    u.T = Tp.p;
    const bool launched = ok(launch_update(S, dX, ldX, m, &u, 1, st));
    const bool done = ok(hipStreamSynchronize(st));
    return launched && done ? 0 : 2;
}

int orth_stream_probe(int mode, double *dst, const double *src, int64_t bytes, double *ms_out)
{
    if (!src || !ms_out || (mode == 0 && !dst) || mode < 0 || mode > 1 || bytes < 16) return 2;
    tl_hip_error = 0;
    const int dev = current_device();
    DevCtx &c = g_ctx[dev];
    std::lock_guard<std::mutex> lk(c.mu);
    if (!ctx_init(c)) return 2;
    // Eight workgroups per CU (a grid sweep on MI355X showed no gain beyond).
    const int nwg = 8 * std::max(1, c.ncu);
    if (mode == 1 && !grow(c.sc, c.sc.part, c.sc.part_n, (size_t)nwg)) return 2;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (!ok(hipEventCreate(&e0)) || !ok(hipEventCreate(&e1))) return 2;
    bool good = ok(hipDeviceSynchronize()) && ok(hipEventRecord(e0, nullptr)) &&
                ok(launch_stream_probe(mode, dst, src, bytes, c.sc.part, nwg, nullptr)) &&
                ok(hipEventRecord(e1, nullptr)) && ok(hipEventSynchronize(e1));
    float ms = 0.f;
    good = good && ok(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *ms_out = ms;
    return good ? 0 : 2;
}

int orth_sync(void)
{
    const int dev = current_device();
    DevCtx &c = g_ctx[dev];
    std::lock_guard<std::mutex> lk(c.mu);
    if (c.done_valid && !ok(hipEventSynchronize(c.done))) return 2;
    return 0;
}

int orth_release(void)
{
    const int dev = current_device();
    DevCtx &c = g_ctx[dev];
    std::lock_guard<std::mutex> lk(c.mu);
    if (c.done_valid) (void)hipEventSynchronize(c.done);
    auto free_scratch = [](DevCtx &x) {
        double **bufs[] = {&x.sc.part, &x.sc.P[0], &x.sc.P[1], &x.sc.Rd, &x.sc.Tbig, &x.sc.Cbuf,
                           &x.Xd};
        for (double **b : bufs) {
            if (*b) (void)hipFree(*b);
            *b = nullptr;
        }
        x.sc.part_n = x.sc.P_n = x.sc.Rd_n = x.sc.Tbig_n = x.sc.Cbuf_n = x.Xd_n = 0;
    };
    free_scratch(c);
    // The batched form's captured graphs and per-stream scratch (the batch
    // calls hold c.mu as well).
    BatchCtx &b = g_batch[dev];
    if (b.init) {
        for (int k = 0; k < kMaxSlots; k++)
            if (b.st[k]) (void)hipStreamSynchronize(b.st[k]);
        if (b.main) (void)hipStreamSynchronize(b.main);
        drop_graphs(b);
        for (int k = 0; k < kMaxSlots; k++) free_scratch(b.slot[k]);
        if (b.info_d) (void)hipFree(b.info_d);
        if (b.info_h) (void)hipHostFree(b.info_h);
        b.info_d = nullptr;
        b.info_h = nullptr;
        b.info_n = 0;
    }
    return 0;
}

}  // extern "C"
