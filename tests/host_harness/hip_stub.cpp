// Implementation of the host-only HIP stand-in (hip/hip_runtime.h in this
// folder) and do-nothing launch wrappers with the signatures of
// one-synchronization-bcgs_amd/csrc/orth_kernels.h.  Device buffers are host
// allocations (checked by AddressSanitizer); capture rules of the real
// runtime that the library relies on are enforced and reported through
// stub_fail(): no synchronous call, allocation or free while a capture is
// active, no capture on a stream that already captures, and EndCapture
// only when every stream that joined the capture has been joined back.
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <map>
#include <mutex>
#include <set>

#include "orth_kernels.h"

int g_stub_failures = 0;

static void stub_fail(const char *what)
{
    fprintf(stderr, "HIP stub: %s\n", what);
    g_stub_failures++;
}

struct ihipStream_t {
    int capture = 0;              // id of the capture this stream belongs to (0 = none)
    bool origin = false;          // the stream that began the capture
    unsigned long last_work = 0;  // sequence number of its latest work
    unsigned long joined = 0;     // latest work the origin has waited for
};
struct ihipEvent_t {
    hipStream_t stream = nullptr;
    unsigned long seq = 0;
    int capture = 0;
    bool recorded = false;
};
struct ihipGraph {
    int id;
};
struct hipGraphExec {
    int id;
};

static std::mutex g_mu;
static unsigned long g_seq = 0;
static int g_next_capture = 1;
static int g_active_captures = 0;
static std::set<hipStream_t> g_streams;
static std::set<hipEvent_t> g_events;
static std::map<void *, size_t> g_dev;  // "device" allocations
static unsigned g_time = 12345;

static void work(hipStream_t s)
{
    if (!s) return;  // the null stream: never captured by the library
    s->last_work = ++g_seq;
}

// Launch failure injection: STUB_FAIL_LAUNCH=k makes the k-th launch (and
// only it) report an error, to exercise the library's error paths.
static long g_launches = 0;
static long g_fail_at = -1;

hipError_t stub_launch(hipStream_t s)
{
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_fail_at < 0) {
        const char *e = getenv("STUB_FAIL_LAUNCH");
        g_fail_at = e ? atol(e) : 0;
    }
    if (g_fail_at > 0 && ++g_launches == g_fail_at) return hipErrorInvalidValue;
    work(s);
    return hipSuccess;
}

hipError_t hipMalloc(void **p, size_t n)
{
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_active_captures) stub_fail("hipMalloc while a stream captures");
    *p = calloc(1, n ? n : 1);
    if (!*p) return hipErrorOutOfMemory;
    g_dev[*p] = n;
    return hipSuccess;
}
hipError_t hipFree(void *p)
{
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_active_captures) stub_fail("hipFree while a stream captures");
    if (!p) return hipSuccess;
    if (!g_dev.erase(p)) stub_fail("hipFree of a pointer hipMalloc did not return");
    free(p);
    return hipSuccess;
}
hipError_t hipHostMalloc(void **p, size_t n, unsigned)
{
    *p = calloc(1, n ? n : 1);
    return *p ? hipSuccess : hipErrorOutOfMemory;
}
hipError_t hipHostFree(void *p)
{
    free(p);
    return hipSuccess;
}
hipError_t hipMemcpy(void *d, const void *s, size_t n, hipMemcpyKind)
{
    {
        std::lock_guard<std::mutex> lk(g_mu);
        if (g_active_captures) stub_fail("synchronous hipMemcpy while a stream captures");
    }
    if (n) memmove(d, s, n);
    return hipSuccess;
}
hipError_t hipMemcpyAsync(void *d, const void *s, size_t n, hipMemcpyKind, hipStream_t st)
{
    if (n) memmove(d, s, n);
    std::lock_guard<std::mutex> lk(g_mu);
    work(st);
    return hipSuccess;
}
hipError_t hipMemcpy2DAsync(void *d, size_t dpitch, const void *s, size_t spitch, size_t width,
                            size_t height, hipMemcpyKind, hipStream_t st)
{
    if (width > dpitch || width > spitch) {
        stub_fail("hipMemcpy2DAsync with a row wider than its pitch");
        return hipErrorInvalidValue;
    }
    for (size_t r = 0; r < height; r++)
        memmove((char *)d + r * dpitch, (const char *)s + r * spitch, width);
    std::lock_guard<std::mutex> lk(g_mu);
    work(st);
    return hipSuccess;
}
hipError_t hipStreamCreateWithFlags(hipStream_t *s, unsigned)
{
    std::lock_guard<std::mutex> lk(g_mu);
    *s = new ihipStream_t;
    g_streams.insert(*s);
    return hipSuccess;
}
hipError_t hipStreamDestroy(hipStream_t s)
{
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_streams.erase(s)) stub_fail("hipStreamDestroy of an unknown stream");
    delete s;
    return hipSuccess;
}
hipError_t hipStreamSynchronize(hipStream_t s)
{
    std::lock_guard<std::mutex> lk(g_mu);
    if (s && s->capture) stub_fail("hipStreamSynchronize on a capturing stream");
    return hipSuccess;
}
hipError_t hipStreamWaitEvent(hipStream_t s, hipEvent_t e, unsigned)
{
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_events.count(e)) {
        stub_fail("hipStreamWaitEvent on an unknown event");
        return hipErrorInvalidValue;
    }
    if (!s) return hipSuccess;
    if (e->capture) {
        if (s->capture == 0) {
            s->capture = e->capture;  // joins the capture
        } else if (s->capture != e->capture) {
            stub_fail("hipStreamWaitEvent across two captures");
        }
        if (s->origin && e->stream && e->stream != s)
            e->stream->joined = e->seq > e->stream->joined ? e->seq : e->stream->joined;
    }
    work(s);
    return hipSuccess;
}
hipError_t hipStreamBeginCapture(hipStream_t s, hipStreamCaptureMode)
{
    std::lock_guard<std::mutex> lk(g_mu);
    if (s->capture) {
        stub_fail("hipStreamBeginCapture on a capturing stream");
        return hipErrorIllegalState;
    }
    s->capture = g_next_capture++;
    s->origin = true;
    g_active_captures++;
    return hipSuccess;
}
hipError_t hipStreamEndCapture(hipStream_t s, hipGraph_t *g)
{
    std::lock_guard<std::mutex> lk(g_mu);
    if (!s->capture || !s->origin) {
        stub_fail("hipStreamEndCapture on a stream that did not begin a capture");
        return hipErrorIllegalState;
    }
    const int id = s->capture;
    hipError_t r = hipSuccess;
    for (hipStream_t o : g_streams) {
        if (o == s || o->capture != id) continue;
        if (o->last_work > o->joined) r = hipErrorStreamCaptureUnjoined;
        o->capture = 0;
        o->joined = 0;
    }
    s->capture = 0;
    s->origin = false;
    g_active_captures--;
    if (r != hipSuccess) {
        *g = nullptr;
        return r;
    }
    *g = new ihipGraph{id};
    return hipSuccess;
}
hipError_t hipStreamIsCapturing(hipStream_t s, hipStreamCaptureStatus *st)
{
    std::lock_guard<std::mutex> lk(g_mu);
    *st = s && s->capture ? hipStreamCaptureStatusActive : hipStreamCaptureStatusNone;
    return hipSuccess;
}
hipError_t hipEventCreate(hipEvent_t *e)
{
    std::lock_guard<std::mutex> lk(g_mu);
    *e = new ihipEvent_t;
    g_events.insert(*e);
    return hipSuccess;
}
hipError_t hipEventCreateWithFlags(hipEvent_t *e, unsigned)
{
    return hipEventCreate(e);
}
hipError_t hipEventDestroy(hipEvent_t e)
{
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_events.erase(e)) {
        stub_fail("hipEventDestroy of an unknown event");
        return hipErrorInvalidValue;
    }
    delete e;
    return hipSuccess;
}
hipError_t hipEventRecord(hipEvent_t e, hipStream_t s)
{
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_events.count(e)) {
        stub_fail("hipEventRecord of an unknown event");
        return hipErrorInvalidValue;
    }
    work(s);
    e->stream = s;
    e->seq = s ? s->last_work : 0;
    e->capture = s ? s->capture : 0;
    e->recorded = true;
    return hipSuccess;
}
hipError_t hipEventSynchronize(hipEvent_t e)
{
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_events.count(e)) {
        stub_fail("hipEventSynchronize of an unknown event");
        return hipErrorInvalidValue;
    }
    if (e->capture) stub_fail("hipEventSynchronize on an event recorded in a capture");
    return hipSuccess;
}
hipError_t hipEventElapsedTime(float *ms, hipEvent_t a, hipEvent_t b)
{
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_events.count(a) || !g_events.count(b) || !a->recorded || !b->recorded) {
        stub_fail("hipEventElapsedTime on an unrecorded event");
        return hipErrorInvalidValue;
    }
    g_time = g_time * 1103515245u + 12345u;
    *ms = 1.0f + (float)((g_time >> 8) % 1000) / 100.0f;
    return hipSuccess;
}
hipError_t hipGraphInstantiate(hipGraphExec_t *x, hipGraph_t g, hipGraphNode_t *, char *, size_t)
{
    if (!g) {
        stub_fail("hipGraphInstantiate of a null graph");
        return hipErrorInvalidValue;
    }
    *x = new hipGraphExec{g->id};
    return hipSuccess;
}
hipError_t hipGraphLaunch(hipGraphExec_t x, hipStream_t s)
{
    if (!x) {
        stub_fail("hipGraphLaunch of a null executable graph");
        return hipErrorInvalidValue;
    }
    std::lock_guard<std::mutex> lk(g_mu);
    work(s);
    return hipSuccess;
}
hipError_t hipGraphExecDestroy(hipGraphExec_t x)
{
    delete x;
    return hipSuccess;
}
hipError_t hipGraphDestroy(hipGraph_t g)
{
    delete g;
    return hipSuccess;
}
hipError_t hipGetLastError(void)
{
    return hipSuccess;
}
hipError_t hipGetDevice(int *d)
{
    *d = 0;
    return hipSuccess;
}
hipError_t hipDeviceGetAttribute(int *v, hipDeviceAttribute_t, int)
{
    *v = 256;
    return hipSuccess;
}
hipError_t hipGetDeviceProperties(hipDeviceProp_t *p, int)
{
    memset(p, 0, sizeof(*p));
    strcpy(p->name, "stand-in");
    strcpy(p->gcnArchName, "gfx950:sramecc+:xnack-");
    p->multiProcessorCount = 256;
    return hipSuccess;
}
hipError_t hipDeviceSynchronize(void)
{
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_active_captures) stub_fail("hipDeviceSynchronize while a stream captures");
    return hipSuccess;
}
hipError_t hipPointerGetAttributes(hipPointerAttribute_t *a, const void *p)
{
    std::lock_guard<std::mutex> lk(g_mu);
    memset(a, 0, sizeof(*a));
    auto it = g_dev.upper_bound((void *)p);
    if (it != g_dev.begin()) {
        --it;
        if ((const char *)p < (const char *)it->first + it->second) a->type = hipMemoryTypeDevice;
    }
    return hipSuccess;
}
const char *hipGetErrorString(hipError_t)
{
    return "stub";
}

// ---- launch wrappers and plans of orth_kernels.h: no device work.
namespace orth {

hipError_t launch_update(int, double *, int64_t, int64_t, const Upd *, int, hipStream_t st)
{
    return stub_launch(st);
}
ProjPlan plan_proj(int64_t m, int64_t na, int64_t nb, int target_wg)
{
    ProjPlan p{};
    p.Ma = (int)((na + 15) / 16);
    p.Nb = (int)((nb + 15) / 16);
    p.ga = 1;
    p.gb = 1;
    p.ngy = p.Ma * p.Nb;
    int64_t nwg = target_wg / p.ngy;
    if (nwg < 1) nwg = 1;
    if (nwg > (m + 15) / 16) nwg = (m + 15) / 16;
    p.rows_per = ((m + nwg - 1) / nwg + 15) / 16 * 16;
    p.nwg = (int)((m + p.rows_per - 1) / p.rows_per);
    p.lds_bytes = 1024;
    return p;
}
hipError_t launch_proj(const ProjPlan &, const double *, int64_t, int64_t, int64_t, int64_t,
                       int64_t, int64_t, double *, hipStream_t st)
{
    return stub_launch(st);
}
hipError_t launch_reduce(const double *, int, int64_t, int64_t, double *, int64_t, hipStream_t st)
{
    return stub_launch(st);
}
hipError_t launch_small(const SmallArgs &, hipStream_t st)
{
    return stub_launch(st);
}
hipError_t launch_reduce_small(const double *, int, int64_t, int64_t, double *, int *,
                               const SmallArgs &, hipStream_t st)
{
    return stub_launch(st);
}
size_t pass_lds_bytes(const PassArgs &a, int F)
{
    return (size_t)a.wl * 32 * F * 8 <= 160 * 1024 ? (size_t)a.wl * 32 * F * 8 + 8 : 0;
}
hipError_t launch_pass(int, const PassArgs &, int, hipStream_t st)
{
    return stub_launch(st);
}
int pass_wg_per_cu(int, const PassArgs &, bool)
{
    return 2;
}
int pass_rows(int, const PassArgs &)
{
    return 32;
}
int pass16b_plan(const PassArgs &a, int *mt, int *nb)
{
    if (a.n1 != 32) return 0;
    *mt = (((a.na + 15) / 16) + 7) / 8;
    if (*mt < 1) *mt = 1;
    *nb = (a.nb + 15) / 16;
    if (*nb < 1) *nb = 1;
    return a.ncols <= 64 ? 128 : a.ncols <= 160 ? 64 : 32;
}
size_t pass16b_lds_bytes(const PassArgs &a, int ts)
{
    return (size_t)a.wl * (ts + 1) * 8 + 4096;
}
hipError_t launch_pass16b(const PassArgs &, int, hipStream_t st)
{
    return stub_launch(st);
}
int pass16d_plan(const PassArgs &a, int *mt, int *nb, bool any_width, int min_cols)
{
    const int minc = any_width ? 1 : min_cols > 0 ? min_cols : 128;
    if (a.n1 != 32 || a.seq || a.ncols < minc || a.wl > 256) return 0;
    *mt = 1;
    *nb = a.na > 0 ? (a.nb + 15) / 16 : 1;
    return a.wl <= 128 ? 8 : 16;
}
hipError_t launch_pass16d(const PassArgs &, int, int, int, hipStream_t st)
{
    return stub_launch(st);
}
hipError_t launch_pass16d_mixed(const PassArgs &, int, int, int, int, hipStream_t st)
{
    return stub_launch(st);
}
size_t pass16m_lds_bytes(const PassArgs &a)
{
    return (size_t)a.wl * 33 * 8;
}
hipError_t launch_pass16m(const PassArgs &, int, int, hipStream_t st)
{
    return stub_launch(st);
}
int pass16bm_plan(const PassArgs &a)
{
    return a.n1 == 32 && a.ldX32 % 4 == 0 ? 32 : 0;
}
hipError_t launch_pass16bm(const PassArgs &, int, int, hipStream_t st)
{
    return stub_launch(st);
}
hipError_t launch_stream_probe(int, double *, const double *, int64_t, double *, int, hipStream_t st)
{
    return stub_launch(st);
}
hipError_t launch_widen(const float *, int64_t, double *, int64_t, int64_t, int64_t, hipStream_t st)
{
    return stub_launch(st);
}
hipError_t launch_fill(double *p, size_t n, hipStream_t st)
{
    memset(p, 0, n * sizeof(double));
    return stub_launch(st);
}
hipError_t launch_fill_int(int *p, size_t n, hipStream_t st)
{
    memset(p, 0, n * sizeof(int));
    return stub_launch(st);
}
int resid_wg_count(int64_t)
{
    return 64;
}
int gram_dd_chunks(int64_t m)
{
    return m >= 4096 ? 64 : 1;
}
hipError_t launch_gram_dd(const double *, int64_t, int64_t, int64_t, double *, hipStream_t st)
{
    return stub_launch(st);
}
hipError_t launch_resid(int64_t, int64_t, const double *, int64_t, const double *, int64_t,
                        const double *, int64_t, double *part2, hipStream_t st)
{
    memset(part2, 0, 2 * 64 * sizeof(double));
    return stub_launch(st);
}
}  // namespace orth
