// abi.hip — the C ABI of libprr.so (include/prr_abi.h).
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "../../include/prr_abi.h"
#include "device/dev_traverse.hpp"
#include "device/launch.hpp"
#include "device/launch_photon.hpp"
#include "device/launch_render.hpp"
#include "context.hpp"
#include "host/scene_build.hpp"

using namespace prr;

static thread_local std::string g_lastError = "";

// Per device: its address of the static spill array behind
// prr_trace_closest_noalloc, looked up once (hipGetSymbolAddress with the
// device current) by the first prr_create on that device.
static std::mutex g_noAllocMu;
static uint2 *g_noAllocSpill[256] = {};

int fail(prr_context *c, int code, const std::string &msg) {
    if (c) c->err = msg;
    g_lastError = msg;
    return code;
}

static void freePhotons(prr_scene *s) {
    for (void *p : s->ph.allocs) (void) hipFree(p);
    s->ph = PhotonMap();
    s->view.ph_count = 0;
    s->view.ph_pos = s->view.ph_dir = s->view.ph_pow = nullptr;
    s->view.ph_start = s->view.ph_end = nullptr;
}

template <class T>
static hipError_t upload(prr_scene *s, T **dst, const std::vector<T> &src, hipStream_t st) {
    const size_t bytes = std::max<size_t>(sizeof(T), src.size() * sizeof(T));
    void *p = nullptr;
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess) return e;
    s->allocs.push_back(p);
    s->bytes += (int64_t) bytes;
    *dst = (T *) p;
    if (!src.empty()) e = hipMemcpyAsync(p, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice, st);
    return e;
}

// AllenwpParams(1.25, 16, 1) (ToneMapping.hpp:125-147), evaluated on the host.
static AgxConst agxConst() {
    const float crossover = 0.1841865f, contrast = 1.25f, outMax = 1.0f;
    const float high_clip = std::max(16.0f, outMax);
    AgxConst p;
    p.toe_a = ((1.f / crossover) - 1.f) * std::pow(crossover, contrast);
    const float denom = std::pow(crossover, contrast) + p.toe_a;
    p.slope = (contrast * std::pow(crossover, contrast - 1.f) * p.toe_a) / (denom * denom);
    p.shoulder_max = outMax - crossover;
    float w = high_clip - crossover;
    w = w * w;
    w = w / p.shoulder_max;
    w = w * p.slope;
    p.w = w;
    return p;
}

namespace prr_host {

void freeRenderBufs(prr_context *c) {
    for (void *p : c->pbAllocs) (void) hipFree(p);
    c->pbAllocs.clear();
    c->pbSlots = 0;
    c->maxDepthAlloc = 0;
}

// ---------------------------------------------------------------------------
// render
// ---------------------------------------------------------------------------

int ensureRenderBufs(prr_context *c, int64_t slots, int maxDepth) {
    if (slots <= c->pbSlots && maxDepth <= c->maxDepthAlloc) return PRR_OK;
    HIPCK(c, hipStreamSynchronize(c->stream));
    freeRenderBufs(c);
    PathBufs &p = c->pb;
    p = PathBufs{};
    p.slots = slots;
    hipError_t e = hipSuccess;
    auto A = [&](auto **ptr, size_t count) {
        using T = typename std::remove_reference<decltype(**ptr)>::type;
        if (e != hipSuccess) return;
        void *q = nullptr;
        e = hipMalloc(&q, std::max<size_t>(1, count) * sizeof(T));
        if (e == hipSuccess) {
            c->pbAllocs.push_back(q);
            *ptr = (T *) q;
        }
    };
    A(&p.L, slots);
    A(&p.beta, slots);
    A(&p.prevP, slots);
    A(&p.pend, 4 * slots);
    A(&p.rng, slots);
    A(&p.vis, 3 * slots);
    A(&p.result, slots);
    // segmented queues: kQSegs segments of qcap (ray) / sqcap (shadow) entries
    p.qcap = segCapacity(slots, 1);
    p.sqcap = segCapacity(slots, 3);
    const size_t qEntries = (size_t) kQSegs * p.qcap, sqEntries = (size_t) kQSegs * p.sqcap;
    for (int k = 0; k < 2; k++) {
        A(&p.qSlot[k], qEntries);
        A(&p.qOrg[k], 3 * qEntries);
        A(&p.qDir[k], 3 * qEntries);
    }
    A(&p.hit, qEntries);
    A(&p.sqSlot, sqEntries);
    A(&p.sqOrg, sqEntries);
    A(&p.sqDir, 3 * sqEntries);
    A(&p.qcnt, (size_t) maxDepth + 2);
    A(&p.sqcnt, (size_t) maxDepth + 2);
    A(&p.segc, ((size_t) maxDepth + 2) * kQSegs);
    A(&p.qpre, ((size_t) maxDepth + 2) * (kQSegs + 1));
    A(&p.sqpre, ((size_t) maxDepth + 2) * (kQSegs + 1));
    A(&p.fetch, 2 * ((size_t) maxDepth + 2));
    c->bins = ShadeBins{};
    if (c->shadeSort) {
        c->bins.cap = slots;  // a list holds at most the queue's entries
        c->bins.blocks = 4 * c->numCU;
        A(&c->bins.lists, 2 * slots);
        A(&c->bins.counts, 2);
        A(&c->bins.blockCounts, 2 * (size_t) c->bins.blocks);
    }
    c->sb = ShadowSortBufs{};
    if (c->shadowSortFrom > 0) {
        // as many entries as the shadow queue holds (up to three NEE rays per path), in order
        ShadowSortBufs &b = c->sb;
        b.cap = 3 * slots;
        b.blocks = 4 * c->numCU;
        A(&b.slot, 3 * slots);
        A(&b.org, 3 * slots);
        A(&b.dir, 9 * slots);
        A(&b.hist, (size_t) shadowBinCount(c->shadowSortBits) * b.blocks);
    }
    if (e != hipSuccess) {
        freeRenderBufs(c);
        c->sb = ShadowSortBufs{};
        return fail(c, PRR_ERR_NOMEM, std::string("render buffers: ") + hipGetErrorString(e));
    }
    p.stats = c->stats;
    c->pbSlots = slots;
    c->maxDepthAlloc = maxDepth;
    return PRR_OK;
}


// Runs bounces 0..maxDepth of one chunk whose queue 0 is already filled.
int runBounces(prr_context *c, const SceneView &view, const ChunkInfo &ci, Timing &tm, hipStream_t st) {
    auto evPair = [&](int kind) -> int {
        if (!tm.on) return -1;
        if (tm.next + 2 > c->evPool.size()) {
            for (int k = 0; k < 256; k++) {
                hipEvent_t e;
                if (hipEventCreate(&e) != hipSuccess) return -1;
                c->evPool.push_back(e);
            }
        }
        const int i = (int) tm.next;
        tm.next += 2;
        tm.timed.emplace_back(kind, i);
        return i;
    };
#define TIMED(kind, strm, launch)                                         \
    do {                                                                  \
        int ei_ = evPair(kind);                                           \
        if (ei_ >= 0) HIPCK(c, hipEventRecord(c->evPool[ei_], strm));     \
        HIPCK(c, launch);                                                 \
        if (ei_ >= 0) HIPCK(c, hipEventRecord(c->evPool[ei_ + 1], strm)); \
    } while (0)
    // Bounce b: extension rays of b and shadow rays of b - 1 (both written by
    // shade b - 1, independent of each other), then shade b, which reads both.
    // With path completion on, kFinish is enqueued at every bounce too: the
    // first bounce whose queue holds at most tailPaths entries runs there and
    // kExtend / kShade of it and of every later bounce do nothing. Each
    // bounce's count is copied to the host behind an event. Before enqueuing
    // bounce b the host waits for the count of b - 1 (final once shade b - 2
    // ended, while bounce b - 1's kernels still occupy the GPU) and stops
    // once it shows the switch, so no launch of a later bounce is wasted and
    // the GPU never waits for the host.
    // Shadow rays on the second stream beside the extension rays hide each
    // launch's tail behind the other's bulk, which pays for small chunks;
    // large chunks keep the GPU full with one kernel at a time (the same frame
    // time, measured) and each kernel's time is its own.
    const bool conc = c->concurrentShadow && c->aux && (int64_t) ci.pixels * ci.spc <= c->concurrentMaxPaths;
    const int wwShadow = c->shadowWide ? c->whileWhileAny : c->whileWhile;
    const bool tail = c->tailPaths > 0 && c->cntHost && ci.maxDepth + 1 < prr_context::kMaxBounces;
    // A radiance-cache render stops its other paths at the query bounce, and
    // only its training paths (one in train_every) go on: completing those in
    // one launch once they number at most 2^19 saves their per-bounce
    // launches (measured: 12.2 -> 12.0 ms per 1080p 4-spp frame with
    // full-depth training; plain frames are best at the default).
    const int64_t tailAt = ci.cache.query ? std::max<int64_t>(c->tailPaths, int64_t(1) << 19) : c->tailPaths;
    ChunkInfo cit = ci;
    cit.tailPaths = tail ? tailAt : 0;
    for (int b = 0; b <= ci.maxDepth; b++) {
        if (tail && b > 0) {
            HIPCK(c, hipEventSynchronize(c->evCnt[b - 1]));
            if ((int64_t) c->cntHost[b - 1] <= tailAt) break;
        }
        if (tail) {
            HIPCK(c, hipMemcpyAsync(c->cntHost + b, c->pb.qcnt + b, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
            HIPCK(c, hipEventRecord(c->evCnt[b], st));
        }
        // shadow rays of bounce b - 1, binned first (from bounce shadowSortFrom
        // on) on the render stream, ahead of the extension rays: a short pass
        // that would otherwise wait for the persistent traversal grid's CUs
        PathBufs pbS = c->pb;
        bool denseS = false;
        if (b > 0 && c->shadowSortFrom > 0 && b - 1 >= c->shadowSortFrom && c->sb.cap > 0) {
            pbS.sqSlot = c->sb.slot;
            pbS.sqOrg = c->sb.org;
            pbS.sqDir = c->sb.dir;
            denseS = true;
            TIMED(4, st, launchShadowSort(view, c->pb, b - 1, c->shadowSortBits, c->sb, st));
        }
        if (b > 0 && conc) {
            HIPCK(c, hipEventRecord(c->evFork, st));
            HIPCK(c, hipStreamWaitEvent(c->aux, c->evFork, 0));
            TIMED(1, c->aux,
                  launchShadow(view, pbS, b - 1, c->spillAux, c->traceBlocks, c->refillAt, c->stackDepth, wwShadow,
                               c->fetchRun, c->shadowWide, denseS, c->aux));
            HIPCK(c, hipEventRecord(c->evJoin, c->aux));
        }
        TIMED(0, st,
              launchExtend(view, c->pb, b, c->spill, c->traceBlocks, c->refillAt, c->stackDepth, c->whileWhile,
                           c->speculate, c->twoLevel, c->fetchRun, cit.tailPaths, st));
        if (b > 0 && !conc)
            TIMED(1, st,
                  launchShadow(view, pbS, b - 1, c->spill, c->traceBlocks, c->refillAt, c->stackDepth, wwShadow,
                               c->fetchRun, c->shadowWide, denseS, st));
        if (b > 0 && conc) HIPCK(c, hipStreamWaitEvent(st, c->evJoin, 0));
        if (tail)
            TIMED(3, st,
                  launchFinish(view, c->pb, cit, b, c->spill, c->traceBlocks, c->stackDepth, c->whileWhile,
                               c->shadowWide, c->numCU, st));
        if (c->shadeSort)
            TIMED(2, st, launchShadeSorted(view, c->pb, cit, b, c->shadeBlocks, c->shadeWaves, c->shadeLightWaves,
                                           c->bins, st));
        else
            TIMED(2, st, launchShade(view, c->pb, cit, b, c->shadeBlocks, c->shadeWaves, st));
        HIPCK(c, launchQueueScan(c->pb, b, st));
    }
#undef TIMED
    return PRR_OK;
}

int resetChunk(prr_context *c, int maxDepth, hipStream_t st) {
    HIPCK(c, hipMemsetAsync(c->pb.qcnt, 0, ((size_t) maxDepth + 2) * sizeof(uint32_t), st));
    HIPCK(c, hipMemsetAsync(c->pb.sqcnt, 0, ((size_t) maxDepth + 2) * sizeof(uint32_t), st));
    HIPCK(c, hipMemsetAsync(c->pb.segc, 0, ((size_t) maxDepth + 2) * kQSegs * sizeof(unsigned long long), st));
    HIPCK(c, hipMemsetAsync(c->pb.fetch, 0, 2 * ((size_t) maxDepth + 2) * sizeof(uint32_t), st));
    return PRR_OK;
}

int finishStats(prr_context *c, const Timing &tm, uint64_t photons, prr_stats *stats) {
    if (!stats) return PRR_OK;
    unsigned long long h[8];
    HIPCK(c, hipMemcpy(h, c->stats, sizeof(h), hipMemcpyDeviceToHost));
    float ms = 0.f;
    HIPCK(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
    stats->rays_closest = h[0];
    stats->rays_any = h[1];
    stats->samples = h[2];
    stats->photons_stored = photons;
    stats->ms_total = ms;
    stats->rays_closest_finish = h[4];
    stats->nan_samples = h[5];
    stats->rays_any_finish = h[6];
    stats->rays_any_skipped = h[7];
    stats->ms_extend = stats->ms_shadow = stats->ms_shade = stats->ms_finish = stats->ms_sort = 0.0;
    stats->launches_extend = stats->launches_shadow = stats->launches_finish = 0;
    for (const auto &t : tm.timed) {
        float dt = 0.f;
        HIPCK(c, hipEventElapsedTime(&dt, c->evPool[t.second], c->evPool[t.second + 1]));
        if (t.first == 0) {
            stats->ms_extend += dt;
            stats->launches_extend++;
        } else if (t.first == 1) {
            stats->ms_shadow += dt;
            stats->launches_shadow++;
        } else if (t.first == 2) {
            stats->ms_shade += dt;
        } else if (t.first == 4) {
            stats->ms_sort += dt;
        } else {
            stats->ms_finish += dt;
            stats->launches_finish++;
        }
    }
    return PRR_OK;
}

// Framebuffer of the render: running sums, the adaptive sampler's Welford
// state and its active-pixel list, in one allocation sized for `pixels`.
int ensureFrame(prr_context *c, int64_t pixels, AdaptiveBufs &ab) {
    if (c->fbPixels < pixels) {
        HIPCK(c, hipStreamSynchronize(c->stream));
        if (c->fb) HIPCK(c, hipFree(c->fb));
        c->fb = nullptr;
        c->fbPixels = 0;
        const size_t bytes = (size_t) pixels * (3 * sizeof(float4) + sizeof(int32_t)) + 256;
        void *q = nullptr;
        const hipError_t e = hipMalloc(&q, bytes);
        if (e != hipSuccess) return fail(c, PRR_ERR_NOMEM, std::string("framebuffer: ") + hipGetErrorString(e));
        c->fb = (float4 *) q;
        c->fbPixels = pixels;
    }
    ab.mean = c->fb + c->fbPixels;
    ab.m2 = ab.mean + c->fbPixels;
    ab.list = (int32_t *) (ab.m2 + c->fbPixels);
    ab.count = (uint32_t *) (ab.list + c->fbPixels);
    return PRR_OK;
}

// Paths per adaptive pass: enough to keep every CU busy through the deep
// bounces; the samples a pixel takes past its stop point are discarded.
static const int64_t kAdaptivePassSlots = (int64_t) 1 << 22;

}  // namespace prr_host

using namespace prr_host;

extern "C" {

const char *prr_last_error(const prr_context *ctx) { return ctx ? ctx->err.c_str() : g_lastError.c_str(); }

int prr_create(int device, prr_context **out) {
    if (!out) return fail(nullptr, PRR_ERR_INVALID, "out is NULL");
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return fail(nullptr, PRR_ERR_DEVICE, "no HIP device");
    if (device < 0 || device >= n) return fail(nullptr, PRR_ERR_INVALID, "device index out of range");
    prr_context *c = new (std::nothrow) prr_context();
    if (!c) return fail(nullptr, PRR_ERR_NOMEM, "out of host memory");
    c->device = device;
    hipError_t e = hipSetDevice(device);
    hipDeviceProp_t prop;
    if (e == hipSuccess) e = hipGetDeviceProperties(&prop, device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->own, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->aux, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreate(&c->ev0);
    if (e == hipSuccess) e = hipEventCreate(&c->ev1);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&c->evFork, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&c->evJoin, hipEventDisableTiming);
    for (int b = 0; b < prr_context::kMaxBounces && e == hipSuccess; b++)
        e = hipEventCreateWithFlags(&c->evCnt[b], hipEventDisableTiming);
    if (e == hipSuccess)
        e = hipHostMalloc((void **) &c->cntHost, prr_context::kMaxBounces * sizeof(uint32_t), hipHostMallocDefault);
    if (e != hipSuccess) {
        std::string m = std::string("device init: ") + hipGetErrorString(e);
        delete c;
        return fail(nullptr, PRR_ERR_DEVICE, m);
    }
    if (e == hipSuccess && device < 256) {
        std::lock_guard<std::mutex> lk(g_noAllocMu);
        if (!g_noAllocSpill[device]) {
            uint2 *p = nullptr;
            e = noAllocScratch(&p);
            if (e == hipSuccess) g_noAllocSpill[device] = p;
        }
        c->noAllocSpill = g_noAllocSpill[device];
    }
    if (e != hipSuccess) {
        std::string m = std::string("device init: ") + hipGetErrorString(e);
        prr_destroy(c);
        return fail(nullptr, PRR_ERR_DEVICE, m);
    }
    c->stream = c->own;
    c->numCU = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 1;
    // persistent traversal grid: 8 blocks of 256 lanes per CU (with 8-entry
    // LDS stacks, 16 KB per block), a wave refills once 24 of its lanes are
    // idle, at most 8 interior steps (4 for shadow rays) per work check
    // (measured on the 1M-tri scene). PRR_TRACE_BLOCKS_PER_CU, PRR_STACK_LDS,
    // PRR_REFILL_AT, PRR_WHILE_WHILE(_ANY), PRR_FETCH_RUN, PRR_SHADOW_WIDE, PRR_SHADOW_FAST,
    // PRR_CONCURRENT_SHADOW, PRR_SHADE_WAVES, PRR_TAIL_PATHS and PRR_BAND_ROWS override the
    // measured defaults for tuning sweeps.
    int tbpc = 8;
    if (const char *ev = getenv("PRR_TRACE_BLOCKS_PER_CU")) tbpc = std::max(1, std::min(16, atoi(ev)));
    c->traceBlocks = c->numCU * tbpc;
    if (const char *ev = getenv("PRR_REFILL_AT")) c->refillAt = std::max(1, std::min(64, atoi(ev)));
    {
        size_t freeB = 0, totalB = 0;
        if (hipMemGetInfo(&freeB, &totalB) == hipSuccess && freeB > 0)
            c->maxSlots = std::max<int64_t>((int64_t) 1 << 20,
                                            std::min<int64_t>(kMaxSlots, (int64_t) (0.4 * (double) freeB) / kPathBytes));
        if (const char *ev = getenv("PRR_MAX_SLOTS"))
            c->maxSlots = std::max<int64_t>(1 << 16, std::min<int64_t>(kMaxSlots, atoll(ev)));
    }
    if (const char *ev = getenv("PRR_WHILE_WHILE")) c->whileWhile = std::max(0, atoi(ev));
    if (const char *ev = getenv("PRR_WHILE_WHILE_ANY")) c->whileWhileAny = std::max(0, atoi(ev));
    if (const char *ev = getenv("PRR_SHADOW_WIDE")) c->shadowWide = std::max(0, std::min(2, atoi(ev)));
    if (const char *ev = getenv("PRR_SHADOW_FAST")) c->shadowFast = atoi(ev) != 0;
    if (const char *ev = getenv("PRR_SPECULATE")) c->speculate = atoi(ev) != 0;
    if (const char *ev = getenv("PRR_TWO_LEVEL")) c->twoLevel = atoi(ev) != 0;
    if (const char *ev = getenv("PRR_CONCURRENT_SHADOW")) c->concurrentShadow = atoi(ev) != 0;
    if (const char *ev = getenv("PRR_CONCURRENT_MAX_PATHS")) c->concurrentMaxPaths = std::max<int64_t>(0, atoll(ev));
    if (const char *ev = getenv("PRR_FETCH_RUN")) c->fetchRun = std::max(64, std::min(1 << 16, atoi(ev)));
    if (const char *ev = getenv("PRR_SHADE_WAVES")) c->shadeWaves = std::max(1, std::min(5, atoi(ev)));
    if (const char *ev = getenv("PRR_SHADE_SORT")) c->shadeSort = atoi(ev) != 0;
    if (const char *ev = getenv("PRR_SHADE_LIGHT_WAVES")) {
        const int v = atoi(ev);
        c->shadeLightWaves = v >= 8 ? 8 : (v >= 6 ? 6 : 4);
    }
    if (const char *ev = getenv("PRR_TAIL_PATHS")) c->tailPaths = std::max<int64_t>(0, atoll(ev));
    if (const char *ev = getenv("PRR_BAND_ROWS")) c->bandRows = std::max(1, std::min(64, atoi(ev)));
    if (const char *ev = getenv("PRR_SHADOW_SORT")) c->shadowSortFrom = std::max(0, std::min(64, atoi(ev)));
    if (const char *ev = getenv("PRR_SHADOW_SORT_BITS")) c->shadowSortBits = std::max(1, std::min(3, atoi(ev)));
    if (const char *ev = getenv("PRR_STACK_LDS")) {  // LDS stack entries per lane: 2, 4, 8, 16 or 32
        const int v = atoi(ev);
        if (v == 2 || v == 4 || v == 8 || v == 16 || v == 32) c->stackDepth = v;
    }
    c->shadeBlocks = c->numCU * 8;
    c->spillLanes = c->traceBlocks * kTraceBlock;
    e = hipMalloc(&c->spill, (size_t) c->spillLanes * kSpill * sizeof(uint2));
    if (e == hipSuccess) e = hipMalloc(&c->spillAux, (size_t) c->spillLanes * kSpill * sizeof(uint2));
    if (e == hipSuccess) e = hipMalloc(&c->stats, 8 * sizeof(unsigned long long));
    if (e != hipSuccess) {
        std::string m = std::string("device alloc: ") + hipGetErrorString(e);
        prr_destroy(c);
        return fail(nullptr, PRR_ERR_NOMEM, m);
    }
    *out = c;
    return PRR_OK;
}

int prr_destroy(prr_context *c) {
    if (!c) return PRR_OK;
    (void) hipSetDevice(c->device);
    (void) hipStreamSynchronize(c->stream);
    if (c->aux) (void) hipStreamSynchronize(c->aux);
    freeRenderBufs(c);
    if (c->fb) (void) hipFree(c->fb);
    if (c->spill) (void) hipFree(c->spill);
    if (c->spillAux) (void) hipFree(c->spillAux);
    if (c->stats) (void) hipFree(c->stats);
    if (c->ev0) (void) hipEventDestroy(c->ev0);
    if (c->ev1) (void) hipEventDestroy(c->ev1);
    for (hipEvent_t e : c->evPool) (void) hipEventDestroy(e);
    if (c->evFork) (void) hipEventDestroy(c->evFork);
    if (c->evJoin) (void) hipEventDestroy(c->evJoin);
    for (hipEvent_t e : c->evCnt)
        if (e) (void) hipEventDestroy(e);
    if (c->cntHost) (void) hipHostFree(c->cntHost);
    if (c->own) (void) hipStreamDestroy(c->own);
    if (c->aux) (void) hipStreamDestroy(c->aux);
    delete c;
    return PRR_OK;
}

int prr_set_stream(prr_context *c, void *stream) {
    if (!c) return fail(nullptr, PRR_ERR_INVALID, "context is NULL");
    c->stream = stream ? (hipStream_t) stream : c->own;
    return PRR_OK;
}

int prr_synchronize(prr_context *c) {
    if (!c) return fail(nullptr, PRR_ERR_INVALID, "context is NULL");
    HIPCK(c, hipStreamSynchronize(c->stream));
    return PRR_OK;
}

int prr_scene_create(prr_context *c, const prr_scene_desc *d, prr_scene **out) {
    if (!c || !out) return fail(c, PRR_ERR_INVALID, "NULL argument");
    *out = nullptr;
    prr_scene *s = new (std::nothrow) prr_scene();
    if (!s) return fail(c, PRR_ERR_NOMEM, "out of host memory");
    std::string err;
    if (!host::buildHostScene(d, s->hs, err)) {
        delete s;
        return fail(c, PRR_ERR_INVALID, err);
    }
    HIPCK(c, hipSetDevice(c->device));
    const host::HostScene &h = s->hs;
    SceneView &v = s->view;
    hipError_t e = hipSuccess;
#define UP(dst, src)                                                               \
    if (e == hipSuccess) e = upload(s, (std::remove_const<std::remove_reference<decltype(*dst)>::type>::type **) &dst, src, c->stream)
    UP(v.nodes, h.nodes);
    if (c->twoLevel && !h.nodes2.empty()) UP(v.nodes2, h.nodes2);  // read only with PRR_TWO_LEVEL
    UP(v.nodes4, h.nodes4);
    UP(v.nodes4q, h.nodes4q);
    UP(v.leavesq, h.leavesq);
    UP(v.tris, h.tris);
    UP(v.sph, h.sph);
    UP(v.sph_id, h.sphId);
    UP(v.prim_n, h.primN);
    UP(v.sph_by_id, h.sphById);
    UP(v.mats, h.mats);
    UP(v.emitters, h.emitters);
    UP(v.lnodes, h.lnodes);
    UP(v.ltri, h.ltri);
    UP(v.env_px, h.envPx);
    UP(v.env_cdf, h.envCdf);
    UP(v.env_marg, h.envMarg);
    UP(v.env_gm, h.envGuideM);
    UP(v.env_gc, h.envGuideC);
    if (!h.tattr.empty()) UP(v.tattr, h.tattr);
    if (!h.matTex.empty()) {
        UP(v.mat_tex, h.matTex);
        UP(v.tex_info, h.texInfo);
        UP(v.texels, h.texels);
    }
#undef UP
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) {
        for (void *p : s->allocs) (void) hipFree(p);
        delete s;
        return fail(c, PRR_ERR_NOMEM, std::string("scene upload: ") + hipGetErrorString(e));
    }
    v.root = h.root;
    v.root4 = h.root4;
    v.root4q = h.root4q;
    v.num_tris = h.numTris;
    v.num_emitters = h.totalEmitArea > 0.f ? (int) h.emitters.size() : 0;
    v.total_emit_area = h.totalEmitArea;
    for (int k = 0; k < 6; k++) v.root_box[k] = h.rootBox[k];
    {
        float R = 0.f;  // the extent the quantiser's margin was sized for
        for (int a = 0; a < 3; a++) R = std::max(R, h.rootBox[2 * a + 1] - h.rootBox[2 * a]);
        for (int a = 0; a < 3; a++) {
            v.safe_box[2 * a] = h.rootBox[2 * a] - R;
            v.safe_box[2 * a + 1] = h.rootBox[2 * a + 1] + R;
        }
        v.quant_fast = c->shadowFast ? 1 : 0;
    }
    v.env_w = h.envW;
    v.env_h = h.envH;
    v.env_total = h.envTotal;
    v.env_gk = h.envGuideK;
    for (int k = 0; k < 3; k++) {
        v.background[k] = h.background[k];
        v.point_pos[k] = h.pointPos[k];
        v.point_int[k] = h.pointInt[k];
    }
    v.num_points = h.numPoints;
    v.cam = h.cam;
    if (!v.nodes2) std::vector<DNode2>().swap(s->hs.nodes2);  // host copy not kept when not uploaded
    s->agx = agxConst();
    *out = s;
    return PRR_OK;
}

int prr_scene_free(prr_context *c, prr_scene *s) {
    if (!s) return PRR_OK;
    if (c) {
        (void) hipSetDevice(c->device);
        (void) hipStreamSynchronize(c->stream);
    }
    freePhotons(s);
    for (void *p : s->allocs) (void) hipFree(p);
    delete s;
    return PRR_OK;
}

int prr_scene_get_info(const prr_scene *s, prr_scene_info *o) {
    if (!s || !o) return fail(nullptr, PRR_ERR_INVALID, "NULL argument");
    const host::HostScene &h = s->hs;
    o->num_nodes = (int32_t) h.nodes.size();
    o->num_leaf_prims = (int32_t) (h.tris.size() / 3);
    o->max_depth = h.maxDepth;
    o->num_emitters = h.totalEmitArea > 0.f ? (int32_t) h.emitters.size() : 0;
    o->total_emit_area = h.totalEmitArea;
    for (int a = 0; a < 3; a++) {
        o->bounds[a] = h.rootBox[2 * a];
        o->bounds[3 + a] = h.rootBox[2 * a + 1];
    }
    o->device_bytes = s->bytes;
    o->num_wide_nodes = (int32_t) h.nodes4.size();
    o->wide_levels_kept = h.wideLevelsKept;
    o->two_level = s->view.nodes2 ? 1 : 0;
    o->reserved = 0;
    return PRR_OK;
}

int prr_scene_export_level(const prr_scene *s, int32_t level, int32_t cap, float *b6, uint32_t *fc, uint32_t *meta,
                           int32_t *count) {
    if (!s || !count) return fail(nullptr, PRR_ERR_INVALID, "NULL argument");
    const host::RefBVH *b = nullptr;
    if (level < 0)
        b = &s->hs.top;
    else if (level < (int32_t) s->hs.meshes.size() && !s->hs.meshes[level].nodes.empty())
        b = &s->hs.meshes[level];
    if (!b) return fail(nullptr, PRR_ERR_INVALID, "no BVH at this level");
    const int n = (int) b->nodes.size();
    *count = n;
    for (int i = 0; i < n && i < cap; i++) {
        if (b6)
            for (int k = 0; k < 6; k++) b6[6 * i + k] = b->nodes[i].b[k];
        if (fc) fc[i] = b->nodes[i].first;
        if (meta) meta[i] = b->nodes[i].meta;
    }
    return PRR_OK;
}

int prr_trace_closest(prr_context *c, const prr_scene *s, int64_t n, const float *org, const float *dir,
                      int32_t *prim, float *t, float *u, float *v) {
    if (!c || !s) return fail(c, PRR_ERR_INVALID, "NULL context or scene");
    if (n < 0 || (n > 0 && (!org || !dir || !prim || !t))) return fail(c, PRR_ERR_INVALID, "bad arrays");
    HIPCK(c, hipSetDevice(c->device));
    HIPCK(c, launchTraceClosest(s->view, n, org, dir, prim, t, u, v, c->spill, c->spillLanes, c->stackDepth, c->stream));
    return PRR_OK;
}

int prr_trace_closest_noalloc(prr_context *c, const prr_scene *s, int64_t n, const float *org, const float *dir,
                              int32_t *prim, float *t, float *u, float *v) {
    if (!c || !s) return fail(c, PRR_ERR_INVALID, "NULL context or scene");
    if (n < 0 || n > PRR_NOALLOC_MAX_RAYS)
        return fail(c, PRR_ERR_INVALID, "prr_trace_closest_noalloc: n outside [0, PRR_NOALLOC_MAX_RAYS]");
    if (n > 0 && (!org || !dir || !prim || !t)) return fail(c, PRR_ERR_INVALID, "bad arrays");
    if (!c->noAllocSpill) return fail(c, PRR_ERR_DEVICE, "no static spill array on this device");
    HIPCK(c, hipSetDevice(c->device));
    HIPCK(c, launchTraceClosestNoAlloc(s->view, n, org, dir, prim, t, u, v, c->noAllocSpill, c->stackDepth,
                                       c->stream));
    return PRR_OK;
}

int prr_trace_any(prr_context *c, const prr_scene *s, int64_t n, const float *org, const float *dir,
                  const float *tmax, uint8_t *hit) {
    if (!c || !s) return fail(c, PRR_ERR_INVALID, "NULL context or scene");
    if (n < 0 || (n > 0 && (!org || !dir || !tmax || !hit))) return fail(c, PRR_ERR_INVALID, "bad arrays");
    HIPCK(c, hipSetDevice(c->device));
    HIPCK(c, launchTraceAny(s->view, n, org, dir, tmax, hit, c->spill, c->spillLanes, c->stackDepth,
                            c->shadowWide == 2, c->stream));
    return PRR_OK;
}

int prr_bsdf_eval(prr_context *c, int64_t n, const float *wi, const float *wo, const float *nrm, const float *params,
                  float *f, float *pdf) {
    if (!c) return fail(nullptr, PRR_ERR_INVALID, "context is NULL");
    if (n < 0 || (n > 0 && (!wi || !wo || !nrm || !params || !f || !pdf))) return fail(c, PRR_ERR_INVALID, "bad arrays");
    HIPCK(c, hipSetDevice(c->device));
    HIPCK(c, launchBsdfEval(n, wi, wo, nrm, params, f, pdf, c->stream));
    return PRR_OK;
}

int prr_medium_intersect(prr_context *c, int64_t n, const float *org, const float *dir, const float *bounds,
                         const int32_t *bounded, float *t_enter, float *t_exit, uint8_t *hit) {
    if (!c) return fail(nullptr, PRR_ERR_INVALID, "context is NULL");
    if (n < 0 || (n > 0 && (!org || !dir || !bounds || !bounded || !t_enter || !t_exit || !hit)))
        return fail(c, PRR_ERR_INVALID, "bad arrays");
    HIPCK(c, hipSetDevice(c->device));
    HIPCK(c, launchMedium(n, org, dir, bounds, bounded, t_enter, t_exit, hit, c->stream));
    return PRR_OK;
}

int prr_phase_hg(prr_context *c, int64_t n, const float *cos_theta, const float *g, float *out) {
    if (!c) return fail(nullptr, PRR_ERR_INVALID, "context is NULL");
    if (n < 0 || (n > 0 && (!cos_theta || !g || !out))) return fail(c, PRR_ERR_INVALID, "bad arrays");
    HIPCK(c, hipSetDevice(c->device));
    HIPCK(c, launchPhaseHG(n, cos_theta, g, out, c->stream));
    return PRR_OK;
}

int prr_phase_hg_sample(prr_context *c, int64_t n, const float *wo, const float *g, const float *u2, float *wi) {
    if (!c) return fail(nullptr, PRR_ERR_INVALID, "context is NULL");
    if (n < 0 || (n > 0 && (!wo || !g || !u2 || !wi))) return fail(c, PRR_ERR_INVALID, "bad arrays");
    HIPCK(c, hipSetDevice(c->device));
    HIPCK(c, launchPhaseHGSample(n, wo, g, u2, wi, c->stream));
    return PRR_OK;
}

int prr_bsdf_sample(prr_context *c, int64_t n, const float *wo, const float *nrm, const float *params, const float *u,
                    float *wi, float *f, float *pdf, int32_t *lobe) {
    if (!c) return fail(nullptr, PRR_ERR_INVALID, "context is NULL");
    if (n < 0 || (n > 0 && (!wo || !nrm || !params || !u || !wi || !f || !pdf || !lobe)))
        return fail(c, PRR_ERR_INVALID, "bad arrays");
    HIPCK(c, hipSetDevice(c->device));
    HIPCK(c, launchBsdfSample(n, wo, nrm, params, u, wi, f, pdf, lobe, c->stream));
    return PRR_OK;
}

int prr_sample_light(prr_context *c, const prr_scene *s, int64_t n, const float *u4, float *pos, float *nrm,
                     int32_t *emitter, float *pdf) {
    if (!c || !s) return fail(c, PRR_ERR_INVALID, "NULL context or scene");
    if (s->view.num_emitters < 1 || !(s->view.total_emit_area > 0.f))
        return fail(c, PRR_ERR_UNSUPPORTED, "scene has no emitter (Scene::emissives is empty)");
    if (n < 0 || (n > 0 && (!u4 || !pos || !nrm))) return fail(c, PRR_ERR_INVALID, "bad arrays");
    HIPCK(c, hipSetDevice(c->device));
    HIPCK(c, launchSampleLight(s->view, n, u4, pos, nrm, emitter, pdf, c->stream));
    return PRR_OK;
}

int prr_env_sample(prr_context *c, const prr_scene *s, int64_t n, const float *u2, float *dir, float *pdf) {
    if (!c || !s) return fail(c, PRR_ERR_INVALID, "NULL context or scene");
    if (s->view.env_w <= 0) return fail(c, PRR_ERR_UNSUPPORTED, "scene has no environment map");
    if (n < 0 || (n > 0 && (!u2 || !dir || !pdf))) return fail(c, PRR_ERR_INVALID, "bad arrays");
    HIPCK(c, hipSetDevice(c->device));
    HIPCK(c, launchEnvSample(s->view, n, u2, dir, pdf, c->stream));
    return PRR_OK;
}

int prr_env_eval(prr_context *c, const prr_scene *s, int64_t n, const float *dir, float *rad, float *pdf) {
    if (!c || !s) return fail(c, PRR_ERR_INVALID, "NULL context or scene");
    if (s->view.env_w <= 0) return fail(c, PRR_ERR_UNSUPPORTED, "scene has no environment map");
    if (n < 0 || (n > 0 && (!dir || !rad || !pdf))) return fail(c, PRR_ERR_INVALID, "bad arrays");
    HIPCK(c, hipSetDevice(c->device));
    HIPCK(c, launchEnvEval(s->view, n, dir, rad, pdf, c->stream));
    return PRR_OK;
}

int prr_camera_rays(prr_context *c, const prr_scene *s, int32_t W, int32_t H, int64_t n, const int32_t *pix,
                    const float *offs, float *org, float *dir) {
    if (!c || !s) return fail(c, PRR_ERR_INVALID, "NULL context or scene");
    if (W <= 0 || H <= 0 || n < 0 || (n > 0 && (!pix || !offs || !org || !dir)))
        return fail(c, PRR_ERR_INVALID, "bad arguments");
    HIPCK(c, hipSetDevice(c->device));
    HIPCK(c, launchCameraRays(s->view, W, H, n, pix, offs, org, dir, c->stream));
    return PRR_OK;
}

int prr_tonemap(prr_context *c, int32_t op, int64_t n, float exposure, const float *lin, float *disp, uint8_t *rgb8) {
    if (!c) return fail(nullptr, PRR_ERR_INVALID, "context is NULL");
    if (op < PRR_TONEMAP_RAW || op > PRR_TONEMAP_ACES) return fail(c, PRR_ERR_INVALID, "unknown tone-map operator");
    if (n < 0 || (n > 0 && !lin)) return fail(c, PRR_ERR_INVALID, "bad arrays");
    HIPCK(c, hipSetDevice(c->device));
    HIPCK(c, launchTonemap(op, n, exposure, lin, disp, rgb8, agxConst(), c->stream));
    return PRR_OK;
}

// ---------------------------------------------------------------------------
// caustic photon map (Scene::buildPhotonMaps, Scene.cpp:8-13; Photon.cpp)
// ---------------------------------------------------------------------------
int prr_photon_build(prr_context *c, prr_scene *s, int64_t num, float radius, int32_t maxDepth, uint64_t seed,
                     int64_t *stored) {
    if (!c || !s) return fail(c, PRR_ERR_INVALID, "NULL context or scene");
    if (num <= 0 || num > ((int64_t) 1 << 31) || !(radius > 0.f) || maxDepth < 0 || maxDepth > 64)
        return fail(c, PRR_ERR_INVALID, "bad photon parameters");
    if (s->view.num_points < 1) return fail(c, PRR_ERR_UNSUPPORTED, "photon emission needs a point light");
    HIPCK(c, hipSetDevice(c->device));
    hipStream_t st = c->stream;
    HIPCK(c, hipStreamSynchronize(st));
    freePhotons(s);
    PhotonMap &pm = s->ph;
    auto A = [&](size_t bytes) -> void * {
        void *p = nullptr;
        if (hipMalloc(&p, std::max<size_t>(bytes, 16)) != hipSuccess) return nullptr;
        pm.allocs.push_back(p);
        return p;
    };
    // per-photon records (scratch, freed after the build)
    std::vector<void *> tmp;
    auto T = [&](size_t bytes) -> void * {
        void *p = nullptr;
        if (hipMalloc(&p, std::max<size_t>(bytes, 16)) != hipSuccess) return nullptr;
        tmp.push_back(p);
        return p;
    };
    auto freeTmp = [&]() {
        for (void *p : tmp) (void) hipFree(p);
    };
    float4 *pos = (float4 *) T(num * sizeof(float4)), *dir = (float4 *) T(num * sizeof(float4));
    float4 *pw = (float4 *) T(num * sizeof(float4));
    uint32_t *cnt = (uint32_t *) T(sizeof(uint32_t));
    if (!pos || !dir || !pw || !cnt) {
        freeTmp();
        return fail(c, PRR_ERR_NOMEM, "photon buffers");
    }
    HIPCK(c, hipMemsetAsync(cnt, 0, sizeof(uint32_t), st));
    SceneView v = s->view;
    v.ph_count = 0;
    hipError_t e = launchPhotonTrace(v, num, seed, maxDepth, pos, dir, pw, cnt, c->spill, c->spillLanes,
                                     c->stackDepth, st);
    uint32_t m = 0;
    if (e == hipSuccess) e = hipMemcpyAsync(&m, cnt, sizeof(uint32_t), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) {
        freeTmp();
        return fail(c, PRR_ERR_DEVICE, std::string("photon trace: ") + hipGetErrorString(e));
    }
    uint32_t buckets = 16;
    while (buckets < 2u * std::max<uint32_t>(m, 1u) && buckets < (1u << 30)) buckets <<= 1;
    const float cell = 2.f * radius;
    uint32_t *kin = (uint32_t *) T(num * 4), *vin = (uint32_t *) T(num * 4);
    uint32_t *kout = (uint32_t *) T(num * 4), *vout = (uint32_t *) T(num * 4);
    float4 *spos = (float4 *) A((size_t) m * sizeof(float4)), *sdir = (float4 *) A((size_t) m * sizeof(float4));
    float4 *spw = (float4 *) A((size_t) m * sizeof(float4));
    uint32_t *bs = (uint32_t *) A((size_t) buckets * 4), *be = (uint32_t *) A((size_t) buckets * 4);
    if (!kin || !vin || !kout || !vout || !spos || !sdir || !spw || !bs || !be) {
        freeTmp();
        freePhotons(s);
        return fail(c, PRR_ERR_NOMEM, "photon grid buffers");
    }
    e = launchPhotonKeys(num, pos, cell, buckets - 1u, kin, vin, st);
    size_t tb = 0;
    if (e == hipSuccess) e = photonSort(nullptr, tb, kin, kout, vin, vout, (int) num, st);
    void *temp = (e == hipSuccess) ? T(tb) : nullptr;
    if (e == hipSuccess && !temp) e = hipErrorOutOfMemory;
    if (e == hipSuccess) e = photonSort(temp, tb, kin, kout, vin, vout, (int) num, st);
    if (e == hipSuccess) e = hipMemsetAsync(bs, 0, (size_t) buckets * 4, st);
    if (e == hipSuccess) e = hipMemsetAsync(be, 0, (size_t) buckets * 4, st);
    if (e == hipSuccess) e = launchPhotonFinish(m, kout, vout, pos, dir, pw, spos, sdir, spw, bs, be, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    freeTmp();
    if (e != hipSuccess) {
        freePhotons(s);
        return fail(c, PRR_ERR_DEVICE, std::string("photon grid: ") + hipGetErrorString(e));
    }
    pm.emitted = num;
    pm.stored = m;
    pm.radius = radius;
    pm.maxDepth = maxDepth;
    pm.seed = seed;
    s->view.ph_pos = spos;
    s->view.ph_dir = sdir;
    s->view.ph_pow = spw;
    s->view.ph_start = bs;
    s->view.ph_end = be;
    s->view.ph_count = m;
    s->view.ph_mask = buckets - 1u;
    s->view.ph_radius = radius;
    s->view.ph_cell = cell;
    if (stored) *stored = m;
    return PRR_OK;
}

// PhotonMap::emit + PhotonMap::build (Photon.cpp:156-190) as Scene::buildPhotonMaps
// calls them (Scene.cpp:8-13), with the reference's own rules.
int prr_photon_emit(prr_context *c, prr_scene *s, int64_t num, float radius, int32_t maxDepth, uint64_t seed,
                    int64_t *stored) {
    if (!c || !s) return fail(c, PRR_ERR_INVALID, "NULL context or scene");
    if (num <= 0 || num > ((int64_t) 1 << 31) || !(radius > 0.f) || maxDepth < 0 || maxDepth > 64)
        return fail(c, PRR_ERR_INVALID, "bad photon parameters");
    if (!(s->view.total_emit_area > 0.f) || s->view.num_emitters < 1)
        return fail(c, PRR_ERR_UNSUPPORTED, "photon emission needs an emissive object (Scene::emissives is empty)");
    HIPCK(c, hipSetDevice(c->device));
    hipStream_t st = c->stream;
    HIPCK(c, hipStreamSynchronize(st));
    freePhotons(s);
    PhotonMap &pm = s->ph;
    std::vector<void *> tmp;
    auto T = [&](size_t bytes) -> void * {
        void *p = nullptr;
        if (hipMalloc(&p, std::max<size_t>(bytes, 16)) != hipSuccess) return nullptr;
        tmp.push_back(p);
        return p;
    };
    auto A = [&](size_t bytes) -> void * {
        void *p = nullptr;
        if (hipMalloc(&p, std::max<size_t>(bytes, 16)) != hipSuccess) return nullptr;
        pm.allocs.push_back(p);
        return p;
    };
    auto freeTmp = [&]() {
        for (void *p : tmp) (void) hipFree(p);
        tmp.clear();
    };
    SceneView v = s->view;
    v.ph_count = 0;
    // records: a photon stores at most once per diffuse hit; start at two per
    // photon and trace again with the exact capacity if that overflows (the
    // trace is deterministic)
    uint32_t cap = (uint32_t) std::min<int64_t>(std::max<int64_t>(2 * num, 1024), 0x7fffffff), m = 0;
    float4 *pos = nullptr, *dir = nullptr, *pw = nullptr;
    uint64_t *key = nullptr;
    uint32_t *cnt = (uint32_t *) T(sizeof(uint32_t));
    for (int pass = 0; pass < 2; pass++) {
        pos = (float4 *) T((size_t) cap * sizeof(float4));
        dir = (float4 *) T((size_t) cap * sizeof(float4));
        pw = (float4 *) T((size_t) cap * sizeof(float4));
        key = (uint64_t *) T((size_t) cap * sizeof(uint64_t));
        if (!cnt || !pos || !dir || !pw || !key) {
            freeTmp();
            return fail(c, PRR_ERR_NOMEM, "photon buffers");
        }
        hipError_t e = hipMemsetAsync(cnt, 0, sizeof(uint32_t), st);
        if (e == hipSuccess)
            e = launchPhotonTraceRef(v, num, seed, maxDepth, pos, dir, pw, key, cap, cnt, c->spill, c->spillLanes,
                                     c->stackDepth, st);
        if (e == hipSuccess) e = hipMemcpyAsync(&m, cnt, sizeof(uint32_t), hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        if (e != hipSuccess) {
            freeTmp();
            return fail(c, PRR_ERR_DEVICE, std::string("photon trace: ") + hipGetErrorString(e));
        }
        if (m <= cap) break;
        if (pass == 1 || m > 0x7fffffffu) {
            freeTmp();
            return fail(c, PRR_ERR_NOMEM, "photon records exceed the capacity");
        }
        cap = m;
    }
    uint32_t buckets = 16;
    while (buckets < 2u * std::max<uint32_t>(m, 1u) && buckets < (1u << 30)) buckets <<= 1;
    const float cell = 2.f * radius;
    // 1. records in (photon, depth) order
    uint32_t *v0 = (uint32_t *) T((size_t) m * 4), *v1 = (uint32_t *) T((size_t) m * 4);
    uint64_t *ksorted = (uint64_t *) T((size_t) m * 8);
    float4 *rpos = (float4 *) T((size_t) m * sizeof(float4)), *rdir = (float4 *) T((size_t) m * sizeof(float4));
    float4 *rpw = (float4 *) T((size_t) m * sizeof(float4));
    uint64_t *rkey = (uint64_t *) T((size_t) m * 8);
    uint32_t *bk = (uint32_t *) T((size_t) m * 4), *bkOut = (uint32_t *) T((size_t) m * 4);
    // 2. the grid, in bucket order
    float4 *spos = (float4 *) A((size_t) m * sizeof(float4)), *sdir = (float4 *) A((size_t) m * sizeof(float4));
    float4 *spw = (float4 *) A((size_t) m * sizeof(float4));
    uint64_t *skey = (uint64_t *) A((size_t) m * 8);
    uint32_t *bs = (uint32_t *) A((size_t) buckets * 4), *be = (uint32_t *) A((size_t) buckets * 4);
    if (!v0 || !v1 || !ksorted || !rpos || !rdir || !rpw || !rkey || !bk || !bkOut || !spos || !sdir || !spw || !skey ||
        !bs || !be) {
        freeTmp();
        freePhotons(s);
        return fail(c, PRR_ERR_NOMEM, "photon grid buffers");
    }
    hipError_t e = launchPhotonIota(m, v0, st);
    size_t tb = 0, tb2 = 0;
    if (e == hipSuccess) e = photonSortByKey(nullptr, tb, key, ksorted, v0, v1, (int) m, st);
    if (e == hipSuccess) e = photonSort(nullptr, tb2, bk, bkOut, v0, v1, (int) m, st);
    void *temp = (e == hipSuccess) ? T(std::max(tb, tb2)) : nullptr;
    if (e == hipSuccess && !temp) e = hipErrorOutOfMemory;
    if (e == hipSuccess) e = photonSortByKey(temp, tb, key, ksorted, v0, v1, (int) m, st);
    if (e == hipSuccess) e = launchPhotonGather(m, v1, pos, dir, pw, key, rpos, rdir, rpw, rkey, st);
    if (e == hipSuccess) e = launchPhotonKeys(m, rpos, cell, buckets - 1u, bk, v0, st);
    if (e == hipSuccess) e = photonSort(temp, tb2, bk, bkOut, v0, v1, (int) m, st);
    if (e == hipSuccess) e = launchPhotonGather(m, v1, rpos, rdir, rpw, rkey, spos, sdir, spw, skey, st);
    if (e == hipSuccess) e = hipMemsetAsync(bs, 0, (size_t) buckets * 4, st);
    if (e == hipSuccess) e = hipMemsetAsync(be, 0, (size_t) buckets * 4, st);
    if (e == hipSuccess) e = launchPhotonRanges(m, bkOut, bs, be, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    freeTmp();
    if (e != hipSuccess) {
        freePhotons(s);
        return fail(c, PRR_ERR_DEVICE, std::string("photon grid: ") + hipGetErrorString(e));
    }
    pm.emitted = num;
    pm.stored = m;
    pm.radius = radius;
    pm.maxDepth = maxDepth;
    pm.seed = seed;
    pm.rules = 1;
    pm.keys = skey;
    s->view.ph_pos = spos;
    s->view.ph_dir = sdir;
    s->view.ph_pow = spw;
    s->view.ph_start = bs;
    s->view.ph_end = be;
    s->view.ph_count = m;
    s->view.ph_mask = buckets - 1u;
    s->view.ph_radius = radius;
    s->view.ph_cell = cell;
    if (stored) *stored = m;
    return PRR_OK;
}

int prr_photon_export_owners(prr_context *c, const prr_scene *s, int64_t cap, int64_t *owner, int32_t *depth,
                             int64_t *count) {
    if (!c || !s || !count) return fail(c, PRR_ERR_INVALID, "NULL argument");
    if (s->ph.rules != 1 || !s->ph.keys)
        return fail(c, PRR_ERR_UNSUPPORTED, "no photon map of the reference's rules (prr_photon_emit)");
    *count = s->ph.stored;
    const int64_t k = std::min<int64_t>(cap, s->ph.stored);
    if (k <= 0) return PRR_OK;
    if (!owner || !depth) return fail(c, PRR_ERR_INVALID, "NULL output");
    std::vector<uint64_t> h((size_t) k);
    HIPCK(c, hipSetDevice(c->device));
    HIPCK(c, hipMemcpyAsync(h.data(), s->ph.keys, (size_t) k * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCK(c, hipStreamSynchronize(c->stream));
    for (int64_t i = 0; i < k; i++) {
        owner[i] = (int64_t) (h[i] >> 7);
        depth[i] = (int32_t) (h[i] & 127u);
    }
    return PRR_OK;
}

int prr_photon_irradiance(prr_context *c, const prr_scene *s, int64_t n, const float *pos, const float *normal,
                          float *out) {
    (void) normal;  // unused, as in PhotonMap::estimateIrradiance
    if (!c || !s) return fail(c, PRR_ERR_INVALID, "NULL context or scene");
    if (s->ph.stored <= 0 && s->ph.emitted <= 0) return fail(c, PRR_ERR_UNSUPPORTED, "no photon map");
    if (n < 0 || (n > 0 && (!pos || !out))) return fail(c, PRR_ERR_INVALID, "bad arrays");
    HIPCK(c, hipSetDevice(c->device));
    HIPCK(c, launchPhotonIrradiance(s->view, n, pos, out, c->stream));
    return PRR_OK;
}

int prr_photon_export(prr_context *c, const prr_scene *s, int64_t cap, float *out9, int64_t *count) {
    if (!c || !s || !count) return fail(c, PRR_ERR_INVALID, "NULL argument");
    *count = s->ph.stored;
    const int64_t k = std::min<int64_t>(cap, s->ph.stored);
    if (k <= 0) return PRR_OK;
    if (!out9) return fail(c, PRR_ERR_INVALID, "NULL output");
    HIPCK(c, hipSetDevice(c->device));
    // pos, dir, power as 9 packed floats per photon (bucket order)
    const size_t pitch = 9 * sizeof(float);
    HIPCK(c, hipMemcpy2DAsync(out9, pitch, s->view.ph_pos, sizeof(float4), 3 * sizeof(float), k,
                              hipMemcpyDeviceToDevice, c->stream));
    HIPCK(c, hipMemcpy2DAsync(out9 + 3, pitch, s->view.ph_dir, sizeof(float4), 3 * sizeof(float), k,
                              hipMemcpyDeviceToDevice, c->stream));
    HIPCK(c, hipMemcpy2DAsync(out9 + 6, pitch, s->view.ph_pow, sizeof(float4), 3 * sizeof(float), k,
                              hipMemcpyDeviceToDevice, c->stream));
    return PRR_OK;
}

int prr_photon_radiance(prr_context *c, const prr_scene *s, int64_t n, const float *pos, const float *wo,
                        const float *normal, const float *params, float *out) {
    if (!c || !s) return fail(c, PRR_ERR_INVALID, "NULL context or scene");
    if (s->ph.stored <= 0) return fail(c, PRR_ERR_UNSUPPORTED, "no photon map (prr_photon_build)");
    if (n < 0 || (n > 0 && (!pos || !wo || !normal || !params || !out))) return fail(c, PRR_ERR_INVALID, "bad arrays");
    HIPCK(c, hipSetDevice(c->device));
    HIPCK(c, launchPhotonRadiance(s->view, n, pos, wo, normal, params, out, c->stream));
    return PRR_OK;
}

int prr_render(prr_context *c, const prr_scene *s, const prr_render_params *p, float *rgb, int32_t *spp_out,
               prr_stats *stats) {
    if (!c || !s || !p || !rgb) return fail(c, PRR_ERR_INVALID, "NULL argument");
    if (p->width <= 0 || p->height <= 0 || p->spp <= 0 || p->max_depth <= 0 || p->max_depth > 64 ||
        p->sample_offset < 0 || p->min_spp < 0 ||
        (int64_t) p->sample_offset + p->spp > INT32_MAX)
        return fail(c, PRR_ERR_INVALID, "bad render parameters");
    const int64_t pixels = (int64_t) p->width * p->height;
    if (pixels > c->maxSlots) return fail(c, PRR_ERR_UNSUPPORTED, "image larger than the context's path slots");
    if (p->num_photons > 0 && s->ph.stored <= 0 && s->ph.emitted <= 0)
        return fail(c, PRR_ERR_INVALID, "num_photons > 0 but the scene has no photon map (prr_photon_build)");
    if (p->num_photons > 0 && s->ph.rules == 1)
        return fail(c, PRR_ERR_UNSUPPORTED,
                    "the reference's photon map (prr_photon_emit) is not gathered by castRay; use prr_photon_build");
    HIPCK(c, hipSetDevice(c->device));
    // Welford stopping can only end a pixel early when the threshold is
    // positive and min_spp < spp (Renderer.cpp:92-96); otherwise every pixel
    // takes spp samples.
    const bool adaptive = p->variance_threshold > 0.f && p->min_spp < p->spp;
    const int spc = (int) std::max<int64_t>(1, std::min<int64_t>(p->spp, c->maxSlots / pixels));
    int rc = ensureRenderBufs(c, adaptive ? std::max<int64_t>(pixels * spc, std::min(c->maxSlots, kAdaptivePassSlots))
                                          : pixels * spc,
                              p->max_depth);
    if (rc != PRR_OK) return rc;
    AdaptiveBufs ab;
    if ((rc = ensureFrame(c, pixels, ab)) != PRR_OK) return rc;
    hipStream_t st = c->stream;
    HIPCK(c, hipMemsetAsync(c->stats, 0, 8 * sizeof(unsigned long long), st));
    HIPCK(c, hipMemsetAsync(c->fb, 0, (size_t) pixels * sizeof(float4), st));
    if (adaptive) {
        HIPCK(c, hipMemsetAsync(ab.mean, 0, (size_t) pixels * sizeof(float4), st));
        HIPCK(c, hipMemsetAsync(ab.m2, 0, (size_t) pixels * sizeof(float4), st));
    }
    HIPCK(c, hipEventRecord(c->ev0, st));
    SceneView view = s->view;
    if (p->num_photons <= 0) view.ph_count = 0;  // photon gather only on request
    Timing tm;
    tm.on = (p->flags & PRR_RENDER_KERNEL_TIMING) != 0;
    ChunkInfo ci;
    ci.W = p->width;
    ci.H = p->height;
    ci.seed = p->seed;
    ci.maxDepth = p->max_depth;
    ci.clampInd = p->indirect_clamp;
    // Renders samples [s0, s0 + count) of the listed pixels (all when list is
    // null) in chunks of at most the context's path slots.
    auto renderSamples = [&](const int32_t *list, int64_t npix, int s0, int count) -> int {
        const int per = (int) std::max<int64_t>(1, std::min<int64_t>(count, c->pbSlots / npix));
        for (int k = 0; k < count; k += per) {
            ci.pixList = list;
            ci.bandRows = list ? 0 : c->bandRows;
            ci.pixels = (int) npix;
            ci.spc = std::min(per, count - k);
            ci.sampleBase = p->sample_offset + s0 + k;
            int r;
            if ((r = resetChunk(c, p->max_depth, st)) != PRR_OK) return r;
            HIPCK(c, launchRaygen(view, c->pb, ci, c->shadeBlocks, st));
            if ((r = runBounces(c, view, ci, tm, st)) != PRR_OK) return r;
            if (adaptive)
                HIPCK(c, launchAccumulateAdaptive(c->pb, ci, c->fb, ab, p->min_spp, p->variance_threshold, st));
            else
                HIPCK(c, launchAccumulate(c->pb, ci, c->fb, st));
        }
        return PRR_OK;
    };
    if (!adaptive) {
        if ((rc = renderSamples(nullptr, pixels, 0, p->spp)) != PRR_OK) return rc;
        HIPCK(c, launchResolve(c->fb, (int) pixels, p->spp, rgb, st));
        if (spp_out) HIPCK(c, hipMemsetD32Async((hipDeviceptr_t) spp_out, p->spp, (size_t) pixels, st));
    } else {
        // First pass: min_spp samples everywhere (no pixel can stop earlier).
        // Later passes: the pixels still sampling, several samples each when
        // few remain. Every active pixel has taken the same number of samples.
        int done = std::max(1, p->min_spp);
        if ((rc = renderSamples(nullptr, pixels, 0, done)) != PRR_OK) return rc;
        while (done < p->spp) {
            HIPCK(c, hipMemsetAsync(ab.count, 0, sizeof(uint32_t), st));
            HIPCK(c, launchCompactActive(ab, (int) pixels, p->spp, st));
            uint32_t active = 0;
            HIPCK(c, hipMemcpyAsync(&active, ab.count, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
            HIPCK(c, hipStreamSynchronize(st));
            if (active == 0) break;
            const int64_t want = (kAdaptivePassSlots + active - 1) / active;
            const int batch = (int) std::min<int64_t>(p->spp - done, std::max<int64_t>(1, want));
            if ((rc = renderSamples(ab.list, active, done, batch)) != PRR_OK) return rc;
            done += batch;
        }
        HIPCK(c, launchResolveAdaptive(c->fb, ab, (int) pixels, rgb, spp_out, st));
    }
    HIPCK(c, hipEventRecord(c->ev1, st));
    HIPCK(c, hipStreamSynchronize(st));
    return finishStats(c, tm, p->num_photons > 0 ? (uint64_t) s->ph.stored : 0, stats);
}

int prr_cast_rays(prr_context *c, const prr_scene *s, int64_t n, const float *org, const float *dir, uint64_t seed,
                  int32_t max_depth, float indirect_clamp, float *L, prr_stats *stats) {
    if (!c || !s) return fail(c, PRR_ERR_INVALID, "NULL context or scene");
    if (n < 0 || max_depth <= 0 || max_depth > 64 || (n > 0 && (!org || !dir || !L)))
        return fail(c, PRR_ERR_INVALID, "bad arguments");
    if (n == 0) return PRR_OK;
    HIPCK(c, hipSetDevice(c->device));
    const int64_t chunk = std::min<int64_t>(n, c->maxSlots);
    int rc = ensureRenderBufs(c, chunk, max_depth);
    if (rc != PRR_OK) return rc;
    hipStream_t st = c->stream;
    HIPCK(c, hipMemsetAsync(c->stats, 0, 8 * sizeof(unsigned long long), st));
    HIPCK(c, hipEventRecord(c->ev0, st));
    SceneView view = s->view;
    view.ph_count = 0;
    Timing tm;
    for (int64_t first = 0; first < n; first += chunk) {
        const int64_t m = std::min(chunk, n - first);
        ChunkInfo ci;
        ci.W = (int) m;
        ci.H = 1;
        ci.pixels = (int) m;
        ci.spc = 1;
        ci.sampleBase = 0;
        ci.seed = seed;
        ci.maxDepth = max_depth;
        ci.clampInd = indirect_clamp;
        if ((rc = resetChunk(c, max_depth, st)) != PRR_OK) return rc;
        HIPCK(c, launchRaygenRays(c->pb, m, org, dir, first, seed, c->shadeBlocks, st));
        if ((rc = runBounces(c, view, ci, tm, st)) != PRR_OK) return rc;
        HIPCK(c, launchCopyResults(c->pb, m, first, L, st));
    }
    HIPCK(c, hipEventRecord(c->ev1, st));
    HIPCK(c, hipStreamSynchronize(st));
    return finishStats(c, tm, 0, stats);
}

}  // extern "C"
