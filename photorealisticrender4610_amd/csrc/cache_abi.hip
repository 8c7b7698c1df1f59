// cache_abi.hip — C ABI of the learned radiance cache (prr_cache_*).
#include <algorithm>
#include <cmath>
#include <new>
#include <string>
#include <vector>

#include "context.hpp"
#include "device/cache.hpp"
#include "device/launch_cache.hpp"
#include "device/launch_render.hpp"

using namespace prr;
using namespace prr_host;

struct prr_cache {
    prr_cache_params p{};
    float *master = nullptr, *m1 = nullptr, *m2 = nullptr;
    _Float16 *wt = nullptr, *wn = nullptr;
    float *partial = nullptr;   // per training block: PRR_CACHE_WEIGHTS floats
    float *lossPart = nullptr;  // per training block
    float *scalar = nullptr;    // loss sum
    float *grad = nullptr;      // gradient readout
    int maxBlocks = 0;
    int64_t step = 0;           // Adam steps taken
    bool diverged = false;
    std::vector<void *> allocs;
    // frame buffers of prr_cache_render
    int64_t qCap = 0, tCap = 0;
    uint32_t *counts = nullptr;  // [0] queries, [1] training records
    uint32_t *qSlot = nullptr, *tSlot = nullptr;
    float *q = nullptr, *tq = nullptr, *tTarget = nullptr;
    float4 *qBeta = nullptr, *tBeta = nullptr, *tLpre = nullptr;
    std::vector<void *> frameAllocs;
    hipEvent_t ev[3] = {nullptr, nullptr, nullptr};  // render end, inference start / end
    CacheView view() const {
        CacheView v;
        v.wt = wt;
        v.wn = wn;
        for (int a = 0; a < 3; a++) {
            v.bmin[a] = p.bounds[a];
            const float ext = p.bounds[3 + a] - p.bounds[a];
            v.inv_ext[a] = ext > 0.f ? 1.f / ext : 0.f;
        }
        return v;
    }
};

static_assert(PRR_CACHE_WEIGHTS == kCacheWeights, "weight count");
static_assert(PRR_CACHE_QUERY == kCacheQ, "query size");

static void freeFrameBufs(prr_cache *k) {
    for (void *q : k->frameAllocs) (void) hipFree(q);
    k->frameAllocs.clear();
    k->qCap = k->tCap = 0;
}

static void freeCache(prr_cache *k) {
    for (void *q : k->allocs) (void) hipFree(q);
    freeFrameBufs(k);
    for (hipEvent_t e : k->ev)
        if (e) (void) hipEventDestroy(e);
    delete k;
}

// He-uniform U(-sqrt(6 / fan_in), +sqrt(6 / fan_in)) per layer from
// SplitMix64 draws of the seed, uploaded as the master weights.
static int initWeights(prr_context *c, prr_cache *k) {
    std::vector<float> w(kCacheWeights);
    uint64_t st = k->p.seed;
    auto next = [&]() {
        uint64_t z = (st += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        return (float) (z >> 40) * (1.f / 16777216.f);
    };
    for (int l = 0; l < kCacheLayers; l++) {
        const float lim = std::sqrt(6.f / (float) cacheIn(l));
        for (int i = 0; i < cacheOut(l) * cacheIn(l); i++) w[cacheOff(l) + i] = (2.f * next() - 1.f) * lim;
    }
    HIPCK(c, hipMemcpyAsync(k->master, w.data(), w.size() * sizeof(float), hipMemcpyHostToDevice, c->stream));
    HIPCK(c, hipStreamSynchronize(c->stream));
    return PRR_OK;
}

// Re-packs the FP16 layouts from the master weights and zeroes Adam state.
static int resetState(prr_context *c, prr_cache *k) {
    HIPCK(c, hipMemsetAsync(k->m1, 0, kCacheWeights * sizeof(float), c->stream));
    HIPCK(c, hipMemsetAsync(k->m2, 0, kCacheWeights * sizeof(float), c->stream));
    CacheOpt o{};
    HIPCK(c, launchCacheAdam(k->master, k->m1, k->m2, nullptr, 0, o, k->wt, k->wn, nullptr, c->stream));
    k->step = 0;
    return PRR_OK;
}

// One training pass (forward + backward) over n samples into the block
// partials; returns the number of blocks.
static int trainPass(prr_context *c, prr_cache *k, int64_t n, const float *q, const float *t, int *blocks,
                     int64_t start = 0, int64_t total = -1, int64_t mult = 1) {
    *blocks = cacheTrainBlocks(n);
    if (*blocks > k->maxBlocks) return fail(c, PRR_ERR_INVALID, "batch larger than train_batch");
    HIPCK(c, launchCacheTrain(k->view(), n, q, t, k->partial, k->lossPart, k->p.grad_clip, start,
                              total > 0 ? total : n, mult, c->stream));
    return PRR_OK;
}

static CacheOpt makeOpt(const prr_cache *k, int64_t n) {
    CacheOpt o;
    o.lr = k->p.learning_rate;
    o.beta1 = k->p.beta1;
    o.beta2 = k->p.beta2;
    o.eps = k->p.epsilon;
    o.bias1 = 1.f - (float) std::pow((double) k->p.beta1, (double) (k->step + 1));
    o.bias2 = 1.f - (float) std::pow((double) k->p.beta2, (double) (k->step + 1));
    o.gscale = 1.f / (3.f * (float) n);
    o.clip = k->p.grad_clip;
    return o;
}

extern "C" {

int prr_cache_create(prr_context *c, const prr_cache_params *p, prr_cache **out) {
    if (!c || !p || !out) return fail(c, PRR_ERR_INVALID, "NULL argument");
    *out = nullptr;
    if (p->train_batch <= 0 || p->learning_rate <= 0.f || !(p->beta1 >= 0.f && p->beta1 < 1.f) ||
        !(p->beta2 >= 0.f && p->beta2 < 1.f) || p->epsilon <= 0.f || p->grad_clip <= 0.f)
        return fail(c, PRR_ERR_INVALID, "bad cache parameters");
    for (int a = 0; a < 3; a++)
        if (!(p->bounds[3 + a] >= p->bounds[a])) return fail(c, PRR_ERR_INVALID, "bad cache bounds");
    HIPCK(c, hipSetDevice(c->device));
    prr_cache *k = new (std::nothrow) prr_cache();
    if (!k) return fail(c, PRR_ERR_NOMEM, "out of host memory");
    k->p = *p;
    k->maxBlocks = cacheTrainBlocks(p->train_batch);
    hipError_t e = hipSuccess;
    auto A = [&](auto **ptr, size_t count) {
        using T = typename std::remove_reference<decltype(**ptr)>::type;
        if (e != hipSuccess) return;
        void *q = nullptr;
        e = hipMalloc(&q, std::max<size_t>(1, count) * sizeof(T));
        if (e == hipSuccess) {
            k->allocs.push_back(q);
            *ptr = (T *) q;
        }
    };
    A(&k->master, kCacheWeights);
    A(&k->m1, kCacheWeights);
    A(&k->m2, kCacheWeights);
    A(&k->wt, kCacheHalfT);
    A(&k->wn, kCacheHalfN);
    A(&k->partial, (size_t) k->maxBlocks * kCacheWeights);
    A(&k->lossPart, (size_t) k->maxBlocks);
    A(&k->scalar, 4);
    A(&k->grad, kCacheWeights);
    if (e != hipSuccess) {
        freeCache(k);
        return fail(c, PRR_ERR_NOMEM, std::string("cache buffers: ") + hipGetErrorString(e));
    }
    // padding of the FP16 layouts stays zero
    e = hipMemsetAsync(k->wt, 0, kCacheHalfT * sizeof(_Float16), c->stream);
    if (e == hipSuccess) e = hipMemsetAsync(k->wn, 0, kCacheHalfN * sizeof(_Float16), c->stream);
    for (int i = 0; i < 3 && e == hipSuccess; i++) e = hipEventCreate(&k->ev[i]);
    if (e != hipSuccess) {
        freeCache(k);
        return fail(c, PRR_ERR_DEVICE, std::string("cache init: ") + hipGetErrorString(e));
    }
    int rc = initWeights(c, k);
    if (rc == PRR_OK) rc = resetState(c, k);
    if (rc != PRR_OK) {
        freeCache(k);
        return rc;
    }
    *out = k;
    return PRR_OK;
}

int prr_cache_free(prr_context *c, prr_cache *k) {
    if (!k) return PRR_OK;
    if (c) {
        (void) hipSetDevice(c->device);
        (void) hipStreamSynchronize(c->stream);
    }
    freeCache(k);
    return PRR_OK;
}

int prr_cache_get_weights(prr_context *c, const prr_cache *k, float *w) {
    if (!c || !k || !w) return fail(c, PRR_ERR_INVALID, "NULL argument");
    HIPCK(c, hipSetDevice(c->device));
    HIPCK(c, hipMemcpyAsync(w, k->master, kCacheWeights * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    HIPCK(c, hipStreamSynchronize(c->stream));
    return PRR_OK;
}

int prr_cache_set_weights(prr_context *c, prr_cache *k, const float *w) {
    if (!c || !k || !w) return fail(c, PRR_ERR_INVALID, "NULL argument");
    HIPCK(c, hipSetDevice(c->device));
    HIPCK(c, hipMemcpyAsync(k->master, w, kCacheWeights * sizeof(float), hipMemcpyHostToDevice, c->stream));
    int rc = resetState(c, k);
    if (rc != PRR_OK) return rc;
    HIPCK(c, hipStreamSynchronize(c->stream));
    return PRR_OK;
}

int prr_cache_infer(prr_context *c, const prr_cache *k, int64_t n, const float *q, float *out) {
    if (!c || !k) return fail(c, PRR_ERR_INVALID, "NULL argument");
    if (n < 0 || (n > 0 && (!q || !out))) return fail(c, PRR_ERR_INVALID, "bad arrays");
    HIPCK(c, hipSetDevice(c->device));
    HIPCK(c, launchCacheInfer(k->view(), n, nullptr, q, out, nullptr, nullptr, nullptr, 0.f, c->numCU * 3, c->stream));
    return PRR_OK;
}

int prr_cache_gradient(prr_context *c, prr_cache *k, int64_t n, const float *q, const float *t, float *grad,
                       float *loss) {
    if (!c || !k) return fail(c, PRR_ERR_INVALID, "NULL argument");
    if (n <= 0 || !q || !t) return fail(c, PRR_ERR_INVALID, "bad arrays");
    HIPCK(c, hipSetDevice(c->device));
    int blocks = 0;
    int rc = trainPass(c, k, n, q, t, &blocks);
    if (rc != PRR_OK) return rc;
    const CacheOpt o = makeOpt(k, n);
    HIPCK(c, launchCacheAdam(k->master, k->m1, k->m2, k->partial, blocks, o, k->wt, k->wn, k->grad, c->stream));
    HIPCK(c, launchCacheLossSum(k->lossPart, blocks, k->scalar, c->stream));
    float ls = 0.f;
    if (grad) HIPCK(c, hipMemcpyAsync(grad, k->grad, kCacheWeights * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    HIPCK(c, hipMemcpyAsync(&ls, k->scalar, sizeof(float), hipMemcpyDeviceToHost, c->stream));
    HIPCK(c, hipStreamSynchronize(c->stream));
    if (loss) *loss = ls / (3.f * (float) n);
    return PRR_OK;
}

int prr_cache_train(prr_context *c, prr_cache *k, int64_t n, const float *q, const float *t, int32_t steps,
                    float *loss) {
    if (!c || !k) return fail(c, PRR_ERR_INVALID, "NULL argument");
    if (n <= 0 || steps < 0 || !q || !t) return fail(c, PRR_ERR_INVALID, "bad arrays");
    HIPCK(c, hipSetDevice(c->device));
    const int64_t B = k->p.train_batch;
    int64_t lastN = 0;
    int lastBlocks = 0;
    for (int s = 0; s < steps; s++) {
        const int64_t a = ((int64_t) s * B) % n;
        const int64_t m = std::min(B, n - a);
        int blocks = 0;
        int rc = trainPass(c, k, m, q + a * kCacheQ, t + a * 3, &blocks);
        if (rc != PRR_OK) return rc;
        const CacheOpt o = makeOpt(k, m);
        HIPCK(c, launchCacheAdam(k->master, k->m1, k->m2, k->partial, blocks, o, k->wt, k->wn, nullptr, c->stream));
        k->step++;
        lastN = m;
        lastBlocks = blocks;
    }
    if (loss) {
        *loss = 0.f;
        if (lastBlocks > 0) {
            float ls = 0.f;
            HIPCK(c, launchCacheLossSum(k->lossPart, lastBlocks, k->scalar, c->stream));
            HIPCK(c, hipMemcpyAsync(&ls, k->scalar, sizeof(float), hipMemcpyDeviceToHost, c->stream));
            HIPCK(c, hipStreamSynchronize(c->stream));
            *loss = ls / (3.f * (float) lastN);
        }
    }
    return PRR_OK;
}

static int ensureFrameBufs(prr_context *c, prr_cache *k, int64_t slots, int64_t tCap) {
    if (slots <= k->qCap && tCap <= k->tCap) return PRR_OK;
    HIPCK(c, hipStreamSynchronize(c->stream));
    freeFrameBufs(k);
    slots = std::max(slots, k->qCap);
    tCap = std::max(tCap, k->tCap);
    hipError_t e = hipSuccess;
    auto A = [&](auto **ptr, size_t count) {
        using T = typename std::remove_reference<decltype(**ptr)>::type;
        if (e != hipSuccess) return;
        void *q = nullptr;
        e = hipMalloc(&q, std::max<size_t>(1, count) * sizeof(T));
        if (e == hipSuccess) {
            k->frameAllocs.push_back(q);
            *ptr = (T *) q;
        }
    };
    A(&k->counts, 2);
    A(&k->qSlot, slots);
    A(&k->q, (size_t) slots * kCacheQ);
    A(&k->qBeta, slots);
    A(&k->tSlot, tCap);
    A(&k->tq, (size_t) tCap * kCacheQ);
    A(&k->tBeta, tCap);
    A(&k->tLpre, tCap);
    A(&k->tTarget, (size_t) tCap * 3);
    if (e != hipSuccess) {
        freeFrameBufs(k);
        return fail(c, PRR_ERR_NOMEM, std::string("cache frame buffers: ") + hipGetErrorString(e));
    }
    k->qCap = slots;
    k->tCap = tCap;
    return PRR_OK;
}

static int64_t gcd64(int64_t a, int64_t b) {
    while (b) {
        const int64_t t = a % b;
        a = b;
        b = t;
    }
    return a;
}

int prr_cache_render(prr_context *c, const prr_scene *s, prr_cache *k, const prr_render_params *p,
                     const prr_cache_render_params *cp, float *rgb, prr_cache_stats *cs) {
    if (!c || !s || !k || !p || !cp || !rgb) return fail(c, PRR_ERR_INVALID, "NULL argument");
    if (p->width <= 0 || p->height <= 0 || p->spp <= 0 || p->max_depth <= 0 || p->max_depth > 64 ||
        p->sample_offset < 0 || (int64_t) p->sample_offset + p->spp > INT32_MAX)
        return fail(c, PRR_ERR_INVALID, "bad render parameters");
    if (p->variance_threshold > 0.f && p->min_spp < p->spp)
        return fail(c, PRR_ERR_UNSUPPORTED, "adaptive sampling is not available with the cache");
    if (p->num_photons > 0) return fail(c, PRR_ERR_UNSUPPORTED, "photon gather is not available with the cache");
    if (cp->train_every <= 0 || (cp->train_every & (cp->train_every - 1)) || cp->query_bounce < 1 ||
        cp->max_train_steps < 0 || cp->train_capacity < 0 || cp->train_suffix < 0)
        return fail(c, PRR_ERR_INVALID, "bad cache render parameters");
    const int64_t pixels = (int64_t) p->width * p->height;
    if (pixels > c->maxSlots) return fail(c, PRR_ERR_UNSUPPORTED, "image larger than the context's path slots");
    HIPCK(c, hipSetDevice(c->device));
    const int spc = (int) std::max<int64_t>(1, std::min<int64_t>(p->spp, c->maxSlots / pixels));
    int rc = ensureRenderBufs(c, pixels * spc, p->max_depth);
    if (rc != PRR_OK) return rc;
    AdaptiveBufs ab;
    if ((rc = ensureFrame(c, pixels, ab)) != PRR_OK) return rc;
    const int64_t tCap = cp->train_capacity > 0 ? cp->train_capacity : ((int64_t) 1 << 22);
    if ((rc = ensureFrameBufs(c, k, c->pbSlots, tCap)) != PRR_OK) return rc;

    const bool fellBack = cp->use_cache && k->diverged;
    if (k->diverged) {  // fallback frame: plain paths, a fresh network
        if ((rc = initWeights(c, k)) != PRR_OK) return rc;
        if ((rc = resetState(c, k)) != PRR_OK) return rc;
        k->diverged = false;
    }
    const bool useCache = cp->use_cache && !fellBack;
    CacheRec cr;
    cr.query = useCache ? 1 : 0;
    cr.train = cp->train ? 1 : 0;
    cr.queryBounce = cp->query_bounce;
    cr.trainStop = useCache && cp->train_suffix > 0 ? cp->query_bounce + cp->train_suffix : 0;
    cr.trainMask = (uint32_t) cp->train_every - 1u;
    cr.tCap = (uint32_t) std::min<int64_t>(tCap, 0xFFFFFFFFll);
    cr.qCount = k->counts;
    cr.tCount = k->counts + 1;
    cr.qSlot = k->qSlot;
    cr.q = k->q;
    cr.qBeta = k->qBeta;
    cr.tSlot = k->tSlot;
    cr.tq = k->tq;
    cr.tBeta = k->tBeta;
    cr.tLpre = k->tLpre;

    hipStream_t st = c->stream;
    HIPCK(c, hipMemsetAsync(c->stats, 0, 8 * sizeof(unsigned long long), st));
    HIPCK(c, hipMemsetAsync(c->fb, 0, (size_t) pixels * sizeof(float4), st));
    HIPCK(c, hipMemsetAsync(k->counts, 0, 2 * sizeof(uint32_t), st));
    HIPCK(c, hipEventRecord(c->ev0, st));
    const SceneView view = [&] {
        SceneView v = s->view;
        v.ph_count = 0;
        return v;
    }();
    Timing tm;
    tm.on = (p->flags & PRR_RENDER_KERNEL_TIMING) != 0;
    ChunkInfo ci;
    ci.W = p->width;
    ci.H = p->height;
    ci.seed = p->seed;
    ci.maxDepth = p->max_depth;
    ci.clampInd = p->indirect_clamp;
    ci.cache = cr;
    uint64_t queries = 0;
    uint32_t tDone = 0, tSeen = 0;
    double msInfer = 0.0;
    for (int s0 = 0; s0 < p->spp; s0 += spc) {
        ci.pixels = (int) pixels;
        ci.spc = std::min(spc, p->spp - s0);
        ci.sampleBase = p->sample_offset + s0;
        if ((rc = resetChunk(c, p->max_depth, st)) != PRR_OK) return rc;
        HIPCK(c, hipMemsetAsync(k->counts, 0, sizeof(uint32_t), st));
        HIPCK(c, launchRaygen(view, c->pb, ci, c->shadeBlocks, st));
        if ((rc = runBounces(c, view, ci, tm, st)) != PRR_OK) return rc;
        if (useCache) {
            HIPCK(c, hipEventRecord(k->ev[1], st));
            HIPCK(c, launchCacheInfer(k->view(), pixels * ci.spc, k->counts, k->q, nullptr, k->qSlot, k->qBeta,
                                      c->pb.result, p->indirect_clamp, c->numCU * 3, st));
            HIPCK(c, hipEventRecord(k->ev[2], st));
        }
        uint32_t cnt[2] = {0, 0};
        HIPCK(c, hipMemcpyAsync(cnt, k->counts, sizeof(cnt), hipMemcpyDeviceToHost, st));
        HIPCK(c, hipStreamSynchronize(st));
        if (useCache) {
            float dt = 0.f;
            HIPCK(c, hipEventElapsedTime(&dt, k->ev[1], k->ev[2]));
            msInfer += dt;
        }
        queries += cnt[0];
        tSeen = cnt[1];
        const uint32_t tNow = (uint32_t) std::min<int64_t>(cnt[1], tCap);
        HIPCK(c, launchCacheTargets(c->pb, cr, tDone, tNow, k->tTarget, st));
        tDone = tNow;
        HIPCK(c, launchAccumulate(c->pb, ci, c->fb, st));
    }
    HIPCK(c, launchResolve(c->fb, (int) pixels, p->spp, rgb, st));
    HIPCK(c, hipEventRecord(k->ev[0], st));

    // training: batches of train_batch spread over the frame's records
    int steps = 0;
    float loss = 0.f;
    const int64_t n = tDone;
    if (cp->train && n > 0 && cp->max_train_steps > 0) {
        const int64_t B = k->p.train_batch;
        steps = (int) std::min<int64_t>(cp->max_train_steps, (n + B - 1) / B);
        int64_t mult = 2654435761ll % n;
        if (mult == 0) mult = 1;
        while (gcd64(mult, n) != 1) mult++;
        int lastBlocks = 0;
        int64_t lastN = 0;
        for (int t = 0; t < steps; t++) {
            const int64_t j0 = (int64_t) t * B;
            const int64_t m = std::min(B, n - j0);
            const int64_t start = (int64_t) (((unsigned __int128) (j0 % n) * (uint64_t) mult) % (uint64_t) n);
            int blocks = 0;
            if ((rc = trainPass(c, k, m, k->tq, k->tTarget, &blocks, start, n, mult)) != PRR_OK) return rc;
            const CacheOpt o = makeOpt(k, m);
            HIPCK(c, launchCacheAdam(k->master, k->m1, k->m2, k->partial, blocks, o, k->wt, k->wn, nullptr, st));
            k->step++;
            lastBlocks = blocks;
            lastN = m;
        }
        float ls = 0.f;
        HIPCK(c, launchCacheLossSum(k->lossPart, lastBlocks, k->scalar, st));
        HIPCK(c, hipMemcpyAsync(&ls, k->scalar, sizeof(float), hipMemcpyDeviceToHost, st));
        HIPCK(c, hipStreamSynchronize(st));
        loss = ls / (3.f * (float) lastN);
        if (!std::isfinite(loss) || loss > cp->divergence_loss) k->diverged = true;
    }
    HIPCK(c, hipEventRecord(c->ev1, st));
    HIPCK(c, hipStreamSynchronize(st));
    if (cs) {
        *cs = prr_cache_stats{};
        if ((rc = finishStats(c, tm, 0, &cs->render)) != PRR_OK) return rc;
        float a = 0.f, b = 0.f;
        HIPCK(c, hipEventElapsedTime(&a, c->ev0, k->ev[0]));
        HIPCK(c, hipEventElapsedTime(&b, k->ev[0], c->ev1));
        cs->ms_render = a;
        cs->ms_train = b;
        cs->ms_infer = msInfer;
        cs->queries = queries;
        cs->train_samples = tSeen;
        cs->train_steps = steps;
        cs->fell_back = fellBack ? 1 : 0;
        cs->train_loss = loss;
        cs->diverged = k->diverged ? 1 : 0;
    }
    return PRR_OK;
}

}  // extern "C"
