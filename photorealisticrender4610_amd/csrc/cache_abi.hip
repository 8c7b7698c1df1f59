// cache_abi.hip — C ABI of the learned radiance cache (prr_cache_*).
#include <algorithm>
#include <cmath>
#include <new>
#include <string>
#include <vector>

#include "context.hpp"
#include "device/cache.hpp"
#include "device/launch_cache.hpp"

using namespace prr;

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
    std::vector<void *> allocs;
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

static void freeCache(prr_cache *k) {
    for (void *q : k->allocs) (void) hipFree(q);
    delete k;
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
static int trainPass(prr_context *c, prr_cache *k, int64_t n, const float *q, const float *t, int *blocks) {
    *blocks = cacheTrainBlocks(n);
    if (*blocks > k->maxBlocks) return fail(c, PRR_ERR_INVALID, "batch larger than train_batch");
    HIPCK(c, launchCacheTrain(k->view(), n, q, t, k->partial, k->lossPart, k->p.grad_clip, c->stream));
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
    // He-uniform: U(-sqrt(6 / fan_in), +sqrt(6 / fan_in)) per layer, SplitMix64 draws
    std::vector<float> w(kCacheWeights);
    uint64_t st = p->seed;
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
    if (e == hipSuccess)
        e = hipMemcpyAsync(k->master, w.data(), w.size() * sizeof(float), hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) {
        freeCache(k);
        return fail(c, PRR_ERR_DEVICE, std::string("cache init: ") + hipGetErrorString(e));
    }
    int rc = resetState(c, k);
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
    HIPCK(c, launchCacheInfer(k->view(), n, q, out, nullptr, nullptr, nullptr, c->numCU * 3, c->stream));
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

}  // extern "C"
