// render.hpp — wavefront path state shared by the render kernels and the host
// code (abi.hip).
#pragma once

#include "dev_common.hpp"

namespace prr {

// path flags
constexpr uint32_t F_SPEC = 1u;          // last bounce was a delta bounce (specularBounce)
constexpr uint32_t F_PREV_DIFF = 2u;     // a BSDF-sampled ray from a diffuse vertex is in flight
constexpr uint32_t F_PEND_A = 4u;        // area-light NEE waiting for its shadow ray
constexpr uint32_t F_PEND_E = 8u;        // environment NEE waiting
constexpr uint32_t F_PEND_P = 16u;       // point-light NEE waiting
constexpr uint32_t F_TERMINATING = 32u;  // path ended; only pending NEE remains
constexpr uint32_t F_PEND_C = 64u;       // caustic photon estimate (no shadow ray)
constexpr uint32_t F_TRAIN = 128u;       // radiance-cache training path: full depth, records its vertices
constexpr uint32_t F_PEND_ALL = F_PEND_A | F_PEND_E | F_PEND_P | F_PEND_C;

// Segmented queues. A bounce's ray queue (and its shadow queue) is split into
// kQSegs segments: segment k holds its count[k] entries at
// [k * cap, k * cap + count[k]). A producer wave appends its rays and shadow
// rays to the segment of its input position, with one 64-bit atomic on that
// segment's counter pair (appends of every wave of the GPU to one counter
// serialise on that address, ~9 ns per wave, which bounded shading).
// Consumers see the queue as the dense range [0, total) and find an entry
// through the prefix of the counts (kQueueScan writes it, with the total,
// after each shading pass). Queues written in order (camera rays) carry the
// prefix of a queue filled segment by segment, which maps w to entry w.
constexpr int kQSegs = 64;

struct PathBufs {
    int64_t slots;           // capacity (paths per chunk)
    float4 *L;               // radiance so far (xyz), w = the path's flags (bits)
    float4 *beta;            // throughput (xyz; after a diffuse vertex, times its BSDF factor f/pdf),
                             // w = that vertex's BSDF pdf
    float4 *prevP;           // diffuse vertex that shot the BSDF ray (xyz; only with area emitters)
    float4 *pend;            // 4 * slots: area, env, point NEE and caustic-photon contributions
    uint64_t *rng;           // PCG32 state
    uint8_t *vis;            // 3 * slots: visibility of the pending NEE rays
    float4 *result;          // final L per slot
    uint32_t *qSlot[2];      // ray queue (ping-pong)
    float *qOrg[2];          // origin, 3 floats per entry; a NaN x marks an entry without a ray
                             // (a path that comes back only for its pending NEE)
    float *qDir[2];          // direction, 3 floats per entry
    float4 *hit;             // t, u, v, prim (bits), aligned with the queue being shaded
    uint32_t *sqSlot;        // shadow queue: slot | kind << 30
    float4 *sqOrg;           // origin, tmax
    float *sqDir;            // direction, 3 floats per entry
    uint32_t *qcnt;          // ray-queue count per bounce (total over its segments)
    uint32_t *sqcnt;         // shadow-queue count per bounce
    unsigned long long *segc;  // per shading pass b and segment: entries appended to the ray queue of
                               // b + 1 (low 32 bits) and to the shadow queue of b (high 32 bits)
    uint32_t *qpre;          // kQSegs + 1 prefix counts per bounce (ray queues)
    uint32_t *sqpre;         // the same for shadow queues
    int64_t qcap;            // entries per ray-queue segment
    int64_t sqcap;           // entries per shadow-queue segment
    uint32_t *fetch;         // persistent-kernel work counters: [2b] extend, [2b+1] shadow
    unsigned long long *stats;  // [0] closest rays, [1] any rays, [2] samples, [3] photons stored,
                                // [4] closest rays traced by kFinish, [5] samples replaced by 0
                                // (non-finite), [6] shadow rays traced by kFinish, [7] NEE terms exactly
                                // zero, whose shadow rays are not traced
};

// Buffers of the shadow-queue binning (k_sort.hip): the binned queue (cap
// entries, as many as the shadow queue) and the per-block bin histograms /
// offsets of its counting sort (bins x blocks).
struct ShadowSortBufs {
    int64_t cap = 0;
    int blocks = 0;
    uint32_t *slot = nullptr;
    float4 *org = nullptr;
    float *dir = nullptr;  // 3 floats per entry
    uint32_t *hist = nullptr;
};

// Shading bins (kShadeBin): queue indices of the entries whose hit is not a
// diffuse surface in lists[0, counts[0]), of those whose hit is, in
// lists[cap, cap + counts[1]).
struct ShadeBins {
    int64_t cap = 0;
    int blocks = 0;                   // blocks of the counting sort (chunks of the queue)
    uint32_t *lists = nullptr;        // 2 * cap
    uint32_t *counts = nullptr;       // 2
    uint32_t *blockCounts = nullptr;  // 2 * blocks
};

constexpr int kCacheQueryFloats = 13;  // position, wo, normal, roughness, albedo

// Radiance-cache mode of a render. Training paths (hash of pixel and sample
// index & trainMask == 0) run to full depth and record each diffuse vertex
// from bounce 1 on; the other paths stop at their first diffuse vertex at
// bounce >= queryBounce and record a query there.
struct CacheRec {
    int query = 0;               // paths other than training paths stop and query
    int train = 0;               // training paths are selected and recorded
    int queryBounce = 2;
    int trainStop = 0;           // training paths stop and query at bounce >= trainStop (0: full depth)
    uint32_t trainMask = 15;
    uint32_t tCap = 0;           // training-record capacity
    uint32_t *qCount = nullptr;  // queries: slot, 13 query floats, throughput
    uint32_t *qSlot = nullptr;
    float *q = nullptr;
    float4 *qBeta = nullptr;
    uint32_t *tCount = nullptr;  // training records: slot, 13 query floats, throughput, radiance so far
    uint32_t *tSlot = nullptr;
    float *tq = nullptr;
    float4 *tBeta = nullptr;
    float4 *tLpre = nullptr;
};

// Serving: one request's share of a batch, samples [s0, s0 + ns) of pixels
// [p0, p0 + np) in sample-major slot order from slot0. A request has at most
// one segment per batch, so each of its pixels takes its samples in order.
struct Segment {
    CamView cam;
    int W, H;
    int p0, np, s0, ns;
    int64_t slot0;
    uint64_t seed;
    float4 *fb;  // the request's running sums
};

struct ChunkInfo {
    int W, H;
    int pixels;          // W * H
    int spc;             // samples of this chunk per pixel
    int sampleBase;      // absolute index of the chunk's first sample
    int bandRows = 0;    // camera-path bands (pixelOf); in the padding before seed
    uint64_t seed;
    int maxDepth;
    float clampInd;      // Integrator::indirectClamp
    // Pixel of slot entry e = slot % pixels (pixelOf): pixList[e] (adaptive
    // passes render only the pixels still sampling); with no list, e in
    // bands of bandRows image rows, column by column within a band, so the
    // 64 lanes of a wave take an 8 x 8 block of pixels at bandRows = 8
    // (bandRows <= 1: scanline order).
    const int32_t *pixList = nullptr;
    CacheRec cache;
    // per-slot sample index (serving batches); null: sampleBase + slot / pixels
    const uint32_t *sampleOf = nullptr;
    // Path completion: the first bounce whose queue holds at most tailPaths
    // entries (0: never) is run by kFinish; kExtend and kShade skip it and
    // every later bounce.
    int64_t tailPaths = 0;
};

PRR_DEV int pixelOf(const ChunkInfo &ci, int e) {
    if (ci.pixList) return ci.pixList[e];
    if (ci.bandRows <= 1) return e;
    const int span = ci.bandRows * ci.W;  // entries of a full band
    const int band = e / span, local = e - band * span;
    const int rows = min(ci.bandRows, ci.H - band * ci.bandRows);
    const int x = local / rows;
    return (band * ci.bandRows + (local - x * rows)) * ci.W + x;
}

// Per-pixel Welford state of the adaptive sampler (Renderer.hpp:14-40):
// mean.xyz / m2.xyz per channel, mean.w = sample count (int bits),
// m2.w = stopped flag.
struct AdaptiveBufs {
    float4 *mean;
    float4 *m2;
    int32_t *list;       // active pixel list
    uint32_t *count;     // active pixel count
};

// Three-float records (queue origins and directions).
PRR_DEV V3 ldv3(const float *a, int64_t e) {
    const float *p = a + 3 * e;
    return V3(p[0], p[1], p[2]);
}
PRR_DEV void stv3(float *a, int64_t e, const V3 &v) {
    float *p = a + 3 * e;
    p[0] = v.x;
    p[1] = v.y;
    p[2] = v.z;
}
// Origin of a ray-queue entry without a ray.
PRR_DEV V3 noRayOrigin() { return V3(__builtin_nanf(""), 0.f, 0.f); }

// Entries per segment of a queue whose producer appends at most perEntry
// entries for each of at most `entries` input positions. The waves that
// append to segment k are those that shade input positions
// [64 j, 64 j + 64) with j % kQSegs == k; one producer pass has at most
// ceil(ceil(entries / 64) / kQSegs) of them. Two slack groups cover the two
// list passes of material-sorted shading, whose positions both start at 0.
inline int64_t segCapacity(int64_t entries, int perEntry) {
    const int64_t groups = (entries + 63) / 64;
    return ((groups + kQSegs - 1) / kQSegs + 2) * 64 * perEntry;
}

// Copies a queue's kQSegs + 1 prefix counts to LDS. Every thread of the
// block calls it (it ends with a barrier).
PRR_DEV void loadPrefix(uint32_t *spre, const uint32_t *pre) {
    for (int k = threadIdx.x; k <= kQSegs; k += blockDim.x) spre[k] = pre[k];
    __syncthreads();
}

// Entry index of dense position w (< total) of a segmented queue: the last
// segment whose prefix is at most w (an empty segment's prefix equals the
// next one's, so it is never chosen), then w's offset in it. Queue indices
// fit in 32 bits (at most kQSegs * cap entries, a little over the 2^27 path
// slots of a chunk), so the arithmetic is 32-bit.
PRR_DEV int64_t segEntry(const uint32_t *spre, int64_t cap, int64_t w) {
    const uint32_t w32 = (uint32_t) w;
    uint32_t k = 0;
#pragma unroll
    for (uint32_t step = kQSegs / 2; step > 0; step >>= 1)
        if (spre[k + step] <= w32) k += step;
    return (int64_t) (k * (uint32_t) cap + (w32 - spre[k]));
}

}  // namespace prr
