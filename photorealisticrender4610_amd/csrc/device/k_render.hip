// k_render.hip — wavefront integrator kernels.
//
// One chunk = `spc` samples for every pixel, one path per (pixel, sample)
// slot, slot = s * pixels + pixel. Per bounce b (0..maxDepth):
//   kExtend  closest hit of the ray queue of bounce b
//   kShadow  any hit of the shadow queue emitted at bounce b-1
//   kShade   Integrator::castRay's loop body for bounce b (Integrator.cpp:122-257)
// The NEE terms of bounce b are added at bounce b+1, once their shadow rays
// are resolved, in the reference's order (area, env, then the BSDF-MIS term),
// so the float sums match the reference's. kAccumulate adds the chunk's
// results per pixel in sample order (Renderer.cpp:84-99). Once few paths
// remain, kFinish runs the rest of each remaining path in one launch.
#include <type_traits>

#include "dev_traverse.hpp"
#include "launch_render.hpp"
#include "render.hpp"

namespace prr {

constexpr int kShadeBlock = 256;

PRR_DEV uint32_t laneId() { return __lane_id(); }

// Wave-aggregated queue append: one atomic per wave. All lanes that reach
// this point take part; lanes with want == false get no index.
PRR_DEV uint32_t waveAppend(uint32_t *counter, bool want) {
    const uint64_t mask = __ballot(want);
    if (mask == 0) return 0;
    const int leader = __ffsll((unsigned long long) mask) - 1;
    uint32_t base = 0;
    if ((int) laneId() == leader) base = atomicAdd(counter, (uint32_t) __popcll(mask));
    base = __shfl(base, leader);
    const uint64_t below = mask & ((1ull << laneId()) - 1ull);
    return base + (uint32_t) __popcll(below);
}

// Adds the wave's count of replaced (non-finite) samples to stats[5]: one
// atomic per wave. Every lane of the wave calls it.
PRR_DEV void countNan(const PathBufs &pb, uint32_t n) {
    for (int o = 32; o > 0; o >>= 1) n += __shfl_xor(n, o);
    if (laneId() == 0 && n) atomicAdd(&pb.stats[5], (unsigned long long) n);
}

PRR_DEV uint64_t pcgInc(uint32_t sample) { return ((uint64_t) sample << 1) | 1u; }

// Pixel/sample hash selecting radiance-cache training paths (lowbias32 mix).
PRR_DEV uint32_t trainHash(uint32_t p, uint32_t s) {
    uint32_t h = p * 0x9E3779B1u + s * 0x85EBCA77u;
    h ^= h >> 16;
    h *= 0x7FEB352Du;
    h ^= h >> 15;
    h *= 0x846CA68Bu;
    h ^= h >> 16;
    return h;
}

// Flags of a path at its first bounce: a radiance-cache training path when
// the render trains the cache and the hash of its pixel and sample selects it.
// Training renders always run the shading kernels compiled with the cache
// (CACHE), which find the pixel through the camera-path bands. The kernels
// compiled without it never see ci.cache.train set; their copy keeps the
// plain entry-to-pixel lookup, which is never reached there and which the
// compiler schedules without spills, where the band lookup (or none) left
// 16-24 B per lane of spills in kShade.
template <bool CACHE>
PRR_DEV uint32_t pathFlags0(const ChunkInfo &ci, uint32_t slot, uint32_t sIdx) {
    if (!ci.cache.train) return 0u;
    const int e = (int) (slot % (uint32_t) ci.pixels);
    const int p = CACHE ? pixelOf(ci, e) : (ci.pixList ? ci.pixList[e] : e);
    return (trainHash((uint32_t) p, sIdx) & ci.cache.trainMask) == 0u ? F_TRAIN : 0u;
}

// One radiance-cache query / training record: position, wo, shading normal,
// roughness, albedo (the cache's 13 query floats).
PRR_DEV void putQuery(float *q, const V3 &P, const V3 &wo, const V3 &n, float rough, const V3 &alb) {
    q[0] = P.x, q[1] = P.y, q[2] = P.z;
    q[3] = wo.x, q[4] = wo.y, q[5] = wo.z;
    q[6] = n.x, q[7] = n.y, q[8] = n.z;
    q[9] = rough;
    q[10] = alb.x, q[11] = alb.y, q[12] = alb.z;
}

PRR_DEV V3 clampInd(const V3 &c, float k) {  // Integrator.cpp:105-111
    if (k <= 0.f) return c;
    float lum = (c.x + c.y + c.z) / 3;
    return (lum > k) ? c * (k / lum) : c;
}
PRR_DEV float mis(float a, float b) { return (a * a) / (a * a + b * b + 1e-6f); }  // Integrator.hpp:30-32

// Queue 0 holds the n camera (or caller) rays in order, entry w at index w:
// its prefix is that of a queue filled segment by segment. Also counts the
// rays and samples.
PRR_DEV void queue0Written(const PathBufs &pb, int64_t n) {
    if (blockIdx.x != 0) return;
    for (int k = threadIdx.x; k <= kQSegs; k += blockDim.x) {
        const int64_t p = (int64_t) k * pb.qcap;
        pb.qpre[k] = (uint32_t) (p < n ? p : n);
    }
    if (threadIdx.x == 0) {
        pb.qcnt[0] = (uint32_t) n;
        atomicAdd(&pb.stats[0], (unsigned long long) n);
        atomicAdd(&pb.stats[2], (unsigned long long) n);
    }
}

// After shading pass b: the prefix counts and totals of the ray queue of
// bounce b + 1 (block 0) and of the shadow queue of bounce b (block 1), each
// count clamped to the segment capacity (a producer never passes it).
__global__ __launch_bounds__(64) void kQueueScan(PathBufs pb, int bounce) {
    static_assert(kQSegs == 64, "one lane per segment");
    const bool sh = blockIdx.x == 1;
    const int b = sh ? bounce : bounce + 1;
    const int64_t cap = sh ? pb.sqcap : pb.qcap;
    const int k = (int) threadIdx.x;
    const unsigned long long pair = pb.segc[(int64_t) bounce * kQSegs + k];
    const uint32_t c = sh ? (uint32_t) (pair >> 32) : (uint32_t) pair;
    uint32_t v = (int64_t) c < cap ? c : (uint32_t) cap;
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t t = __shfl_up(v, o);
        if (k >= o) v += t;
    }
    uint32_t *pre = (sh ? pb.sqpre : pb.qpre) + (int64_t) b * (kQSegs + 1);
    pre[k + 1] = v;
    if (k == 0) pre[0] = 0u;
    if (k == 63) (sh ? pb.sqcnt : pb.qcnt)[b] = v;
}

// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kShadeBlock) void kRaygen(SceneView s, PathBufs pb, ChunkInfo ci) {
    const int64_t n = (int64_t) ci.pixels * ci.spc;
    const int64_t stride = (int64_t) gridDim.x * kShadeBlock;
    for (int64_t slot = (int64_t) blockIdx.x * kShadeBlock + threadIdx.x; slot < n; slot += stride) {
        const int p = pixelOf(ci, (int) (slot % ci.pixels));
        const int sIdx = ci.sampleBase + (int) (slot / ci.pixels);
        Pcg rng = Pcg::make(ci.seed, (uint32_t) p, (uint32_t) sIdx);
        const float ox = rng.next();  // Renderer.cpp:75-76
        const float oy = rng.next();
        V3 o, d;
        cameraRay(s.cam, p % ci.W, p / ci.W, ci.W, ci.H, ox, oy, rng, o, d);
        pb.rng[slot] = rng.state;
        pb.qSlot[0][slot] = (uint32_t) slot;
        stv3(pb.qOrg[0], slot, o);
        stv3(pb.qDir[0], slot, d);
    }
    queue0Written(pb, n);
}

// ---------------------------------------------------------------------------
// Serving batches: slot -> segment by binary search over slot0 (segments in
// increasing slot0), then the segment's pixel and sample; camera rays from the
// request's own camera, PCG32 stream of (request seed, pixel, sample).
PRR_DEV int findSegment(const Segment *segs, int nseg, int64_t slot) {
    int lo = 0, hi = nseg - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (segs[mid].slot0 <= slot)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

__global__ __launch_bounds__(kShadeBlock) void kRaygenSeg(PathBufs pb, const Segment *__restrict__ segs, int nseg,
                                                          int64_t n, uint32_t *__restrict__ sampleOf) {
    const int64_t stride = (int64_t) gridDim.x * kShadeBlock;
    for (int64_t slot = (int64_t) blockIdx.x * kShadeBlock + threadIdx.x; slot < n; slot += stride) {
        const Segment &sg = segs[findSegment(segs, nseg, slot)];
        const int64_t local = slot - sg.slot0;
        const int p = sg.p0 + (int) (local % sg.np);
        const int sIdx = sg.s0 + (int) (local / sg.np);
        Pcg rng = Pcg::make(sg.seed, (uint32_t) p, (uint32_t) sIdx);
        const float ox = rng.next();  // Renderer.cpp:75-76
        const float oy = rng.next();
        V3 o, d;
        cameraRay(sg.cam, p % sg.W, p / sg.W, sg.W, sg.H, ox, oy, rng, o, d);
        pb.rng[slot] = rng.state;
        sampleOf[slot] = (uint32_t) sIdx;
        pb.qSlot[0][slot] = (uint32_t) slot;
        stv3(pb.qOrg[0], slot, o);
        stv3(pb.qDir[0], slot, d);
    }
    queue0Written(pb, n);
}

// Renderer.cpp:84-99 per request pixel: its ns samples of the batch added in
// sample order. pixStart[k] = pixels of segments before k.
__global__ void kAccumulateSeg(PathBufs pb, const Segment *__restrict__ segs, const int64_t *__restrict__ pixStart,
                               int nseg, int64_t total) {
    const int64_t t = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= total) return;
    int lo = 0, hi = nseg - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (pixStart[mid] <= t)
            lo = mid;
        else
            hi = mid - 1;
    }
    const Segment &sg = segs[lo];
    const int64_t pp = t - pixStart[lo];
    float4 *fb = sg.fb + sg.p0 + pp;
    V3 acc = ld3(*fb);
    for (int k = 0; k < sg.ns; k++) acc += ld3(pb.result[sg.slot0 + (int64_t) k * sg.np + pp]);
    *fb = st4(acc, 0.f);
}

// ---------------------------------------------------------------------------
// Integrator::castRay on caller rays: slot k traces ray k with the PCG32
// stream of (seed, k, sample 0); no camera draws.
__global__ __launch_bounds__(kShadeBlock) void kRaygenRays(PathBufs pb, int64_t n, const float *org, const float *dir,
                                                           int64_t first, uint64_t seed) {
    const int64_t stride = (int64_t) gridDim.x * kShadeBlock;
    for (int64_t slot = (int64_t) blockIdx.x * kShadeBlock + threadIdx.x; slot < n; slot += stride) {
        const int64_t k = first + slot;
        const Pcg rng = Pcg::make(seed, (uint32_t) k, 0u);
        pb.rng[slot] = rng.state;
        pb.qSlot[0][slot] = (uint32_t) slot;
        stv3(pb.qOrg[0], slot, V3(org[3 * k], org[3 * k + 1], org[3 * k + 2]));
        stv3(pb.qDir[0], slot, V3(dir[3 * k], dir[3 * k + 1], dir[3 * k + 2]));
    }
    queue0Written(pb, n);
}

__global__ void kCopyResults(PathBufs pb, int64_t n, int64_t first, float *L) {
    const int64_t i = (int64_t) blockIdx.x * 256 + threadIdx.x;
    uint32_t nan = 0;
    if (i < n) {
        const float4 r = pb.result[i];
        L[3 * (first + i)] = r.x;
        L[3 * (first + i) + 1] = r.y;
        L[3 * (first + i) + 2] = r.z;
        nan = r.w != 0.f ? 1u : 0u;
    }
    countNan(pb, nan);
}

// Persistent traversal: idle lanes refill from the queue once `refillAt`
// lanes of the wave are idle (64: only when the whole wave is idle), taking
// consecutive entries from the wave's claimed run (waveTake), which costs an
// atomic on the bounce's counter only when the run is used up. ITER: 0 one
// node per work check, 1 while-while (closestIter), 2 speculative while-while
// (closestIterSpec), 3 while-while on the two-level records (closestIter2).

template <int ITER>
__global__ __launch_bounds__(kTraceBlock) void kExtend(SceneView s, PathBufs pb, int bounce, uint2 *spill, int refillAt,
                                                       int stackDepth, int wwCap, int fetchRun, int64_t tailPaths) {
    extern __shared__ uint2 stk[];
    __shared__ uint32_t spre[kQSegs + 1];
    const int q = bounce & 1;
    const int64_t n = pb.qcnt[bounce];
    if (n <= tailPaths) return;  // kFinish runs this bounce's paths (tailPaths > 0) or there are none
    loadPrefix(spre, pb.qpre + (int64_t) bounce * (kQSegs + 1));
    uint32_t *fetch = &pb.fetch[2 * bounce];
    const int64_t lane = (int64_t) blockIdx.x * kTraceBlock + threadIdx.x;
    LaneStack st;
    st.lds = stk;
    st.lane = threadIdx.x;
    st.spill = spill + lane * kSpill;
    st.depth = stackDepth;
    st.sp = 0;
    TravState ts;
    int64_t idx = -1;
    bool exhausted = false;
    WaveRange wr;
    const int64_t waves = (int64_t) gridDim.x * (kTraceBlock / 64);
    while (true) {
        const bool need = idx < 0 && !exhausted;
        const uint64_t needMask = __ballot(need);
        const bool refill = __popcll(needMask) >= refillAt || __ballot(idx >= 0) == 0;
        const int64_t got = waveTake(wr, fetch, n, need && refill, fetchRun, waves);
        if (need && refill) {
            if (got >= n) {
                exhausted = true;
            } else {
                const int64_t e = segEntry(spre, pb.qcap, got);
                const V3 o = ldv3(pb.qOrg[q], e);
                if (isnan(o.x)) {  // no ray (a ray with a NaN origin hits no primitive: the same record)
                    pb.hit[e] = make_float4(__builtin_inff(), 0.f, 0.f, __int_as_float(-1));
                } else {
                    idx = e;
                    travBegin(ts, o, ldv3(pb.qDir[q], e), 0.f, s.root, st);
                }
            }
        }
        if (__ballot(idx >= 0 || !exhausted) == 0) break;
        if (idx >= 0 && (ITER == 3   ? closestIter2(s, ts, st, wwCap)
                         : ITER == 2 ? closestIterSpec(s, ts, st, wwCap)
                         : ITER == 1 ? closestIter(s, ts, st, wwCap)
                                     : closestStep(s, ts, st))) {
            pb.hit[idx] = make_float4(ts.h.t, ts.h.u, ts.h.v, __int_as_float(ts.h.prim));
            idx = -1;
        }
    }
}

// TREE: 0 the binary tree, 1 the four-wide occlusion tree, 2 its quantised copy.
template <bool WW, int TREE>
__global__ __launch_bounds__(kTraceBlock) void kShadow(SceneView s, PathBufs pb, int bounce, uint2 *spill, int refillAt,
                                                       int stackDepth, int wwCap, int fetchRun, bool dense) {
    extern __shared__ uint2 stk[];
    __shared__ uint32_t spre[kQSegs + 1];
    const int64_t n = pb.sqcnt[bounce];
    loadPrefix(spre, pb.sqpre + (int64_t) bounce * (kQSegs + 1));
    uint32_t *fetch = &pb.fetch[2 * bounce + 1];
    const int64_t lane = (int64_t) blockIdx.x * kTraceBlock + threadIdx.x;
    using Stack = typename std::conditional<TREE != 0, RefStack, LaneStack>::type;
    Stack st = makeStack<Stack>(stk, spill + lane * kSpill, threadIdx.x, stackDepth);
    TravState ts;
    int64_t idx = -1;
    uint32_t sk = 0;
    bool exhausted = false;
    WaveRange wr;
    const int64_t waves = (int64_t) gridDim.x * (kTraceBlock / 64);
    while (true) {
        const bool need = idx < 0 && !exhausted;
        const uint64_t needMask = __ballot(need);
        const bool refill = __popcll(needMask) >= refillAt || __ballot(idx >= 0) == 0;
        const int64_t got = waveTake(wr, fetch, n, need && refill, fetchRun, waves);
        if (need && refill) {
            if (got >= n) {
                exhausted = true;
            } else {
                const int64_t e = dense ? got : segEntry(spre, pb.sqcap, got);
                const uint32_t k = pb.sqSlot[e];
                const float4 o4 = pb.sqOrg[e];
                travBegin(ts, ld3(o4), ldv3(pb.sqDir, e), o4.w, TREE == 2 ? s.root4q : (TREE == 1 ? s.root4 : s.root),
                          st);
                if constexpr (TREE == 2) anyPrepQ(s, ts);
                if (anyBegin(s, ts)) {
                    idx = got;
                    sk = k;
                } else {
                    pb.vis[(int64_t) (k >> 30) * pb.slots + (k & 0x3FFFFFFFu)] = 1;
                }
            }
        }
        if (__ballot(idx >= 0 || !exhausted) == 0) break;
        int res = kAnyGo;
        if (idx >= 0) {
            if constexpr (TREE == 2)
                res = WW ? anyIter4Q(s, ts, st, wwCap) : anyStep4Q(s, ts, st);
            else if constexpr (TREE == 1)
                res = WW ? anyIter4(s, ts, st, wwCap) : anyStep4(s, ts, st);
            else
                res = WW ? anyIter(s, ts, st, wwCap) : anyStep(s, ts, st);
        }
        if (res != kAnyGo) {
            pb.vis[(int64_t) (sk >> 30) * pb.slots + (sk & 0x3FFFFFFFu)] = res == kAnyHit ? 0 : 1;
            idx = -1;
        }
    }
}

// ---------------------------------------------------------------------------
struct HitGeom {
    V3 P, Ng;  // Ng: Intersection::normal (per-vertex normals interpolated where the triangle has them)
    int mat;
    int flags;
};
// pn: the hit primitive's prim_n row.
PRR_DEV HitGeom hitGeom(const SceneView &s, int prim, const float4 &pn, float t, const V3 &o, const V3 &d, float u,
                        float v) {
    HitGeom g;
    const int bits = __float_as_int(pn.w);
    g.mat = bits & kPrimMatMask;
    g.flags = bits >> kPrimFlagShift;
    if (prim < s.num_tris) {
        g.P = o + t * d;  // Triangle::finalize (Triangle.cpp:34-59)
        g.Ng = (g.flags & kPrimSmooth) ? smoothNormal(s, prim, u, v) : ld3(pn);
    } else {
        const float4 sc = s.sph_by_id[prim - s.num_tris];
        g.P = V3(o + d * t);  // Sphere.hpp:34-35
        g.Ng = normalize(V3(g.P - V3(sc.x, sc.y, sc.z)));
    }
    return g;
}

// What shading one path at one bounce emits: its next extension ray (or a
// terminating entry that brings it back once more for its pending NEE) and
// up to three shadow rays (area light, environment, point light).
struct ShadeOut {
    bool emitRay = false, termEntry = false, emitA = false, emitE = false, emitP = false;
    V3 newO, newD, sOA, sDA, sOE, sDE, sOP, sDP;
    float tA = 0.f, tP = 0.f;
    uint32_t skipped = 0;  // NEE terms exactly zero: their shadow rays are not traced
};

// An NEE term that is exactly zero in every channel (the BSDF vanishes for
// the light direction, e.g. below the surface) cannot change the sample: the
// reference adds it to L only when its shadow ray is unoccluded, and L + 0 is
// L (L starts at +0 and is never -0). Such a term's shadow ray is not traced
// and the term is not stored. A NaN term compares unequal to zero and takes
// its shadow ray as before.
PRR_DEV bool zeroTerm(const V3 &c) { return c.x == 0.f && c.y == 0.f && c.z == 0.f; }

// The ray a path traced and its closest hit (t, u, v, prim bits; prim -1 for
// a miss or a terminating entry), as read from the queue entry and the hit
// record, or as a kernel holds them.
struct HeldRay {
    float4 h;
    V3 o, d;
    PRR_DEV float4 hit() const { return h; }
    PRR_DEV V3 org() const { return o; }
    PRR_DEV V3 dir() const { return d; }
};

// Shading of the path in `slot` at `bounce` (Integrator.cpp:104-263, one
// loop iteration) for the ray and hit `ray` gives.
// PHOTON / CACHE: whether the caustic photon gather and the radiance-cache
// query / training code are compiled in (the kernel's registers are those of
// its heaviest branch, so renders without them use a leaner variant).
// KIND: kShadeAny any entry; kShadeLight only entries whose hit is not a
// diffuse surface (miss, terminating entry, emitter, glass, mirror: the
// diffuse vertex code is not compiled in); kShadeDiffuse only entries whose
// hit is a diffuse surface.
constexpr int kShadeAny = 0, kShadeLight = 1, kShadeDiffuse = 2;

template <bool PHOTON, bool CACHE, int KIND, class Ray>
__device__ __forceinline__ void shadePath(const SceneView &s, const PathBufs &pb, const ChunkInfo &ci, int bounce,
                                          uint32_t slot, const Ray &ray, ShadeOut &so) {
    bool done = false;
    const float4 h = ray.hit();
    const int prim = __float_as_int(h.w);
    const V3 o = ray.org(), d = ray.dir();
    // Every load that depends only on the slot and the hit is issued here,
    // together, ahead of the branches that use it: the path state, the hit
    // primitive's row and, past bounce 0, the pending NEE terms of the kinds
    // the scene has and the BSDF factor (each would otherwise wait for the
    // flags, one dependent round trip more).
    const float4 pn = prim >= 0 ? s.prim_n[prim] : make_float4(0.f, 0.f, 0.f, 0.f);
    const int sIdx = ci.sampleOf ? (int) ci.sampleOf[slot] : ci.sampleBase + (int) (slot / (uint32_t) ci.pixels);
    uint32_t fl;
    V3 L, beta;
    float bpdf = 0.f;  // BSDF pdf of the ray a diffuse vertex sampled (F_PREV_DIFF)
    if (bounce == 0) {
        // a path's first bounce starts from L = 0, beta = 1 (Integrator.cpp:
        // castRay's locals); the camera-ray kernels store only the ray and
        // the random-number state
        fl = pathFlags0<CACHE>(ci, slot, (uint32_t) sIdx);
        L = V3(0.f);
        beta = V3(1.f);
    } else {
        const float4 l4 = pb.L[slot];
        L = ld3(l4);
        fl = __float_as_uint(l4.w);
        const float4 b4 = pb.beta[slot];
        beta = ld3(b4);
        bpdf = b4.w;
    }
    Pcg rng;
    rng.state = pb.rng[slot];
    rng.inc = pcgInc((uint32_t) sIdx);
    float4 cA = make_float4(0.f, 0.f, 0.f, 0.f), cE = cA, cP = cA, cC = cA;
    uint8_t vA = 0, vE = 0, vP = 0;
    if (bounce > 0) {
        if (s.total_emit_area > 0.f) {
            cA = pb.pend[slot];
            vA = pb.vis[slot];
        }
        if (s.env_w > 0) {
            cE = pb.pend[pb.slots + slot];
            vE = pb.vis[pb.slots + slot];
        }
        if (s.num_points > 0) {
            cP = pb.pend[2 * pb.slots + slot];
            vP = pb.vis[2 * pb.slots + slot];
        }
        if (PHOTON && s.ph_count > 0) cC = pb.pend[3 * pb.slots + slot];
    }
    // 1. NEE of the previous bounce, in reference order
    if ((fl & F_PEND_A) && vA) L += ld3(cA);
    if ((fl & F_PEND_E) && vE) L += ld3(cE);
    if ((fl & F_PEND_P) && vP) L += ld3(cP);
    if (fl & F_PEND_C) L += ld3(cC);
    fl &= ~F_PEND_ALL;
    done = (fl & F_TERMINATING) != 0u;
    // 2. the BSDF-sampled ray of the previous (diffuse) bounce; the stored
    // throughput is already beta * betaScale (the product every use below
    // starts with, Integrator.cpp:217-246)
    if (!done && (fl & F_PREV_DIFF)) {
        const bool ind = bounce - 1 > 0;
        if (prim < 0) {  // Integrator.cpp:217-229
            if (s.env_w > 0) {
                const V3 envL = envLookup(s, d);
                const float w = mis(bpdf, envPdf(s, d));
                const V3 c = beta * w * envL;
                L += ind ? clampInd(c, ci.clampInd) : c;
            } else {
                const V3 c = beta * V3(s.background[0], s.background[1], s.background[2]);
                L += ind ? clampInd(c, ci.clampInd) : c;
            }
            done = true;
        } else {
            const HitGeom g = hitGeom(s, prim, pn, h.x, o, d, h.y, h.z);
            const DMaterial hm = s.mats[g.mat];
            if ((int) hm.base_type.w == MAT_EMISSIVE) {  // Integrator.cpp:231-242
                const float cosT = smax(0.f, dot(-d, normalized(g.Ng)));
                // light pdf of the hit point (0 without area emitters: the
                // diffuse vertex is then not stored)
                float lpSA = 0.f;
                if (s.total_emit_area > 0.f) {
                    const V3 dd = g.P - ld3(pb.prevP[slot]);
                    const float hd2 = dot(dd, dd);
                    lpSA = (1.f / s.total_emit_area) * hd2 / smax(cosT, 1e-4f);
                }
                const float w = mis(bpdf, lpSA);
                const V3 c = beta * w * ld3(hm.emission);
                L += ind ? clampInd(c, ci.clampInd) : c;
                done = true;
            } else if (!finite3(beta)) {  // beta = beta * betaScale (Integrator.cpp:244-246)
                done = true;
            }
        }
        if (!done && bounce - 1 >= 3) {  // Russian roulette (Integrator.cpp:250-256)
            const float mc = max3(beta);
            const float qq = smax(0.05f, 1.f - mc);
            if (rng.next() < qq)
                done = true;
            else
                beta = beta * (1.f / (1.f - qq));
        }
        fl &= ~F_PREV_DIFF;
    }
    if (!done && bounce >= ci.maxDepth) done = true;
    // 3. this bounce (Integrator.cpp:124-248)
    if (!done) {
        const bool spec = (fl & F_SPEC) != 0u;
        if (bounce == 0 || spec) {
            if (prim < 0) {
                const V3 bg = s.env_w > 0 ? envLookup(s, d)
                                          : V3(s.background[0], s.background[1], s.background[2]);
                const V3 c = beta * bg;
                L += bounce > 0 ? clampInd(c, ci.clampInd) : c;
            }
        }
        if (prim < 0) {
            done = true;
        } else {
            const HitGeom g = hitGeom(s, prim, pn, h.x, o, d, h.y, h.z);
            const DMaterial m = s.mats[g.mat];
            const int mtype = (int) m.base_type.w;
            const V3 base = ld3(m.base_type);
            // Material::buildShadingData (Material.cpp:3-26): constants, or the textures
            ShadeData sd{base, m.params.x, m.params.y, g.Ng};
            if (mtype == MAT_EMISSIVE) {
                if (bounce == 0 || spec) {
                    const V3 c = beta * ld3(m.emission);
                    L += bounce > 0 ? clampInd(c, ci.clampInd) : c;
                }
                // scatters nothing (EmissiveMaterial: f = pdf = 0)
                done = true;
            } else if (mtype == MAT_GLASS || mtype == MAT_MIRROR) {
                if (m.params.w != 0.f) texturedShading(s, g.mat, prim, g.flags, h.y, h.z, sd);
                const BsdfSample bs = mtype == MAT_GLASS ? sampleGlass(-d, sd.N, m.params.z, rng.next())
                                                         : sampleMirror(-d, sd.N, base);
                beta = beta * bs.f / bs.pdf;
                const bool trans = dot(bs.wi, g.Ng) < 0.f;
                const V3 off = trans ? -g.Ng : g.Ng;
                if (trans) beta = beta * sd.base;  // Integrator.cpp:161-165
                fl |= F_SPEC;
                if (bounce + 1 >= ci.maxDepth) {
                    done = true;  // the loop ends before this ray is used
                } else {
                    so.emitRay = true;
                    so.newO = g.P + off * kEps;
                    so.newD = bs.wi;
                }
            } else if constexpr (KIND != kShadeLight) {
                fl &= ~F_SPEC;
                sd.N = dot(d, g.Ng) < 0.f ? g.Ng : -g.Ng;
                if (m.params.w != 0.f) texturedShading(s, g.mat, prim, g.flags, h.y, h.z, sd);
                const V3 sN = sd.N;
                const V3 wo = -d;
                const CacheRec &cr = ci.cache;
                bool cacheQuery = false;
                if (CACHE && (cr.query | cr.train)) {
                    const int stopAt = (fl & F_TRAIN) ? cr.trainStop : cr.queryBounce;
                    if (cr.query && stopAt > 0 && bounce >= stopAt) {
                        // the path ends here; the cache adds beta * L_out(P, wo) (for a
                        // training path, also to the targets of its recorded vertices)
                        cacheQuery = true;
                        done = true;
                        const uint32_t k = waveAppend(cr.qCount, true);
                        putQuery(cr.q + (int64_t) k * kCacheQueryFloats, g.P, wo, sN, sd.rough, sd.base);
                        cr.qSlot[k] = slot;
                        cr.qBeta[k] = st4(beta, 0.f);
                    } else if ((fl & F_TRAIN) && bounce >= 1 && min3(beta) > 1e-3f) {
                        // target later: (final L - L so far) / beta
                        const uint32_t k = waveAppend(cr.tCount, true);
                        if (k < cr.tCap) {
                            putQuery(cr.tq + (int64_t) k * kCacheQueryFloats, g.P, wo, sN, sd.rough, sd.base);
                            cr.tSlot[k] = slot;
                            cr.tBeta[k] = st4(beta, 0.f);
                            cr.tLpre[k] = st4(L, 0.f);
                        }
                    }
                }
                if (!cacheQuery) {
                const Frame fr = makeFrame(sN);
                const Lobes lb = makeLobes(sd.base, sd.rough, sd.metal, m.params.z);
                const bool ind = bounce > 0;
                // NEE area (Integrator.cpp:14-34, 78-87)
                if (s.total_emit_area > 0.f) {
                    V3 lp, ln;
                    int lm;
                    sampleLight(s, rng, lp, ln, lm);
                    const V3 toL = lp - g.P;
                    const float d2 = dot(toL, toL);
                    const float dist = sqrtf(d2);
                    const V3 wi = toL / dist;
                    const float cosL = smax(0.f, dot(-wi, normalized(ln)));
                    if (cosL > 0.f) {
                        const float pdfSA = (1.0f / s.total_emit_area) * d2 / cosL;
                        V3 f;
                        float bp;
                        evalPdf(lb, fr, wi, wo, f, bp);
                        const float w = mis(pdfSA, bp);
                        V3 c = beta * (w * ld3(s.mats[lm].emission) * f / pdfSA);
                        c = ind ? clampInd(c, ci.clampInd) : c;
                        if (zeroTerm(c)) {
                            so.skipped++;
                        } else {
                            pb.pend[slot] = st4(c, 0.f);
                            fl |= F_PEND_A;
                            so.emitA = true;
                            so.sOA = g.P + wi * kEps;
                            so.sDA = wi;
                            so.tA = dist - 2.f * kEps;
                        }
                    }
                }
                // NEE environment (Integrator.cpp:39-66, 93-102)
                if (s.env_w > 0) {
                    const float r1 = rng.next(), r2 = rng.next();
                    float pe;
                    const V3 dir = envImportanceSample(s, r1, r2, pe);
                    if (!(pe < 1e-6f)) {
                        const V3 Le = envLookup(s, dir);
                        V3 f;
                        float bp;
                        evalPdf(lb, fr, dir, wo, f, bp);
                        const float w = mis(pe, bp);
                        V3 c = beta * (w * Le * f / (pe + 1e-6f));
                        c = ind ? clampInd(c, ci.clampInd) : c;
                        if (zeroTerm(c)) {
                            so.skipped++;
                        } else {
                            pb.pend[pb.slots + slot] = st4(c, 0.f);
                            fl |= F_PEND_E;
                            so.emitE = true;
                            so.sOE = g.P + dir * kEps;
                            so.sDE = dir;
                        }
                    }
                }
                // NEE point light (build extension; no draws)
                if (s.num_points > 0) {
                    const V3 pp(s.point_pos[0], s.point_pos[1], s.point_pos[2]);
                    const V3 toL = pp - g.P;
                    const float d2 = dot(toL, toL);
                    const float dist = sqrtf(d2);
                    const V3 wi = toL / dist;
                    V3 f;
                    float bp;
                    evalPdf(lb, fr, wi, wo, f, bp);
                    V3 c = beta * (V3(s.point_int[0], s.point_int[1], s.point_int[2]) * f / d2);
                    c = ind ? clampInd(c, ci.clampInd) : c;
                    if (zeroTerm(c)) {
                        so.skipped++;
                    } else {
                        pb.pend[2 * pb.slots + slot] = st4(c, 0.f);
                        fl |= F_PEND_P;
                        so.emitP = true;
                        so.sOP = g.P + wi * kEps;
                        so.sDP = wi;
                        so.tP = dist - 2.f * kEps;
                    }
                }
                // caustic photon estimate at camera / delta-reached vertices
                // (build extension; no draws)
                if (PHOTON && s.ph_count > 0 && (bounce == 0 || spec)) {
                    V3 c = beta * photonRadiance(s, g.P, wo, fr, lb);
                    c = ind ? clampInd(c, ci.clampInd) : c;
                    pb.pend[3 * pb.slots + slot] = st4(c, 0.f);
                    fl |= F_PEND_C;
                }
                // BSDF sampling (Integrator.cpp:207-215)
                const BsdfSample bs = sampleDiffuse(lb, fr, wo, rng);
                if (bs.pdf < 1e-6f) {
                    done = true;
                } else {
                    const V3 bscale = bs.f / bs.pdf;
                    const V3 off = dot(bs.wi, sN) >= 0.f ? sN : -sN;
                    // the next bounce's throughput, beta * betaScale, and the pdf
                    beta = beta * bscale;
                    bpdf = bs.pdf;
                    if (s.total_emit_area > 0.f) pb.prevP[slot] = st4(g.P, 0.f);
                    fl |= F_PREV_DIFF;
                    so.emitRay = true;
                    so.newO = g.P + off * kEps;
                    so.newD = bs.wi;
                }
                }  // !cacheQuery
            }
        }
    }
    // a path that ends with NEE still in flight comes back once more
    if (done && (fl & F_PEND_ALL) && !(fl & F_TERMINATING)) {
        fl |= F_TERMINATING;
        so.emitRay = true;
        done = false;
    }
    if (done) {
        // Integrator.cpp:259-261; w = 1 marks the replaced sample (counted at accumulation)
        const bool fin = finite3(L);
        pb.result[slot] = st4(fin ? L : V3(0.f), fin ? 0.f : 1.f);
    } else {
        pb.L[slot] = st4(L, __uint_as_float(fl));
        pb.beta[slot] = st4(beta, bpdf);
        pb.rng[slot] = rng.state;
    }
    so.termEntry = (fl & F_TERMINATING) != 0u;
}

// Shades the entries of bounce's queue (list == null), or the count[0]
// entries list[0..) of it (a shading bin).
template <bool PHOTON, bool CACHE, int KIND>
__device__ __forceinline__ void shadeImpl(const SceneView &s, const PathBufs &pb, const ChunkInfo &ci, int bounce,
                                          const uint32_t *list, const uint32_t *count) {
    __shared__ uint32_t spre[kQSegs + 1];
    const int q = bounce & 1;
    const int64_t cnt = list ? (int64_t) *count : (int64_t) pb.qcnt[bounce];
    // at most tailPaths entries: kFinish runs this bounce's paths (an empty loop here)
    const int64_t n = pb.qcnt[bounce] > ci.tailPaths ? cnt : 0;
    if (!list) loadPrefix(spre, pb.qpre + (int64_t) bounce * (kQSegs + 1));
    const int64_t stride = (int64_t) gridDim.x * kShadeBlock;
    const int64_t i0 = (int64_t) blockIdx.x * kShadeBlock + threadIdx.x;
    // uniform trip count per wave so every lane reaches the appends
    const int64_t waveBase = i0 - (threadIdx.x & 63);
    // the wave's query counts (wave-uniform: sums of ballot popcounts)
    uint32_t nRays = 0, nShadow = 0, nSkip = 0;
    for (int64_t wb = waveBase; wb < n; wb += stride) {
        const int64_t i = wb + (threadIdx.x & 63);
        uint32_t slot = 0;
        ShadeOut so;
        if (i < n) {
            const int64_t e = list ? (int64_t) list[i] : segEntry(spre, pb.qcap, i);
            // queue 0 holds the camera rays in slot order (entry e is slot e)
            slot = bounce == 0 ? (uint32_t) e : pb.qSlot[q][e];
            const HeldRay ray{pb.hit[e], ldv3(pb.qOrg[q], e), ldv3(pb.qDir[q], e)};
            shadePath<PHOTON, CACHE, KIND>(s, pb, ci, bounce, slot, ray, so);
        }
        const bool emitRay = so.emitRay, termEntry = so.termEntry, emitA = so.emitA, emitE = so.emitE,
                   emitP = so.emitP;
        // queue appends (all lanes of the wave take part): the wave's rays
        // and shadow rays (area, env, point) go to the segment of its input
        // position, reserved with one atomic on the segment's counter pair
        {
            const int seg = (int) ((wb >> 6) % kQSegs);
            const uint64_t mR = __ballot(emitRay);
            const uint64_t mA = __ballot(emitA), mE = __ballot(emitE), mP = __ballot(emitP);
            const uint32_t nR = (uint32_t) __popcll(mR);
            const uint32_t nA = (uint32_t) __popcll(mA), nE = (uint32_t) __popcll(mE);
            const uint32_t total = nA + nE + (uint32_t) __popcll(mP);
            nShadow += total;
            unsigned long long pair = 0;
            if (nR | total) {
                if (laneId() == 0)
                    pair = atomicAdd(&pb.segc[(int64_t) bounce * kQSegs + seg],
                                     ((unsigned long long) total << 32) | nR);
                pair = __shfl(pair, 0);
            }
            const uint64_t below = (1ull << laneId()) - 1ull;
            const int64_t qi = (int64_t) (uint32_t) pair + __popcll(mR & below);
            if (emitRay && qi < pb.qcap) {
                const int64_t e = (int64_t) seg * pb.qcap + qi;
                pb.qSlot[q ^ 1][e] = slot;
                stv3(pb.qOrg[q ^ 1], e, termEntry ? noRayOrigin() : so.newO);
                stv3(pb.qDir[q ^ 1], e, termEntry ? V3(0.f) : so.newD);
            }
            const uint32_t base = (uint32_t) (pair >> 32);
            const int64_t s0 = (int64_t) seg * pb.sqcap;
            if (emitA) {
                const int64_t ia = (int64_t) base + __popcll(mA & below);
                if (ia < pb.sqcap) {
                    pb.sqSlot[s0 + ia] = slot;
                    pb.sqOrg[s0 + ia] = st4(so.sOA, so.tA);
                    stv3(pb.sqDir, s0 + ia, so.sDA);
                }
            }
            if (emitE) {
                const int64_t ie = (int64_t) base + nA + __popcll(mE & below);
                if (ie < pb.sqcap) {
                    pb.sqSlot[s0 + ie] = slot | (1u << 30);
                    pb.sqOrg[s0 + ie] = st4(so.sOE, __builtin_inff());
                    stv3(pb.sqDir, s0 + ie, so.sDE);
                }
            }
            if (emitP) {
                const int64_t ip = (int64_t) base + nA + nE + __popcll(mP & below);
                if (ip < pb.sqcap) {
                    pb.sqSlot[s0 + ip] = slot | (2u << 30);
                    pb.sqOrg[s0 + ip] = st4(so.sOP, so.tP);
                    stv3(pb.sqDir, s0 + ip, so.sDP);
                }
            }
        }
        nRays += (uint32_t) __popcll(__ballot(emitRay && !termEntry));
        const uint32_t sk = so.skipped;
        nSkip += (uint32_t) (__popcll(__ballot(sk >= 1u)) + __popcll(__ballot(sk >= 2u)) + __popcll(__ballot(sk >= 3u)));
    }
    // query counts: one atomic per wave and launch (contended device-scope
    // atomics per wave-iteration would serialise the kernel)
    if (laneId() == 0) {
        if (nRays) atomicAdd(&pb.stats[0], (unsigned long long) nRays);
        if (nShadow) atomicAdd(&pb.stats[1], (unsigned long long) nShadow);
        if (nSkip) atomicAdd(&pb.stats[7], (unsigned long long) nSkip);
    }
}

// kShade: one shading pass, compiled for at least W waves per SIMD (fewer
// registers, spilled if need be, for more latency hiding).
template <int W, bool PHOTON, bool CACHE>
__global__ __launch_bounds__(kShadeBlock, W) void kShade(SceneView s, PathBufs pb, ChunkInfo ci, int bounce) {
    shadeImpl<PHOTON, CACHE, kShadeAny>(s, pb, ci, bounce, nullptr, nullptr);
}

// Shading sorted by material (the hit's material class): the bounce's queue
// is split into the entries whose hit is a diffuse surface and the rest, and
// kShadeLightList / kShadeDiffuseList shade the two lists. The light kernel
// holds none of the diffuse vertex's code (NEE, BSDF sampling, photon
// gather, cache query) and runs at more waves per SIMD; the diffuse kernel's
// waves no longer wait on lanes that escaped or hit glass. Every entry is
// shaded by the same code as in kShade, so images do not change.
//
// The split is a counting sort without global atomics (appends through one
// counter per list serialise on it when every wave of a light kernel hits
// it): block b takes chunk b of the queue; kShadeBinCount counts each
// block's entries per list, kShadeBinScan turns the counts into offsets,
// kShadeBinScatter writes each block's entries in chunk order (wave
// ballots + prefix sums within the block).
PRR_DEV bool diffuseHit(const SceneView &s, const PathBufs &pb, int q, int64_t i) {
    // an entry without a ray has a miss record (prim -1)
    const int prim = __float_as_int(pb.hit[i].w);
    if (prim < 0) return false;
    const int mat = __float_as_int(s.prim_n[prim].w) & kPrimMatMask;
    return (int) s.mats[mat].base_type.w == MAT_DIFFUSE;
}

PRR_DEV void shadeChunk(int64_t n, int b, int nb, int64_t &c0, int64_t &c1) {
    const int64_t per = ((n + nb - 1) / nb + 63) & ~(int64_t) 63;  // whole waves
    c0 = (int64_t) b * per;
    c0 = c0 < n ? c0 : n;
    c1 = c0 + per < n ? c0 + per : n;
}

// per block: [0] entries of the light list, [1] of the diffuse list
__global__ __launch_bounds__(kShadeBlock) void kShadeBinCount(SceneView s, PathBufs pb, ChunkInfo ci, int bounce,
                                                              uint32_t *blockCounts) {
    __shared__ uint32_t part[2][kShadeBlock / 64];
    __shared__ uint32_t spre[kQSegs + 1];
    const int q = bounce & 1;
    const int64_t cnt = pb.qcnt[bounce];
    const int64_t n = cnt > ci.tailPaths ? cnt : 0;
    loadPrefix(spre, pb.qpre + (int64_t) bounce * (kQSegs + 1));
    int64_t c0, c1;
    shadeChunk(n, blockIdx.x, gridDim.x, c0, c1);
    uint32_t nl = 0, nd = 0;
    for (int64_t i = c0 + threadIdx.x; i < c1; i += kShadeBlock) {
        const bool d = diffuseHit(s, pb, q, segEntry(spre, pb.qcap, i));
        nd += d ? 1u : 0u;
        nl += d ? 0u : 1u;
    }
    for (int o = 32; o > 0; o >>= 1) {
        nl += __shfl_xor(nl, o);
        nd += __shfl_xor(nd, o);
    }
    if (laneId() == 0) {
        part[0][threadIdx.x >> 6] = nl;
        part[1][threadIdx.x >> 6] = nd;
    }
    __syncthreads();
    if (threadIdx.x < 2) {
        uint32_t t = 0;
        for (int w = 0; w < kShadeBlock / 64; w++) t += part[threadIdx.x][w];
        blockCounts[(size_t) threadIdx.x * gridDim.x + blockIdx.x] = t;
    }
}

// Exclusive scan of the 2 * nb block counts (light blocks, then diffuse
// blocks) into offsets within each list; counts[0..1] = list sizes.
__global__ __launch_bounds__(1024) void kShadeBinScan(uint32_t *blockCounts, int nb, uint32_t *counts) {
    __shared__ uint32_t part[1024];
    for (int list = 0; list < 2; list++) {
        uint32_t *c = blockCounts + (size_t) list * nb;
        const int t = threadIdx.x;
        const int per = (nb + 1023) / 1024;
        const int a = t * per, e = a + per < nb ? a + per : nb;
        uint32_t sum = 0;
        for (int i = a; i < e; i++) sum += c[i];
        part[t] = sum;
        __syncthreads();
        for (int off = 1; off < 1024; off <<= 1) {
            const uint32_t v = t >= off ? part[t - off] : 0u;
            __syncthreads();
            part[t] += v;
            __syncthreads();
        }
        uint32_t run = part[t] - sum;
        for (int i = a; i < e; i++) {
            const uint32_t x = c[i];
            c[i] = run;
            run += x;
        }
        if (t == 1023) counts[list] = part[1023];
        __syncthreads();
    }
}

__global__ __launch_bounds__(kShadeBlock) void kShadeBinScatter(SceneView s, PathBufs pb, ChunkInfo ci, int bounce,
                                                                const uint32_t *blockCounts, uint32_t *lists,
                                                                int64_t cap) {
    __shared__ uint32_t wcnt[2][kShadeBlock / 64];
    __shared__ uint32_t spre[kQSegs + 1];
    const int q = bounce & 1;
    const int64_t cnt = pb.qcnt[bounce];
    const int64_t n = cnt > ci.tailPaths ? cnt : 0;
    loadPrefix(spre, pb.qpre + (int64_t) bounce * (kQSegs + 1));
    int64_t c0, c1;
    shadeChunk(n, blockIdx.x, gridDim.x, c0, c1);
    uint32_t base[2] = {blockCounts[blockIdx.x], blockCounts[(size_t) gridDim.x + blockIdx.x]};
    const int w = threadIdx.x >> 6;
    const uint64_t below = (1ull << laneId()) - 1ull;
    // block-uniform trip count (the barrier below)
    for (int64_t bb = c0; bb < c1; bb += kShadeBlock) {
        const int64_t i = bb + threadIdx.x;
        const bool valid = i < c1;
        const int64_t e = valid ? segEntry(spre, pb.qcap, i) : 0;
        const bool d = valid && diffuseHit(s, pb, q, e);
        const uint64_t md = __ballot(d), ml = __ballot(valid && !d);
        if (laneId() == 0) {
            wcnt[0][w] = (uint32_t) __popcll(ml);
            wcnt[1][w] = (uint32_t) __popcll(md);
        }
        __syncthreads();
        uint32_t offl = base[0], offd = base[1], totl = 0, totd = 0;
        for (int k = 0; k < kShadeBlock / 64; k++) {
            const uint32_t a = wcnt[0][k], b = wcnt[1][k];
            offl += k < w ? a : 0u;
            offd += k < w ? b : 0u;
            totl += a;
            totd += b;
        }
        __syncthreads();
        if (valid) {
            if (d)
                lists[cap + offd + (uint32_t) __popcll(md & below)] = (uint32_t) e;
            else
                lists[offl + (uint32_t) __popcll(ml & below)] = (uint32_t) e;
        }
        base[0] += totl;
        base[1] += totd;
    }
}

template <int W, bool PHOTON, bool CACHE>
__global__ __launch_bounds__(kShadeBlock, W) void kShadeLightList(SceneView s, PathBufs pb, ChunkInfo ci, int bounce,
                                                                  const uint32_t *lists, const uint32_t *counts) {
    shadeImpl<PHOTON, CACHE, kShadeLight>(s, pb, ci, bounce, lists, counts);
}
template <int W, bool PHOTON, bool CACHE>
__global__ __launch_bounds__(kShadeBlock, W) void kShadeDiffuseList(SceneView s, PathBufs pb, ChunkInfo ci, int bounce,
                                                                    const uint32_t *lists, int64_t cap,
                                                                    const uint32_t *counts) {
    shadeImpl<PHOTON, CACHE, kShadeDiffuse>(s, pb, ci, bounce, lists + cap, counts + 1);
}

// ---------------------------------------------------------------------------
// Path completion. Each bounce of the wavefront costs three launches, and
// each launch lasts as long as its slowest ray; once few paths remain, those
// tails cost more than the work. kFinish takes the entries of bounce b0's
// queue (the shadow rays of bounce b0 - 1 already resolved) and runs each
// path to its end in one lane: closest hit, shading, then the shadow rays
// shading emitted, traced at once instead of in the next launch. A path
// meets the same traversals and the same shading code, bounce by bounce, as
// in the wavefront kernels, so every slot ends with the same result.

// IntersectP of one shadow ray on the chosen tree (0 the binary tree, 1 the
// four-wide occlusion tree, 2 its quantised copy), traced to its end.
PRR_DEV bool occluded(const SceneView &s, const V3 &o, const V3 &d, float tmax, int tree, int cap, LaneStack &lst,
                      RefStack &rst) {
    if (tree == 2) return traceAny4Q(s, o, d, tmax, rst);
    if (tree == 1) return traceAny4(s, o, d, tmax, rst);
    TravState ts;
    travBegin(ts, o, d, tmax, s.root, lst);
    if (!anyBegin(s, ts)) return false;
    int r;
    while ((r = anyIter(s, ts, lst, cap)) == kAnyGo) {
    }
    return r == kAnyHit;
}

// Compiled for 3 waves per SIMD (at most 168 VGPRs): latency hiding for the
// traversal loads matters more here than registers for shading.
template <bool PHOTON, bool CACHE>
__global__ __launch_bounds__(kTraceBlock, 3) void kFinish(SceneView s, PathBufs pb, ChunkInfo ci, int bounce0,
                                                       uint2 *spill, int stackDepth, int wwCap, int tree,
                                                       int64_t residentWaves) {
    extern __shared__ uint2 stk[];
    __shared__ uint32_t spre[kQSegs + 1];
    const int q = bounce0 & 1;
    const int64_t n = pb.qcnt[bounce0];
    // runs the first bounce whose queue holds at most tailPaths entries
    if (n > ci.tailPaths || (bounce0 > 0 && pb.qcnt[bounce0 - 1] <= ci.tailPaths)) return;
    loadPrefix(spre, pb.qpre + (int64_t) bounce0 * (kQSegs + 1));
    uint32_t *fetch = &pb.fetch[2 * bounce0];
    const int64_t lane = (int64_t) blockIdx.x * kTraceBlock + threadIdx.x;
    // Two LDS stack areas: the closest-hit stack (uint2 entries) and the
    // shadow-ray stack (uint32 entries). Sharing one area would overlap a
    // lane's entries with those of another lane (lane / 2 at twice the
    // entry size), which may belong to a different wave of the block that is
    // traversing at the same time. A lane runs one traversal at a time, so
    // its two stacks share its spill region.
    LaneStack lst = makeStack<LaneStack>(stk, spill + lane * kSpill, threadIdx.x, stackDepth);
    RefStack rst = makeStack<RefStack>(stk + stackDepth * kTraceBlock, spill + lane * kSpill, threadIdx.x, stackDepth);
    WaveRange wr;
    const int64_t waves = (int64_t) gridDim.x * (kTraceBlock / 64);
    uint32_t nRays = 0, nShadow = 0, nTraced = 0, nSkip = 0;
    // Only the first `lanes` lanes of a wave take paths: a wave waits at each
    // bounce for its slowest path, so when few paths remain they are spread
    // over the waves the GPU holds at once, fewer per wave (a power of two,
    // 1..64).
    int lanes = 1;
    while (lanes < 64 && (int64_t) lanes * residentWaves < n) lanes *= 2;
    const bool mine = (int) laneId() < lanes;
    while (mine) {
        // every lane that reaches this point takes one entry (>= n: none left)
        const int64_t w = waveTake(wr, fetch, n, true, lanes, waves, lanes);
        if (w >= n) break;
        const int64_t i = segEntry(spre, pb.qcap, w);
        const uint32_t slot = pb.qSlot[q][i];
        V3 o = ldv3(pb.qOrg[q], i), d = ldv3(pb.qDir[q], i);
        for (int b = bounce0;; b++) {
            float4 h = make_float4(__builtin_inff(), 0.f, 0.f, __int_as_float(-1));
            if (!isnan(o.x)) {  // a NaN origin: no ray (terminating entry)
                TravState ts;
                travBegin(ts, o, d, 0.f, s.root, lst);
                while (!closestIter(s, ts, lst, wwCap)) {
                }
                h = make_float4(ts.h.t, ts.h.u, ts.h.v, __int_as_float(ts.h.prim));
                nTraced++;
            }
            ShadeOut so;
            shadePath<PHOTON, CACHE, kShadeAny>(s, pb, ci, b, slot, HeldRay{h, o, d}, so);
            if (so.emitA) pb.vis[slot] = occluded(s, so.sOA, so.sDA, so.tA, tree, wwCap, lst, rst) ? 0 : 1;
            if (so.emitE)
                pb.vis[pb.slots + slot] =
                    occluded(s, so.sOE, so.sDE, __builtin_inff(), tree, wwCap, lst, rst) ? 0 : 1;
            if (so.emitP) pb.vis[2 * pb.slots + slot] = occluded(s, so.sOP, so.sDP, so.tP, tree, wwCap, lst, rst) ? 0 : 1;
            nRays += (so.emitRay && !so.termEntry) ? 1u : 0u;
            nShadow += (so.emitA ? 1u : 0u) + (so.emitE ? 1u : 0u) + (so.emitP ? 1u : 0u);
            nSkip += so.skipped;
            // the wavefront shades bounces 0..maxDepth, never the queue past it
            if (!so.emitRay || b >= ci.maxDepth) break;
            o = so.termEntry ? noRayOrigin() : so.newO;
            d = so.newD;
        }
    }
    for (int o = 32; o > 0; o >>= 1) {
        nRays += __shfl_xor(nRays, o);
        nShadow += __shfl_xor(nShadow, o);
        nTraced += __shfl_xor(nTraced, o);
        nSkip += __shfl_xor(nSkip, o);
    }
    if (laneId() == 0) {
        if (nRays) atomicAdd(&pb.stats[0], (unsigned long long) nRays);
        if (nShadow) {
            atomicAdd(&pb.stats[1], (unsigned long long) nShadow);
            atomicAdd(&pb.stats[6], (unsigned long long) nShadow);
        }
        if (nTraced) atomicAdd(&pb.stats[4], (unsigned long long) nTraced);
        if (nSkip) atomicAdd(&pb.stats[7], (unsigned long long) nSkip);
    }
}

// Radiance-cache training targets of records [r0, r1): the radiance the
// path gathered after the vertex, per unit throughput.
__global__ void kCacheTargets(PathBufs pb, CacheRec cr, uint32_t r0, uint32_t r1, float *target) {
    const uint32_t r = r0 + blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= r1) return;
    const V3 res = ld3(pb.result[cr.tSlot[r]]);
    const V3 lp = ld3(cr.tLpre[r]);
    const V3 b = ld3(cr.tBeta[r]);
    target[3 * (int64_t) r] = (res.x - lp.x) / b.x;
    target[3 * (int64_t) r + 1] = (res.y - lp.y) / b.y;
    target[3 * (int64_t) r + 2] = (res.z - lp.z) / b.z;
}

// Renderer.cpp:84-99: framebuffer[m] += s in sample order.
__global__ void kAccumulate(PathBufs pb, ChunkInfo ci, float4 *fb) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t nan = 0;
    if (e < ci.pixels) {
        const int p = pixelOf(ci, e);
        V3 acc = ld3(fb[p]);
        for (int k = 0; k < ci.spc; k++) {
            const float4 r = pb.result[(int64_t) k * ci.pixels + e];
            acc += ld3(r);
            nan += r.w != 0.f ? 1u : 0u;
        }
        fb[p] = st4(acc, 0.f);
    }
    countNan(pb, nan);
}

// Welford::update (Renderer.hpp:20-26) on one channel.
PRR_DEV void welfordUpdate(float x, int n, float &mean, float &m2) {
    const float delta = x - mean;
    mean += delta / (float) n;
    m2 += delta * (x - mean);
}
// Welford::ci (Renderer.hpp:33-37): 95% confidence half-width / |mean|.
PRR_DEV float welfordCi(int n, float mean, float m2) {
    if (n < 2 || fabsf(mean) < 1e-6f) return 1.f;
    const float var = m2 / (float) (n - 1);
    return 1.96f * sqrtf(var / (float) n) / fabsf(mean);
}

// Renderer.cpp:74-99 with the adaptive stop rule, for one pixel: samples of
// the chunk are taken in sample order until n >= minSpp and all three
// channel CIs are below thr; later samples of a stopped pixel are discarded.
// Returns how many of the taken samples were replaced (non-finite).
PRR_DEV uint32_t accumulateAdaptive(const PathBufs &pb, const ChunkInfo &ci, float4 *fb, const AdaptiveBufs &ab,
                                    int minSpp, float thr, int e) {
    const int p = pixelOf(ci, e);
    float4 m = ab.mean[p], q = ab.m2[p];
    if (q.w != 0.f) return 0u;
    int n = __float_as_int(m.w);
    V3 acc = ld3(fb[p]);
    bool stop = false;
    uint32_t nan = 0;
    for (int k = 0; k < ci.spc && !stop; k++) {
        const float4 r4 = pb.result[(int64_t) k * ci.pixels + e];
        const V3 sv = ld3(r4);
        nan += r4.w != 0.f ? 1u : 0u;
        acc += sv;
        n++;
        welfordUpdate(sv.x, n, m.x, q.x);
        welfordUpdate(sv.y, n, m.y, q.y);
        welfordUpdate(sv.z, n, m.z, q.z);
        if (n >= minSpp)
            stop = welfordCi(n, m.x, q.x) < thr && welfordCi(n, m.y, q.y) < thr && welfordCi(n, m.z, q.z) < thr;
    }
    fb[p] = st4(acc, 0.f);
    m.w = __int_as_float(n);
    q.w = stop ? 1.f : 0.f;
    ab.mean[p] = m;
    ab.m2[p] = q;
    return nan;
}

__global__ void kAccumulateAdaptive(PathBufs pb, ChunkInfo ci, float4 *fb, AdaptiveBufs ab, int minSpp, float thr) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t nan = 0;
    if (e < ci.pixels) nan = accumulateAdaptive(pb, ci, fb, ab, minSpp, thr, e);
    countNan(pb, nan);
}

// Active pixels (not stopped, fewer than spp samples) in pixel order within
// each wave; one atomic per wave.
__global__ void kCompactActive(AdaptiveBufs ab, int pixels, int spp) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    bool act = false;
    if (p < pixels) act = ab.m2[p].w == 0.f && __float_as_int(ab.mean[p].w) < spp;
    const uint32_t idx = waveAppend(ab.count, act);
    if (act) ab.list[idx] = p;
}

__global__ void kResolveAdaptive(const float4 *fb, const AdaptiveBufs ab, int pixels, float *rgb, int32_t *sppOut) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= pixels) return;
    const int n = __float_as_int(ab.mean[p].w);
    const V3 m = ld3(fb[p]) / (float) n;  // Renderer.cpp:99
    rgb[3 * p] = m.x;
    rgb[3 * p + 1] = m.y;
    rgb[3 * p + 2] = m.z;
    if (sppOut) sppOut[p] = n;
}

__global__ void kResolve(const float4 *fb, int pixels, int spp, float *rgb) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= pixels) return;
    const V3 m = ld3(fb[p]) / (float) spp;  // Renderer.cpp:99
    rgb[3 * p] = m.x;
    rgb[3 * p + 1] = m.y;
    rgb[3 * p + 2] = m.z;
}

// ---------------------------------------------------------------------------
hipError_t launchRaygen(const SceneView &s, const PathBufs &pb, const ChunkInfo &ci, int grid, hipStream_t st) {
    hipLaunchKernelGGL(kRaygen, dim3(grid), dim3(kShadeBlock), 0, st, s, pb, ci);
    return hipGetLastError();
}
hipError_t launchRaygenRays(const PathBufs &pb, int64_t n, const float *org, const float *dir, int64_t first,
                            uint64_t seed, int grid, hipStream_t st) {
    hipLaunchKernelGGL(kRaygenRays, dim3(grid), dim3(kShadeBlock), 0, st, pb, n, org, dir, first, seed);
    return hipGetLastError();
}
hipError_t launchCopyResults(const PathBufs &pb, int64_t n, int64_t first, float *L, hipStream_t st) {
    hipLaunchKernelGGL(kCopyResults, dim3((unsigned) ((n + 255) / 256)), dim3(256), 0, st, pb, n, first, L);
    return hipGetLastError();
}
hipError_t launchExtend(const SceneView &s, const PathBufs &pb, int bounce, uint2 *spill, int traceGrid,
                        int refillAt, int stackDepth, int whileWhile, bool speculate, bool twoLevel, int fetchRun,
                        int64_t tailPaths, hipStream_t st) {
    const int wwCap = whileWhile > 1 ? whileWhile : 0;
    const size_t lds = stackDepth * kTraceBlock * sizeof(uint2);
    if (whileWhile && twoLevel && s.nodes2)
        hipLaunchKernelGGL(kExtend<3>, dim3(traceGrid), dim3(kTraceBlock), lds, st, s, pb, bounce, spill, refillAt,
                           stackDepth, wwCap, fetchRun, tailPaths);
    else if (whileWhile && speculate)
        hipLaunchKernelGGL(kExtend<2>, dim3(traceGrid), dim3(kTraceBlock), lds, st, s, pb, bounce, spill, refillAt,
                           stackDepth, wwCap, fetchRun, tailPaths);
    else if (whileWhile)
        hipLaunchKernelGGL(kExtend<1>, dim3(traceGrid), dim3(kTraceBlock), lds, st, s, pb, bounce, spill, refillAt,
                           stackDepth, wwCap, fetchRun, tailPaths);
    else
        hipLaunchKernelGGL(kExtend<0>, dim3(traceGrid), dim3(kTraceBlock), lds, st, s, pb, bounce, spill, refillAt,
                           stackDepth, wwCap, fetchRun, tailPaths);
    return hipGetLastError();
}
hipError_t launchShadow(const SceneView &s, const PathBufs &pb, int bounce, uint2 *spill, int traceGrid,
                        int refillAt, int stackDepth, int whileWhile, int fetchRun, int tree, bool dense,
                        hipStream_t st) {
    const int wwCap = whileWhile > 1 ? whileWhile : 0;
    const size_t lds = stackDepth * kTraceBlock * sizeof(uint2);
#define PRR_SHADOW(W, V)                                                                                             \
    hipLaunchKernelGGL((kShadow<W, V>), dim3(traceGrid), dim3(kTraceBlock), lds, st, s, pb, bounce, spill, refillAt, \
                       stackDepth, wwCap, fetchRun, dense)
    if (whileWhile && tree == 2)
        PRR_SHADOW(true, 2);
    else if (whileWhile && tree == 1)
        PRR_SHADOW(true, 1);
    else if (whileWhile)
        PRR_SHADOW(true, 0);
    else if (tree == 2)
        PRR_SHADOW(false, 2);
    else if (tree == 1)
        PRR_SHADOW(false, 1);
    else
        PRR_SHADOW(false, 0);
#undef PRR_SHADOW
    return hipGetLastError();
}
hipError_t launchQueueScan(const PathBufs &pb, int bounce, hipStream_t st) {
    hipLaunchKernelGGL(kQueueScan, dim3(2), dim3(64), 0, st, pb, bounce);
    return hipGetLastError();
}
template <bool P, bool C>
static void shadeLaunch(const SceneView &s, const PathBufs &pb, const ChunkInfo &ci, int bounce, int grid, int waves,
                        hipStream_t st) {
    if (waves >= 5)
        hipLaunchKernelGGL((kShade<5, P, C>), dim3(grid), dim3(kShadeBlock), 0, st, s, pb, ci, bounce);
    else if (waves == 4)
        hipLaunchKernelGGL((kShade<4, P, C>), dim3(grid), dim3(kShadeBlock), 0, st, s, pb, ci, bounce);
    else
        hipLaunchKernelGGL((kShade<1, P, C>), dim3(grid), dim3(kShadeBlock), 0, st, s, pb, ci, bounce);
}

template <bool P, bool C>
static void shadeSortedLaunch(const SceneView &s, const PathBufs &pb, const ChunkInfo &ci, int bounce, int grid,
                              int waves, int lightWaves, const ShadeBins &sb, hipStream_t st) {
    hipLaunchKernelGGL(kShadeBinCount, dim3(sb.blocks), dim3(kShadeBlock), 0, st, s, pb, ci, bounce, sb.blockCounts);
    hipLaunchKernelGGL(kShadeBinScan, dim3(1), dim3(1024), 0, st, sb.blockCounts, sb.blocks, sb.counts);
    hipLaunchKernelGGL(kShadeBinScatter, dim3(sb.blocks), dim3(kShadeBlock), 0, st, s, pb, ci, bounce,
                       sb.blockCounts, sb.lists, sb.cap);
    if (lightWaves >= 8)
        hipLaunchKernelGGL((kShadeLightList<8, P, C>), dim3(grid), dim3(kShadeBlock), 0, st, s, pb, ci, bounce,
                           sb.lists, sb.counts);
    else if (lightWaves >= 6)
        hipLaunchKernelGGL((kShadeLightList<6, P, C>), dim3(grid), dim3(kShadeBlock), 0, st, s, pb, ci, bounce,
                           sb.lists, sb.counts);
    else
        hipLaunchKernelGGL((kShadeLightList<4, P, C>), dim3(grid), dim3(kShadeBlock), 0, st, s, pb, ci, bounce,
                           sb.lists, sb.counts);
    if (waves >= 5)
        hipLaunchKernelGGL((kShadeDiffuseList<5, P, C>), dim3(grid), dim3(kShadeBlock), 0, st, s, pb, ci, bounce,
                           sb.lists, sb.cap, sb.counts);
    else
        hipLaunchKernelGGL((kShadeDiffuseList<4, P, C>), dim3(grid), dim3(kShadeBlock), 0, st, s, pb, ci, bounce,
                           sb.lists, sb.cap, sb.counts);
}

hipError_t launchShadeSorted(const SceneView &s, const PathBufs &pb, const ChunkInfo &ci, int bounce, int grid,
                             int waves, int lightWaves, const ShadeBins &sb, hipStream_t st) {
    const bool photon = s.ph_count > 0, cache = ci.cache.query || ci.cache.train;
    if (photon)
        cache ? shadeSortedLaunch<true, true>(s, pb, ci, bounce, grid, waves, lightWaves, sb, st)
              : shadeSortedLaunch<true, false>(s, pb, ci, bounce, grid, waves, lightWaves, sb, st);
    else
        cache ? shadeSortedLaunch<false, true>(s, pb, ci, bounce, grid, waves, lightWaves, sb, st)
              : shadeSortedLaunch<false, false>(s, pb, ci, bounce, grid, waves, lightWaves, sb, st);
    return hipGetLastError();
}

hipError_t launchShade(const SceneView &s, const PathBufs &pb, const ChunkInfo &ci, int bounce, int grid, int waves,
                       hipStream_t st) {
    const bool photon = s.ph_count > 0, cache = ci.cache.query || ci.cache.train;
    if (photon)
        cache ? shadeLaunch<true, true>(s, pb, ci, bounce, grid, waves, st)
              : shadeLaunch<true, false>(s, pb, ci, bounce, grid, waves, st);
    else
        cache ? shadeLaunch<false, true>(s, pb, ci, bounce, grid, waves, st)
              : shadeLaunch<false, false>(s, pb, ci, bounce, grid, waves, st);
    return hipGetLastError();
}
// Waves of kFinish the GPU holds at once.
static int64_t finishResidentWaves(const void *kernel, size_t lds, int numCU) {
    int perCU = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCU, kernel, kTraceBlock, lds) != hipSuccess || perCU < 1)
        perCU = 1;
    return (int64_t) perCU * (kTraceBlock / 64) * (numCU > 0 ? numCU : 1);
}
template <bool P, bool C>
static void finishLaunch(const SceneView &s, const PathBufs &pb, const ChunkInfo &ci, int bounce0, uint2 *spill,
                         int grid, int stackDepth, int wwCap, int tree, int numCU, hipStream_t st) {
    // two stack areas (closest-hit and shadow rays)
    const size_t lds = 2 * stackDepth * kTraceBlock * sizeof(uint2);
    const int64_t resident = finishResidentWaves(reinterpret_cast<const void *>(&kFinish<P, C>), lds, numCU);
    // persistent: the blocks the GPU holds at once (fewer to start when the launch has nothing to do)
    const int64_t blocks = resident / (kTraceBlock / 64);
    grid = blocks < grid ? (int) blocks : grid;
    hipLaunchKernelGGL((kFinish<P, C>), dim3(grid), dim3(kTraceBlock), lds, st, s, pb, ci, bounce0, spill, stackDepth,
                       wwCap, tree, resident);
}
hipError_t launchFinish(const SceneView &s, const PathBufs &pb, const ChunkInfo &ci, int bounce0, uint2 *spill,
                        int traceGrid, int stackDepth, int whileWhile, int tree, int numCU, hipStream_t st) {
    const int wwCap = whileWhile > 1 ? whileWhile : 0;
    // two stack areas of 2 KB per LDS entry: at most 64 KB per block
    stackDepth = stackDepth < 16 ? stackDepth : 16;
    const bool photon = s.ph_count > 0, cache = ci.cache.query || ci.cache.train;
#define PRR_FINISH(P, C) finishLaunch<P, C>(s, pb, ci, bounce0, spill, traceGrid, stackDepth, wwCap, tree, numCU, st)
    if (photon)
        cache ? PRR_FINISH(true, true) : PRR_FINISH(true, false);
    else
        cache ? PRR_FINISH(false, true) : PRR_FINISH(false, false);
#undef PRR_FINISH
    return hipGetLastError();
}
hipError_t launchAccumulate(const PathBufs &pb, const ChunkInfo &ci, float4 *fb, hipStream_t st) {
    hipLaunchKernelGGL(kAccumulate, dim3((ci.pixels + 255) / 256), dim3(256), 0, st, pb, ci, fb);
    return hipGetLastError();
}
hipError_t launchAccumulateAdaptive(const PathBufs &pb, const ChunkInfo &ci, float4 *fb, const AdaptiveBufs &ab,
                                    int minSpp, float thr, hipStream_t st) {
    hipLaunchKernelGGL(kAccumulateAdaptive, dim3((ci.pixels + 255) / 256), dim3(256), 0, st, pb, ci, fb, ab, minSpp,
                       thr);
    return hipGetLastError();
}
hipError_t launchCompactActive(const AdaptiveBufs &ab, int pixels, int spp, hipStream_t st) {
    hipLaunchKernelGGL(kCompactActive, dim3((pixels + 255) / 256), dim3(256), 0, st, ab, pixels, spp);
    return hipGetLastError();
}
hipError_t launchResolveAdaptive(const float4 *fb, const AdaptiveBufs &ab, int pixels, float *rgb, int32_t *sppOut,
                                 hipStream_t st) {
    hipLaunchKernelGGL(kResolveAdaptive, dim3((pixels + 255) / 256), dim3(256), 0, st, fb, ab, pixels, rgb, sppOut);
    return hipGetLastError();
}
hipError_t launchRaygenSeg(const PathBufs &pb, const Segment *segs, int nseg, int64_t n, uint32_t *sampleOf, int grid,
                           hipStream_t st) {
    hipLaunchKernelGGL(kRaygenSeg, dim3(grid), dim3(kShadeBlock), 0, st, pb, segs, nseg, n, sampleOf);
    return hipGetLastError();
}
hipError_t launchAccumulateSeg(const PathBufs &pb, const Segment *segs, const int64_t *pixStart, int nseg,
                               int64_t total, hipStream_t st) {
    if (total <= 0) return hipSuccess;
    hipLaunchKernelGGL(kAccumulateSeg, dim3((unsigned) ((total + 255) / 256)), dim3(256), 0, st, pb, segs, pixStart,
                       nseg, total);
    return hipGetLastError();
}
hipError_t launchCacheTargets(const PathBufs &pb, const CacheRec &cr, uint32_t r0, uint32_t r1, float *target,
                              hipStream_t st) {
    if (r1 <= r0) return hipSuccess;
    hipLaunchKernelGGL(kCacheTargets, dim3((r1 - r0 + 255) / 256), dim3(256), 0, st, pb, cr, r0, r1, target);
    return hipGetLastError();
}
hipError_t launchResolve(const float4 *fb, int pixels, int spp, float *rgb, hipStream_t st) {
    hipLaunchKernelGGL(kResolve, dim3((pixels + 255) / 256), dim3(256), 0, st, fb, pixels, spp, rgb);
    return hipGetLastError();
}

}  // namespace prr
