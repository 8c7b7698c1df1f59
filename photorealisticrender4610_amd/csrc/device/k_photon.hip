// k_photon.hip — photon maps on the device. Two rule sets: the build's
// caustic map from the point light (Photon.cpp:96-190 changed: point-light
// emission, delta chains with castRay's outward-normal glass, storage at the
// first diffuse hit after >= 1 delta bounce) and the reference's own rules
// from the area emitters (kPhotonTraceRef). Both are kept in a hash grid
// instead of the reference's unordered_map of grid cells.
//
// 1. kPhotonTrace   one photon per lane (grid-stride), PCG32 stream per
//                   photon index; record written at the photon's index
// 2. kPhotonKeys    bucket hash of the stored photons' cells (cell = 2r)
// 3. radix sort     (hipCUB, stable) of (bucket, photon index): the layout is
//                   deterministic
// 4. kPhotonScatter gather records into bucket order; kPhotonRanges: bucket
//                   start/end
#include <hipcub/hipcub.hpp>

#include "dev_traverse.hpp"
#include "launch_photon.hpp"

namespace prr {

constexpr int kPhBlock = 256;

__global__ __launch_bounds__(kTraceBlock) void kPhotonTrace(SceneView s, int64_t n, uint64_t seed, int maxDepth,
                                                            float4 *pos, float4 *dir, float4 *pw, uint32_t *stored,
                                                            uint2 *spill, int stackDepth) {
    extern __shared__ uint2 stk[];
    const int64_t lane = (int64_t) blockIdx.x * kTraceBlock + threadIdx.x;
    LaneStack st;
    st.lds = stk;
    st.lane = threadIdx.x;
    st.spill = spill + lane * kSpill;
    st.depth = stackDepth;
    st.sp = 0;
    const int64_t stride = (int64_t) gridDim.x * kTraceBlock;
    const V3 lp(s.point_pos[0], s.point_pos[1], s.point_pos[2]);
    const V3 li(s.point_int[0], s.point_int[1], s.point_int[2]);
    for (int64_t i = lane; i < n; i += stride) {
        Pcg rng = Pcg::make(seed, (uint32_t) (i & 0xffffffffu), 0x7000000u + (uint32_t) (i >> 32));
        const float u1 = rng.next(), u2 = rng.next();
        const float z = 1.f - 2.f * u1;
        const float rr = sqrtf(smax(0.f, 1.f - z * z));
        const float phi = 2.f * kPi * u2;
        V3 d(rr * cosf(phi), z, rr * sinf(phi));
        V3 o = lp;
        V3 power = li * (4.f * kPi / (float) n);
        bool caustic = false, keep = false;
        for (int depth = 0; depth <= maxDepth; depth++) {
            const HitOut h = traceClosest(s, o, d, st);
            if (h.prim < 0) break;
            const float4 pn = s.prim_n[h.prim];
            const int bits = __float_as_int(pn.w), pflags = bits >> kPrimFlagShift;
            const DMaterial m = s.mats[bits & kPrimMatMask];
            const int mtype = (int) m.base_type.w;
            V3 P, N;
            if (h.prim < s.num_tris) {
                P = o + h.t * d;
                N = (pflags & kPrimSmooth) ? smoothNormal(s, h.prim, h.u, h.v) : ld3(pn);
            } else {
                const float4 sc = s.sph_by_id[h.prim - s.num_tris];
                P = V3(o + d * h.t);
                N = normalize(V3(P - V3(sc.x, sc.y, sc.z)));
            }
            if (mtype == MAT_EMISSIVE) break;
            if (mtype == MAT_GLASS || mtype == MAT_MIRROR) {
                const V3 base = ld3(m.base_type);
                ShadeData sd{base, m.params.x, m.params.y, N};
                if (m.params.w != 0.f) texturedShading(s, bits & kPrimMatMask, h.prim, pflags, h.u, h.v, sd);
                const BsdfSample bs = mtype == MAT_GLASS ? sampleGlass(-d, sd.N, m.params.z, rng.next())
                                                         : sampleMirror(-d, sd.N, base);
                power = power * bs.f / bs.pdf;
                const bool trans = dot(bs.wi, N) < 0.f;
                if (trans) power = power * sd.base;
                o = P + (trans ? -N : N) * kEps;
                d = bs.wi;
                caustic = true;
                continue;
            }
            if (caustic && depth > 0) {
                pos[i] = st4(P, 1.f);
                dir[i] = st4(d, 0.f);
                pw[i] = st4(power, 0.f);
                keep = true;
            }
            break;
        }
        if (!keep) pos[i] = make_float4(0.f, 0.f, 0.f, 0.f);
        const uint64_t m = __ballot(keep);
        if ((threadIdx.x & 63) == 0 && m) atomicAdd(stored, (uint32_t) __popcll(m));
    }
}

// The reference's photon map rules (PhotonMap::emit's per-photon body,
// Photon.cpp:166-180, and PhotonMap::trace, Photon.cpp:96-152): emission
// from Scene::emissives through Scene::sampleLight, a cosine hemisphere about
// the light normal and power Le * A / N; at a delta vertex the normal flipped
// to face the photon, power *= f (glass therefore always "entered": eta =
// 1 / ior, no tint); every diffuse hit after a delta vertex stores the photon
// (depth > 0), then Russian roulette on max(power) and a diffuse
// continuation with weight f * cos / pdf. One photon per lane with the PCG32
// stream of its index; stored photons are appended (slot from one atomic)
// with key photon * 128 + depth, which orders them deterministically later.
__global__ __launch_bounds__(kTraceBlock) void kPhotonTraceRef(SceneView s, int64_t n, uint64_t seed, int maxDepth,
                                                               float4 *pos, float4 *dir, float4 *pw, uint64_t *key,
                                                               uint32_t cap, uint32_t *count, uint2 *spill,
                                                               int stackDepth) {
    extern __shared__ uint2 stk[];
    const int64_t lane = (int64_t) blockIdx.x * kTraceBlock + threadIdx.x;
    LaneStack st;
    st.lds = stk;
    st.lane = threadIdx.x;
    st.spill = spill + lane * kSpill;
    st.depth = stackDepth;
    st.sp = 0;
    const int64_t stride = (int64_t) gridDim.x * kTraceBlock;
    const float area = 1.0f / (1.0f / s.total_emit_area);  // 1 / pdf, pdf = 1 / totalEmitArea
    for (int64_t i = lane; i < n; i += stride) {
        Pcg rng = Pcg::make(seed, (uint32_t) (i & 0xffffffffu), 0x7000000u + (uint32_t) (i >> 32));
        V3 lp, ln;
        int lm;
        sampleLight(s, rng, lp, ln, lm);
        // sampleCosineHemisphere (BRDFUtils.hpp:66-73)
        const float r1 = rng.next();
        const float r2 = rng.next();
        const float r = sqrtf(r1);
        const float phi = 2.f * kPi * r2;
        V3 d = toWorld(makeFrame(ln), V3(r * cosf(phi), r * sinf(phi), sqrtf(smax(0.f, 1.f - r1))));
        V3 power = ld3(s.mats[lm].emission) * area / (float) n;
        V3 o = lp + ln * kEps;
        bool caustic = false;
        for (int depth = 0; depth <= maxDepth; depth++) {
            const HitOut h = traceClosest(s, o, d, st);
            if (h.prim < 0) break;
            const float4 pn = s.prim_n[h.prim];
            const int bits = __float_as_int(pn.w), pflags = bits >> kPrimFlagShift, mi = bits & kPrimMatMask;
            const DMaterial m = s.mats[mi];
            const int mtype = (int) m.base_type.w;
            if (mtype == MAT_EMISSIVE) break;
            V3 P, N;
            if (h.prim < s.num_tris) {
                P = o + h.t * d;
                N = (pflags & kPrimSmooth) ? smoothNormal(s, h.prim, h.u, h.v) : ld3(pn);
            } else {
                const float4 sc = s.sph_by_id[h.prim - s.num_tris];
                P = V3(o + d * h.t);
                N = normalize(V3(P - V3(sc.x, sc.y, sc.z)));
            }
            const V3 geoN = dot(d, N) < 0.f ? N : -N;
            if (mtype == MAT_GLASS || mtype == MAT_MIRROR) {
                const BsdfSample bs = mtype == MAT_GLASS ? sampleGlass(-d, geoN, m.params.z, rng.next())
                                                         : sampleMirror(-d, geoN, ld3(m.base_type));
                caustic = true;
                power = power * bs.f;
                o = P + (dot(bs.wi, geoN) > 0.f ? geoN : -geoN) * kEps;
                d = bs.wi;
                continue;
            }
            if (depth > 0 && caustic) {
                const uint32_t k = atomicAdd(count, 1u);
                if (k < cap) {
                    pos[k] = st4(P, 1.f);
                    dir[k] = st4(d, 0.f);
                    pw[k] = st4(power, 0.f);
                    key[k] = (uint64_t) i * 128u + (uint64_t) depth;
                }
            }
            float survival = fminf(1.0f, max3(power));
            if (!(rng.next() < survival && depth < maxDepth)) break;
            ShadeData sd{ld3(m.base_type), m.params.x, m.params.y, geoN};
            if (m.params.w != 0.f) texturedShading(s, mi, h.prim, pflags, h.u, h.v, sd);
            const Frame fr = makeFrame(sd.N);
            const Lobes lb = makeLobes(sd.base, sd.rough, sd.metal, m.params.z);
            const BsdfSample bs = sampleDiffuse(lb, fr, -d, rng);
            if (bs.pdf < 1e-6f) break;
            const float cosTheta = smax(0.f, dot(bs.wi, geoN));
            const V3 weight = bs.f * cosTheta / bs.pdf;
            power = power * weight;
            survival = fminf(1.0f, max3(power));
            if (rng.next() > survival) break;
            power = power / survival;
            o = P + geoN * kEps;
            d = bs.wi;
        }
    }
}

// Records [0, m) in key order: out[j] = record vals[j].
__global__ void kPhotonGather(uint32_t m, const uint32_t *vals, const float4 *pos, const float4 *dir,
                              const float4 *pw, const uint64_t *key, float4 *opos, float4 *odir, float4 *opw,
                              uint64_t *okey) {
    const uint32_t j = blockIdx.x * kPhBlock + threadIdx.x;
    if (j >= m) return;
    const uint32_t i = vals[j];
    opos[j] = pos[i];
    odir[j] = dir[i];
    opw[j] = pw[i];
    okey[j] = key[i];
}

__global__ void kPhotonIota(uint32_t m, uint32_t *vals) {
    const uint32_t j = blockIdx.x * kPhBlock + threadIdx.x;
    if (j < m) vals[j] = j;
}

__global__ void kPhotonIrradiance(SceneView s, int64_t n, const float *pos, float *out) {
    const int64_t i = (int64_t) blockIdx.x * kPhBlock + threadIdx.x;
    if (i >= n) return;
    const V3 e = photonIrradiance(s, V3(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]));
    out[3 * i] = e.x;
    out[3 * i + 1] = e.y;
    out[3 * i + 2] = e.z;
}

__global__ void kPhotonKeys(int64_t n, const float4 *pos, float cell, uint32_t mask, uint32_t *keys, uint32_t *vals) {
    const int64_t i = (int64_t) blockIdx.x * kPhBlock + threadIdx.x;
    if (i >= n) return;
    const float4 p = pos[i];
    uint32_t k = 0xFFFFFFFFu;  // unstored photons sort last
    if (p.w > 0.f)
        k = photonHash((int) floorf(p.x / cell), (int) floorf(p.y / cell), (int) floorf(p.z / cell), mask);
    keys[i] = k;
    vals[i] = (uint32_t) i;
}

__global__ void kPhotonScatter(uint32_t m, const uint32_t *vals, const float4 *pos, const float4 *dir,
                               const float4 *pw, float4 *spos, float4 *sdir, float4 *spw) {
    const uint32_t j = blockIdx.x * kPhBlock + threadIdx.x;
    if (j >= m) return;
    const uint32_t i = vals[j];
    spos[j] = pos[i];
    sdir[j] = dir[i];
    spw[j] = pw[i];
}

__global__ void kPhotonRanges(uint32_t m, const uint32_t *keys, uint32_t *start, uint32_t *end) {
    const uint32_t j = blockIdx.x * kPhBlock + threadIdx.x;
    if (j >= m) return;
    const uint32_t k = keys[j];
    if (j == 0 || keys[j - 1] != k) start[k] = j;
    if (j == m - 1 || keys[j + 1] != k) end[k] = j + 1;
}

__global__ void kPhotonRadiance(SceneView s, int64_t n, const float *pos, const float *wo, const float *nrm,
                                const float *params, float *out) {
    const int64_t i = (int64_t) blockIdx.x * kPhBlock + threadIdx.x;
    if (i >= n) return;
    const float *p = params + 8 * i;
    const Frame fr = makeFrame(V3(nrm[3 * i], nrm[3 * i + 1], nrm[3 * i + 2]));
    const Lobes lb = makeLobes(V3(p[1], p[2], p[3]), p[4], p[5], p[6]);
    const V3 r = photonRadiance(s, V3(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]),
                                V3(wo[3 * i], wo[3 * i + 1], wo[3 * i + 2]), fr, lb);
    out[3 * i] = r.x;
    out[3 * i + 1] = r.y;
    out[3 * i + 2] = r.z;
}

hipError_t launchPhotonRadiance(const SceneView &s, int64_t n, const float *pos, const float *wo, const float *nrm,
                                const float *params, float *out, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(kPhotonRadiance, dim3((unsigned) ((n + kPhBlock - 1) / kPhBlock)), dim3(kPhBlock), 0, st, s, n,
                       pos, wo, nrm, params, out);
    return hipGetLastError();
}

hipError_t launchPhotonTrace(const SceneView &s, int64_t n, uint64_t seed, int maxDepth, float4 *pos, float4 *dir,
                             float4 *pw, uint32_t *stored, uint2 *spill, int spillLanes, int stackDepth,
                             hipStream_t st) {
    int64_t blocks = (n + kTraceBlock - 1) / kTraceBlock;
    if (blocks > spillLanes / kTraceBlock) blocks = spillLanes / kTraceBlock;
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(kPhotonTrace, dim3((unsigned) blocks), dim3(kTraceBlock),
                       stackDepth * kTraceBlock * sizeof(uint2), st, s, n, seed, maxDepth, pos, dir, pw, stored, spill,
                       stackDepth);
    return hipGetLastError();
}

hipError_t launchPhotonTraceRef(const SceneView &s, int64_t n, uint64_t seed, int maxDepth, float4 *pos, float4 *dir,
                                float4 *pw, uint64_t *key, uint32_t cap, uint32_t *count, uint2 *spill, int spillLanes,
                                int stackDepth, hipStream_t st) {
    int64_t blocks = (n + kTraceBlock - 1) / kTraceBlock;
    if (blocks > spillLanes / kTraceBlock) blocks = spillLanes / kTraceBlock;
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(kPhotonTraceRef, dim3((unsigned) blocks), dim3(kTraceBlock),
                       stackDepth * kTraceBlock * sizeof(uint2), st, s, n, seed, maxDepth, pos, dir, pw, key, cap,
                       count, spill, stackDepth);
    return hipGetLastError();
}

hipError_t photonSortByKey(void *temp, size_t &tempBytes, const uint64_t *kin, uint64_t *kout, const uint32_t *vin,
                           uint32_t *vout, int n, hipStream_t st) {
    return hipcub::DeviceRadixSort::SortPairs(temp, tempBytes, kin, kout, vin, vout, n, 0, 64, st);
}

hipError_t launchPhotonIota(uint32_t m, uint32_t *vals, hipStream_t st) {
    if (m == 0) return hipSuccess;
    hipLaunchKernelGGL(kPhotonIota, dim3((m + kPhBlock - 1) / kPhBlock), dim3(kPhBlock), 0, st, m, vals);
    return hipGetLastError();
}

hipError_t launchPhotonGather(uint32_t m, const uint32_t *vals, const float4 *pos, const float4 *dir, const float4 *pw,
                              const uint64_t *key, float4 *opos, float4 *odir, float4 *opw, uint64_t *okey,
                              hipStream_t st) {
    if (m == 0) return hipSuccess;
    hipLaunchKernelGGL(kPhotonGather, dim3((m + kPhBlock - 1) / kPhBlock), dim3(kPhBlock), 0, st, m, vals, pos, dir,
                       pw, key, opos, odir, opw, okey);
    return hipGetLastError();
}

hipError_t launchPhotonRanges(uint32_t m, const uint32_t *keys, uint32_t *start, uint32_t *end, hipStream_t st) {
    if (m == 0) return hipSuccess;
    hipLaunchKernelGGL(kPhotonRanges, dim3((m + kPhBlock - 1) / kPhBlock), dim3(kPhBlock), 0, st, m, keys, start, end);
    return hipGetLastError();
}

hipError_t launchPhotonIrradiance(const SceneView &s, int64_t n, const float *pos, float *out, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(kPhotonIrradiance, dim3((unsigned) ((n + kPhBlock - 1) / kPhBlock)), dim3(kPhBlock), 0, st, s,
                       n, pos, out);
    return hipGetLastError();
}

hipError_t launchPhotonKeys(int64_t n, const float4 *pos, float cell, uint32_t mask, uint32_t *keys, uint32_t *vals,
                            hipStream_t st) {
    hipLaunchKernelGGL(kPhotonKeys, dim3((unsigned) ((n + kPhBlock - 1) / kPhBlock)), dim3(kPhBlock), 0, st, n, pos,
                       cell, mask, keys, vals);
    return hipGetLastError();
}

hipError_t photonSort(void *temp, size_t &tempBytes, const uint32_t *kin, uint32_t *kout, const uint32_t *vin,
                      uint32_t *vout, int n, hipStream_t st) {
    return hipcub::DeviceRadixSort::SortPairs(temp, tempBytes, kin, kout, vin, vout, n, 0, 32, st);
}

hipError_t launchPhotonFinish(uint32_t m, const uint32_t *keys, const uint32_t *vals, const float4 *pos,
                              const float4 *dir, const float4 *pw, float4 *spos, float4 *sdir, float4 *spw,
                              uint32_t *start, uint32_t *end, hipStream_t st) {
    if (m == 0) return hipSuccess;
    const dim3 g((m + kPhBlock - 1) / kPhBlock);
    hipLaunchKernelGGL(kPhotonScatter, g, dim3(kPhBlock), 0, st, m, vals, pos, dir, pw, spos, sdir, spw);
    hipLaunchKernelGGL(kPhotonRanges, g, dim3(kPhBlock), 0, st, m, keys, start, end);
    return hipGetLastError();
}

}  // namespace prr
