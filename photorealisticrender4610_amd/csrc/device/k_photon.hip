// k_photon.hip — caustic photon map on the device (build extension of
// Photon.cpp:96-190: emission from the point light, delta-chain tracing,
// storage at the first diffuse hit after >= 1 delta bounce, hash-grid
// storage instead of the reference's unordered_map of grid cells).
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
