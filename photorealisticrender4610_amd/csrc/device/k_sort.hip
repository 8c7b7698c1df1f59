// k_sort.hip — coherence binning of a bounce's shadow queue before the
// any-hit traversal. Shadow rays of the later bounces start at scattered
// points (queue order follows the paths, not space), while most of them head
// the same way (environment NEE rays towards a bright sun). Binning them by
// (direction octant, Morton code of the origin's cell in a coarse grid over
// the scene box) gives each wave rays that start close together and go the
// same way, so its lanes walk the same nodes: fewer divergent steps and more
// cache reuse. Any-hit answers do not depend on the order in which rays are
// traced, and each ray's answer is written to its own path slot, so images
// do not change.
//
// One counting-sort pass over the bin key (a radix-sort digit pass):
//   kBinHist     per block, a histogram of its chunk of the queue (LDS)
//   kBinScan     exclusive scan of the histograms in (bin, block) order
//   kBinScatter  each ray to its bin's range for its block (LDS cursors)
// LDS counters are updated once per distinct bin of a wave. Within a bin and
// block the order follows the waves' atomics (not stable); the traversal
// does not depend on it.
#include "launch_render.hpp"

namespace prr {

constexpr int kBinBlock = 256;
constexpr int kBinScanThreads = 1024;

// The low 10 bits of x spread to every third bit (Morton interleave).
PRR_DEV uint32_t spread3(uint32_t x) {
    x &= 0x3ffu;
    x = (x | (x << 16)) & 0x030000FFu;
    x = (x | (x << 8)) & 0x0300F00Fu;
    x = (x | (x << 4)) & 0x030C30C3u;
    x = (x | (x << 2)) & 0x09249249u;
    return x;
}

PRR_DEV uint32_t cellOf(float x, float lo, float scale, uint32_t top) {
    const float c = (x - lo) * scale;
    return c > 0.f ? (c < (float) top ? (uint32_t) c : top) : 0u;  // NaN -> 0
}

struct BinGrid {
    float lo[3], scale[3];
    uint32_t top;
    int bits;
};

PRR_DEV uint32_t binOf(const BinGrid &g, const float4 &o, const V3 &d) {
    const uint32_t mort = spread3(cellOf(o.x, g.lo[0], g.scale[0], g.top)) |
                          (spread3(cellOf(o.y, g.lo[1], g.scale[1], g.top)) << 1) |
                          (spread3(cellOf(o.z, g.lo[2], g.scale[2], g.top)) << 2);
    const uint32_t oct = (d.x < 0.f ? 1u : 0u) | (d.y < 0.f ? 2u : 0u) | (d.z < 0.f ? 4u : 0u);
    return (oct << (3 * g.bits)) | mort;
}

// Entries [c0, c1) of the queue's m entries handled by block b of nb.
PRR_DEV void chunkOf(int64_t m, int b, int nb, int64_t &c0, int64_t &c1) {
    const int64_t per = (m + nb - 1) / nb;
    c0 = (int64_t) b * per;
    c1 = c0 + per < m ? c0 + per : m;
}

// Adds the wave's lanes with valid to the LDS counters of their bins, one
// atomic per distinct bin in the wave (lanes of a wave mostly share a bin),
// and returns each such lane's previous counter value plus its rank among
// the wave's lanes of that bin. Every lane of the wave calls it.
PRR_DEV uint32_t waveBinAdd(uint32_t *cnt, uint32_t bin, bool valid) {
    const int lane = (int) __lane_id();
    const uint64_t below = (1ull << lane) - 1ull;
    uint64_t pending = __ballot(valid);
    uint32_t mine = 0;
    while (pending) {
        const int leader = __ffsll((unsigned long long) pending) - 1;
        const uint32_t lb = __shfl(bin, leader);
        const uint64_t m = __ballot(bin == lb) & pending;
        uint32_t b = 0;
        if (lane == leader) b = atomicAdd(&cnt[lb], (uint32_t) __popcll(m));
        b = __shfl(b, leader);
        if ((m >> lane) & 1ull) mine = b + (uint32_t) __popcll(m & below);
        pending &= ~m;
    }
    return mine;
}

__global__ __launch_bounds__(kBinBlock) void kBinHist(PathBufs pb, int bounce, BinGrid g, int bins,
                                                      uint32_t *__restrict__ hist) {
    extern __shared__ uint32_t cnt[];
    __shared__ uint32_t spre[kQSegs + 1];
    for (int k = threadIdx.x; k < bins; k += kBinBlock) cnt[k] = 0u;
    loadPrefix(spre, pb.sqpre + (int64_t) bounce * (kQSegs + 1));
    int64_t c0, c1;
    chunkOf(pb.sqcnt[bounce], blockIdx.x, gridDim.x, c0, c1);
    // wave-uniform trip count: every lane reaches the ballots
    for (int64_t w = c0 + (threadIdx.x & ~63); w < c1; w += kBinBlock) {
        const int64_t i = w + (threadIdx.x & 63);
        const bool valid = i < c1;
        const int64_t e = valid ? segEntry(spre, pb.sqcap, i) : 0;
        const uint32_t bin = valid ? binOf(g, pb.sqOrg[e], ldv3(pb.sqDir, e)) : 0u;
        (void) waveBinAdd(cnt, bin, valid);
    }
    __syncthreads();
    for (int k = threadIdx.x; k < bins; k += kBinBlock) hist[(size_t) k * gridDim.x + blockIdx.x] = cnt[k];
}

// Exclusive scan of n counts in place, one block: each thread sums a run of
// consecutive counts, the run sums are scanned in LDS, then each thread
// writes its run's offsets.
__global__ __launch_bounds__(kBinScanThreads) void kBinScan(uint32_t *__restrict__ hist, int64_t n) {
    __shared__ uint32_t part[kBinScanThreads];
    const int t = threadIdx.x;
    const int64_t per = (n + kBinScanThreads - 1) / kBinScanThreads;
    const int64_t a = (int64_t) t * per, e = a + per < n ? a + per : n;
    uint32_t sum = 0;
    for (int64_t i = a; i < e; i++) sum += hist[i];
    part[t] = sum;
    __syncthreads();
    for (int off = 1; off < kBinScanThreads; off <<= 1) {  // Hillis-Steele inclusive scan
        const uint32_t v = t >= off ? part[t - off] : 0u;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    uint32_t run = part[t] - sum;
    for (int64_t i = a; i < e; i++) {
        const uint32_t c = hist[i];
        hist[i] = run;
        run += c;
    }
}

__global__ __launch_bounds__(kBinBlock) void kBinScatter(PathBufs pb, int bounce, BinGrid g, int bins,
                                                         const uint32_t *__restrict__ base, uint32_t *__restrict__ slot,
                                                         float4 *__restrict__ org, float *__restrict__ dir) {
    extern __shared__ uint32_t cur[];
    __shared__ uint32_t spre[kQSegs + 1];
    for (int k = threadIdx.x; k < bins; k += kBinBlock) cur[k] = base[(size_t) k * gridDim.x + blockIdx.x];
    loadPrefix(spre, pb.sqpre + (int64_t) bounce * (kQSegs + 1));
    int64_t c0, c1;
    chunkOf(pb.sqcnt[bounce], blockIdx.x, gridDim.x, c0, c1);
    for (int64_t w = c0 + (threadIdx.x & ~63); w < c1; w += kBinBlock) {
        const int64_t i = w + (threadIdx.x & 63);
        const bool valid = i < c1;
        float4 o = make_float4(0.f, 0.f, 0.f, 0.f);
        V3 d(0.f);
        uint32_t sl = 0;
        if (valid) {
            const int64_t e = segEntry(spre, pb.sqcap, i);
            o = pb.sqOrg[e];
            d = ldv3(pb.sqDir, e);
            sl = pb.sqSlot[e];
        }
        const uint32_t j = waveBinAdd(cur, valid ? binOf(g, o, d) : 0u, valid);
        if (valid) {
            slot[j] = sl;
            org[j] = o;
            stv3(dir, j, d);
        }
    }
}

static BinGrid binGrid(const SceneView &s, int bits) {
    BinGrid g;
    g.bits = bits;
    g.top = (1u << bits) - 1u;
    for (int a = 0; a < 3; a++) {
        g.lo[a] = s.root_box[2 * a];
        const float ext = s.root_box[2 * a + 1] - s.root_box[2 * a];
        g.scale[a] = ext > 0.f ? (float) (1u << bits) / ext : 0.f;
    }
    return g;
}

int shadowBinCount(int bits) { return 8 << (3 * bits); }

hipError_t launchShadowSort(const SceneView &s, const PathBufs &pb, int bounce, int bits, const ShadowSortBufs &sb,
                            hipStream_t st) {
    const BinGrid g = binGrid(s, bits);
    const int bins = shadowBinCount(bits);
    const size_t lds = (size_t) bins * sizeof(uint32_t);
    hipLaunchKernelGGL(kBinHist, dim3(sb.blocks), dim3(kBinBlock), lds, st, pb, bounce, g, bins, sb.hist);
    hipLaunchKernelGGL(kBinScan, dim3(1), dim3(kBinScanThreads), 0, st, sb.hist, (int64_t) bins * sb.blocks);
    hipLaunchKernelGGL(kBinScatter, dim3(sb.blocks), dim3(kBinBlock), lds, st, pb, bounce, g, bins, sb.hist, sb.slot,
                       sb.org, sb.dir);
    return hipGetLastError();
}

}  // namespace prr
