// k_trace.hip — batch closest-hit / any-hit kernels (prr_trace_closest /
// prr_trace_any), one ray per lane, persistent grid-stride blocks so the
// global spill region is sized by resident lanes, not by rays.
#include "dev_traverse.hpp"
#include "launch.hpp"

namespace prr {

__global__ __launch_bounds__(kTraceBlock) void kTraceClosest(SceneView s, int64_t n, const float *__restrict__ org,
                                                             const float *__restrict__ dir, int32_t *__restrict__ prim,
                                                             float *__restrict__ tout, float *__restrict__ uout,
                                                             float *__restrict__ vout, uint2 *__restrict__ spill,
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
    for (int64_t i = lane; i < n; i += stride) {
        const V3 o(org[3 * i], org[3 * i + 1], org[3 * i + 2]);
        const V3 d(dir[3 * i], dir[3 * i + 1], dir[3 * i + 2]);
        const HitOut h = traceClosest(s, o, d, st);
        prim[i] = h.prim;
        tout[i] = h.prim >= 0 ? h.t : __builtin_inff();
        if (uout) uout[i] = h.u;
        if (vout) vout[i] = h.v;
    }
}

__global__ __launch_bounds__(kTraceBlock) void kTraceAny(SceneView s, int64_t n, const float *__restrict__ org,
                                                         const float *__restrict__ dir, const float *__restrict__ tmax,
                                                         uint8_t *__restrict__ hit, uint2 *__restrict__ spill,
                                                         int stackDepth, int quant) {
    extern __shared__ uint2 stk[];
    const int64_t lane = (int64_t) blockIdx.x * kTraceBlock + threadIdx.x;
    RefStack st = makeStack<RefStack>(stk, spill + lane * kSpill, threadIdx.x, stackDepth);
    const int64_t stride = (int64_t) gridDim.x * kTraceBlock;
    for (int64_t i = lane; i < n; i += stride) {
        const V3 o(org[3 * i], org[3 * i + 1], org[3 * i + 2]);
        const V3 d(dir[3 * i], dir[3 * i + 1], dir[3 * i + 2]);
        hit[i] = (quant ? traceAny4Q(s, o, d, tmax[i], st) : traceAny4(s, o, d, tmax[i], st)) ? 1 : 0;
    }
}

static int gridFor(int64_t n, int spillLanes) {
    int64_t blocks = (n + kTraceBlock - 1) / kTraceBlock;
    const int64_t maxBlocks = spillLanes / kTraceBlock;
    if (blocks > maxBlocks) blocks = maxBlocks;
    return (int) (blocks < 1 ? 1 : blocks);
}

hipError_t launchTraceClosest(const SceneView &s, int64_t n, const float *org, const float *dir, int32_t *prim,
                              float *t, float *u, float *v, uint2 *spill, int spillLanes, int stackDepth,
                              hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(kTraceClosest, dim3(gridFor(n, spillLanes)), dim3(kTraceBlock),
                       stackDepth * kTraceBlock * sizeof(uint2), st, s, n, org, dir, prim, t, u, v, spill, stackDepth);
    return hipGetLastError();
}

// Spill region of prr_trace_closest_noalloc: the stacks of kNoAllocBlocks
// blocks of kTraceBlock lanes, kSpill entries each (32 MiB), a file-scope
// device array that every GPU the library runs on holds until the process
// ends. Host code never passes the array's name to a kernel: that name is a
// host-side stand-in, not a device address. noAllocScratch() returns the
// current device's own address of it, which the launcher below takes.
static __device__ __align__(256) uint2 scratch[(size_t) kNoAllocBlocks * kTraceBlock * kSpill];

hipError_t noAllocScratch(uint2 **out) { return hipGetSymbolAddress((void **) out, HIP_SYMBOL(scratch)); }

hipError_t launchTraceClosestNoAlloc(const SceneView &s, int64_t n, const float *org, const float *dir, int32_t *prim,
                                     float *t, float *u, float *v, uint2 *scratchPtr, int stackDepth, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    // at most kNoAllocBlocks blocks: each lane's spill region lies inside the array
    hipLaunchKernelGGL(kTraceClosest, dim3(gridFor(n, kNoAllocBlocks * kTraceBlock)), dim3(kTraceBlock),
                       stackDepth * kTraceBlock * sizeof(uint2), st, s, n, org, dir, prim, t, u, v, scratchPtr,
                       stackDepth);
    return hipGetLastError();
}

hipError_t launchTraceAny(const SceneView &s, int64_t n, const float *org, const float *dir, const float *tmax,
                          uint8_t *hit, uint2 *spill, int spillLanes, int stackDepth, int quant, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(kTraceAny, dim3(gridFor(n, spillLanes)), dim3(kTraceBlock),
                       stackDepth * kTraceBlock * sizeof(uint2), st, s, n, org, dir, tmax, hit, spill, stackDepth, quant);
    return hipGetLastError();
}

}  // namespace prr
