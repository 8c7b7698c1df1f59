// context.hpp — host-side state behind the C ABI handles (prr_context,
// prr_scene) and the render helpers shared by abi.hip (plain renders)
// and cache_abi.hip (radiance-cache renders).
#pragma once

#include <string>
#include <utility>
#include <vector>

#include "../../include/prr_abi.h"
#include "device/dev_common.hpp"
#include "device/render.hpp"
#include "host/scene_build.hpp"

struct prr_context {
    int device = 0;
    hipStream_t own = nullptr;
    hipStream_t stream = nullptr;
    std::string err;
    int numCU = 1;
    int traceBlocks = 1, shadeBlocks = 1;
    int refillAt = 24;
    int stackDepth = 8;
    int whileWhile = 8;     // traversal iteration: 0 one node per work check, 1 interior nodes then a leaf,
                            // n > 1 the same with at most n interior steps per check
    int whileWhileAny = 4;  // the same for shadow rays (on the four-wide tree)
    int fetchRun = 256;     // traversal work fetch: a wave claims at most this many queue entries per atomic
    int shadowWide = 2;     // shadow rays: 0 the binary tree, 1 the four-wide occlusion tree, 2 its quantised copy
    bool shadowFast = true; // quantised tree: short form of the node test for rays in the scene's safe region
    bool speculate = false; // extension rays: speculative while-while iteration
    bool twoLevel = false;  // extension rays: two binary levels per step on the two-level records
    int shadeWaves = 4;     // shading kernel compiled for at least this many waves per SIMD (1, 4, 5)
    // shading sorted by material: each bounce's queue split into diffuse-hit
    // entries and the rest, shaded by two specialised kernels (the rest at
    // shadeLightWaves waves per SIMD: 4, 6 or 8)
    int shadeSort = 0;
    int shadeLightWaves = 8;
    prr::ShadeBins bins{};
    int64_t maxSlots = 1 << 24;  // paths per chunk (set at creation from free HBM)
    uint2 *spill = nullptr;
    int spillLanes = 0;
    // this device's address of the static spill array of
    // prr_trace_closest_noalloc (looked up once per device at prr_create)
    uint2 *noAllocSpill = nullptr;
    // shadow rays of bounce b - 1 are traced on `aux` while the extension
    // rays of bounce b are traced on the render stream (each launch's tail
    // overlaps the other's bulk); the aux launches use their own spill region
    int concurrentShadow = 1;
    int64_t concurrentMaxPaths = (int64_t) 1 << 24;  // only chunks of at most this many paths
    hipStream_t aux = nullptr;
    hipEvent_t evFork = nullptr, evJoin = nullptr;
    uint2 *spillAux = nullptr;
    // path completion: once a bounce's queue holds at most tailPaths entries
    // (0: never), one kFinish launch runs those paths to their end instead of
    // three launches per remaining bounce. The kernels choose from the queue
    // counts on the device; the host copies each bounce's count into cntHost
    // (pinned, kMaxBounces entries) behind event evCnt[b] and stops
    // enqueuing bounces once a copied count shows the switch.
    static constexpr int kMaxBounces = 66;
    int64_t tailPaths = 1 << 18;
    // camera paths of a full-image render in bands of bandRows image rows,
    // column by column within a band (ChunkInfo::bandRows; 1: scanline order)
    int bandRows = 8;
    uint32_t *cntHost = nullptr;
    hipEvent_t evCnt[kMaxBounces] = {};
    // shadow-queue binning before the any-hit traversal of bounces >=
    // shadowSortFrom (0: off) on a grid of 2^shadowSortBits cells per axis
    int shadowSortFrom = 0;
    int shadowSortBits = 2;
    prr::ShadowSortBufs sb{};
    // render buffers
    prr::PathBufs pb{};
    int64_t pbSlots = 0;
    std::vector<void *> pbAllocs;
    int maxDepthAlloc = 0;
    float4 *fb = nullptr;
    int64_t fbPixels = 0;
    unsigned long long *stats = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    std::vector<hipEvent_t> evPool;  // per-launch timing events (PRR_RENDER_KERNEL_TIMING)
};

struct PhotonMap {
    std::vector<void *> allocs;
    int64_t emitted = 0, stored = 0;
    float radius = 0.f;
    int maxDepth = 0;
    uint64_t seed = 0;
    // 0: the build's caustic map from the point light (prr_photon_build),
    // gathered by renders; 1: the reference's own map from the area emitters
    // (prr_photon_emit), which castRay never gathers
    int rules = 0;
    uint64_t *keys = nullptr;  // rules 1: emitting photon * 128 + trace depth per stored photon (bucket order)
};

struct prr_scene {
    prr::host::HostScene hs;
    prr::SceneView view{};
    std::vector<void *> allocs;
    int64_t bytes = 0;
    prr::AgxConst agx{};
    PhotonMap ph;
};

// Records msg as the context's (and the thread's) last error; returns code.
int fail(prr_context *c, int code, const std::string &msg);

#define HIPCK(ctx, expr)                                                                          \
    do {                                                                                          \
        hipError_t e_ = (expr);                                                                   \
        if (e_ != hipSuccess)                                                                     \
            return fail(ctx, PRR_ERR_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

namespace prr_host {

// Paths in flight per chunk, at most. Path state is kPathBytes per slot, so
// 2^27 slots hold ~47 GB of HBM: a whole 1080p frame at 64 spp in one chunk
// (with shadow binning on, 108 B more per slot); the shading bins take 8 B.
// Large chunks pay the per-bounce tail of the traversal kernels (the last few
// long rays of each bounce) once per frame instead of once per chunk. The
// context's own limit (prr_context::maxSlots) also keeps path state within
// 40 % of the HBM free when the context is created.
constexpr int64_t kMaxSlots = (int64_t) 1 << 27;
constexpr int64_t kPathBytes = 360;

struct Timing {
    bool on = false;
    std::vector<std::pair<int, int>> timed;  // (kind 0 extend / 1 shadow / 2 shade / 3 finish / 4 shadow sort,
                                             // first event)
    size_t next = 0;
};

void freeRenderBufs(prr_context *c);
int ensureRenderBufs(prr_context *c, int64_t slots, int maxDepth);
// Runs bounces 0..maxDepth of one chunk whose queue 0 is already filled.
int runBounces(prr_context *c, const prr::SceneView &view, const prr::ChunkInfo &ci, Timing &tm, hipStream_t st);
int resetChunk(prr_context *c, int maxDepth, hipStream_t st);
int finishStats(prr_context *c, const Timing &tm, uint64_t photons, prr_stats *stats);
int ensureFrame(prr_context *c, int64_t pixels, prr::AdaptiveBufs &ab);

}  // namespace prr_host
