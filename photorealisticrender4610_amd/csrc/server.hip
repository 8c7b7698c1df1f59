// server.hip — continuous-batching render server (prr_server_*).
//
// A worker thread owns the context. Each loop: admit pending requests in
// arrival order while their framebuffers fit the HBM budget; plan one batch
// of at most path_slots paths with one segment per admitted request (the
// share of each set by the server's policy); run it as one part per scene
// (segment ray generation, the wavefront bounces, per-request accumulation in
// sample order), the part of the most urgent request first; after each part,
// resolve, copy out and hand over the requests it completed. Requests
// submitted meanwhile wait only for the current batch to end.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <new>
#include <string>
#include <thread>
#include <vector>

#include "context.hpp"
#include "device/launch_render.hpp"
#include "host/scene_build.hpp"

using namespace prr;
using namespace prr_host;

namespace {

using Clock = std::chrono::steady_clock;

struct Request {
    int64_t id = 0;
    prr_request_desc d{};
    CamView cam{};
    int64_t pixels = 0;
    int64_t bytes = 0;
    float4 *fb = nullptr;
    int sCur = 0;       // next sample round
    int64_t pCur = 0;   // first pixel of the round not yet rendered
    bool done = false, failed = false;
    double tSubmit = 0, tAdmit = 0, tDone = 0;
    int batches = 0;
    std::vector<float> rgb;
    std::string err;
    int64_t remaining() const { return (int64_t) (d.spp - sCur) * pixels - pCur; }
};

struct Planned {
    Request *r;
    Segment seg;
    bool finishes;
};

}  // namespace

struct prr_server {
    prr_context *ctx = nullptr;
    std::vector<const prr_scene *> scenes;
    std::vector<int> depth;
    std::vector<float> clampInd;
    prr_server_params p{};
    int64_t budget = 0;
    std::mutex mu;
    std::condition_variable cvWork, cvDone;
    std::deque<Request *> pending;
    std::vector<Request *> active;
    std::map<int64_t, Request *> all;
    int64_t nextId = 1;
    int64_t admittedBytes = 0;
    bool stop = false;
    std::thread worker;
    Clock::time_point t0;
    prr_server_stats stats{};
    std::string err;
    // device buffers
    uint32_t *sampleOf = nullptr;
    Segment *dSegs = nullptr;
    int64_t *dPixStart = nullptr;
    int segCap = 0;
    double aging = 1.0;  // PRR_SERVE_AGING: weight of the waiting credit in PRR_SERVE_AGED
    float *dRgb = nullptr;
    int64_t rgbCap = 0;
    double now() const { return std::chrono::duration<double, std::milli>(Clock::now() - t0).count(); }
};


// Bookkeeping of plan entries whose scene part has run (worker thread, under
// the lock): sample progress, and for the requests the batch completes their
// completion time, release of their budget, removal from the active list and
// the wake-up of their waiters. A completed request may be deleted by its
// waiter as soon as the lock is released, so nothing touches it afterwards.
static void completeEntries(prr_server *S, std::vector<Planned *> &done, double t) {
    for (Planned *pl : done) {
        Request *r = pl->r;
        r->batches++;
        S->stats.samples += (int64_t) pl->seg.np * pl->seg.ns;
        if (pl->seg.ns == 1 && pl->seg.p0 + pl->seg.np < r->pixels) {
            r->pCur = pl->seg.p0 + pl->seg.np;
        } else {
            r->pCur = 0;
            r->sCur += pl->seg.ns;
        }
        if (pl->finishes) {
            r->done = true;
            r->tDone = t;
            S->admittedBytes -= r->bytes;
            S->stats.requests_done++;
        }
        pl->r = nullptr;
    }
    S->active.erase(std::remove_if(S->active.begin(), S->active.end(), [](Request *r) { return r->done; }),
                    S->active.end());
    S->cvDone.notify_all();
}

// Runs one planned batch: one part per scene (segment ray generation, the
// wavefront bounces, per-request accumulation in sample order), the parts in
// the order of their most urgent request (the plan's order). The requests a
// part completes are resolved and handed to their waiters as soon as that
// part has run, not at the end of the batch. Called without the lock; takes
// it only to complete entries.
static int runBatch(prr_server *S, std::vector<Planned> &plan, std::unique_lock<std::mutex> &lk, int64_t &rays) {
    prr_context *c = S->ctx;
    hipStream_t st = c->stream;
    const int nseg = (int) plan.size();
    if (nseg > S->segCap) return fail(c, PRR_ERR_INVALID, "too many segments");
    HIPCK(c, hipMemsetAsync(c->stats, 0, 8 * sizeof(unsigned long long), st));
    std::vector<int> sceneOrder;
    for (auto &pl : plan)
        if (std::find(sceneOrder.begin(), sceneOrder.end(), pl.r->d.scene) == sceneOrder.end())
            sceneOrder.push_back(pl.r->d.scene);
    std::vector<Segment> segs;
    std::vector<int64_t> pixStart;
    std::vector<Planned *> part;
    int off = 0;
    for (const int sc : sceneOrder) {
        segs.clear();
        pixStart.clear();
        part.clear();
        int64_t slots = 0, pix = 0;
        for (auto &pl : plan) {
            if (!pl.r || pl.r->d.scene != sc) continue;  // null: completed by an earlier part
            pl.seg.slot0 = slots;
            segs.push_back(pl.seg);
            pixStart.push_back(pix);
            part.push_back(&pl);
            slots += (int64_t) pl.seg.np * pl.seg.ns;
            pix += pl.seg.np;
        }
        Segment *dS = S->dSegs + off;
        int64_t *dP = S->dPixStart + off;
        off += (int) segs.size();
        HIPCK(c, hipMemcpyAsync(dS, segs.data(), segs.size() * sizeof(Segment), hipMemcpyHostToDevice, st));
        HIPCK(c, hipMemcpyAsync(dP, pixStart.data(), pixStart.size() * sizeof(int64_t), hipMemcpyHostToDevice, st));
        SceneView view = S->scenes[sc]->view;
        view.ph_count = 0;
        ChunkInfo ci;
        ci.W = ci.H = 1;
        ci.pixels = (int) slots;
        ci.spc = 1;
        ci.sampleBase = 0;
        ci.seed = 0;
        ci.maxDepth = S->depth[sc];
        ci.clampInd = S->clampInd[sc];
        ci.sampleOf = S->sampleOf;
        int rc;
        if ((rc = resetChunk(c, ci.maxDepth, st)) != PRR_OK) return rc;
        HIPCK(c, launchRaygenSeg(c->pb, dS, (int) segs.size(), slots, S->sampleOf, c->shadeBlocks, st));
        Timing tm;
        if ((rc = runBounces(c, view, ci, tm, st)) != PRR_OK) return rc;
        HIPCK(c, launchAccumulateSeg(c->pb, dS, dP, (int) segs.size(), pix, st));
        // completed requests of this part: mean = sums / spp (Renderer.cpp:99), copied out
        for (Planned *pl : part) {
            if (!pl->finishes) continue;
            Request *r = pl->r;
            if (r->pixels * 3 > S->rgbCap) {
                HIPCK(c, hipStreamSynchronize(st));
                if (S->dRgb) HIPCK(c, hipFree(S->dRgb));
                S->dRgb = nullptr;
                S->rgbCap = 0;
                HIPCK(c, hipMalloc(&S->dRgb, (size_t) r->pixels * 3 * sizeof(float)));
                S->rgbCap = r->pixels * 3;
            }
            HIPCK(c, launchResolve(r->fb, (int) r->pixels, r->d.spp, S->dRgb, st));
            r->rgb.resize((size_t) r->pixels * 3);
            HIPCK(c, hipMemcpyAsync(r->rgb.data(), S->dRgb, r->rgb.size() * sizeof(float), hipMemcpyDeviceToHost, st));
            HIPCK(c, hipStreamSynchronize(st));
            HIPCK(c, hipFree(r->fb));
            r->fb = nullptr;
        }
        const double t = S->now();
        lk.lock();
        completeEntries(S, part, t);
        lk.unlock();
    }
    unsigned long long h[3] = {0, 0, 0};
    HIPCK(c, hipMemcpyAsync(h, c->stats, sizeof(h), hipMemcpyDeviceToHost, st));
    HIPCK(c, hipStreamSynchronize(st));
    rays = (int64_t) (h[0] + h[1]);
    return PRR_OK;
}

// One segment for request r with room paths left in the batch: the rest of a
// partial sample round, else whole rounds, else the start of a round.
static bool planSegment(Request *r, int64_t room, Planned &pl) {
    const int64_t P = r->pixels;
    Segment sg;
    sg.cam = r->cam;
    sg.W = r->d.width;
    sg.H = r->d.height;
    sg.seed = r->d.seed;
    sg.fb = r->fb;
    sg.slot0 = 0;
    sg.s0 = r->sCur;
    if (r->pCur > 0) {
        sg.p0 = (int) r->pCur;
        sg.np = (int) std::min(P - r->pCur, room);
        sg.ns = 1;
    } else if (room >= P) {
        sg.p0 = 0;
        sg.np = (int) P;
        sg.ns = (int) std::min<int64_t>(r->d.spp - r->sCur, room / P);
    } else {
        sg.p0 = 0;
        sg.np = (int) room;
        sg.ns = 1;
    }
    if (sg.np <= 0 || sg.ns <= 0) return false;
    pl.r = r;
    pl.seg = sg;
    const bool roundEnds = sg.p0 + sg.np == P;
    pl.finishes = roundEnds && sg.s0 + sg.ns >= r->d.spp;
    return true;
}

// One batch of at most path_slots paths, one segment per admitted request.
//   PRR_SERVE_FAIR      max-min fair shares: requests in order of remaining
//                       work each take min(their remaining work, an equal
//                       share of the room left), so small requests finish in
//                       the next batch and large ones split the rest evenly
//   PRR_SERVE_OLDEST    arrival order, each takes all it can
//   PRR_SERVE_SHORTEST  shortest remaining work first, each takes all it can
//   PRR_SERVE_AGED      shortest remaining work first after crediting each
//                       request with the samples the server renders in the
//                       time it has waited: waiting requests gain priority
//                       at the service rate, so large ones are not starved
static void planBatch(prr_server *S, std::vector<Planned> &plan) {
    std::vector<Request *> order(S->active.begin(), S->active.end());
    if (S->p.policy == PRR_SERVE_AGED) {
        const double now = S->now();
        const double rate = S->stats.busy_ms > 0.0 ? (double) S->stats.samples / S->stats.busy_ms : 2e5;
        std::vector<std::pair<double, Request *>> key;
        for (Request *r : order) key.emplace_back((double) r->remaining() - S->aging * rate * (now - r->tSubmit), r);
        std::stable_sort(key.begin(), key.end(),
                         [](const std::pair<double, Request *> &a, const std::pair<double, Request *> &b) {
                             return a.first < b.first;
                         });
        for (size_t i = 0; i < key.size(); i++) order[i] = key[i].second;
    } else if (S->p.policy != PRR_SERVE_OLDEST) {
        std::stable_sort(order.begin(), order.end(),
                         [](const Request *a, const Request *b) { return a->remaining() < b->remaining(); });
    }
    int64_t room = S->p.path_slots;
    const int n = (int) std::min<size_t>(order.size(), (size_t) S->segCap);
    for (int i = 0; i < n && room > 0; i++) {
        int64_t give = room;
        if (S->p.policy == PRR_SERVE_FAIR) give = std::max<int64_t>(1, room / (n - i));
        Planned pl;
        if (!planSegment(order[i], give, pl)) continue;
        room -= (int64_t) pl.seg.np * pl.seg.ns;
        plan.push_back(pl);
    }
}

static void workerLoop(prr_server *S) {
    (void) hipSetDevice(S->ctx->device);
    std::unique_lock<std::mutex> lk(S->mu);
    while (true) {
        // admission in arrival order while the budget allows
        while (!S->pending.empty()) {
            Request *r = S->pending.front();
            if (S->admittedBytes + r->bytes > S->budget) break;
            void *q = nullptr;
            hipError_t e = hipMalloc(&q, (size_t) r->bytes);
            if (e == hipSuccess) e = hipMemsetAsync(q, 0, (size_t) r->bytes, S->ctx->stream);
            if (e != hipSuccess) {
                if (q) (void) hipFree(q);
                if (!S->active.empty()) break;  // retry once running requests free memory
                // nothing running will free memory: the request fails
                r->failed = r->done = true;
                r->err = std::string("request framebuffer: ") + hipGetErrorString(e);
                r->tDone = S->now();
                S->pending.pop_front();
                S->cvDone.notify_all();
                continue;
            }
            r->fb = (float4 *) q;
            r->tAdmit = S->now();
            S->admittedBytes += r->bytes;
            S->pending.pop_front();
            S->active.push_back(r);
            S->stats.peak_admitted_bytes = std::max(S->stats.peak_admitted_bytes, S->admittedBytes);
            S->stats.peak_active = std::max(S->stats.peak_active, (int32_t) S->active.size());
        }
        if (S->active.empty()) {
            if (S->stop && S->pending.empty()) break;
            S->cvWork.wait(lk);
            continue;
        }
        std::vector<Planned> plan;
        planBatch(S, plan);
        lk.unlock();
        const double tb = S->now();
        int64_t rays = 0;
        const int rc = runBatch(S, plan, lk, rays);
        const double te = S->now();
        lk.lock();
        S->stats.batches++;
        S->stats.busy_ms += te - tb;
        S->stats.rays += rays;
        if (rc != PRR_OK) {
            // the device failed: every request still admitted fails with the message
            S->err = S->ctx->err;
            for (Request *r : S->active) {
                r->failed = r->done = true;
                r->err = S->err;
                r->tDone = te;
                if (r->fb) (void) hipFree(r->fb);
                r->fb = nullptr;
            }
            S->active.clear();
            S->admittedBytes = 0;
            S->cvDone.notify_all();
            continue;
        }
        S->cvDone.notify_all();
    }
}

extern "C" {

int prr_server_create(prr_context *c, const prr_scene *const *scenes, const int32_t *max_depth,
                      const float *indirect_clamp, int32_t num_scenes, const prr_server_params *p,
                      prr_server **out) {
    if (!c || !scenes || !max_depth || !indirect_clamp || !p || !out || num_scenes <= 0)
        return fail(c, PRR_ERR_INVALID, "NULL argument");
    *out = nullptr;
    if (p->path_slots <= 0 || p->path_slots > c->maxSlots || p->hbm_budget < 0 || p->policy < PRR_SERVE_AGED ||
        p->policy > PRR_SERVE_FAIR)
        return fail(c, PRR_ERR_INVALID, "bad server parameters");
    int maxDepth = 1;
    for (int i = 0; i < num_scenes; i++) {
        if (!scenes[i] || max_depth[i] <= 0 || max_depth[i] > 64) return fail(c, PRR_ERR_INVALID, "bad scene");
        maxDepth = std::max(maxDepth, (int) max_depth[i]);
    }
    HIPCK(c, hipSetDevice(c->device));
    int rc = ensureRenderBufs(c, p->path_slots, maxDepth);
    if (rc != PRR_OK) return rc;
    prr_server *S = new (std::nothrow) prr_server();
    if (!S) return fail(c, PRR_ERR_NOMEM, "out of host memory");
    S->ctx = c;
    S->p = *p;
    if (const char *ev = getenv("PRR_SERVE_AGING")) S->aging = std::max(0.0, atof(ev));
    for (int i = 0; i < num_scenes; i++) {
        S->scenes.push_back(scenes[i]);
        S->depth.push_back(max_depth[i]);
        S->clampInd.push_back(indirect_clamp[i]);
    }
    size_t freeB = 0, totalB = 0;
    (void) hipMemGetInfo(&freeB, &totalB);
    S->budget = p->hbm_budget > 0 ? p->hbm_budget : (int64_t) (freeB / 2);
    S->segCap = 1024;
    hipError_t e = hipMalloc(&S->sampleOf, (size_t) p->path_slots * sizeof(uint32_t));
    if (e == hipSuccess) e = hipMalloc(&S->dSegs, (size_t) S->segCap * sizeof(Segment));
    if (e == hipSuccess) e = hipMalloc(&S->dPixStart, (size_t) S->segCap * sizeof(int64_t));
    if (e != hipSuccess) {
        if (S->sampleOf) (void) hipFree(S->sampleOf);
        if (S->dSegs) (void) hipFree(S->dSegs);
        delete S;
        return fail(c, PRR_ERR_NOMEM, std::string("server buffers: ") + hipGetErrorString(e));
    }
    S->t0 = Clock::now();
    S->worker = std::thread(workerLoop, S);
    *out = S;
    return PRR_OK;
}

int prr_server_destroy(prr_server *S) {
    if (!S) return PRR_OK;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        S->stop = true;
    }
    S->cvWork.notify_all();
    if (S->worker.joinable()) S->worker.join();
    (void) hipSetDevice(S->ctx->device);
    for (auto &kv : S->all) {
        if (kv.second->fb) (void) hipFree(kv.second->fb);
        delete kv.second;
    }
    if (S->sampleOf) (void) hipFree(S->sampleOf);
    if (S->dSegs) (void) hipFree(S->dSegs);
    if (S->dPixStart) (void) hipFree(S->dPixStart);
    if (S->dRgb) (void) hipFree(S->dRgb);
    delete S;
    return PRR_OK;
}

int prr_server_submit(prr_server *S, const prr_request_desc *d, int64_t *id) {
    if (!S || !d || !id) return PRR_ERR_INVALID;
    if (d->scene < 0 || d->scene >= (int) S->scenes.size() || d->width <= 0 || d->height <= 0 || d->spp <= 0 ||
        (int64_t) d->width * d->height > ((int64_t) 1 << 30)) {
        std::lock_guard<std::mutex> lk(S->mu);
        S->err = "bad request";
        return PRR_ERR_INVALID;
    }
    const int64_t pixels = (int64_t) d->width * d->height;
    if (pixels * (int64_t) sizeof(float4) > S->budget) {
        std::lock_guard<std::mutex> lk(S->mu);
        S->err = "request larger than the server's budget";
        return PRR_ERR_NOMEM;
    }
    Request *r = new (std::nothrow) Request();
    if (!r) return PRR_ERR_NOMEM;
    r->d = *d;
    r->cam = host::makeCamView(d->camera);
    r->pixels = pixels;
    r->bytes = pixels * (int64_t) sizeof(float4);
    {
        std::lock_guard<std::mutex> lk(S->mu);
        r->id = S->nextId++;
        r->tSubmit = S->now();
        S->pending.push_back(r);
        S->all[r->id] = r;
        *id = r->id;
    }
    S->cvWork.notify_one();
    return PRR_OK;
}

int prr_server_wait(prr_server *S, int64_t id, float *rgb, prr_request_stats *st) {
    if (!S) return PRR_ERR_INVALID;
    std::unique_lock<std::mutex> lk(S->mu);
    auto it = S->all.find(id);
    if (it == S->all.end()) {
        S->err = "unknown request id";
        return PRR_ERR_INVALID;
    }
    Request *r = it->second;
    S->cvDone.wait(lk, [&] { return r->done; });
    int rc = PRR_OK;
    if (r->failed) {
        S->err = r->err;
        rc = PRR_ERR_DEVICE;
    } else if (rgb) {
        std::memcpy(rgb, r->rgb.data(), r->rgb.size() * sizeof(float));
    }
    if (st) {
        st->t_submit_ms = r->tSubmit;
        st->t_admit_ms = r->tAdmit;
        st->t_done_ms = r->tDone;
        st->samples = r->pixels * r->d.spp;
        st->batches = r->batches;
        st->reserved = 0;
    }
    S->all.erase(it);
    delete r;
    return rc;
}

int prr_server_get_stats(prr_server *S, prr_server_stats *st) {
    if (!S || !st) return PRR_ERR_INVALID;
    std::lock_guard<std::mutex> lk(S->mu);
    *st = S->stats;
    return PRR_OK;
}

// The message of the last failed call; read it from the thread that made the
// call, before further calls.
const char *prr_server_last_error(const prr_server *S) { return S ? S->err.c_str() : ""; }

}  // extern "C"
