// dev_traverse.hpp — BVH closest-hit and any-hit traversal on gfx950.
//
// The reference's two-level BVH (scene level over objects, mesh level over
// triangles; BVH.cpp:311-473) is merged on the host into one tree of 64-byte
// nodes that hold both children's boxes, so an interior expansion is one
// 64-byte read. Visit order, box test arithmetic, culling and the primitive
// tests are those of BVHAccel::Intersect / IntersectP and Triangle::hitTest /
// intersectT, which makes closest-hit ids (including ties) and t bit-identical
// to the reference.
//
// Stack: each lane keeps the top kShort entries (node ref, box entry t) in
// LDS, laid out [depth][lane] so a wave's 64 lanes hit 64 distinct banks;
// deeper entries spill to a per-lane region of global memory.
#pragma once

#include "dev_common.hpp"

namespace prr {

constexpr int kShort = 16;        // default LDS stack entries per lane
constexpr int kSpill = 64;        // spill entries per lane (the reference's stack holds 64)
constexpr int kTraceBlock = 256;  // threads per block of traversal kernels

struct RayPre {
    float o[3], inv[3];
    int oct[3];
};
// BVHAccel::precompute (BVH.cpp:290-299)
PRR_DEV RayPre precompute(const V3 &o, const V3 &d) {
    RayPre r;
    const float dv[3] = {d.x, d.y, d.z};
    const float ov[3] = {o.x, o.y, o.z};
#pragma unroll
    for (int a = 0; a < 3; a++) {
        r.o[a] = ov[a];
        r.inv[a] = (dv[a] != 0.f) ? (1.f / dv[a]) : __builtin_inff();
        r.oct[a] = (dv[a] < 0.f) ? 1 : 0;
    }
    return r;
}
// nodeIntersect (BVH.hpp:115-124) on one box [minX,maxX,minY,maxY,minZ,maxZ].
PRR_DEV float slab(const RayPre &r, float x0, float x1, float y0, float y1, float z0, float z1) {
    float tmin = 0.f, tmax = __builtin_inff();
    float nx = r.oct[0] ? x1 : x0, fx = r.oct[0] ? x0 : x1;
    float ny = r.oct[1] ? y1 : y0, fy = r.oct[1] ? y0 : y1;
    float nz = r.oct[2] ? z1 : z0, fz = r.oct[2] ? z0 : z1;
    float t0 = (nx - r.o[0]) * r.inv[0], t1 = (fx - r.o[0]) * r.inv[0];
    tmin = fmaxf(tmin, t0);  // == std::max(tmin, t0) for non-NaN tmin
    tmax = fminf(tmax, t1);
    t0 = (ny - r.o[1]) * r.inv[1];
    t1 = (fy - r.o[1]) * r.inv[1];
    tmin = fmaxf(tmin, t0);
    tmax = fminf(tmax, t1);
    t0 = (nz - r.o[2]) * r.inv[2];
    t1 = (fz - r.o[2]) * r.inv[2];
    tmin = fmaxf(tmin, t0);
    tmax = fminf(tmax, t1);
    return tmin <= tmax ? tmin : __builtin_inff();
}

struct TriRec {
    V3 v0, e1, e2;
    int id;
};
PRR_DEV TriRec loadTri(const float4 *tris, uint32_t slot) {
    const float4 a = tris[3 * slot], b = tris[3 * slot + 1], c = tris[3 * slot + 2];
    TriRec t;
    t.v0 = V3(a.x, a.y, a.z);
    t.e1 = V3(a.w, b.x, b.y);
    t.e2 = V3(b.z, b.w, c.x);
    t.id = __float_as_int(c.y);
    return t;
}
// Triangle::hitTest (Triangle.hpp:46-64)
PRR_DEV bool triHit(const TriRec &tr, const V3 &o, const V3 &d, float &t, float &u, float &v) {
    V3 P = cross(d, tr.e2);
    float det = dot(P, tr.e1);
    if (fabsf(det) < 1e-6f) return false;
    float inv = 1.f / det;
    V3 T = o - tr.v0;
    u = dot(P, T) * inv;
    if (u < 0.f) return false;
    V3 Q = cross(T, tr.e1);
    v = dot(Q, d) * inv;
    if (v < 0.f || u + v > 1.f) return false;
    t = dot(Q, tr.e2) * inv;
    return t > 1e-3f;
}
// Triangle::intersectT (Triangle.hpp:75-92): occluder in (1e-3, tMax)
PRR_DEV bool triAny(const TriRec &tr, const V3 &o, const V3 &d, float tMax) {
    V3 P = cross(d, tr.e2);
    float det = dot(P, tr.e1);
    if (fabsf(det) < 1e-6f) return false;
    float inv = 1.f / det;
    V3 T = o - tr.v0;
    float u = dot(P, T) * inv;
    if (u < 0.f || u > 1.f) return false;
    V3 Q = cross(T, tr.e1);
    float v = dot(Q, d) * inv;
    if (v < 0.f || u + v > 1.f) return false;
    float t = dot(Q, tr.e2) * inv;
    return t > 1e-3f && t < tMax;
}
// Sphere::getIntersection (Sphere.hpp:22-44) with solveQuadratic
// (global.hpp:19-33) whose sqrt is the double overload.
PRR_DEV bool sphHit(const float4 &s, const V3 &o, const V3 &d, float &t) {
    V3 c(s.x, s.y, s.z);
    V3 L = o - c;
    float a = dot(d, d);
    float b = 2 * dot(d, L);
    float cc = dot(L, L) - s.w * s.w;
    float discr = b * b - 4 * a * cc;
    if (discr < 0) return false;
    float x0, x1;
    if (discr == 0) {
        x0 = x1 = (float) (-0.5 * (double) b / (double) a);
    } else {
        double sq = sqrt((double) discr);
        float q = (float) ((b > 0) ? -0.5 * ((double) b + sq) : -0.5 * ((double) b - sq));
        x0 = q / a;
        x1 = cc / q;
    }
    if (x0 > x1) {
        float tt = x0;
        x0 = x1;
        x1 = tt;
    }
    if (x0 < 0) x0 = x1;
    if (x0 < 0) return false;
    t = x0;
    return true;
}

// Per-lane stack: the top `depth` entries (a power of two) live in LDS laid
// out [slot][lane] (uint2 node ref + box t), the rest in a global spill
// region of kSpill entries per lane.
struct LaneStack {
    uint2 *lds;    // base of this block's LDS stack
    uint2 *spill;  // this lane's spill region (kSpill entries)
    int lane;      // threadIdx.x
    int sp;
    int depth;     // LDS entries per lane (power of two)
    PRR_DEV void push(uint32_t ref, float t) {
        const uint2 e = make_uint2(ref, __float_as_uint(t));
        const int slot = (sp & (depth - 1)) * kTraceBlock + lane;
        if (sp >= depth) spill[sp - depth] = lds[slot];
        lds[slot] = e;
        sp++;
    }
    PRR_DEV uint2 pop() {
        sp--;
        const int slot = (sp & (depth - 1)) * kTraceBlock + lane;
        uint2 e = lds[slot];
        if (sp >= depth) lds[slot] = spill[sp - depth];
        return e;
    }
};

// Any-hit stack: node references only (an any-hit pop tests no entry t),
// so the same LDS bytes hold twice the entries; [slot][lane] layout and a
// global spill region of 2 * kSpill entries per lane (the same bytes as a
// LaneStack's).
struct RefStack {
    uint32_t *lds;
    uint32_t *spill;
    int lane;
    int sp;
    int depth;  // LDS entries per lane (power of two)
    PRR_DEV void push(uint32_t ref) {
        const int slot = (sp & (depth - 1)) * kTraceBlock + lane;
        if (sp >= depth) spill[sp - depth] = lds[slot];
        lds[slot] = ref;
        sp++;
    }
    // Pushes the first m (0..3) of a, b, c. Without wrap-around (sp + 3 <=
    // depth) the three LDS slots above the top are free, so all three are
    // written and sp advances by m: no branch on m.
    PRR_DEV void push3(uint32_t a, uint32_t b, uint32_t c, int m) {
        if (sp + 3 <= depth) {
            uint32_t *q = lds + sp * kTraceBlock + lane;
            q[0] = a;
            q[kTraceBlock] = b;
            q[2 * kTraceBlock] = c;
            sp += m;
        } else {
            if (m > 0) push(a);
            if (m > 1) push(b);
            if (m > 2) push(c);
        }
    }
    PRR_DEV uint32_t pop() {
        sp--;
        const int slot = (sp & (depth - 1)) * kTraceBlock + lane;
        const uint32_t e = lds[slot];
        if (sp >= depth) lds[slot] = spill[sp - depth];
        return e;
    }
};

// Stack over a block's LDS stack area (ldsDepth uint2 entries per lane) and
// a lane's spill region (kSpill uint2 entries).
template <class S>
PRR_DEV S makeStack(uint2 *lds, uint2 *spill, int lane, int ldsDepth);
template <>
PRR_DEV LaneStack makeStack<LaneStack>(uint2 *lds, uint2 *spill, int lane, int ldsDepth) {
    LaneStack st;
    st.lds = lds;
    st.spill = spill;
    st.lane = lane;
    st.depth = ldsDepth;
    st.sp = 0;
    return st;
}
template <>
PRR_DEV RefStack makeStack<RefStack>(uint2 *lds, uint2 *spill, int lane, int ldsDepth) {
    RefStack st;
    st.lds = reinterpret_cast<uint32_t *>(lds);
    st.spill = reinterpret_cast<uint32_t *>(spill);
    st.lane = lane;
    st.depth = 2 * ldsDepth;
    st.sp = 0;
    return st;
}

struct HitOut {
    float t, u, v;
    int prim;  // global primitive id or -1
};

// BVHAccel::Intersect on the merged tree. Visits nodes in the reference's
// order; the running closest t persists across mesh boundaries, which only
// removes nodes the reference would reject at scene level (strict <), so the
// winner is unchanged.
PRR_DEV HitOut traceClosest(const SceneView &s, const V3 &o, const V3 &d, LaneStack &st) {
    HitOut h;
    h.t = __builtin_inff();
    h.u = h.v = 0.f;
    h.prim = -1;
    const RayPre r = precompute(o, d);
    st.sp = 0;
    uint32_t cur = s.root;
    while (true) {
        if (refIsLeaf(cur)) {
            const uint32_t first = refFirst(cur), cnt = refCount(cur);
            if (refSphere(cur)) {
                for (uint32_t i = 0; i < cnt; i++) {
                    float t;
                    if (sphHit(s.sph[first + i], o, d, t) && t < h.t) {
                        h.t = t;
                        h.u = h.v = 0.f;
                        h.prim = s.sph_id[first + i];
                    }
                }
            } else {
                for (uint32_t i = 0; i < cnt; i++) {
                    const TriRec tr = loadTri(s.tris, first + i);
                    float t, u, v;
                    if (triHit(tr, o, d, t, u, v) && t < h.t) {
                        h.t = t;
                        h.u = u;
                        h.v = v;
                        h.prim = tr.id;
                    }
                }
            }
        } else {
            const DNode n = s.nodes[cur];
            const float tL = slab(r, n.a.x, n.a.y, n.a.z, n.a.w, n.b.x, n.b.y);
            const float tR = slab(r, n.b.z, n.b.w, n.c.x, n.c.y, n.c.z, n.c.w);
            const uint32_t L = (uint32_t) n.d.x, R = (uint32_t) n.d.y;
            if (tL <= tR) {
                if (tL < h.t) {
                    if (tR < h.t) st.push(R, tR);
                    cur = L;
                    continue;
                }
            } else {
                if (tR < h.t) {
                    if (tL < h.t) st.push(L, tL);
                    cur = R;
                    continue;
                }
            }
        }
        bool found = false;
        while (st.sp > 0) {
            const uint2 e = st.pop();
            if (__uint_as_float(e.y) < h.t) {
                cur = e.x;
                found = true;
                break;
            }
        }
        if (!found) break;
    }
    return h;
}

// ---------------------------------------------------------------------------
// Resumable traversal for persistent kernels: one node per call, so a lane
// whose ray finished can fetch a new ray while the rest of its wave keeps
// traversing (while-while with per-lane replenishment).
// ---------------------------------------------------------------------------
struct TravState {
    RayPre r;
    V3 o, d;
    float tmax;  // any-hit range
    uint32_t cur;
    float curT;  // closest hit, speculative iteration: entry t of cur's box
    HitOut h;
    // quantised occlusion tree: this ray may take the short form of the node
    // test (anyPrepQ), with 1/d clamped to +-2^100
    bool fast;
    float invc[3];
};

template <class Stack>
PRR_DEV void travBegin(TravState &ts, const V3 &o, const V3 &d, float tmax, uint32_t root, Stack &st) {
    ts.o = o;
    ts.d = d;
    ts.tmax = tmax;
    ts.r = precompute(o, d);
    ts.cur = root;
    ts.curT = 0.f;
    ts.h.t = __builtin_inff();
    ts.h.u = ts.h.v = 0.f;
    ts.h.prim = -1;
    st.sp = 0;
}

// ts.cur after a pop whose entry fails the t test: the next step pops again.
// Never a node index (the node array is far smaller) and not a leaf ref.
constexpr uint32_t kCulled = 0x7FFFFFFEu;

// One pop of BVHAccel::Intersect's stack (BVH.cpp:399-407): the entry is
// visited next if its box entry t is below the closest hit so far, else the
// next step pops again; true when the stack is empty (the ray is finished).
// Popping one entry per step keeps the traversal loop free of a nested loop.
PRR_DEV bool popOne(TravState &ts, LaneStack &st) {
    if (st.sp == 0) return true;
    const uint2 e = st.pop();
    ts.cur = __uint_as_float(e.y) < ts.h.t ? e.x : kCulled;
    return false;
}

// BVHAccel::Intersect's primitive tests at leaf `cur` (in leaf order).
PRR_DEV void closestPrims(const SceneView &s, TravState &ts, uint32_t cur) {
    const uint32_t first = refFirst(cur), cnt = refCount(cur);
    if (refSphere(cur)) {
        for (uint32_t i = 0; i < cnt; i++) {
            float t;
            if (sphHit(s.sph[first + i], ts.o, ts.d, t) && t < ts.h.t) {
                ts.h.t = t;
                ts.h.u = ts.h.v = 0.f;
                ts.h.prim = s.sph_id[first + i];
            }
        }
    } else {
        // four triangles' loads in flight at once, tested in leaf order
        for (uint32_t i = 0; i < cnt; i += 4) {
            TriRec tr[4];
#pragma unroll
            for (int k = 0; k < 4; k++)
                if (i + k < cnt) tr[k] = loadTri(s.tris, first + i + k);
#pragma unroll
            for (int k = 0; k < 4; k++) {
                float t, u, v;
                if (i + k < cnt && triHit(tr[k], ts.o, ts.d, t, u, v) && t < ts.h.t) {
                    ts.h.t = t;
                    ts.h.u = u;
                    ts.h.v = v;
                    ts.h.prim = tr[k].id;
                }
            }
        }
    }
}

// BVHAccel::Intersect at a leaf (ts.cur): test its primitives, then pop.
PRR_DEV bool closestLeaf(const SceneView &s, TravState &ts, LaneStack &st) {
    closestPrims(s, ts, ts.cur);
    return popOne(ts, st);
}

// BVHAccel::Intersect at an interior node: near child first, the far one
// pushed with its entry t, or pop when neither box is closer than the hit.
// Near is the left child when tL <= tR; since tNear <= tFar, a far child that
// passes implies a near one that passes, as in the reference's nesting.
PRR_DEV bool closestInterior(const SceneView &s, TravState &ts, LaneStack &st) {
    if (ts.cur == kCulled) return popOne(ts, st);
    const DNode n = s.nodes[ts.cur];
    const float tL = slab(ts.r, n.a.x, n.a.y, n.a.z, n.a.w, n.b.x, n.b.y);
    const float tR = slab(ts.r, n.b.z, n.b.w, n.c.x, n.c.y, n.c.z, n.c.w);
    const uint32_t L = (uint32_t) n.d.x, R = (uint32_t) n.d.y;
    const bool lFirst = tL <= tR;
    const float tNear = lFirst ? tL : tR, tFar = lFirst ? tR : tL;
    const uint32_t rNear = lFirst ? L : R, rFar = lFirst ? R : L;
    if (tFar < ts.h.t) st.push(rFar, tFar);
    if (tNear < ts.h.t) {
        ts.cur = rNear;
        return false;
    }
    return popOne(ts, st);
}

// One step of BVHAccel::Intersect; returns true when the ray is finished.
PRR_DEV bool closestStep(const SceneView &s, TravState &ts, LaneStack &st) {
    return refIsLeaf(ts.cur) ? closestLeaf(s, ts, st) : closestInterior(s, ts, st);
}

// While-while iteration (Aila and Laine 2009): interior steps until this
// lane stands on a leaf, then that leaf, so a wave tests its leaves
// together. Each ray visits the same nodes in the same order as with
// closestStep. At most cap interior steps per call (cap <= 0: no limit); the
// descent then resumes at the next call. The interior step is written with
// one pop site so the loop carries few divergent exits.
PRR_DEV bool closestIter(const SceneView &s, TravState &ts, LaneStack &st, int cap) {
    // Loop-carried state in locals; the closest t does not change during
    // interior steps (only leaves update it).
    uint32_t cur = ts.cur;
    int sp = st.sp;
    const float best = ts.h.t;
    const char *nodes = reinterpret_cast<const char *>(s.nodes);
    uint2 *lds = st.lds + st.lane;
    const int dmask = st.depth - 1;
    bool done = false;
    const int lim = cap > 0 ? cap : -1;  // k counts up from 0: -1 is never reached
    for (int k = 0; !done && k != lim && !refIsLeaf(cur); k++) {  // one loop exit
        bool pop = true;
        if (cur != kCulled) {
            // node byte offset fits 32 bits (at most 2^26 nodes of 64 B)
            const DNode n = *reinterpret_cast<const DNode *>(nodes + (cur << 6));
            const float tL = slab(ts.r, n.a.x, n.a.y, n.a.z, n.a.w, n.b.x, n.b.y);
            const float tR = slab(ts.r, n.b.z, n.b.w, n.c.x, n.c.y, n.c.z, n.c.w);
            const bool lFirst = tL <= tR;
            const float tNear = lFirst ? tL : tR, tFar = lFirst ? tR : tL;
            const uint32_t rNear = (uint32_t) (lFirst ? n.d.x : n.d.y), rFar = (uint32_t) (lFirst ? n.d.y : n.d.x);
            if (tFar < best) {  // LaneStack::push
                uint2 *slot = lds + (sp & dmask) * kTraceBlock;
                if (sp >= st.depth) st.spill[sp - st.depth] = *slot;
                *slot = make_uint2(rFar, __float_as_uint(tFar));
                sp++;
            }
            pop = !(tNear < best);
            cur = rNear;
        }
        if (pop) {
            done = sp == 0;
            if (!done) {  // LaneStack::pop
                sp--;
                uint2 *slot = lds + (sp & dmask) * kTraceBlock;
                const uint2 e = *slot;
                if (sp >= st.depth) *slot = st.spill[sp - st.depth];
                cur = __uint_as_float(e.y) < best ? e.x : kCulled;
            }
        }
    }
    ts.cur = cur;
    st.sp = sp;
    if (done) return true;
    if (!refIsLeaf(cur)) return false;  // step cap reached
    return closestLeaf(s, ts, st);
}

// Entry and exit times of a box's six planes for one ray, as slab() forms
// them ((bound - o) * inv, near bound chosen by the ray's octant), with the
// NaN of a plane through the origin of a ray parallel to it (0 * inf) read as
// -inf for a near plane and +inf for a far one: slab() ignores such a term,
// and so do these values in its max / min.
struct Planes {
    float nx, fx, ny, fy, nz, fz;
};
PRR_DEV Planes planesOf(const RayPre &r, float x0, float x1, float y0, float y1, float z0, float z1) {
    Planes p;
    p.nx = fmaxf(((r.oct[0] ? x1 : x0) - r.o[0]) * r.inv[0], -__builtin_inff());
    p.fx = fminf(((r.oct[0] ? x0 : x1) - r.o[0]) * r.inv[0], __builtin_inff());
    p.ny = fmaxf(((r.oct[1] ? y1 : y0) - r.o[1]) * r.inv[1], -__builtin_inff());
    p.fy = fminf(((r.oct[1] ? y0 : y1) - r.o[1]) * r.inv[1], __builtin_inff());
    p.nz = fmaxf(((r.oct[2] ? z1 : z0) - r.o[2]) * r.inv[2], -__builtin_inff());
    p.fz = fminf(((r.oct[2] ? z0 : z1) - r.o[2]) * r.inv[2], __builtin_inff());
    return p;
}
// slab()'s result from the plane times.
PRR_DEV float entryOf(const Planes &p) {
    const float tmin = fmaxf(fmaxf(0.f, p.nx), fmaxf(p.ny, p.nz));
    const float tmax = fminf(p.fx, fminf(p.fy, p.fz));
    return tmin <= tmax ? tmin : __builtin_inff();
}
// Plane times of the union of two boxes: (bound - o) * inv is monotone in the
// bound (non-decreasing for inv > 0, non-increasing for inv < 0, and so are
// the NaN readings above), so the union's near plane time is the smaller of
// the two near times and its far time the larger, bit for bit what slab()
// computes from the union's own bounds.
PRR_DEV Planes unionOf(const Planes &a, const Planes &b) {
    Planes p;
    p.nx = fminf(a.nx, b.nx);
    p.fx = fmaxf(a.fx, b.fx);
    p.ny = fminf(a.ny, b.ny);
    p.fy = fmaxf(a.fy, b.fy);
    p.nz = fminf(a.nz, b.nz);
    p.fz = fmaxf(a.fz, b.fz);
    return p;
}

// closestIter on the two-level records (DNode2): one 128-byte line per step
// for up to two levels of BVHAccel::Intersect. At binary node cur the
// children's boxes are the unions of their pairs' (unionOf), so tL and tR are
// slab()'s; the far child is pushed with its entry t and the near one
// entered as in closestIter, and when the near child is interior its own
// step follows at once from the same line (its children are the near pair),
// with the same closest t, since no leaf lies between the two steps. Pushes,
// pops and leaves are therefore the reference's, in its order.
PRR_DEV bool closestIter2(const SceneView &s, TravState &ts, LaneStack &st, int cap) {
    uint32_t cur = ts.cur;
    int sp = st.sp;
    const float best = ts.h.t;
    uint2 *lds = st.lds + st.lane;
    const int dmask = st.depth - 1;
    bool done = false;
    const int lim = cap > 0 ? cap : -1;
    for (int k = 0; !done && k != lim && !refIsLeaf(cur); k++) {
        bool pop = true;
        if (cur != kCulled) {
            const DNode2 n = s.nodes2[cur];
            const Planes g0 = planesOf(ts.r, n.a0.x, n.a0.y, n.a0.z, n.a0.w, n.b0.x, n.b0.y);
            const Planes g1 = planesOf(ts.r, n.b0.z, n.b0.w, n.c0.x, n.c0.y, n.c0.z, n.c0.w);
            const Planes g2 = planesOf(ts.r, n.a1.x, n.a1.y, n.a1.z, n.a1.w, n.b1.x, n.b1.y);
            const Planes g3 = planesOf(ts.r, n.b1.z, n.b1.w, n.c1.x, n.c1.y, n.c1.z, n.c1.w);
            const float tL = entryOf(unionOf(g0, g1)), tR = entryOf(unionOf(g2, g3));
            const bool lFirst = tL <= tR;
            const float tNear = lFirst ? tL : tR, tFar = lFirst ? tR : tL;
            const uint32_t rNear = lFirst ? n.cref.x : n.cref.y, rFar = lFirst ? n.cref.y : n.cref.x;
            if (tFar < best) {  // LaneStack::push
                uint2 *slot = lds + (sp & dmask) * kTraceBlock;
                if (sp >= st.depth) st.spill[sp - st.depth] = *slot;
                *slot = make_uint2(rFar, __float_as_uint(tFar));
                sp++;
            }
            if (tNear < best) {
                pop = false;
                cur = rNear;
                if (!refIsLeaf(rNear)) {  // the near child's step: its children are the near pair
                    const float tA = entryOf(lFirst ? g0 : g2), tB = entryOf(lFirst ? g1 : g3);
                    const uint32_t rA = lFirst ? n.gref.x : n.gref.z, rB = lFirst ? n.gref.y : n.gref.w;
                    const bool aFirst = tA <= tB;
                    const float tN2 = aFirst ? tA : tB, tF2 = aFirst ? tB : tA;
                    if (tF2 < best) {
                        uint2 *slot = lds + (sp & dmask) * kTraceBlock;
                        if (sp >= st.depth) st.spill[sp - st.depth] = *slot;
                        *slot = make_uint2(aFirst ? rB : rA, __float_as_uint(tF2));
                        sp++;
                    }
                    if (tN2 < best)
                        cur = aFirst ? rA : rB;
                    else
                        pop = true;
                }
            }
        }
        if (pop) {
            done = sp == 0;
            if (!done) {  // LaneStack::pop
                sp--;
                uint2 *slot = lds + (sp & dmask) * kTraceBlock;
                const uint2 e = *slot;
                if (sp >= st.depth) *slot = st.spill[sp - st.depth];
                cur = __uint_as_float(e.y) < best ? e.x : kCulled;
            }
        }
    }
    ts.cur = cur;
    st.sp = sp;
    if (done) return true;
    if (!refIsLeaf(cur)) return false;  // step cap reached
    return closestLeaf(s, ts, st);
}

// Speculative while-while iteration (Aila and Laine 2009): a lane that
// reaches a leaf sets it aside and keeps stepping through interior nodes
// until it reaches a second leaf (or its stack runs out, or the step cap),
// so fewer lanes sit idle while the rest of the wave finishes its descent;
// then it tests the set-aside leaf and the second one, in that order.
//
// The result is BVHAccel::Intersect's, ids and ties included. The steps taken
// after the first leaf use the closest t from before that leaf's tests, which
// is never smaller than the reference's at the same point, so they visit a
// superset of the reference's nodes in the same order (near/far choices
// depend on boxes only) and push no entry the reference pops uncullled. A
// leaf is tested only if its own entry t is still below the closest hit when
// its turn comes: the slab test is monotone in the box bounds and every box
// of the tree contains its children's, so its ancestors' entry t are no
// larger and the reference, whose closest t at their visits was no smaller,
// entered them and reaches the leaf with the same closest t. Both therefore
// test the same leaves in the same order with the same running closest hit.
PRR_DEV bool closestIterSpec(const SceneView &s, TravState &ts, LaneStack &st, int cap) {
    uint32_t cur = ts.cur;
    float curT = ts.curT;  // entry t of cur's box (0: the root)
    uint32_t aside = kCulled;  // set-aside leaf (kCulled: none)
    float asideT = 0.f;
    int sp = st.sp;
    const float best = ts.h.t;
    const char *nodes = reinterpret_cast<const char *>(s.nodes);
    uint2 *lds = st.lds + st.lane;
    const int dmask = st.depth - 1;
    bool done = false;
    const int lim = cap > 0 ? cap : -1;
    for (int k = 0; !done && k != lim; k++) {
        bool pop = true;
        if (refIsLeaf(cur)) {
            if (aside != kCulled) break;  // two leaves in hand
            aside = cur;
            asideT = curT;
        } else if (cur != kCulled) {
            const DNode n = *reinterpret_cast<const DNode *>(nodes + (cur << 6));
            const float tL = slab(ts.r, n.a.x, n.a.y, n.a.z, n.a.w, n.b.x, n.b.y);
            const float tR = slab(ts.r, n.b.z, n.b.w, n.c.x, n.c.y, n.c.z, n.c.w);
            const bool lFirst = tL <= tR;
            const float tNear = lFirst ? tL : tR, tFar = lFirst ? tR : tL;
            const uint32_t rNear = (uint32_t) (lFirst ? n.d.x : n.d.y), rFar = (uint32_t) (lFirst ? n.d.y : n.d.x);
            if (tFar < best) {  // LaneStack::push
                uint2 *slot = lds + (sp & dmask) * kTraceBlock;
                if (sp >= st.depth) st.spill[sp - st.depth] = *slot;
                *slot = make_uint2(rFar, __float_as_uint(tFar));
                sp++;
            }
            pop = !(tNear < best);
            cur = rNear;
            curT = tNear;
        }
        if (pop) {
            done = sp == 0;
            if (!done) {  // LaneStack::pop
                sp--;
                uint2 *slot = lds + (sp & dmask) * kTraceBlock;
                const uint2 e = *slot;
                if (sp >= st.depth) *slot = st.spill[sp - st.depth];
                curT = __uint_as_float(e.y);
                cur = curT < best ? e.x : kCulled;
            }
        }
    }
    st.sp = sp;
    if (aside != kCulled && asideT < ts.h.t) closestPrims(s, ts, aside);
    if (!done && refIsLeaf(cur)) {
        if (curT < ts.h.t) closestPrims(s, ts, cur);
        // the next entry (BVH.cpp:399-407), culled against the updated closest t
        done = st.sp == 0;
        if (!done) {
            const uint2 e = st.pop();
            curT = __uint_as_float(e.y);
            cur = curT < ts.h.t ? e.x : kCulled;
        }
    }
    ts.cur = cur;
    ts.curT = curT;
    return done;
}

// Root test of BVHAccel::IntersectP; false = the ray misses the scene box.
PRR_DEV bool anyBegin(const SceneView &s, TravState &ts) {
    const float tb = slab(ts.r, s.root_box[0], s.root_box[1], s.root_box[2], s.root_box[3], s.root_box[4],
                          s.root_box[5]);
    return tb < ts.tmax;
}

// Any-hit steps return kAnyGo (continue at ts.cur), kAnyClear (stack empty:
// not occluded) or kAnyHit (occluded). The answer travels in the return
// value, not in TravState, so the compiler never merges stores to different
// TravState fields into one address-selected store (which would put the
// traversal state in scratch memory).
constexpr int kAnyGo = 0, kAnyClear = 1, kAnyHit = 2;

// BVHAccel::IntersectP at a leaf.
PRR_DEV uint32_t popRef(LaneStack &st) { return st.pop().x; }
PRR_DEV uint32_t popRef(RefStack &st) { return st.pop(); }

template <class Stack>
PRR_DEV int anyPop(TravState &ts, Stack &st) {
    if (st.sp == 0) return kAnyClear;
    ts.cur = popRef(st);
    return kAnyGo;
}

template <class Stack>
PRR_DEV int anyLeaf(const SceneView &s, TravState &ts, Stack &st) {
    const uint32_t cur = ts.cur;
    const uint32_t first = refFirst(cur), cnt = refCount(cur);
    if (refSphere(cur)) {
        for (uint32_t i = 0; i < cnt; i++) {
            float t;
            if (sphHit(s.sph[first + i], ts.o, ts.d, t) && t < ts.tmax) return kAnyHit;
        }
    } else {
        // four triangles' loads in flight at once
        for (uint32_t i = 0; i < cnt; i += 4) {
            TriRec tr[4];
#pragma unroll
            for (int k = 0; k < 4; k++)
                if (i + k < cnt) tr[k] = loadTri(s.tris, first + i + k);
            bool hit = false;
#pragma unroll
            for (int k = 0; k < 4; k++) hit = hit || (i + k < cnt && triAny(tr[k], ts.o, ts.d, ts.tmax));
            if (hit) return kAnyHit;
        }
    }
    return anyPop(ts, st);
}

PRR_DEV int anyInterior(const SceneView &s, TravState &ts, LaneStack &st) {
    const DNode n = s.nodes[ts.cur];
    const float tL = slab(ts.r, n.a.x, n.a.y, n.a.z, n.a.w, n.b.x, n.b.y);
    const float tR = slab(ts.r, n.b.z, n.b.w, n.c.x, n.c.y, n.c.z, n.c.w);
    const bool hl = tL < ts.tmax, hr = tR < ts.tmax;
    const uint32_t L = (uint32_t) n.d.x, R = (uint32_t) n.d.y;
    if (hl && hr) {
        const bool lf = tL <= tR;
        st.push(lf ? R : L, 0.f);
        ts.cur = lf ? L : R;
        return kAnyGo;
    }
    if (hl || hr) {
        ts.cur = hl ? L : R;
        return kAnyGo;
    }
    return anyPop(ts, st);
}

// One step of BVHAccel::IntersectP.
PRR_DEV int anyStep(const SceneView &s, TravState &ts, LaneStack &st) {
    return refIsLeaf(ts.cur) ? anyLeaf(s, ts, st) : anyInterior(s, ts, st);
}

// While-while iteration of IntersectP (see closestIter).
PRR_DEV int anyIter(const SceneView &s, TravState &ts, LaneStack &st, int cap) {
    int k = 0;
    while (!refIsLeaf(ts.cur)) {
        const int r = anyInterior(s, ts, st);
        if (r != kAnyGo) return r;
        if (++k == cap) return kAnyGo;
    }
    return anyLeaf(s, ts, st);
}

// Children of a four-wide node entered with box entry t0..t3 (infinity: not
// entered): the nearest next, the other entered ones pushed in child order,
// or a pop when none is entered before tmax.
PRR_DEV int anyChoose4(TravState &ts, RefStack &st, const uint4 rf, float t0, float t1, float t2, float t3) {
    const float tm = ts.tmax;
    // nearest entered child (slot index, so equal refs never arise)
    int bk = -1;
    float bt = tm;
    bk = t0 < bt ? 0 : bk;
    bt = t0 < bt ? t0 : bt;
    bk = t1 < bt ? 1 : bk;
    bt = t1 < bt ? t1 : bt;
    bk = t2 < bt ? 2 : bk;
    bt = t2 < bt ? t2 : bt;
    bk = t3 < bt ? 3 : bk;
    if (bk < 0) return anyPop(ts, st);
    const bool h0 = t0 < tm && bk != 0, h1 = t1 < tm && bk != 1, h2 = t2 < tm && bk != 2, h3 = t3 < tm && bk != 3;
    // compact the other entered children into p0..p2
    const uint32_t p0 = h0 ? rf.x : (h1 ? rf.y : (h2 ? rf.z : rf.w));
    const uint32_t p1 = h0 ? (h1 ? rf.y : (h2 ? rf.z : rf.w)) : (h1 ? (h2 ? rf.z : rf.w) : rf.w);
    const uint32_t p2 = bk == 3 ? rf.z : rf.w;  // the third, when all three others were entered
    const int m = (int) h0 + (int) h1 + (int) h2 + (int) h3;
    st.push3(p0, p1, p2, m);
    ts.cur = bk == 0 ? rf.x : (bk == 1 ? rf.y : (bk == 2 ? rf.z : rf.w));
    return kAnyGo;
}

// Any-hit at a node of the occlusion tree (DNode4): the nearest child whose
// box is entered before tmax next, the other entered children pushed (in
// child order), or a pop when none is entered. Empty slots hold an inverted
// box, which the slab test never enters. The answer equals
// BVHAccel::IntersectP's on the binary tree: the occlusion tree keeps its
// leaves and skips only boxes that contain their children's boxes.
PRR_DEV int anyInterior4(const SceneView &s, TravState &ts, RefStack &st) {
    const DNode4 *n = s.nodes4 + ts.cur;
    const float4 lx = n->lox, hx = n->hix, ly = n->loy, hy = n->hiy, lz = n->loz, hz = n->hiz;
    const float t0 = slab(ts.r, lx.x, hx.x, ly.x, hy.x, lz.x, hz.x);
    const float t1 = slab(ts.r, lx.y, hx.y, ly.y, hy.y, lz.y, hz.y);
    const float t2 = slab(ts.r, lx.z, hx.z, ly.z, hy.z, lz.z, hz.z);
    const float t3 = slab(ts.r, lx.w, hx.w, ly.w, hy.w, lz.w, hz.w);
    return anyChoose4(ts, st, n->ref, t0, t1, t2, t3);
}

PRR_DEV int anyStep4(const SceneView &s, TravState &ts, RefStack &st) {
    return refIsLeaf(ts.cur) ? anyLeaf(s, ts, st) : anyInterior4(s, ts, st);
}

// While-while iteration on the occlusion tree (see closestIter).
PRR_DEV int anyIter4(const SceneView &s, TravState &ts, RefStack &st, int cap) {
    int k = 0;
    while (!refIsLeaf(ts.cur)) {
        const int r = anyInterior4(s, ts, st);
        if (r != kAnyGo) return r;
        if (++k == cap) return kAnyGo;
    }
    return anyLeaf(s, ts, st);
}

// BVHAccel::IntersectP through the occlusion tree; the scene box test first
// (BVH.cpp:437-438).
PRR_DEV bool traceAny4(const SceneView &s, const V3 &o, const V3 &d, float tMax, RefStack &st) {
    TravState ts;
    travBegin(ts, o, d, tMax, s.root4, st);
    if (!anyBegin(s, ts)) return false;
    int r;
    while ((r = anyStep4(s, ts, st)) == kAnyGo) {
    }
    return r == kAnyHit;
}

// slab() with the near and far plane of each axis already chosen.
PRR_DEV float slabNF(const RayPre &r, float nx, float fx, float ny, float fy, float nz, float fz) {
    float tmin = 0.f, tmax = __builtin_inff();
    float t0 = (nx - r.o[0]) * r.inv[0], t1 = (fx - r.o[0]) * r.inv[0];
    tmin = fmaxf(tmin, t0);
    tmax = fminf(tmax, t1);
    t0 = (ny - r.o[1]) * r.inv[1];
    t1 = (fy - r.o[1]) * r.inv[1];
    tmin = fmaxf(tmin, t0);
    tmax = fminf(tmax, t1);
    t0 = (nz - r.o[2]) * r.inv[2];
    t1 = (fz - r.o[2]) * r.inv[2];
    tmin = fmaxf(tmin, t0);
    tmax = fminf(tmax, t1);
    return tmin <= tmax ? tmin : __builtin_inff();
}

// Bound of byte k of a quantised word: fma(q, scale, origin), exact.
PRR_DEV float decQ(uint32_t w, int k, float s, float o) {
    return __builtin_fmaf((float) ((w >> (8 * k)) & 255u), s, o);
}

// Any-hit at a node of the quantised occlusion tree (DNode4Q): the decoded
// boxes contain the children's boxes, so every child the exact test would
// enter is entered (the slab test is monotone in the bounds); leaves repeat
// the exact test (anyLeafQ), so the reached leaves, and the answer, are the
// binary tree's. The near/far choice per axis is made once per node on the
// packed words. Child order and pushes as anyInterior4.
PRR_DEV int anyInterior4Q(const SceneView &s, TravState &ts, RefStack &st) {
    const DNode4Q nq = s.nodes4q[ts.cur];
    const float ox = __uint_as_float(nq.w0.x), oy = __uint_as_float(nq.w0.y), oz = __uint_as_float(nq.w0.z);
    const float sx = __uint_as_float(nq.w0.w), sy = __uint_as_float(nq.w1.x), sz = __uint_as_float(nq.w1.y);
    const uint32_t nxw = ts.r.oct[0] ? nq.w1.w : nq.w1.z, fxw = ts.r.oct[0] ? nq.w1.z : nq.w1.w;
    const uint32_t nyw = ts.r.oct[1] ? nq.w2.y : nq.w2.x, fyw = ts.r.oct[1] ? nq.w2.x : nq.w2.y;
    const uint32_t nzw = ts.r.oct[2] ? nq.w2.w : nq.w2.z, fzw = ts.r.oct[2] ? nq.w2.z : nq.w2.w;
#define PRR_TQ(k)                                                                                             \
    slabNF(ts.r, decQ(nxw, k, sx, ox), decQ(fxw, k, sx, ox), decQ(nyw, k, sy, oy), decQ(fyw, k, sy, oy), \
           decQ(nzw, k, sz, oz), decQ(fzw, k, sz, oz))
    const float t0 = PRR_TQ(0), t1 = PRR_TQ(1), t2 = PRR_TQ(2), t3 = PRR_TQ(3);
#undef PRR_TQ
    return anyChoose4(ts, st, nq.w3, t0, t1, t2, t3);
}

// Whether a shadow ray may take the short form of the quantised node test:
// its origin inside the scene's safe box (NaN fails every comparison) and its
// largest direction component in [2^-10, 2^10]. Also sets the clamped 1/d.
PRR_DEV void anyPrepQ(const SceneView &s, TravState &ts) {
    const float m = fmaxf(fmaxf(fabsf(ts.d.x), fabsf(ts.d.y)), fabsf(ts.d.z));
    const bool in = ts.o.x >= s.safe_box[0] && ts.o.x <= s.safe_box[1] && ts.o.y >= s.safe_box[2] &&
                    ts.o.y <= s.safe_box[3] && ts.o.z >= s.safe_box[4] && ts.o.z <= s.safe_box[5];
    ts.fast = s.quant_fast != 0 && in && m >= 0x1p-10f && m <= 0x1p10f;
#pragma unroll
    for (int a = 0; a < 3; a++) ts.invc[a] = fminf(fmaxf(ts.r.inv[a], -0x1p100f), 0x1p100f);
}

// Plane k of a quantised word in the short form: fma(q, scale / d, (origin - o) / d).
PRR_DEV float planeF(uint32_t w, int k, float a, float b) {
    return __builtin_fmaf((float) ((w >> (8 * k)) & 255u), a, b);
}

// anyInterior4Q in the short form, for rays with ts.fast: per axis
// a = scale * (1/d) and b = (origin - o) * (1/d) once per node, then each of
// a child's six planes is one fma of its byte instead of a decode, a
// subtraction and a product. Against the reference's own test of any leaf
// below (t = (plane - o) * (1/d), rounded twice), this form's entry and exit
// times for the decoded box differ from the exact ones for that box by at
// most about 6 ulps of the distances involved (the decoded plane is at most
// the node extent from the node origin, which is within the scene extent R
// plus the largest coordinate of the origin), and the quantiser widened
// every decoded box by 32 ulps of those distances, so the decoded box's
// entry is never later and its exit never earlier than the reference's for a
// leaf it contains. A 1/d clamped to +-2^100 (d = 0 or tiny) keeps the
// planes' order on that axis with exit times above 2^80 whenever the origin
// lies in the slab, while the largest direction component (>= 2^-10) bounds
// every exit time the reference could compare against below that. Every leaf
// the exact test enters is therefore reached; the leaf test is exact, so the
// answer is unchanged.
PRR_DEV int anyInterior4QF(const SceneView &s, TravState &ts, RefStack &st) {
    const DNode4Q nq = s.nodes4q[ts.cur];
    const float ax = __uint_as_float(nq.w0.w) * ts.invc[0], bx = (__uint_as_float(nq.w0.x) - ts.r.o[0]) * ts.invc[0];
    const float ay = __uint_as_float(nq.w1.x) * ts.invc[1], by = (__uint_as_float(nq.w0.y) - ts.r.o[1]) * ts.invc[1];
    const float az = __uint_as_float(nq.w1.y) * ts.invc[2], bz = (__uint_as_float(nq.w0.z) - ts.r.o[2]) * ts.invc[2];
    const uint32_t nxw = ts.r.oct[0] ? nq.w1.w : nq.w1.z, fxw = ts.r.oct[0] ? nq.w1.z : nq.w1.w;
    const uint32_t nyw = ts.r.oct[1] ? nq.w2.y : nq.w2.x, fyw = ts.r.oct[1] ? nq.w2.x : nq.w2.y;
    const uint32_t nzw = ts.r.oct[2] ? nq.w2.w : nq.w2.z, fzw = ts.r.oct[2] ? nq.w2.z : nq.w2.w;
    float t[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const float tmin = fmaxf(fmaxf(0.f, planeF(nxw, k, ax, bx)), fmaxf(planeF(nyw, k, ay, by), planeF(nzw, k, az, bz)));
        const float tmax = fminf(planeF(fxw, k, ax, bx), fminf(planeF(fyw, k, ay, by), planeF(fzw, k, az, bz)));
        t[k] = tmin <= tmax ? tmin : __builtin_inff();
    }
    return anyChoose4(ts, st, nq.w3, t[0], t[1], t[2], t[3]);
}

// A leaf of the quantised tree: its exact box first (the binary tree's test
// of that leaf, BVH.cpp:445-459), then its primitives. A triangle leaf's
// record holds its box and its triangles, read together (up to four
// triangles at once) and tested in leaf order.
PRR_DEV int anyLeafQ(const SceneView &s, TravState &ts, RefStack &st) {
    const float4 *rec = s.leavesq + refFirst(ts.cur);
    const uint32_t cnt = refCount(ts.cur);
    const float4 a = rec[0], b = rec[1];
    if (refSphere(ts.cur)) {
        if (!(slab(ts.r, a.x, a.y, a.z, a.w, b.x, b.y) < ts.tmax)) return anyPop(ts, st);
        ts.cur = __float_as_uint(b.z);
        return anyLeaf(s, ts, st);
    }
    const float4 *tris = rec + kLeafQHead;
    for (uint32_t i = 0; i < cnt; i += 4) {
        TriRec tr[4];
#pragma unroll
        for (int k = 0; k < 4; k++)
            if (i + k < cnt) tr[k] = loadTri(tris, i + k);
        if (i == 0 && !(slab(ts.r, a.x, a.y, a.z, a.w, b.x, b.y) < ts.tmax)) return anyPop(ts, st);
        bool hit = false;
#pragma unroll
        for (int k = 0; k < 4; k++) hit = hit || (i + k < cnt && triAny(tr[k], ts.o, ts.d, ts.tmax));
        if (hit) return kAnyHit;
    }
    return anyPop(ts, st);
}

PRR_DEV int anyNode4Q(const SceneView &s, TravState &ts, RefStack &st) {
    return ts.fast ? anyInterior4QF(s, ts, st) : anyInterior4Q(s, ts, st);
}

// Steps of IntersectP on the quantised occlusion tree; the ray's TravState
// was prepared by anyPrepQ after travBegin.
PRR_DEV int anyStep4Q(const SceneView &s, TravState &ts, RefStack &st) {
    return refIsLeaf(ts.cur) ? anyLeafQ(s, ts, st) : anyNode4Q(s, ts, st);
}

PRR_DEV int anyIter4Q(const SceneView &s, TravState &ts, RefStack &st, int cap) {
    int k = 0;
    while (!refIsLeaf(ts.cur)) {
        const int r = anyNode4Q(s, ts, st);
        if (r != kAnyGo) return r;
        if (++k == cap) return kAnyGo;
    }
    return anyLeafQ(s, ts, st);
}

PRR_DEV bool traceAny4Q(const SceneView &s, const V3 &o, const V3 &d, float tMax, RefStack &st) {
    TravState ts;
    travBegin(ts, o, d, tMax, s.root4q, st);
    anyPrepQ(s, ts);
    if (!anyBegin(s, ts)) return false;
    int r;
    while ((r = anyStep4Q(s, ts, st)) == kAnyGo) {
    }
    return r == kAnyHit;
}

// Work fetch through a wave-private range of queue entries. A wave claims a
// run of consecutive entries from *counter with one atomic and hands them to
// its idle lanes over several refills; the run is sized by guided
// self-scheduling, max(minRun, min(maxRun, remaining / (4 * waves))), so that
// the atomics on the one shared counter fall from one per refill to a few
// per wave while the runs shrink to a single refill's worth near the end of
// the queue (no wave sits on a long run while others are idle). `remaining`
// uses the counter value the wave saw on its last claim. Entries come out in
// queue order within a wave. Lanes with need get an index (>= n once the
// queue is exhausted); others get -1.
struct WaveRange {
    int64_t lo = 0, hi = 0;  // entries [lo, hi) claimed and not yet handed out
    int64_t seen = 0;        // counter value after this wave's last claim
};

// minRun must be at least the number of lanes that can need an entry at once.
PRR_DEV int64_t waveTake(WaveRange &r, uint32_t *counter, int64_t n, bool need, int64_t maxRun, int64_t waves,
                         int64_t minRun = 64) {
    const uint64_t mask = __ballot(need);
    if (mask == 0) return -1;
    const int lane = (int) __lane_id();
    const int64_t k = __popcll(mask);
    const int64_t rank = __popcll(mask & ((1ull << lane) - 1ull));
    const int64_t have = r.hi - r.lo;
    int64_t idx;
    if (have >= k) {
        idx = r.lo + rank;
        r.lo += k;
    } else if (r.seen >= n) {  // queue exhausted when this wave last claimed
        idx = rank < have ? r.lo + rank : n;
        r.lo = r.hi;
    } else {
        int64_t run = (n - r.seen) / (4 * waves);
        run = run > maxRun ? maxRun : run;
        run = run < minRun ? minRun : run;
        const int leader = __ffsll((unsigned long long) mask) - 1;
        uint32_t b = 0;
        if (lane == leader) b = atomicAdd(counter, (uint32_t) run);
        const int64_t base = (int64_t) __shfl(b, leader);
        r.seen = base + run;
        const int64_t end = r.seen < n ? r.seen : n;
        idx = rank < have ? r.lo + rank : base + (rank - have);
        r.lo = base + (k - have);
        r.hi = end > r.lo ? end : r.lo;
    }
    return need ? idx : -1;
}

}  // namespace prr
