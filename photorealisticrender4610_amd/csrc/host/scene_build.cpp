// scene_build.cpp — see scene_build.hpp.
#include "scene_build.hpp"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <limits>
#include <thread>

namespace prr {
namespace host {

static const float kInfF = std::numeric_limits<float>::infinity();

Box::Box() {
    for (int a = 0; a < 3; a++) {
        mn[a] = kInfF;   // (float) std::numeric_limits<double>::max()    (Bounds3.hpp:11-16)
        mx[a] = -kInfF;  // (float) std::numeric_limits<double>::lowest()
    }
}
Box::Box(const float *a, const float *b) {
    for (int k = 0; k < 3; k++) {
        mn[k] = fminf(a[k], b[k]);
        mx[k] = fmaxf(a[k], b[k]);
    }
}
void Box::unite(const Box &o) {  // Union(Bounds3, Bounds3) (Bounds3.hpp:111-117)
    for (int k = 0; k < 3; k++) {
        mn[k] = std::min(mn[k], o.mn[k]);
        mx[k] = std::max(mx[k], o.mx[k]);
    }
}
void Box::unite(const float *p) {  // Union(Bounds3, Vector3f) (Bounds3.hpp:119-125)
    for (int k = 0; k < 3; k++) {
        mn[k] = std::min(mn[k], p[k]);
        mx[k] = std::max(mx[k], p[k]);
    }
}
float Box::area() const {  // Bounds3::SurfaceArea (Bounds3.hpp:37-41)
    float d0 = mx[0] - mn[0], d1 = mx[1] - mn[1], d2 = mx[2] - mn[2];
    return 2 * (d0 * d1 + d0 * d2 + d1 * d2);
}
void Box::centroid(float *c) const {
    for (int k = 0; k < 3; k++) c[k] = mn[k] * 0.5f + mx[k] * 0.5f;
}
int Box::maxExtent() const {  // Bounds3.hpp:26-35
    float d0 = mx[0] - mn[0], d1 = mx[1] - mn[1], d2 = mx[2] - mn[2];
    if (d0 > d1 && d0 > d2) return 0;
    if (d1 > d2) return 1;
    return 2;
}

// ---------------------------------------------------------------------------
// Binned SAH (BVH.cpp:66-256) + depth-first flatten (BVH.cpp:263-280)
// ---------------------------------------------------------------------------
namespace {
struct Builder {
    const std::vector<Box> &bounds;
    const std::vector<float> &areas;
    int maxPrims;
    struct BN {
        Box b;
        int child[2] = {-1, -1};
        float area = 0.f;
        int axis = 0, primOffset = 0, primCount = 0;
    };
    std::vector<BN> pool;
    std::vector<int> ordered;

    Builder(const std::vector<Box> &b, const std::vector<float> &a, int mp)
        : bounds(b), areas(a), maxPrims(std::min(255, mp)) {}

    int rec(std::vector<int> &prims, int start, int end) {
        pool.emplace_back();
        const int ni = (int) pool.size() - 1;
        const int count = end - start;
        std::vector<float> pc(3 * (size_t) count);
        Box bb, cb;
        for (int i = 0; i < count; i++) {
            const Box &pb = bounds[prims[start + i]];
            pb.centroid(&pc[3 * i]);
            bb.unite(pb);
            cb.unite(&pc[3 * i]);
        }
        pool[ni].b = bb;
        if (count == 1) {
            pool[ni].primOffset = (int) ordered.size();
            pool[ni].primCount = 1;
            pool[ni].area = areas[prims[start]];
            ordered.push_back(prims[start]);
            return ni;
        }
        if (count <= maxPrims) {
            pool[ni].primOffset = (int) ordered.size();
            pool[ni].primCount = count;
            float a = 0.f;
            for (int i = start; i < end; i++) {
                a += areas[prims[i]];
                ordered.push_back(prims[i]);
            }
            pool[ni].area = a;
            return ni;
        }
        const int axis = cb.maxExtent();
        pool[ni].axis = axis;
        int mid = (start + end) / 2;
        if (cb.mx[axis] == cb.mn[axis]) {
            // all centroids coincide: median split, no reordering
        } else if (count > 4) {
            constexpr int NB = 12;
            float bestCost = kInfF;
            int bestAxis = axis, bestSplit = NB / 2;
            const float parentArea = bb.area();
            for (int a = 0; a < 3; a++) {
                const float lo = cb.mn[a], hi = cb.mx[a];
                if (lo == hi) continue;
                int bcnt[NB] = {0};
                Box bbox[NB];
                const float span = hi - lo;
                for (int i = 0; i < count; i++) {
                    int b = std::clamp((int) (NB * (pc[3 * i + a] - lo) / span), 0, NB - 1);
                    bcnt[b]++;
                    bbox[b].unite(bounds[prims[start + i]]);
                }
                float lA[NB - 1], rA[NB - 1];
                int lC[NB - 1], rC[NB - 1];
                {
                    Box bl;
                    int c = 0;
                    for (int i = 0; i < NB - 1; i++) {
                        bl.unite(bbox[i]);
                        c += bcnt[i];
                        lA[i] = bl.area();
                        lC[i] = c;
                    }
                }
                {
                    Box br;
                    int c = 0;
                    for (int i = NB - 2; i >= 0; i--) {
                        br.unite(bbox[i + 1]);
                        c += bcnt[i + 1];
                        rA[i] = br.area();
                        rC[i] = c;
                    }
                }
                for (int i = 0; i < NB - 1; i++) {
                    if (lC[i] == 0 || rC[i] == 0) continue;
                    float cost = 1.f + (lC[i] * lA[i] + rC[i] * rA[i]) / parentArea;
                    if (cost < bestCost) {
                        bestCost = cost;
                        bestAxis = a;
                        bestSplit = i;
                    }
                }
            }
            const float lo = cb.mn[bestAxis];
            const float span = cb.mx[bestAxis] - lo;
            const float edge = lo + (bestSplit + 1) * span / NB;
            std::vector<int> tmp;
            tmp.reserve(count);
            int nl = 0;
            for (int i = 0; i < count; i++)
                if (pc[3 * i + bestAxis] < edge) {
                    tmp.push_back(prims[start + i]);
                    nl++;
                }
            for (int i = 0; i < count; i++)
                if (!(pc[3 * i + bestAxis] < edge)) tmp.push_back(prims[start + i]);
            if (nl == 0 || nl == count) {
                mid = (start + end) / 2;
            } else {
                std::copy(tmp.begin(), tmp.end(), prims.begin() + start);
                mid = start + nl;
            }
        } else {
            const std::vector<Box> &bd = bounds;
            std::nth_element(prims.data() + start, prims.data() + mid, prims.data() + end,
                             [axis, &bd](int x, int y) {
                                 float cx[3], cy[3];
                                 bd[x].centroid(cx);
                                 bd[y].centroid(cy);
                                 return cx[axis] < cy[axis];
                             });
        }
        const int c0 = rec(prims, start, mid);
        const int c1 = rec(prims, mid, end);
        pool[ni].child[0] = c0;
        pool[ni].child[1] = c1;
        pool[ni].area = pool[c0].area + pool[c1].area;
        return ni;
    }

    uint32_t flatten(RefBVH &out, int bi, uint32_t *off) {
        const BN &b = pool[bi];
        const uint32_t my = (*off)++;
        RefNode &n = out.nodes[my];
        n.b[0] = b.b.mn[0];
        n.b[1] = b.b.mx[0];
        n.b[2] = b.b.mn[1];
        n.b[3] = b.b.mx[1];
        n.b[4] = b.b.mn[2];
        n.b[5] = b.b.mx[2];
        out.nodeAreas[my] = b.area;
        if (b.primCount > 0) {
            out.nodes[my].first = (uint32_t) b.primOffset;
            out.nodes[my].meta = (uint32_t) b.primCount & 0x00FFFFFFu;
        } else {
            out.nodes[my].meta = (uint32_t) b.axis << 24;
            flatten(out, b.child[0], off);
            const uint32_t r = flatten(out, b.child[1], off);
            out.nodes[my].first = r;
        }
        return my;
    }
};
}  // namespace

RefBVH buildRefBVH(const std::vector<Box> &bounds, const std::vector<float> &areas, int maxPrims) {
    RefBVH out;
    if (bounds.empty()) return out;
    Builder b(bounds, areas, maxPrims);
    b.pool.reserve(2 * bounds.size());
    b.ordered.reserve(bounds.size());
    std::vector<int> prims(bounds.size());
    for (size_t i = 0; i < prims.size(); i++) prims[i] = (int) i;
    const int root = b.rec(prims, 0, (int) prims.size());
    out.nodes.resize(b.pool.size());
    out.nodeAreas.resize(b.pool.size());
    uint32_t off = 0;
    b.flatten(out, root, &off);
    out.totalArea = b.pool[root].area;
    out.ordered = std::move(b.ordered);
    return out;
}

// ---------------------------------------------------------------------------
// Scene
// ---------------------------------------------------------------------------
namespace {
struct TriPre {
    V3 v0, v1, v2, e1, e2, n;
    float area;
};
TriPre prepTri(const float *t) {  // Triangle ctor (Triangle.hpp:28-34)
    TriPre p;
    p.v0 = V3(t[0], t[1], t[2]);
    p.v1 = V3(t[3], t[4], t[5]);
    p.v2 = V3(t[6], t[7], t[8]);
    p.e1 = p.v1 - p.v0;
    p.e2 = p.v2 - p.v0;
    p.n = normalize(cross(p.e1, p.e2));
    p.area = norm(cross(p.e1, p.e2)) * 0.5f;
    return p;
}
Box triBox(const TriPre &p) {  // Triangle::getBounds (Triangle.hpp:182-184)
    float a[3] = {p.v0.x, p.v0.y, p.v0.z}, b[3] = {p.v1.x, p.v1.y, p.v1.z}, c[3] = {p.v2.x, p.v2.y, p.v2.z};
    Box bx(a, b);
    bx.unite(c);
    return bx;
}

struct Merger {
    HostScene &S;
    const prr_scene_desc *d;
    const std::vector<TriPre> &tp;
    std::vector<int> sphSlot;  // per sphere row: leaf slot

    uint32_t leafTri(int obj, const RefNode &n, int maxDepthHere) {
        (void) maxDepthHere;
        const RefBVH &m = S.meshes[obj];
        const int first = d->objects[obj].first;
        const uint32_t slot0 = (uint32_t) (S.tris.size() / 3);
        for (uint32_t i = 0; i < n.count(); i++) {
            const int id = first + m.ordered[n.first + i];
            const TriPre &t = tp[id];
            S.tris.push_back(make_float4(t.v0.x, t.v0.y, t.v0.z, t.e1.x));
            S.tris.push_back(make_float4(t.e1.y, t.e1.z, t.e2.x, t.e2.y));
            S.tris.push_back(make_float4(t.e2.z, __builtin_bit_cast(float, id), 0.f, 0.f));
        }
        return makeLeafRef(slot0, n.count(), false);
    }
    static void setBoxes(DNode &dn, const float *l, const float *r) {
        dn.a = make_float4(l[0], l[1], l[2], l[3]);
        dn.b = make_float4(l[4], l[5], r[0], r[1]);
        dn.c = make_float4(r[2], r[3], r[4], r[5]);
    }
    uint32_t mesh(int obj, int j, int depth) {
        const RefBVH &m = S.meshes[obj];
        const RefNode &n = m.nodes[j];
        if (n.leaf()) return leafTri(obj, n, depth);
        const int di = (int) S.nodes.size();
        S.nodes.emplace_back();
        S.maxDepth = std::max(S.maxDepth, depth + 1);
        const int l = j + 1, r = (int) n.first;
        const uint32_t L = mesh(obj, l, depth + 1);
        const uint32_t R = mesh(obj, r, depth + 1);
        setBoxes(S.nodes[di], m.nodes[l].b, m.nodes[r].b);
        S.nodes[di].d = make_int4((int) L, (int) R, 0, 0);
        return (uint32_t) di;
    }
    uint32_t object(int k, int depth) {
        if (d->objects[k].kind == PRR_OBJ_SPHERE) {
            const int row = d->objects[k].first;
            const uint32_t slot = (uint32_t) S.sph.size();
            S.sph.push_back(S.sphById[row]);
            S.sphId.push_back(S.numTris + row);
            return makeLeafRef(slot, 1, true);
        }
        return mesh(k, 0, depth);
    }
    uint32_t top(int i, int depth) {
        const RefNode &n = S.top.nodes[i];
        if (n.leaf()) return object(S.top.ordered[n.first], depth);
        const int di = (int) S.nodes.size();
        S.nodes.emplace_back();
        S.maxDepth = std::max(S.maxDepth, depth + 1);
        const int l = i + 1, r = (int) n.first;
        const uint32_t L = top(l, depth + 1);
        const uint32_t R = top(r, depth + 1);
        setBoxes(S.nodes[di], S.top.nodes[l].b, S.top.nodes[r].b);
        S.nodes[di].d = make_int4((int) L, (int) R, 0, 0);
        return (uint32_t) di;
    }
};

// Two-level records of the merged binary tree (DNode2): for every interior
// node, its children's children. The device rebuilds a child's box as the
// union of its pair's, which must equal the child's own box bit for bit (the
// reference's node bounds are unions of its primitives' bounds, so they do);
// false if some box differs, and then no record is kept.
bool buildTwoLevel(HostScene &S) {
    const float inf = std::numeric_limits<float>::infinity();
    S.nodes2.assign(S.nodes.size(), DNode2{});
    auto boxesOf = [](const DNode &n, float *l, float *r) {
        const float lb[6] = {n.a.x, n.a.y, n.a.z, n.a.w, n.b.x, n.b.y};
        const float rb[6] = {n.b.z, n.b.w, n.c.x, n.c.y, n.c.z, n.c.w};
        std::memcpy(l, lb, sizeof(lb));
        std::memcpy(r, rb, sizeof(rb));
    };
    for (size_t i = 0; i < S.nodes.size(); i++) {
        const DNode &n = S.nodes[i];
        float cb[2][6];
        boxesOf(n, cb[0], cb[1]);
        const uint32_t cr[2] = {(uint32_t) n.d.x, (uint32_t) n.d.y};
        float pb[2][2][6];
        uint32_t gr[2][2];
        for (int c = 0; c < 2; c++) {
            if (refIsLeaf(cr[c])) {
                std::memcpy(pb[c][0], cb[c], sizeof(cb[c]));
                for (int a = 0; a < 3; a++) {
                    pb[c][1][2 * a] = inf;
                    pb[c][1][2 * a + 1] = -inf;
                }
                gr[c][0] = cr[c];
                gr[c][1] = cr[c];
            } else {
                const DNode &m = S.nodes[cr[c]];
                boxesOf(m, pb[c][0], pb[c][1]);
                gr[c][0] = (uint32_t) m.d.x;
                gr[c][1] = (uint32_t) m.d.y;
                for (int a = 0; a < 3; a++)
                    if (std::min(pb[c][0][2 * a], pb[c][1][2 * a]) != cb[c][2 * a] ||
                        std::max(pb[c][0][2 * a + 1], pb[c][1][2 * a + 1]) != cb[c][2 * a + 1]) {
                        S.nodes2.clear();
                        return false;
                    }
            }
        }
        DNode2 &q = S.nodes2[i];
        q.a0 = make_float4(pb[0][0][0], pb[0][0][1], pb[0][0][2], pb[0][0][3]);
        q.b0 = make_float4(pb[0][0][4], pb[0][0][5], pb[0][1][0], pb[0][1][1]);
        q.c0 = make_float4(pb[0][1][2], pb[0][1][3], pb[0][1][4], pb[0][1][5]);
        q.a1 = make_float4(pb[1][0][0], pb[1][0][1], pb[1][0][2], pb[1][0][3]);
        q.b1 = make_float4(pb[1][0][4], pb[1][0][5], pb[1][1][0], pb[1][1][1]);
        q.c1 = make_float4(pb[1][1][2], pb[1][1][3], pb[1][1][4], pb[1][1][5]);
        q.gref = make_uint4(gr[0][0], gr[0][1], gr[1][0], gr[1][1]);
        q.cref = make_uint4(cr[0], cr[1], 0u, 0u);
    }
    return true;
}

// Occlusion tree: the merged binary tree with levels collapsed into nodes of
// up to four children, keeping its leaves. An any-hit query only asks whether
// some reached leaf holds an occluder. The slab test's arithmetic is monotone
// in the box bounds, so a box passes whenever a box inside it passes, and a
// leaf is then reached in the binary tree exactly when its own box passes.
// Skipping the test of a box that contains its children's boxes therefore
// changes no answer; a level where a child's box sticks out is kept.
struct Widener {
    HostScene &S;
    int maxPush = 0;
    struct Cand {
        float b[6];  // minX maxX minY maxY minZ maxZ
        uint32_t ref;
    };
    static void split(const DNode &n, Cand &l, Cand &r) {
        const float lb[6] = {n.a.x, n.a.y, n.a.z, n.a.w, n.b.x, n.b.y};
        const float rb[6] = {n.b.z, n.b.w, n.c.x, n.c.y, n.c.z, n.c.w};
        for (int k = 0; k < 6; k++) {
            l.b[k] = lb[k];
            r.b[k] = rb[k];
        }
        l.ref = (uint32_t) n.d.x;
        r.ref = (uint32_t) n.d.y;
    }
    static bool inside(const float *in, const float *out) {
        return in[0] >= out[0] && in[1] <= out[1] && in[2] >= out[2] && in[3] <= out[3] && in[4] >= out[4] &&
               in[5] <= out[5];
    }
    static double area(const float *b) {
        const double x = (double) b[1] - b[0], y = (double) b[3] - b[2], z = (double) b[5] - b[4];
        return x * y + y * z + z * x;
    }
    // pushes: stack entries a traversal may hold on arriving here
    uint32_t build(uint32_t ref, int pushes) {
        if (refIsLeaf(ref)) return ref;
        Cand c[4];
        int n = 2;
        split(S.nodes[ref], c[0], c[1]);
        while (n < 4) {  // open the largest interior child whose children lie inside its box
            int best = -1;
            double bestA = -1.0;
            for (int k = 0; k < n; k++) {
                if (refIsLeaf(c[k].ref)) continue;
                Cand l, r;
                split(S.nodes[c[k].ref], l, r);
                if (!inside(l.b, c[k].b) || !inside(r.b, c[k].b)) continue;
                const double a = area(c[k].b);
                if (a > bestA) {
                    bestA = a;
                    best = k;
                }
            }
            if (best < 0) break;
            Cand l, r;
            split(S.nodes[c[best].ref], l, r);
            c[best] = l;
            c[n++] = r;
        }
        for (int k = 0; k < n; k++)
            if (!refIsLeaf(c[k].ref) && n < 4) S.wideLevelsKept++;
        const uint32_t idx = (uint32_t) S.nodes4.size();
        S.nodes4.emplace_back();
        const int below = pushes + n - 1;
        maxPush = std::max(maxPush, below);
        uint32_t refs[4] = {kNoChild, kNoChild, kNoChild, kNoChild};
        for (int k = 0; k < n; k++) refs[k] = build(c[k].ref, below);
        const float inf = std::numeric_limits<float>::infinity();
        float lo[3][4], hi[3][4];
        for (int k = 0; k < 4; k++)
            for (int a = 0; a < 3; a++) {
                lo[a][k] = k < n ? c[k].b[2 * a] : inf;
                hi[a][k] = k < n ? c[k].b[2 * a + 1] : -inf;
            }
        DNode4 &w = S.nodes4[idx];
        w.lox = make_float4(lo[0][0], lo[0][1], lo[0][2], lo[0][3]);
        w.hix = make_float4(hi[0][0], hi[0][1], hi[0][2], hi[0][3]);
        w.loy = make_float4(lo[1][0], lo[1][1], lo[1][2], lo[1][3]);
        w.hiy = make_float4(hi[1][0], hi[1][1], hi[1][2], hi[1][3]);
        w.loz = make_float4(lo[2][0], lo[2][1], lo[2][2], lo[2][3]);
        w.hiz = make_float4(hi[2][0], hi[2][1], hi[2][2], hi[2][3]);
        w.ref = make_uint4(refs[0], refs[1], refs[2], refs[3]);
        w.pad = make_uint4(0, 0, 0, 0);
        return idx;
    }
};

// Quantised copy of the occlusion tree (DNode4Q, leaf records). Per node and axis
// the origin is the children's minimum and the scale a power of two with
// extent / scale <= 255; each child's bytes are moved outwards until the
// decoded bound fma(q, scale, origin), evaluated in float as on the device,
// lies outside the child's box widened by a margin, and the scale doubles if
// a byte runs out. The margin, 2^-19 * (node extent + scene extent R + largest
// scene coordinate), covers the rounding of the device's short form of the
// slab test (t = fma(q, scale / d, (origin - o) / d), anyInterior4QF) for ray
// origins within R of the scene box: it is at least 32 float ulps of every
// distance involved, while the short form errs by at most about 6 (see the
// device code), so a leaf the reference's exact test enters is entered by
// every decoded ancestor either way.
struct Quantiser {
    HostScene &S;
    static float dec(uint32_t q, float s, float o) { return std::fma((float) q, s, o); }
    // bytes of one axis for the n valid children; false if the scale is too fine
    static bool axis(const float *lo, const float *hi, int n, float o, float s, uint32_t &wlo, uint32_t &whi) {
        wlo = whi = 0;
        for (int k = 0; k < 4; k++) {
            uint32_t ql = 255, qh = 0;  // empty slot: inverted
            if (k < n) {
                double fl = std::floor(((double) lo[k] - o) / s), fh = std::ceil(((double) hi[k] - o) / s);
                ql = (uint32_t) std::min(255.0, std::max(0.0, fl));
                qh = (uint32_t) std::min(255.0, std::max(0.0, fh));
                while (ql > 0 && !(dec(ql, s, o) <= lo[k])) ql--;
                while (qh < 255 && !(dec(qh, s, o) >= hi[k])) qh++;
                if (!(dec(ql, s, o) <= lo[k]) || !(dec(qh, s, o) >= hi[k])) return false;
            }
            wlo |= ql << (8 * k);
            whi |= qh << (8 * k);
        }
        return true;
    }
    // leaf record: exact box, the binary tree's leaf ref, then a copy of a
    // triangle leaf's triangles; the child ref keeps the leaf's count and
    // sphere bit and points at the record
    bool leaf(const float *b, uint32_t ref, uint32_t &out, std::string &err) {
        const size_t off = S.leavesq.size();
        if (off > 0x07FFFFFFu) {
            err = "occlusion tree: leaf records exceed 2^27 float4s";
            return false;
        }
        S.leavesq.push_back(make_float4(b[0], b[1], b[2], b[3]));
        S.leavesq.push_back(make_float4(b[4], b[5], __builtin_bit_cast(float, ref), 0.f));
        if (!refSphere(ref))
            for (uint32_t i = 0; i < refCount(ref); i++)
                for (int j = 0; j < 3; j++) S.leavesq.push_back(S.tris[3 * (size_t) (refFirst(ref) + i) + j]);
        out = makeLeafRef((uint32_t) off, refCount(ref), refSphere(ref));
        return true;
    }
    // a float at or below x (at or above when up)
    static float roundOut(double x, bool up) {
        float f = (float) x;
        if (up ? (double) f < x : (double) f > x) f = std::nextafter(f, up ? HUGE_VALF : -HUGE_VALF);
        return f;
    }
    bool run(std::string &err) {
        S.nodes4q.assign(S.nodes4.size(), DNode4Q{});
        S.leavesq.clear();
        double R = 0.0, M = 0.0;  // scene extent, largest coordinate magnitude
        for (int a = 0; a < 3; a++) {
            R = std::max(R, (double) S.rootBox[2 * a + 1] - (double) S.rootBox[2 * a]);
            M = std::max(M, std::max(std::fabs((double) S.rootBox[2 * a]), std::fabs((double) S.rootBox[2 * a + 1])));
        }
        for (size_t i = 0; i < S.nodes4.size(); i++) {
            const DNode4 &w = S.nodes4[i];
            const float lo3[3][4] = {{w.lox.x, w.lox.y, w.lox.z, w.lox.w},
                                     {w.loy.x, w.loy.y, w.loy.z, w.loy.w},
                                     {w.loz.x, w.loz.y, w.loz.z, w.loz.w}};
            const float hi3[3][4] = {{w.hix.x, w.hix.y, w.hix.z, w.hix.w},
                                     {w.hiy.x, w.hiy.y, w.hiy.z, w.hiy.w},
                                     {w.hiz.x, w.hiz.y, w.hiz.z, w.hiz.w}};
            const uint32_t rf[4] = {w.ref.x, w.ref.y, w.ref.z, w.ref.w};
            int n = 0;
            while (n < 4 && rf[n] != kNoChild) n++;
            float org[3] = {0.f, 0.f, 0.f}, scl[3] = {1.f, 1.f, 1.f};
            uint32_t wq[6] = {0xFFFFFFFFu, 0u, 0xFFFFFFFFu, 0u, 0xFFFFFFFFu, 0u};
            for (int a = 0; a < 3 && n > 0; a++) {
                float mn0 = lo3[a][0], mx0 = hi3[a][0];
                for (int k = 1; k < n; k++) {
                    mn0 = std::min(mn0, lo3[a][k]);
                    mx0 = std::max(mx0, hi3[a][k]);
                }
                const double margin = std::ldexp((double) mx0 - (double) mn0 + R + M, -19);
                float lo[4], hi[4];  // the children's bounds widened by the margin
                for (int k = 0; k < n; k++) {
                    lo[k] = roundOut((double) lo3[a][k] - margin, false);
                    hi[k] = roundOut((double) hi3[a][k] + margin, true);
                }
                float mn = lo[0], mx = hi[0];
                for (int k = 1; k < n; k++) {
                    mn = std::min(mn, lo[k]);
                    mx = std::max(mx, hi[k]);
                }
                org[a] = mn;
                const double ext = (double) mx - (double) mn;
                int e = 0;
                if (ext > 0.0) {
                    std::frexp(ext / 255.0, &e);  // ext / 255 < 2^e
                    e = std::max(e, -126);
                }
                bool ok = false;
                for (int tries = 0; tries < 8 && !ok; tries++, e++) {
                    scl[a] = std::ldexp(1.f, e);
                    ok = axis(lo, hi, n, org[a], scl[a], wq[2 * a], wq[2 * a + 1]);
                }
                if (!ok) {
                    err = "occlusion tree: a child box does not fit its node's grid";
                    return false;
                }
            }
            uint32_t refs[4] = {kNoChild, kNoChild, kNoChild, kNoChild};
            for (int k = 0; k < n; k++) {
                const float b[6] = {lo3[0][k], hi3[0][k], lo3[1][k], hi3[1][k], lo3[2][k], hi3[2][k]};
                refs[k] = rf[k];
                if (refIsLeaf(rf[k]) && !leaf(b, rf[k], refs[k], err)) return false;
            }
            DNode4Q &q = S.nodes4q[i];
            q.w0 = make_uint4(__builtin_bit_cast(uint32_t, org[0]), __builtin_bit_cast(uint32_t, org[1]),
                              __builtin_bit_cast(uint32_t, org[2]), __builtin_bit_cast(uint32_t, scl[0]));
            q.w1 = make_uint4(__builtin_bit_cast(uint32_t, scl[1]), __builtin_bit_cast(uint32_t, scl[2]), wq[0], wq[1]);
            q.w2 = make_uint4(wq[2], wq[3], wq[4], wq[5]);
            q.w3 = make_uint4(refs[0], refs[1], refs[2], refs[3]);
        }
        S.root4q = S.root4;
        if (refIsLeaf(S.root4) && !leaf(S.rootBox, S.root4, S.root4q, err)) return false;
        if (S.nodes4q.empty()) S.nodes4q.emplace_back();
        if (S.leavesq.empty()) S.leavesq.emplace_back();
        return true;
    }
};

bool hasEmission(const prr_material_desc &m) {  // Material::hasEmission (Material.hpp:113-115)
    return norm(V3(m.emission[0], m.emission[1], m.emission[2])) > kEps;
}
}  // namespace

bool buildHostScene(const prr_scene_desc *d, HostScene &S, std::string &err) {
    if (!d) {
        err = "scene descriptor is NULL";
        return false;
    }
    if (d->num_objects <= 0 || !d->objects) {
        err = "scene has no objects";
        return false;
    }
    if (d->num_materials <= 0 || !d->materials) {
        err = "scene has no materials";
        return false;
    }
    if (d->num_triangles < 0 || d->num_spheres < 0 || (d->num_triangles > 0 && !d->triangles) ||
        (d->num_spheres > 0 && !d->spheres)) {
        err = "bad triangle / sphere arrays";
        return false;
    }
    if ((int64_t) d->num_triangles + d->num_spheres >= (int64_t) 1 << 27) {
        err = "too many primitives (limit 2^27)";
        return false;
    }
    if (d->num_point_lights > 1 || (d->num_point_lights > 0 && !d->point_lights)) {
        err = "at most one point light is supported";
        return false;
    }
    for (int k = 0; k < d->num_objects; k++) {
        const prr_object_desc &o = d->objects[k];
        if (o.material < 0 || o.material >= d->num_materials) {
            err = "object " + std::to_string(k) + ": material index out of range";
            return false;
        }
        if (o.kind == PRR_OBJ_MESH) {
            if (o.count <= 0 || o.first < 0 || (int64_t) o.first + o.count > d->num_triangles) {
                err = "object " + std::to_string(k) + ": triangle range out of bounds";
                return false;
            }
        } else if (o.kind == PRR_OBJ_SPHERE) {
            if (o.first < 0 || o.first >= d->num_spheres) {
                err = "object " + std::to_string(k) + ": sphere row out of range";
                return false;
            }
        } else {
            err = "object " + std::to_string(k) + ": unknown kind";
            return false;
        }
    }
    for (int i = 0; i < d->num_materials; i++) {
        const int t = d->materials[i].type;
        if (t < PRR_MAT_DIFFUSE || t > PRR_MAT_EMISSIVE) {
            err = "material " + std::to_string(i) + ": unknown type";
            return false;
        }
    }
    if (d->num_materials >= (1 << kPrimFlagShift)) {
        err = "too many materials";
        return false;
    }
    if (d->tri_flags && d->num_triangles > 0) {
        uint8_t any = 0;
        for (int i = 0; i < d->num_triangles; i++) any |= d->tri_flags[i];
        if (any & ~(PRR_TRI_SMOOTH_NORMALS | PRR_TRI_TANGENTS | PRR_TRI_UVS)) {
            err = "tri_flags: unknown bits";
            return false;
        }
        if (((any & PRR_TRI_SMOOTH_NORMALS) && !d->tri_normals) || ((any & PRR_TRI_UVS) && !d->tri_uvs) ||
            ((any & PRR_TRI_TANGENTS) && !d->tri_tangents)) {
            err = "tri_flags: attribute array missing";
            return false;
        }
    }
    if (d->num_textures < 0 || (d->num_textures > 0 && !d->textures)) {
        err = "bad texture list";
        return false;
    }
    for (int i = 0; i < d->num_textures; i++) {
        const prr_texture_desc &t = d->textures[i];
        if (t.width <= 0 || t.height <= 0 || t.channels < 1 || !t.data ||
            (int64_t) t.width * t.height * t.channels > ((int64_t) 1 << 31)) {
            err = "texture " + std::to_string(i) + ": bad size or data";
            return false;
        }
    }
    if (d->material_textures) {
        for (int i = 0; i < d->num_materials; i++)
            for (int j = 0; j < 4; j++) {
                const int t = d->material_textures[4 * i + j];
                if (t < -1 || t >= d->num_textures) {
                    err = "material " + std::to_string(i) + ": texture row out of range";
                    return false;
                }
                // colour, metallic-roughness and normal lookups read three channels
                if (t >= 0 && j < 3 && d->textures[t].channels < 3) {
                    err = "material " + std::to_string(i) + ": texture " + std::to_string(t) +
                          " needs at least 3 channels";
                    return false;
                }
            }
    }
    if (d->env.width < 0 || d->env.height < 0 || (d->env.width > 0 && (d->env.height <= 0 || !d->env.pixels))) {
        err = "bad environment map";
        return false;
    }
    S.numTris = d->num_triangles;
    S.numSpheres = d->num_spheres;

    // triangles (Triangle ctor), in parallel
    std::vector<TriPre> tp(d->num_triangles);
    const unsigned nth = std::max(1u, std::min(32u, std::thread::hardware_concurrency()));
    {
        std::vector<std::thread> th;
        for (unsigned w = 0; w < nth; w++)
            th.emplace_back([&, w]() {
                for (int64_t i = w; i < d->num_triangles; i += nth) tp[i] = prepTri(d->triangles + 9 * i);
            });
        for (auto &t : th) t.join();
    }
    // spheres (Sphere ctor, Sphere.hpp:16-17)
    S.sphById.resize(d->num_spheres);
    for (int i = 0; i < d->num_spheres; i++) {
        const float *p = d->spheres + 4 * (size_t) i;
        if (!(p[3] > 0.f)) {
            err = "sphere " + std::to_string(i) + ": radius must be > 0";
            return false;
        }
        S.sphById[i] = make_float4(p[0], p[1], p[2], p[3]);
    }
    // per-object bounds / area and mesh sub-BVHs (maxPrims 4), meshes in parallel
    const int nObj = d->num_objects;
    S.meshes.assign(nObj, RefBVH());
    S.objBounds.assign(nObj, Box());
    S.objArea.assign(nObj, 0.f);
    std::atomic<int> next(0);
    {
        std::vector<std::thread> th;
        for (unsigned w = 0; w < nth; w++)
            th.emplace_back([&]() {
                for (int k = next++; k < nObj; k = next++) {
                    const prr_object_desc &o = d->objects[k];
                    if (o.kind == PRR_OBJ_SPHERE) {
                        const float4 s = S.sphById[o.first];
                        float a[3] = {s.x - s.w, s.y - s.w, s.z - s.w}, b[3] = {s.x + s.w, s.y + s.w, s.z + s.w};
                        S.objBounds[k] = Box(a, b);                 // Sphere::getBounds (Sphere.hpp:47-50)
                        S.objArea[k] = 4 * kPi * s.w * s.w;         // Sphere ctor
                        continue;
                    }
                    float mn[3] = {kInfF, kInfF, kInfF}, mx[3] = {-kInfF, -kInfF, -kInfF};
                    std::vector<Box> bx(o.count);
                    std::vector<float> ar(o.count);
                    float area = 0.f;
                    for (int i = 0; i < o.count; i++) {
                        const TriPre &t = tp[o.first + i];
                        const V3 vs[3] = {t.v0, t.v1, t.v2};
                        for (const V3 &p : vs) {
                            mn[0] = std::min(mn[0], p.x);
                            mn[1] = std::min(mn[1], p.y);
                            mn[2] = std::min(mn[2], p.z);
                            mx[0] = std::max(mx[0], p.x);
                            mx[1] = std::max(mx[1], p.y);
                            mx[2] = std::max(mx[2], p.z);
                        }
                        bx[i] = triBox(t);
                        ar[i] = t.area;
                        area += t.area;  // MeshTriangle ctor (Triangle.hpp:116-124)
                    }
                    S.objBounds[k] = Box(mn, mx);
                    S.objArea[k] = area;
                    S.meshes[k] = buildRefBVH(bx, ar, 4);
                }
            });
        for (auto &t : th) t.join();
    }
    // scene BVH over objects (maxPrims 1, SAH)
    S.top = buildRefBVH(S.objBounds, S.objArea, 1);

    // merged device tree
    S.primN.assign((size_t) d->num_triangles + d->num_spheres, make_float4(0, 0, 0, 0));
    for (int k = 0; k < nObj; k++) {
        const prr_object_desc &o = d->objects[k];
        const float mbits = __builtin_bit_cast(float, o.material);
        if (o.kind == PRR_OBJ_MESH) {
            for (int i = 0; i < o.count; i++) {
                const TriPre &t = tp[o.first + i];
                const int fl = d->tri_flags ? d->tri_flags[o.first + i] : 0;
                S.primN[o.first + i] =
                    make_float4(t.n.x, t.n.y, t.n.z, __builtin_bit_cast(float, o.material | fl << kPrimFlagShift));
            }
        } else {
            S.primN[(size_t) d->num_triangles + o.first] = make_float4(0, 0, 0, mbits);
        }
    }
    S.nodes.clear();
    S.tris.clear();
    S.tris.reserve(3 * (size_t) d->num_triangles);
    S.maxDepth = 0;
    Merger mg{S, d, tp, {}};
    S.root = mg.top(0, 0);
    for (int k = 0; k < 6; k++) S.rootBox[k] = S.top.nodes[0].b[k];
    if (S.maxDepth + 1 > 64) {
        err = "BVH deeper than the 64-entry traversal stack";
        return false;
    }
    S.nodes4.clear();
    S.wideLevelsKept = 0;
    {
        Widener wd{S};
        S.root4 = wd.build(S.root, 0);
        if (wd.maxPush > 64) {  // the traversal stack's spill region holds 64 entries
            err = "occlusion tree needs more than 64 stack entries";
            return false;
        }
    }
    if (S.nodes4.empty()) S.nodes4.emplace_back();
    if (!Quantiser{S}.run(err)) return false;
    buildTwoLevel(S);  // no records (binary steps only) if a box is not the union of its children's
    if (S.tris.empty()) S.tris.push_back(make_float4(0, 0, 0, 0));  // keep buffers non-empty
    if (S.sph.empty()) {
        S.sph.push_back(make_float4(0, 0, 0, 0));
        S.sphId.push_back(-1);
    }
    if (S.nodes.empty()) S.nodes.emplace_back();
    if (S.sphById.empty()) S.sphById.push_back(make_float4(0, 0, 0, 0));

    // per-vertex attributes: n0 n1 n2 | t0 t1 t2 | tangent xyz + w per vertex
    S.tattr.clear();
    if (d->tri_flags) {
        S.tattr.resize((size_t) kTriAttrRows * d->num_triangles);
        for (int64_t i = 0; i < d->num_triangles; i++) {
            const int fl = d->tri_flags[i];
            float a[28] = {0};
            if (fl & PRR_TRI_SMOOTH_NORMALS) memcpy(a, d->tri_normals + 9 * i, 9 * sizeof(float));
            if (fl & PRR_TRI_UVS) memcpy(a + 9, d->tri_uvs + 6 * i, 6 * sizeof(float));
            if (fl & PRR_TRI_TANGENTS) memcpy(a + 16, d->tri_tangents + 12 * i, 12 * sizeof(float));
            for (int r = 0; r < kTriAttrRows; r++)
                S.tattr[(size_t) kTriAttrRows * i + r] = make_float4(a[4 * r], a[4 * r + 1], a[4 * r + 2], a[4 * r + 3]);
        }
    }
    // textures: one byte pool
    S.texInfo.clear();
    S.texels.clear();
    for (int i = 0; i < d->num_textures; i++) {
        const prr_texture_desc &t = d->textures[i];
        const size_t n = (size_t) t.width * t.height * t.channels;
        if (S.texels.size() + n > ((size_t) 1 << 31)) {
            err = "textures exceed 2 GiB";
            return false;
        }
        S.texInfo.push_back(make_int4((int) S.texels.size(), t.width, t.height, t.channels));
        S.texels.insert(S.texels.end(), t.data, t.data + n);
    }
    // materials; params.w = 1 when a colour, metallic-roughness or normal
    // texture is attached (the emissive texture is not read by castRay)
    S.mats.resize(d->num_materials);
    S.matTex.clear();
    bool textured = false;
    for (int i = 0; i < d->num_materials; i++) {
        const prr_material_desc &m = d->materials[i];
        const int4 tx = d->material_textures ? make_int4(d->material_textures[4 * i], d->material_textures[4 * i + 1],
                                                         d->material_textures[4 * i + 2], d->material_textures[4 * i + 3])
                                             : make_int4(-1, -1, -1, -1);
        const bool tex = tx.x >= 0 || tx.y >= 0 || tx.z >= 0;
        textured |= tex;
        S.mats[i].base_type = make_float4(m.base_color[0], m.base_color[1], m.base_color[2], (float) m.type);
        S.mats[i].params = make_float4(m.roughness, m.metallic, m.ior, tex ? 1.f : 0.f);
        S.mats[i].emission = make_float4(m.emission[0], m.emission[1], m.emission[2], 0.f);
        S.matTex.push_back(tx);
    }
    if (!textured) S.matTex.clear();
    // emitters (Scene::buildBVH, Scene.cpp:19-25) and light trees
    S.totalEmitArea = 0.f;
    S.emitters.clear();
    S.lnodes.clear();
    S.ltri.clear();
    for (int k = 0; k < nObj; k++) {
        const prr_object_desc &o = d->objects[k];
        if (!hasEmission(d->materials[o.material])) continue;
        DEmitter e{};
        e.area = S.objArea[k];
        e.kind = o.kind == PRR_OBJ_SPHERE ? 1 : 0;
        e.material = o.material;
        if (o.kind == PRR_OBJ_SPHERE) {
            e.lt_root = o.first;
        } else {
            const RefBVH &m = S.meshes[k];
            const int nodeBase = (int) S.lnodes.size();
            const int slotBase = (int) (S.ltri.size() / 4);
            e.lt_root = nodeBase;
            e.total = m.totalArea;
            for (size_t j = 0; j < m.nodes.size(); j++) {
                const RefNode &n = m.nodes[j];
                DLightNode ln{};
                if (n.leaf()) {
                    ln.first = slotBase + (int) n.first;
                    ln.count = (int) n.count();
                } else {
                    ln.first = nodeBase + (int) n.first;
                    ln.count = 0;
                    ln.left_area = m.nodeAreas[j + 1];
                }
                S.lnodes.push_back(ln);
            }
            for (int li : m.ordered) {
                const TriPre &t = tp[o.first + li];
                S.ltri.push_back(make_float4(t.v0.x, t.v0.y, t.v0.z, t.area));
                S.ltri.push_back(make_float4(t.v1.x, t.v1.y, t.v1.z, 0.f));
                S.ltri.push_back(make_float4(t.v2.x, t.v2.y, t.v2.z, 0.f));
                // Triangle::Sample's normal: normalize(n0 + n1 + n2) for a triangle
                // with per-vertex normals (GLTFLoader.cpp:250), else the face normal
                V3 n = t.n;
                const int64_t g = o.first + li;
                if (d->tri_flags && (d->tri_flags[g] & PRR_TRI_SMOOTH_NORMALS)) {
                    const float *q = d->tri_normals + 9 * g;
                    n = normalize(V3(q[0], q[1], q[2]) + V3(q[3], q[4], q[5]) + V3(q[6], q[7], q[8]));
                }
                S.ltri.push_back(make_float4(n.x, n.y, n.z, 0.f));
            }
        }
        S.emitters.push_back(e);
        S.totalEmitArea += S.objArea[k];
    }
    if (S.emitters.empty()) S.emitters.push_back(DEmitter{});
    if (S.lnodes.empty()) S.lnodes.push_back(DLightNode{});
    if (S.ltri.empty()) S.ltri.assign(4, make_float4(0, 0, 0, 0));

    // environment (EnvMap.hpp:189-224)
    S.envW = d->env.width;
    S.envH = d->env.height;
    if (S.envW > 0) {
        const int W = S.envW, H = S.envH;
        S.envPx.resize((size_t) W * H);
        S.envCdf.assign((size_t) W * H, 0.f);
        S.envMarg.assign(H, 0.f);
        std::vector<float> rowSums(H, 0.f);
        for (size_t i = 0; i < (size_t) W * H; i++) {
            const float *p = d->env.pixels + 3 * i;
            const float lum = 0.2126f * p[0] + 0.7152f * p[1] + 0.0722f * p[2];
            S.envPx[i] = make_float4(p[0], p[1], p[2], lum);
        }
        for (int row = 0; row < H; row++) {
            const float theta = (row + 0.5f) / (float) H * kPi;
            const float sinT = std::sin(theta);
            float rs = 0.f;
            for (int col = 0; col < W; col++) {
                rs += S.envPx[(size_t) row * W + col].w * sinT;
                S.envCdf[(size_t) row * W + col] = rs;
            }
            if (rs > 0.f)
                for (int col = 0; col < W; col++) S.envCdf[(size_t) row * W + col] /= rs;
            rowSums[row] = rs;
        }
        float t = 0.f;
        for (int row = 0; row < H; row++) {
            t += rowSums[row];
            S.envMarg[row] = t;
        }
        S.envTotal = t;
        if (t > 0.f)
            for (float &v : S.envMarg) v /= t;
        buildEnvGuides(S);
    } else {
        S.envPx.assign(1, make_float4(0, 0, 0, 0));
        S.envCdf.assign(1, 0.f);
        S.envMarg.assign(1, 0.f);
        S.envTotal = 0.f;
    }
    if (S.envGuideK == 0) {
        S.envGuideM.assign(1, 0);
        S.envGuideC.assign(1, 0);
    }

    S.cam = makeCamView(d->camera);
    for (int k = 0; k < 3; k++) S.background[k] = d->background[k];
    S.numPoints = d->num_point_lights;
    if (S.numPoints > 0)
        for (int k = 0; k < 3; k++) {
            S.pointPos[k] = d->point_lights[0].position[k];
            S.pointInt[k] = d->point_lights[0].intensity[k];
        }
    return true;
}

// Guide tables for the two lower_bound searches of EnvMap::importanceSample
// (EnvMap.hpp:145-169). For a non-decreasing CDF a and K a power of two, a
// draw x in [0, 1) lies in bucket k = floor(x K) (exact: x K only scales the
// exponent), and lower_bound(a, x) lies in [g[k], g[k + 1]] with g[k] =
// lower_bound(a, k / K): the device searches that range only and finds the
// same index. A CDF that is not non-decreasing (NaN, negative luminance)
// leaves the guides off.
void buildEnvGuides(HostScene &S) {
    const int K = 256, W = S.envW, H = S.envH;
    auto monotone = [](const float *a, int n) {
        for (int i = 0; i + 1 < n; i++)
            if (!(a[i] <= a[i + 1])) return false;
        return n == 0 || a[0] == a[0];
    };
    if (!monotone(S.envMarg.data(), H)) return;
    for (int r = 0; r < H; r++)
        if (!monotone(S.envCdf.data() + (size_t) r * W, W)) return;
    auto guides = [&](const float *a, int n, int32_t *g) {
        for (int k = 0; k <= K; k++)
            g[k] = (int32_t) (std::lower_bound(a, a + n, (float) k / (float) K) - a);
    };
    S.envGuideM.assign(K + 1, 0);
    S.envGuideC.assign((size_t) H * (K + 1), 0);
    guides(S.envMarg.data(), H, S.envGuideM.data());
    for (int r = 0; r < H; r++) guides(S.envCdf.data() + (size_t) r * W, W, S.envGuideC.data() + (size_t) r * (K + 1));
    S.envGuideK = K;
}

CamView makeCamView(const prr_camera_desc &c) {
    V3 eye(c.eye[0], c.eye[1], c.eye[2]), tgt(c.target[0], c.target[1], c.target[2]), up(c.up[0], c.up[1], c.up[2]);
    V3 fwd = normalize(tgt - eye);
    V3 right = normalize(cross(fwd, up));
    V3 cu = cross(right, fwd);
    const float halfFovH = c.fovh * kPi / 180.0f * 0.5f;
    float focus = c.focus_distance;
    if (focus < 0.0f) {
        V3 dd = tgt - eye;
        focus = std::sqrt(dd.x * dd.x + dd.y * dd.y + dd.z * dd.z);
    }
    CamView v;
    const V3 *vs[4] = {&eye, &fwd, &right, &cu};
    float *dst[4] = {v.eye, v.fwd, v.right, v.up};
    for (int i = 0; i < 4; i++) {
        dst[i][0] = vs[i]->x;
        dst[i][1] = vs[i]->y;
        dst[i][2] = vs[i]->z;
    }
    v.aperture = c.aperture;
    v.focus = focus;
    v.tanhalf = std::tan(halfFovH);
    return v;
}

}  // namespace host
}  // namespace prr
