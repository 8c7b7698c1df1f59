// oracle_scene.hpp — TEST INFRASTRUCTURE ONLY (see oracle_core.hpp).
//
// Restates the reference's geometry (Triangle.hpp, Sphere.hpp), two-level
// binned-SAH BVH (BVH.cpp), Scene (Scene.cpp) and Integrator::castRay
// (Integrator.cpp) with the same structure: a scene BVH over objects whose
// leaves call into per-mesh sub-BVHs.
#pragma once

#include "oracle_core.hpp"

#include <array>
#include <cstdio>
#include <functional>

namespace orc {

// Triangle ctor (Triangle.hpp:28-34)
struct Tri {
    V3 v0, v1, v2, e1, e2, n;
    float area;
    // per-vertex attributes (Triangle.hpp:16-22; GLTFLoader.cpp:244-268)
    V3 n0, n1, n2, tan0, tan1, tan2;
    float t0[2] = {0, 0}, t1[2] = {0, 0}, t2[2] = {0, 0};
    float tw0 = 1.f, tw1 = 1.f, tw2 = 1.f;
    bool smooth = false, hasTan = false;
    Tri() : area(0) {}
    Tri(const V3 &a, const V3 &b, const V3 &c) : v0(a), v1(b), v2(c) {
        e1 = v1 - v0;
        e2 = v2 - v0;
        n = normalize(cross(e1, e2));
        area = norm(cross(e1, e2)) * 0.5f;
    }
    Box bounds() const { return unite(Box(v0, v1), v2); }  // Triangle.hpp:182-184
};

// Sphere ctor (Sphere.hpp:16-17)
struct Sph {
    V3 c;
    float r, r2, area;
    Box bounds() const {  // Sphere.hpp:47-50
        return Box(V3(c.x - r, c.y - r, c.z - r), V3(c.x + r, c.y + r, c.z + r));
    }
};

struct Counters {
    uint64_t inner = 0;    // interior node expansions (two child boxes read)
    uint64_t tris = 0;     // triangle tests
    uint64_t spheres = 0;  // sphere tests
    uint64_t closest = 0;  // scene-level closest-hit queries
    uint64_t any = 0;      // scene-level any-hit queries
    uint64_t innerAny = 0; // interior expansions of any-hit queries
    uint64_t trisAny = 0;  // triangle tests of any-hit queries
    void add(const Counters &o) {
        inner += o.inner;
        tris += o.tris;
        spheres += o.spheres;
        closest += o.closest;
        any += o.any;
    }
};

struct Hit {
    bool happened = false;
    V3 coords, normal;
    float t = kInf;
    int prim = -1;  // global primitive id
    int obj = -1;   // object index
    float u = 0.f, v = 0.f;
    float tc[2] = {0.f, 0.f};  // Intersection::tcoords
    V3 tangent;                // Intersection::tangent (hasTangent)
    float hand = 1.f;
    bool hasTan = false;
};

// Triangle::finalize's attribute part (Triangle.cpp:42-56)
inline void finalizeTri(const Tri &tr, float u, float v, Hit &h) {
    const float w0 = 1.f - u - v;
    h.tc[0] = tr.t0[0] * w0 + tr.t1[0] * u + tr.t2[0] * v;
    h.tc[1] = tr.t0[1] * w0 + tr.t1[1] * u + tr.t2[1] * v;
    h.normal = tr.smooth ? normalize(tr.n0 * w0 + tr.n1 * u + tr.n2 * v) : tr.n;
    if (tr.hasTan) {
        h.tangent = normalize(tr.tan0 * w0 + tr.tan1 * u + tr.tan2 * v);
        const float w = tr.tw0 * w0 + tr.tw1 * u + tr.tw2 * v;
        h.hand = w >= 0.f ? 1.f : -1.f;
        h.hasTan = true;
    }
}

// Triangle::hitTest (Triangle.hpp:46-64)
inline bool triHit(const Tri &tri, const Ray &r, float &t, float &u, float &v) {
    V3 P = cross(r.d, tri.e2);
    float det = dot(P, tri.e1);
    if (std::fabs(det) < 1e-6f) return false;
    float inv = 1.f / det;
    V3 T = r.o - tri.v0;
    u = dot(P, T) * inv;
    if (u < 0.f) return false;
    V3 Q = cross(T, tri.e1);
    v = dot(Q, r.d) * inv;
    if (v < 0.f || u + v > 1.f) return false;
    t = dot(Q, tri.e2) * inv;
    if (t <= 1e-3f) return false;
    return true;
}
// Triangle::intersectT (Triangle.hpp:75-92)
inline float triAnyT(const Tri &tri, const Ray &r, float tMax) {
    V3 P = cross(r.d, tri.e2);
    float det = dot(P, tri.e1);
    if (std::fabs(det) < 1e-6f) return -1.f;
    float inv = 1.f / det;
    V3 T = r.o - tri.v0;
    float u = dot(P, T) * inv;
    if (u < 0.f || u > 1.f) return -1.f;
    V3 Q = cross(T, tri.e1);
    float v = dot(Q, r.d) * inv;
    if (v < 0.f || u + v > 1.f) return -1.f;
    float t = dot(Q, tri.e2) * inv;
    return (t > 1e-3f && t < tMax) ? t : -1.f;
}
// solveQuadratic (global.hpp:19-33); the unqualified sqrt there resolves to
// the double overload, so q is formed in double.
inline bool solveQuadratic(float a, float b, float c, float &x0, float &x1) {
    float discr = b * b - 4 * a * c;
    if (discr < 0) return false;
    if (discr == 0) {
        x0 = x1 = (float) (-0.5 * (double) b / (double) a);
    } else {
        double s = std::sqrt((double) discr);
        float q = (float) ((b > 0) ? -0.5 * ((double) b + s) : -0.5 * ((double) b - s));
        x0 = q / a;
        x1 = c / q;
    }
    if (x0 > x1) std::swap(x0, x1);
    return true;
}
// Sphere::getIntersection (Sphere.hpp:22-44): nearest t >= 0, no epsilon.
inline bool sphHit(const Sph &s, const Ray &r, float &t) {
    V3 L = r.o - s.c;
    float a = dot(r.d, r.d);
    float b = 2 * dot(r.d, L);
    float c = dot(L, L) - s.r2;
    float t0, t1;
    if (!solveQuadratic(a, b, c, t0, t1)) return false;
    if (t0 < 0) t0 = t1;
    if (t0 < 0) return false;
    t = t0;
    return true;
}

// ---------------------------------------------------------------------------
// BVHAccel (BVH.hpp:34-145, BVH.cpp:13-514), generic over primitive handles.
// ---------------------------------------------------------------------------
struct BNode {
    float b[6];  // [minX,maxX,minY,maxY,minZ,maxZ]
    uint32_t first;
    uint32_t meta;  // [31:24] split axis, [23:0] prim count (BVH.hpp:37)
    uint32_t count() const { return meta & 0x00FFFFFFu; }
    bool leaf() const { return count() > 0; }
    int axis() const { return (int) (meta >> 24); }
};

struct RayData {  // BVH.cpp:290-299
    float org[3], inv[3];
    int oct[3];
};
inline RayData precompute(const Ray &r) {
    RayData rd;
    for (int a = 0; a < 3; a++) {
        float d = r.d[a];
        rd.org[a] = r.o[a];
        rd.inv[a] = (d != 0.f) ? (1.f / d) : kInf;
        rd.oct[a] = (d < 0.f) ? 1 : 0;
    }
    return rd;
}
// nodeIntersect (BVH.hpp:115-124)
inline float nodeHit(const BNode &n, const RayData &rd) {
    float tMin = 0.f, tMax = kInf;
    for (int a = 0; a < 3; a++) {
        float t0 = (n.b[a * 2 + rd.oct[a]] - rd.org[a]) * rd.inv[a];
        float t1 = (n.b[a * 2 + 1 - rd.oct[a]] - rd.org[a]) * rd.inv[a];
        tMin = std::max(tMin, t0);
        tMax = std::min(tMax, t1);
    }
    return tMin <= tMax ? tMin : kInf;
}

struct BVH {
    int maxPrims = 1;
    std::vector<BNode> nodes;
    std::vector<int> ordered;
    std::vector<float> nodeAreas;
    float totalArea = 0.f;

    struct BuildNode {
        Box bounds;
        int child[2] = {-1, -1};
        float area = 0.f;
        int axis = 0, primOffset = 0, primCount = 0;
    };
    std::vector<BuildNode> pool;
    std::function<Box(int)> getBounds;
    std::function<float(int)> getArea;

    // BVHAccel ctor (BVH.cpp:13-54)
    void build(std::vector<int> prims, int maxPrimsInNode, std::function<Box(int)> gb,
               std::function<float(int)> ga) {
        maxPrims = std::min(255, maxPrimsInNode);
        getBounds = gb;
        getArea = ga;
        nodes.clear();
        ordered.clear();
        nodeAreas.clear();
        pool.clear();
        if (prims.empty()) return;
        pool.reserve(2 * prims.size());
        ordered.reserve(prims.size());
        int root = rec(prims, 0, (int) prims.size());
        nodes.resize(pool.size());
        nodeAreas.resize(pool.size());
        uint32_t off = 0;
        flatten(root, &off);
        totalArea = pool[root].area;
        pool.clear();
        pool.shrink_to_fit();
    }

    // recursiveBuild (BVH.cpp:66-256)
    int rec(std::vector<int> &prims, int start, int end) {
        pool.emplace_back();
        int ni = (int) pool.size() - 1;
        const int count = end - start;
        std::vector<Box> pb(count);
        std::vector<V3> pc(count);
        Box bounds, cb;
        for (int i = 0; i < count; i++) {
            pb[i] = getBounds(prims[start + i]);
            pc[i] = pb[i].centroid();
            bounds = unite(bounds, pb[i]);
            cb = unite(cb, pc[i]);
        }
        pool[ni].bounds = bounds;
        if (count == 1) {
            pool[ni].primOffset = (int) ordered.size();
            pool[ni].primCount = 1;
            pool[ni].area = getArea(prims[start]);
            ordered.push_back(prims[start]);
            return ni;
        }
        if (count <= maxPrims) {
            pool[ni].primOffset = (int) ordered.size();
            pool[ni].primCount = count;
            pool[ni].area = 0.f;
            for (int i = start; i < end; i++) {
                pool[ni].area += getArea(prims[i]);
                ordered.push_back(prims[i]);
            }
            return ni;
        }
        const int axis = cb.maxExtent();
        pool[ni].axis = axis;
        int mid = (start + end) / 2;
        if (cb.mx[axis] == cb.mn[axis]) {
            // degenerate: median split without reordering (count > maxPrims here)
        } else if (count > 4) {
            constexpr int NB = 12;
            struct Bucket {
                int count = 0;
                Box bounds;
            };
            float bestCost = kInf;
            int bestAxis = axis, bestSplit = NB / 2;
            const float parentArea = bounds.area();
            for (int a = 0; a < 3; a++) {
                float lo = cb.mn[a], hi = cb.mx[a];
                if (lo == hi) continue;
                Bucket bk[NB];
                const float span = hi - lo;
                for (int i = 0; i < count; i++) {
                    int b = std::clamp((int) (NB * (pc[i][a] - lo) / span), 0, NB - 1);
                    bk[b].count++;
                    bk[b].bounds = unite(bk[b].bounds, pb[i]);
                }
                float lA[NB - 1], rA[NB - 1];
                int lC[NB - 1], rC[NB - 1];
                {
                    Box bl;
                    int c = 0;
                    for (int i = 0; i < NB - 1; i++) {
                        bl = unite(bl, bk[i].bounds);
                        c += bk[i].count;
                        lA[i] = bl.area();
                        lC[i] = c;
                    }
                }
                {
                    Box br;
                    int c = 0;
                    for (int i = NB - 2; i >= 0; i--) {
                        br = unite(br, bk[i + 1].bounds);
                        c += bk[i + 1].count;
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
            float lo = cb.mn[bestAxis];
            float span = cb.mx[bestAxis] - lo;
            float edge = lo + (bestSplit + 1) * span / NB;
            std::vector<int> li, ri;
            li.reserve(count);
            ri.reserve(count);
            for (int i = 0; i < count; i++) {
                if (pc[i][bestAxis] < edge)
                    li.push_back(i);
                else
                    ri.push_back(i);
            }
            if (li.empty() || ri.empty()) {
                mid = (start + end) / 2;
            } else {
                std::vector<int> tmp(count);
                int w = 0;
                for (int idx : li) tmp[w++] = prims[start + idx];
                for (int idx : ri) tmp[w++] = prims[start + idx];
                for (int i = 0; i < count; i++) prims[start + i] = tmp[i];
                mid = start + (int) li.size();
            }
        } else {
            auto gb = getBounds;
            std::nth_element(prims.data() + start, prims.data() + mid, prims.data() + end,
                             [axis, &gb](int a, int b) {
                                 return gb(a).centroid()[axis] < gb(b).centroid()[axis];
                             });
        }
        int c0 = rec(prims, start, mid);
        int c1 = rec(prims, mid, end);
        pool[ni].child[0] = c0;
        pool[ni].child[1] = c1;
        pool[ni].area = pool[c0].area + pool[c1].area;
        return ni;
    }

    // flatten (BVH.cpp:263-280): left child at parent+1, right at first_child
    uint32_t flatten(int bi, uint32_t *off) {
        BNode &n = nodes[*off];
        const BuildNode &b = pool[bi];
        n.b[0] = b.bounds.mn.x;
        n.b[1] = b.bounds.mx.x;
        n.b[2] = b.bounds.mn.y;
        n.b[3] = b.bounds.mx.y;
        n.b[4] = b.bounds.mn.z;
        n.b[5] = b.bounds.mx.z;
        nodeAreas[*off] = b.area;
        uint32_t my = (*off)++;
        if (b.primCount > 0) {
            nodes[my].first = (uint32_t) b.primOffset;
            nodes[my].meta = (uint32_t) b.primCount & 0x00FFFFFFu;
        } else {
            nodes[my].meta = (uint32_t) b.axis << 24;
            flatten(b.child[0], off);
            nodes[my].first = flatten(b.child[1], off);
        }
        return my;
    }

    // Intersect (BVH.cpp:311-415). leafFn(prim, closestT) returns a new
    // closest t (< closestT) on a hit and records the winner, or closestT.
    template <class LeafFn>
    void intersect(const Ray &r, LeafFn leafFn, Counters *cnt) const {
        if (nodes.empty()) return;
        const RayData rd = precompute(r);
        struct SE {
            uint32_t idx;
            float tBox;
        } stack[64];
        int sp = 0;
        uint32_t ni = 0;
        float closestT = kInf;
        while (true) {
            const BNode &n = nodes[ni];
            if (n.leaf()) {
                for (uint32_t i = 0; i < n.count(); i++) closestT = leafFn(ordered[n.first + i], closestT);
            } else {
                if (cnt) cnt->inner++;
                uint32_t left = ni + 1, right = n.first;
                float tL = nodeHit(nodes[left], rd);
                float tR = nodeHit(nodes[right], rd);
                if (tL <= tR) {
                    if (tL < closestT) {
                        if (tR < closestT) stack[sp++] = {right, tR};
                        ni = left;
                        continue;
                    }
                } else {
                    if (tR < closestT) {
                        if (tL < closestT) stack[sp++] = {left, tL};
                        ni = right;
                        continue;
                    }
                }
            }
            bool found = false;
            while (sp > 0) {
                ni = stack[--sp].idx;
                if (stack[sp].tBox < closestT) {
                    found = true;
                    break;
                }
            }
            if (!found) break;
        }
    }

    // IntersectP (BVH.cpp:421-473). anyFn(prim) returns true if occluded.
    template <class AnyFn>
    bool intersectP(const Ray &r, float tMax, AnyFn anyFn, Counters *cnt) const {
        if (nodes.empty()) return false;
        const RayData rd = precompute(r);
        uint32_t stack[64];
        int sp = 0;
        uint32_t ni = 0;
        while (true) {
            const BNode &n = nodes[ni];
            float tBox = nodeHit(n, rd);
            if (tBox < tMax) {
                if (n.leaf()) {
                    for (uint32_t i = 0; i < n.count(); i++)
                        if (anyFn(ordered[n.first + i])) return true;
                } else {
                    if (cnt) cnt->innerAny++;
                    uint32_t left = ni + 1, right = n.first, nearC, farC;
                    if (rd.oct[n.axis()] == 0) {
                        nearC = left;
                        farC = right;
                    } else {
                        nearC = right;
                        farC = left;
                    }
                    if (nodeHit(nodes[farC], rd) < tMax) stack[sp++] = farC;
                    ni = nearC;
                    continue;
                }
            }
            if (sp == 0) break;
            ni = stack[--sp];
        }
        return false;
    }

    // getSample walk (BVH.cpp:478-508); returns the chosen primitive.
    int samplePrim(float p) const {
        uint32_t ni = 0;
        while (true) {
            const BNode &n = nodes[ni];
            if (n.leaf()) {
                for (uint32_t i = 0; i < n.count(); i++) {
                    int prim = ordered[n.first + i];
                    p -= getArea(prim);
                    if (p <= 0.f || i == n.count() - 1) return prim;
                }
                return -1;
            }
            const float leftArea = nodeAreas[ni + 1];
            if (p < leftArea)
                ni = ni + 1;
            else {
                p -= leftArea;
                ni = n.first;
            }
        }
    }
};

// ---------------------------------------------------------------------------
// Scene (Scene.hpp / Scene.cpp)
// ---------------------------------------------------------------------------
enum ObjKind { OBJ_MESH = 0, OBJ_SPHERE = 1 };

struct Obj {
    int kind = OBJ_MESH, mat = 0, first = 0, count = 0;
    Box bounds;
    float area = 0.f;
    BVH sub;  // meshes only (maxPrims 4, Triangle.hpp:173-179)
};

struct PointLight {
    V3 pos, intensity;
};

// Material::buildShadingData (Material.cpp:3-26) with the glTF overload's
// tangent frame (Material.hpp:61-74) and TextureUtils lookups.
struct Tex {
    int w = 0, h = 0, c = 0;
    std::vector<uint8_t> d;
};
inline void uvToTexel(float u, float v, int w, int h, int &x0, int &y0, int &x1, int &y1, float &fx, float &fy) {
    u = u - std::floor(u);  // TextureUtils.hpp:24-39
    v = v - std::floor(v);
    float px = u * w, py = v * h;
    x0 = (int) px % w;
    y0 = (int) py % h;
    x1 = (x0 + 1) % w;
    y1 = (y0 + 1) % h;
    fx = px - std::floor(px);
    fy = py - std::floor(py);
}
inline V3 texLookup(const Tex &t, const float uv[2], bool snorm) {
    int x0, y0, x1, y1;
    float fx, fy;
    uvToTexel(uv[0], uv[1], t.w, t.h, x0, y0, x1, y1, fx, fy);
    auto at = [&](int x, int y) {
        const uint8_t *q = t.d.data() + ((size_t) y * t.w + x) * t.c;
        if (snorm) return V3(q[0] / 255.f * 2.f - 1.f, q[1] / 255.f * 2.f - 1.f, q[2] / 255.f * 2.f - 1.f);
        return V3(q[0] / 255.f, q[1] / 255.f, q[2] / 255.f);
    };
    V3 c00 = at(x0, y0), c10 = at(x1, y0), c01 = at(x0, y1), c11 = at(x1, y1);
    V3 top = c00 + (c10 - c00) * fx;  // TextureUtils.hpp:61-66
    V3 bottom = c01 + (c11 - c01) * fx;
    return top + (bottom - top) * fy;
}
inline float srgbToLinear(float x) {  // TextureUtils.hpp:51-54
    x = std::max(0.f, x);
    return x <= 0.04045f ? x / 12.92f : std::pow((x + 0.055f) / 1.055f, 2.4f);
}

struct Scene {
    std::vector<Tri> tris;
    std::vector<Tex> texs;
    std::vector<std::array<int, 4>> matTex;  // per material (empty: untextured)

    // ShadingData for `mat` at hit h, with Ng the normal the caller shades with
    Shading shading(int mat, const Hit &h, const V3 &Ng) const {
        Shading sd = makeShading(mats[mat], Ng);
        if (matTex.empty()) return sd;
        const std::array<int, 4> &tx = matTex[mat];
        if (tx[0] >= 0) {
            V3 c = texLookup(texs[tx[0]], h.tc, false);
            sd.base = V3(srgbToLinear(c.x), srgbToLinear(c.y), srgbToLinear(c.z));
        }
        if (tx[1] >= 0) {
            V3 c = texLookup(texs[tx[1]], h.tc, false);
            sd.roughness = std::max(c.y, 0.08f);
            sd.metallic = c.z;
        }
        if (tx[2] >= 0) {
            V3 T, B;
            bool ok = false;
            if (h.hasTan) {
                V3 tg = normalize(h.tangent);  // castRay normalises again (Integrator.cpp:145, 179)
                T = tg - dot(Ng, tg) * Ng;
                float tlen = std::sqrt(T.x * T.x + T.y * T.y + T.z * T.z);
                if (tlen >= 1e-6f) {
                    T = T / tlen;
                    B = cross(Ng, T) * h.hand;
                    ok = true;
                }
            }
            if (!ok) buildTBN(Ng, T, B);
            V3 tn = normalize(texLookup(texs[tx[2]], h.tc, true));
            sd.N = normalize(tn.x * T + tn.y * B + tn.z * Ng);
        }
        return sd;
    }
    std::vector<Sph> sphs;
    std::vector<Material> mats;
    std::vector<Obj> objs;
    BVH top;  // maxPrims 1, SAH (Scene.cpp:17)
    std::vector<int> emissives;
    float totalEmitArea = 0.f;
    Env env;
    V3 background = V3(0.235294f, 0.67451f, 0.843137f);  // Scene.hpp:17
    Camera camera;
    std::vector<PointLight> points;

    int numTris() const { return (int) tris.size(); }

    // MeshTriangle ctor + Scene::buildBVH (Triangle.hpp:116-124, Scene.cpp:15-27)
    void finalizeBuild() {
        for (int k = 0; k < (int) objs.size(); k++) {
            Obj &o = objs[k];
            if (o.kind == OBJ_MESH) {
                V3 mn(kInf), mx(-kInf);
                o.area = 0;
                std::vector<int> ptrs;
                ptrs.reserve(o.count);
                for (int i = 0; i < o.count; i++) {
                    const Tri &t = tris[o.first + i];
                    const V3 vs[3] = {t.v0, t.v1, t.v2};
                    for (const V3 &p : vs) {
                        mn = V3(std::min(mn.x, p.x), std::min(mn.y, p.y), std::min(mn.z, p.z));
                        mx = V3(std::max(mx.x, p.x), std::max(mx.y, p.y), std::max(mx.z, p.z));
                    }
                    o.area += t.area;
                    ptrs.push_back(o.first + i);
                }
                o.bounds = Box(mn, mx);
                const std::vector<Tri> *tv = &tris;
                o.sub.build(
                    ptrs, 4, [tv](int i) { return (*tv)[i].bounds(); },
                    [tv](int i) { return (*tv)[i].area; });
            } else {
                const Sph &s = sphs[o.first];
                o.bounds = s.bounds();
                o.area = s.area;
            }
        }
        std::vector<int> ids(objs.size());
        for (size_t i = 0; i < objs.size(); i++) ids[i] = (int) i;
        const std::vector<Obj> *ov = &objs;
        top.build(
            ids, 1, [ov](int i) { return (*ov)[i].bounds; }, [ov](int i) { return (*ov)[i].area; });
        totalEmitArea = 0.f;
        emissives.clear();
        for (int k = 0; k < (int) objs.size(); k++) {
            if (mats[objs[k].mat].hasEmission()) {
                emissives.push_back(k);
                totalEmitArea += objs[k].area;
            }
        }
    }

    // Object::getIntersection for one object (Triangle.hpp:130-135 + BVH fast
    // path + Triangle.cpp:34-59; Sphere.hpp:22-44)
    Hit objectHit(int k, const Ray &r, Counters *cnt) const {
        const Obj &o = objs[k];
        Hit h;
        if (o.kind == OBJ_MESH) {
            int best = -1;
            float bt = 0, bu = 0, bv = 0;
            o.sub.intersect(
                r,
                [&](int ti, float closestT) {
                    if (cnt) cnt->tris++;
                    float t, u, v;
                    if (triHit(tris[ti], r, t, u, v) && t < closestT) {
                        best = ti;
                        bt = t;
                        bu = u;
                        bv = v;
                        return t;
                    }
                    return closestT;
                },
                cnt);
            if (best >= 0) {
                const float t = bt;
                h.happened = true;
                h.t = t;
                h.coords = r.o + t * r.d;
                h.prim = best;
                h.obj = k;
                h.u = bu;
                h.v = bv;
                finalizeTri(tris[best], bu, bv, h);
            }
        } else {
            if (cnt) cnt->spheres++;
            const Sph &s = sphs[o.first];
            float t;
            if (sphHit(s, r, t)) {
                h.happened = true;
                h.t = t;
                h.coords = V3(r.o + r.d * t);
                h.normal = normalize(V3(h.coords - s.c));
                h.prim = numTris() + o.first;
                h.obj = k;
            }
        }
        return h;
    }

    // Scene::intersect (Scene.cpp:29-31) -> scene BVH general path
    Hit intersect(const Ray &r, Counters *cnt) const {
        if (cnt) cnt->closest++;
        Hit result;
        top.intersect(
            r,
            [&](int k, float closestT) {
                Hit tmp = objectHit(k, r, cnt);
                if (tmp.happened && tmp.t < closestT) {
                    result = tmp;
                    return tmp.t;
                }
                return closestT;
            },
            cnt);
        return result;
    }

    // Scene::intersectP (Scene.cpp:33-35)
    bool intersectP(const Ray &r, float tMax, Counters *cnt) const {
        if (cnt) cnt->any++;
        return top.intersectP(
            r, tMax,
            [&](int k) {
                const Obj &o = objs[k];
                float t;
                if (o.kind == OBJ_MESH) {
                    bool occ = o.sub.intersectP(
                        r, tMax,
                        [&](int ti) {
                            if (cnt) cnt->trisAny++;
                            float tt = triAnyT(tris[ti], r, tMax);
                            return tt >= 0.f && tt < tMax;
                        },
                        cnt);
                    t = occ ? 1e-3f : -1.f;  // MeshTriangle::intersectT (Triangle.hpp:141-146)
                } else {
                    if (cnt) cnt->spheres++;
                    float ts;  // Object::intersectT default (Object.hpp:22-27)
                    t = (sphHit(sphs[o.first], r, ts) && ts < tMax) ? ts : -1.f;
                }
                return t >= 0.f && t < tMax;
            },
            cnt);
    }

    // Scene::sampleLight (Scene.cpp:37-58) incl. MeshTriangle::Sample ->
    // BVHAccel::Sample (BVH.cpp:510-514) -> Triangle::Sample (Triangle.hpp:98-103)
    // and Sphere::Sample (Sphere.hpp:54-63). Returns false if no emitter.
    template <class R>
    bool sampleLight(R &rng, V3 &pos, V3 &nrm, int &mat, float &pdf) const {
        if (totalEmitArea <= 0.f) return false;
        float p = rng.next() * totalEmitArea;
        float acc = 0;
        int chosen = emissives.empty() ? -1 : emissives.back();
        for (int k : emissives) {
            acc += objs[k].area;
            if (p <= acc) {
                chosen = k;
                break;
            }
        }
        if (chosen < 0) return false;
        const Obj &o = objs[chosen];
        if (o.kind == OBJ_MESH) {
            float q = std::sqrt(rng.next()) * o.sub.totalArea;
            int ti = o.sub.samplePrim(q);
            const Tri &t = tris[ti];
            float x = std::sqrt(rng.next()), y = rng.next();
            pos = t.v0 * (1.0f - x) + t.v1 * (x * (1.0f - y)) + t.v2 * (x * y);
            nrm = t.n;
        } else {
            const Sph &s = sphs[o.first];
            float theta = (float) (2.0 * (double) kPi * (double) rng.next());
            float phi = kPi * rng.next();
            V3 dir(std::cos(phi), std::sin(phi) * std::cos(theta), std::sin(phi) * std::sin(theta));
            pos = s.c + s.r * dir;
            nrm = dir;
        }
        mat = o.mat;
        pdf = 1.0f / totalEmitArea;
        return true;
    }
};

// ---------------------------------------------------------------------------
// Integrator::castRay (Integrator.cpp:104-263)
//
// Build extensions, both inactive for scenes without them (the reference has
// neither): NEE to point lights after the environment NEE (no draws), and a
// caustic-photon radiance estimate at diffuse vertices reached from the camera
// directly or through delta bounces only (photonFn != nullptr).
// ---------------------------------------------------------------------------
struct IntegratorParams {
    int maxDepth = 8;
    float indirectClamp = 10.0f;
    // > 0: the path ends at its first scattering (diffuse) vertex at bounce >=
    // truncateAt, before that vertex's NEE (a radiance-cache query path whose
    // cache contributes nothing)
    int truncateAt = 0;
};

using PhotonFn = std::function<V3(const V3 &pos, const V3 &wo, const Shading &sd)>;

template <class R>
V3 castRay(const Scene &sc, const Ray &ray0, R &rng, const IntegratorParams &ip, Counters *cnt,
           const PhotonFn *photonFn = nullptr) {
    auto clampIndirect = [&](const V3 &c) -> V3 {
        if (ip.indirectClamp <= 0.f) return c;
        float lum = (c.x + c.y + c.z) / 3;
        return (lum > ip.indirectClamp) ? c * (ip.indirectClamp / lum) : c;
    };
    V3 L(0.f), beta(1.f);
    Ray cur = ray0;
    bool specular = false;
    Hit inter = sc.intersect(cur, cnt);
    for (int bounce = 0; bounce < ip.maxDepth; bounce++) {
        if (bounce == 0 || specular) {
            if (!inter.happened) {
                V3 bg = !sc.env.empty() ? sc.env.sample(cur.d) : sc.background;
                V3 c = beta * bg;
                L += (bounce > 0) ? clampIndirect(c) : c;
            } else if (sc.mats[sc.objs[inter.obj].mat].isEmissive()) {
                V3 c = beta * sc.mats[sc.objs[inter.obj].mat].emission;
                L += (bounce > 0) ? clampIndirect(c) : c;
            }
        }
        if (!inter.happened) break;
        const Material &mat = sc.mats[sc.objs[inter.obj].mat];
        V3 hitPoint = inter.coords;
        V3 geoN = inter.normal;
        if (mat.isDelta()) {
            const Shading dsd = sc.shading(sc.objs[inter.obj].mat, inter, geoN);
            BSDFSample bs = (mat.type == MAT_GLASS) ? glassSample(-cur.d, dsd.N, mat.ior, rng)
                                                     : mirrorSample(-cur.d, dsd.N, mat.base);
            if (bs.pdf < 1e-6f) break;
            beta = beta * bs.f / bs.pdf;
            bool trans = dot(bs.wi, geoN) < 0.f;
            V3 off = trans ? -geoN : geoN;
            cur = Ray{hitPoint + off * kEps, bs.wi};
            inter = sc.intersect(cur, cnt);
            if (trans) beta = beta * dsd.base;  // Integrator.cpp:161-165 (sd.baseColor)
            specular = true;
            continue;
        }
        {
            bool camOrSpec = (bounce == 0) || specular;
            specular = false;
            V3 sN = dot(cur.d, geoN) < 0.f ? geoN : -geoN;
            Shading sd = sc.shading(sc.objs[inter.obj].mat, inter, sN);
            V3 wo = -cur.d;
            bool indirect = bounce > 0;
            const bool scatters = (mat.type == MAT_DIFFUSE);  // EmissiveMaterial: f = pdf = 0
            if (scatters && ip.truncateAt > 0 && bounce >= ip.truncateAt) break;
            // NEE area (Integrator.cpp:14-34, 78-87)
            if (sc.totalEmitArea > 0.f) {
                V3 lp, ln;
                int lm;
                float pdfA;
                V3 contrib(0.f);
                if (sc.sampleLight(rng, lp, ln, lm, pdfA)) {
                    V3 toL = lp - hitPoint;
                    float d2 = dot(toL, toL);
                    float dist = std::sqrt(d2);
                    V3 wi = toL / dist;
                    float cosL = std::max(0.f, dot(-wi, normalized(ln)));
                    if (cosL > 0.f) {
                        float pdfSA = pdfA * d2 / cosL;
                        Ray sr{hitPoint + wi * kEps, wi};
                        bool vis = !sc.intersectP(sr, dist - 2.f * kEps, cnt);
                        if (vis) {
                            V3 f = scatters ? diffuseEval(wi, wo, sd) : V3(0.f);
                            float bp = scatters ? diffusePdf(wi, wo, sd) : 0.f;
                            float w = (pdfSA * pdfSA) / (pdfSA * pdfSA + bp * bp + 1e-6f);
                            contrib = beta * (w * sc.mats[lm].emission * f / pdfSA);
                        }
                    }
                }
                L += indirect ? clampIndirect(contrib) : contrib;
            }
            // NEE env (Integrator.cpp:39-66, 93-102)
            if (!sc.env.empty()) {
                float r1 = rng.next(), r2 = rng.next();
                float pdfE;
                V3 dir = sc.env.importanceSample(r1, r2, pdfE);
                V3 contrib(0.f);
                if (!(pdfE < 1e-6f)) {
                    Ray sr{hitPoint + dir * kEps, dir};
                    if (!sc.intersectP(sr, kInf, cnt)) {
                        V3 Le = sc.env.sample(dir);
                        V3 f = scatters ? diffuseEval(dir, wo, sd) : V3(0.f);
                        float bp = scatters ? diffusePdf(dir, wo, sd) : 0.f;
                        float w = (pdfE * pdfE) / (pdfE * pdfE + bp * bp + 1e-6f);
                        contrib = beta * (w * Le * f / (pdfE + 1e-6f));
                    }
                }
                L += indirect ? clampIndirect(contrib) : contrib;
            }
            // NEE point lights (build extension; PointLight.hpp is unused upstream)
            for (const PointLight &pl : sc.points) {
                V3 toL = pl.pos - hitPoint;
                float d2 = dot(toL, toL);
                float dist = std::sqrt(d2);
                V3 wi = toL / dist;
                V3 contrib(0.f);
                Ray sr{hitPoint + wi * kEps, wi};
                if (scatters && !sc.intersectP(sr, dist - 2.f * kEps, cnt)) {
                    V3 f = diffuseEval(wi, wo, sd);
                    contrib = beta * (pl.intensity * f / d2);
                }
                L += indirect ? clampIndirect(contrib) : contrib;
            }
            // caustic photon estimate (build extension)
            if (photonFn && camOrSpec && scatters) {
                V3 c = beta * (*photonFn)(hitPoint, wo, sd);
                L += indirect ? clampIndirect(c) : c;
            }
            // BSDF sampling (Integrator.cpp:207-248)
            BSDFSample bs;
            if (scatters) bs = diffuseSample(wo, sd, rng);
            if (bs.pdf < 1e-6f) break;
            V3 betaScale = bs.f / bs.pdf;
            V3 off = dot(bs.wi, sd.N) >= 0.f ? sd.N : -sd.N;
            cur = Ray{hitPoint + off * kEps, bs.wi};
            inter = sc.intersect(cur, cnt);
            if (!inter.happened) {
                if (!sc.env.empty()) {
                    V3 envL = sc.env.sample(bs.wi);
                    float envPdf = sc.env.pdf(bs.wi);
                    float w = (bs.pdf * bs.pdf) / (bs.pdf * bs.pdf + envPdf * envPdf + 1e-6f);
                    V3 c = beta * betaScale * w * envL;
                    L += indirect ? clampIndirect(c) : c;
                } else {
                    V3 c = beta * betaScale * sc.background;
                    L += indirect ? clampIndirect(c) : c;
                }
                break;
            }
            const Material &hm = sc.mats[sc.objs[inter.obj].mat];
            if (hm.isEmissive()) {
                float cosT = std::max(0.f, dot(-bs.wi, normalized(inter.normal)));
                V3 d = inter.coords - hitPoint;
                float hd2 = dot(d, d);
                float lpA = sc.totalEmitArea > 0.f ? 1.f / sc.totalEmitArea : 0.f;
                float lpSA = lpA * hd2 / std::max(cosT, 1e-4f);
                float w = (bs.pdf * bs.pdf) / (bs.pdf * bs.pdf + lpSA * lpSA + 1e-6f);
                V3 c = beta * betaScale * w * hm.emission;
                L += indirect ? clampIndirect(c) : c;
                break;
            }
            beta = beta * betaScale;
            if (!finite3(beta)) break;
        }
        if (bounce >= 3) {
            float mc = std::max({beta.x, beta.y, beta.z});
            float q = std::max(0.05f, 1.f - mc);
            if (rng.next() < q) break;
            beta = beta * (1.f / (1.f - q));
        }
    }
    if (!finite3(L)) return V3(0.f);
    return L;
}

}  // namespace orc
