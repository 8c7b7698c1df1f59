// scene_build.hpp — host side of Scene::buildBVH for the device library.
//
// Builds the reference's two-level BVH with its binned-SAH rules (scene
// level maxPrims 1 over objects, mesh level maxPrims 4 over triangles;
// BVH.cpp:13-280, Scene.cpp:15-27, Triangle.hpp:173-179), keeps each level in
// the reference's flattened layout for inspection, and produces the device
// arrays: one merged tree of two-child 64-byte nodes, leaf-ordered triangle
// records, emitter light trees (BVH.cpp:478-514) and the environment CDF
// (EnvMap.hpp:189-224).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "../../../include/prr_abi.h"
#include "../device/dev_common.hpp"

namespace prr {
namespace host {

struct Box {
    float mn[3], mx[3];
    Box();
    Box(const float *a, const float *b);  // Bounds3(p1, p2): fmin / fmax
    void unite(const Box &o);             // Union(Bounds3, Bounds3): std::min / std::max
    void unite(const float *p);           // Union(Bounds3, Vector3f)
    float area() const;                   // Bounds3::SurfaceArea (float expression)
    void centroid(float *c) const;        // 0.5 * pMin + 0.5 * pMax
    int maxExtent() const;
};

// One level of the reference BVH in its flattened layout (BVH.hpp:34-67).
struct RefNode {
    float b[6];
    uint32_t first;
    uint32_t meta;
    bool leaf() const { return (meta & 0x00FFFFFFu) > 0u; }
    uint32_t count() const { return meta & 0x00FFFFFFu; }
};
struct RefBVH {
    std::vector<RefNode> nodes;
    std::vector<int> ordered;      // local primitive index in leaf order
    std::vector<float> nodeAreas;  // BVH.hpp:135
    float totalArea = 0.f;
};

// Binned-SAH build of n primitives given their boxes and areas.
RefBVH buildRefBVH(const std::vector<Box> &bounds, const std::vector<float> &areas, int maxPrims);

struct HostScene {
    // reference-layout levels
    RefBVH top;                  // over objects
    std::vector<RefBVH> meshes;  // per object (empty for spheres)
    std::vector<Box> objBounds;
    std::vector<float> objArea;
    // device staging
    std::vector<DNode> nodes;
    std::vector<DNode2> nodes2;    // two levels per record, per binary node (empty: a box not the union of its children's)
    std::vector<DNode4> nodes4;    // occlusion tree
    std::vector<DNode4Q> nodes4q;  // the same tree quantised
    std::vector<float4> leavesq;   // its leaf records: exact box, then the leaf's triangles
    std::vector<float4> tris;      // 3 per leaf slot
    std::vector<float4> sph;       // leaf-order spheres
    std::vector<int> sphId;        // leaf-order sphere -> global prim id
    std::vector<float4> primN;     // per global prim id
    std::vector<float4> sphById;   // per sphere row
    std::vector<DMaterial> mats;
    std::vector<DEmitter> emitters;
    std::vector<DLightNode> lnodes;
    std::vector<float4> ltri;      // 4 per light leaf slot
    // per-vertex attributes (empty: none): kTriAttrRows float4 per triangle
    std::vector<float4> tattr;
    std::vector<int4> matTex;      // per material: texture rows (-1 none); empty: untextured scene
    std::vector<int4> texInfo;     // per texture: first texel byte, width, height, channels
    std::vector<uint8_t> texels;
    std::vector<float4> envPx;
    std::vector<float> envCdf, envMarg;
    float envTotal = 0.f;
    int envW = 0, envH = 0;
    // guide tables: entry k of a CDF = its lower_bound of k / envGuideK
    // (marginal: K + 1 entries; conditional: K + 1 per row); K = 0: none
    int envGuideK = 0;
    std::vector<int32_t> envGuideM, envGuideC;
    uint32_t root = 0;
    uint32_t root4 = 0;
    uint32_t root4q = 0;
    int wideLevelsKept = 0;        // binary levels not collapsed (a child box not inside its parent's)
    float rootBox[6] = {0, 0, 0, 0, 0, 0};
    int maxDepth = 0;
    float totalEmitArea = 0.f;
    int numTris = 0, numSpheres = 0;
    CamView cam{};
    float background[3] = {0, 0, 0};
    int numPoints = 0;
    float pointPos[3] = {0, 0, 0}, pointInt[3] = {0, 0, 0};
};

// Validates the descriptor and builds everything. Returns false with a
// message on invalid input.
bool buildHostScene(const prr_scene_desc *d, HostScene &out, std::string &err);
// Guide tables of the environment CDFs (sets envGuideK, or leaves it 0).
void buildEnvGuides(HostScene &S);
// Camera::init (Camera.hpp:43-56): basis, focus distance, tan(fovh / 2).
CamView makeCamView(const prr_camera_desc &c);

}  // namespace host
}  // namespace prr
