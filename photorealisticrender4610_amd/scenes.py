"""Scene description on the host and the procedural workload scenes.

A :class:`SceneData` holds exactly what ``prr_scene_desc`` (include/prr_abi.h)
carries: an ordered object list (the reference's ``Scene::objects``,
Scene.hpp:18), triangle rows, sphere rows, materials, environment, camera,
background and point lights. :class:`SceneBuilder` mirrors the reference's
scene-script pattern (``scene.Add(new MeshTriangle(...))``,
scenes/museum_final.cpp) for procedural content.

Workloads
---------
``box_room``      256x256 closed box, GGX conductor + glass spheres, quad light.
``sphere_field``  4096 UV spheres x 256 tris (1,048,576) + ground quad, GGX
                  materials, 2048x1024 sky with a 2-degree sun, 1920x1080.
``glass_torus``   65,536-tri glass torus on a Lambert plane, point light.
"""
from __future__ import annotations

import dataclasses
import math
from typing import List, Optional

import numpy as np

MAT_DIFFUSE, MAT_GLASS, MAT_MIRROR, MAT_EMISSIVE = 0, 1, 2, 3
OBJ_MESH, OBJ_SPHERE = 0, 1


@dataclasses.dataclass
class SceneData:
    objects: np.ndarray      # int32 [n, 4]: kind, material, first, count
    tris: np.ndarray         # float32 [nt, 9]: v0, v1, v2
    spheres: np.ndarray      # float32 [ns, 4]: center, radius
    materials: np.ndarray    # float32 [nm, 12]: type, base3, rough, metal, ior, emission3, pad2
    background: np.ndarray   # float32 [3]
    camera: np.ndarray       # float32 [12]: eye3, target3, up3, fovh, aperture, focus
    env: Optional[np.ndarray] = None          # float32 [H, W, 3], row 0 = zenith
    point_lights: Optional[np.ndarray] = None  # float32 [np, 6]: position, intensity
    name: str = "scene"
    # per-vertex attributes (all or none): prr_tri_flag bits, n0..n2, t0..t2, tangent xyzw x3
    tri_flags: Optional[np.ndarray] = None     # uint8 [nt]
    tri_normals: Optional[np.ndarray] = None   # float32 [nt, 9]
    tri_uvs: Optional[np.ndarray] = None       # float32 [nt, 6]
    tri_tangents: Optional[np.ndarray] = None  # float32 [nt, 12]
    mat_tex: Optional[np.ndarray] = None       # int32 [nm, 4]: base colour, metal-rough, normal, emissive (-1 none)
    textures: Optional[list] = None            # uint8 [H, W, C] each

    @property
    def num_triangles(self):
        return int(self.tris.shape[0])

    @property
    def num_spheres(self):
        return int(self.spheres.shape[0])

    def bundle_arrays(self):
        """Arrays in the golden-harness bundle layout."""
        out = {
            "objects": self.objects.astype(np.int32).ravel(),
            "tris": self.tris.astype(np.float32).ravel(),
            "spheres": self.spheres.astype(np.float32).ravel(),
            "materials": self.materials.astype(np.float32).ravel(),
            "background": self.background.astype(np.float32),
            "camera": self.camera.astype(np.float32),
        }
        if self.env is not None:
            h, w, _ = self.env.shape
            out["env_wh"] = np.array([w, h], np.int32)
            out["env_pixels"] = self.env.astype(np.float32).ravel()
        if self.tri_flags is not None:
            out["tri_flags"] = self.tri_flags.astype(np.int32)
            out["tri_attr"] = np.concatenate([self.tri_normals, self.tri_uvs, self.tri_tangents], 1) \
                .astype(np.float32).ravel()
        if self.mat_tex is not None:
            out["mat_tex"] = self.mat_tex.astype(np.int32).ravel()
            out["tex_whc"] = np.array([[t.shape[1], t.shape[0], t.shape[2]] for t in self.textures],
                                      np.int32).reshape(-1)
            out["texels"] = np.concatenate([t.ravel() for t in self.textures]) if self.textures \
                else np.zeros(0, np.uint8)
        return out


class SceneBuilder:
    def __init__(self):
        self._objs: List[tuple] = []
        self._tris: List[np.ndarray] = []
        self._ntris = 0
        self._sph: List[List[float]] = []
        self._mats: List[List[float]] = []
        self.background = np.array([0.235294, 0.67451, 0.843137], np.float32)  # Scene.hpp:17
        self.env = None
        self.points: List[List[float]] = []
        self.camera = np.array([0, 0, -1, 0, 0, 0, 0, 1, 0, 40.0, 0.0, -1.0], np.float32)
        self._attrs: List[tuple] = []   # per mesh: (flags, normals, uvs, tangents) or None
        self._mat_tex: List[List[int]] = []
        self._textures: List[np.ndarray] = []

    def material(self, mtype=MAT_DIFFUSE, base=(0.8, 0.8, 0.8), roughness=0.5, metallic=0.0,
                 ior=1.5, emission=(0.0, 0.0, 0.0), textures=(-1, -1, -1, -1)) -> int:
        """textures: rows from :meth:`texture` for base colour, metallic-roughness
        (G = roughness, B = metallic), normal map, emissive; -1 none."""
        self._mats.append([float(mtype), *map(float, base), float(roughness), float(metallic),
                           float(ior), *map(float, emission), 0.0, 0.0])
        self._mat_tex.append([int(x) for x in textures])
        return len(self._mats) - 1

    def texture(self, img) -> int:
        """8-bit texture [H, W, C] (TextureUtils.hpp:9-19); returns its row."""
        t = np.ascontiguousarray(img, np.uint8)
        if t.ndim != 3 or t.shape[2] < 1:
            raise ValueError("texture must be [H, W, C] uint8")
        self._textures.append(t)
        return len(self._textures) - 1

    def mesh(self, tris, material: int, normals=None, uvs=None, tangents=None) -> int:
        """Triangles [n, 9]; optional per-vertex normals [n, 9] (normalised per
        vertex, Triangle::hasSmoothNormals), uvs [n, 6], tangents [n, 12]
        (xyz + handedness per vertex, Triangle::hasTangents)."""
        t = np.asarray(tris, np.float32).reshape(-1, 9)
        if t.shape[0] == 0:
            raise ValueError("empty mesh")
        n = t.shape[0]
        if normals is None and uvs is None and tangents is None:
            self._attrs.append(None)
        else:
            fl = np.zeros(n, np.uint8)
            def rows(a, k, bit):
                if a is None:
                    return np.zeros((n, k), np.float32)
                fl[:] |= bit
                return np.asarray(a, np.float32).reshape(n, k)
            self._attrs.append((fl, rows(normals, 9, 1), rows(uvs, 6, 4), rows(tangents, 12, 2)))
        self._objs.append((OBJ_MESH, material, self._ntris, n))
        self._tris.append(t)
        self._ntris += n
        return len(self._objs) - 1

    def gltf(self, path, material_type=MAT_DIFFUSE, material: Optional[int] = None) -> List[int]:
        """GLTFLoader::load (GLTFLoader.cpp:140-281) + scene.Add of each mesh:
        one mesh per triangle primitive with its own material (or `material`,
        the loader's override). Returns the object rows."""
        from . import abi
        g = abi.load_gltf(path, material_type)
        base = len(self._textures)
        self._textures += g["textures"]
        out = []
        for first, count, mrow in g["meshes"]:
            if material is None:
                m = g["materials"][mrow]
                tx = [x + base if x >= 0 else -1 for x in g["mat_tex"][mrow]]
                mid = self.material(int(m[0]), m[1:4], m[4], m[5], m[6], m[7:10], tx)
            else:
                mid = material
            sl = slice(first, first + count)
            fl = g["flags"][sl]
            out.append(self.mesh(g["positions"][sl], mid,
                                 normals=g["normals"][sl], uvs=g["uvs"][sl], tangents=g["tangents"][sl]))
            self._attrs[-1][0][:] = fl
        return out

    def sphere(self, center, radius, material: int) -> int:
        self._objs.append((OBJ_SPHERE, material, len(self._sph), 1))
        self._sph.append([*map(float, center), float(radius)])
        return len(self._objs) - 1

    def point_light(self, position, intensity):
        self.points.append([*map(float, position), *map(float, intensity)])

    def set_camera(self, eye, target, up=(0, 1, 0), fovh=40.0, aperture=0.0, focus=-1.0):
        self.camera = np.array([*eye, *target, *up, fovh, aperture, focus], np.float32)

    def build(self, name="scene") -> SceneData:
        tris = np.concatenate(self._tris, 0) if self._tris else np.zeros((0, 9), np.float32)
        return SceneData(
            objects=np.array(self._objs, np.int32).reshape(-1, 4),
            tris=np.ascontiguousarray(tris, np.float32),
            spheres=np.array(self._sph, np.float32).reshape(-1, 4),
            materials=np.array(self._mats, np.float32).reshape(-1, 12),
            background=np.asarray(self.background, np.float32),
            camera=self.camera.astype(np.float32),
            env=None if self.env is None else np.ascontiguousarray(self.env, np.float32),
            point_lights=np.array(self.points, np.float32).reshape(-1, 6) if self.points else None,
            name=name,
            **self._attr_fields(),
        )

    def _attr_fields(self):
        out = {}
        if any(a is not None for a in self._attrs):
            parts = [a if a is not None else (np.zeros(t.shape[0], np.uint8), np.zeros((t.shape[0], 9), np.float32),
                                              np.zeros((t.shape[0], 6), np.float32),
                                              np.zeros((t.shape[0], 12), np.float32))
                     for a, t in zip(self._attrs, self._tris)]
            out["tri_flags"] = np.concatenate([p[0] for p in parts])
            out["tri_normals"] = np.concatenate([p[1] for p in parts])
            out["tri_uvs"] = np.concatenate([p[2] for p in parts])
            out["tri_tangents"] = np.concatenate([p[3] for p in parts])
        if self._textures or any(any(x >= 0 for x in r) for r in self._mat_tex):
            out["mat_tex"] = np.array(self._mat_tex, np.int32).reshape(-1, 4)
            out["textures"] = list(self._textures)
        return out


# ---------------------------------------------------------------------------
# geometry helpers
# ---------------------------------------------------------------------------
def quad(p0, p1, p2, p3):
    """Two triangles (p0,p1,p2), (p0,p2,p3); normal = (p1-p0)x(p2-p0)."""
    p0, p1, p2, p3 = (np.asarray(p, np.float32) for p in (p0, p1, p2, p3))
    return np.stack([np.concatenate([p0, p1, p2]), np.concatenate([p0, p2, p3])]).astype(np.float32)


def uv_sphere(segments=16, bands=9):
    """Unit UV sphere: triangle fans at the poles, quads (2 tris) in between.
    16 segments x 9 bands = 16 + 16 + 7*32 = 256 triangles, outward winding."""
    th = np.linspace(0.0, math.pi, bands + 1)
    ph = np.linspace(0.0, 2 * math.pi, segments + 1)
    def P(i, j):
        t, p = th[i], ph[j % segments]
        return np.array([math.sin(t) * math.cos(p), math.cos(t), math.sin(t) * math.sin(p)])
    tris = []
    for i in range(bands):
        for j in range(segments):
            a, b, c, d = P(i, j), P(i, j + 1), P(i + 1, j + 1), P(i + 1, j)
            if i == 0:
                tris.append(np.concatenate([a, c, d]))
            elif i == bands - 1:
                tris.append(np.concatenate([a, b, d]))
            else:
                tris.append(np.concatenate([a, b, c]))
                tris.append(np.concatenate([a, c, d]))
    out = np.array(tris, np.float64)
    # orient outward: normal (v1-v0)x(v2-v0) must point away from the centre
    v0, v1, v2 = out[:, 0:3], out[:, 3:6], out[:, 6:9]
    n = np.cross(v1 - v0, v2 - v0)
    flip = np.einsum("ij,ij->i", n, (v0 + v1 + v2) / 3) < 0
    out[flip] = np.concatenate([v0[flip], v2[flip], v1[flip]], 1)
    return out


def torus(R=1.0, r=0.35, nu=256, nv=128):
    """Torus around the y axis, (nu x nv) grid, 2 triangles per cell."""
    u = np.arange(nu) * (2 * math.pi / nu)
    v = np.arange(nv) * (2 * math.pi / nv)
    def P(iu, iv):
        uu, vv = u[iu % nu], v[iv % nv]
        return np.stack([(R + r * np.cos(vv)) * np.cos(uu), r * np.sin(vv),
                         (R + r * np.cos(vv)) * np.sin(uu)], -1)
    iu, iv = np.meshgrid(np.arange(nu), np.arange(nv), indexing="ij")
    iu, iv = iu.ravel(), iv.ravel()
    a, b, c, d = P(iu, iv), P(iu + 1, iv), P(iu + 1, iv + 1), P(iu, iv + 1)
    t1 = np.concatenate([a, c, b], 1)
    t2 = np.concatenate([a, d, c], 1)
    out = np.stack([t1, t2], 1).reshape(-1, 9)
    v0, v1, v2 = out[:, 0:3], out[:, 3:6], out[:, 6:9]
    n = np.cross(v1 - v0, v2 - v0)
    cen = (v0 + v1 + v2) / 3
    ring = cen.copy()
    rr = np.linalg.norm(cen[:, [0, 2]], axis=1, keepdims=True)
    ring[:, 0:1] = cen[:, 0:1] / rr * R
    ring[:, 2:3] = cen[:, 2:3] / rr * R
    ring[:, 1] = 0
    flip = np.einsum("ij,ij->i", n, cen - ring) < 0
    out[flip] = np.concatenate([v0[flip], v2[flip], v1[flip]], 1)
    return out


def random_rotation(rng):
    """Uniform random rotation matrix from a Shoemake quaternion."""
    u1, u2, u3 = rng.random(3)
    q = np.array([math.sqrt(1 - u1) * math.sin(2 * math.pi * u2), math.sqrt(1 - u1) * math.cos(2 * math.pi * u2),
                  math.sqrt(u1) * math.sin(2 * math.pi * u3), math.sqrt(u1) * math.cos(2 * math.pi * u3)])
    x, y, z, w = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def sky(width, height, sun_dir=(0.45, 0.55, 0.35), sun_sigma_deg=2.0, sun_peak=1.0e4,
        zenith=(0.25, 0.45, 0.9), horizon=(0.85, 0.9, 1.0), ground=(0.25, 0.22, 0.2)):
    """Equirectangular sky in EnvMap coordinates (EnvMap.hpp:71-89: y up,
    row 0 = zenith, phi = atan2(z, x)): zenith->horizon gradient above the
    horizon, constant ground below, plus a Gaussian sun lobe
    sun_peak * exp(-angle^2 / (2 sigma^2)) of angular sigma sun_sigma_deg."""
    v = (np.arange(height) + 0.5) / height
    u = (np.arange(width) + 0.5) / width
    theta = v * math.pi
    phi = u * 2 * math.pi - math.pi
    st = np.sin(theta)[:, None]
    d = np.stack([st * np.cos(phi)[None, :], np.cos(theta)[:, None] * np.ones((1, width)),
                  st * np.sin(phi)[None, :]], -1)
    up = np.clip(d[..., 1], 0, 1)[..., None]
    zen, hor, gnd = (np.array(c, np.float64) for c in (zenith, horizon, ground))
    t = np.power(up, 0.5)
    col = hor * (1 - t) + zen * t
    col = np.where(d[..., 1:2] >= 0, col, gnd)
    s = np.asarray(sun_dir, np.float64)
    s = s / np.linalg.norm(s)
    ang = np.arccos(np.clip(d @ s, -1, 1))
    sig = math.radians(sun_sigma_deg)
    col = col + sun_peak * np.exp(-(ang ** 2) / (2 * sig * sig))[..., None] * np.array([1.0, 0.95, 0.85])
    return col.astype(np.float32)


# ---------------------------------------------------------------------------
# workload scenes
# ---------------------------------------------------------------------------
def box_room() -> SceneData:
    """Closed box room: 5 wall quads (10 tris; Lambert-like DiffuseMaterial
    m=0 r=1 ior=1.5), GGX conductor sphere (r 0.3, F0 0.95/0.64/0.54), glass
    sphere (ior 1.5, tint 1), ceiling quad light (17, 12, 4), fovh 40."""
    b = SceneBuilder()
    b.background = np.zeros(3, np.float32)
    grey = b.material(MAT_DIFFUSE, (0.73, 0.73, 0.73), 1.0, 0.0, 1.5)
    red = b.material(MAT_DIFFUSE, (0.65, 0.05, 0.05), 1.0, 0.0, 1.5)
    green = b.material(MAT_DIFFUSE, (0.12, 0.45, 0.15), 1.0, 0.0, 1.5)
    light = b.material(MAT_EMISSIVE, (0.8, 0.8, 0.8), 0.5, 0.0, 1.5, (17.0, 12.0, 4.0))
    metal = b.material(MAT_DIFFUSE, (0.95, 0.64, 0.54), 0.3, 1.0, 1.5)
    glass = b.material(MAT_GLASS, (1.0, 1.0, 1.0), 0.5, 0.0, 1.5)
    b.mesh(quad((-1, 0, -1), (-1, 0, 1), (1, 0, 1), (1, 0, -1)), grey)      # floor
    b.mesh(quad((-1, 2, -1), (1, 2, -1), (1, 2, 1), (-1, 2, 1)), grey)      # ceiling
    b.mesh(quad((-1, 0, -1), (1, 0, -1), (1, 2, -1), (-1, 2, -1)), grey)    # back
    b.mesh(quad((-1, 0, -1), (-1, 2, -1), (-1, 2, 1), (-1, 0, 1)), red)     # left
    b.mesh(quad((1, 0, -1), (1, 0, 1), (1, 2, 1), (1, 2, -1)), green)       # right
    b.mesh(quad((-0.25, 1.98, -0.25), (0.25, 1.98, -0.25), (0.25, 1.98, 0.25), (-0.25, 1.98, 0.25)), light)
    b.sphere((-0.45, 0.4, -0.3), 0.4, metal)
    b.sphere((0.45, 0.4, 0.25), 0.4, glass)
    b.set_camera((0, 1, 3.9), (0, 1, 0), (0, 1, 0), 40.0)
    return b.build("box_room")


def sphere_field(num_spheres=4096, seed=0xC1, env_size=(2048, 1024)) -> SceneData:
    """num_spheres UV spheres (256 tris) with seeded position U[0,100]^3,
    scale U[0.5,3], uniform rotation, roughness U[0.05,1], metallic
    Bernoulli(0.5), base colour U[0,1]^3; ground quad y=0 over [-50,150]^2;
    sky with a 2-degree Gaussian sun of peak 1e4."""
    rng = np.random.default_rng(seed)
    unit = uv_sphere()
    b = SceneBuilder()
    ground = b.material(MAT_DIFFUSE, (0.5, 0.5, 0.5), 1.0, 0.0, 1.5)
    for _ in range(num_spheres):
        pos = rng.random(3) * 100.0
        scale = 0.5 + rng.random() * 2.5
        rot = random_rotation(rng)
        rough = 0.05 + rng.random() * 0.95
        metal = 1.0 if rng.random() < 0.5 else 0.0
        base = rng.random(3)
        m = b.material(MAT_DIFFUSE, base, rough, metal, 1.5)
        v = unit.reshape(-1, 3) @ rot.T * scale + pos
        b.mesh(v.reshape(-1, 9), m)
    b.mesh(quad((-50, 0, -50), (-50, 0, 150), (150, 0, 150), (150, 0, -50)), ground)
    b.env = sky(*env_size)
    b.set_camera((-60.0, 45.0, -70.0), (50.0, 30.0, 50.0), (0, 1, 0), 60.0)
    return b.build("sphere_field")


def glass_torus(nu=256, nv=128) -> SceneData:
    """Glass torus (R 1, r 0.35, ior 1.5) on a Lambert plane (0.8) at
    y=-0.35, lit by a point light of power 1000 W at (0, 5, 0)."""
    b = SceneBuilder()
    b.background = np.zeros(3, np.float32)
    glass = b.material(MAT_GLASS, (1.0, 1.0, 1.0), 0.5, 0.0, 1.5)
    plane = b.material(MAT_DIFFUSE, (0.8, 0.8, 0.8), 1.0, 0.0, 1.5)
    b.mesh(torus(1.0, 0.35, nu, nv), glass)
    b.mesh(quad((-10, -0.35, -10), (-10, -0.35, 10), (10, -0.35, 10), (10, -0.35, -10)), plane)
    intensity = 1000.0 / (4 * math.pi)
    b.point_light((0.0, 5.0, 0.0), (intensity,) * 3)
    b.set_camera((0.0, 3.2, 4.2), (0.0, -0.2, 0.0), (0, 1, 0), 45.0)
    return b.build("glass_torus")
