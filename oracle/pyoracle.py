"""ctypes wrapper of the CPU oracle (oracle/build/liboracle*.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's CPU-baseline leg, as the checker. The ctypes structures below
restate include/prr_abi.h independently of the product package.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_LIBS = {}


class MaterialDesc(C.Structure):
    _fields_ = [("type", C.c_int32), ("base_color", C.c_float * 3), ("roughness", C.c_float),
                ("metallic", C.c_float), ("ior", C.c_float), ("emission", C.c_float * 3)]


class ObjectDesc(C.Structure):
    _fields_ = [("kind", C.c_int32), ("material", C.c_int32), ("first", C.c_int32), ("count", C.c_int32)]


class CameraDesc(C.Structure):
    _fields_ = [("eye", C.c_float * 3), ("target", C.c_float * 3), ("up", C.c_float * 3), ("fovh", C.c_float),
                ("aperture", C.c_float), ("focus_distance", C.c_float)]


class EnvDesc(C.Structure):
    _fields_ = [("width", C.c_int32), ("height", C.c_int32), ("pixels", C.POINTER(C.c_float))]


class PointLightDesc(C.Structure):
    _fields_ = [("position", C.c_float * 3), ("intensity", C.c_float * 3)]


class TextureDesc(C.Structure):
    _fields_ = [("width", C.c_int32), ("height", C.c_int32), ("channels", C.c_int32),
                ("data", C.POINTER(C.c_uint8))]


class SceneDesc(C.Structure):
    _fields_ = [("num_objects", C.c_int32), ("objects", C.POINTER(ObjectDesc)),
                ("num_triangles", C.c_int32), ("triangles", C.POINTER(C.c_float)),
                ("num_spheres", C.c_int32), ("spheres", C.POINTER(C.c_float)),
                ("num_materials", C.c_int32), ("materials", C.POINTER(MaterialDesc)),
                ("env", EnvDesc), ("background", C.c_float * 3),
                ("num_point_lights", C.c_int32), ("point_lights", C.POINTER(PointLightDesc)),
                ("camera", CameraDesc),
                ("tri_flags", C.POINTER(C.c_uint8)), ("tri_normals", C.POINTER(C.c_float)),
                ("tri_uvs", C.POINTER(C.c_float)), ("tri_tangents", C.POINTER(C.c_float)),
                ("material_textures", C.POINTER(C.c_int32)), ("num_textures", C.c_int32),
                ("textures", C.POINTER(TextureDesc))]


class RenderParams(C.Structure):
    _fields_ = [("width", C.c_int32), ("height", C.c_int32), ("spp", C.c_int32), ("max_depth", C.c_int32),
                ("indirect_clamp", C.c_float), ("sample_offset", C.c_int32), ("seed", C.c_uint64),
                ("num_photons", C.c_int64), ("photon_radius", C.c_float), ("photon_max_depth", C.c_int32),
                ("flags", C.c_int32), ("min_spp", C.c_int32), ("variance_threshold", C.c_float),
                ("reserved", C.c_int32)]


def _f(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _p(a):
    return C.c_void_p(a.ctypes.data) if a is not None else None


def scene_desc(sc):
    """(SceneDesc, keepalive) from a SceneData-like object."""
    keep = []
    objs = np.ascontiguousarray(sc.objects, np.int32)
    tris = np.ascontiguousarray(sc.tris, np.float32)
    sph = np.ascontiguousarray(sc.spheres, np.float32).reshape(-1, 4)
    mats = np.ascontiguousarray(sc.materials, np.float32).reshape(-1, 12)
    ma = (MaterialDesc * max(1, len(mats)))()
    for i, m in enumerate(mats):
        ma[i].type = int(m[0])
        ma[i].base_color[:] = [float(x) for x in m[1:4]]
        ma[i].roughness, ma[i].metallic, ma[i].ior = float(m[4]), float(m[5]), float(m[6])
        ma[i].emission[:] = [float(x) for x in m[7:10]]
    oa = (ObjectDesc * max(1, len(objs)))()
    for i, o in enumerate(objs):
        oa[i].kind, oa[i].material, oa[i].first, oa[i].count = (int(x) for x in o)
    d = SceneDesc()
    d.num_objects = len(objs)
    d.objects = C.cast(oa, C.POINTER(ObjectDesc))
    d.num_triangles = tris.shape[0]
    d.triangles = _f(tris)
    d.num_spheres = sph.shape[0]
    d.spheres = _f(sph)
    d.num_materials = len(mats)
    d.materials = C.cast(ma, C.POINTER(MaterialDesc))
    env = getattr(sc, "env", None)
    if env is not None:
        env = np.ascontiguousarray(env, np.float32)
        d.env.height, d.env.width = env.shape[0], env.shape[1]
        d.env.pixels = _f(env)
    d.background[:] = [float(x) for x in sc.background]
    pl = getattr(sc, "point_lights", None)
    if pl is not None and len(pl):
        pa = (PointLightDesc * len(pl))()
        for i, p in enumerate(pl):
            pa[i].position[:] = [float(x) for x in p[:3]]
            pa[i].intensity[:] = [float(x) for x in p[3:6]]
        d.num_point_lights = len(pl)
        d.point_lights = C.cast(pa, C.POINTER(PointLightDesc))
        keep.append(pa)
    if getattr(sc, "tri_flags", None) is not None:
        fl = np.ascontiguousarray(sc.tri_flags, np.uint8)
        nrm = np.ascontiguousarray(sc.tri_normals, np.float32)
        uvs = np.ascontiguousarray(sc.tri_uvs, np.float32)
        tan = np.ascontiguousarray(sc.tri_tangents, np.float32)
        d.tri_flags = fl.ctypes.data_as(C.POINTER(C.c_uint8))
        d.tri_normals, d.tri_uvs, d.tri_tangents = _f(nrm), _f(uvs), _f(tan)
        keep += [fl, nrm, uvs, tan]
    if getattr(sc, "mat_tex", None) is not None:
        mt = np.ascontiguousarray(sc.mat_tex, np.int32)
        d.material_textures = mt.ctypes.data_as(C.POINTER(C.c_int32))
        texs = [np.ascontiguousarray(t, np.uint8) for t in sc.textures]
        ta = (TextureDesc * max(1, len(texs)))()
        for i, t in enumerate(texs):
            ta[i].height, ta[i].width, ta[i].channels = t.shape[0], t.shape[1], t.shape[2]
            ta[i].data = t.ctypes.data_as(C.POINTER(C.c_uint8))
        d.num_textures = len(texs)
        d.textures = C.cast(ta, C.POINTER(TextureDesc))
        keep += [mt, ta, texs]
    cam = np.asarray(sc.camera, np.float32)
    d.camera.eye[:] = [float(x) for x in cam[0:3]]
    d.camera.target[:] = [float(x) for x in cam[3:6]]
    d.camera.up[:] = [float(x) for x in cam[6:9]]
    d.camera.fovh, d.camera.aperture, d.camera.focus_distance = float(cam[9]), float(cam[10]), float(cam[11])
    keep += [objs, tris, sph, mats, ma, oa, env]
    return d, keep


def lib(fast=False):
    name = "liboracle_fast.so" if fast else "liboracle.so"
    if name in _LIBS:
        return _LIBS[name]
    path = os.path.join(HERE, "build", name)
    if not os.path.exists(path):
        raise RuntimeError("oracle library %s is not built (run __graft_entry__.build())" % path)
    L = C.CDLL(path)
    vp = C.c_void_p
    L.orc_scene_create.restype = vp
    L.orc_scene_create.argtypes = [C.POINTER(SceneDesc)]
    L.orc_scene_free.argtypes = [vp]
    L.orc_export_level.restype = C.c_int
    L.orc_export_level.argtypes = [vp, C.c_int, C.c_int, vp, vp, vp]
    L.orc_export_order.restype = C.c_int
    L.orc_export_order.argtypes = [vp, C.c_int, C.c_int, vp, vp]
    L.orc_sample_light.argtypes = [vp, C.c_int64, vp, vp, vp, vp]
    L.orc_total_emit_area.restype = C.c_float
    L.orc_total_emit_area.argtypes = [vp]
    L.orc_trace_closest.argtypes = [vp, C.c_int64, vp, vp, vp, vp, vp, vp, vp]
    L.orc_trace_any.argtypes = [vp, C.c_int64, vp, vp, vp, vp, vp]
    L.orc_brute_closest.argtypes = [vp, C.c_int64, vp, vp, vp, vp]
    L.orc_brute_any.argtypes = [vp, C.c_int64, vp, vp, vp, vp]
    L.orc_bsdf_eval.argtypes = [C.c_int64, vp, vp, vp, vp, vp, vp]
    L.orc_bsdf_sample.argtypes = [C.c_int64, vp, vp, vp, vp, vp, vp, vp, vp]
    L.orc_ggx_terms.argtypes = [C.c_int64, vp, vp]
    L.orc_env_build.argtypes = [C.c_int, C.c_int, vp, vp, vp, vp]
    L.orc_env_sample.argtypes = [C.c_int, C.c_int, vp, C.c_int64, vp, vp, vp]
    L.orc_env_eval.argtypes = [C.c_int, C.c_int, vp, C.c_int64, vp, vp, vp]
    L.orc_camera_rays.argtypes = [C.POINTER(CameraDesc), C.c_int, C.c_int, C.c_int64, vp, vp, vp, vp]
    L.orc_tonemap.argtypes = [C.c_int, C.c_int64, C.c_float, vp, vp, vp]
    L.orc_medium.argtypes = [C.c_int64, vp, vp, vp, vp, vp, vp, vp]
    L.orc_phase_hg.argtypes = [C.c_int64, vp, vp, vp]
    L.orc_phase_hg_sample.argtypes = [C.c_int64, vp, vp, vp, vp]
    L.orc_castray_splitmix.argtypes = [vp, C.c_int, C.c_int, C.c_int64, vp, vp, C.c_int, C.c_float, vp]
    L.orc_castray_pcg.argtypes = [vp, C.c_int64, vp, vp, C.c_uint64, C.c_int, C.c_float, vp]
    L.orc_render.argtypes = [vp, C.POINTER(RenderParams), C.c_int, C.c_int, vp, C.c_int, vp, vp, vp, vp]
    L.orc_welford.argtypes = [C.c_int, C.c_int, vp, vp, vp, vp, vp]
    L.orc_render_truncated.argtypes = [vp, C.POINTER(RenderParams), C.c_int, C.c_int, C.c_int, C.c_int, C.c_uint32,
                                       C.c_int, vp, vp]
    L.orc_to_half.argtypes = [C.c_int64, vp, vp]
    L.orc_cache_encode.argtypes = [C.c_int64, vp, vp, vp]
    L.orc_cache_forward.argtypes = [C.c_int64, vp, vp, vp, C.c_int, vp]
    L.orc_cache_gradient.argtypes = [C.c_int64, vp, vp, vp, vp, C.c_int, C.c_float, vp]
    L.orc_cache_gradient.restype = C.c_double
    L.orc_cache_adam.argtypes = [vp, vp, vp, vp, C.c_int64, C.c_float, C.c_float, C.c_float, C.c_float]
    L.orc_photon_build.restype = vp
    L.orc_photon_build.argtypes = [vp, C.c_int64, C.c_float, C.c_int, C.c_uint64, C.c_int]
    L.orc_photon_free.argtypes = [vp]
    L.orc_photon_emit.restype = vp
    L.orc_photon_emit.argtypes = [vp, C.c_int64, C.c_float, C.c_int, C.c_uint64, C.c_int]
    L.orc_photon_emit_splitmix.restype = C.c_int64
    L.orc_photon_emit_splitmix.argtypes = [vp, C.c_int64, vp, C.c_float, C.c_int, C.c_int64, vp, vp]
    L.orc_photon_map_from.restype = vp
    L.orc_photon_map_from.argtypes = [C.c_int64, vp, C.c_float]
    L.orc_photon_owners.restype = C.c_int64
    L.orc_photon_owners.argtypes = [vp, C.c_int64, vp, vp]
    L.orc_pcg32_sequence.argtypes = [C.c_uint64, C.c_uint64, C.c_int, vp, vp]
    L.orc_pcg32_stream.argtypes = [C.c_uint64, C.c_uint32, C.c_uint32, C.c_int, vp]
    L.orc_photon_count.restype = C.c_int64
    L.orc_photon_count.argtypes = [vp]
    L.orc_photon_export.restype = C.c_int64
    L.orc_photon_export.argtypes = [vp, C.c_int64, vp]
    L.orc_photon_irradiance.argtypes = [vp, C.c_int64, vp, vp]
    L.orc_photon_radiance.argtypes = [vp, C.c_int64, vp, vp, vp, vp, vp]
    _LIBS[name] = L
    return L


def _c(a, dt=np.float32):
    return np.ascontiguousarray(a, dt)


class OracleScene:
    def __init__(self, sc, fast=False):
        self.L = lib(fast)
        self.sc = sc
        d, keep = scene_desc(sc)
        self.h = self.L.orc_scene_create(C.byref(d))
        del keep

    def __del__(self):
        try:
            self.L.orc_scene_free(self.h)
        except Exception:
            pass

    def export_level(self, level):
        n = self.L.orc_export_level(self.h, level, 0, None, None, None)
        if n < 0:
            return None
        b = np.zeros((n, 6), np.float32)
        fc = np.zeros(n, np.uint32)
        meta = np.zeros(n, np.uint32)
        self.L.orc_export_level(self.h, level, n, _p(b), _p(fc), _p(meta))
        return b, fc, meta

    def export_order(self, level):
        n = self.L.orc_export_order(self.h, level, 0, None, None)
        if n < 0:
            return None
        nodes = self.export_level(level)[0].shape[0]
        order = np.zeros(n, np.int32)
        areas = np.zeros(max(n, nodes), np.float32)
        self.L.orc_export_order(self.h, level, max(n, nodes), _p(order), _p(areas))
        return order, areas[:nodes]

    def sample_light(self, seeds_u32):
        s = np.asarray(seeds_u32, np.uint64)
        s64 = _c((s << np.uint64(32)) | s, np.uint64)
        n = s64.shape[0]
        pos = np.zeros((n, 3), np.float32)
        nrm = np.zeros((n, 3), np.float32)
        pdf = np.zeros(n, np.float32)
        self.L.orc_sample_light(self.h, n, _p(s64), _p(pos), _p(nrm), _p(pdf))
        return pos, nrm, pdf

    def total_emit_area(self):
        return float(self.L.orc_total_emit_area(self.h))

    def closest(self, org, d, counters=False):
        org, d = _c(org), _c(d)
        n = org.shape[0]
        prim = np.zeros(n, np.int32)
        t, u, v = (np.zeros(n, np.float32) for _ in range(3))
        cnt = np.zeros((n, 3), np.uint32) if counters else None
        self.L.orc_trace_closest(self.h, n, _p(org), _p(d), _p(prim), _p(t), _p(u), _p(v), _p(cnt))
        return (prim, t, u, v, cnt) if counters else (prim, t, u, v)

    def any(self, org, d, tmax, counters=False):
        org, d, tmax = _c(org), _c(d), _c(tmax)
        n = org.shape[0]
        hit = np.zeros(n, np.uint8)
        cnt = np.zeros((n, 3), np.uint32) if counters else None
        self.L.orc_trace_any(self.h, n, _p(org), _p(d), _p(tmax), _p(hit), _p(cnt))
        return (hit, cnt) if counters else hit

    def brute_closest(self, org, d):
        org, d = _c(org), _c(d)
        n = org.shape[0]
        prim = np.zeros(n, np.int32)
        t = np.zeros(n, np.float32)
        self.L.orc_brute_closest(self.h, n, _p(org), _p(d), _p(prim), _p(t))
        return prim, t

    def brute_any(self, org, d, tmax):
        org, d, tmax = _c(org), _c(d), _c(tmax)
        hit = np.zeros(org.shape[0], np.uint8)
        self.L.orc_brute_any(self.h, org.shape[0], _p(org), _p(d), _p(tmax), _p(hit))
        return hit

    def castray_splitmix(self, w, h, pix, seeds_u32, max_depth, clamp=10.0):
        pix = _c(pix, np.int32)
        s = np.asarray(seeds_u32, np.uint64)
        s64 = _c((s << np.uint64(32)) | s, np.uint64)
        L = np.zeros((pix.shape[0], 3), np.float32)
        self.L.orc_castray_splitmix(self.h, w, h, pix.shape[0], _p(pix), _p(s64), max_depth, clamp, _p(L))
        return L

    def castray_pcg(self, org, d, seed, max_depth, clamp=10.0):
        org, d = _c(org), _c(d)
        L = np.zeros((org.shape[0], 3), np.float32)
        self.L.orc_castray_pcg(self.h, org.shape[0], _p(org), _p(d), seed, max_depth, clamp, _p(L))
        return L

    def render(self, width, height, spp, max_depth, seed, clamp=10.0, sample_offset=0, rows=None,
               photon_map=None, threads=0, want_var=False, min_spp=0, variance_threshold=0.0, want_spp=False):
        p = RenderParams(width=width, height=height, spp=spp, max_depth=max_depth, indirect_clamp=clamp,
                         sample_offset=sample_offset, seed=seed, min_spp=min_spp,
                         variance_threshold=variance_threshold)
        mean = np.zeros((height, width, 3), np.float32)
        var = np.zeros((height, width, 3), np.float32) if want_var else None
        sppo = np.zeros((height, width), np.int32) if want_spp else None
        cnt = np.zeros(7, np.uint64)
        r0, r1 = rows if rows else (0, 0)
        self.L.orc_render(self.h, C.byref(p), r0, r1, photon_map.h if photon_map else None, threads,
                          _p(mean), _p(var), _p(sppo), _p(cnt))
        keys = ["closest", "any", "samples", "inner", "tris", "inner_any", "tris_any"]
        stats = dict(zip(keys, (int(x) for x in cnt)))
        if want_spp:
            stats["spp"] = sppo
        return mean, var, stats


    def render_truncated(self, width, height, spp, max_depth, seed, query_bounce, train_every, clamp=10.0,
                         sample_offset=0, threads=0, train_suffix=0):
        """The cache render with a zero network: non-training paths end at
        their first diffuse vertex at bounce >= query_bounce (train_every
        None: no training paths); training paths end at their first diffuse
        vertex at bounce >= query_bounce + train_suffix (0: full depth)."""
        p = RenderParams(width=width, height=height, spp=spp, max_depth=max_depth, indirect_clamp=clamp,
                         sample_offset=sample_offset, seed=seed)
        mean = np.zeros((height, width, 3), np.float32)
        cnt = np.zeros(7, np.uint64)
        mask = 0xFFFFFFFF if train_every is None else train_every - 1
        stop = query_bounce + train_suffix if train_suffix > 0 else 0
        self.L.orc_render_truncated(self.h, C.byref(p), 0, 0, threads, query_bounce, mask, stop, _p(mean), _p(cnt))
        keys = ["closest", "any", "samples", "inner", "tris", "inner_any", "tris_any"]
        return mean, dict(zip(keys, (int(x) for x in cnt)))


def train_hash(pixel, sample):
    """Training-path selection hash of the cache render (prr_oracle.cpp
    trainHash): lowbias32 mix of pixel and sample index; a path is a training
    path when train_hash & (train_every - 1) == 0."""
    m = 0xFFFFFFFF
    h = (pixel * 0x9E3779B1 + sample * 0x85EBCA77) & m
    h ^= h >> 16
    h = (h * 0x7FEB352D) & m
    h ^= h >> 15
    h = (h * 0x846CA68B) & m
    h ^= h >> 16
    return h


class OraclePhotonMap:
    """rules "point": the build's caustic map from the point light
    (orc_photon_build); "reference": PhotonMap::emit's own rules from the
    area emitters (orc_photon_emit)."""

    def __init__(self, scene: OracleScene, num, radius, max_depth, seed, threads=0, rules="point"):
        self.L = scene.L
        fn = self.L.orc_photon_emit if rules == "reference" else self.L.orc_photon_build
        self.h = fn(scene.h, num, radius, max_depth, seed, threads)

    def __del__(self):
        try:
            self.L.orc_photon_free(self.h)
        except Exception:
            pass

    def count(self):
        return int(self.L.orc_photon_count(self.h))

    def export(self):
        n = self.count()
        out = np.zeros((n, 9), np.float32)
        self.L.orc_photon_export(self.h, n, _p(out))
        return out

    def radiance(self, pos, wo, nrm, params):
        pos, wo, nrm, params = _c(pos), _c(wo), _c(nrm), _c(params)
        out = np.zeros_like(pos)
        self.L.orc_photon_radiance(self.h, pos.shape[0], _p(pos), _p(wo), _p(nrm), _p(params), _p(out))
        return out

    @classmethod
    def from_photons(cls, ph9, radius):
        """A map over given photons (rows pos, dir, power)."""
        self = cls.__new__(cls)
        self.L = lib()
        self._keep = np.ascontiguousarray(ph9, np.float32).reshape(-1, 9)
        self.h = self.L.orc_photon_map_from(self._keep.shape[0], _p(self._keep), radius)
        return self

    def owners(self):
        """(emitting photon, trace depth) per stored photon (reference rules)."""
        n = self.count()
        own = np.zeros(n, np.int64)
        dep = np.zeros(n, np.int32)
        self.L.orc_photon_owners(self.h, n, _p(own), _p(dep))
        return own, dep

    def irradiance(self, pos):
        pos = _c(pos)
        out = np.zeros_like(pos)
        self.L.orc_photon_irradiance(self.h, pos.shape[0], _p(pos), _p(out))
        return out


def photon_emit_splitmix(scene: OracleScene, states_u64, num_photons, max_depth):
    """The reference's photon rules per photon k from SplitMix64 state
    states_u64[k]: (photons [m, 9], owner [m])."""
    st = np.ascontiguousarray(states_u64, np.uint64)
    cap = 64 * max(1, st.shape[0])
    out = np.zeros((cap, 9), np.float32)
    own = np.zeros(cap, np.int32)
    m = scene.L.orc_photon_emit_splitmix(scene.h, st.shape[0], _p(st), float(num_photons), max_depth, cap, _p(out),
                                         _p(own))
    assert m <= cap
    return out[:m], own[:m]


def pcg32_sequence(initstate, initseq, n):
    """(uint32 outputs, floats) after pcg32_srandom_r(initstate, initseq)."""
    L = lib()
    out = np.zeros(n, np.uint32)
    f = np.zeros(n, np.float32)
    L.orc_pcg32_sequence(initstate, initseq, n, _p(out), _p(f))
    return out, f


def pcg32_stream(seed, pixel, sample, n):
    f = np.zeros(n, np.float32)
    lib().orc_pcg32_stream(seed, pixel, sample, n, _p(f))
    return f


def bsdf_eval(wi, wo, n, params):
    L = lib()
    wi, wo, n, params = _c(wi), _c(wo), _c(n), _c(params)
    k = wi.shape[0]
    f = np.zeros((k, 3), np.float32)
    pdf = np.zeros(k, np.float32)
    L.orc_bsdf_eval(k, _p(wi), _p(wo), _p(n), _p(params), _p(f), _p(pdf))
    return f, pdf


def bsdf_sample(wo, n, params, u3):
    L = lib()
    wo, n, params, u3 = _c(wo), _c(n), _c(params), _c(u3)
    k = wo.shape[0]
    wi = np.zeros((k, 3), np.float32)
    f = np.zeros((k, 3), np.float32)
    pdf = np.zeros(k, np.float32)
    lobe = np.zeros(k, np.int32)
    L.orc_bsdf_sample(k, _p(wo), _p(n), _p(params), _p(u3), _p(wi), _p(f), _p(pdf), _p(lobe))
    return wi, f, pdf, lobe


def medium_intersect(org, d, bounds, bounded, fill=-7.0):
    """Medium::intersect per entry: (t_enter, t_exit, hit); t values stay `fill` on a miss."""
    L = lib()
    org, d, bounds = _c(org), _c(d), _c(bounds)
    bounded = np.ascontiguousarray(bounded, np.int32)
    k = org.shape[0]
    te = np.full(k, fill, np.float32)
    tx = np.full(k, fill, np.float32)
    hit = np.zeros(k, np.uint8)
    L.orc_medium(k, _p(org), _p(d), _p(bounds), _p(bounded), _p(te), _p(tx), _p(hit))
    return te, tx, hit


def phase_hg(cos, g):
    L = lib()
    cos, g = _c(cos), _c(g)
    out = np.zeros(cos.shape[0], np.float32)
    L.orc_phase_hg(cos.shape[0], _p(cos), _p(g), _p(out))
    return out


def phase_hg_sample(wo, g, u2):
    L = lib()
    wo, g, u2 = _c(wo), _c(g), _c(u2)
    wi = np.zeros((wo.shape[0], 3), np.float32)
    L.orc_phase_hg_sample(wo.shape[0], _p(wo), _p(g), _p(u2), _p(wi))
    return wi


def ggx_terms(in4):
    L = lib()
    in4 = _c(in4)
    out = np.zeros((in4.shape[0], 6), np.float32)
    L.orc_ggx_terms(in4.shape[0], _p(in4), _p(out))
    return out


def env_build(px):
    L = lib()
    px = _c(px)
    h, w, _ = px.shape
    cdf = np.zeros(w * h, np.float32)
    marg = np.zeros(h, np.float32)
    tot = np.zeros(1, np.float32)
    L.orc_env_build(w, h, _p(px), _p(cdf), _p(marg), _p(tot))
    return cdf, marg, float(tot[0])


def env_sample(px, u2):
    L = lib()
    px, u2 = _c(px), _c(u2)
    h, w, _ = px.shape
    n = u2.shape[0]
    d = np.zeros((n, 3), np.float32)
    pdf = np.zeros(n, np.float32)
    L.orc_env_sample(w, h, _p(px), n, _p(u2), _p(d), _p(pdf))
    return d, pdf


def env_eval(px, dirs):
    L = lib()
    px, dirs = _c(px), _c(dirs)
    h, w, _ = px.shape
    n = dirs.shape[0]
    rad = np.zeros((n, 3), np.float32)
    pdf = np.zeros(n, np.float32)
    L.orc_env_eval(w, h, _p(px), n, _p(dirs), _p(rad), _p(pdf))
    return rad, pdf


def camera_rays(cam12, w, h, pix, offs):
    L = lib()
    cd = CameraDesc()
    cam = np.asarray(cam12, np.float32)
    cd.eye[:] = [float(x) for x in cam[0:3]]
    cd.target[:] = [float(x) for x in cam[3:6]]
    cd.up[:] = [float(x) for x in cam[6:9]]
    cd.fovh, cd.aperture, cd.focus_distance = float(cam[9]), float(cam[10]), float(cam[11])
    pix, offs = _c(pix, np.int32), _c(offs)
    n = pix.shape[0]
    o = np.zeros((n, 3), np.float32)
    d = np.zeros((n, 3), np.float32)
    L.orc_camera_rays(C.byref(cd), w, h, n, _p(pix), _p(offs), _p(o), _p(d))
    return o, d


def tonemap(op, lin, exposure=1.0):
    L = lib()
    lin = _c(lin)
    n = lin.shape[0]
    disp = np.zeros((n, 3), np.float32)
    rgb8 = np.zeros((n, 3), np.uint8)
    L.orc_tonemap(op, n, exposure, _p(lin), _p(disp), _p(rgb8))
    return disp, rgb8


def splitmix_draws(seed_u32, k):
    """Reference RNG (global.hpp:35-50) from state (v<<32)|v, in numpy."""
    s = (np.uint64(seed_u32) << np.uint64(32)) | np.uint64(seed_u32)
    out = []
    with np.errstate(over="ignore"):
        for _ in range(k):
            s = s + np.uint64(0x9E3779B97F4A7C15)
            z = s
            z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
            z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
            z = z ^ (z >> np.uint64(31))
            out.append(np.float32(int(z >> np.uint64(40))) * np.float32(2.0 ** -24))
    return np.array(out, np.float32)


def welford(seq, min_spp, thr):
    """Renderer.cpp:71-99 stop rule on per-sample values seq (nseq, len, 3);
    min_spp / thr per sequence. Returns (samples taken, mean)."""
    seq = _c(seq)
    nseq, ln = seq.shape[0], seq.shape[1]
    ms = _c(np.broadcast_to(np.asarray(min_spp, np.int32), (nseq,)), np.int32)
    th = _c(np.broadcast_to(np.asarray(thr, np.float32), (nseq,)))
    cnt = np.zeros(nseq, np.int32)
    mean = np.zeros((nseq, 3), np.float32)
    lib().orc_welford(nseq, ln, _p(seq), _p(ms), _p(th), _p(cnt), _p(mean))
    return cnt, mean


# ---------------------------------------------------------------------------
# radiance-cache network
CACHE_WEIGHTS = 22720


def to_half(a):
    a = _c(a)
    out = np.empty_like(a)
    lib().orc_to_half(a.size, _p(a), _p(out))
    return out


def cache_encode(q, bounds):
    q, b = _c(q), _c(bounds)
    x = np.zeros((q.shape[0], 96), np.float32)
    lib().orc_cache_encode(q.shape[0], _p(q), _p(b), _p(x))
    return x


def cache_forward(q, bounds, w, half=True):
    q, b, w = _c(q), _c(bounds), _c(w)
    y = np.zeros((q.shape[0], 3), np.float32)
    lib().orc_cache_forward(q.shape[0], _p(q), _p(b), _p(w), int(half), _p(y))
    return y


def cache_gradient(q, t, bounds, w, half=True, clip=100.0):
    q, t, b, w = _c(q), _c(t), _c(bounds), _c(w)
    g = np.zeros(CACHE_WEIGHTS, np.float32)
    loss = lib().orc_cache_gradient(q.shape[0], _p(q), _p(t), _p(b), _p(w), int(half), clip, _p(g))
    return g, loss


def cache_adam(w, m1, m2, grad, step, lr=1e-2, b1=0.9, b2=0.99, eps=1e-8):
    """One Adam step in place on float32 arrays."""
    lib().orc_cache_adam(_p(w), _p(m1), _p(m2), _p(_c(grad)), step, lr, b1, b2, eps)
