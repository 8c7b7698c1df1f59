"""Small seeded scenes and query sets used by the golden fixtures and tests.

mini_field  a scaled-down sphere_field: UV spheres with GGX materials, a
            ground quad, a 64x32 sky with a wide sun, and a 12-triangle
            emissive mesh whose sub-BVH has interior nodes (so the light
            sample walk of BVH.cpp:478-508 is exercised).
soup        random triangle meshes + spheres + degenerate / duplicated /
            edge-sharing triangles, for traversal edge cases.
"""
import math

import numpy as np

from photorealisticrender4610_amd import scenes as S


def mini_field(seed=7, n=24):
    rng = np.random.default_rng(seed)
    unit = S.uv_sphere()
    b = S.SceneBuilder()
    ground = b.material(S.MAT_DIFFUSE, (0.5, 0.5, 0.5), 1.0, 0.0, 1.5)
    for _ in range(n):
        pos = rng.random(3) * 10.0
        scale = 0.3 + rng.random() * 1.2
        rot = S.random_rotation(rng)
        m = b.material(S.MAT_DIFFUSE, rng.random(3), 0.05 + rng.random() * 0.95,
                       1.0 if rng.random() < 0.5 else 0.0, 1.5)
        b.mesh((unit.reshape(-1, 3) @ rot.T * scale + pos).reshape(-1, 9), m)
    b.mesh(S.quad((-5, 0, -5), (-5, 0, 15), (15, 0, 15), (15, 0, -5)), ground)
    # glass and mirror spheres
    b.sphere((3.0, 1.0, 3.0), 1.0, b.material(S.MAT_GLASS, (1, 1, 1), 0.5, 0.0, 1.5))
    b.sphere((7.0, 1.2, 6.0), 1.2, b.material(S.MAT_MIRROR, (0.9, 0.9, 0.9), 0.5, 0.0, 1.5))
    # 12-triangle light panel facing down, irregular triangle areas
    lp = []
    xs = [2.0, 2.7, 3.1, 4.2, 5.0]
    for k in range(len(xs) - 1):
        x0, x1 = xs[k], xs[k + 1]
        lp.append([x0, 12, 2, x1, 12, 2, x1, 12, 4])
        lp.append([x0, 12, 2, x1, 12, 4, x0, 12, 4])
        lp.append([x0, 12, 4, x1, 12, 4, x1, 12, 5.5])
    light = b.material(S.MAT_EMISSIVE, (0.8, 0.8, 0.8), 0.5, 0.0, 1.5, (40.0, 35.0, 30.0))
    b.mesh(np.array(lp, np.float32), light)
    b.env = S.sky(64, 32, sun_sigma_deg=8.0, sun_peak=60.0)
    b.set_camera((-6.0, 6.0, -8.0), (5.0, 2.0, 5.0), (0, 1, 0), 55.0)
    return b.build("mini_field")


def soup(seed=11):
    rng = np.random.default_rng(seed)
    b = S.SceneBuilder()
    mats = [b.material(S.MAT_DIFFUSE, rng.random(3), 0.5, 0.0, 1.5) for _ in range(4)]
    for k in range(6):
        n = int(rng.integers(5, 300))
        c = rng.random(3) * 4.0
        v = c + (rng.random((n, 3, 3)) - 0.5) * rng.uniform(0.2, 2.0)
        b.mesh(v.reshape(-1, 9), mats[k % 4])
    # a grid mesh with shared edges (ties along edges/vertices)
    g = []
    for i in range(6):
        for j in range(6):
            p0 = (i * 0.5, 0.0, j * 0.5)
            g.append(S.quad(p0, (p0[0], 0.0, p0[2] + 0.5), (p0[0] + 0.5, 0.0, p0[2] + 0.5), (p0[0] + 0.5, 0.0, p0[2])))
    b.mesh(np.concatenate(g, 0), mats[1])
    # duplicated triangles (exact ties) and degenerate ones
    t = np.array([[1.0, 1.0, 1.0, 2.0, 1.0, 1.0, 1.0, 2.0, 1.5]], np.float32)
    degen = np.array([[0.5, 0.5, 0.5, 1.0, 1.0, 1.0, 1.5, 1.5, 1.5],
                      [2.0, 2.0, 2.0, 2.0, 2.0, 2.0, 2.0, 2.0, 2.0]], np.float32)
    b.mesh(np.concatenate([t, t, degen, t], 0), mats[2])
    b.mesh(np.array([[3.0, 0.2, 3.0, 3.5, 0.2, 3.0, 3.0, 0.8, 3.0]], np.float32), mats[3])  # 1-tri mesh
    for k in range(3):
        b.sphere(rng.random(3) * 4.0, rng.uniform(0.2, 0.7), mats[k])
    b.set_camera((-3, 3, -3), (2, 1, 2), (0, 1, 0), 50.0)
    return b.build("soup")


def scene_bounds(sc):
    pts = [sc.tris.reshape(-1, 3)]
    if sc.spheres.shape[0]:
        pts.append(sc.spheres[:, :3] - sc.spheres[:, 3:4])
        pts.append(sc.spheres[:, :3] + sc.spheres[:, 3:4])
    p = np.concatenate(pts, 0)
    return p.min(0), p.max(0)


def random_rays(sc, n, seed):
    """Rays for traversal parity: random origins in the enlarged scene box
    with random (unnormalized and normalized) directions, axis-aligned
    directions with signed zeros, rays aimed at triangle vertices/edges and
    rays starting on surfaces; tmax spans [0, 2 * diagonal]."""
    rng = np.random.default_rng(seed)
    lo, hi = scene_bounds(sc)
    ext = hi - lo
    lo2, hi2 = lo - 0.25 * ext, hi + 0.25 * ext
    org = lo2 + rng.random((n, 3)) * (hi2 - lo2)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    k = n // 8
    # axis-aligned with signed zeros
    ax = np.zeros((k, 3))
    idx = rng.integers(0, 3, k)
    ax[np.arange(k), idx] = rng.choice([-1.0, 1.0], k)
    ax[ax == 0] = np.where(rng.random(int((ax == 0).sum())) < 0.5, 0.0, -0.0)
    d[:k] = ax
    # aim at vertices and edge midpoints
    nt = sc.tris.shape[0]
    ti = rng.integers(0, nt, k)
    tv = sc.tris[ti].reshape(k, 3, 3).astype(np.float64)
    w = rng.integers(0, 3, k)
    tgt = np.where((rng.random(k) < 0.5)[:, None], tv[np.arange(k), w], 0.5 * (tv[:, 0] + tv[:, 1]))
    dd = tgt - org[k:2 * k]
    d[k:2 * k] = dd / np.linalg.norm(dd, axis=1, keepdims=True)
    # unnormalized directions
    d[2 * k:3 * k] *= rng.uniform(0.1, 10.0, (k, 1))
    # origins on triangle surfaces (barycentric points), random directions
    ti = rng.integers(0, nt, k)
    tv = sc.tris[ti].reshape(k, 3, 3).astype(np.float64)
    a, bb = rng.random(k), rng.random(k)
    m = a + bb > 1
    a[m], bb[m] = 1 - a[m], 1 - bb[m]
    org[3 * k:4 * k] = tv[:, 0] + a[:, None] * (tv[:, 1] - tv[:, 0]) + bb[:, None] * (tv[:, 2] - tv[:, 0])
    tmax = rng.random(n) * 2.0 * float(np.linalg.norm(ext))
    tmax[: n // 16] = np.inf
    return org.astype(np.float32), d.astype(np.float32), tmax.astype(np.float32)


def bsdf_queries(n, seed):
    """Random shading queries: normals, wo (some below the horizon), wi, and
    material params (8 floats: type, base rgb, roughness, metallic, ior, pad)
    covering diffuse (roughness below the 0.08 clamp, metallic 0/1/mixed),
    glass (both sides) and mirror."""
    rng = np.random.default_rng(seed)
    def unit(m):
        v = rng.normal(size=(m, 3))
        return v / np.linalg.norm(v, axis=1, keepdims=True)
    N = unit(n)
    wo = unit(n)
    wi = unit(n)
    # bias most wo/wi to the upper hemisphere
    up = rng.random(n) < 0.85
    wo[up] = np.where((np.einsum("ij,ij->i", wo, N) < 0)[:, None], -wo, wo)[up]
    up = rng.random(n) < 0.85
    wi[up] = np.where((np.einsum("ij,ij->i", wi, N) < 0)[:, None], -wi, wi)[up]
    params = np.zeros((n, 8), np.float32)
    typ = rng.choice([0, 0, 0, 0, 1, 2], n)
    params[:, 0] = typ
    params[:, 1:4] = rng.random((n, 3))
    params[:, 4] = rng.choice([0.01, 0.05, 0.08, 0.3, 0.7, 1.0], n) * np.where(rng.random(n) < 0.5, 1.0, rng.random(n))
    params[:, 5] = rng.choice([0.0, 1.0, 0.5], n) * np.where(rng.random(n) < 0.7, 1.0, rng.random(n))
    params[:, 6] = rng.choice([1.5, 1.33, 2.4, 1.0001], n)
    return (N.astype(np.float32), wo.astype(np.float32), wi.astype(np.float32), params)


def sphere_dirs(n, seed):
    rng = np.random.default_rng(seed)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    special = np.array([[0, 1, 0], [0, -1, 0], [-1, 0, 0], [-1, 0, 1e-7], [-1, 0, -1e-7],
                        [1, 0, 0], [0, 0, 1], [0, 0, -1], [0.5, 1e-8, 0.5]], np.float64)
    special /= np.linalg.norm(special, axis=1, keepdims=True)
    d[: len(special)] = special
    return d.astype(np.float32)


def attr_sphere(center, radius, segments=12, bands=8):
    """UV-sphere mesh with per-vertex normals, uvs and tangents (glTF-style
    attributes): (tris [n,9], normals [n,9], uvs [n,6], tangents [n,12])."""
    c = np.asarray(center, np.float64)
    tris, nrm, uvs, tan = [], [], [], []

    def vert(i, j):
        th, ph = math.pi * i / bands, 2 * math.pi * j / segments
        n = np.array([math.sin(th) * math.cos(ph), math.cos(th), math.sin(th) * math.sin(ph)])
        t = np.array([-math.sin(ph), 0.0, math.cos(ph)])
        return c + radius * n, n, (j / segments * 2.0, i / bands), np.append(t, 1.0 if j % 2 else -1.0)
    for i in range(bands):
        for j in range(segments):
            a, b, cc, d = vert(i, j), vert(i, j + 1), vert(i + 1, j + 1), vert(i + 1, j)
            for f in ((a, d, cc), (a, cc, b)):
                tris.append(np.concatenate([v[0] for v in f]))
                nrm.append(np.concatenate([v[1] for v in f]))
                uvs.append(np.concatenate([v[2] for v in f]))
                tan.append(np.concatenate([v[3] for v in f]))
    f32 = lambda x: np.asarray(x, np.float32)
    return f32(tris), f32(nrm), f32(uvs), f32(tan)


def textured_room(seed=17, edges=False):
    """box_room with glTF-style content: a textured floor (sRGB colour
    checker, metallic-roughness and normal maps, tangents), a smooth-normal
    GGX sphere mesh with a metallic-roughness map and a normal map without
    tangents (buildTBN frame), a smooth-normal glass sphere mesh tinted by a
    colour texture, a textured analytic sphere (uv 0), and a ceiling light
    whose per-vertex normals tilt (light-sample normal normalize(n0+n1+n2)).

    edges=True (textured_edges): the tangent frame's edge cases under the
    normal maps (Material.hpp:62-74). The floor's left half has tangents
    parallel to its normal (Gram-Schmidt leaves nothing: the buildTBN
    fallback); its right half has handedness -1 on one triangle and mixed
    signs on the other (the interpolated w changes sign inside it); the GGX
    sphere carries tangents whose handedness alternates per column."""
    rng = np.random.default_rng(seed)
    b = S.SceneBuilder()
    b.background = np.zeros(3, np.float32)
    checker = np.zeros((8, 8, 3), np.uint8)
    checker[(np.add.outer(np.arange(8), np.arange(8)) % 2) == 0] = (230, 200, 60)
    checker[(np.add.outer(np.arange(8), np.arange(8)) % 2) == 1] = (40, 90, 200)
    t_col = b.texture(checker)
    t_mr = b.texture(rng.integers(0, 256, (5, 7, 4)).astype(np.uint8))
    nm = rng.normal(scale=0.35, size=(6, 6, 3)) + np.array([0, 0, 1.0])
    nm /= np.linalg.norm(nm, axis=2, keepdims=True)
    t_nrm = b.texture(np.clip((nm * 0.5 + 0.5) * 255 + 0.5, 0, 255).astype(np.uint8))
    t_tint = b.texture(rng.integers(120, 256, (4, 4, 3)).astype(np.uint8))
    grey = b.material(S.MAT_DIFFUSE, (0.73, 0.73, 0.73), 1.0, 0.0, 1.5)
    red = b.material(S.MAT_DIFFUSE, (0.65, 0.05, 0.05), 1.0, 0.0, 1.5)
    green = b.material(S.MAT_DIFFUSE, (0.12, 0.45, 0.15), 1.0, 0.0, 1.5)
    floor_m = b.material(S.MAT_DIFFUSE, (1, 1, 1), 0.6, 0.2, 1.5, textures=(t_col, t_mr, t_nrm, -1))
    ggx_m = b.material(S.MAT_DIFFUSE, (0.95, 0.64, 0.54), 0.3, 1.0, 1.5, textures=(-1, t_mr, t_nrm, -1))
    glass_m = b.material(S.MAT_GLASS, (1, 1, 1), 0.5, 0.0, 1.5, textures=(t_tint, -1, -1, -1))
    ball_m = b.material(S.MAT_DIFFUSE, (0.5, 0.5, 0.5), 0.5, 0.0, 1.5, textures=(t_col, -1, -1, -1))
    light = b.material(S.MAT_EMISSIVE, (0.8, 0.8, 0.8), 0.5, 0.0, 1.5, (17.0, 12.0, 4.0))
    fl_n = np.tile(np.array([0, 1, 0], np.float32), (2, 3))
    if not edges:
        fl = S.quad((-1, 0, -1), (-1, 0, 1), (1, 0, 1), (1, 0, -1))
        fl_uv = np.array([[0, 0, 0, 3, 3, 3], [0, 0, 3, 3, 3, 0]], np.float32)
        fl_t = np.tile(np.array([1, 0, 0, 1], np.float32), (2, 3))
        b.mesh(fl, floor_m, normals=fl_n, uvs=fl_uv, tangents=fl_t)
    else:
        def uv_of(q):
            return np.stack([1.5 * (q[:, 0::3] + 1), 1.5 * (q[:, 2::3] + 1)], 2).reshape(-1, 6).astype(np.float32)
        left = S.quad((-1, 0, -1), (-1, 0, 1), (0, 0, 1), (0, 0, -1))
        right = S.quad((0, 0, -1), (0, 0, 1), (1, 0, 1), (1, 0, -1))
        t_par = np.tile(np.array([0, 1, 0, 1], np.float32), (2, 3))
        t_neg = np.array([[0.8, 0, 0.6, -1] * 3, [1, 0, 0, 1, 1, 0, 0, -1, 1, 0, 0, -1]], np.float32)
        b.mesh(left, floor_m, normals=fl_n, uvs=uv_of(left), tangents=t_par)
        b.mesh(right, floor_m, normals=fl_n, uvs=uv_of(right), tangents=t_neg)
    b.mesh(S.quad((-1, 2, -1), (1, 2, -1), (1, 2, 1), (-1, 2, 1)), grey)
    b.mesh(S.quad((-1, 0, -1), (1, 0, -1), (1, 2, -1), (-1, 2, -1)), grey)
    b.mesh(S.quad((-1, 0, -1), (-1, 2, -1), (-1, 2, 1), (-1, 0, 1)), red)
    b.mesh(S.quad((1, 0, -1), (1, 0, 1), (1, 2, 1), (1, 2, -1)), green)
    lq = S.quad((-0.25, 1.98, -0.25), (0.25, 1.98, -0.25), (0.25, 1.98, 0.25), (-0.25, 1.98, 0.25))
    ln = np.array([[0.2, -1, 0.1], [-0.1, -1, 0.3], [0.05, -1, -0.2]], np.float32)
    ln /= np.linalg.norm(ln, axis=1, keepdims=True)
    b.mesh(lq, light, normals=np.tile(ln.ravel(), (2, 1)))
    t, n, uv, tg = attr_sphere((-0.45, 0.4, -0.3), 0.4)
    b.mesh(t, ggx_m, normals=n, uvs=uv, tangents=tg if edges else None)
    t, n, uv, tg = attr_sphere((0.45, 0.4, 0.25), 0.4)
    b.mesh(t, glass_m, normals=n, uvs=uv, tangents=tg)
    b.sphere((0.3, 1.5, -0.5), 0.2, ball_m)
    b.set_camera((0, 1, 3.9), (0, 1, 0), (0, 1, 0), 40.0)
    return b.build("textured_edges" if edges else "textured_room")


def textured_edges(seed=17):
    return textured_room(seed, edges=True)
