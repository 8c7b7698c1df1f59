"""GPU renders against the oracle (same PCG32 streams: per-path parity) and
against the reference's own estimator image (per-pixel 3-sigma agreement and
relMSE against a converged image at equal spp)."""
import numpy as np
import pytest

from conftest import fixture_scene, load_golden
import pyoracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from photorealisticrender4610_amd import abi
    c = abi.Context(0)
    yield c
    c.close()


def _render_pair(ctx, sc, w, h, spp, depth, seed):
    scene = ctx.scene(sc)
    img, st = ctx.render(scene, w, h, spp, depth, seed=seed)
    g = img.cpu().numpy()
    ref, _, cnt = O.OracleScene(sc).render(w, h, spp, depth, seed)
    return g, ref, st, cnt


@pytest.mark.parametrize("which", ["box_room", "mini_field"])
def test_per_path_parity_with_oracle(ctx, which):
    """Same seeds, same PCG32 streams, same draw order: each pixel's path is
    the oracle's path; only ulp-level differences of sin/cos/pow can flip a
    branch (lobe choice, RR), so nearly all pixels agree to 1e-4."""
    from photorealisticrender4610_amd import scenes
    import testscenes as TS
    sc = scenes.box_room() if which == "box_room" else TS.mini_field()
    w, h, depth = (128, 128, 8) if which == "box_room" else (160, 120, 12)
    print("case", which, flush=True)
    g, ref, st, cnt = _render_pair(ctx, sc, w, h, 1, depth, seed=3)
    close = np.all(np.abs(g - ref) <= 1e-4 * np.maximum(np.abs(ref), 1e-2), axis=2)
    assert close.mean() > 0.98, close.mean()
    assert abs(g.mean() - ref.mean()) <= 0.02 * ref.mean()
    # ray counts agree with the restated reference loop (its shadow queries:
    # those traced plus those whose NEE term is exactly zero, not traced)
    assert abs(st["rays_closest"] - cnt["closest"]) <= 0.01 * cnt["closest"]
    assert abs(st["rays_any"] + st["rays_any_skipped"] - cnt["any"]) <= 0.01 * cnt["any"]
    if which == "mini_field":  # sun-lit spheres: NEE directions below the surface on their far sides
        assert st["rays_any_skipped"] > 0.05 * cnt["any"], (st["rays_any_skipped"], cnt["any"])
    assert st["samples"] == w * h


def test_multi_sample_accumulation_order(ctx):
    """spp > 1 with several samples of a pixel in one chunk: the per-pixel
    mean is summed in sample order like Renderer.cpp:84-99."""
    from photorealisticrender4610_amd import scenes
    sc = scenes.box_room()
    g, ref, st, _ = _render_pair(ctx, sc, 48, 40, 6, 8, seed=11)
    close = np.all(np.abs(g - ref) <= 1e-4 * np.maximum(np.abs(ref), 1e-2), axis=2)
    assert close.mean() > 0.95, close.mean()
    assert st["samples"] == 48 * 40 * 6


def _estimator_vs_reference(ctx, sc, ref, w, h, spp, depth, label):
    """The GPU estimator against the reference's own image (ref: its
    spp-sample mean per pixel, [h, w, 3]).

    64 independent GPU renders at the same spp give, per pixel, the
    distribution of the spp-sample mean. Under the null hypothesis (the same
    estimator) the reference's image is one more draw from it, exchangeable
    with the 64, whatever the tails (caustics and the sun make them heavy,
    so a Gaussian 3-sigma count would reject the GPU against itself):
    * per pixel, the rank of the reference's luminance among the 65 values is
      uniform: it falls outside the 64 GPU values with probability 2/65, and
      the rank histogram passes a chi-square test (13 bins, p = 1e-4);
    * the image means agree to 0.5%;
    * the reference's relMSE against a 16384-spp GPU image is within 5% of
      the mean relMSE of the GPU renders at equal spp.
    The Gaussian 3-sigma exceedance (reference vs GPU, and GPU vs GPU as its
    null rate) is printed for information."""
    from scipy import stats
    scene = ctx.scene(sc)
    runs = np.stack([ctx.render(scene, w, h, spp, depth, seed=1000 + k)[0].cpu().numpy() for k in range(64)])
    runs = runs.astype(np.float64)
    ref = ref.astype(np.float64)
    gpu_mean = runs.mean(0)
    lum = lambda x: x @ np.array([0.2126, 0.7152, 0.0722])  # noqa: E731
    rl, gl = lum(ref), lum(runs)
    less = (gl < rl[None]).sum(0)
    eq = (gl == rl[None]).sum(0)
    rng = np.random.default_rng(7)
    rank = (less + rng.integers(0, eq + 1)).ravel()  # ties broken at random
    outside = np.mean((rank == 0) | (rank == 64))
    p = 2.0 / 65.0
    npx = rank.size
    hist = np.bincount(rank // 5, minlength=13)
    chi2 = ((hist - npx / 13.0) ** 2 / (npx / 13.0)).sum()
    sig = runs.std(0, ddof=1)
    z = np.abs(runs[0] - ref) / np.maximum(np.sqrt(2.0) * sig, 1e-12)
    zg = np.abs(runs[1] - runs[0]) / np.maximum(np.sqrt(2.0) * sig, 1e-12)
    print("%s: outside the 64 GPU values %.4f (null %.4f); rank chi2 %.1f; 3-sigma ref %.4f gpu-vs-gpu %.4f; "
          "mean ref %.5f gpu %.5f" % (label, outside, p, chi2, np.mean(z > 3), np.mean(zg > 3), ref.mean(),
                                      gpu_mean.mean()), flush=True)
    assert outside <= p + 4.0 * np.sqrt(p * (1 - p) / npx), outside
    assert chi2 < stats.chi2.ppf(1 - 1e-4, 12), (chi2, hist)
    assert abs(ref.mean() - gpu_mean.mean()) <= 0.005 * gpu_mean.mean()
    conv = np.zeros_like(gpu_mean)
    for k in range(16):  # 16384 spp in 16 renders
        conv += ctx.render(scene, w, h, 1024, depth, seed=99, sample_offset=(1 << 20) + 1024 * k)[0].cpu().numpy()
    conv /= 16

    def relmse(x):
        return np.mean(np.sum((x - conv) ** 2 / (conv ** 2 + 1e-2), axis=2))

    r_gpu = np.array([relmse(r) for r in runs])
    r_ref = relmse(ref)
    print("%s: relMSE gpu mean %.5f std %.5f ref %.5f" % (label, r_gpu.mean(), r_gpu.std(), r_ref))
    assert abs(r_gpu.mean() - r_ref) <= 0.05 * r_gpu.mean(), (r_gpu.mean(), r_ref)
    assert abs(conv.mean() - gpu_mean.mean()) <= 0.005 * conv.mean()


def test_estimator_vs_reference_image(ctx):
    """The box room (64x64, 256 spp, depth 8) against the compiled
    reference's own image."""
    g = load_golden("image_box_room.npz")
    w, h, depth, spp = (int(x) for x in g["cfg"])
    _estimator_vs_reference(ctx, fixture_scene(g), g["ref_mean"].reshape(h, w, 3), w, h, spp, depth, "box_room")


def test_estimator_vs_reference_image_sphere_field(ctx):
    """A 512-sphere instance of the 1M-triangle workload's scene with its
    full-size sky (2048 x 1024, 2-degree sun of peak 1e4), 96x64 at 256 spp,
    depth 12, against the reference's own image of it. The fixture holds the
    image and a digest of the scene, which is regenerated from its seed."""
    import hashlib
    from photorealisticrender4610_amd import scenes
    g = load_golden("image_sphere_field.npz")
    w, h, depth, spp = (int(x) for x in g["cfg"])
    sc = scenes.sphere_field(num_spheres=int(g["num_spheres"][0]), env_size=tuple(int(x) for x in g["env_size"]))
    hsh = hashlib.sha256()
    for k, v in sorted(sc.bundle_arrays().items()):
        hsh.update(k.encode())
        hsh.update(np.ascontiguousarray(v).tobytes())
    assert np.array_equal(np.frombuffer(hsh.digest(), np.uint8), g["digest"]), "scene generator changed"
    _estimator_vs_reference(ctx, sc, g["ref_mean"].reshape(h, w, 3), w, h, spp, depth, "sphere_field")


def test_shadow_trees_agree(ctx, monkeypatch):
    """Shadow rays traced on the quantised four-wide occlusion tree (the
    default), on its full-precision copy and on the binary tree give the same
    image bit for bit, and batch any-hit queries agree with the oracle on
    each: an any-hit answer does not depend on the tree, the occlusion tree
    skips only boxes that contain their children's boxes, and the quantised
    boxes contain the exact ones with the exact leaf test kept. The same holds
    with the short form of the quantised node test switched off
    (PRR_SHADOW_FAST=0: every ray on the full decode)."""
    from photorealisticrender4610_amd import abi
    import testscenes as TS
    sc = TS.mini_field()
    print("case shadow trees", flush=True)
    scene = ctx.scene(sc)
    info = scene.info()
    print("occlusion tree: %d nodes, %d levels kept (binary tree: %d nodes)"
          % (info["num_wide_nodes"], info["wide_levels_kept"], info["num_nodes"]), flush=True)
    assert 0 < info["num_wide_nodes"] < info["num_nodes"]
    img_q, st_q = ctx.render(scene, 160, 120, 4, 12, seed=5)
    img_q = img_q.cpu().numpy()
    rng = np.random.default_rng(12)
    n = 50000
    b = np.asarray(info["bounds"], np.float32)
    org = (b[:3] + rng.random((n, 3)) * (b[3:] - b[:3])).astype(np.float32)
    d = rng.normal(size=(n, 3)).astype(np.float32)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    tmax = (rng.random(n) * np.linalg.norm(b[3:] - b[:3])).astype(np.float32)
    want = O.OracleScene(sc).any(org, d, tmax)
    assert np.array_equal(ctx.trace_any(scene, org, d, tmax).cpu().numpy(), want)
    for var, mode in (("PRR_SHADOW_WIDE", "1"), ("PRR_SHADOW_WIDE", "0"), ("PRR_SHADOW_FAST", "0")):
        monkeypatch.delenv("PRR_SHADOW_WIDE", raising=False)
        monkeypatch.setenv(var, mode)
        c2 = abi.Context(0)
        try:
            s2 = c2.scene(sc)
            img, st = c2.render(s2, 160, 120, 4, 12, seed=5)
            img = img.cpu().numpy()
            anyh = c2.trace_any(s2, org, d, tmax).cpu().numpy()
            del s2
        finally:
            c2.close()
        assert np.array_equal(img_q, img), (var, mode)
        assert st_q["rays_any"] == st["rays_any"] and st_q["rays_closest"] == st["rays_closest"]
        assert np.array_equal(anyh, want), (var, mode)


@pytest.mark.parametrize("var", ["PRR_SPECULATE", "PRR_TWO_LEVEL"])
def test_extension_ray_iterations_agree(ctx, monkeypatch, var):
    """Extension rays traced with speculative while-while iteration
    (PRR_SPECULATE=1: a lane that reaches a leaf sets it aside and steps on to
    a second one, each leaf re-checked against the closest hit when its turn
    comes) or on the two-level records (PRR_TWO_LEVEL=1: two binary levels per
    128-byte line, the children's boxes rebuilt as unions of their pairs) give
    the same images and ray counts bit for bit as the plain while-while
    iteration, on scenes with triangles and spheres, glass and mirror chains,
    and the full 1M-triangle scene."""
    from photorealisticrender4610_amd import abi, scenes
    import testscenes as TS
    print("case extension-ray iteration %s" % var, flush=True)
    cases = [(scenes.box_room(), 128, 96, 8, 8), (TS.mini_field(), 160, 120, 4, 12),
             (scenes.sphere_field(), 320, 180, 4, 12), (scenes.glass_torus(), 128, 128, 4, 12)]
    # every bounce in the per-bounce kernels (no path-completion launch), so
    # all extension rays go through kExtend
    monkeypatch.setenv("PRR_TAIL_PATHS", "0")
    c1 = abi.Context(0)
    want = []
    try:
        for sc, w, h, spp, depth in cases:
            scene = c1.scene(sc)
            assert not scene.info()["two_level"]  # uploaded only for a context that reads them
            img, st = c1.render(scene, w, h, spp, depth, seed=21)
            want.append((img.cpu().numpy(), st["rays_closest"], st["rays_any"]))
            del scene
    finally:
        c1.close()
    monkeypatch.setenv(var, "1")
    c2 = abi.Context(0)
    try:
        for (sc, w, h, spp, depth), (img0, rc, ra) in zip(cases, want):
            s2 = c2.scene(sc)
            if var == "PRR_TWO_LEVEL":
                assert s2.info()["two_level"], sc.name
            img, st = c2.render(s2, w, h, spp, depth, seed=21)
            assert np.array_equal(img.cpu().numpy(), img0), sc.name
            assert st["rays_closest"] == rc and st["rays_any"] == ra, sc.name
            del s2
    finally:
        c2.close()


def test_shadow_short_form_boundaries(ctx):
    """The short form of the quantised node test (fma of a plane byte with
    scale / d and (origin - o) / d, against boxes widened by a margin) at its
    boundaries, against the oracle's IntersectP: origins on leaf-box planes
    (triangle vertices) with axis directions, which make the reference's
    1/d infinite and some of its plane times NaN; direction components of
    1e-35 (1/d clamped to +-2^100); origins far outside the region the margin
    covers and directions scaled by 1e-6 or 1e6 (those rays keep the full
    decode); ranges of 0, the hit distance and infinity."""
    import testscenes as TS
    print("case shadow short form", flush=True)
    sc = TS.mini_field()
    scene = ctx.scene(sc)
    orc = O.OracleScene(sc)
    info = scene.info()
    b = np.asarray(info["bounds"], np.float32)
    ext = float(np.max(b[3:] - b[:3]))
    rng = np.random.default_rng(41)
    tris = np.asarray(sc.tris, np.float32).reshape(-1, 3, 3)
    verts = tris[rng.integers(0, len(tris), 3000), rng.integers(0, 3, 3000)]
    axes = np.array([(1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1),
                     (0.70710677, 0, -0.70710677), (0, 0.70710677, 0.70710677)], np.float32)
    o1 = np.repeat(verts, len(axes), 0)
    d1 = np.tile(axes, (len(verts), 1))
    n2 = 20000
    o2 = (b[:3] + rng.random((n2, 3)) * (b[3:] - b[:3])).astype(np.float32)
    d2 = rng.normal(size=(n2, 3)).astype(np.float32)
    d2 /= np.linalg.norm(d2, axis=1, keepdims=True)
    d2[: n2 // 4, rng.integers(0, 3)] = 1e-35
    d2[n2 // 4: n2 // 2] *= np.float32(1e-6)
    d2[n2 // 2: 3 * n2 // 4] *= np.float32(1e6)
    far = (b[:3] + b[3:]) / 2 + rng.normal(size=(n2 // 4, 3)).astype(np.float32) * 5 * ext
    o2[3 * n2 // 4:] = far
    tgt = b[:3] + rng.random((n2 // 4, 3)) * (b[3:] - b[:3])
    d2[3 * n2 // 4:] = (tgt - far) / np.linalg.norm(tgt - far, axis=1, keepdims=True)
    org = np.concatenate([o1, o2]).astype(np.float32)
    d = np.concatenate([d1, d2]).astype(np.float32)
    n = len(org)
    prim, t, _, _ = (x.cpu().numpy() for x in ctx.trace_closest(scene, org, d))
    for tm in (np.zeros(n), t, np.full(n, np.inf), rng.random(n) * 2 * ext):
        tm = np.asarray(tm, np.float32)
        got = ctx.trace_any(scene, org, d, tm).cpu().numpy()
        assert np.array_equal(got, orc.any(org, d, tm))
    assert (prim >= 0).mean() > 0.3


def test_render_deterministic(ctx):
    from photorealisticrender4610_amd import scenes
    scene = ctx.scene(scenes.box_room())
    a, _ = ctx.render(scene, 64, 64, 8, 8, seed=7)
    b, _ = ctx.render(scene, 64, 64, 8, 8, seed=7)
    assert np.array_equal(a.cpu().numpy(), b.cpu().numpy())


def test_cast_rays_vs_oracle(ctx):
    """prr_cast_rays (Integrator::castRay on caller rays) against the oracle's
    castRay on the same PCG32 streams."""
    import testscenes as TS
    sc = TS.mini_field()
    scene = ctx.scene(sc)
    rng = np.random.default_rng(8)
    n = 20000
    org = np.tile(np.array([[-6.0, 6.0, -8.0]], np.float32), (n, 1)) + rng.normal(scale=0.5, size=(n, 3)).astype(np.float32)
    d = (np.array([[5.0, 2.0, 5.0]]) - org + rng.normal(scale=3.0, size=(n, 3))).astype(np.float32)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    g = ctx.cast_rays(scene, org, d, 12, seed=21).cpu().numpy()
    o = O.OracleScene(sc).castray_pcg(org, d, 21, 12)
    close = np.all(np.abs(g - o) <= 1e-4 * np.maximum(np.abs(o), 1e-2), axis=1)
    assert close.mean() > 0.98, close.mean()
    assert (o.sum(1) > 0).mean() > 0.5


def _per_sample(ctx, scene, w, h, n, depth, seed):
    """Sample k of every pixel alone: a 1-spp render with sample_offset k uses
    the PCG32 stream of (seed, pixel, k), the same path the adaptive render
    takes for that sample."""
    out = np.zeros((h * w, n, 3), np.float32)
    for k in range(n):
        img, _ = ctx.render(scene, w, h, 1, depth, seed=seed, sample_offset=k)
        out[:, k] = img.cpu().numpy().reshape(-1, 3)
    return out


def test_adaptive_sampling_follows_stop_rule(ctx):
    """Renderer::Render's adaptive sampler on the device: each pixel's sample
    count and mean equal the Welford stop rule (oracle, pinned to the
    reference's Welford) applied to that pixel's own per-sample radiance, and
    the later passes over a shrinking pixel list change nothing else."""
    from photorealisticrender4610_amd import scenes
    sc = scenes.box_room()
    scene = ctx.scene(sc)
    w, h, spp, depth = 40, 32, 24, 8
    print("case adaptive box_room 40x32 spp 24 min 3 thr 0.15", flush=True)
    img, st = ctx.render(scene, w, h, spp, depth, seed=5, min_spp=3, variance_threshold=0.15, counts=True)
    g = img.cpu().numpy().reshape(-1, 3)
    n = st["spp"].cpu().numpy().ravel()
    seq = _per_sample(ctx, scene, w, h, spp, depth, 5)
    en, em = O.welford(seq, 3, 0.15)
    assert np.array_equal(n, en)
    assert np.array_equal(g, em)
    assert 3 <= n.min() and n.max() == spp and (n < spp).mean() > 0.1
    assert st["samples"] >= n.sum()


def test_adaptive_threshold_zero_is_fixed_spp(ctx):
    from photorealisticrender4610_amd import scenes
    sc = scenes.box_room()
    scene = ctx.scene(sc)
    a, _ = ctx.render(scene, 32, 24, 8, 8, seed=9)
    b, st = ctx.render(scene, 32, 24, 8, 8, seed=9, min_spp=2, variance_threshold=0.0, counts=True)
    assert np.array_equal(a.cpu().numpy(), b.cpu().numpy())
    assert np.all(st["spp"].cpu().numpy() == 8)


def test_adaptive_vs_oracle(ctx):
    """The oracle's adaptive render on the same streams: counts agree except
    where an ulp-level radiance difference moves a CI across the threshold."""
    from photorealisticrender4610_amd import scenes
    sc = scenes.box_room()
    scene = ctx.scene(sc)
    w, h = 48, 40
    img, st = ctx.render(scene, w, h, 32, 8, seed=13, min_spp=4, variance_threshold=0.1, counts=True)
    ref, _, ost = O.OracleScene(sc).render(w, h, 32, 8, 13, min_spp=4, variance_threshold=0.1, want_spp=True)
    n, on = st["spp"].cpu().numpy(), ost["spp"]
    same = n == on
    assert same.mean() > 0.97, same.mean()
    g = img.cpu().numpy()
    close = np.all(np.abs(g - ref) <= 1e-4 * np.maximum(np.abs(ref), 1e-2), axis=2)
    assert close[same].mean() > 0.95


def test_shadow_binning_keeps_images(ctx, monkeypatch):
    """Shadow rays binned by origin cell and direction octant before
    traversal (PRR_SHADOW_SORT=b: from bounce b on) give the image of
    unsorted traversal bit for bit: any-hit answers do not depend on trace
    order."""
    from photorealisticrender4610_amd import abi
    import testscenes as TS
    sc = TS.mini_field()
    print("case shadow binning", flush=True)
    scene = ctx.scene(sc)
    # large enough that bounces 1 and 2 run in the per-bounce kernels
    img, st = ctx.render(scene, 320, 240, 16, 12, seed=5, kernel_timing=True)
    assert st["ms_sort"] == 0  # off by default
    for mode in ("1", "2"):
        monkeypatch.setenv("PRR_SHADOW_SORT", mode)
        c2 = abi.Context(0)
        try:
            s2 = c2.scene(sc)
            img2, st2 = c2.render(s2, 320, 240, 16, 12, seed=5, kernel_timing=True)
            img2 = img2.cpu().numpy()
            del s2
        finally:
            c2.close()
        assert np.array_equal(img.cpu().numpy(), img2), mode
        assert st2["rays_any"] == st["rays_any"]
        if mode == "1":
            assert st2["ms_sort"] > 0


@pytest.mark.parametrize("which", ["mini_field", "textured_room", "glass_torus"])
def test_material_sorted_shading_keeps_images(ctx, monkeypatch, which):
    """Shading sorted by material (PRR_SHADE_SORT=1: each bounce's queue split
    into diffuse-hit entries and the rest, two specialised kernels) gives the
    image of one shading kernel over the queue in order, bit for bit, with
    the same ray counts; also with the caustic photon gather."""
    from photorealisticrender4610_amd import abi, scenes
    import testscenes as TS
    sc = {"mini_field": TS.mini_field, "textured_room": TS.textured_room,
          "glass_torus": lambda: scenes.glass_torus(nu=64, nv=32)}[which]()
    photons = which == "glass_torus"
    print("case material-sorted shading %s" % which, flush=True)
    imgs = []
    for mode in ("1", "0"):
        monkeypatch.setenv("PRR_SHADE_SORT", mode)
        c2 = abi.Context(0)
        try:
            s2 = c2.scene(sc)
            if photons:
                c2.photon_build(s2, 100000, 0.05, 30, 9)
            img, st = c2.render(s2, 320, 240, 16, 12, seed=5, photons=photons)
            imgs.append((img.cpu().numpy(), st["rays_closest"], st["rays_any"]))
            del s2
        finally:
            c2.close()
    assert np.array_equal(imgs[0][0], imgs[1][0])
    assert imgs[0][1:] == imgs[1][1:]
