"""The CPU oracle against the reference's own outputs (golden fixtures written
by tests/golden/make_golden.py from the compiled reference) and against
brute force. These pin the oracle before it is trusted as the GPU checker."""
import numpy as np
import pytest

from conftest import fixture_scene, load_golden
import pyoracle as O

TRACE_SETS = ["trace_box_room.npz", "trace_mini_field.npz", "trace_soup.npz"]


def _rel(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return np.abs(a - b) / np.maximum(np.abs(b), 1e-30)


@pytest.mark.parametrize("name", TRACE_SETS)
def test_trace_matches_reference(name):
    g = load_golden(name)
    sc = O.OracleScene(fixture_scene(g))
    prim, t, u, v = sc.closest(g["org"], g["dir"])
    assert np.array_equal(prim, g["ref_prim"]), "closest-hit ids differ from BVH::Intersect"
    hit = prim >= 0
    # same arithmetic in the same order: bit-exact
    assert np.array_equal(t[hit], g["ref_t"][hit])
    assert np.array_equal(u[hit], g["ref_u"][hit])
    assert np.array_equal(v[hit], g["ref_v"][hit])
    anyh = sc.any(g["org"], g["dir"], g["tmax"])
    assert np.array_equal(anyh, g["ref_any"]), "any-hit differs from BVH::IntersectP"
    assert sc.total_emit_area() == float(g["ref_total_emit_area"][0])


@pytest.mark.parametrize("name", TRACE_SETS)
def test_trace_matches_brute_force(name):
    g = load_golden(name)
    sc = O.OracleScene(fixture_scene(g))
    prim, t, _, _ = sc.closest(g["org"], g["dir"])
    bp, bt = sc.brute_closest(g["org"], g["dir"])
    # Rays k..2k of the set are aimed exactly at triangle vertices / edge
    # midpoints: there the reference's slab test (BVH.hpp:115-124) can round a
    # zero-thickness box closed and miss a hit brute force finds. Every other
    # ray must agree with the tree-independent minimum exactly.
    n = t.shape[0]
    k = n // 8
    grazing = np.zeros(n, bool)
    grazing[k:2 * k] = True
    assert np.array_equal(t[~grazing], bt[~grazing]), "closest t must equal the minimum over all primitives"
    assert np.mean(t[grazing] == bt[grazing]) > 0.85
    # ids equal except exact ties (same t from another primitive)
    diff = (prim != bp) & ~grazing
    assert np.all(t[diff] == bt[diff])
    anyh = sc.any(g["org"], g["dir"], g["tmax"])
    ba = sc.brute_any(g["org"], g["dir"], g["tmax"])
    assert np.array_equal(anyh[~grazing], ba[~grazing])
    assert (prim >= 0).mean() > 0.2 and anyh.mean() > 0.05  # the ray set actually hits things


@pytest.mark.parametrize("name", TRACE_SETS)
def test_sah_tree_matches_reference(name):
    g = load_golden(name)
    fs = fixture_scene(g)
    sc = O.OracleScene(fs)
    levels = [-1] + [k for k in range(fs.objects.shape[0]) if fs.objects[k, 0] == 0]
    for lv in levels:
        pfx = "ref_top" if lv < 0 else "ref_mesh%d" % lv
        b, fc, meta = sc.export_level(lv)
        assert np.array_equal(b.ravel(), g[pfx + "_bounds"]), pfx
        assert np.array_equal(fc, g[pfx + "_first"]), pfx
        assert np.array_equal(meta, g[pfx + "_meta"]), pfx
        order, areas = sc.export_order(lv)
        assert np.array_equal(order, g[pfx + "_order"]), pfx
        assert np.array_equal(areas, g[pfx + "_areas"]), pfx


@pytest.mark.parametrize("name", TRACE_SETS)
def test_light_sampling_matches_reference(name):
    g = load_golden(name)
    sc = O.OracleScene(fixture_scene(g))
    pos, nrm, pdf = sc.sample_light(g["light_seeds"])
    assert np.array_equal(pos.ravel(), g["ref_light_pos"])
    assert np.array_equal(nrm.ravel(), g["ref_light_nrm"])
    assert np.array_equal(pdf, g["ref_light_pdf"])


def test_light_pick_distribution_q4():
    """Scene::sampleLight picks a triangle of a 2-triangle quad light by
    p = sqrt(u) * A walked over the leaf (BVH.cpp:478-514): P(tri 0) =
    P(sqrt(u) * A <= A/2) = 1/4 for equal halves."""
    g = load_golden("trace_box_room.npz")
    sc = O.OracleScene(fixture_scene(g))
    seeds = np.arange(1, 20001, dtype=np.uint32)
    pos, nrm, pdf = sc.sample_light(seeds)
    # light quad: tri 0 = (p0,p1,p2) covers x >= z side, tri 1 the other
    x, z = pos[:, 0], pos[:, 2]
    frac0 = np.mean(x - z > 0)
    assert abs(frac0 - 0.25) < 0.015
    assert np.allclose(pdf, 1.0 / 0.25)


def test_splitmix_matches_reference_draws():
    g = load_golden("bsdf.npz")
    for i in range(50):
        d = O.splitmix_draws(int(g["seeds"][i]), 3)
        assert np.array_equal(d, g["ref_draws"].reshape(-1, 3)[i])


def test_bsdf_matches_reference():
    g = load_golden("bsdf.npz")
    n = g["seeds"].shape[0]
    f, pdf = O.bsdf_eval(g["wi"], g["wo"], g["n"], g["params"])
    assert np.array_equal(f.ravel(), g["ref_eval_f"])
    assert np.array_equal(pdf, g["ref_eval_pdf"])
    u = g["ref_draws"].reshape(n, 3)
    wi, sf, spdf, lobe = O.bsdf_sample(g["wo"], g["n"], g["params"], u)
    assert np.array_equal(wi.ravel(), g["ref_sample_wi"])
    assert np.array_equal(sf.ravel(), g["ref_sample_f"])
    assert np.array_equal(spdf, g["ref_sample_pdf"])
    ok = spdf > 0
    assert np.array_equal(lobe[ok], g["ref_sample_lobe"][ok])
    assert ok.mean() > 0.5


def test_env_matches_reference():
    g = load_golden("env.npz")
    for name in ("sky", "blocky"):
        px = g[name + "_pixels"]
        cdf, marg, tot = O.env_build(px)
        assert np.array_equal(cdf, g[name + "_ref_cdf"])
        assert np.array_equal(marg, g[name + "_ref_marg"])
        assert tot == float(g[name + "_ref_total"][0])
        u2 = g[name + "_ref_draws"].reshape(-1, 2)
        d, pdf = O.env_sample(px, u2)
        assert np.array_equal(d.ravel(), g[name + "_ref_is_dir"])
        assert np.array_equal(pdf, g[name + "_ref_is_pdf"])
        rad, p2 = O.env_eval(px, g[name + "_dirs"])
        assert np.array_equal(rad.ravel(), g[name + "_ref_rad"])
        assert np.array_equal(p2, g[name + "_ref_pdf"])


def test_env_sample_pdf_consistency():
    """importanceSample's pdf equals importanceSamplePdf at the returned
    texel-centre direction (EnvMap.hpp:124-169), up to the sin(theta) factor
    that cancels analytically."""
    g = load_golden("env.npz")
    px = g["sky_pixels"]
    rng = np.random.default_rng(3)
    u2 = rng.random((4000, 2)).astype(np.float32)
    d, pdf = O.env_sample(px, u2)
    _, p2 = O.env_eval(px, d)
    assert np.max(_rel(p2, pdf)) < 1e-4
    # the sampled texel distribution follows lum * sin(theta)
    h, w, _ = px.shape
    lum = px @ np.array([0.2126, 0.7152, 0.0722], np.float32)
    st = np.sin((np.arange(h) + 0.5) / h * np.pi)[:, None]
    p = (lum * st).ravel()
    p /= p.sum()
    cdf, marg, tot = O.env_build(px)
    assert abs(tot - float((lum * st).sum(dtype=np.float64))) / tot < 1e-4


def test_camera_matches_reference():
    g = load_golden("camera.npz")
    for name in ("box", "oblique"):
        w, h = (int(x) for x in g[name + "_wh"])
        o, d = O.camera_rays(g[name + "_camera"], w, h, g[name + "_pix"], g[name + "_offs"])
        assert np.array_equal(o.ravel(), g[name + "_ref_org"])
        assert np.array_equal(d.ravel(), g[name + "_ref_dir"])


def lens_offsets(g):
    """Pixel offsets plus the two lens uniforms each reference ray drew from
    its seeded generator (Camera.hpp:87-88)."""
    uv = np.stack([O.splitmix_draws(int(s), 2) for s in g["lens_seeds"]])
    return np.concatenate([g["lens_offs"], uv], 1).astype(np.float32)


def test_thin_lens_camera_matches_reference():
    """Camera::generateRay with aperture > 0 (concentric disc sample on the
    lens, ray through the focus plane; Camera.hpp:58-94) bit for bit."""
    g = load_golden("camera.npz")
    w, h = (int(x) for x in g["lens_wh"])
    o, d = O.camera_rays(g["lens_camera"], w, h, g["lens_pix"], lens_offsets(g))
    assert np.array_equal(o.ravel(), g["lens_ref_org"])
    assert np.array_equal(d.ravel(), g["lens_ref_dir"])
    # the lens spreads the origins over the aperture
    spread = np.linalg.norm(o - np.array([0, 1, 3.9], np.float32), axis=1)
    assert spread.max() <= 0.08 * 1.0001 and spread.mean() > 0.03


def test_tonemap_matches_reference():
    g = load_golden("tonemap.npz")
    for op, name in enumerate(["raw", "standard", "agx", "pbrneutral"]):
        disp, _ = O.tonemap(op, g["linear"], 1.0)
        assert np.array_equal(disp.ravel(), g["ref_" + name]), name


def test_tonemap_closed_forms():
    # AgX curve fixes the crossover; PBR Neutral below compression is
    # c - offset(min c) then the OETF; ACES against its closed form.
    lin = np.array([[0.3, 0.2, 0.1], [0.5, 0.4, 0.7]], np.float32)
    disp, _ = O.tonemap(3, lin)
    x = lin.min(1, keepdims=True)
    off = np.where(x < 0.08, x - 6.25 * x * x, 0.04)
    c = np.clip(lin - off, 0, 1)
    oetf = np.where(c < 0.0031308, 12.92 * c, 1.055 * np.power(c, 1 / 2.4) - 0.055)
    assert np.allclose(disp, oetf, atol=2e-6)
    v = np.linspace(0, 20, 200, dtype=np.float32)
    lin = np.stack([v, v, v], 1)
    disp, _ = O.tonemap(4, lin)
    a = np.clip(v * (2.51 * v + 0.03) / (v * (2.43 * v + 0.59) + 0.14), 0, 1)
    oetf = np.where(a < 0.0031308, 12.92 * a, 1.055 * np.power(a, 1 / 2.4) - 0.055)
    assert np.allclose(disp[:, 0], oetf, atol=2e-6)


@pytest.mark.parametrize("name", ["castray_box_room.npz", "castray_mini_field.npz", "castray_textured_room.npz",
                                  "castray_textured_edges.npz"])
def test_castray_matches_reference(name):
    """Full per-path parity with the reference castRay on identical
    SplitMix64 streams: same draws consumed in the same order."""
    g = load_golden(name)
    sc = O.OracleScene(fixture_scene(g))
    w, h, depth = (int(x) for x in g["cfg"])
    L = sc.castray_splitmix(w, h, g["pix"], g["seeds"], depth)
    ref = g["ref_L"].reshape(-1, 3)
    exact = np.all(L == ref, axis=1)
    close = np.all(np.abs(L - ref) <= 1e-5 * np.maximum(np.abs(ref), 1e-3), axis=1)
    assert exact.mean() > 0.995, exact.mean()
    assert close.mean() > 0.999, close.mean()
    assert (ref.sum(1) > 0).mean() > 0.3


def test_castray_textured_features_matter():
    """The textured fixture exercises what it is for: the same paths without
    the textures, or without the per-vertex attributes, leave the
    reference's radiance on many paths."""
    g = load_golden("castray_textured_room.npz")
    w, h, depth = (int(x) for x in g["cfg"])
    ref = g["ref_L"].reshape(-1, 3)
    for strip in ("mat_tex", "tri_flags"):
        sc = fixture_scene(g)
        setattr(sc, strip, None)
        L = O.OracleScene(sc).castray_splitmix(w, h, g["pix"], g["seeds"], depth)
        same = np.all(np.abs(L - ref) <= 1e-5 * np.maximum(np.abs(ref), 1e-3), axis=1)
        assert same.mean() < 0.9, (strip, same.mean())


def test_castray_tangent_edges_matter():
    """The tangent edge-case fixture reaches its cases: paths meet the floor
    half whose tangents are parallel to its normal (the buildTBN fallback,
    Material.hpp:67-70) and the tangents of handedness -1 and of mixed sign
    (Triangle.cpp:52-54); flipping every handedness, or giving the
    degenerate half a usable tangent, moves the reference's radiance on many
    paths, so the oracle's treatment of each is pinned by the fixture."""
    g = load_golden("castray_textured_edges.npz")
    w, h, depth = (int(x) for x in g["cfg"])
    ref = g["ref_L"].reshape(-1, 3)
    L = O.OracleScene(fixture_scene(g)).castray_splitmix(w, h, g["pix"], g["seeds"], depth)
    base = np.all(np.abs(L - ref) <= 1e-5 * np.maximum(np.abs(ref), 1e-3), axis=1)
    for change in ("flip", "tangent"):
        sc = fixture_scene(g)
        tt = sc.tri_tangents.reshape(-1, 3, 4).copy()
        if change == "flip":
            tt[..., 3] = -tt[..., 3]
        else:
            par = np.all(np.abs(tt[..., :3] - np.array([0, 1, 0], np.float32)) == 0, axis=2).all(1)
            assert par.sum() == 2
            tt[par, :, :3] = np.array([1, 0, 0], np.float32)
        sc.tri_tangents = tt.reshape(sc.tri_tangents.shape)
        L2 = O.OracleScene(sc).castray_splitmix(w, h, g["pix"], g["seeds"], depth)
        same = np.all(np.abs(L2 - ref) <= 1e-5 * np.maximum(np.abs(ref), 1e-3), axis=1)
        assert same.mean() < base.mean() - 0.02, (change, same.mean(), base.mean())


def test_welford_stop_rule_matches_reference():
    """Adaptive sampling (Renderer.cpp:71-99 with Welford, Renderer.hpp:14-40):
    samples taken and the mean, bit-exact against the reference's own Welford
    on the same sequences."""
    g = load_golden("welford.npz")
    n, mean = O.welford(g["seq"], g["min_spp"], g["thr"])
    assert np.array_equal(n, g["ref_n"])
    assert np.array_equal(mean, g["ref_mean"])
    ln = g["seq"].shape[1]
    # the fixture reaches the branches that matter: early stops, stops exactly
    # at min_spp, and sequences that never stop
    assert (g["ref_n"] < ln).mean() > 0.2 and (g["ref_n"] == ln).mean() > 0.2
    assert np.any((g["ref_n"] == g["min_spp"]) & (g["min_spp"] >= 2))


def test_welford_closed_forms():
    """Zero variance stops at max(min_spp, 2); |mean| < 1e-6 (and n < 2) gives
    CI 1, so it stops only for thresholds above 1, at the first sample;
    threshold <= 0 never stops."""
    ln = 32
    const = np.full((4, ln, 3), 0.5, np.float32)
    n, m = O.welford(const, np.array([0, 1, 2, 9], np.int32), 0.05)
    assert n.tolist() == [2, 2, 2, 9]
    assert np.all(m == 0.5)
    zero = np.zeros((2, ln, 3), np.float32)
    n, _ = O.welford(zero, 0, np.array([0.5, 1.5], np.float32))
    assert n.tolist() == [ln, 1]
    rng = np.random.default_rng(3)
    noisy = rng.random((3, ln, 3)).astype(np.float32)
    n, m = O.welford(noisy, 0, 0.0)
    assert np.all(n == ln)
    np.testing.assert_allclose(m, noisy.mean(1), rtol=1e-5)
    # the CI rule itself, in float64 on the stopping prefix: the last sample
    # is the first whose CIs are all below the threshold
    seq = rng.lognormal(0.0, 0.5, (200, 64, 3)).astype(np.float32)
    n, _ = O.welford(seq, 4, 0.2)
    for q in range(200):
        def ci(k):
            x = seq[q, :k].astype(np.float64)
            return 1.96 * np.sqrt(x.var(0, ddof=1) / k) / np.abs(x.mean(0))
        k = int(n[q])
        if k < 64:
            assert np.all(ci(k) < 0.2 + 1e-4)
        if k > 4:
            assert not np.all(ci(k - 1) < 0.2 - 1e-4)


def test_media_match_reference():
    """Medium::intersect, phaseHG and samplePhaseHG (Volume.hpp:17-44,
    Volume.cpp:3-24) against the reference's own outputs: the box test bit for
    bit, NaN entry / exit times included (rays on a box plane and parallel to
    it), the phase function and the phase samples at the reference's draws."""
    g = load_golden("volume.npz")
    te, tx, hit = O.medium_intersect(g["org"], g["dir"], g["bounds"], g["bounded"])
    assert np.array_equal(hit, g["ref_hit"].astype(np.uint8))
    assert np.array_equal(te, g["ref_t_enter"], equal_nan=True)
    assert np.array_equal(tx, g["ref_t_exit"], equal_nan=True)
    assert np.isnan(g["ref_t_exit"]).any() and (g["bounded"] == 0).any() and (hit == 0).any()
    assert np.array_equal(O.phase_hg(g["cos"], g["g"]), g["ref_phase"])
    wi = O.phase_hg_sample(g["wo"], g["g"], g["ref_draws"].reshape(-1, 2))
    assert np.array_equal(wi, g["ref_sample_wi"].reshape(-1, 3))


def test_media_closed_forms():
    """The HG phase function integrates to one over the sphere and is 1/(4 pi)
    at g = 0; its samples' cos theta about wo follow the inverse of the HG
    CDF (mean cos theta = g)."""
    rng = np.random.default_rng(3)
    for gv in (-0.7, -0.2, 0.0, 0.3, 0.85):
        mu = np.linspace(-1, 1, 200001).astype(np.float32)
        p = O.phase_hg(mu, np.full_like(mu, gv)).astype(np.float64)
        integral = 2 * np.pi * np.trapezoid(p, mu.astype(np.float64))
        assert abs(integral - 1.0) < 2e-3, (gv, integral)
        n = 200000
        wo = np.tile(np.array([[0.0, 0.0, 1.0]], np.float32), (n, 1))
        wi = O.phase_hg_sample(wo, np.full(n, gv, np.float32), rng.random((n, 2)).astype(np.float32))
        assert abs(wi[:, 2].mean() - gv) < 5e-3, (gv, wi[:, 2].mean())
        assert np.allclose(np.linalg.norm(wi, axis=1), 1.0, atol=1e-5)
    assert abs(O.phase_hg(np.zeros(1, np.float32), np.zeros(1, np.float32))[0] - 1 / (4 * np.pi)) < 1e-8
