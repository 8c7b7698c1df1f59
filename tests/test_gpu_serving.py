"""The continuous-batching server (prr_server_*): a request renders the image
prr_render renders, bit for bit, alone and in a mix of requests that split
into many batches; admission keeps admitted memory under the HBM budget."""
import threading

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from photorealisticrender4610_amd import abi
    c = abi.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="module")
def scenes(ctx):
    from photorealisticrender4610_amd import scenes as S
    import testscenes as TS
    box, field = S.box_room(), TS.mini_field()
    return [(box, ctx.scene(box), 8), (field, ctx.scene(field), 12)]


def _cam(sc, abi):
    c = np.asarray(sc.camera, np.float32)
    return abi.camera_desc(c[0:3], c[3:6], c[6:9], c[9], c[10], c[11])


def _orbit(abi, center, radius, k, fovh=40.0):
    rng = np.random.default_rng(1000 + k)
    v = rng.normal(size=3)
    v[1] = abs(v[1]) + 0.2
    v /= np.linalg.norm(v)
    return abi.camera_desc(center + radius * v, center, (0.0, 1.0, 0.0), fovh)


def test_request_equals_prr_render(ctx, scenes):
    from photorealisticrender4610_amd import abi
    print("case serving request == prr_render", flush=True)
    srv = ctx.server([s[1] for s in scenes], [s[2] for s in scenes], path_slots=1 << 16)
    try:
        for k, (sc, scene, depth) in enumerate(scenes):
            rid = srv.submit(k, 80, 60, 6, _cam(sc, abi), seed=11 + k)
            img, st = srv.wait(rid)
            ref, _ = ctx.render(scene, 80, 60, 6, depth, seed=11 + k)
            assert np.array_equal(img, ref.cpu().numpy())
            assert st["samples"] == 80 * 60 * 6 and st["t_done_ms"] >= st["t_admit_ms"] >= st["t_submit_ms"]
    finally:
        srv.close()


def test_same_alone_and_in_mix(ctx, scenes):
    """Seven requests of both scenes, sizes and sample counts, submitted from
    two threads into a server whose batches hold 20,000 paths (so requests
    split into partial sample rounds and share batches): each image equals
    the image of the same request rendered alone."""
    from photorealisticrender4610_amd import abi
    print("case serving alone vs mix", flush=True)
    reqs = []
    for k in range(7):
        si = k % 2
        sc = scenes[si][0]
        b = np.asarray(sc.camera, np.float32)
        center = b[3:6]
        radius = 2.5 if si == 0 else 12.0
        reqs.append((si, [64, 96, 48][k % 3], [48, 64, 40][k % 3], [1, 4, 9][k % 3], _orbit(abi, center, radius, k),
                     100 + k))
    srv = ctx.server([s[1] for s in scenes], [s[2] for s in scenes], path_slots=20000)
    try:
        alone = []
        for r in reqs:
            alone.append(srv.wait(srv.submit(*r[:5], seed=r[5]))[0])
        ids = [None] * len(reqs)

        def feed(idx):
            for i in idx:
                ids[i] = srv.submit(*reqs[i][:5], seed=reqs[i][5])
        t1 = threading.Thread(target=feed, args=([0, 2, 4, 6],))
        t2 = threading.Thread(target=feed, args=([1, 3, 5],))
        t1.start()
        t2.start()
        t1.join()
        t2.join()
        mixed = [srv.wait(i) for i in ids]
        st = srv.stats()
    finally:
        srv.close()
    for k in range(len(reqs)):
        assert np.array_equal(mixed[k][0], alone[k]), k
    assert max(m[1]["batches"] for m in mixed) > 1
    assert st["peak_active"] >= 2


def test_admission_by_budget(ctx, scenes):
    """Budget for two 64x64 framebuffers: at most two requests are admitted
    at once, later ones wait for memory, and all finish."""
    from photorealisticrender4610_amd import abi
    print("case serving admission", flush=True)
    sc = scenes[0][0]
    fb = 64 * 64 * 16
    srv = ctx.server([s[1] for s in scenes], [s[2] for s in scenes], path_slots=1 << 14, hbm_budget=2 * fb)
    try:
        ids = [srv.submit(0, 64, 64, 8, _cam(sc, abi), seed=k) for k in range(5)]
        outs = [srv.wait(i) for i in ids]
        st = srv.stats()
        with pytest.raises(Exception):
            srv.submit(0, 128, 128, 1, _cam(sc, abi))  # larger than the whole budget
    finally:
        srv.close()
    assert st["peak_active"] <= 2 and st["peak_admitted_bytes"] <= 2 * fb
    assert st["requests_done"] == 5
    # the third request is admitted only after one of the first two is done
    assert outs[2][1]["t_admit_ms"] >= min(outs[0][1]["t_done_ms"], outs[1][1]["t_done_ms"])


@pytest.fixture(scope="module")
def workload_scenes(ctx):
    """The serving workload's resident scenes: box_room (depth 8) and the
    1,048,578-triangle sphere_field with its 2048 x 1024 sky (depth 12)."""
    from photorealisticrender4610_amd import scenes as S
    box, field = S.box_room(), S.sphere_field()
    return [(box, ctx.scene(box), 8), (field, ctx.scene(field), 12)]


def test_workload_scenes_alone_and_in_mix(ctx, workload_scenes):
    """The serving workload's own scenes and request shapes: resolutions
    256^2 and 512^2, 1 / 4 / 16 spp, cameras on a sphere of 2.5x each
    scene's bounding radius aimed at its centre (as bench.py draws them).
    Six requests go in from two threads with batches of 2^20 paths, so the
    larger ones split over batches shared with others: each image equals
    the same request served alone, and a sphere_field request at the scene's
    own camera equals prr_render."""
    from photorealisticrender4610_amd import abi
    print("case serving workload scenes alone vs mix", flush=True)
    rng = np.random.default_rng(0xC3)
    reqs = []
    for k in range(6):
        si = k % 2
        sc, scene, _ = workload_scenes[si]
        b = np.asarray(scene.info()["bounds"], np.float64)
        center = 0.5 * (b[:3] + b[3:])
        radius = 0.5 * np.linalg.norm(b[3:] - b[:3])
        v = rng.normal(size=3)
        v /= np.linalg.norm(v)
        size, spp = [256, 512, 256][k % 3], [4, 1, 16][k % 3]
        cam = abi.camera_desc(center + 2.5 * radius * v, center, (0.0, 1.0, 0.0), float(sc.camera[9]))
        reqs.append((si, size, size, spp, cam, 300 + k))
    srv = ctx.server([s[1] for s in workload_scenes], [s[2] for s in workload_scenes], path_slots=1 << 20)
    try:
        field, fscene, fdepth = workload_scenes[1]
        img, _ = srv.wait(srv.submit(1, 256, 256, 2, _cam(field, abi), seed=7))
        ref, _ = ctx.render(fscene, 256, 256, 2, fdepth, seed=7)
        assert np.array_equal(img, ref.cpu().numpy())
        alone = [srv.wait(srv.submit(*r[:5], seed=r[5]))[0] for r in reqs]
        ids = [None] * len(reqs)

        def feed(idx):
            for i in idx:
                ids[i] = srv.submit(*reqs[i][:5], seed=reqs[i][5])
        t1 = threading.Thread(target=feed, args=([0, 2, 4],))
        t2 = threading.Thread(target=feed, args=([1, 3, 5],))
        t1.start()
        t2.start()
        t1.join()
        t2.join()
        mixed = [srv.wait(i) for i in ids]
        st = srv.stats()
    finally:
        srv.close()
    for k in range(len(reqs)):
        # box_room seen from outside is black (its light is inside the box);
        # the sun-lit sphere_field never is
        assert np.isfinite(alone[k]).all() and (reqs[k][0] == 0 or alone[k].mean() > 0), k
        assert np.array_equal(mixed[k][0], alone[k]), k
    assert max(m[1]["batches"] for m in mixed) > 1
    assert st["peak_active"] >= 2
