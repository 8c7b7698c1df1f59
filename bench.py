"""Benchmark: Mrays/s of the MI355X path tracer on the 1M-triangle
sphere_field workload (4096 UV spheres x 256 tris + ground quad, 2048x1024
sky with a 2-degree sun, 1920x1080, 64 spp, depth 12, ACES display map).

One step = one full render (Renderer::Render equivalent: every camera,
extension and shadow ray of all 64 spp) plus the ACES tone map to 8 bits, with
the scene already resident in HBM. Prints ONE JSON line.

Usage: python bench.py [--gpus 1] [--steps 3] [--warmup 1]
                       [--workload sphere_field|glass_torus|box_room|radiance_cache|serving]
                       [--no-cpu] [--spp N] [--ref-spp N] [--frames N] [--requests N] [--rate R]

--workload radiance_cache prints the equal-time relMSE line of the learned
radiance cache against the plain tracer (60 frames of 4 spp, plain reference
of --ref-spp samples per pixel). --workload serving prints the p50/p99
latency line of the continuous-batching server against running the same
Poisson request stream one request at a time.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, 8 TB/s


def log(*a):
    print(*a, file=sys.stderr, flush=True)


WORKLOADS = {
    # name: (scene factory, width, height, spp, depth, description)
    "sphere_field": ("sphere_field", 1920, 1080, 64, 12,
                     "sphere_field: 4096 UV spheres x 256 tris + ground (1,048,578 tris), 2048x1024 sky with "
                     "2-degree sun (peak 1e4), GGX materials, 1920x1080, 64 spp, depth 12, ACES"),
    "glass_torus": ("glass_torus", 1024, 1024, 32, 12,
                    "glass_torus: 65,536-tri glass torus (ior 1.5) on a Lambert plane, point light 1000 W; "
                    "10M caustic photons (radius 0.01, hash grid) + path tracing, 1024x1024, 32 spp, depth 12"),
    "box_room": ("box_room", 256, 256, 16, 8,
                 "box_room: 10-tri box, GGX conductor + glass spheres, quad light (17,12,4), 256x256, 16 spp, "
                 "depth 8, NEE+MIS"),
}
# radiance cache on the sphere_field scene: frames of 4 spp, plain vs cache
CACHE_WORKLOAD = ("sphere_field", 1920, 1080, 4, 12, 60,
                  "radiance_cache: sphere_field (1,048,578 tris), 1920x1080, 4 spp per frame x 60 frames, depth 12; "
                  "FP16 MLP 96-64x5-3 trained online (Adam lr 1e-2, batches of 32768, at most 16 steps per frame) "
                  "on 1/16 of the paths, which stop one bounce later and take the cache's value as their tail "
                  "(self-training); other paths stop at bounce 2 and "
                  "query (full_depth_training: the same with training paths run to full depth); relMSE against a "
                  "plain reference image")
CACHE_RENDER = {"query_bounce": 2, "train_every": 16, "train_suffix": 1}
MFMA_FP16_PEAK_TFLOPS = 2500.0  # MI355X dense FP16 MFMA
CACHE_FLOPS_PER_QUERY = 2 * 22720  # one multiply-add per weight


def reference_baseline(exe, sc, w, h, depth, target_s=12.0):
    """The reference renderer itself (oracle/_ref/ref_bench*: its Integrator,
    BVH, materials and Render sample loop compiled from its own sources with
    its release flags), timed on this host's cores: rows of the image at 16
    spp, top to bottom, until target_s has passed (threads stop taking new
    rows). The timed build is ref_bench_lto (the reference's link-time
    optimisation), which cannot count queries; ref_bench, which counts
    closest / any-hit queries at Scene::intersect / intersectP, then renders
    the same rows to give the queries per sample. Without ref_bench_lto,
    ref_bench is timed and counts in one run."""
    import subprocess
    import tempfile
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from bundle import read_bundle, write_bundle
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
    spp = 16
    lto = os.path.join(os.path.dirname(exe), "ref_bench_lto")
    arr = sc.bundle_arrays()

    def run(binary, r0, r1, budget_s):
        arr["cfg"] = np.array([w, h, depth, spp, r0, r1, threads, int(budget_s * 1000)], np.int32)
        with tempfile.TemporaryDirectory() as td:
            fin, fout = os.path.join(td, "in.bin"), os.path.join(td, "out.bin")
            write_bundle(fin, arr)
            env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PROC_BIND="spread")
            subprocess.run([binary, "bench", fin, fout], check=True, env=env, timeout=600,
                           stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)
            out = read_bundle(fout)
        c = [int(x) for x in out["counts"]]
        return (float(out["seconds"][0]), float(out["seconds"][1]), (c[0] << 32) | c[1], (c[2] << 32) | c[3],
                (c[4] << 32) | c[5], c[6])

    if os.path.exists(lto):
        sec, build_s, _, _, samples, used = run(lto, 0, h, target_s)
        rows = samples // (w * spp)
        _, _, closest, anyq, csamples, _ = run(exe, 0, rows, 0.0)
        rays = (closest + anyq) * samples / max(csamples, 1)
        what = ("oracle/_ref/ref_bench_lto: -O3 -ffast-math -march=x86-64-v3 -flto, OpenMP, OMP_PROC_BIND=spread; rows 0..%d of the "
                "%dx%d image at %d spp: %d samples in %.1f s; queries per sample (%.3f closest, %.3f any-hit) counted "
                "by oracle/_ref/ref_bench on the same rows" % (rows, w, h, spp, samples, sec, closest / csamples,
                                                               anyq / csamples))
    else:
        sec, build_s, closest, anyq, samples, used = run(exe, 0, h, target_s)
        rays = closest + anyq
        what = ("oracle/_ref/ref_bench: -O3 -ffast-math -march=x86-64-v3, OpenMP, no LTO; rows 0.. of the %dx%d "
                "image at %d spp: %d samples, %d rays (%d closest, %d any-hit) in %.1f s"
                % (w, h, spp, samples, rays, closest, anyq, sec))
    return {"value": round(rays / sec / 1e6, 3), "unit": "Mrays/s", "cores": used, "kind": "reference",
            "msamples_per_s": round(samples / sec / 1e6, 3), "cpu_model": cpu_model(),
            "sample": "the reference's own castRay / BVH / Render loop (%s); its scene build (%.1f s) excluded"
                      % (what, build_s)}


def cpu_baseline(sc, w, h, depth, seed, target_s=12.0):
    """The CPU baseline: the reference renderer itself when its build
    (oracle/_ref/ref_bench) is present, else the oracle built with the
    reference's release flags (-O3 -ffast-math, OpenMP), timed on this
    host's cores on a bounded sample of the same render (a band of rows, or
    the whole frame at a few spp), sized to ~target_s."""
    exe = os.path.join(ROOT, "oracle", "_ref", "ref_bench")
    if os.path.exists(exe):
        try:
            return reference_baseline(exe, sc, w, h, depth, target_s)
        except Exception as e:  # the port below is the fallback
            log("reference baseline failed (%s); timing the port" % e)
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as O
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
    t0 = time.time()
    osc = O.OracleScene(sc, fast=True)
    build_s = time.time() - t0
    r0, rows = h // 2 - 8, 16
    t0 = time.time()
    osc.render(w, h, 1, depth, seed, rows=(r0, r0 + rows), threads=threads)
    per_row = max((time.time() - t0) / rows, 1e-5)
    rows = int(min(h, max(16, target_s / per_row)))
    spp = int(max(1, min(64, target_s / (per_row * rows))))
    r0 = max(0, h // 2 - rows // 2)
    t0 = time.time()
    _, _, cnt = osc.render(w, h, spp, depth, seed, rows=(r0, r0 + rows), threads=threads)
    dt = time.time() - t0
    rays = cnt["closest"] + cnt["any"]
    return {"value": round(rays / dt / 1e6, 3), "unit": "Mrays/s", "cores": threads, "kind": "port",
            "msamples_per_s": round(cnt["samples"] / dt / 1e6, 3), "cpu_model": cpu_model(),
            "sample": "rows %d-%d of the %dx%d image at %d spp (%d samples, %d rays) in %.1f s; "
                      "oracle scene build %.1f s excluded" % (r0, r0 + rows, w, h, spp, cnt["samples"], rays, dt,
                                                              build_s)}


def traversal_bytes(sc, w, h, depth, seed):
    """Algorithmic bytes per ray, from the oracle's restatement of the
    reference's traversal on a band of the same render: 64 B per interior
    expansion (two child boxes) + 48 B per triangle test (v0, e1, e2 + ids)
    + 48 B ray in / result out. Closest hit: BVHAccel::Intersect's counts;
    any hit: BVHAccel::IntersectP's on the same binary tree (the device
    traces shadow rays on a four-wide tree built from it, whose leaves are
    the same)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as O
    osc = O.OracleScene(sc)
    r0 = h // 2 - 4
    _, _, cnt = osc.render(w, h, 1, depth, seed, rows=(r0, r0 + 8))
    per_ray = (64.0 * cnt["inner"] + 48.0 * cnt["tris"]) / max(cnt["closest"], 1) + 48.0
    per_any = (64.0 * cnt["inner_any"] + 48.0 * cnt["tris_any"]) / max(cnt["any"], 1) + 48.0
    return per_ray, per_any, cnt


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def calibrate_copy(mib):
    """A plain device copy of known size for calibrating rocprofv3's
    FETCH_SIZE / WRITE_SIZE counters: a copy of B bytes fetches and writes B
    bytes (run under rocprofv3 --pmc)."""
    import torch
    n = mib * (1 << 20) // 4
    x = torch.rand(n, device="cuda")
    y = torch.empty_like(x)
    torch.cuda.synchronize()
    for _ in range(3):
        y.copy_(x)
    torch.cuda.synchronize()
    log("copied %d MiB x3" % mib)


def profiled_traffic(workload, kernel="kExtend"):
    """HBM bytes per launch of `kernel` on this workload, measured by
    rocprofv3 --pmc (FETCH_SIZE and WRITE_SIZE per dispatch, each scaled by
    what the counter reads for a plain device copy of known size) and kept in
    the committed summary profiles/pmc_traversal_<workload>.md; None when no
    summary covers the workload."""
    import re
    path = os.path.join(ROOT, "profiles", "pmc_traversal_%s.md" % workload)
    if not os.path.exists(path):
        return None
    m = re.search(r"^%s HBM bytes per launch [^:]*: (\d+)$" % kernel, open(path).read(), re.M)
    return int(m.group(1)) if m else None


def profiled_pipeline(workload, kernel="kExtend"):
    """Busy fractions of the vector-memory pipeline for `kernel` on this
    workload, from the memory-pipeline table of the same committed counter
    summary (rocprofv3 --pmc TA_TA_BUSY_sum, TD_TD_BUSY_sum,
    TCP_TOTAL_CACHE_ACCESSES_sum over GRBM_GUI_ACTIVE): the texture address and
    data units' busy share of the kernel's cycles and the vector L1's tag
    lookups per CU per cycle (its peak is one). None without such a table."""
    import re
    path = os.path.join(ROOT, "profiles", "pmc_traversal_%s.md" % workload)
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("| `%s<" % kernel) and line.count("|") == 11:
            c = [x.strip() for x in line.strip().strip("|").split("|")]
            try:
                return {"ta_busy": round(float(c[2]) / 100, 3), "td_busy": round(float(c[4]) / 100, 3),
                        "l1_lookups_per_cu_cycle": float(c[6]), "l1_lookups_per_load": float(c[7]),
                        "source": "rocprofv3 --pmc, profiles/pmc_traversal_%s.md" % workload}
            except ValueError:
                return None
    return None


def rel_mse(x, ref):
    """mean over pixels of sum over channels of (x - ref)^2 / (ref^2 + 1e-2)."""
    return float((((x - ref) ** 2) / (ref ** 2 + 1e-2)).sum(-1).mean())


def cache_bench(args):
    """Equal-time comparison of the plain tracer and the radiance cache:
    progressive 4-spp frames of each method, wall time per frame (render +
    training), relMSE of the running mean after every frame against a plain
    reference of --ref-spp samples per pixel."""
    import torch
    from photorealisticrender4610_amd import abi, scenes
    fac, W, H, SPF, DEPTH, FRAMES, desc = CACHE_WORKLOAD
    if args.frames > 0:
        FRAMES = args.frames
    seed = 0xC4
    sc = getattr(scenes, fac)()
    ctx = abi.Context(0)
    scene = ctx.scene(sc)
    bounds = scene.info()["bounds"]
    # reference: ref_spp samples in chunks of 256 from sample index 1 << 20 on
    ref = torch.zeros((H, W, 3), dtype=torch.float64, device="cuda")
    chunk = min(256, args.ref_spp)
    t0 = time.time()
    for k in range(0, args.ref_spp, chunk):
        img, _ = ctx.render(scene, W, H, chunk, DEPTH, seed=seed, sample_offset=(1 << 20) + k)
        ref += img.double()
        if (k // chunk) % 8 == 0:
            log("reference %d/%d spp (%.0f s)" % (k + chunk, args.ref_spp, time.time() - t0))
    ref = (ref / (args.ref_spp // chunk)).float()

    def run(use_cache, render_args=None):
        render_args = CACHE_RENDER if render_args is None else render_args
        cache = ctx.cache(bounds, seed=seed) if use_cache else None
        acc = torch.zeros((H, W, 3), dtype=torch.float32, device="cuda")
        curve = []
        t = 0.0
        infer_ms = train_ms = 0.0
        queries = rays = 0
        for f in range(FRAMES):
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            if use_cache:
                img, st = cache.render(scene, W, H, SPF, DEPTH, seed=seed, sample_offset=f * SPF, **render_args)
                infer_ms += st["ms_infer"]
                train_ms += st["ms_train"]
                queries += st["queries"]
            else:
                img, st = ctx.render(scene, W, H, SPF, DEPTH, seed=seed, sample_offset=f * SPF)
            torch.cuda.synchronize()
            t += time.perf_counter() - t1
            rays += st["rays_closest"] + st["rays_any"]
            acc += img
            curve.append((t, rel_mse(acc / (f + 1), ref)))
        log("%s: %d frames in %.3f s, final relMSE %.5g" % ("cache" if use_cache else "plain", FRAMES, t,
                                                             curve[-1][1]))
        return curve, {"infer_ms": infer_ms, "train_ms": train_ms, "queries": queries, "rays": rays}

    plain, pst = run(False)
    cached, cst = run(True)
    full, fst = run(True, dict(CACHE_RENDER, train_suffix=0))  # training paths to full depth
    t_plain, e_plain = plain[-1]

    def reach(curve):  # time a curve needs to reach the plain tracer's final relMSE
        for (ta, ea), (tb, eb) in zip([(0.0, float("inf"))] + curve[:-1], curve):
            if eb <= e_plain:
                return ta + (tb - ta) * ((ea - e_plain) / (ea - eb) if np.isfinite(ea) and ea > eb else 1.0)
        return None
    t_reach, t_reach_full = reach(cached), reach(full)

    def at(curve, t):  # relMSE of the running mean at time t (last frame done by t)
        e = None
        for tt, ee in curve:
            if tt <= t:
                e = ee
        return e
    t_eq = min(t_plain, cached[-1][0])
    infer_tflops = (cst["queries"] * CACHE_FLOPS_PER_QUERY / 1e12) / (cst["infer_ms"] / 1e3) if cst["infer_ms"] else 0
    line = {
        "metric": "equal-time relMSE (radiance cache vs plain)", "value": at(cached, t_eq), "unit": "relMSE",
        "n_gpus": 1, "steps": FRAMES, "warmup": 0, "ms_per_step": round(cached[-1][0] * 1e3 / FRAMES, 2),
        "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "f32 (MLP f16, f32 accumulate)",
        "data": "synthetic (seeded procedural scene)",
        "config": {"workload": desc, "width": W, "height": H, "spp_per_frame": SPF, "frames": FRAMES,
                   "max_depth": DEPTH, "reference_spp": args.ref_spp},
        "plain_relmse_at_equal_time": at(plain, t_eq), "equal_time_s": round(t_eq, 4),
        "plain_final": {"relmse": e_plain, "time_s": round(t_plain, 4), "rays": pst["rays"]},
        "cache_final": {"relmse": cached[-1][1], "time_s": round(cached[-1][0], 4), "rays": cst["rays"],
                        "queries": cst["queries"], "train_ms": round(cst["train_ms"], 2),
                        "infer_ms": round(cst["infer_ms"], 2)},
        "cache_time_to_plain_relmse_ratio": round(t_reach / t_plain, 4) if t_reach is not None else None,
        "full_depth_training": {"relmse": full[-1][1], "time_s": round(full[-1][0], 4), "rays": fst["rays"],
                                "time_to_plain_relmse_ratio": round(t_reach_full / t_plain, 4)
                                if t_reach_full is not None else None},
        "roofline": {"bound": "mfma", "achieved": round(infer_tflops, 2), "peak": MFMA_FP16_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": round(infer_tflops / MFMA_FP16_PEAK_TFLOPS, 5), "traffic": None,
                     "kernel": "kCacheInfer (fused MLP forward, %d FLOP/query)" % CACHE_FLOPS_PER_QUERY},
        "cpu_baseline": None,
    }
    print(json.dumps(line), flush=True)


SERVING = {"requests": 400, "rate": 20.0, "sizes": (256, 512, 1024), "spps": (1, 4, 16, 64), "seed": 0xC3,
           "desc": "serving: box_room (depth 8) and sphere_field (1,048,578 tris, depth 12) resident; Poisson "
                   "arrivals at 20 req/s; resolution uniform in {256^2, 512^2, 1024^2}, spp uniform in {1, 4, 16, "
                   "64}, camera on a sphere of 2.5x the scene's bounding radius aimed at its centre; 400 requests"}


def serving_requests(n, infos, seed, rate):
    """The seeded request list: (arrival s, scene, size, spp, eye, center, fovh)."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for i in range(n):
        t += rng.exponential(1.0 / rate)
        si = int(rng.integers(0, 2))
        size = int(rng.choice(SERVING["sizes"]))
        spp = int(rng.choice(SERVING["spps"]))
        b = np.asarray(infos[si]["bounds"], np.float64)
        center = 0.5 * (b[:3] + b[3:])
        radius = 0.5 * np.linalg.norm(b[3:] - b[:3])
        v = rng.normal(size=3)
        v /= np.linalg.norm(v)
        out.append((t, si, size, spp, center + 2.5 * radius * v, center, infos[si]["fovh"]))
    return out


def serving_bench(args):
    """Continuous batching vs one request at a time on the same request list:
    p50 / p99 latency (completion - arrival) and aggregate Msamples/s."""
    import threading
    from photorealisticrender4610_amd import abi, scenes
    ctx = abi.Context(0)
    scs = [scenes.box_room(), scenes.sphere_field()]
    dev = [ctx.scene(sc) for sc in scs]
    infos = [dict(d.info(), fovh=float(sc.camera[9])) for d, sc in zip(dev, scs)]
    n = args.requests if args.requests > 0 else SERVING["requests"]
    rate = args.rate if args.rate > 0 else SERVING["rate"]
    reqs = serving_requests(n, infos, SERVING["seed"], rate)
    depths = [8, 12]

    def submit(srv, k):
        t, si, size, spp, eye, center, fovh = reqs[k]
        cam = abi.camera_desc(eye, center, (0.0, 1.0, 0.0), fovh)
        return srv.submit(si, size, size, spp, cam, seed=k)

    def run(mode, policy="fair"):
        srv = ctx.server(dev, depths, path_slots=1 << 22, policy=policy)
        ids = [None] * n
        arrive = [None] * n  # server clock, ms
        done = []
        t0 = time.perf_counter()
        if mode == "batched":
            def feeder():
                for k in range(n):
                    dt = reqs[k][0] - (time.perf_counter() - t0)
                    if dt > 0:
                        time.sleep(dt)
                    ids[k] = submit(srv, k)
                    if k % 50 == 49:
                        log("batched: submitted %d/%d" % (k + 1, n))
            th = threading.Thread(target=feeder)
            th.start()
            for k in range(n):
                while ids[k] is None:
                    time.sleep(0.0005)
                _, st = srv.wait(ids[k], image=False)
                arrive[k] = st["t_submit_ms"]
                done.append(st)
            th.join()
        else:
            anchor = None
            for k in range(n):
                dt = reqs[k][0] - (time.perf_counter() - t0)
                if dt > 0:
                    time.sleep(dt)
                _, st = srv.wait(submit(srv, k), image=False)
                if anchor is None:
                    anchor = st["t_submit_ms"] - reqs[0][0] * 1e3
                arrive[k] = anchor + reqs[k][0] * 1e3
                done.append(st)
                if k % 50 == 49:
                    log("one at a time: %d/%d done" % (k + 1, n))
        sst = srv.stats()
        srv.close()
        lat = np.array([d["t_done_ms"] - a for d, a in zip(done, arrive)])
        span = (max(d["t_done_ms"] for d in done) - min(arrive)) / 1e3
        samples = sum(d["samples"] for d in done)
        res = {"p50_ms": round(float(np.percentile(lat, 50)), 2), "p99_ms": round(float(np.percentile(lat, 99)), 2),
               "mean_ms": round(float(lat.mean()), 2), "msamples_per_s": round(samples / span / 1e6, 2),
               "span_s": round(span, 3), "samples": int(samples), "rays": int(sst["rays"]),
               "batches": int(sst["batches"]), "peak_active": int(sst["peak_active"]),
               "busy_s": round(sst["busy_ms"] / 1e3, 3)}
        log("%s %s: %s" % (mode, policy, res))
        return res

    batched = run("batched", "aged")
    others = {p: run("batched", p) for p in ("oldest", "shortest", "fair")}
    single = run("one_at_a_time")
    line = {
        "metric": "p50/p99 request latency", "value": batched["p99_ms"], "unit": "ms (p99)", "n_gpus": 1,
        "steps": n, "warmup": 0, "ms_per_step": round(batched["span_s"] * 1e3 / n, 2), "higher_is_better": False,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic (seeded procedural scenes)",
        "config": {"workload": SERVING["desc"] if rate == SERVING["rate"] else
                   SERVING["desc"].replace("%g req/s" % SERVING["rate"], "%g req/s" % rate), "requests": n,
                   "rate_per_s": rate},
        "p50_ms": batched["p50_ms"], "p99_ms": batched["p99_ms"],
        "msamples_per_s": batched["msamples_per_s"],
        "mrays_per_s": round(batched["rays"] / batched["span_s"] / 1e6, 2),
        "policy": "aged", "batched": batched, "one_at_a_time": single,
        "other_policies": {p: {k: r[k] for k in ("p50_ms", "p99_ms", "mean_ms", "msamples_per_s")}
                           for p, r in others.items()},
        "p99_speedup_vs_one_at_a_time": round(single["p99_ms"] / batched["p99_ms"], 2),
        "roofline": None, "cpu_baseline": None,
    }
    print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="sphere_field",
                    choices=sorted(WORKLOADS) + ["radiance_cache", "serving"])
    ap.add_argument("--requests", type=int, default=0, help="serving: override the request count")
    ap.add_argument("--rate", type=float, default=0.0, help="serving: override the arrival rate (requests/s)")
    ap.add_argument("--ref-spp", type=int, default=16384, help="radiance_cache: reference samples per pixel")
    ap.add_argument("--frames", type=int, default=0, help="radiance_cache: override the frame count")
    ap.add_argument("--spp", type=int, default=0, help="override spp (profiling runs only)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB (counter calibration under rocprofv3 --pmc)")
    args = ap.parse_args()
    if args.calibrate_copy > 0:
        return calibrate_copy(args.calibrate_copy)
    if args.gpus != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path (the render would shard "
              "by pixels/samples with one final reduce). Use --gpus 1.", file=sys.stderr)
        sys.exit(2)

    if args.workload == "radiance_cache":
        return cache_bench(args)
    if args.workload == "serving":
        return serving_bench(args)
    import torch
    from photorealisticrender4610_amd import abi, scenes

    fac, W, H, SPP, DEPTH, desc = WORKLOADS[args.workload]
    if args.spp > 0:
        SPP = args.spp
    seed = 0xC1
    t0 = time.time()
    sc = getattr(scenes, fac)()
    log("scene generated in %.1f s: %d tris" % (time.time() - t0, sc.num_triangles))
    ctx = abi.Context(0)
    t0 = time.time()
    scene = ctx.scene(sc)
    log("host SAH build + upload in %.1f s: %s" % (time.time() - t0, scene.info()))
    out = torch.empty((H, W, 3), dtype=torch.float32, device="cuda")
    photons = args.workload == "glass_torus"

    def photon_pass(k):
        if photons:
            return ctx.photon_build(scene, 10_000_000, 0.01, 30, seed + 1000 + k)
        return 0

    for k in range(args.warmup):
        photon_pass(k)
        _, st = ctx.render(scene, W, H, SPP, DEPTH, seed=seed + k, out=out, photons=photons)
        ctx.tonemap("aces", out.view(-1, 3), 1.0)
        log("warmup %d: %.1f ms, %d rays" % (k, st["ms"], st["rays_closest"] + st["rays_any"]))
    torch.cuda.synchronize()
    rays = closest = skipped = 0
    ms_ext = 0.0
    launches = 0
    t0 = time.perf_counter()
    stored = 0
    anyr = 0
    ms_sh = 0.0
    launches_sh = 0
    for k in range(args.steps):
        stored = photon_pass(100 + k)
        _, st = ctx.render(scene, W, H, SPP, DEPTH, seed=seed + 100 + k, out=out, kernel_timing=True,
                           photons=photons)
        ctx.tonemap("aces", out.view(-1, 3), 1.0)
        rays += st["rays_closest"] + st["rays_any"]
        skipped += st["rays_any_skipped"]  # NEE terms exactly zero: shadow rays not traced
        closest += st["rays_closest"] - st["rays_closest_finish"]  # the rays kExtend traced
        anyr += st["rays_any"] - st["rays_any_finish"]  # the rays kShadow traced
        ms_ext += st["ms_extend"]
        launches += st["launches_extend"]
        ms_sh += st["ms_shadow"]
        launches_sh += st["launches_shadow"]
        log("step %d: %.1f ms (extend %.1f, shadow %.1f, shadow binning %.1f, shade %.1f, finish %.1f), %d rays" % (
            k, st["ms"], st["ms_extend"], st["ms_shadow"], st["ms_sort"], st["ms_shade"], st["ms_finish"],
            st["rays_closest"] + st["rays_any"]))
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    value = rays / dt / 1e6

    per_ray, per_any, ocnt = traversal_bytes(sc, W, H, DEPTH, seed)

    def kernel_roofline(name, desc, ms, nlaunch, nrays, bytes_per_ray):
        """achieved / frac: bytes the kernel moved between L2 and the memory
        fabric per launch (rocprofv3 FETCH_SIZE + WRITE_SIZE, copy-calibrated,
        from the committed counter summary of this workload) over the average
        launch time measured here with HIP events; the algorithmic bytes
        (what a cacheless machine would move) are reported beside them."""
        avg_ms = ms / max(nlaunch, 1)
        traffic = profiled_traffic(args.workload, name) if args.spp == 0 else None
        alg = nrays / max(nlaunch, 1) * bytes_per_ray
        ach = traffic / (avg_ms / 1e3) / 1e9 if traffic and avg_ms > 0 else None
        return {"bound": "hbm", "achieved": round(ach, 1) if ach else None, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(ach / HBM_PEAK_GBS, 4) if ach else None, "traffic": traffic,
                "traffic_unit": "fabric bytes per launch (rocprofv3 --pmc FETCH_SIZE, WRITE_SIZE; copy-calibrated)",
                "kernel": desc, "avg_launch_ms": round(avg_ms, 4), "launches_per_step": nlaunch // args.steps,
                "summed_ms_per_step": round(ms / args.steps, 2),
                "algorithmic": {"bytes_per_ray": round(bytes_per_ray, 1), "bytes_per_launch": int(alg),
                                "GBps": round(alg / (avg_ms / 1e3) / 1e9, 1) if avg_ms > 0 else None},
                # what binds the kernel: each lane's 16-B record load is one L1
                # lookup, and the vector-memory units run near their rate
                "pipeline": profiled_pipeline(args.workload, name) if args.spp == 0 else None}
    rl_ext = kernel_roofline("kExtend", "kExtend (closest-hit traversal, extension rays)", ms_ext, launches, closest,
                             per_ray)
    rl_sh = kernel_roofline("kShadow", "kShadow (any-hit traversal, shadow rays)", ms_sh, launches_sh, anyr, per_any)
    # the dominant kernel by summed launch time (the two overlap on two streams)
    roofline = rl_sh if ms_sh >= ms_ext else rl_ext
    cpu = None
    if not args.no_cpu:
        try:
            cpu = cpu_baseline(sc, W, H, DEPTH, seed)
        except Exception as e:  # the baseline is reported, never required
            log("cpu baseline failed:", e)
    line = {
        "metric": "Mrays/s", "value": round(value, 2), "unit": "Mrays/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(dt * 1e3 / args.steps, 2),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded procedural scene)",
        "config": {"workload": desc, "width": W, "height": H, "spp": SPP, "max_depth": DEPTH,
                   "triangles": sc.num_triangles},
        "msamples_per_s": round(W * H * SPP * args.steps / dt / 1e6, 2),
        "rays_per_step": rays // args.steps,
        # shadow queries of the reference's loop whose NEE term is exactly zero
        # (the sample cannot depend on them): not traced, not counted in value
        "shadow_queries_skipped_per_step": skipped // args.steps,
        "photons_stored": stored if photons else None,
        "roofline": roofline,
        "roofline_kernels": [rl_sh, rl_ext],
        "cpu_baseline": cpu,
    }
    if cpu:
        line["gpu_over_cpu"] = round(value / cpu["value"], 1)
        if cpu.get("msamples_per_s"):  # the same image work on both sides
            line["gpu_over_cpu_samples"] = round(line["msamples_per_s"] / cpu["msamples_per_s"], 1)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
