"""Benchmark: Mrays/s of the MI355X path tracer on the 1M-triangle
sphere_field workload (4096 UV spheres x 256 tris + ground quad, 2048x1024
sky with a 2-degree sun, 1920x1080, 64 spp, depth 12, ACES display map).

One step = one full render (Renderer::Render equivalent: every camera,
extension and shadow ray of all 64 spp) plus the ACES tone map to 8 bits, with
the scene already resident in HBM. Prints ONE JSON line.

Usage: python bench.py [--gpus 1] [--steps 3] [--warmup 1] [--workload sphere_field|box_room]
                       [--no-cpu] [--spp N]
"""
import argparse
import json
import os
import sys
import time

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


def cpu_baseline(sc, w, h, depth, seed, target_s=12.0):
    """The oracle built with the reference's release flags (-O3 -ffast-math,
    OpenMP), timed on this host's cores on a bounded sample of the same render
    (a band of rows, or the whole frame at a few spp), sized to ~target_s."""
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
            "sample": "rows %d-%d of the %dx%d image at %d spp (%d samples, %d rays) in %.1f s; "
                      "oracle scene build %.1f s excluded" % (r0, r0 + rows, w, h, spp, cnt["samples"], rays, dt,
                                                              build_s)}


def traversal_bytes(sc, w, h, depth, seed):
    """Algorithmic bytes per closest-hit ray, from the oracle's restatement of
    BVHAccel::Intersect on a band of the same render: 64 B per interior
    expansion (two child boxes) + 48 B per triangle test (v0, e1, e2 + ids)
    + 32 B per sphere test + 48 B ray in / hit out."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as O
    osc = O.OracleScene(sc)
    r0 = h // 2 - 4
    _, _, cnt = osc.render(w, h, 1, depth, seed, rows=(r0, r0 + 8))
    per_ray = (64.0 * cnt["inner"] + 48.0 * cnt["tris"]) / max(cnt["closest"], 1) + 48.0
    return per_ray, cnt


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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="sphere_field", choices=sorted(WORKLOADS))
    ap.add_argument("--spp", type=int, default=0, help="override spp (profiling runs only)")
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path (the render would shard "
              "by pixels/samples with one final reduce). Use --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import numpy as np
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
    rays = closest = 0
    ms_ext = 0.0
    launches = 0
    t0 = time.perf_counter()
    stored = 0
    for k in range(args.steps):
        stored = photon_pass(100 + k)
        _, st = ctx.render(scene, W, H, SPP, DEPTH, seed=seed + 100 + k, out=out, kernel_timing=True,
                           photons=photons)
        ctx.tonemap("aces", out.view(-1, 3), 1.0)
        rays += st["rays_closest"] + st["rays_any"]
        closest += st["rays_closest"]
        ms_ext += st["ms_extend"]
        launches += st["launches_extend"]
        log("step %d: %.1f ms (extend %.1f, shadow %.1f, shade %.1f), %d rays" % (
            k, st["ms"], st["ms_extend"], st["ms_shadow"], st["ms_shade"], st["rays_closest"] + st["rays_any"]))
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    value = rays / dt / 1e6

    per_ray, ocnt = traversal_bytes(sc, W, H, DEPTH, seed)
    achieved = (closest * per_ray / 1e9) / (ms_ext / 1e3) if ms_ext > 0 else 0.0
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": profiled_traffic(args.workload) if args.spp == 0 else None,
                "traffic_unit": "HBM bytes per launch (rocprofv3 --pmc, copy-calibrated)",
                "kernel": "kExtend (closest-hit traversal)",
                "bytes_per_ray": round(per_ray, 1),
                "avg_launch_ms": round(ms_ext / max(launches, 1), 4)}
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
        "photons_stored": stored if photons else None,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    if cpu:
        line["gpu_over_cpu"] = round(value / cpu["value"], 1)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
