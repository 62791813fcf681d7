#!/usr/bin/env python3
"""Single-frame throughput of the serialized-SVO treelet ray caster.

Workload: depth-11 value-noise terrain (2048^3 leaf grid) in 4096-node
treelets, all resident in HBM; one 1920x1080 FP32 primary-ray frame from a
seeded camera 0.1 above the surface; treelet (OpenCL renderer) traversal
policy, full depth (no LOD stop).  A step is one launch that renders the
frame with the fused colour / depth / node / treelet epilogue; inputs are
resident in HBM before the timed region starts.

Prints one JSON line (rank 0, one GPU).  Progress goes to stderr.
"""
import argparse
import json
import os
import statistics
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Mrays/s (depth-11 SVO, 1080p) + p50/p99 request latency; % HBM peak"
HBM_PEAK_GBS = 8000.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_baseline(scene, cam, threads, frames=5):
    """Oracle restatement of CpuRaycasterSingleRay3 (the reference CPU caster)
    timed on the host: OpenMP over rows on `threads` threads, median of
    `frames` full frames after one warm-up, plus one single-thread frame."""
    import oracle
    with tempfile.TemporaryDirectory() as d:
        base = os.path.join(d, "scene")
        scene.write(base)
        info = scene.info
        ora = oracle.OracleScene(base, info["treelet_bytes"] * info["num_treelets"])
    rays = cam.width * cam.height
    ora.render(cam, "cpu3", threads=threads, extras=False)
    times = []
    for _ in range(frames):
        t0 = time.perf_counter()
        ora.render(cam, "cpu3", threads=threads, extras=False)
        times.append(time.perf_counter() - t0)
    med = statistics.median(times)
    t0 = time.perf_counter()
    ora.render(cam, "cpu3", threads=1, extras=False)
    single = time.perf_counter() - t0
    t0 = time.perf_counter()
    ora.render(cam, "treelet", threads=threads, extras=False)
    pol_a = time.perf_counter() - t0
    return {
        "value": rays / med / 1e6,
        "unit": "Mrays/s",
        "cores": threads,
        "kind": "port",
        "sample": (f"oracle CpuRaycasterSingleRay3 policy, same scene and camera, {cam.width}x{cam.height} rays, "
                   f"median of {frames} frames after 1 warm-up on {threads} threads; single thread "
                   f"{rays / single / 1e6:.2f} Mrays/s; treelet policy {rays / pol_a / 1e6:.2f} Mrays/s"),
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--policy", default="treelet", choices=["treelet", "cpu3"])
    ap.add_argument("--lod", type=float, default=0.0, help="tScaleRatio multiplier; 0 = full depth")
    ap.add_argument("--depth", type=int, default=11)
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--copy-calibration-mb", type=int, default=0,
                    help="after the timed launches, run one plain device copy of this many MiB "
                         "(known traffic for scaling FETCH_SIZE / WRITE_SIZE counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (the single-GPU build has no "
              f"multi-GPU path)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import libsvo_amd as svo
    from libsvo_amd import workloads

    if svo.device_count() < 1:
        log("no HIP device visible")
        sys.exit(1)
    t0 = time.perf_counter()
    scene = workloads.terrain_scene(depth=args.depth)
    info = scene.info
    log(f"scene built in {time.perf_counter() - t0:.1f}s: {info}")
    mm = svo.TreeletMemoryManagerHip(scene, workloads.incore_bytes_for(scene), device=0, resident_all=True)
    log(f"resident: {mm.info}")
    cam = workloads.terrain_camera(workloads.TERRAIN_CAMERA_SEED, args.width, args.height, lod=args.lod)
    rays = cam.width * cam.height
    stream = svo.Stream()

    # untimed launch with per-ray statistics -> algorithmic bytes
    stat = svo.Frame.alloc(cam.width, cam.height, pushes=True)
    svo.render_to_buffer(mm, [cam], [stat], args.policy, stream)
    stream.synchronize()
    res = stat.fetch()
    alg_bytes = workloads.algorithmic_bytes(res["pushes"], res["node"])
    hits = int(np.count_nonzero(res["node"] != svo.MISS))
    log(f"hits {hits}/{rays}, mean pushes {res['pushes'].mean():.2f}, algorithmic bytes/launch {alg_bytes}")

    frame = svo.Frame.alloc(cam.width, cam.height)
    if args.warmup > 0:
        svo.render_timed(mm, [cam], [frame], args.policy, stream, iters=args.warmup)
    svo._capi.check(svo.lib.svo_device_synchronize(), "sync")
    w0 = time.perf_counter()
    ms = svo.render_timed(mm, [cam], [frame], args.policy, stream, iters=args.steps)
    svo._capi.check(svo.lib.svo_device_synchronize(), "sync")
    wall = time.perf_counter() - w0
    value = rays / (ms / 1e3) / 1e6
    achieved = alg_bytes / (ms / 1e3) / 1e9
    log(f"{args.steps} launches: {ms:.4f} ms/launch (events), wall {wall * 1e3 / args.steps:.4f} ms/launch, "
        f"{value:.1f} Mrays/s, {achieved:.1f} GB/s algorithmic")

    if args.copy_calibration_mb > 0:
        nbytes = args.copy_calibration_mb << 20
        src, dst = svo.DeviceBuffer(nbytes), svo.DeviceBuffer(nbytes)
        src.fill(1)
        svo._capi.check(svo.lib.svo_device_synchronize(), "sync")
        svo._capi.check(svo.lib.svo_debug_copy(dst.ptr, src.ptr, nbytes, stream.handle), "svo_debug_copy")
        stream.synchronize()
        log(f"calibration copy of {nbytes} bytes done")

    out = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"terrain depth-{args.depth} SVO, 4096-node treelets all resident, "
                        f"{cam.width}x{cam.height} primary rays, {args.policy} policy, "
                        f"lod multiplier {args.lod}",
            "treelets": info["num_treelets"],
            "tree_nodes": info["tree_nodes"],
            "hit_rays": hits,
            "mean_pushes": round(float(res["pushes"].mean()), 3),
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 2),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 5),
            "traffic": None,
        },
    }
    if not args.no_cpu_baseline:
        threads = min(16, os.cpu_count() or 1)
        log(f"cpu baseline on {threads} threads")
        out["cpu_baseline"] = cpu_baseline(scene, cam, threads)
        log(f"cpu baseline {out['cpu_baseline']['value']:.2f} Mrays/s; GPU/CPU = "
            f"{value / out['cpu_baseline']['value']:.1f}x")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
