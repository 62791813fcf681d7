#!/usr/bin/env python3
"""Throughput of the serialized-SVO treelet ray caster on one GPU.

Headline (``value``, launch shape from the autotuner): depth-11 value-noise terrain (2048^3 leaf grid) in
4096-node treelets, all resident in HBM; one 1920x1080 FP32 primary-ray frame
from a seeded camera 0.1 above the surface; treelet (OpenCL renderer)
traversal policy, full depth (no LOD stop).  A step is one launch that
renders the frame with the fused colour / depth / node / treelet epilogue;
inputs are resident in HBM before the timed region starts.

Also reported in the same line:
* ``policies``: the same frame under both GPU traversal policies (the
  treelet renderer's and the CPU caster's), each with its own tuned shape;
* ``reference_case``: the reference's own CPU-runnable case (level-4 Menger
  sponge at depth 8 in one treelet, 512x512 rays from a seeded point on a
  sphere of radius 2.5) under both GPU policies, next to the oracle's
  restatement of the CPU caster on the host;
* ``batched``: 64 views over the depth-11 terrain and a depth-10 sphere
  cloud, resolutions uniform over {256^2, 512^2, 720p, 1080p}, one launch
  captured in a hipGraph, launch shape from the autotuner;
* ``ooc``: depth-13 terrain in 16384-node treelets (larger than the pool)
  behind an 8 GiB device pool, 300 1080p frames along a seeded fly-through;
  each frame runs the analyse pass, the feedback update (page-ins,
  evictions), the upload and the render: frames/s and page-ins per frame,
  one stage after the other, and pipelined (frame i's render overlapping
  the host update of frame i+1);
* ``serving``: Poisson request streams over four resident scenes (Menger
  depth 9, terrain depth 11, sphere cloud depth 10, terrain depth 12) in a
  64 GB HBM budget, served by the continuous-batching server and by the plain
  one-frame-per-launch loop: frames/s, Mrays/s, p50/p99 latency, from low
  rates past saturation;
* ``autotune``: the headline's autotune cache key, the tuning time and the
  time of a second tuning call of the same workload class (a cache hit);
* ``cpu_baseline``, ``boundary``: the CPU caster's restatement on the host
  and the headline frame's boundary-ray and differing-pixel counts.

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


PROFILE = os.path.join(ROOT, "profiles", "trace_kernel_terrain_d11_1080p.md")


def profiled_traffic(shape):
    """HBM bytes per launch of the 1080p depth-11 frame for a launch shape,
    from the FETCH_SIZE / WRITE_SIZE counter passes summarised in the
    committed profile (None when the shape has no row there)."""
    import re
    try:
        text = open(PROFILE).read()
    except OSError:
        return None
    key = f"| ({shape[0]}, {shape[1]}, {shape[2]}) |"
    section = text.find("## HBM traffic per launch shape")
    if section < 0:
        return None
    for line in text[section:].splitlines():
        if line.startswith(key):
            m = re.findall(r"\|\s*([\d,]+)\s*", line[len(key) - 1:])
            if len(m) >= 3:
                return int(m[2].replace(",", ""))
    return None


def profiled_l2_hit(shape):
    """L2 hit rate TCC_HIT / (TCC_HIT + TCC_MISS) of the 1080p depth-11 frame
    for a launch shape, from the same table of the committed profile (None
    when the shape has no row there)."""
    try:
        text = open(PROFILE).read()
    except OSError:
        return None
    key = f"| ({shape[0]}, {shape[1]}, {shape[2]}) |"
    section = text.find("## HBM traffic per launch shape")
    if section < 0:
        return None
    for line in text[section:].splitlines():
        if line.startswith(key):
            cells = [c.strip() for c in line.strip().strip("|").split("|")]
            try:
                return float(cells[4])
            except (IndexError, ValueError):
                return None
    return None


VALU_CEILING_PER_CU_CYCLE = 2.0  # gfx950: one wave64 VALU instruction per SIMD every 2 cycles, 4 SIMDs per CU


def profiled_issue(shape):
    """VALU issue of the 1080p depth-11 frame for a launch shape, from the
    issue-counter table of the committed profile: instructions per CU-cycle
    against the 2.0 ceiling, SIMD busy fraction and lane use (None when the
    shape has no row there)."""
    import re
    try:
        text = open(PROFILE).read()
    except OSError:
        return None
    section = text.find("## Issue counters per launch shape")
    if section < 0:
        return None
    key = f"| ({shape[0]}, {shape[1]}, {shape[2]}) |"
    for line in text[section:].splitlines():
        if line.startswith(key):
            cells = [c.strip() for c in line.strip().strip("|").split("|")]
            try:
                valu, busy, lane = float(cells[3]), float(cells[4]), float(cells[5])
            except (IndexError, ValueError):
                return None
            return {"bound": "valu_issue", "achieved": valu, "peak": VALU_CEILING_PER_CU_CYCLE,
                    "unit": "VALU instructions per CU-cycle", "frac": round(valu / VALU_CEILING_PER_CU_CYCLE, 3),
                    "simd_busy": busy, "frac_while_busy": round(valu / busy / VALU_CEILING_PER_CU_CYCLE, 3),
                    "lane_use": lane}
    return None


def stream_copy_gbs(stream, nbytes=1 << 30, iters=10):
    """HBM bandwidth of a plain device copy (16 B per lane, grid-stride) of
    `nbytes`, read plus written, over `iters` back-to-back copies after two
    warm-ups: the measured counterpart of the 8 TB/s vendor peak."""
    import libsvo_amd as svo
    src, dst = svo.DeviceBuffer(nbytes), svo.DeviceBuffer(nbytes)
    for _ in range(2):
        svo._capi.check(svo.lib.svo_debug_copy(dst.ptr, src.ptr, nbytes, stream.handle), "svo_debug_copy")
    svo._capi.check(svo.lib.svo_device_synchronize(), "sync")
    t0 = time.perf_counter()
    for _ in range(iters):
        svo._capi.check(svo.lib.svo_debug_copy(dst.ptr, src.ptr, nbytes, stream.handle), "svo_debug_copy")
    svo._capi.check(svo.lib.svo_device_synchronize(), "sync")
    dt = time.perf_counter() - t0
    return 2.0 * nbytes * iters / dt / 1e9


def host_cpus():
    """CPU threads the baseline may use: the process's affinity mask, capped
    by OMP_NUM_THREADS when it is set (the CPU share granted to the process
    that drives one GPU), with the machine's core count for the record."""
    affinity = len(os.sched_getaffinity(0))
    omp = os.environ.get("OMP_NUM_THREADS")
    threads = affinity
    if omp and omp.isdigit() and int(omp) > 0:
        threads = min(threads, int(omp))
    return {"nproc": os.cpu_count(), "affinity_cpus": affinity, "omp_num_threads": omp, "threads": max(1, threads)}


def _oracle_scene(scene):
    import oracle
    with tempfile.TemporaryDirectory() as d:
        base = os.path.join(d, "scene")
        scene.write(base)
        info = scene.info
        return oracle.OracleScene(base, info["treelet_bytes"] * info["num_treelets"])


def cpu_baseline(scene, cam, cpus, frames=5, ora=None):
    """Oracle restatement of CpuRaycasterSingleRay3 (the reference CPU caster)
    timed on the host: OpenMP over rows on every thread the process may use,
    median of `frames` full frames after one warm-up, plus one single-thread
    frame and one frame of the treelet renderer's policy."""
    ora = ora or _oracle_scene(scene)
    threads = cpus["threads"]
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
        "nproc": cpus["nproc"],
        "affinity_cpus": cpus["affinity_cpus"],
        "single_thread_mrays_per_s": round(rays / single / 1e6, 3),
        "treelet_policy_mrays_per_s": round(rays / pol_a / 1e6, 3),
        "sample": (f"oracle CpuRaycasterSingleRay3 policy, same scene and camera, {cam.width}x{cam.height} rays, "
                   f"median of {frames} frames after 1 warm-up on {threads} threads (affinity mask "
                   f"{cpus['affinity_cpus']} CPUs, OMP_NUM_THREADS={cpus['omp_num_threads']}, nproc "
                   f"{cpus['nproc']}); single thread {rays / single / 1e6:.2f} Mrays/s; treelet policy "
                   f"{rays / pol_a / 1e6:.2f} Mrays/s on {threads} threads"),
    }


def cpu_parity(ora, cam, gpu, threads):
    """The CPU leg's check of the headline frame: the oracle renders it under
    both policies with its boundary-ray flags (a ray passing within 1e-6 of
    a child face, edge or pop threshold, a direction component below the
    clamp, or a maxLoops trip); the GPU frames of both policies are compared
    with them pixel by pixel (node, treelet, rgba, t bits)."""
    import numpy as np
    ref = {p: ora.render(cam, p, threads=threads, extras=True) for p in ("treelet", "cpu3")}
    flagged = ref["treelet"]["boundary"].astype(bool) | ref["cpu3"]["boundary"].astype(bool)

    def differ(a, b):
        d = np.zeros(a["node"].shape, bool)
        for k in ("node", "treelet", "rgba"):
            d |= a[k] != b[k]
        return d | (a["t"].view(np.uint32) != b["t"].view(np.uint32))

    a_a, c_c = differ(gpu["treelet"], ref["treelet"]), differ(gpu["cpu3"], ref["cpu3"])
    a_c = differ(gpu["treelet"], ref["cpu3"])
    node_a_c = gpu["treelet"]["node"] != ref["cpu3"]["node"]
    return {
        "pixels": int(flagged.size), "boundary_flagged": int(flagged.sum()),
        "gpu_treelet_vs_oracle_treelet_differ": int(a_a.sum()),
        "gpu_cpu3_vs_oracle_cpu3_differ": int(c_c.sum()),
        "gpu_treelet_vs_oracle_cpu3_differ": int(a_c.sum()),
        "gpu_treelet_vs_oracle_cpu3_node_differ": int(node_a_c.sum()),
        "of_which_boundary_flagged": int((a_c & flagged).sum()),
    }


def reference_case_section(stream, cpus, cpu=True, launches=50, shape=None):
    """The reference's own CPU-runnable case: a level-4 Menger sponge at depth
    8 in one treelet, 512x512 primary rays from a seeded point on a sphere of
    radius 2.5 (60 degree field of view), both GPU policies autotuned, next
    to the oracle's restatement of the CPU caster on the host."""
    import libsvo_amd as svo
    from libsvo_amd import workloads
    sc = workloads.menger_scene(depth=8, level=4, treelet_bytes=0)
    mm = svo.TreeletMemoryManagerHip(sc, workloads.incore_bytes_for(sc), device=0, resident_all=True)
    cam = workloads.sphere_camera(workloads.MENGER_CAMERA_SEED, 512, 512)
    rays = cam.width * cam.height
    view = [svo.View(mm, cam, svo.Frame.alloc(cam.width, cam.height))]
    res = {"workload": "Menger sponge level 4, depth 8, one treelet, 512x512 primary rays, LOD off",
           "tree_nodes": sc.info["tree_nodes"]}
    given = shape
    for policy in ("treelet", "cpu3"):
        r = svo.Renderer(0)
        if given:
            r.config = given
            shape = r.config
        else:
            shape, _ = r.autotune(view, policy, stream, iters=10)
        r.timed(view, policy, stream, iters=5)
        ms = r.timed(view, policy, stream, iters=launches)
        res[policy] = {"ms_per_launch": round(ms, 5), "mrays_per_s": round(rays / ms / 1e3, 1), "launch_shape": list(shape)}
        log(f"reference case, {policy} policy: {shape} {ms:.4f} ms/launch = {rays / ms / 1e3:.1f} Mrays/s")
    if cpu:
        base = cpu_baseline(sc, cam, cpus)
        res["cpu_baseline"] = base
        res["gpu_over_cpu"] = round(res["cpu3"]["mrays_per_s"] / base["value"], 1)
        res["gpu_over_single_cpu_thread"] = round(res["cpu3"]["mrays_per_s"] / base["single_thread_mrays_per_s"], 1)
        log(f"reference case on the CPU: {base['value']:.2f} Mrays/s on {cpus['threads']} threads; "
            f"GPU (same policy) / CPU = {res['gpu_over_cpu']}x")
    return res


def _scene(cache, key):
    """Builds (once) and makes resident one of the named benchmark scenes."""
    import libsvo_amd as svo
    from libsvo_amd import workloads
    if key not in cache:
        t0 = time.perf_counter()
        sc = {"terrain11": lambda: workloads.terrain_scene(depth=11),
              "terrain12": lambda: workloads.terrain_scene(depth=12),
              "spheres10": lambda: workloads.sphere_cloud_scene(depth=10),
              "menger9": lambda: workloads.menger_scene(depth=9, treelet_bytes=32768)}[key]()
        mm = svo.TreeletMemoryManagerHip(sc, workloads.incore_bytes_for(sc), device=0, resident_all=True)
        log(f"{key}: built and resident in {time.perf_counter() - t0:.1f}s, {sc.info['num_treelets']} treelets, "
            f"{mm.info['device_bytes'] / 2**30:.2f} GiB in HBM")
        cache[key] = (sc, mm)
    return cache[key]


def batched_section(cache, stream, policy, launches=10):
    """64 views over two scenes in one launch, replayed from a hipGraph."""
    import libsvo_amd as svo
    from libsvo_amd import workloads
    kinds = ["terrain", "spheres"]
    mms = [_scene(cache, "terrain11")[1], _scene(cache, "spheres10")[1]]
    spec = workloads.batched_views(kinds)
    views = [svo.View(mms[k], cam, svo.Frame.alloc(cam.width, cam.height)) for k, cam in spec]
    rays = sum(v.camera.width * v.camera.height for v in views)
    r = svo.Renderer(0)
    default_ms = r.timed(views, policy, stream, iters=3)
    t0 = time.perf_counter()
    shape, tuned_ms = r.autotune(views, policy, stream, iters=2)
    log(f"batched: {len(views)} views, {rays} rays; default shape {default_ms:.3f} ms, autotuned {shape} "
        f"{tuned_ms:.3f} ms (tuning took {time.perf_counter() - t0:.1f}s)")
    b = svo.Batch(r, views, policy, stream)
    b.launch(stream)
    stream.synchronize()
    t0 = time.perf_counter()
    for _ in range(launches):
        b.launch(stream)
    stream.synchronize()
    graph_ms = (time.perf_counter() - t0) * 1e3 / launches
    # the same views, one launch per view
    t0 = time.perf_counter()
    for v in views:
        r.render([v], policy, stream)
    stream.synchronize()
    single_ms = (time.perf_counter() - t0) * 1e3
    log(f"batched: graph replay {graph_ms:.3f} ms/launch = {rays / graph_ms / 1e3:.1f} Mrays/s; one launch per "
        f"view {single_ms:.3f} ms = {rays / single_ms / 1e3:.1f} Mrays/s")
    return {
        "workload": "64 views, depth-11 terrain + depth-10 sphere cloud (2000 spheres), resolutions uniform over "
                    "{256^2, 512^2, 1280x720, 1920x1080}, one hipGraph-captured launch",
        "views": len(views), "rays_per_launch": rays, "launch_shape": list(shape),
        "ms_per_launch": round(graph_ms, 4), "mrays_per_s": round(rays / graph_ms / 1e3, 1),
        "default_shape_ms": round(default_ms, 4), "autotuned_ms": round(tuned_ms, 4),
        "one_launch_per_view_mrays_per_s": round(rays / single_ms / 1e3, 1),
        "seed": workloads.BATCH_SEED,
    }


def _ooc_flight(mm, cams, policy, divide, renderer, stream, fused=False):
    """One fly-through from the manager's current residency, one stage
    after the other per frame: the analyse pass (or, fused, nothing: the
    previous frame's render recorded feedback for every divide-th pixel per
    axis), the feedback update (visibility, page-ins, evictions), the upload
    and the render, each timed."""
    import numpy as np
    import libsvo_amd as svo
    from libsvo_amd import workloads
    rng = np.random.default_rng(workloads.FLIGHT_SEED)
    frame = svo.Frame.alloc(cams[0].width, cams[0].height)
    view = [svo.View(mm, cams[0], frame, feedback_stride=divide if fused else 0)]
    pages, evictions, failed, upload_bytes, moved_bytes = [], [], [], [], []
    t_feedback = t_analyse = t_upload = t_render = 0.0
    wall0 = time.perf_counter()
    for i, cam in enumerate(cams):
        a = time.perf_counter()
        jit = tuple(float(v) for v in rng.uniform(-1.0, 1.0, 2))
        if not fused:
            mm.performAnalysePass(cam, divide, jit, policy, stream)
            stream.synchronize()
        a2 = time.perf_counter()
        st = mm.updateClientSideIncoreBuffer(stream)
        b = time.perf_counter()
        up = mm.updateDeviceMemory(stream)
        c = time.perf_counter()
        view[0].camera = cam
        renderer.render(view, policy, stream)
        stream.synchronize()
        d = time.perf_counter()
        t_feedback += b - a2  # readback and host update, without the analyse pass
        t_analyse += a2 - a
        t_upload += c - b
        t_render += d - c
        pages.append(st["inserted"])
        evictions.append(st["evicted"])
        failed.append(st["insert_failed"])
        upload_bytes.append(up["bytes"])
        moved_bytes.append(up["bytes_moved"])
        if i % 100 == 0:
            log(f"ooc frame {i}{' (fused feedback)' if fused else ''}: {st['inserted']} page-ins, {st['evicted']} "
                f"evictions, {mm.info['resident_treelets']} resident, {up['bytes'] / 2**20:.1f} MiB uploaded")
    wall = time.perf_counter() - wall0
    n = len(cams)
    return {
        "frames": n, "frames_per_s": round(n / wall, 1), "ms_per_frame": round(wall * 1e3 / n, 3),
        "ms_feedback": round(t_feedback * 1e3 / n, 3), "ms_analyse_pass": round(t_analyse * 1e3 / n, 3),
        "ms_upload": round(t_upload * 1e3 / n, 3), "ms_render": round(t_render * 1e3 / n, 3),
        "page_ins_per_frame": round(float(np.mean(pages)), 2), "page_ins_max": int(max(pages)),
        "evictions_per_frame": round(float(np.mean(evictions)), 2), "evictions_total": int(sum(evictions)),
        "insert_failed_total": int(sum(failed)),
        "upload_mib_per_frame": round(float(np.mean(upload_bytes)) / 2**20, 2),
        "upload_mib_moved_per_frame": round(float(np.mean(moved_bytes)) / 2**20, 2),
        "resident_treelets_end": mm.info["resident_treelets"],
    }, pages


def ooc_section(policy, frames, depth=13, treelet_bytes=131072, pool_bytes=8 << 30, small_pool_bytes=1 << 30,
                divide=8):
    """Out-of-core fly-through: analyse -> feedback update -> upload -> render
    per frame, starting from the root treelet alone; then the same flight
    pipelined, with feedback recorded by the render itself, and behind a
    pool smaller than the flight's working set (evictions every frame)."""
    import libsvo_amd as svo
    from libsvo_amd import workloads
    t0 = time.perf_counter()
    sc = workloads.terrain_scene(depth=depth, treelet_bytes=treelet_bytes)
    info = sc.info
    N = info["nodes_per_treelet"]
    slots = pool_bytes // (N * 16)  # nodes (8 B) + colour (4 B) + normal (4 B) per node
    mm = svo.TreeletMemoryManagerHip(sc, slots * info["treelet_bytes"], device=0)
    log(f"ooc: depth-{depth} terrain built in {time.perf_counter() - t0:.1f}s: {info['num_treelets']} treelets "
        f"({info['num_treelets'] * treelet_bytes / 2**30:.1f} GiB of nodes), pool {slots} slots "
        f"({mm.info['device_bytes'] / 2**30:.2f} GiB in HBM)")
    cams = workloads.flythrough_cameras(frames, lod=1.0)
    r = svo.Renderer(0)
    s = svo.Stream()
    res, pages = _ooc_flight(mm, cams, policy, divide, r, s)
    res["workload"] = (f"terrain depth {depth}, {N}-node treelets ({info['num_treelets']} treelets, "
                       f"{info['num_treelets'] * treelet_bytes / 2**30:.1f} GiB of nodes), {slots}-slot pool "
                       f"({mm.info['device_bytes'] / 2**30:.2f} GiB HBM), {frames} frames "
                       f"{cams[0].width}x{cams[0].height} along a seeded Catmull-Rom fly-through, LOD 1 pixel, "
                       f"analyse pass 1/{divide} per axis, starting from the root treelet")
    res["page_ins_after_frame_30"] = round(sum(pages[30:]) / len(pages[30:]), 2) if frames > 30 else None
    res["seed"] = workloads.FLIGHT_SEED
    log(f"ooc: {res['frames_per_s']} frames/s, {res['page_ins_per_frame']} page-ins/frame, "
        f"{res['upload_mib_per_frame']} MiB/frame in merged ranges ({res['upload_mib_moved_per_frame']} MiB moved), "
        f"render {res['ms_render']} ms, "
        f"feedback {res['ms_feedback']} ms (analyse pass {res['ms_analyse_pass']} ms), upload {res['ms_upload']} ms")
    resident_end = mm.info["resident_treelets"]
    # the same manager again, back to the root treelet alone (the reset's
    # upload zeroes what the fly-through wrote; not timed)
    mm.resetIncoreBuffer()
    mm.updateDeviceMemory(s)
    res["pipelined"] = _ooc_pipelined(mm, cams, divide, policy, pages, resident_end)
    log(f"ooc pipelined: {res['pipelined']}")
    # feedback recorded by each frame's render (every divide-th pixel per
    # axis, no jitter) drives the next frame's update: no analyse launch
    mm.resetIncoreBuffer()
    mm.updateDeviceMemory(s)
    fused, _ = _ooc_flight(mm, cams, policy, divide, r, s, fused=True)
    fused["note"] = (f"feedback recorded during the render of frame i (every {divide}th pixel per axis, no jitter) "
                     f"drives the update before frame i+1; the reference's separate jittered 1/{divide} analyse "
                     f"pass is not run, so residency lags one frame and differs from the line above")
    res["fused_feedback"] = fused
    log(f"ooc fused feedback: {fused}")
    del mm
    # a pool smaller than the flight's working set: freeIncorePosition evicts
    small = small_pool_bytes // (N * 16)
    mm2 = svo.TreeletMemoryManagerHip(sc, small * info["treelet_bytes"], device=0)
    ev, _ = _ooc_flight(mm2, cams, policy, divide, r, s)
    ev["workload"] = (f"the same flight behind a {small}-slot pool ({mm2.info['device_bytes'] / 2**30:.2f} GiB HBM), "
                      f"smaller than the flight's working set")
    res["evicting_pool"] = ev
    log(f"ooc evicting pool: {ev}")
    return res


def _ooc_pipelined(mm, cams, divide, policy, seq_pages, seq_resident_end):
    """The same fly-through with frame i's render overlapping the host
    residency update of frame i+1, on one stream: after the upload of frame
    i (the host waits for it), the analyse pass of frame i+1, the readback of
    its records (fetchFeedback) and render i are enqueued back to back; the
    host waits for the readback only and applies the records while render i
    runs.  Both frames read the residency frame i renders with, so every
    frame sees the same residency as in the sequential loop; the upload of
    frame i+1 follows render i on the stream."""
    import numpy as np
    import libsvo_amd as svo
    from libsvo_amd import workloads
    rng = np.random.default_rng(workloads.FLIGHT_SEED)
    jit = [tuple(float(v) for v in rng.uniform(-1.0, 1.0, 2)) for _ in cams]
    frame = svo.Frame.alloc(cams[0].width, cams[0].height)
    r = svo.Renderer(0)
    s = svo.Stream()
    view = [svo.View(mm, cams[0], frame)]
    pages = []
    wall0 = time.perf_counter()
    mm.performAnalysePass(cams[0], divide, jit[0], policy, s)
    st = mm.updateClientSideIncoreBuffer(s)
    for i, cam in enumerate(cams):
        mm.updateDeviceMemory(s)  # after render i-1 on the stream
        pages.append(st["inserted"])
        if i + 1 < len(cams):
            mm.performAnalysePass(cams[i + 1], divide, jit[i + 1], policy, s)
            mm.fetchFeedback(s)
        view[0].camera = cam
        r.render(view, policy, s)
        if i + 1 < len(cams):
            st = mm.updateClientSideIncoreBuffer(s)
    s.synchronize()
    wall = time.perf_counter() - wall0
    same = pages == list(seq_pages) and mm.info["resident_treelets"] == seq_resident_end
    return {"frames_per_s": round(len(cams) / wall, 1), "ms_per_frame": round(wall * 1e3 / len(cams), 3),
            "same_page_ins_as_sequential": bool(same)}


def serving_section(cache, rates, seconds, policy):
    """Poisson request streams over four resident scenes: continuous batching
    against the plain one-frame-per-launch loop."""
    import libsvo_amd as svo
    from libsvo_amd import workloads
    keys = ["menger9", "terrain11", "spheres10", "terrain12"]
    kinds = ["menger", "terrain", "spheres", "terrain"]
    budget = 64 << 30
    r = svo.Renderer(0)
    modes = {"continuous": dict(max_batch_views=64, max_inflight=4, max_batch_rays=16 << 20),
             "plain": dict(max_batch_views=1, max_inflight=1)}
    servers = {}
    for name, kw in modes.items():
        srv = svo.Server(r, hbm_budget=budget, frame_pool=4 << 30, policy=policy, **kw)
        for k in keys:
            srv.add_scene(_scene(cache, k)[1])
        servers[name] = srv
    scene_bytes = sum(_scene(cache, k)[1].info["device_bytes"] for k in keys)
    def row_of(st):
        return {"frames_per_s": round(st["frames_per_s"], 1), "mrays_per_s": round(st["mrays_per_s"], 1),
                "p50_ms": round(st["p50_ms"], 3), "p99_ms": round(st["p99_ms"], 3), "max_ms": round(st["max_ms"], 3),
                "mean_batch_views": round(st["mean_batch_views"], 2), "max_batch_views": st["max_batch_views_seen"],
                "max_inflight": st["max_inflight_seen"]}

    results = []
    for rate in rates:
        reqs = workloads.serving_requests(kinds, rate, max(50, int(rate * seconds)))
        span = reqs[-1][2] - reqs[0][2]
        # the rate this sample of the Poisson process offers (requests over
        # the span of their arrivals)
        row = {"rate_req_s": rate, "requests": len(reqs),
               "offered_req_s": round(len(reqs) / span, 1) if span > 0 else None}
        for name, srv in servers.items():
            st, comp = srv.replay(reqs)
            row[name] = row_of(st)
            log(f"serving {rate:.0f} req/s {name}: {row[name]}")
        results.append(row)
    # saturation: every request queued at t=0 (latency is then mostly queueing)
    reqs = workloads.serving_requests(kinds, 1e9, 2000)
    sat = {}
    for name, srv in servers.items():
        st, _ = srv.replay(reqs)
        sat[name] = row_of(st)
        log(f"serving saturation {name}: {sat[name]}")
    # the highest rate of the sweep each loop keeps up with (throughput
    # within 5 % of the offered rate) at p99 <= 20 ms
    sustained = {}
    for name in servers:
        ok = [row["rate_req_s"] for row in results
              if row[name]["p99_ms"] <= 20.0 and row["offered_req_s"]
              and row[name]["frames_per_s"] >= 0.95 * row["offered_req_s"]]
        sustained[name] = max(ok) if ok else None
    return {
        "workload": "Menger d9 + terrain d11 + sphere cloud d10 + terrain d12 resident, Poisson arrivals, "
                    "resolutions 256^2 40% / 720p 40% / 1080p 15% / 4K 5%, random cameras; latency = "
                    "completion on device - arrival",
        "hbm_budget_bytes": budget, "scene_bytes": scene_bytes, "rates": results, "saturation": sat,
        "max_rate_req_s_at_p99_le_20ms": sustained,
        "seed": workloads.SERVING_SEED, "seconds_per_rate": seconds,
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
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--no-ooc", action="store_true")
    ap.add_argument("--no-reference-case", action="store_true",
                    help="skip the reference's own case (Menger depth 8 in one treelet, 512x512)")
    ap.add_argument("--ooc-frames", type=int, default=300)
    ap.add_argument("--serving-seconds", type=float, default=2.0, help="length of each request stream")
    ap.add_argument("--serving-rates", default="200,1000,2000,4000,6000,8000,10000,12000,14000",
                    help="Poisson arrival rates (req/s)")
    ap.add_argument("--launch-shape", default="",
                    help="tile_width,waves,stack (e.g. 8,1,lds): use this shape for the headline instead of the "
                         "autotuner's pick (profiling runs)")
    ap.add_argument("--valu-calibration-blocks", type=int, default=0,
                    help="after the timed launches, run the VALU issue probe with this many 256-lane "
                         "workgroups (the VALU-busy reading of a kernel that only issues VALU instructions)")
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
    view = [svo.View(mm, cam, frame)]
    renderer = svo.Renderer(0)
    if args.launch_shape:
        t, w, k = args.launch_shape.split(",")
        renderer.config = (int(t), int(w), k)
        shape = renderer.config
        log(f"launch shape {shape} (given)")
    else:
        tune_cache = svo.TuneCache()
        t0 = time.perf_counter()
        shape, tune_ms = renderer.autotune(view, args.policy, stream, iters=20, cache=tune_cache)  # shapes lie within a few %
        t1 = time.perf_counter()
        renderer.autotune(view, args.policy, stream, iters=20, cache=tune_cache)  # the same class again: a cache hit
        t2 = time.perf_counter()
        autotune = {"key": renderer.tune_key(view, args.policy), "shape": list(shape),
                    "tuning_s": round(t1 - t0, 3), "cached_lookup_ms": round((t2 - t1) * 1e3, 3),
                    "second_call_hit": tune_cache.last_hit}
        log(f"autotuned launch shape {shape} ({tune_ms:.4f} ms in tuning, {t1 - t0:.2f} s); cache hit "
            f"{tune_cache.last_hit} in {(t2 - t1) * 1e3:.2f} ms")
    if args.warmup > 0:
        renderer.timed(view, args.policy, stream, iters=args.warmup)
    svo._capi.check(svo.lib.svo_device_synchronize(), "sync")
    w0 = time.perf_counter()
    ms = renderer.timed(view, args.policy, stream, iters=args.steps)
    svo._capi.check(svo.lib.svo_device_synchronize(), "sync")
    wall = time.perf_counter() - w0
    value = rays / (ms / 1e3) / 1e6
    achieved = alg_bytes / (ms / 1e3) / 1e9
    log(f"{args.steps} launches: {ms:.4f} ms/launch (events), wall {wall * 1e3 / args.steps:.4f} ms/launch, "
        f"{value:.1f} Mrays/s, {achieved:.1f} GB/s algorithmic")

    # the other traversal policy on the same frame (autotuned on its own)
    other = "cpu3" if args.policy == "treelet" else "treelet"
    if args.launch_shape:
        oshape = shape
    else:
        oshape, _ = renderer.autotune(view, other, stream, iters=20)
    ms_other = renderer.timed(view, other, stream, iters=max(10, args.steps // 2))
    log(f"{other} policy: {oshape} {ms_other:.4f} ms/launch = {rays / ms_other / 1e3:.1f} Mrays/s")
    policies = {args.policy: {"mrays_per_s": round(value, 1), "ms_per_launch": round(ms, 5), "launch_shape": list(shape)},
                other: {"mrays_per_s": round(rays / ms_other / 1e3, 1), "ms_per_launch": round(ms_other, 5),
                        "launch_shape": list(oshape)}}

    if args.valu_calibration_blocks > 0:
        blocks = args.valu_calibration_blocks
        buf = svo.DeviceBuffer(blocks * 256 * 4)
        svo._capi.check(svo.lib.svo_debug_valu(buf.ptr, blocks, 4096, stream.handle), "svo_debug_valu")
        stream.synchronize()
        log(f"VALU probe of {blocks} workgroups x 4096 iterations done")
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
            "launch_shape": list(shape),
        },
        "policies": policies,
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 2),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 5),
            "traffic": profiled_traffic(shape),
            "issue": profiled_issue(shape),
        },
    }
    traffic = out["roofline"]["traffic"]
    # the counters' HBM bytes per launch (profile) over this run's launch time,
    # and the L2 hit rate of the same shape
    out["roofline"]["hbm_gbs_measured"] = round(traffic / (ms / 1e3) / 1e9, 1) if traffic else None
    out["roofline"]["l2_hit_rate"] = profiled_l2_hit(shape)
    copy_gbs = stream_copy_gbs(stream)
    out["roofline"]["measured_copy_gbs"] = round(copy_gbs, 1)
    out["roofline"]["frac_of_measured_copy"] = round(achieved / copy_gbs, 5)
    log(f"stream copy {copy_gbs:.0f} GB/s (read + write); algorithmic rate {achieved / copy_gbs:.3f} of it")
    if not args.launch_shape:
        out["autotune"] = autotune
    # both policies' frames (untimed) for the CPU leg's pixel check
    gpu_frames = {}
    for pol in ("treelet", "cpu3"):
        f = svo.Frame.alloc(cam.width, cam.height)
        svo.render_to_buffer(mm, [cam], [f], pol, stream)
        stream.synchronize()
        gpu_frames[pol] = f.fetch()
    cpus = host_cpus()
    extra_scenes = {"terrain11": (scene, mm)}
    if not args.no_reference_case:
        out["reference_case"] = reference_case_section(stream, cpus, cpu=not args.no_cpu_baseline,
                                                       shape=shape if args.launch_shape else None)
    if not args.no_batched:
        out["batched"] = batched_section(extra_scenes, stream, args.policy)
    if not args.no_ooc:
        out["ooc"] = ooc_section(args.policy, args.ooc_frames)
    if not args.no_serving:
        rates = [float(r) for r in args.serving_rates.split(",") if r]
        out["serving"] = serving_section(extra_scenes, rates, args.serving_seconds, args.policy)
    if not args.no_cpu_baseline:
        log(f"cpu baseline on {cpus['threads']} threads ({cpus})")
        ora = _oracle_scene(scene)
        out["cpu_baseline"] = cpu_baseline(scene, cam, cpus, ora=ora)
        out["gpu_over_cpu"] = round(value / out["cpu_baseline"]["value"], 1)
        out["gpu_over_single_cpu_thread"] = round(value / out["cpu_baseline"]["single_thread_mrays_per_s"], 1)
        log(f"cpu baseline {out['cpu_baseline']['value']:.2f} Mrays/s; GPU/CPU = {out['gpu_over_cpu']}x")
        out["boundary"] = cpu_parity(ora, cam, gpu_frames, cpus["threads"])
        log(f"boundary rays and differing pixels of the headline frame: {out['boundary']}")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
