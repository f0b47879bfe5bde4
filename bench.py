#!/usr/bin/env python3
"""Ray-query throughput on one MI355X: incoherent rays at 1.31M triangles.

Workload: 8,388,608 cosine-weighted bounce rays from the primary hits of a
1920x1080 view of a displaced icosphere (level 8, 1,310,720 triangles) above
a 2-triangle ground plane, split into 64 lognormal batches (median 65,536,
clipped to 4k..1M).  A step traces every batch once, in one persistent
hipGraph launch, through the SAH BVH4 with ENCORE's predicate; rays and hit
buffers are resident in HBM before timing.

Prints one JSON line: Mrays/s; the roofline of the traversal kernel
(algorithmic bytes per ray = 32 + 16 + 128 N_node + 48 N_tri, counted by an
instrumented copy of the kernel); the CPU baseline (ENCORE's kd-tree
algorithm restated in oracle/, all host threads, on a strided ray sample);
parity of the timed launch's own hits with that baseline on the sample
(every difference adjudicated) and bitwise against one launch per batch on
every ray; the FP64 error of a ~1% sample; the exact path (ENCORE's kd-tree
on the GPU, bit for bit) in the same persistent launch.

    python bench.py [--steps K] [--warmup W] [--mode plain|batched] [--no-cpu]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--mode", choices=["plain", "batched"], default="batched")
    ap.add_argument("--width", type=int, choices=[2, 4], default=4, help="BVH children per node")
    ap.add_argument("--leaf", type=int, default=0, help="BVH leaf size 1..16 (0 = the builder default, 4)")
    ap.add_argument("--level", type=int, default=8)
    ap.add_argument("--rays", type=int, default=8388608)
    ap.add_argument("--batches", type=int, default=64)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-sample", type=int, default=524288)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-counts", action="store_true")
    ap.add_argument("--no-exact", action="store_true", help="skip the bit-exact kd-path measurement")
    ap.add_argument("--dump", default="", help="write the adjudicated parity differences to this .npz")
    ap.add_argument("--workload", choices=["c3", "c1", "c2", "c4", "c5", "fg", "ic", "pt", "refit"], default="c3",
                    help="c3 (default): 8M incoherent rays at 1.31M triangles; c1: Cornell box eye rays; c2: "
                         "1080p eye + shadow rays at 1.31M triangles; c4: photon gather; c5: serving; fg: final "
                         "gather; ic: irradiance-cache lookups; pt: photon emission; refit: scene update")
    ap.add_argument("--pt-sample", type=int, default=100000,
                    help="pt: global-map capacity of the CPU baseline and parity sample")
    ap.add_argument("--seed", type=int, default=7, help="pt: seed of the emission draws")
    ap.add_argument("--requests", type=int, default=1000, help="c5: requests per arrival-rate point")
    ap.add_argument("--copy-probe", type=int, default=0,
                    help="run one device-to-device copy of this many bytes first (counter calibration)")
    ap.add_argument("--batch-config", default="", help="stack,blocks_per_cu,fetch_threshold: skip autotuning")
    ap.add_argument("--no-serving", action="store_true", help="c3: skip the config-5 serving figures")
    return ap.parse_args()


def cpu_threads(req):
    """Threads of the CPU baseline: all CPUs in this process's affinity, up
    to the 16 host cores a one-GPU share of the MI355X box provides (the
    line also states the affinity count)."""
    n = len(os.sched_getaffinity(0))
    if req > 0:
        return req
    return max(1, min(16, n))


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py measures one GPU: ray queries need no collective, so every configuration here runs on a "
              "single MI355X (--gpus %d is not supported)" % args.gpus, file=sys.stderr)
        return 2
    if args.workload != "c3":
        return run_workload(args)
    import numpy as np
    import torch

    from encore_amd import rq, scenes

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    torch.cuda.init()
    dev = rq.Device(0)
    log("device open:", torch.cuda.get_device_name(0))

    t = time.time()
    xyz, idx = scenes.config2_scene(level=args.level)
    log("scene: %d triangles (%.1fs)" % (len(idx), time.time() - t))
    t = time.time()
    scene = rq.Scene(dev, xyz, idx, accel="bvh", predicate="encore", width=args.width, leaf_size=args.leaf)
    st = scene.stats()
    log("bvh%d: %d nodes, depth %d, %.1f MB, build %.2fs" % (args.width, st["nnodes"], st["max_depth"],
                                                            st["device_bytes"] / 1e6, st["build_seconds"]))
    prim = scenes.config2_primary()
    d_prim = torch.from_numpy(prim).cuda()
    d_ph = torch.empty((len(prim), 4), dtype=torch.float32, device="cuda")
    scene.intersect(d_prim, d_ph)
    torch.cuda.synchronize()
    ph = d_ph.cpu().numpy()
    del d_prim, d_ph
    t = time.time()
    rays, sizes = scenes.config3_rays(xyz, idx, prim, ph, count=args.rays, nbatches=args.batches)
    log("rays: %d in %d batches (%d..%d), generated in %.1fs" % (len(rays), len(sizes), sizes.min(), sizes.max(),
                                                                 time.time() - t))
    N = len(rays)
    d_rays = torch.from_numpy(rays).cuda()
    # The timed launch writes d_hits; the plain per-batch leg has its own
    # buffer, so parity below reads what the timed launch produced.
    d_hits = torch.full((N, 4), -7.0, dtype=torch.float32, device="cuda")
    d_hits_plain = torch.full((N, 4), -9.0, dtype=torch.float32, device="cuda")
    offs = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    stream = torch.cuda.current_stream()
    nb = len(sizes)

    def events(n):
        return [torch.cuda.Event(enable_timing=True) for _ in range(n)]

    def plain_step(timed=None, out=None):
        # timed: a list of 2 * nb events (start, end per launch), or None
        out = d_hits if out is None else out
        for b in range(nb):
            o, n = int(offs[b]), int(sizes[b])
            if timed:
                timed[2 * b].record(stream)
            scene.intersect(d_rays[o:o + n], out[o:o + n], stream=stream)
            if timed:
                timed[2 * b + 1].record(stream)
        return nb

    if args.copy_probe:
        src = torch.empty(args.copy_probe // 4, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        src.fill_(1.0)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        log("copy probe: %d bytes" % args.copy_probe)
        del src, dst
    batch = None
    tuned = None
    if args.mode == "batched":
        batch = rq.Batch(dev)
        batch.set([(scene, d_rays[int(offs[b]):int(offs[b + 1])], d_hits[int(offs[b]):int(offs[b + 1])])
                   for b in range(nb)])
        t = time.time()
        if args.batch_config:
            st_, bpc_, th_ = (int(x) for x in args.batch_config.split(","))
            batch.set_config(st_, bpc_, th_)
            tuned = dict(batch.config(), ms=None)
        else:
            tuned = batch.autotune(stream=stream)
        log("launch config in %.1fs: %s" % (time.time() - t, tuned))

    def batched_step(timed=None):
        if timed:
            timed[0].record(stream)
        batch.launch(graph=True, stream=stream)
        if timed:
            timed[1].record(stream)
        return 1

    def run(step, steps, warmup, launches):
        # The steps run back to back (each with its own events), bracketed
        # by a synchronisation on both sides; the per-launch event times are
        # read after the second one.
        for _ in range(warmup):
            step()
        torch.cuda.synchronize()
        evs = [events(2 * launches) for _ in range(steps)]
        t0 = time.perf_counter()
        for i in range(steps):
            step(timed=evs[i])
        torch.cuda.synchronize()
        secs = time.perf_counter() - t0
        launch_ms = [e[2 * b].elapsed_time(e[2 * b + 1]) for e in evs for b in range(launches)]
        return secs, launch_ms

    step = batched_step if args.mode == "batched" else plain_step
    secs, launch_ms = run(step, args.steps, args.warmup, 1 if args.mode == "batched" else nb)
    # wall time over the steps, host launches included
    ms_per_step = secs * 1000.0 / args.steps
    value = N * args.steps / secs / 1e6
    kernel_ms_total = float(np.sum(launch_ms)) / args.steps
    launches_per_step = len(launch_ms) // args.steps
    log("timed (%s): %.2f ms/step, %.1f Mrays/s; launch time %.2f ms/step over %d launches" % (
        args.mode, ms_per_step, value, kernel_ms_total, launches_per_step))
    hits = d_hits.cpu().numpy()
    plain_value = None
    batched_equals_plain = None
    if args.mode == "batched":
        pstep = lambda timed=None: plain_step(timed, out=d_hits_plain)
        psecs, _ = run(pstep, max(1, min(args.steps, 3)), 1, nb)
        plain_value = N * max(1, min(args.steps, 3)) / psecs / 1e6
        # Bitwise, every ray: the timed persistent launch against one launch
        # per batch of the same kernel family.
        batched_equals_plain = bool(torch.equal(d_hits.view(torch.int32), d_hits_plain.view(torch.int32)))
        log("plain per-batch launches: %.1f Mrays/s; batched == plain on all %d rays: %s" % (
            plain_value, N, batched_equals_plain))
    del d_hits_plain

    exact = None
    kd_hits = None
    kd_batched_hits = None
    if not args.no_exact:
        # The bit-exact drop-in: ENCORE's own kd-tree traversed on the GPU.
        t = time.time()
        kscene = rq.Scene(dev, xyz, idx, accel="kd", predicate="encore")
        ks = kscene.stats()
        log("kd: %d nodes, %d refs, depth %d, build %.2fs" % (ks["nnodes"], ks["nrefs"], ks["max_depth"],
                                                            time.time() - t))
        d_kh = torch.empty((N, 4), dtype=torch.float32, device="cuda")
        kscene.intersect(d_rays, d_kh, stream=stream)
        torch.cuda.synchronize()
        kt0 = time.perf_counter()
        ksteps = max(1, min(args.steps, 3))
        for _ in range(ksteps):
            for b in range(nb):
                o, n = int(offs[b]), int(sizes[b])
                kscene.intersect(d_rays[o:o + n], d_kh[o:o + n], stream=stream)
        torch.cuda.synchronize()
        ksecs = time.perf_counter() - kt0
        kd_hits = d_kh.cpu().numpy()
        # the same rays in one launch (the batches' tails overlap)
        kscene.intersect(d_rays, d_kh, stream=stream)
        torch.cuda.synchronize()
        kt1 = time.perf_counter()
        for _ in range(ksteps):
            kscene.intersect(d_rays, d_kh, stream=stream)
        torch.cuda.synchronize()
        k1secs = time.perf_counter() - kt1
        assert np.array_equal(d_kh.cpu().numpy().view(np.uint32), kd_hits.view(np.uint32))
        # The 64 batches in one persistent launch (replayed hipGraph,
        # autotuned), as the BVH headline runs them.
        d_kb = torch.full((N, 4), -5.0, dtype=torch.float32, device="cuda")
        kb = rq.Batch(dev)
        kb.set([(kscene, d_rays[int(offs[b]):int(offs[b + 1])], d_kb[int(offs[b]):int(offs[b + 1])])
                for b in range(nb)])
        t = time.time()
        ktuned = kb.autotune(stream=stream)
        log("kd batch config in %.1fs: %s" % (time.time() - t, ktuned))
        d_kb.fill_(-5.0)
        kb.launch(graph=True, stream=stream)
        torch.cuda.synchronize()
        kt2 = time.perf_counter()
        for _ in range(ksteps):
            kb.launch(graph=True, stream=stream)
        torch.cuda.synchronize()
        kbsecs = time.perf_counter() - kt2
        kd_batched_hits = d_kb.cpu().numpy()
        kd_batched_equals_plain = bool(np.array_equal(kd_batched_hits.view(np.uint32), kd_hits.view(np.uint32)))
        kb.destroy()
        del d_kb
        kn = kscene.counts(d_rays) if not args.no_counts else None
        exact = {"accel": "kd", "value": round(N * ksteps / kbsecs / 1e6, 2), "unit": "Mrays/s",
                 "ms_per_step": round(kbsecs * 1000 / ksteps, 2),
                 "mode": "one persistent hipGraph launch per step, autotuned", "autotuned": ktuned,
                 "batched_equals_per_batch_launches": kd_batched_equals_plain,
                 "per_batch_launches_value": round(N * ksteps / ksecs / 1e6, 2),
                 "one_launch_value": round(N * ksteps / k1secs / 1e6, 2)}
        if kn:
            exact["nodes_per_ray"] = round(kn[1] / kn[0], 2)
            exact["tris_per_ray"] = round(kn[2] / kn[0], 2)
        log("kd exact path: %.1f Mrays/s batched (graph), %.1f per batch, %.1f in one launch; batched == per-batch: %s"
            % (exact["value"], exact["per_batch_launches_value"], exact["one_launch_value"], kd_batched_equals_plain))
        del d_kh
        kscene.destroy()

    roof = None
    if not args.no_counts:
        n_rays, n_nodes, n_tris = scene.counts(d_rays)
        node_bytes = 128 if args.width == 4 else 64
        alg_bytes = n_rays * (32 + 16) + node_bytes * n_nodes + 48 * n_tris
        achieved = alg_bytes / (kernel_ms_total / 1000.0) / 1e9
        traffic = None
        tf = os.path.join(ROOT, "profiles", "bvh_trace_config3_traffic.json")
        if os.path.exists(tf):
            with open(tf) as f:
                tj = json.load(f).get("by_width", {}).get(str(args.width))
            # Counters were read for the batched launch of each node width.
            if tj and args.mode == "batched":
                traffic = tj.get("hbm_bytes_per_launch")
        roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                "alg_bytes_per_launch": int(alg_bytes / launches_per_step), "nodes_per_ray": round(n_nodes / n_rays, 2),
                "tris_per_ray": round(n_tris / n_rays, 2),
                "avg_launch_ms": round(kernel_ms_total / launches_per_step, 4),
                "node_bytes": node_bytes,
                "kernel": "bvh_trace<PredEncore, closest, stack %s, batched, width %d>"
                % (tuned["stack"] if tuned else 16, args.width)
                if args.mode == "batched" else "bvh_trace<PredEncore, closest, stack 16, width %d>" % args.width}
        log("roofline: %.1f nodes/ray, %.1f tris/ray, %.1f GB/s algorithmic" % (n_nodes / n_rays, n_tris / n_rays,
                                                                             achieved))

    cpu = None
    parity = None
    fp64 = None
    if not args.no_cpu:
        import oracle

        thr = cpu_threads(args.cpu_threads)
        stride = max(1, N // args.cpu_sample)
        sample = np.ascontiguousarray(rays[::stride][: args.cpu_sample])
        t = time.time()
        kd = oracle.KdTree(xyz, idx)
        log("oracle kd-tree built in %.1fs" % (time.time() - t))
        small = np.ascontiguousarray(sample[:8192])
        kd.intersect(small[:256], nthreads=thr)  # warm
        t = time.perf_counter()
        kd.intersect(small, nthreads=1)
        one = len(small) / (time.perf_counter() - t) / 1e6
        t = time.perf_counter()
        ref = kd.intersect(sample, nthreads=thr)
        allc = len(sample) / (time.perf_counter() - t) / 1e6
        log("cpu kd: %.4f Mrays/s on 1 thread, %.4f Mrays/s on %d threads" % (one, allc, thr))
        cpu = {"value": round(allc, 5), "unit": "Mrays/s", "cores": thr, "kind": "port",
               "host_cpus_in_affinity": len(os.sched_getaffinity(0)),
               "sample": "%d of the %d bounce rays (every %dth), ENCORE kd-tree (Kdtree.cpp) restated, %d threads; "
                         "1 thread: %.5f Mrays/s on %d rays" % (len(sample), N, stride, thr, one, len(small)),
               "one_core": round(one, 5)}
        # Parity of the timed launch's own output on the sample, every
        # difference adjudicated (brute force over all triangles, then FP64).
        g = hits[::stride][: args.cpu_sample]
        parity = {"sample": len(sample), "checked": "batched (timed) launch output" if args.mode == "batched"
                  else "plain launch output"}
        parity["bvh"] = _classify(g, ref, xyz, idx, sample, thr, dump=args.dump)
        parity["bvh"]["prim_equal"] = int((g[:, 3].view(np.uint32) == ref[:, 3].view(np.uint32)).sum())
        if batched_equals_plain is not None:
            parity["batched_equals_plain"] = batched_equals_plain
        if kd_hits is not None:
            kg = kd_hits[::stride][: args.cpu_sample]
            parity["kd"] = {"bit_equal": int(np.all(kg.view(np.uint32) == ref.view(np.uint32), axis=1).sum())}
        if kd_batched_hits is not None:
            kg = kd_batched_hits[::stride][: args.cpu_sample]
            parity["kd_batched"] = {"bit_equal": int(np.all(kg.view(np.uint32) == ref.view(np.uint32), axis=1).sum())}
        log("parity on sample:", parity)
        # FP64 brute force (box-culled, identical results) on a ~1% sample:
        # every 6th ray of the parity sample.
        sub = slice(None, None, 6)
        t = time.time()
        fp64 = {"sample": "every %dth ray (%d of %d)" % (6 * stride, len(sample[sub]), N),
                "gpu": _fp64_report(xyz, idx, np.ascontiguousarray(sample[sub]), g[sub], thr),
                "reference_cpu": _fp64_report(xyz, idx, np.ascontiguousarray(sample[sub]), ref[sub], thr)}
        log("fp64 1%% sample (%.1fs): %s" % (time.time() - t, fp64))

    out = {
        "metric": "Mrays/s incoherent @1M tris; p50/p99 batch latency, served vs plain launch",
        "value": round(value, 2),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: seeded displaced icosphere and seeded bounce rays",
        "config": {"workload": "8,388,608 incoherent bounce rays in %d lognormal batches on a displaced icosphere "
                               "(level %d) + ground, %d triangles; BVH%d SAH, ENCORE predicate; %s" % (
                                   nb, args.level, len(idx), args.width,
                                   "one launch per batch" if args.mode == "plain" else
                                   "one persistent hipGraph launch per step, autotuned"),
                   "rays": N, "batches": nb, "triangles": int(len(idx)), "accel": "bvh%d" % args.width,
                   "predicate": "encore",
                   "mode": args.mode, "autotuned": tuned},
        "roofline": roof,
        "cpu_baseline": cpu,
        "parity": parity,
        "fp64_1pct": fp64,
    }
    if plain_value is not None:
        out["plain_per_batch_value"] = round(plain_value, 2)
    if exact:
        out["exact_path"] = exact
    if cpu:
        out["speedup_vs_cpu_all_cores"] = round(value / cpu["value"], 1)
        if exact:
            exact["speedup_vs_cpu_all_cores"] = round(exact["value"] / cpu["value"], 1)
    if batch is not None:
        batch.destroy()
    scene.destroy()
    del d_rays, d_hits
    torch.cuda.empty_cache()
    if not args.no_serving:
        # The serving half of the metric (config 5): served vs plain launch,
        # Mrays/s and p50/p99 latency, in the same line.
        res = _serving(args, dev, _trace_fn)
        res["config"] = res["config"]["workload"]
        out["serving"] = res
    print(json.dumps(out), flush=True)
    dev.close()
    return 0


# ---------------------------------------------------------------------------
# Other configurations (--workload): each prints one JSON line of its own.
# ---------------------------------------------------------------------------
def _line(metric, value, unit, ms, config, higher=True, **extra):
    out = {"metric": metric, "value": round(value, 3), "unit": unit, "n_gpus": 1, "ms_per_step": round(ms, 4),
           "higher_is_better": higher, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
           "data": "synthetic (seeded)", "config": config}
    out.update(extra)
    print(json.dumps(out), flush=True)


def _time_gpu(fn, steps, warmup):
    import torch

    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    s = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record(s)
    for _ in range(steps):
        fn()
    e1.record(s)
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / steps, e0.elapsed_time(e1) / steps


def _outside_box(xyz, idx, rays, t, prim):
    """Distance (world units) by which each hit point o + t d lies outside the
    bounding box of triangle `prim`, less a rounding allowance of 1e-4 of the
    point's magnitude; > 0 means a BVH, whose boxes enclose their triangles,
    cannot reach that hit."""
    import numpy as np

    tri = xyz[idx[prim]].astype(np.float64)
    P = rays[:, :3].astype(np.float64) + t.astype(np.float64)[:, None] * rays[:, 4:7].astype(np.float64)
    out = (np.maximum(tri.min(1) - P, 0) + np.maximum(P - tri.max(1), 0)).max(1)
    return out - 1e-4 * np.maximum(1.0, np.abs(P).max(1))


def _predicate_hit(xyz, idx, rays, g):
    """True where g's (t, prim) is exactly what ENCORE's predicate returns for
    that ray and triangle."""
    import numpy as np

    import oracle

    gp = np.ascontiguousarray(g[:, 3]).view(np.uint32)
    ok = gp != 0xFFFFFFFF
    res = np.zeros(len(g), bool)
    if ok.any():
        tri = np.ascontiguousarray(xyz[idx[gp[ok]]].reshape(-1, 9), np.float32)
        h = oracle.tri_intersect_pairs(np.ascontiguousarray(rays[ok]), tri)
        res[ok] = (h[:, 3] != 0) & (h[:, 0].view(np.uint32) == np.ascontiguousarray(g[ok, 0]).view(np.uint32))
    return res


def _adjudicate(xyz, idx, rays, g, ref, thr, dump=""):
    """Closest-hit differences that are not ties, every one decided by brute
    force over every triangle (ENCORE's predicate) and FP64: counts of kd
    misses the GPU finds, reference flush phantoms (the reference's triangle
    is not crossed in FP64), and unexplained rows."""
    import numpy as np

    import oracle

    out = {"kd_miss_gpu_matches_brute": 0, "reference_flush_phantom_fp64_off_triangle": 0,
           "reference_phantom_outside_its_triangle_box": 0, "unexplained": 0, "adjudicated": 0}
    if len(rays) == 0:
        return out
    rays = np.ascontiguousarray(rays)
    gp = np.ascontiguousarray(g[:, 3]).view(np.uint32)
    rp = np.ascontiguousarray(ref[:, 3]).view(np.uint32)
    br = oracle.brute_closest(xyz, idx, rays, nthreads=thr)
    bp = np.ascontiguousarray(br[:, 3]).view(np.uint32)
    fr = oracle.fp64_eval(xyz, idx, rays, rp)
    fg = oracle.fp64_eval(xyz, idx, rays, gp)

    def crossed(f):
        return (f[:, 1] >= 0) & (f[:, 2] >= 0) & (f[:, 1] + f[:, 2] <= 1) & (f[:, 0] > 0)

    ref_off = (rp != 0xFFFFFFFF) & ~crossed(fr)
    gpu_ok = (gp == 0xFFFFFFFF) | crossed(fg)
    # The reference's hit is a phantom lying outside its own triangle's box
    # (so no BVH can reach it) and the GPU's answer is an exact predicate hit
    # farther along, or no hit.
    ref_box_out = (rp != 0xFFFFFFFF) & (_outside_box(xyz, idx, rays, ref[:, 0], np.where(rp == 0xFFFFFFFF, 0, rp)) > 0)
    gpu_exact = (gp == 0xFFFFFFFF) | _predicate_hit(xyz, idx, rays, g)
    unexplained = []
    for i in range(len(rays)):
        if gp[i] == bp[i] or (gp[i] != 0xFFFFFFFF and bp[i] != 0xFFFFFFFF and
                              abs(g[i, 0] - br[i, 0]) <= 1e-6 * abs(br[i, 0])):
            out["kd_miss_gpu_matches_brute"] += 1
        elif ref_off[i] and gpu_ok[i]:
            # the reference (and brute force, same predicate) report a hit on a
            # triangle the ray does not cross; the GPU's answer is genuine
            out["reference_flush_phantom_fp64_off_triangle"] += 1
        elif ref_off[i] and ref_box_out[i] and gpu_exact[i]:
            out["reference_phantom_outside_its_triangle_box"] += 1
        else:
            out["unexplained"] += 1
            unexplained.append(i)
    out["adjudicated"] = int(len(rays))
    if dump and unexplained:
        u = np.array(unexplained)
        np.savez(dump, rays=rays[u], gpu=g[u], kd=ref[u], brute=br[u])
    return out


def _classify(g, ref, xyz=None, idx=None, rays=None, thr=1, dump=""):
    """Every ray of the sample in exactly one class: bit-equal, same prim with
    other bits (never expected: the predicate is ENCORE's), a tie (another
    genuine prim within 1e-6 t), or -- adjudicated ray by ray when the scene
    is given -- a kd miss, a reference flush phantom, a phantom outside its
    triangle's box, or unexplained.  The classes are checked to sum to n."""
    import numpy as np

    gp = np.ascontiguousarray(g[:, 3]).view(np.uint32)
    rp = np.ascontiguousarray(ref[:, 3]).view(np.uint32)
    bits = np.all(g.view(np.uint32) == ref.view(np.uint32), axis=1)
    diff = np.nonzero(~bits)[0]
    same_prim = gp[diff] == rp[diff]
    both = (gp[diff] != 0xFFFFFFFF) & (rp[diff] != 0xFFFFFFFF)
    with np.errstate(invalid="ignore"):
        is_tie = ~same_prim & both & (np.abs(g[diff, 0] - ref[diff, 0]) <= 1e-6 * np.abs(ref[diff, 0]))
    out = {"n": int(len(g)), "bit_equal": int(bits.sum()), "same_prim_other_bits": int(same_prim.sum()),
           "ties_1e-6t": int(is_tie.sum())}
    rest = diff[~same_prim & ~is_tie]
    if xyz is not None:
        out.update(_adjudicate(xyz, idx, rays[rest], g[rest], ref[rest], thr, dump=dump))
        parts = ("bit_equal", "same_prim_other_bits", "ties_1e-6t", "kd_miss_gpu_matches_brute",
                 "reference_flush_phantom_fp64_off_triangle", "reference_phantom_outside_its_triangle_box",
                 "unexplained")
        assert sum(out[k] for k in parts) == out["n"], out
    else:
        out["other"] = int(len(rest))
    return out


def _fp64_report(xyz, idx, rays, hits, thr):
    """Error of float32 hits against the FP64 closest hit over every triangle
    (no flush): for the rays whose prim equals the FP64 prim, the largest
    relative t error and the largest |du|, |dv|, also relative to
    max(|u|, 1e-3) as the 1e-5-relative rule reads near zero."""
    import numpy as np

    import oracle

    f, fp = oracle.fp64_closest(xyz, idx, rays, nthreads=thr, culled=True)
    gp = np.ascontiguousarray(hits[:, 3]).view(np.uint32)
    same = (fp != 0xFFFFFFFF) & (gp == fp)
    out = {"rays": int(len(rays)), "fp64_hits": int((fp != 0xFFFFFFFF).sum()), "prim_equal_fp64": int(same.sum()),
           "both_miss": int(((fp == 0xFFFFFFFF) & (gp == 0xFFFFFFFF)).sum())}
    if same.any():
        t, u, v = (hits[same, k].astype(np.float64) for k in range(3))
        du, dv = np.abs(u - f[same, 1]), np.abs(v - f[same, 2])
        rel_t = np.abs(t - f[same, 0]) / np.abs(f[same, 0])
        ru = du / np.maximum(np.abs(f[same, 1]), 1e-3)
        rv = dv / np.maximum(np.abs(f[same, 2]), 1e-3)
        out.update({"max_rel_t": float(rel_t.max()), "max_abs_du": float(du.max()), "max_abs_dv": float(dv.max()),
                    "max_rel_uv_floor_1e-3": float(max(ru.max(), rv.max())),
                    "p99_rel_t": float(np.percentile(rel_t, 99)),
                    "within_1e-5_rel": int(((rel_t <= 1e-5) & (ru <= 1e-5) & (rv <= 1e-5)).sum())})
    return out


def _final_gather_workload(args, dev, thr, trace_fn):
    """Final gather (PhotonMapperCPU.cpp:1110-1167) with ENCORE's shipped
    settings scaled to the 1/20 Cornell box: 2M photons split into the direct
    map (first hits) and the indirect map, 8x8 jittered hemisphere samples
    per shading point, indirect estimate (100 photons, r = 3/20), direct
    estimate (100, r = 2/20)."""
    import numpy as np
    import torch

    import oracle
    from encore_amd import rq, scenes

    xyz, idx = scenes.config4_scene()
    bvh = rq.Scene(dev, xyz, idx, accel="bvh")
    t = time.time()
    P, N, W, B = scenes.trace_photons(xyz, idx, trace_fn(bvh), 2000000, batch=1 << 18, with_bounce=True)
    ind, dr = B >= 1, B == 0
    log("photons: %d traced in %.1fs (%d direct, %d indirect)" % (len(P), time.time() - t, dr.sum(), ind.sum()))
    Q, QN = scenes.photon_queries(xyz, idx, trace_fn(bvh), res=256)
    ns = 8
    nper = ns * ns
    dirs = scenes.hemisphere_dirs(QN, ns, scenes.Stream(scenes.SEED_BASE + 6))
    diffuse = np.random.default_rng(6).uniform(0.2, 0.9, (len(idx), 3)).astype(np.float32)
    k_ind, r_ind, k_dir, r_dir = 100, 0.15, 100, 0.10
    log("shading points: %d x %d samples = %d rays" % (len(Q), nper, len(Q) * nper))
    gi = rq.PhotonMap(dev, P[ind], N[ind], W[ind])
    gd = rq.PhotonMap(dev, P[dr], N[dr], W[dr])
    d_q, d_dirs = torch.from_numpy(Q).cuda(), torch.from_numpy(dirs).cuda()
    d_dif = torch.from_numpy(diffuse).cuda()
    rgb = torch.empty((len(Q), 3), device="cuda")
    ri = torch.empty((len(Q),), device="cuda")
    out = {}
    kdsc = rq.Scene(dev, xyz, idx, accel="kd")
    for name, sc in (("bvh", bvh), ("kd", kdsc)):
        fn = lambda sc=sc: sc.final_gather(gi, k_ind, r_ind, gd, k_dir, r_dir, d_dif, d_q, d_dirs, nper, rgb, ri)
        wall, ev = _time_gpu(fn, args.steps, args.warmup)
        out[name] = {"Mrays_s": round(len(Q) * nper / (ev / 1000.0) / 1e6, 2), "ms": round(ev, 3),
                     "Mpoints_s": round(len(Q) / (ev / 1000.0) / 1e6, 4)}
        out[name + "_rgb"] = rgb.cpu().numpy()
        out[name + "_ri"] = ri.cpu().numpy()
        log("final gather %s: %s" % (name, out[name]))
    # ENCORE's default configuration (m_bPrecomputeIrr, PrecompIrrValue 1):
    # the indirect map replaced by its precomputed irradiance
    # (PhotonMap::PrecomputeIrradiance), and the indirect term at a sample hit
    # the power of the nearest photon within 10 (1/20: 0.5), which is the
    # k = 1 estimate of that map (PhotonMapperCPU.cpp:1129-1140).
    r_pre = 0.5
    torch.cuda.synchronize()
    t = time.perf_counter()
    pre, chosen, power = gi.precompute_irradiance(P[ind], N[ind], 1, k_ind, r_ind)
    pre_s = time.perf_counter() - t
    log("precomputed irradiance: %d photons in %.3fs" % (len(chosen), pre_s))
    for name, sc in (("bvh", bvh), ("kd", kdsc)):
        fn = lambda sc=sc: sc.final_gather(pre, 1, r_pre, gd, k_dir, r_dir, d_dif, d_q, d_dirs, nper, rgb, ri)
        wall, ev = _time_gpu(fn, args.steps, args.warmup)
        key = name + "_precomputed"
        out[key] = {"Mrays_s": round(len(Q) * nper / (ev / 1000.0) / 1e6, 2), "ms": round(ev, 3),
                    "Mpoints_s": round(len(Q) / (ev / 1000.0) / 1e6, 4)}
        out[key + "_rgb"] = rgb.cpu().numpy()
        out[key + "_ri"] = ri.cpu().numpy()
        log("final gather %s: %s" % (key, out[key]))
    out["precompute_irradiance_s"] = round(pre_s, 4)
    # CPU baseline and parity on a strided sample of shading points.
    stride = max(1, len(Q) // 2048)
    sel = np.arange(0, len(Q), stride)
    sdirs = dirs.reshape(len(Q), nper, 3)[sel].reshape(-1, 3)
    kd = oracle.KdTree(xyz, idx)
    oi = oracle.PhotonMap(P[ind], N[ind], W[ind])
    od = oracle.PhotonMap(P[dr], N[dr], W[dr])
    t = time.perf_counter()
    o_rgb, o_ri = oracle.final_gather(xyz, idx, kd, oi, k_ind, r_ind, od, k_dir, r_dir, diffuse, Q[sel], sdirs, nper,
                                      nthreads=thr)
    cpu_s = time.perf_counter() - t
    # the precomputed branch: the oracle gathers over the photons and powers
    # the GPU's precomputation chose
    oi_pre = oracle.PhotonMap(np.ascontiguousarray(P[ind][chosen]), np.ascontiguousarray(N[ind][chosen]), power)
    t = time.perf_counter()
    o_rgb_pre, o_ri_pre = oracle.final_gather(xyz, idx, kd, oi_pre, 1, r_pre, od, k_dir, r_dir, diffuse, Q[sel],
                                              sdirs, nper, nthreads=thr)
    cpu_pre_s = time.perf_counter() - t
    parity = {"sample_points": int(len(sel))}
    for name, ref_rgb, ref_ri in (("kd", o_rgb, o_ri), ("bvh", o_rgb, o_ri), ("kd_precomputed", o_rgb_pre, o_ri_pre),
                                  ("bvh_precomputed", o_rgb_pre, o_ri_pre)):
        g, gr = out.pop(name + "_rgb")[sel], out.pop(name + "_ri")[sel]
        parity[name] = {"bit_equal": int(np.all(g.view(np.uint32) == ref_rgb.view(np.uint32), axis=1).sum()),
                        "within_1e-3": int((np.all(np.isclose(g, ref_rgb, rtol=1e-3, atol=1e-6), axis=1) &
                                            np.isclose(gr, ref_ri, rtol=1e-3)).sum())}
    log("parity", parity)
    out["cpu_precomputed_Mrays_s"] = round(len(sel) * nper / cpu_pre_s / 1e6, 5)
    pre.destroy()
    _line("final-gather Mrays/s (8x8 samples, indirect + direct estimates)", out["bvh"]["Mrays_s"], "Mrays/s",
          out["bvh"]["ms"],
          {"workload": "final gather at %d shading points x %d hemisphere samples on the Cornell box at 1/20 "
                       "scale tessellated to %d triangles; %d indirect + %d direct photons, k = 100, r = 0.15 / 0.10"
                       % (len(Q), nper, len(idx), ind.sum(), dr.sum()), "accel": "bvh4"},
          paths=out, parity=parity,
          cpu_baseline={"value": round(len(sel) * nper / cpu_s / 1e6, 5), "unit": "Mrays/s", "cores": thr,
                        "kind": "port", "sample": "%d shading points x %d samples" % (len(sel), nper)})
    for o in (gi, gd, kdsc, bvh):
        o.destroy()
    return 0


def _irrcache_workload(args, dev, thr, trace_fn):
    """Irradiance-cache lookups (IrrCache::GetIrrValuesNearPoint and the final
    gather's weighted average, IrrCacheOctree.h:294-326, PhotonMapperCPU.cpp:
    1087-1103) at the ~1M hit points of a 1000x1000 view of the Cornell box at
    1/20 scale (34,816 triangles).  The cache holds what the final gather
    (rq_final_gather: 8x8 samples, indirect and direct estimates as the fg
    workload) computes at the 128x128 view's hit points, with ENCORE's
    FGTolerance a = 0.08 and root box (-75, -75, -75) + 150^3 scaled by 1/20."""
    import numpy as np
    import torch

    import oracle
    from encore_amd import rq, scenes

    xyz, idx = scenes.config4_scene()
    bvh = rq.Scene(dev, xyz, idx, accel="bvh")
    P, N, W, B = scenes.trace_photons(xyz, idx, trace_fn(bvh), 2000000, batch=1 << 18, with_bounce=True)
    ind, dr = B >= 1, B == 0
    gi = rq.PhotonMap(dev, P[ind], N[ind], W[ind])
    gd = rq.PhotonMap(dev, P[dr], N[dr], W[dr])
    Qc, QNc = scenes.photon_queries(xyz, idx, trace_fn(bvh), res=128)
    nper = 64
    dirs = scenes.hemisphere_dirs(QNc, 8, scenes.Stream(scenes.SEED_BASE + 6))
    diffuse = np.random.default_rng(6).uniform(0.2, 0.9, (len(idx), 3)).astype(np.float32)
    rgb_c = torch.empty((len(Qc), 3), device="cuda")
    ri_c = torch.empty((len(Qc),), device="cuda")
    bvh.final_gather(gi, 100, 0.15, gd, 100, 0.10, torch.from_numpy(diffuse).cuda(), torch.from_numpy(Qc).cuda(),
                     torch.from_numpy(dirs).cuda(), nper, rgb_c, ri_c)
    torch.cuda.synchronize()
    val, ri = rgb_c.cpu().numpy(), ri_c.cpu().numpy()
    a, lo, size = 0.08, (-3.75, -3.75, -3.75), (7.5, 7.5, 7.5)
    cache = rq.IrrCache(dev, a, lo, size)
    st = cache.add(val, Qc, QNc, ri)
    Q, QN = scenes.photon_queries(xyz, idx, trace_fn(bvh), res=1000)
    dq, dqn = torch.from_numpy(Q).cuda(), torch.from_numpy(QN).cuda()
    rgb = torch.empty((len(Q), 3), device="cuda")
    ws = torch.empty((len(Q),), device="cuda")
    cnt = torch.empty((len(Q),), dtype=torch.int32, device="cuda")
    t = time.perf_counter()
    cache.query(dq, dqn, rgb, ws, cnt)
    torch.cuda.synchronize()
    first_ms = (time.perf_counter() - t) * 1e3  # includes building the index
    wall, ev = _time_gpu(lambda: cache.query(dq, dqn, rgb, ws, cnt), args.steps, args.warmup)
    info = cache.info()
    n_look, n_exam, n_cnt = cache.counts(dq, dqn, torch.empty_like(rgb))
    # HBM bytes of one call: per lookup its position and normal (24), outputs
    # (rgb, weight sum, count: 20) and its cell's range (8); the cache and its
    # index once (48 per value, 4 per cell reference, 4 per cell start and
    # per global value).  The cache (16K values, 1.4 MB with its index) stays
    # in L2, so the per-lookup reads of cached records -- 4 (index) + 32
    # (position and Ri, normal) per value examined, 16 (value) per value
    # counted -- are reported beside it as L2-served bytes.
    alg = 52.0 * n_look + 48.0 * info["count"] + 4.0 * (info["index_refs"] + info["index_cells"] + 1 +
                                                         info["global_count"])
    achieved = alg / (ev / 1e3) / 1e9
    rec = 36.0 * n_exam + 16.0 * n_cnt
    log("irrcache: %d values (%d global), index %d cells; %.2f values examined and %.2f counted per lookup; "
        "%.3f ms per call (first %.1f ms)" % (info["count"], info["global_count"], info["index_cells"],
                                             n_exam / n_look, n_cnt / n_look, ev, first_ms))
    # CPU baseline: the reference's walk of the whole list per lookup, on a
    # strided sample; bit-exact parity on the same sample.
    o = oracle.IrrCache(a, lo, size)
    ost = o.add(val, Qc, QNc, ri)
    stride = max(1, len(Q) // 20000)
    sq, sqn = np.ascontiguousarray(Q[::stride]), np.ascontiguousarray(QN[::stride])
    t = time.perf_counter()
    orgb, ows, ocnt = o.query(sq, sqn, nthreads=thr)
    cpu_s = time.perf_counter() - t
    g = (rgb.cpu().numpy()[::stride], ws.cpu().numpy()[::stride], cnt.cpu().numpy().view(np.uint32)[::stride])
    eq = np.all(g[0].view(np.uint32) == orgb.view(np.uint32), axis=1) & (g[1].view(np.uint32) == ows.view(np.uint32)) & (
        g[2] == ocnt)
    parity = {"sample": int(len(sq)), "bit_equal": int(eq.sum()), "status_equal": bool(np.array_equal(st, ost)),
              "lookups_with_values": int((g[2] > 0).sum())}
    log("parity", parity)
    _line("Mlookups/s irradiance cache (final gather, a = 0.08)", len(Q) / (ev / 1e3) / 1e6, "Mlookups/s", ev,
          {"workload": "irradiance-cache lookups at %d hit points of the Cornell box at 1/20 scale (%d triangles); "
                       "cache of %d final-gather values (128x128 view, 8x8 samples), a = 0.08" % (
                           len(Q), len(idx), info["count"])},
          roofline={"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                    "hbm_bytes_per_lookup": round(alg / n_look, 1),
                    "l2_record_bytes_per_lookup": round(rec / n_look, 1),
                    "l2_record_GBps": round(rec / (ev / 1e3) / 1e9, 1),
                    "value_tests_per_s": float("%.4g" % (n_exam / (ev / 1e3))),
                    "values_examined_per_lookup": round(n_exam / n_look, 3),
                    "values_counted_per_lookup": round(n_cnt / n_look, 3), "kernel": "ic_query"},
          index=info, first_call_ms_incl_index=round(first_ms, 2), parity=parity,
          cpu_baseline={"value": round(len(sq) / cpu_s / 1e6, 5), "unit": "Mlookups/s", "cores": thr, "kind": "port",
                        "sample": "%d lookups (every %dth), the whole list per lookup as IrrCache's root list" % (
                            len(sq), stride)})
    for x in (cache, gi, gd, bvh):
        x.destroy()
    return 0


def _photon_emission_workload(args, dev, thr):
    """Photon emission (PhotonMapperCPU::EmitPhotons / TracePhoton,
    PhotonMapperCPU.cpp:503-775) into the global (1,000,000), direct
    (2,000,000) and caustic (0) maps, 3 bounces, from ENCORE's rectangle light
    in the Cornell box at 1/20 scale tessellated to 34,816 triangles with the
    scene file's materials.  One step = one complete emission."""
    import numpy as np

    import oracle
    from encore_amd import rq, scenes

    xyz, idx, mats, pm, lights = scenes.photon_scene(tess=32)
    cap = (1000000, 2000000, 0)
    out, stored = {}, {}
    for accel in ("bvh", "kd"):
        sc = rq.Scene(dev, xyz, idx, accel=accel)
        sc.set_materials(mats, pm)
        res = {}

        def fn(sc=sc, res=res):
            maps, paths = sc.emit_photons(lights, cap, seed=args.seed, extras=False)
            res["paths"] = paths
            res["counts"] = [int(maps[m]["pos"].shape[0]) for m in rq.MAPS]

        steps = args.steps if accel == "bvh" else 1
        wall, ev = _time_gpu(fn, steps, args.warmup if accel == "bvh" else 1)
        n = sum(res["counts"])
        out[accel] = {"Mphotons_s": round(n / wall / 1e6, 3), "Mpaths_s": round(res["paths"] / wall / 1e6, 3),
                      "ms": round(wall * 1e3, 3), "paths": res["paths"], "counts": res["counts"]}
        stored[accel] = sc
        log("emission %s: %s" % (accel, out[accel]))
    # CPU baseline (the oracle's sequential emission, one thread as the
    # reference's loop) and bit-exact parity of the kd scene on a smaller map.
    kd = oracle.KdTree(xyz, idx)
    scap = (args.pt_sample, 2 * args.pt_sample, 0)
    t = time.perf_counter()
    o, o_paths = oracle.emit_photons(kd, mats, pm, lights, scap, seed=args.seed)
    cpu_s = time.perf_counter() - t
    n_cpu = sum(len(o[m]["pos"]) for m in oracle.MAPS)
    parity = {"sample_global_capacity": args.pt_sample}
    def keyed(x):  # photons by (path, position bits, normal bits, shadow)
        return set(zip(x["path"].astype(np.int64).tolist(), map(bytes, x["pos"].view(np.uint32)),
                       map(bytes, x["normal"].view(np.uint32)), x["shadow"].tolist()))

    for accel in ("kd", "bvh"):
        g, paths = stored[accel].emit_photons(lights, scap, seed=args.seed)
        same, exact = 0, True
        for m in oracle.MAPS:
            gm = {k: v.cpu().numpy() for k, v in g[m].items()}
            same += len(keyed(gm) & keyed(o[m]))
            exact = exact and len(gm["pos"]) == len(o[m]["pos"]) and all(
                np.array_equal(gm[k].view(np.uint8), o[m][k].astype(gm[k].dtype).view(np.uint8)) for k in gm)
        parity[accel] = {"paths": paths, "oracle_paths": o_paths, "photons_identical": same, "photons": n_cpu,
                         "maps_bit_equal": bool(exact)}
    log("parity", parity)
    _line("photon-emission Mphotons/s (3 bounces, global 1M + direct maps)", out["bvh"]["Mphotons_s"], "Mphotons/s",
          out["bvh"]["ms"],
          {"workload": "photon emission from the 1.5 x 1.5 ceiling light of the Cornell box at 1/20 scale "
                       "tessellated to %d triangles until the global map holds %d photons (direct map %d), "
                       "3 bounces" % (len(idx), cap[0], cap[1]), "accel": "bvh4"},
          paths=out, parity=parity,
          cpu_baseline={"value": round(n_cpu / cpu_s / 1e6, 5), "unit": "Mphotons/s", "cores": 1, "kind": "port",
                        "sample": "emission until the global map holds %d photons (%d paths)" % (scap[0], o_paths)})
    for sc in stored.values():
        sc.destroy()
    return 0


def _refit_workload(args, dev, thr):
    """Scene update (AccelerationStructure::update) at 1.31M triangles: the
    displaced icosphere of configs 2/3 turned 0.05 rad about y with a
    travelling 1% wobble.  One step = one rq_scene_refit of the BVH4 scene
    (exact boxes from the new vertices).  Also: the config-3 rays' speed on
    the refitted tree against a tree built for the moved vertices, and
    parity on a ray sample against the oracle kd-tree built on them."""
    import math

    import numpy as np
    import torch

    import oracle
    from encore_amd import rq, scenes

    xyz, idx = scenes.config2_scene(level=args.level)
    p = xyz.astype(np.float64)
    a = 0.05
    x = math.cos(a) * p[:, 0] + math.sin(a) * p[:, 2]
    z = -math.sin(a) * p[:, 0] + math.cos(a) * p[:, 2]
    w = 1.0 + 0.01 * np.sin(p[:, 1] / 12.0)
    moved = np.stack([x * w, p[:, 1], z * w], axis=1).astype(np.float32)
    d_rest, d_moved = torch.from_numpy(xyz).cuda(), torch.from_numpy(moved).cuda()
    sc = rq.Scene(dev, xyz, idx, accel="bvh")
    build_s = sc.stats()["build_seconds"]
    flip = [0]

    def step():
        flip[0] ^= 1
        sc.refit(d_moved if flip[0] else d_rest)

    wall, ev = _time_gpu(step, args.steps * 2, args.warmup * 2)
    log("refit: %.4f ms per update of %d triangles (host build %.3f s)" % (ev, len(idx), build_s))
    sc.refit(d_moved)
    torch.cuda.synchronize()
    # Config-3 rays on the moved geometry: refitted tree against a fresh build.
    prim = scenes.config2_primary()
    fresh = rq.Scene(dev, moved, idx, accel="bvh")
    ph = torch.empty((len(prim), 4), device="cuda")
    fresh.intersect(torch.from_numpy(prim).cuda(), ph)
    rays, _ = scenes.config3_rays(moved, idx, prim, ph.cpu().numpy(), count=args.rays, nbatches=args.batches)
    d_rays = torch.from_numpy(rays).cuda()
    hits = torch.empty((len(rays), 4), device="cuda")
    speed = {}
    for name, s_ in (("refitted", sc), ("rebuilt", fresh)):
        _, t = _time_gpu(lambda s_=s_: s_.intersect(d_rays, hits), args.steps, args.warmup)
        speed[name] = round(len(rays) / (t / 1e3) / 1e6, 1)
    # The GPU build (linear BVH, BVH2 layout) of the moved vertices: wall time
    # of rq_scene_create including the upload, and the same rays on its tree.
    builds = []
    for _ in range(3):
        t0 = time.perf_counter()
        lb = rq.Scene(dev, moved, idx, accel="bvh", builder="lbvh")
        builds.append(time.perf_counter() - t0)
        if len(builds) < 3:
            lb.destroy()
    _, t = _time_gpu(lambda: lb.intersect(d_rays, hits), args.steps, args.warmup)
    speed["lbvh"] = round(len(rays) / (t / 1e3) / 1e6, 1)
    lb.intersect(d_rays, hits)
    got_lb = hits.cpu().numpy()
    log("lbvh build %.1f ms, %.1f Mrays/s" % (min(builds) * 1e3, speed["lbvh"]))
    sc.intersect(d_rays, hits)
    got = hits.cpu().numpy()
    stride = max(1, len(rays) // 65536)
    samp = np.ascontiguousarray(rays[::stride])
    kd = oracle.KdTree(moved, idx)
    ref = kd.intersect(samp, nthreads=thr)
    parity = _classify(got[::stride], ref, moved, idx, samp, thr)
    parity_lb = _classify(got_lb[::stride], ref, moved, idx, samp, thr)
    log("refit parity", parity, "lbvh parity", parity_lb)
    # ENCORE's Kdtree::update builds the tree again: the oracle's kd build of
    # the moved vertices, one thread, is the CPU baseline.
    t0 = time.perf_counter()
    oracle.KdTree(moved, idx)
    cpu_s = time.perf_counter() - t0
    out = {"refit_ms": round(ev, 4), "host_build_s": round(build_s, 3), "lbvh_build_ms": round(min(builds) * 1e3, 2),
           "c3_rays_Mrays_s": speed, "rays": len(rays), "lbvh_parity": parity_lb}
    lb.destroy()
    log("refit", out)
    _line("scene-update Mtris/s (BVH4 refit)", len(idx) / (ev / 1e3) / 1e6, "Mtris/s", ev,
          {"workload": "refit of the BVH4 scene of the %d-triangle displaced icosphere to vertices turned 0.05 rad "
                       "with a 1%% wobble, alternating with the rest pose" % len(idx), "accel": "bvh4"},
          paths=out, parity=parity,
          cpu_baseline={"value": round(len(idx) / cpu_s / 1e6, 4), "unit": "Mtris/s", "cores": 1, "kind": "port",
                        "sample": "ENCORE's kd-tree build (Kdtree::update rebuilds) of the moved %d triangles"
                                  % len(idx)})
    sc.destroy()
    fresh.destroy()
    return 0


def _trace_fn(scene):
    """Closest hits of host rays on `scene`, back on the host (set-up work
    outside the timed regions)."""
    import numpy as np
    import torch

    def f(rays):
        d = torch.from_numpy(np.ascontiguousarray(rays, np.float32)).cuda()
        o = torch.empty((len(rays), 4), dtype=torch.float32, device="cuda")
        scene.intersect(d, o)
        torch.cuda.synchronize()
        return o.cpu().numpy()
    return f


def run_workload(args):
    import numpy as np
    import torch

    import oracle
    from encore_amd import rq, scenes

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    torch.cuda.init()
    dev = rq.Device(0)
    thr = cpu_threads(args.cpu_threads)

    trace_fn = _trace_fn

    if args.workload in ("c1", "c2"):
        if args.workload == "c1":
            xyz, idx, rays = scenes.config1(res=512)
            sets = [("primary", rays, False)]
            name = "Cornell box, 34 triangles, 512x512 pixel-centre eye rays"
        else:
            xyz, idx = scenes.config2_scene(level=args.level)
            rays = scenes.config2_primary()
            name = "displaced icosphere level %d + ground (%d triangles), 1920x1080 eye rays + 1 shadow ray per hit" % (
                args.level, len(idx))
            sets = [("primary", rays, False)]
        results = {}
        kd = oracle.KdTree(xyz, idx)
        cpu_ref = {}
        for accel in ("bvh", "kd"):
            sc = rq.Scene(dev, xyz, idx, accel=accel)
            if args.workload == "c2" and accel == "bvh":
                ph = trace_fn(sc)(rays)
                ok, shadow = scenes.config2_shadow(xyz, idx, rays, ph)
                sets = [("primary", rays, False), ("shadow", shadow, True)]
            for label, r, any_hit in sets:
                d = torch.from_numpy(r).cuda()
                o = torch.empty((len(r),) if any_hit else (len(r), 4),
                                dtype=torch.int32 if any_hit else torch.float32, device="cuda")
                fn = (lambda d=d, o=o: sc.occluded(d, o)) if any_hit else (lambda d=d, o=o: sc.intersect(d, o))
                steps = args.steps if accel == "bvh" else 1
                wall, ev = _time_gpu(fn, steps, 1 if accel == "kd" else args.warmup)
                got = o.cpu().numpy()
                stride = max(1, len(r) // args.cpu_sample)
                samp = np.ascontiguousarray(r[::stride])
                if label not in cpu_ref:
                    t = time.perf_counter()
                    ref = kd.occluded(samp, nthreads=thr) if any_hit else kd.intersect(samp, nthreads=thr)
                    cpu_ref[label] = (ref, time.perf_counter() - t)
                ref, cpu_s = cpu_ref[label]
                if any_hit:
                    go = got[::stride].astype(np.uint32)
                    dif = np.nonzero(go != ref)[0]
                    par = {"n": int(len(samp)), "equal": int(len(samp) - len(dif))}
                    if len(dif):
                        # brute force: does any triangle give a hit below tmax?
                        lim = np.nextafter(samp[dif, 7], np.float32(0)).astype(np.float32)
                        bo = (oracle.count_hits_below(xyz, idx, samp[dif], lim, 0.0) > 0).astype(np.uint32)
                        par["gpu_matches_brute_force"] = int((bo == go[dif]).sum())
                        # FP64 without the flush: is any triangle really crossed below tmax?
                        f, fp = oracle.fp64_closest(xyz, idx, samp[dif], nthreads=thr)
                        real = ((fp != 0xFFFFFFFF) & (f[:, 0] < samp[dif, 7])).astype(np.uint32)
                        par["gpu_matches_fp64"] = int((real == go[dif]).sum())
                        # Rows neither explains: the GPU finds no hit while the
                        # reference's predicate hits.  Enumerate every predicate
                        # hit below tmax; if each lies outside its own triangle's
                        # box (a flush phantom no BVH reaches), the GPU answer is
                        # the predicate over the boxes the ray enters.
                        rest = dif[(bo != go[dif]) & (real != go[dif])]
                        par["reference_phantom_outside_its_triangle_box"] = 0
                        T = len(idx)
                        tri9 = np.ascontiguousarray(xyz[idx].reshape(T, 9), np.float32)
                        for j in rest[:16]:
                            if go[j] != 0:
                                continue
                            hh = oracle.tri_intersect_pairs(np.repeat(samp[j:j + 1], T, axis=0), tri9)
                            hit = np.nonzero((hh[:, 3] != 0) & (hh[:, 0] > 0) & (hh[:, 0] < samp[j, 7]))[0]
                            if len(hit) and np.all(_outside_box(xyz, idx, np.repeat(samp[j:j + 1], len(hit), axis=0),
                                                                hh[hit, 0], hit) > 0):
                                par["reference_phantom_outside_its_triangle_box"] += 1
                        if args.dump:
                            np.savez("%s_%s_%s.npz" % (args.dump, accel, label), rays=samp[dif], gpu=go[dif], kd=ref[dif],
                                     brute=bo, fp64=real, fp64_t=f[:, 0])
                else:
                    par = _classify(got[::stride], ref, xyz, idx, samp, thr,
                                    dump=args.dump and "%s_%s_%s.npz" % (args.dump, accel, label)) if accel == "bvh" else \
                        _classify(got[::stride], ref)
                cnt = sc.counts(d, any_hit=any_hit) if accel == "bvh" else None
                results["%s_%s" % (accel, label)] = {
                    "Mrays_s": round(len(r) / (ev / 1000.0) / 1e6, 2), "launch_ms": round(ev, 4),
                    "rays": int(len(r)), "cpu_Mrays_s": round(len(samp) / cpu_s / 1e6, 5), "cpu_sample": int(len(samp)),
                    "parity": par,
                    "nodes_per_ray": round(cnt[1] / cnt[0], 2) if cnt else None,
                    "tris_per_ray": round(cnt[2] / cnt[0], 2) if cnt else None}
                log(accel, label, results["%s_%s" % (accel, label)])
            sc.destroy()
        main = results["bvh_primary"]
        _line("Mrays/s (%s)" % args.workload, main["Mrays_s"], "Mrays/s", main["launch_ms"],
              {"workload": name, "accel": "bvh", "predicate": "encore", "mode": "one launch per ray set"},
              results=results, cpu_cores=thr)
        return 0

    if args.workload == "fg":
        return _final_gather_workload(args, dev, thr, trace_fn)
    if args.workload == "pt":
        return _photon_emission_workload(args, dev, thr)
    if args.workload == "ic":
        return _irrcache_workload(args, dev, thr, trace_fn)
    if args.workload == "refit":
        return _refit_workload(args, dev, thr)
    if args.workload == "c4":
        return _photon_gather_workload(args, dev, thr, trace_fn)

    # ---- c5: serving ---------------------------------------------------
    res = _serving(args, dev, trace_fn)
    _line("served Mrays/s at equal p99 (serving, 16 resident scenes)", res["served_at_equal_p99"]["Mrays_s"]
          if res["served_at_equal_p99"] else 0.0, "Mrays/s", 0.0, res.pop("config"), **res)
    return 0


def _photon_gather_workload(args, dev, thr, trace_fn):
    """Config 4: radiance estimates with neighbour lists (GetNearestNPhotons
    and GetRadianceEstimate, PhotonMap.cpp:76-164, 224-297) at the ~1M hit
    points of a 1000x1000 view, k = 50, over 2,000,000 photons traced into
    the Cornell box tessellated to 34,816 triangles.  The box is built at
    1/20 scale (at 1/100 its 0.03-unit edges would flush Cross()'s normals);
    the headline radius 0.25 is the 1/100 box's r = 0.05 scaled by 5 -- the
    same photons per gather disk, ~10^3 candidates per query -- and r = 0.05
    at 1/20 is reported beside it with the balanced-tree walks."""
    import numpy as np
    import torch

    import oracle
    from encore_amd import rq, scenes

    xyz, idx = scenes.config4_scene()
    sc = rq.Scene(dev, xyz, idx, accel="bvh")
    t = time.time()
    P, N, W = scenes.trace_photons(xyz, idx, trace_fn(sc), 2000000, batch=1 << 20)
    log("photons: %d traced in %.1fs" % (len(P), time.time() - t))
    Q, QN = scenes.photon_queries(xyz, idx, trace_fn(sc), res=1000)
    log("queries: %d" % len(Q))
    k = 50
    dq = torch.from_numpy(Q).cuda()
    dqn = torch.from_numpy(QN).cuda()
    di = torch.empty((len(Q), k), dtype=torch.int32, device="cuda")
    dd = torch.empty((len(Q), k), dtype=torch.float32, device="cuda")
    dc = torch.empty((len(Q),), dtype=torch.int32, device="cuda")
    drgb = torch.empty((len(Q), 3), dtype=torch.float32, device="cuda")
    ref = oracle.PhotonMap(P, N, W)
    perm, _ = ref.order()
    radii = {}
    for r, trees, cpu_queries in ((0.05, True, 65536), (0.25, False, 8192)):
        out = {}
        # The engine's default (photon cell grid), then, at the small radius,
        # the balanced-tree walks (RQ_KNN_CELLS=0) over the spatial and
        # ENCORE's own layout.  The parity sample checks the default path.
        paths = ((("tree_encore_layout", "encore", True), ("tree_spatial_layout", "spatial", True)) if trees
                 else ()) + (("cells", "spatial", False),)
        for name, layout, tree in paths:
            if tree:
                os.environ["RQ_KNN_CELLS"] = "0"
            else:
                os.environ.pop("RQ_KNN_CELLS", None)
            t = time.time()
            pm = rq.PhotonMap(dev, P, N, W, layout=layout)
            build = time.time() - t
            fn = lambda: pm.radiance(dq, dqn, k, r, drgb, di, dd, dc)
            t = time.perf_counter()
            fn()
            torch.cuda.synchronize()
            first_ms = (time.perf_counter() - t) * 1e3
            wall, ev = _time_gpu(fn, args.steps, args.warmup)
            out[name] = {"Mqueries_s": round(len(Q) / (ev / 1000.0) / 1e6, 3), "launch_ms": round(ev, 3),
                         "build_s": round(build, 2)}
            if not tree:
                out[name]["first_call_ms_incl_grid_build"] = round(first_ms, 2)
                # the gather's work, counted by the kernel's instrumented copy
                # (no lists: the same cells, photons and candidates)
                rgb_c = torch.empty_like(drgb)
                work = pm.radiance_counts(dq, dqn, k, r, rgb_c)
                torch.cuda.synchronize()
                out[name]["counts_rgb_equal"] = bool(torch.equal(rgb_c.view(torch.int32), drgb.view(torch.int32)))
            log("r=%g" % r, name, out[name])
            pm.destroy()
        os.environ.pop("RQ_KNN_CELLS", None)
        cnt = dc.cpu().numpy()
        got_i = di.cpu().numpy().view(np.uint32)
        got_d = dd.cpu().numpy()
        rgb = drgb.cpu().numpy()
        stride = max(1, len(Q) // min(args.cpu_sample, cpu_queries))
        sq, sqn = np.ascontiguousarray(Q[::stride]), np.ascontiguousarray(QN[::stride])
        t = time.perf_counter()
        rrgb = ref.radiance(sq, sqn, k, r, nthreads=thr)
        cpu_s = time.perf_counter() - t
        rs, rd, rc = ref.knn(sq, sqn, k, r, nthreads=thr)
        gi, gd, gc = got_i[::stride], got_d[::stride], cnt[::stride]
        same_cnt = int((gc == rc).sum())
        same_set = 0
        for i in range(len(sq)):
            c = int(rc[i])
            a, bset = gi[i, :c], perm[rs[i, :c]]
            if gc[i] == c and np.array_equal(np.sort(gd[i, :c]), np.sort(rd[i, :c])) and (
                    c == 0 or set(a[gd[i, :c] < gd[i, :c].max()].tolist()) ==
                    set(bset[rd[i, :c] < rd[i, :c].max()].tolist())):
                same_set += 1
        rel = np.abs(rgb[::stride] - rrgb) / np.maximum(np.abs(rrgb), 1e-30)
        parity = {"sample": int(len(sq)), "count_equal": same_cnt, "set_equal_tie_aware": same_set,
                  "radiance_max_rel": float(rel.max()), "mean_count": float(gc.mean())}
        log("r=%g parity" % r, parity)
        # Algorithmic bytes of one call: per query 24 B in (position, normal),
        # 8 B per range ranked (a cell's photon range, or a block's 32-byte
        # box counted at the cell rate), 20 B per photon read (position and
        # class 16, slot 4), 28 B per photon gathered (power 16 and insertion
        # index 4 read, index and distance 4 + 4 written) and 16 B out
        # (radiance, count).
        nqf = float(len(Q))
        gathered = float(cnt.astype(np.int64).sum())
        alg = 40.0 * nqf + 8.0 * work["ranges"] + 20.0 * work["photons"] + 28.0 * gathered
        per_q = {"ranges_ranked": round(work["ranges"] / nqf, 2), "photons_read": round(work["photons"] / nqf, 1),
                 "candidates": round(work["candidates"] / nqf, 1), "gathered": round(gathered / nqf, 2),
                 "bytes": round(alg / nqf, 1)}
        log("r=%g per query" % r, per_q)
        achieved = alg / (out["cells"]["launch_ms"] / 1e3) / 1e9
        radii[r] = {"paths": out, "parity": parity, "per_query": per_q,
                    "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                 "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                                 "kernel": "knn_cell_prep + knn_cell_kernel (one call, events around both)"},
                    "cpu_baseline": {"value": round(len(sq) / cpu_s / 1e6, 5), "unit": "Mqueries/s", "cores": thr,
                                     "kind": "port", "sample": "%d gather points (every %dth)" % (len(sq), stride)}}
    main, side = radii[0.25], radii[0.05]
    side["Mqueries_s"] = side["paths"]["cells"]["Mqueries_s"]
    _line("Mqueries/s photon gather (k=50, r=0.05 at 1/100 scale = 0.25 at 1/20)",
          main["paths"]["cells"]["Mqueries_s"], "Mqueries/s", main["paths"]["cells"]["launch_ms"],
          {"workload": "radiance estimate with lists at %d gather points, %d photons (3 bounces, albedo 0.7) on "
                       "the Cornell box at 1/20 scale tessellated to %d triangles; radius 0.25 = the 1/100 box's 0.05 "
                       "scaled by 5" % (len(Q), len(P), len(idx)), "k": k, "max_dist": 0.25},
          paths=main["paths"], parity=main["parity"], per_query=main["per_query"], roofline=main["roofline"],
          cpu_baseline=main["cpu_baseline"], r_0_05=side)
    return 0


def _serving(args, dev, trace_fn, served_cfg=None):
    """Config 5: 16 resident displaced-sphere scenes, 10k to 4M triangles
    (log-uniform; the two ends forced), Poisson arrivals of ray-batch requests
    (lognormal 4,096..2,097,152 rays, median 65,536; a random scene; eye,
    bounce or shadow rays), host<->HBM copies included.  The plain server (one
    request per launch) is run at 50/80/100% of its capacity; its p99 at 80%
    is the latency bar; the served (continuous-batching) server is swept over
    rates and its best throughput with p99 at or under the bar is reported
    with the ratio to plain at 80%."""
    ctx = _serving_setup(args, dev, trace_fn)
    try:
        return _serving_measure(ctx, served_cfg or {})
    finally:
        for sc in ctx["scenes"]:
            sc.destroy()


def _serving_setup(args, dev, trace_fn):
    import numpy as np
    import torch

    from encore_amd import rq, scenes

    t0 = time.time()
    st = scenes.Stream(scenes.SEED_BASE + 5)
    T = np.exp(np.log(1e4) + (np.log(4e6) - np.log(1e4)) * st.draw(16))
    T[np.argmin(T)], T[np.argmax(T)] = 1e4, 4e6
    nreq = args.requests
    rs = scenes.Stream(scenes.SEED_BASE + 55)
    sizes = np.clip(np.exp(np.log(65536) + 1.18 * rs.normal(nreq)), 4096, 2097152).astype(np.int64)
    which = np.minimum((rs.draw(nreq) * 16).astype(np.int64), 15)
    kinds = np.minimum((rs.draw(nreq) * 3).astype(np.int64), 2)
    gaps = -np.log(np.maximum(rs.draw(nreq), 1e-12))  # unit-rate exponential inter-arrivals
    scns, ntris, reqs = [], [], []
    for i, tcount in enumerate(T):
        freq = int(np.ceil(np.sqrt(tcount / 20.0)))
        xyz, idx = scenes.displaced_sphere(freq=freq, seed=scenes.SEED_BASE + 50 + i, ground=False)
        sc = rq.Scene(dev, xyz, idx, accel="bvh")
        scns.append(sc)
        ntris.append(len(idx))
        if not (which == i).any():
            continue
        # Pools as large as this scene's largest request: eye rays, bounce
        # rays from their hits, shadow rays from the hits toward a light.
        need = int(sizes[which == i].max())
        side = int(np.ceil(np.sqrt(need)))
        prim = scenes.eye_rays(side, side, (0.0, 30.0, 350.0), (0.0, 0.0, 0.0), 45.0, 1.0, 0.1)
        ok, p, nrm = scenes.hit_points(xyz, idx, trace_fn(sc)(prim))
        pool = [prim[:need],
                scenes.bounce_rays(p, nrm, prim[ok, 4:7], need, scenes.Stream(scenes.SEED_BASE + 90 + i), 0.1),
                np.tile(scenes.shadow_rays(p, (200.0, 400.0, 200.0), 100.0, scenes.Stream(scenes.SEED_BASE + 70 + i)),
                        (int(np.ceil(need / max(1, ok.sum()))), 1))[:need]]
        # Pinned pools; a request reads the first n rays of its kind's pool.
        pool = [torch.from_numpy(np.ascontiguousarray(x)).pin_memory() for x in pool]
        for r in np.nonzero(which == i)[0]:
            reqs.append((r, pool[kinds[r]][:sizes[r]]))
    reqs = [x for _, x in sorted(reqs, key=lambda t: t[0])]
    log("serving: 16 scenes (%d..%d triangles) and %d requests ready in %.1fs" % (
        min(ntris), max(ntris), nreq, time.time() - t0))
    total_rays = int(sizes.sum())
    out_pool = torch.empty((total_rays, 4), dtype=torch.float32).pin_memory()
    return {"dev": dev, "scenes": scns, "ntris": ntris, "reqs": reqs, "sizes": sizes, "which": which, "kinds": kinds,
            "gaps": gaps, "nreq": nreq, "total_rays": total_rays, "out_pool": out_pool, "t0": t0}


def _serving_measure(ctx, served_cfg):
    import numpy as np
    import torch

    from encore_amd import rq

    dev, scns, ntris, reqs = ctx["dev"], ctx["scenes"], ctx["ntris"], ctx["reqs"]
    sizes, which, kinds, gaps = ctx["sizes"], ctx["which"], ctx["kinds"], ctx["gaps"]
    nreq, total_rays, out_pool, t0 = ctx["nreq"], ctx["total_rays"], ctx["out_pool"], time.time()

    def run_trace(server, rate):
        """Submit the request trace at `rate` requests/s (inf = all at once)."""
        tickets = []
        o = 0
        t_start = time.perf_counter()
        next_t = t_start
        for i in range(nreq):
            if np.isfinite(rate):
                next_t += gaps[i] / rate
                while True:
                    d = next_t - time.perf_counter()
                    if d <= 0:
                        break
                    time.sleep(min(d, 0.0005))
            n = int(sizes[i])
            any_hit = kinds[i] == 2
            out = out_pool[o:o + n] if not any_hit else out_pool[o:o + n, 0].view(torch.int32)
            tickets.append(server.submit(scns[which[i]], reqs[i], out, any_hit))
            o += n
        infos = [server.wait(t, 600000) for t in tickets]
        t_end = time.perf_counter()
        lat = np.array([x["latency_ms"] for x in infos])
        return {"rate_req_s": float(rate), "Mrays_s": total_rays / (t_end - t_start) / 1e6,
                "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
                "launches": int(server.stats()["launches"])}

    plain = rq.Server(dev, plain=True, hbm_budget=8 << 30)
    served = rq.Server(dev, hbm_budget=8 << 30, **served_cfg)
    run_trace(plain, float("inf"))  # warm
    cap = run_trace(plain, float("inf"))
    cap_rate = nreq / (total_rays / (cap["Mrays_s"] * 1e6))
    log("plain capacity: %.1f Mrays/s = %.1f requests/s" % (cap["Mrays_s"], cap_rate))
    run_trace(served, float("inf"))
    sweep = {"plain": [], "served": []}
    for f in (0.5, 0.8, 1.0):
        sweep["plain"].append(run_trace(plain, f * cap_rate))
        log("plain", sweep["plain"][-1])
    for f in (0.5, 0.8, 1.0, 1.25, 1.5, 1.75, 2.0, 2.25, 2.5, 2.75, 3.0, 3.25, 3.5, 4.0):
        sweep["served"].append(run_trace(served, f * cap_rate))
        log("served", sweep["served"][-1])
    served_cap = run_trace(served, float("inf"))
    ref_p99 = sweep["plain"][1]["p99_ms"]  # plain at 80% of its capacity
    ok = [x for x in sweep["served"] if x["p99_ms"] <= ref_p99]
    best = max(ok, key=lambda x: x["Mrays_s"]) if ok else None
    ratio = best["Mrays_s"] / sweep["plain"][1]["Mrays_s"] if best else None
    plain.destroy()
    served.destroy()
    log("serving: served/plain at equal p99 = %s (%.1fs)" % (ratio and round(ratio, 2), time.time() - t0))
    rnd = lambda d: {k: round(v, 3) for k, v in d.items()}
    return {"config": {"workload": "16 displaced-sphere scenes %d-%d triangles, %d Poisson requests of 4k-2M rays "
                                   "(lognormal, median 65536), primary/bounce/shadow, host<->HBM copies included" % (
                                       min(ntris), max(ntris), nreq), "triangles": ntris},
            "plain_at_80pct": rnd(sweep["plain"][1]),
            "served_at_equal_p99": rnd(best) if best else None,
            "served_over_plain_at_equal_p99": round(ratio, 2) if ratio else None,
            "capacity": {"plain_Mrays_s": round(cap["Mrays_s"], 1), "served_Mrays_s": round(served_cap["Mrays_s"], 1)},
            "sweep": {k: [rnd(x) for x in v] for k, v in sweep.items()}}




if __name__ == "__main__":
    sys.exit(main())
