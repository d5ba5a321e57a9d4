"""Benchmark of the MI355X AprilTag detection path.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--batch B]

Workload: 10 seeded synthetic 1920x1080 frames, 20 tag36h11 tags each
(40-300 px, tilt <= 45 deg, occluding bars, value-noise background,
blur 0-1.5, noise 2-8), decimate 2, refine_edges 1, bits_corrected 2.
The frames are resident in HBM before the timed region.

A step is one pass of the detection chain over one batch of B frames
(the 10 frames repeated); `value` is the batched frames/s. Beside it:
the single-frame path (one frame per call) frames/s and its p50/p99
latency, bit-exactness of both against the oracle, per-kernel device
times (HIP events on the chain's stream), the roofline of the longest
kernel of the HBM-bound stages (a)-(c) on the algorithmic bytes of its
stage (the longest kernel of the whole chain named beside it), the
reference CPU detector timed on the host (1 thread and all cores, median
of 3 runs, alternating with the GPU), the 640x480 and 3840x2160 configurations one frame at a
time against the CPU, batches from host memory, the uneven 64-frame batch,
the serving sweep, and the pose of the batch's detections (GPU against the
reference's estimate_tag_pose on one host thread).
"""
import argparse
import json
import os
import sys
import threading
import time
from pathlib import Path

import numpy as np

REPO = Path(__file__).resolve().parent
sys.path.insert(0, str(REPO / "tests"))

HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def frames_config2(atg, n=10):
    return [atg.synth_frame(2, seed)[0] for seed in range(n)]


def pctl(v, q):
    return float(np.percentile(np.asarray(v), q))


def cpu_model():
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


# The kernels behind each timing mark of the chain (the names rocprofv3
# shows), and the stage of the detection path each belongs to: (a)
# decimate and threshold (with the run-length encoding fused into it), (b)
# union-find, (c) roots, points and cluster order, (d) quad fit, (e) decode.
MARK_KERNELS = {
    "decimate": ["k_decimate_tiles"], "blur": ["k_blur_rows", "k_blur_cols", "k_tiles"],
    "threshold_rle": ["k_threshold"], "uf_edges": ["k_uf_edges<true>", "k_uf_edges<false>"],
    "uf_front.sweep": ["k_front<1024>", "k_front<512>", "k_front<256>", "k_front<128>", "k_front<64>",
                       "k_front_local<1024>", "k_front_carry<1024>", "k_front_final<1024>"],
    "uf_front.dense": ["k_front_dense"], "uf_replay.rows": ["k_replay<true>", "k_replay<false>"],
    "uf_replay.wide": ["k_replay_wide<true>", "k_replay_wide<false>", "k_replay_hbm", "k_work_offsets"],
    "uf_stitch.jump": ["k_root_jump"], "uf_stitch.lost": ["k_lost"], "resolve": ["k_resolve"],
    "labels.rows": ["k_labels"],
    "points.count": ["k_point_count", "k_point_scan"], "points.write": ["k_point_rows"],
    "clusters.rank": ["k_cluster_rank"],
    "cluster_sort": ["k_rank_clear", "k_rank_counts", "k_cluster_starts", "k_radix_hist", "k_radix_offsets",
                     "k_radix_scatter"],
}


def stage_of(mark):
    g = mark.split(".")[0]
    if g in ("decimate", "blur", "threshold_rle"):
        return "a"
    if g.startswith("uf_") or g == "resolve":
        return "b"
    if g in ("labels", "points", "clusters", "cluster_sort"):
        return "c"
    if g in ("fit_prep", "angle_sort", "moments", "corners"):
        return "d"
    return "e"


def measured_traffic(batch):
    """HBM bytes per launch of each kernel at this batch size, as the
    rocprofv3 FETCH_SIZE and WRITE_SIZE runs measured them (scaled by the
    copy kernel's readings): the kernel rows of
    profiles/traffic_1080p_batch<B>.txt, if that profile exists."""
    path = REPO / "profiles" / ("traffic_1080p_batch%d.txt" % batch)
    out = {}
    if not path.exists():
        return out
    for line in path.read_text().splitlines():
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cells) >= 5 and cells[0].startswith("kernel:"):
            out[cells[0][len("kernel:"):].strip()] = int(cells[4])
    return out


def frame_counts(ctx, i, W, H):
    """The data-dependent sizes of frame i of the last call: R runs, N
    nodes, E union-find links, P points, C clusters, K points of the
    clusters that reach the quad fit, Q quads."""
    c = ctx.counters(i)
    st = ctx.stages(i)
    qw, qh = st["qw"], st["qh"]
    R = c[0]
    # the cluster sort: 11-bit radix passes over the cluster ranks' bits
    return dict(W=W, H=H, qw=qw, qh=qh, R=R, N=R + qh, E=c[6], P=c[1], C=c[2], K=c[7], Q=c[3],
                S=(c[11] + 10) // 11)


def survey_bytes(n, decimated=True):
    """Algorithmic HBM bytes per frame of stages (a)-(c) (each logical
    array written once by its producer and read once by each consumer):
    (a) W H + 2 qw qh (read the frame; write the quad image and the
    thresholded image; W H + qw qh without decimation), (b) qw qh + 10 R +
    4 N + 4 qh (read t; write and read runs; write roots; row offsets), (c)
    qw qh + 5 R + 4 N + 8 P (read t, runs and roots; write points)."""
    W, H, qw, qh, R, N, P = n["W"], n["H"], n["qw"], n["qh"], n["R"], n["N"], n["P"]
    px = qw * qh
    return {"a": W * H + (2 if decimated else 1) * px, "b": px + 10 * R + 4 * N + 4 * qh,
            "c": px + 5 * R + 4 * N + 8 * P}


def traffic_model(n):
    """The implementation's own byte model per kernel group (every array a
    kernel group reads or writes, once): what the chain moves, which is
    more than the algorithmic bytes (two label planes, radix passes,
    gathers); reported beside them, never as the roofline."""
    W, H, qw, qh = n["W"], n["H"], n["qw"], n["qh"]
    R, N, E, P, K, Q = n["R"], n["N"], n["E"], n["P"], n["K"], n["Q"]
    px, tiles = qw * qh, (qw // 4) * (qh // 4)
    lg, d = 0, 1  # root-jump passes: each follows 4 hops, the resolved distance x5
    while d < qh:
        d, lg = d * 5, lg + 1
    return {
        "decimate": W * H + px + 2 * tiles,
        "threshold_rle": px + 2 * tiles + px + 4 * R + 4 * qh,
        "uf_edges": 8 * R + 4 * E + 4 * qh,
        "uf_front": 4 * R + 4 * E + 4 * R + 4 * N + 12 * qh,
        "uf_replay": 4 * R + 4 * E + 4 * R + 4 * N + 12 * N,
        "uf_stitch": lg * 8 * N + 16 * N,
        "resolve": 16 * N,
        "labels": 4 * N + px + 4 * px,
        "points": 2 * (px + 4 * px) + 12 * P,
        "clusters": 28 * n["C"],
        "cluster_sort": 16 * P + n["S"] * 16 * P + (n["S"] - 1) * 8 * P + 16 * P,
        "fit_prep": 8 * P + 20 * P + 12 * K,
        "angle_sort": 16 * K,
        "moments": 72 * K,
        "corners": 80 * K,
        "decode": Q * 4 * 16 * 25 * 8,
        "finish": Q * 168,
    }


SERVE_SIZES = [(640, 480), (1280, 720), (1920, 1080), (2448, 2048)]


def serving_sweep(atg, ctx, loads, seconds, nstreams=32, nframes=4, check_ref=True):
    """Serving: nstreams camera streams (seeded sizes from SERVE_SIZES, rates
    30-120 frames/s scaled to each offered load), frames of drifting tags
    (config-1 scenes) in pinned host memory, Poisson arrivals fed to the
    server; per load: sustained frames/s and p50/p99 arrival-to-result
    latency. Every served result is compared with the frame's result alone
    (and those with the reference CPU detector)."""
    from concurrent.futures import ThreadPoolExecutor
    from atg_testlib import Reference, default_params, have_ref, same_dets
    fams = [("tag36h11", 2)]
    rng = np.random.default_rng(2024)
    sizes = [SERVE_SIZES[i] for i in rng.integers(0, 4, nstreams)]
    rates = rng.uniform(30.0, 120.0, nstreams)

    def gen(k):
        w, h = sizes[k]
        spec = atg.synth_spec(1, w, h)
        out = []
        for j in range(nframes):
            pf = atg.PinnedFrame(h, w)
            atg.synth_stream_frame(spec, 5000 + k, 8 * j, pf.array)
            out.append(pf)
        return out

    with ThreadPoolExecutor(16) as ex:
        frames = list(ex.map(gen, range(nstreams)))
    ctx.set_families(fams)
    prm = atg.default_params()
    alone = {(k, j): ctx.detect(frames[k][j].array, prm) for k in range(nstreams) for j in range(nframes)}
    ref_ok = None
    if check_ref and have_ref():
        ref = Reference()
        ref_ok = True
        for (k, j), a in alone.items():
            r = ref.detect(frames[k][j].array, default_params(), fams)
            ref_ok = ref_ok and same_dets(r["dets"], a["dets"]) and r["nquads"] == a["nquads"]
    points, all_exact = [], True
    for load in loads:
        srv = atg.Server(fams, lanes=3, max_batch=128, budget_bytes=16 << 30)
        try:
            ids = [srv.add_stream(*sizes[k]) for k in range(nstreams)]
            scale = load / rates.sum()
            arr = []
            for k in range(nstreams):
                t, j = 0.0, 0
                gap = 1e9 / (rates[k] * scale)
                while True:
                    t += rng.exponential(gap)
                    if t >= seconds * 1e9:
                        break
                    arr.append((ids[k], t, frames[k][j % nframes].array, k * 1000000 + j))
                    j += 1
            log("serving: offered %d frames/s, %d arrivals" % (load, len(arr)))
            srv.feed(arr)
            srv.drain(600000)
            res = []
            while True:
                got = srv.poll(4096)
                if not got:
                    break
                res.extend(got)
            counts = srv.counts()
            # dropped arrivals come back with ATG_ERR_LIMIT and no detections
            res = [r for r in res if r["status"] != -3]
        finally:
            srv.close()
        exact = all(r["status"] == 0 and same_dets(r["dets"], alone[(r["tag"] // 1000000, (r["tag"] % 1000000) % nframes)]["dets"])
                    for r in res)
        all_exact = all_exact and exact
        lat = np.array([(r["done_ns"] - r["arrival_ns"]) * 1e-6 for r in res])
        t0 = min(r["arrival_ns"] for r in res)
        t1 = max(r["done_ns"] for r in res)
        points.append(dict(offered_fps=load, arrivals=len(arr), completed=len(res), dropped=counts["dropped"],
                           sustained_fps=round(len(res) / ((t1 - t0) * 1e-9), 1),
                           p50_ms=round(float(np.percentile(lat, 50)), 3),
                           p99_ms=round(float(np.percentile(lat, 99)), 3),
                           mean_batch=round(float(np.mean([r["batch_frames"] for r in res])), 1),
                           batches=counts["batches"], bit_exact_vs_alone=bool(exact)))
    for fr in frames:
        for pf in fr:
            pf.free()
    return dict(streams=nstreams, frame_sizes={"%dx%d" % s: sizes.count(s) for s in SERVE_SIZES},
                rates_fps="30-120 per stream, scaled to the offered load", budget_gib=16, lanes=3, max_batch=128,
                seconds_per_point=seconds, points=points, bit_exact_vs_alone=bool(all_exact),
                alone_bit_exact_vs_cpu_reference=ref_ok)


def batching_config4(atg, ctx, steps):
    """Uneven batch: 64 frames, 16 each of 640x480, 1280x720, 1920x1080
    and 2448x2048 (config-1 scenes with 5-30 tags), resident in HBM, one
    call of the chain; frames/s with the default work splits and with the
    autotuner's, and that both equal each frame's single-frame result."""
    from atg_testlib import same_dets
    sizes = [(640, 480), (1280, 720), (1920, 1080), (2448, 2048)]
    frames = [atg.synth_frame(4, 100 + i, width=sizes[i % 4][0], height=sizes[i % 4][1])[0] for i in range(64)]
    ctx.set_families([("tag36h11", 2)])
    prm = atg.default_params()
    dev = [ctx.to_device(f) for f in frames]
    try:
        alone = [ctx.detect(None, prm, on_device_ptr=d) for d in dev]

        def timed():
            ctx.detect_batch(dev, prm, on_device=True)
            t0 = time.perf_counter()
            for _ in range(steps):
                out = ctx.detect_batch(dev, prm, on_device=True)
            return 64 * steps / (time.perf_counter() - t0), out

        default_tune = ctx.get_tune()
        fps0, out0 = timed()
        tune, _ = ctx.autotune(dev, prm, on_device=True)
        fps1, out1 = timed()
        exact = all(same_dets(a["dets"], b["dets"]) and a["nquads"] == b["nquads"] and
                    same_dets(a["dets"], c["dets"]) for a, b, c in zip(alone, out0, out1))
        ctx.set_tune(*default_tune)
    finally:
        for d in dev:
            ctx.free_device(d[0])
    return dict(frames="64: 16 each of 640x480, 1280x720, 1920x1080, 2448x2048", fps_default=round(fps0, 1),
                fps_autotuned=round(fps1, 1), autotuned=dict(zip(atg.TUNE_FIELDS, tune)),
                default=dict(zip(atg.TUNE_FIELDS, default_tune)),
                bit_exact_vs_single=bool(exact), detections=int(sum(len(a["dets"]) for a in alone)))


class CpuRef:
    """The reference detector (oracle/_ref, built from the reference's own
    sources with its single-thread flags) timed on this host; the C
    restatement (oracle/) where the reference build is absent."""

    def __init__(self):
        from atg_testlib import Oracle, Reference, have_ref
        self.kind = "reference" if have_ref() else "port"
        self.ref = Reference() if have_ref() else None
        self.orc = None if have_ref() else Oracle()
        self.dets = {}

    def runner(self, params, fams, nthreads):
        """A function running the detector on one frame."""
        from atg_testlib import default_params
        p = dict(params, nthreads=nthreads)
        if self.ref is not None:
            key = (json.dumps(p, sort_keys=True), tuple(fams))
            if key not in self.dets:
                self.dets[key] = self.ref.detector(p, fams)
            td, lib = self.dets[key], self.ref.lib
            return lambda img: lib.at_detect(td, img.ctypes.data, img.shape[1], img.shape[0], img.strides[0])
        return lambda img: self.orc.detect(img, default_params(**p), fams)

    def close(self):
        for td in self.dets.values():
            self.ref.lib.at_detector_destroy(td)
        self.dets = {}


def gpu_single_run(ctx, dev, prm, iters):
    """One frame a call over the device-resident frames (cycled): latencies (s)."""
    lat = []
    for k in range(iters):
        t = time.perf_counter()
        ctx.detect(None, prm, on_device_ptr=dev[k % len(dev)])
        lat.append(time.perf_counter() - t)
    return lat


def cpu_run(fn, frames, iters):
    t0 = time.perf_counter()
    for k in range(iters):
        fn(frames[k % len(frames)])
    return iters / (time.perf_counter() - t0)


def compare_single(atg, ctx, cpu, frames, pdict, fams, iters, cpu_iters, cores):
    """A configuration one frame at a time: the GPU single-frame path and
    the CPU detector at 1 thread and at all cores, 3 runs each, the sides
    alternating (runs of `iters` GPU calls and at least 15 CPU frames);
    medians of the runs."""
    ctx.set_families(fams)
    prm = atg.params_from_dict(pdict)
    dev = [ctx.to_device(f) for f in frames]
    try:
        gpu_single_run(ctx, dev, prm, len(dev))  # warm-up: graphs of the call shapes
        one, alln = cpu.runner(pdict, fams, 1), cpu.runner(pdict, fams, cores)
        one(frames[0])
        alln(frames[0])
        lat, g_fps, c1, cn = [], [], [], []
        for _ in range(3):
            t0 = time.perf_counter()
            run = gpu_single_run(ctx, dev, prm, iters)
            g_fps.append(iters / (time.perf_counter() - t0))
            lat += run
            c1.append(cpu_run(one, frames, cpu_iters[0]))
            cn.append(cpu_run(alln, frames, cpu_iters[1]))
    finally:
        for d in dev:
            ctx.free_device(d[0])
    return dict(gpu_fps=round(float(np.median(g_fps)), 2), gpu_p50_ms=round(1e3 * pctl(lat, 50), 3),
                gpu_p99_ms=round(1e3 * pctl(lat, 99), 3), cpu_fps_1thread=round(float(np.median(c1)), 2),
                cpu_fps_all_cores=round(float(np.median(cn)), 2),
                runs="3 alternating runs: %d GPU calls, %d and %d CPU frames (1 thread, %d threads)"
                     % (iters, cpu_iters[0], cpu_iters[1], cores))


def host_input_batches(atg, frames, fams, batch, calls):
    """Batches whose frames are in pinned host memory (copied in by DMA each
    call, PCIe included): two contexts on two threads, so that one batch's
    copies overlap the other's kernels; frames/s over all calls."""
    pins = []
    for f in frames:
        pf = atg.PinnedFrame(*f.shape)
        pf.array[:] = f
        pins.append(pf)
    prm = atg.default_params()
    ctxs = [atg.Context(0) for _ in range(2)]
    try:
        for c in ctxs:
            c.set_families(fams)
        imgs = [pins[i % len(pins)].array for i in range(batch)]
        for c in ctxs:
            c.detect_batch(imgs, prm)  # warm-up

        def work(c):
            for _ in range(calls):
                c.detect_batch(imgs, prm)
        th = [threading.Thread(target=work, args=(c,)) for c in ctxs]
        t0 = time.perf_counter()
        for t in th:
            t.start()
        for t in th:
            t.join()
        wall = time.perf_counter() - t0
    finally:
        for c in ctxs:
            c.close()
        for pf in pins:
            pf.free()
    return dict(fps=round(2 * calls * batch / wall, 1), batch=batch, calls_per_context=calls, contexts=2,
                frame_bytes=int(frames[0].nbytes))


def pose_leg(atg, ctx, dets_per_frame, W, H, batch, cpu_seconds):
    """The pose of every detection of a batched call (the workload's
    detections repeated to the batch): the GPU (atg_estimate_poses, one
    launch, host arrays in and out) against the reference's own
    estimate_tag_pose (oracle/_ref/libposeref.so) on this host, one
    thread, over a bounded sample; the errors of the sample compared."""
    import ctypes as C
    from atg_testlib import POSE_IN_DTYPE, POSE_REF_SO
    rows = [d for k in range(batch) for d in dets_per_frame[k % len(dets_per_frame)]]
    tags = np.zeros(len(rows), POSE_IN_DTYPE)
    for i, d in enumerate(rows):
        tags[i] = (d["H"], d["p"], 0.1, 1400.0, 1400.0, W / 2.0, H / 2.0)
    ctx.estimate_poses(tags[:64])
    reps = 10
    t0 = time.perf_counter()
    for _ in range(reps):
        got = ctx.estimate_poses(tags)
    gpu_s = (time.perf_counter() - t0) / reps
    res = dict(tags=len(tags), gpu_ms_per_call=round(1e3 * gpu_s, 3), gpu_tags_per_s=round(len(tags) / gpu_s, 1),
               n_iters=50, camera="f 1400 px, centre at the frame centre, tag 0.1 m")
    if POSE_REF_SO.exists():
        L = C.CDLL(str(POSE_REF_SO))
        L.ref_estimate_many.argtypes = [C.c_void_p, C.c_int64, C.c_void_p]
        n, done, t0 = 0, [], time.perf_counter()
        while time.perf_counter() - t0 < cpu_seconds and n < len(tags):
            k = min(64, len(tags) - n)
            part = np.ascontiguousarray(tags[n:n + k])
            e = np.zeros(k)
            L.ref_estimate_many(part.ctypes.data, k, e.ctypes.data)
            done.append(e)
            n += k
        cpu_s = time.perf_counter() - t0
        e_ref = np.concatenate(done)
        e_gpu = np.where(got["best"][:n] == 1, got["err1"][:n], got["err2"][:n])
        res.update(cpu_tags_per_s_1thread=round(n / cpu_s, 1), cpu_sample_tags=n,
                   cpu_kind="reference", gpu_over_cpu=round(len(tags) / gpu_s / (n / cpu_s), 1),
                   errors_within_1e9=bool((np.abs(e_gpu - e_ref) <= 1e-9 * np.abs(e_ref) + 1e-18).all()))
    return res


def kernel_profile(ctx, batch, prm, calls=5):
    """Device time of each marked kernel of a batched call: HIP events
    between the kernels on the chain's stream (the calls without a graph),
    averaged over `calls` calls."""
    ctx.set_profiling(2)
    try:
        ctx.detect_batch(batch, prm, on_device=True)
        acc = {}
        for _ in range(calls):
            ctx.detect_batch(batch, prm, on_device=True)
            for name, ms in ctx.stage_times():
                acc[name] = acc.get(name, 0.0) + ms / calls
    finally:
        ctx.set_profiling(0)
    return acc


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--batch", type=int, default=256)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-batching", action="store_true", help="skip the uneven-batch (config 4) measurement")
    ap.add_argument("--no-configs", action="store_true", help="skip the 640x480 and 3840x2160 comparisons")
    ap.add_argument("--no-host-input", action="store_true", help="skip the batches from pinned host memory")
    ap.add_argument("--no-pose", action="store_true", help="skip the pose of the batch's detections")
    ap.add_argument("--streams", type=int, default=3, help="contexts (streams) the batched steps are spread over")
    ap.add_argument("--serve-seconds", type=float, default=2.0)
    ap.add_argument("--serve-loads", default="500,1000,2000,4000,8000,20000")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: frames are independent, so more GPUs would each run this same "
              "single-GPU path on their own share of frames (no collective); --gpus %d is not supported"
              % args.gpus)
        sys.exit(2)

    from atg_testlib import Oracle, default_params, load_atg, same_dets
    atg = load_atg()
    log("generating frames")
    frames = frames_config2(atg)
    W, H = frames[0].shape[1], frames[0].shape[0]
    fams = [("tag36h11", 2)]
    pdict = default_params()
    ctx = atg.Context(0)
    ctx.set_families(fams)
    prm = atg.params_from_dict(pdict)
    dev = [ctx.to_device(f) for f in frames]

    # parity: batch and single-frame outputs against the oracle
    log("checking parity against the oracle")
    orc = Oracle()
    want = [orc.detect(f, pdict, fams) for f in frames]
    got_b = ctx.detect_batch(dev, prm, on_device=True)
    got_s = [ctx.detect(None, prm, on_device_ptr=d) for d in dev]
    bit_exact = all(same_dets(w["dets"], g["dets"]) and w["nquads"] == g["nquads"] for w, g in zip(want, got_b)) and \
        all(same_dets(w["dets"], g["dets"]) and w["nquads"] == g["nquads"] for w, g in zip(want, got_s))
    ndets = sum(len(w["dets"]) for w in want)

    # batched path: the headline
    log("timing the batched path")
    batch = [dev[i % len(dev)] for i in range(args.batch)]
    # one context, calls back to back: each call's host work (planning,
    # launch, results) between the device's chains
    for _ in range(args.warmup):
        ctx.detect_batch(batch, prm, on_device=True)
    step_t = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        t = time.perf_counter()
        ctx.detect_batch(batch, prm, on_device=True)
        step_t.append(time.perf_counter() - t)
    wall1 = time.perf_counter() - t0
    chain_ms = ctx.chain_ms()
    # the headline: the same K steps spread over `streams` contexts, each a
    # chain on its own stream driven by its own host thread, so that one
    # call's host work and its low-occupancy phases overlap another's chain
    extra = [atg.Context(0) for _ in range(args.streams - 1)]
    lanes = [ctx] + extra
    for c in extra:
        c.set_families(fams)
        for _ in range(args.warmup):
            c.detect_batch(batch, prm, on_device=True)
    share = [args.steps // len(lanes) + (1 if i < args.steps % len(lanes) else 0) for i in range(len(lanes))]

    def lane_run(c, k):
        for _ in range(k):
            c.detect_batch(batch, prm, on_device=True)

    threads = [threading.Thread(target=lane_run, args=(lanes[i], share[i])) for i in range(len(lanes))]
    t0 = time.perf_counter()
    for th in threads:
        th.start()
    for th in threads:
        th.join()
    wall = time.perf_counter() - t0
    for c in extra:
        c.close()
    batched_fps = args.batch * args.steps / wall

    # per-kernel device times of the batched call, and the bytes they move
    log("per-kernel device times")
    kt = kernel_profile(ctx, batch, prm)
    ctx.detect_batch(dev, prm, on_device=True)
    counts = [frame_counts(ctx, i, W, H) for i in range(len(dev))]
    reps = args.batch / len(dev)
    sb = {"a": 0.0, "b": 0.0, "c": 0.0}
    model = {}
    for n in counts:
        for k, v in survey_bytes(n).items():
            sb[k] += v * reps
        for k, v in traffic_model(n).items():
            model[k] = model.get(k, 0.0) + v * reps
    # the roofline: the longest kernel of the HBM-bound stages (a)-(c), all
    # of its stage's algorithmic bytes credited to it (stages d-e have no
    # algorithmic byte figure; the longest kernel of the whole chain is
    # reported beside it)
    top, top_ms = max(kt.items(), key=lambda kv: kv[1])
    dom, dom_ms = max(((k, v) for k, v in kt.items() if stage_of(k) in sb), key=lambda kv: kv[1])
    st = stage_of(dom)
    traffic_tab = measured_traffic(args.batch)
    kernels = MARK_KERNELS.get(dom, [])
    traffic = sum(traffic_tab[k] for k in kernels if k in traffic_tab) or None
    dom_bytes = sb[st]
    bytes_note = "stage (%s) algorithmic bytes of the batch (all of them credited to this kernel)" % st
    ach = dom_bytes / (dom_ms * 1e-3) / 1e9
    roof = dict(kernel=dom, kernels=kernels, stage=st, bound="hbm", achieved=round(ach, 3), peak=HBM_PEAK_GBS,
                unit="GB/s", frac=round(ach / HBM_PEAK_GBS, 6), traffic=traffic, bytes_per_launch=int(dom_bytes),
                ms_per_launch=round(dom_ms, 4), bytes=bytes_note,
                longest_kernel={"mark": top, "stage": stage_of(top), "ms": round(top_ms, 4)})
    t_ac = sum(v for k, v in kt.items() if stage_of(k) in sb)
    b_ac = sum(sb.values())
    groups = {}
    for k, v in kt.items():
        groups[k.split(".")[0]] = groups.get(k.split(".")[0], 0.0) + v
    ac_traffic = sum(v for k, v in traffic_tab.items()
                     if any(k in MARK_KERNELS[m] for m in MARK_KERNELS if stage_of(m) in sb)) or None
    stages_a_c = {"bytes": int(b_ac), "bytes_per_stage": {k: int(v) for k, v in sb.items()}, "ms": round(t_ac, 4),
                  "GB/s": round(b_ac / (t_ac * 1e-3) / 1e9, 2), "frac": round(b_ac / (t_ac * 1e-3) / 1e9 / HBM_PEAK_GBS, 6),
                  "measured_traffic": ac_traffic,
                  "traffic_over_algorithmic": round(ac_traffic / b_ac, 2) if ac_traffic else None}
    model_gbs = {g: {"ms": round(groups[g], 4), "model_bytes": int(model[g]),
                     "GB/s": round(model[g] / (groups[g] * 1e-3) / 1e9, 2)} for g in model if g in groups}

    cores = os.cpu_count() or 1
    cpu = None if args.no_cpu else CpuRef()
    # the headline configuration one frame at a time, GPU and CPU alternating
    log("single-frame path against the CPU detector")
    if cpu is not None:
        c2 = compare_single(atg, ctx, cpu, frames, pdict, fams, 10 * max(1, args.steps // 2), (30, 90), cores)
    else:
        ctx.set_families(fams)
        gpu_single_run(ctx, dev, prm, len(dev))
        t0 = time.perf_counter()
        lat = gpu_single_run(ctx, dev, prm, 10 * max(1, args.steps // 2))
        c2 = dict(gpu_fps=round(len(lat) / (time.perf_counter() - t0), 2), gpu_p50_ms=round(1e3 * pctl(lat, 50), 3),
                  gpu_p99_ms=round(1e3 * pctl(lat, 99), 3))
    ctx.set_families(fams)
    configs = None
    if cpu is not None and not args.no_configs:
        log("configurations 1 and 3 one frame at a time")
        c1_frames = [atg.synth_frame(1, s)[0] for s in range(8)]
        c3_frames = [atg.synth_frame(3, s)[0] for s in range(2)]
        configs = {
            "640x480 tag36h11 decimate 2 (8 frames)": compare_single(atg, ctx, cpu, c1_frames, pdict, fams, 80,
                                                                    (60, 120), cores),
            "3840x2160 decimate 1, tag36h11 (2) + tagStandard52h13 (1), clutter (2 frames)": compare_single(
                atg, ctx, cpu, c3_frames, default_params(quad_decimate=1.0),
                [("tag36h11", 2), ("tagStandard52h13", 1)], 20, (16, 30), cores),
        }
        ctx.set_families(fams)
    host_in = None
    if not args.no_host_input:
        log("batches from pinned host memory")
        host_in = host_input_batches(atg, frames, fams, args.batch, max(2, args.steps // 4))
    batching = None
    if not args.no_batching:
        log("config 4: uneven batch")
        batching = batching_config4(atg, ctx, max(2, args.steps // 4))
        ctx.set_families(fams)
    pose = None
    if not args.no_pose:
        log("pose of the batch's detections")
        pose = pose_leg(atg, ctx, [w["dets"] for w in want], W, H, args.batch, 0.0 if args.no_cpu else 3.0)
    # a plain copy kernel of known size: the reference for the traffic
    # counters and for what HBM delivers to a simple kernel here
    copy_gbs, _ = ctx.copy_bandwidth(1 << 29)
    serving = None
    if not args.no_serve:
        serving = serving_sweep(atg, ctx, [int(x) for x in args.serve_loads.split(",")], args.serve_seconds,
                                check_ref=not args.no_cpu)
    cpu_line = None
    if cpu is not None:
        cpu_line = dict(value=c2["cpu_fps_all_cores"], unit="frames/s", cores=cores, kind=cpu.kind,
                        value_1thread=c2["cpu_fps_1thread"], cores_1thread=1, cpu_model=cpu_model(),
                        sample="the 10 workload frames cycled: median of 3 runs of 30 frames (1 thread) and "
                               "90 frames (%d threads), alternating with GPU single-frame runs" % cores)
        cpu.close()
    out = {
        "metric": "frames/s @1080p tag36h11 batched vs single-frame; p50/p99 latency, bit-exact",
        "value": round(batched_fps, 2),
        "unit": "frames/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1000.0 * wall / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8 (stages a-c), f64/f32 as the reference (d-e)",
        "data": "synthetic",
        "config": {"workload": "10 seeded 1920x1080 frames, 20 tag36h11 tags each, decimate 2, "
                               "refine_edges 1, bits_corrected 2; batch of %d frames (the 10 cycled)" % args.batch,
                   "batch": args.batch, "frame": "1920x1080 u8"},
        "single_frame_fps": c2["gpu_fps"],
        "batched_over_single": round(batched_fps / c2["gpu_fps"], 2),
        "latency_ms_single": {"p50": c2["gpu_p50_ms"], "p99": c2["gpu_p99_ms"]},
        "latency_ms_batch": {"p50": round(1000 * pctl(step_t, 50), 3), "p99": round(1000 * pctl(step_t, 99), 3)},
        "batched_streams": args.streams,
        "batched_one_stream": {"fps": round(args.batch * args.steps / wall1, 2),
                               "ms_per_step": round(1000.0 * wall1 / args.steps, 3)},
        "bit_exact": bool(bit_exact),
        "detections_per_10_frames": int(ndets),
        "device_chain_ms": round(chain_ms, 3),
        "kernel_ms": {k: round(v, 4) for k, v in kt.items()},
        "roofline": roof,
        "stages_a_c": stages_a_c,
        "traffic_model": model_gbs,
        "copy_kernel_gbs": round(copy_gbs, 1),
        "single_frame_vs_cpu": c2,
        "configs_single_frame": configs,
        "batched_host_input": host_in,
        "batching_uneven": batching,
        "serving": serving,
        "pose": pose,
        "cpu_baseline": cpu_line,
    }
    for d in dev:
        ctx.free_device(d[0])
    ctx.close()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
