"""Benchmark of the MI355X AprilTag detection path.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--batch B]

Workload: 10 seeded synthetic 1920x1080 frames, 20 tag36h11 tags each
(40-300 px, tilt <= 45 deg, occluding bars, value-noise background,
blur 0-1.5, noise 2-8), decimate 2, refine_edges 1, bits_corrected 2.
The frames are resident in HBM before the timed region.

A step is one pass of the detection chain over one batch of B frames
(the 10 frames repeated); `value` is the batched frames/s. Beside it:
the single-frame path (one frame per call) frames/s and its p50/p99
latency, bit-exactness of both against the oracle, per-stage device
times, the roofline of the dominant kernel, and the reference CPU
detector timed on the host (1 thread and all cores).
"""
import argparse
import json
import os
import sys
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


def measured_traffic(batch):
    """HBM bytes per launch of each kernel group at this batch size, as the
    rocprofv3 FETCH_SIZE and WRITE_SIZE runs measured them (scaled by the
    copy kernel's readings): the table of profiles/traffic_1080p_batch<B>.txt,
    if that profile exists."""
    path = REPO / "profiles" / ("traffic_1080p_batch%d.txt" % batch)
    out = {}
    if not path.exists():
        return out
    for line in path.read_text().splitlines():
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cells) >= 4 and cells[0].startswith("group:"):
            out[cells[0][len("group:"):].strip()] = int(cells[3])
    return out


def frame_counts(ctx, i, W, H):
    """The data-dependent sizes of frame i of the last call: R runs, N
    nodes, E union-find links, P points, C clusters, K points of the
    clusters that reach the quad fit, Q quads."""
    c = ctx.counters(i)
    st = ctx.stages(i)
    qw, qh = st["qw"], st["qh"]
    R = c[0]
    # the cluster sort: 11-bit radix passes over the key's bits; a frame
    # whose key and point index fit one 64-bit word (pcap = qw qh + 64)
    # sorts that word alone
    packed = c[11] + (qw * qh + 64).bit_length() <= 64
    return dict(W=W, H=H, qw=qw, qh=qh, R=R, N=R + qh, E=c[6], P=c[1], C=c[2], K=c[7], Q=c[3],
                S=(c[11] + 10) // 11, SB=24 if packed else 32)


def kernel_bytes(n):
    """Algorithmic HBM bytes per frame of each kernel group: every array
    it reads or writes, once (a 4-byte word per run, node, link and record
    slot; an 8-byte word per point; 48 bytes of moments per fitted point).
    Stages (a)-(c) follow the per-frame formulas of the byte model split by
    kernel, the cluster sort with the frame's own number of radix passes
    (S) and bytes a point per pass (SB: 8-byte words read by the histogram
    and read and written by the scatter, 24, or 12-byte key and index
    pairs, 32); labels count the ok roots' numbering (8N), the gathers of
    run labels and dense numbers (8N), t and the two label planes; points
    read t and plab (count pass) and t, plab and dlab (write pass) and
    write 16 bytes a point; (d)-(e) count the arrays of the fit and, for
    decode, the pixels each quad samples (refine: 4 edges x 16 samples x 25
    steps x 2 bilinear reads of 4 pixels)."""
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
        "labels": 16 * N + px + 8 * px,
        "points": (px + 4 * px) + (px + 8 * px) + 16 * P,
        "cluster_sort": n["S"] * n["SB"] * P,
        "clusters": 24 * P,
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


def cpu_baseline(frames, sample_s):
    """The reference detector (oracle/_ref, built from the reference's own
    sources) timed on this host at 1 thread and at all cores, on a bounded
    sample of the same frames; the C restatement if the reference build
    is absent."""
    from atg_testlib import Oracle, Reference, default_params, have_ref
    cores = os.cpu_count() or 1
    fams = [("tag36h11", 2)]
    res = {}
    if have_ref():
        ref = Reference()
        for nt in (1, cores):
            p = default_params(nthreads=nt)
            td = ref.detector(p, fams)
            n, t0 = 0, time.perf_counter()
            while time.perf_counter() - t0 < sample_s:
                img = frames[n % len(frames)]
                ref.lib.at_detect(td, img.ctypes.data, img.shape[1], img.shape[0], img.strides[0])
                n += 1
            dt = time.perf_counter() - t0
            ref.lib.at_detector_destroy(td)
            res[nt] = (n / dt, n)
        return dict(value=round(res[cores][0], 2), unit="frames/s", cores=cores, kind="reference",
                    value_1thread=round(res[1][0], 2), cores_1thread=1,
                    sample="%d and %d frames (the 10 workload frames, cycled; ~%.0f s each) at %d and 1 threads"
                           % (res[cores][1], res[1][1], sample_s, cores))
    orc = Oracle()
    p = default_params()
    n, t0 = 0, time.perf_counter()
    while time.perf_counter() - t0 < sample_s:
        orc.detect(frames[n % len(frames)], p, fams)
        n += 1
    dt = time.perf_counter() - t0
    return dict(value=round(n / dt, 2), unit="frames/s", cores=1, kind="port",
                sample="%d frames of the workload, ~%.0f s, 1 thread" % (n, sample_s))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--batch", type=int, default=256)
    ap.add_argument("--cpu-sample-s", type=float, default=8.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-batching", action="store_true", help="skip the uneven-batch (config 4) measurement")
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

    # single-frame path
    log("timing the single-frame path")
    for _ in range(args.warmup):
        for d in dev:
            ctx.detect(None, prm, on_device_ptr=d)
    lat = []
    t0 = time.perf_counter()
    for _ in range(max(1, args.steps // 2)):
        for d in dev:
            t = time.perf_counter()
            ctx.detect(None, prm, on_device_ptr=d)
            lat.append(time.perf_counter() - t)
    single_fps = len(lat) / (time.perf_counter() - t0)

    # batched path
    log("timing the batched path")
    batch = [dev[i % len(dev)] for i in range(args.batch)]
    for _ in range(args.warmup):
        ctx.detect_batch(batch, prm, on_device=True)
    step_t = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        t = time.perf_counter()
        ctx.detect_batch(batch, prm, on_device=True)
        step_t.append(time.perf_counter() - t)
    wall = time.perf_counter() - t0
    batched_fps = args.batch * args.steps / wall

    # per-stage device times of one batched call (HIP events on the stream)
    ctx.set_profiling(True)
    ctx.detect_batch(batch, prm, on_device=True)
    stages = ctx.stage_times()
    ctx.set_profiling(False)
    chain_ms = ctx.chain_ms()
    # algorithmic bytes of the batch, from the first 10 frames' counts
    ctx.detect_batch(dev, prm, on_device=True)
    counts = [frame_counts(ctx, i, W, H) for i in range(len(dev))]
    reps = args.batch / len(dev)
    kb = {}
    for n in counts:
        for k, v in kernel_bytes(n).items():
            kb[k] = kb.get(k, 0) + v * reps
    dom_name, dom_ms = max(stages, key=lambda kv: kv[1])
    ach = kb[dom_name] / (dom_ms * 1e-3) / 1e9
    bound = "hbm"
    traffic = measured_traffic(args.batch).get(dom_name)
    roof = dict(kernel=dom_name, bound=bound, achieved=round(ach, 3), peak=HBM_PEAK_GBS, unit="GB/s",
                frac=round(ach / HBM_PEAK_GBS, 6), traffic=traffic, bytes_per_launch=int(kb[dom_name]),
                ms_per_launch=round(dom_ms, 4))
    ac = ("decimate", "threshold_rle", "uf_edges", "uf_front", "uf_replay", "uf_stitch", "resolve", "labels",
          "points", "cluster_sort", "clusters")
    t_ac = sum(v for k, v in stages if k in ac)
    b_ac = sum(kb[k] for k in ac if k in kb)
    per_kernel = {k: {"ms": round(v, 4), "GB/s": round(kb.get(k, 0) / (v * 1e-3) / 1e9, 2)} for k, v in stages}

    batching = None
    if not args.no_batching:
        log("config 4: uneven batch")
        batching = batching_config4(atg, ctx, max(2, args.steps // 4))
        ctx.set_families(fams)
    # a plain copy kernel of known size: the reference for the traffic
    # counters and for what HBM delivers to a simple kernel here
    copy_gbs, _ = ctx.copy_bandwidth(1 << 29)
    serving = None
    if not args.no_serve:
        serving = serving_sweep(atg, ctx, [int(x) for x in args.serve_loads.split(",")], args.serve_seconds,
                                check_ref=not args.no_cpu)
    cpu = None if args.no_cpu else cpu_baseline(frames, args.cpu_sample_s)
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
        "single_frame_fps": round(single_fps, 2),
        "batched_over_single": round(batched_fps / single_fps, 2),
        "latency_ms_single": {"p50": round(1000 * pctl(lat, 50), 3), "p99": round(1000 * pctl(lat, 99), 3)},
        "latency_ms_batch": {"p50": round(1000 * pctl(step_t, 50), 3), "p99": round(1000 * pctl(step_t, 99), 3)},
        "bit_exact": bool(bit_exact),
        "detections_per_10_frames": int(ndets),
        "device_chain_ms": round(chain_ms, 3),
        "stage_ms": {k: round(v, 4) for k, v in stages},
        "kernel_gbs": per_kernel,
        "stages_a_c": {"bytes": int(b_ac), "ms": round(t_ac, 4), "GB/s": round(b_ac / (t_ac * 1e-3) / 1e9, 2)},
        "roofline": roof,
        "copy_kernel_gbs": round(copy_gbs, 1),
        "batching_uneven": batching,
        "serving": serving,
        "cpu_baseline": cpu,
    }
    for d in dev:
        ctx.free_device(d[0])
    ctx.close()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
