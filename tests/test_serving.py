"""The serving loop on the GPU (atg_server_*).

Each frame's result in the mix of streams equals what the single-frame
entry point gives for it alone, bit for bit: detections and quad count.
Streams of the four frame sizes, frames with slowly drifting tags,
Poisson arrivals released by the server's own feeder, several lanes in
flight; then a budget so tight that a lane holds only a few frames; then
the submit path.
"""
import numpy as np
import pytest

from atg_testlib import same_dets

pytestmark = pytest.mark.gpu

SIZES = [(640, 480), (1280, 720), (1920, 1080), (2448, 2048)]
FAMS = [("tag36h11", 2)]


def stream_frames(atg, nstreams, nsteps, seed0=100):
    frames = {}
    for k in range(nstreams):
        w, h = SIZES[k % 4]
        spec = atg.synth_spec(1, w, h)
        for step in range(nsteps):
            frames[(k, step)] = atg.synth_stream_frame(spec, seed0 + k, step)
    return frames


def alone(atg, ctx, frames):
    ctx.set_families(FAMS)
    p = atg.default_params()
    return {key: ctx.detect(img, p) for key, img in frames.items()}


def check(results, want, nframes):
    assert len(results) == nframes
    seen = set()
    for r in results:
        assert r["status"] == 0, r
        key = (r["tag"] // 1000, r["tag"] % 1000)
        seen.add(key)
        w = want[key]
        assert r["nquads"] == w["nquads"], key
        assert same_dets(r["dets"], w["dets"]), key
        assert r["arrival_ns"] <= r["admit_ns"] <= r["done_ns"]
    assert len(seen) == nframes


def test_served_in_the_mix_equals_alone(atg, ctx):
    nstreams, nsteps = 8, 3
    frames = stream_frames(atg, nstreams, nsteps)
    want = alone(atg, ctx, frames)
    assert sum(len(w["dets"]) for w in want.values()) >= 40
    srv = atg.Server(FAMS, lanes=3, max_batch=16)
    try:
        ids = [srv.add_stream(*SIZES[k % 4]) for k in range(nstreams)]
        rng = np.random.default_rng(7)
        t = np.zeros(nstreams)
        arrivals = []
        for step in range(nsteps):
            for k in range(nstreams):
                t[k] += rng.exponential(2e6)  # ~500 frames/s a stream
                arrivals.append((ids[k], t[k], frames[(k, step)], k * 1000 + step))
        print("feeding", len(arrivals), "frames", flush=True)
        srv.feed(arrivals)
        srv.drain(120000)
        res = srv.poll()
        counts = srv.counts()
    finally:
        srv.close()
    check(res, want, nstreams * nsteps)
    assert counts["completed"] == nstreams * nsteps and counts["dropped"] == 0
    assert max(r["batch_frames"] for r in res) > 1  # frames shared batches


def test_tight_budget_and_submit(atg, ctx):
    """Lanes that hold only a few frames (admission by budget), frames
    handed in with atg_server_submit all at once."""
    nstreams, nsteps = 8, 2
    frames = stream_frames(atg, nstreams, nsteps, seed0=300)
    want = alone(atg, ctx, frames)
    srv = atg.Server(FAMS, lanes=2, max_batch=64, budget_bytes=1 << 30)
    try:
        ids = [srv.add_stream(*SIZES[k % 4]) for k in range(nstreams)]
        print("submitting", len(frames), "frames", flush=True)
        for step in range(nsteps):
            for k in range(nstreams):
                assert srv.submit(ids[k], frames[(k, step)], k * 1000 + step) == 0
        srv.drain(120000)
        res = srv.poll()
        counts = srv.counts()
    finally:
        srv.close()
    check(res, want, nstreams * nsteps)
    # a lane's arena holds about a third of the frames' arenas: they run
    # in several batches
    print(counts, flush=True)
    assert counts["lane_arena_bytes"] < (1 << 29)
    assert counts["batches"] >= 4
