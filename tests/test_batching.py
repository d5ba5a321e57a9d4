"""The batched form's work splits and graph replays on the GPU.

Work splits (atg_set_tune, chosen by atg_autotune) only change how the
work is cut into blocks: every split gives the detections of the default.
A call shape that repeats is replayed as a captured HIP graph: every
replay, with other work and synchronisations between them, equals the
eager launch of the same frames.
"""
import numpy as np
import pytest

from atg_testlib import same_dets, synth_case

pytestmark = pytest.mark.gpu

FAMS = [("tag36h11", 2)]


def uneven_frames(atg):
    return [synth_case(atg, 4, s)[0] for s in range(4)] + [synth_case(atg, 1, 0)[0]]


def same(a, b):
    return len(a) == len(b) and all(same_dets(x["dets"], y["dets"]) and x["nquads"] == y["nquads"]
                                    for x, y in zip(a, b))


@pytest.mark.parametrize("tune", [(2, 1, 128, 64, 0), (4, 4, 512, 1024, 1), (1, 2, 256, 32, 3), (2, 4, 128, 512, 64),
                                  (1, 4, 64, 4096, 7), (1, 4, 1024, 2048, 2)])
def test_work_splits_give_the_same_detections(atg, ctx, tune):
    frames = uneven_frames(atg)
    ctx.set_families(FAMS)
    ctx.set_graphs(False)
    base = ctx.detect_batch(frames)
    ctx.set_tune(*tune)
    try:
        print("tune", tune, flush=True)
        got = ctx.detect_batch(frames)
    finally:
        ctx.set_tune(1, 4, 256, 256)
        ctx.set_graphs(True)
    assert same(base, got)
    assert sum(len(b["dets"]) for b in base) > 20


def test_autotune_keeps_detections(atg, ctx):
    frames = uneven_frames(atg)
    ctx.set_families(FAMS)
    base = ctx.detect_batch(frames)
    print("autotune", flush=True)
    tune, ms = ctx.autotune(frames)
    try:
        assert ms > 0 and ctx.get_tune() == tune
        assert same(base, ctx.detect_batch(frames))
    finally:
        ctx.set_tune(1, 4, 256, 256)


def test_graph_replays_equal_eager(atg):
    """Frames A and B of one shape, alternated, with a call of another shape
    (a new graph) and eager calls in between: every replay equals eager."""
    eager = atg.Context()
    graph = atg.Context()
    try:
        for c in (eager, graph):
            c.set_families(FAMS)
        eager.set_graphs(False)
        graph.set_graphs(True)
        a = [synth_case(atg, 1, s)[0] for s in (0, 1, 2)]
        b = [synth_case(atg, 1, s)[0] for s in (3, 4, 5)]
        other = [synth_case(atg, 4, 1)[0]]
        want_a, want_b, want_o = eager.detect_batch(a), eager.detect_batch(b), eager.detect_batch(other)
        for rep in range(4):
            print("replay", rep, flush=True)
            assert same(graph.detect_batch(a), want_a)
            assert same(eager.detect_batch(b), want_b)  # eager work between replays
            assert same(graph.detect_batch(b), want_b)
            if rep == 1:
                assert same(graph.detect_batch(other), want_o)  # another shape, then back
        # single-frame calls of one size (the at_detect pattern)
        for img in a + b:
            assert same([graph.detect(img)], [eager.detect(img)])
    finally:
        eager.close()
        graph.close()


def test_work_split_outside_the_kernels_bounds_is_refused(atg, ctx):
    """Blocks of more than 256 threads for the per-row kernels are refused
    on the host (no launch)."""
    with pytest.raises(RuntimeError):
        ctx.set_tune(1, 8, 256, 256)
    assert ctx.get_tune()[1] <= 4


def test_pinned_host_frames_equal_pageable(atg, ctx):
    """Host frames in page-locked memory are copied in by DMA as they lie;
    pageable ones through the context's staging: the same detections."""
    frames = uneven_frames(atg)
    ctx.set_families(FAMS)
    pins = []
    try:
        for f in frames:
            pf = atg.PinnedFrame(*f.shape)
            pf.array[:] = f
            pins.append(pf)
        a = ctx.detect_batch(frames)
        b = ctx.detect_batch([p.array for p in pins])
        c = ctx.detect_batch([p.array for p in pins])  # the graph of the DMA call shape, replayed
    finally:
        for pf in pins:
            pf.free()
    assert same(a, b) and same(a, c)
