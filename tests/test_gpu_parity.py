"""Parity of the HIP path with the reference, on the GPU.

Bit-exact bar (integer, index and floating-point fields alike): every
detection field, the quad count, and the stage arrays (quad image,
thresholded image, runs, union-find representatives, component sizes,
ok flags, cluster ids and order, points in cluster order, fitted quads)
equal the oracle's. The oracle is pinned against the reference by
tests/test_oracle.py.
"""
import ctypes as C

import numpy as np
import pytest

from atg_testlib import (GOLDEN, SYNTH_CASES, Detector, crowded_frame, default_params, det_rows, frame_hash,
                         golden_cases, parse_case_args, parse_hex, read_pnm, ref_images, same_dets, synth_case)

pytestmark = pytest.mark.gpu

STAGE_KEYS = ["qim", "t", "runs", "row_off", "rep", "csize", "ok", "cluster_ids", "ccount", "pts", "quads"]
CASES = golden_cases()


def gpu_detect(atg, ctx, img, p, fams):
    ctx.set_families(fams)
    return ctx.detect(img, atg.params_from_dict(p))


@pytest.mark.parametrize("name,args", CASES, ids=[c[0] for c in CASES])
def test_reference_golden_hex_on_gpu(atg, ctx, name, args):
    p, fam, ham = parse_case_args(args)
    gold = parse_hex(GOLDEN / "ref_hex" / (name + ".hex"))
    for path in ref_images():
        print("case", name, path.name, flush=True)
        got = det_rows(gpu_detect(atg, ctx, read_pnm(path), p, [(fam, ham)])["dets"])
        assert got == gold.get(path.name, []), (name, path.name)


@pytest.mark.parametrize("cfg,seed", SYNTH_CASES, ids=["c%d_%d" % c for c in SYNTH_CASES])
def test_synth_frames_match_reference_fixtures(atg, ctx, cfg, seed):
    fx = np.load(GOLDEN / "synth_expected.npz")
    tag = "c%d_%d" % (cfg, seed)
    img, p, fams = synth_case(atg, cfg, seed)
    assert int(fx[tag + "_hash"]) == frame_hash(img)
    print("case", tag, img.shape, flush=True)
    g = gpu_detect(atg, ctx, img, p, fams)
    assert g["nquads"] == int(fx[tag + "_nquads"])
    assert g["dets"].tobytes() == fx[tag + "_dets"].tobytes()


@pytest.mark.parametrize("cfg,seed", [(1, 0), (1, 5), (2, 0), (4, 1), (3, 0)])
def test_stages_match_oracle(atg, ctx, oracle, cfg, seed):
    img, p, fams = synth_case(atg, cfg, seed)
    print("case", cfg, seed, img.shape, flush=True)
    o = oracle.detect(img, p, fams, stages=True)
    gpu_detect(atg, ctx, img, p, fams)
    g = ctx.stages(0)
    for k in STAGE_KEYS:
        a, b = o["stages"][k], g[k]
        assert a.shape == b.shape and np.array_equal(a, b), k


def _blocky(w, h, b, seed):
    r = np.random.default_rng(seed)
    small = (r.random(((h + b - 1) // b, (w + b - 1) // b)) > 0.5).astype(np.uint8) * 255
    return np.ascontiguousarray(np.kron(small, np.ones((b, b), np.uint8))[:h, :w])


def _stripes(w, h, seed):
    r = np.random.default_rng(seed)
    img = np.zeros((h, w), np.uint8)
    for k in range(4):
        img[:, k::16] = 255
    n = (r.random((h // 4, w // 4)) > 0.97).astype(np.uint8) * 255
    return np.ascontiguousarray(img ^ np.kron(n, np.ones((4, 4), np.uint8)))


def test_noise_frame_retries_match_oracle(atg, ctx, oracle):
    """Random binary noise at decimate 1 with min_cluster_pixels 0: more
    points than pixels and more clusters than the first cluster table
    holds; the call re-runs with the worst-case sizes and equals the
    oracle stage by stage."""
    r = np.random.default_rng(5)
    img = (r.random((512, 512)) > 0.5).astype(np.uint8) * 255
    p, fams = default_params(quad_decimate=1.0, min_cluster_pixels=0), [("tag36h11", 2)]
    print("case noise 512x512", flush=True)
    o = oracle.detect(img, p, fams, stages=True)
    assert len(o["stages"]["pts"]) > 512 * 512 and len(o["stages"]["cluster_ids"]) > 16384
    g = gpu_detect(atg, ctx, img, p, fams)
    st = ctx.stages(0)
    for k in STAGE_KEYS:
        assert np.array_equal(o["stages"][k], st[k]), k
    assert same_dets(o["dets"], g["dets"]) and o["nquads"] == g["nquads"]


def test_4096_square_decimate_1_matches_oracle(atg, ctx, oracle):
    """A 4096x4096 frame at decimate 1 (over 2^24 union-find nodes)."""
    img, _ = atg.synth_frame(2, 3, width=4096, height=4096)
    p, fams = default_params(quad_decimate=1.0), [("tag36h11", 2)]
    print("case 4096x4096 decimate 1", flush=True)
    o = oracle.detect(img, p, fams)
    g = gpu_detect(atg, ctx, img, p, fams)
    assert o["nquads"] == g["nquads"] and same_dets(o["dets"], g["dets"]) and len(g["dets"]) > 0


@pytest.mark.parametrize("name", ["wide", "tall", "front_scratch"])
def test_extreme_shapes_match_oracle(atg, ctx, oracle, name):
    """Shapes at the union-find's limits: rows of over 1024 runs (past the
    front's register stage and the 704-run replay launch: the widest-row
    launch), 2048 quad rows of stripes spanning all of them (five root-jump
    passes), and a 3600-wide quad image whose front state is too large for
    LDS (HBM scratch without the debug switch)."""
    img, dec = {"wide": (_blocky(4096, 128, 2, 1), 2.0), "tall": (_stripes(256, 4096, 2), 2.0),
                "front_scratch": (_blocky(3600, 96, 2, 3), 1.0)}[name]
    p, fams = default_params(quad_decimate=dec), [("tag36h11", 2)]
    print("case", name, img.shape, flush=True)
    o = oracle.detect(img, p, fams, stages=True)
    g = gpu_detect(atg, ctx, img, p, fams)
    st = ctx.stages(0)
    for k in STAGE_KEYS:
        a, b = o["stages"][k], st[k]
        assert a.shape == b.shape and np.array_equal(a, b), k
    assert same_dets(o["dets"], g["dets"]) and o["nquads"] == g["nquads"]


def test_crowded_frame_matches_oracle(atg, ctx, oracle):
    """15530 quad candidates and 1250 detections in one frame (past any
    fixed per-frame capacity): stages and detections equal the oracle's."""
    img = crowded_frame(atg)
    p, fams = default_params(quad_decimate=1.0), [("tag16h5", 0)]
    print("case crowded 2048x2048 tag16h5", flush=True)
    o = oracle.detect(img, p, fams, stages=True)
    g = gpu_detect(atg, ctx, img, p, fams)
    st = ctx.stages(0)
    for k in STAGE_KEYS:
        assert np.array_equal(o["stages"][k], st[k]), k
    assert o["nquads"] == g["nquads"] > 4096 and len(g["dets"]) > 1024
    assert same_dets(o["dets"], g["dets"])


def test_max_nmaxima_limit(atg, ctx, oracle):
    """max_nmaxima 0..32 runs as the oracle (0 keeps no maxima: no quads;
    17..32 takes the 256-thread corner search for every cluster); outside
    it the call is refused with an error, not run."""
    img, p, fams = synth_case(atg, 2, 0)
    ctx.set_families(fams)
    for m in (0, 3, 17, 32):
        q = dict(p, max_nmaxima=m)
        g, o = ctx.detect(img, atg.params_from_dict(q)), oracle.detect(img, q, fams)
        assert g["nquads"] == o["nquads"] and same_dets(g["dets"], o["dets"]), m
        assert (g["nquads"] > 0) == (m > 3), m
    for bad in (-1, 33):
        with pytest.raises(RuntimeError):
            ctx.detect(img, atg.params_from_dict(dict(p, max_nmaxima=bad)))


@pytest.mark.parametrize("mcp", [0, 1, 3])
def test_small_min_cluster_pixels_match_oracle(atg, ctx, oracle, mcp):
    """min_cluster_pixels below 4: every component holds points and every
    cluster is a candidate; stages and detections equal the oracle's."""
    img, p, fams = synth_case(atg, 2, 1)
    p = dict(p, min_cluster_pixels=mcp)
    print("case mcp", mcp, flush=True)
    o = oracle.detect(img, p, fams, stages=True)
    g = gpu_detect(atg, ctx, img, p, fams)
    st = ctx.stages(0)
    for k in STAGE_KEYS:
        assert np.array_equal(o["stages"][k], st[k]), k
    assert o["nquads"] == g["nquads"] and same_dets(o["dets"], g["dets"])


@pytest.mark.parametrize("case,bands", [("c2", 2), ("c2", 3), ("c2", 16), ("c2", 64), ("c4", 5), ("c4", 64),
                                        ("tall", 64), ("wide", 7), ("c3", 13)])
def test_split_front_matches_oracle(atg, ctx, oracle, case, bands):
    """The union-find front split into row bands (swept from ports, joined
    band by band, swept again): roots, sizes, clusters and detections equal
    the oracle's for every band count."""
    if case == "tall":
        img, p, fams = _stripes(256, 4096, 2), default_params(), [("tag36h11", 2)]
    elif case == "wide":
        img, p, fams = _blocky(4096, 128, 2, 1), default_params(), [("tag36h11", 2)]
    else:
        img, p, fams = synth_case(atg, {"c2": 2, "c4": 4, "c3": 3}[case], {"c2": 1, "c4": 3, "c3": 0}[case])
    print("case split front", case, bands, img.shape, flush=True)
    o = oracle.detect(img, p, fams, stages=True)
    ctx.set_tune(1, 4, 1024, 2048, bands)
    try:
        g = gpu_detect(atg, ctx, img, p, fams)
        st = ctx.stages(0)
    finally:
        ctx.set_tune(1, 4, 1024, 2048, 0)
    for k in STAGE_KEYS:
        assert np.array_equal(o["stages"][k], st[k]), k
    assert same_dets(o["dets"], g["dets"]) and o["nquads"] == g["nquads"]


@pytest.mark.parametrize("cfg,seed", [(1, 1), (4, 3)])
def test_union_find_scratch_path_matches_oracle(atg, ctx, oracle, cfg, seed):
    """The union-find replay with its state in HBM scratch (the path of
    frames too wide for LDS) gives the oracle's roots and detections."""
    img, p, fams = synth_case(atg, cfg, seed)
    print("case", cfg, seed, img.shape, flush=True)
    o = oracle.detect(img, p, fams, stages=True)
    ctx.force_scratch(True)
    try:
        g = gpu_detect(atg, ctx, img, p, fams)
        st = ctx.stages(0)
    finally:
        ctx.force_scratch(False)
    for k in ("rep", "csize", "ok", "cluster_ids"):
        assert np.array_equal(o["stages"][k], st[k]), k
    assert same_dets(o["dets"], g["dets"]) and o["nquads"] == g["nquads"]


def test_batch_of_uneven_frames_equals_single(atg, ctx):
    """Frames of four sizes in one pass give what each gives alone."""
    frames = [synth_case(atg, 4, s)[0] for s in range(4)] + [synth_case(atg, 1, 0)[0]]
    p = default_params()
    ctx.set_families([("tag36h11", 2)])
    prm = atg.params_from_dict(p)
    alone = [ctx.detect(f, prm) for f in frames]
    batch = ctx.detect_batch(frames, prm)
    for a, b in zip(alone, batch):
        assert a["nquads"] == b["nquads"]
        assert same_dets(a["dets"], b["dets"])
    assert sum(len(a["dets"]) for a in alone) > 20


def test_large_batch_of_uneven_frames_equals_single(atg, ctx):
    """40 frames of four sizes in one call (the shapes of calls of more than
    32 frames: one-wave fit kernels, 128-thread moment blocks, the 448-run
    replay rows with the 704-run and widest-row launches for the 2448x2048
    rows) give what each frame gives alone."""
    frames = [synth_case(atg, 4, s)[0] for s in range(40)]
    p = default_params()
    ctx.set_families([("tag36h11", 2)])
    prm = atg.params_from_dict(p)
    print("case large batch: 40 frames, 10 each of 640x480 1280x720 1920x1080 2448x2048", flush=True)
    alone = [ctx.detect(f, prm) for f in frames]
    batch = ctx.detect_batch(frames, prm)
    for k, (a, b) in enumerate(zip(alone, batch)):
        assert a["nquads"] == b["nquads"], k
        assert same_dets(a["dets"], b["dets"]), k
    assert sum(len(a["dets"]) for a in alone) > 200


def test_repeatable(atg, ctx):
    img, p, fams = synth_case(atg, 2, 1)
    a = gpu_detect(atg, ctx, img, p, fams)
    b = gpu_detect(atg, ctx, img, p, fams)
    assert same_dets(a["dets"], b["dets"]) and a["nquads"] == b["nquads"]


def test_native_at_detect_api(atg):
    """at_detector_create / at_add_family / at_detect as a C caller uses them."""
    L = atg.lib()
    L.at_detector_create.restype = C.POINTER(Detector)
    L.at_detector_destroy.argtypes = [C.POINTER(Detector)]
    L.at_family_by_name.restype = C.c_void_p
    L.at_family_by_name.argtypes = [C.c_char_p]
    L.at_add_family.restype = C.c_bool
    L.at_add_family.argtypes = [C.POINTER(Detector), C.c_void_p, C.c_int]
    L.at_detect.argtypes = [C.POINTER(Detector), C.c_void_p, C.c_int64, C.c_int64, C.c_int64]
    td = L.at_detector_create()
    try:
        assert L.at_add_family(td, L.at_family_by_name(b"tag36h11"), 1)
        assert not L.at_add_family(td, L.at_family_by_name(b"tag36h11"), 4)
        gold = parse_hex(GOLDEN / "ref_hex" / "default.hex")
        for path in ref_images():
            img = read_pnm(path)
            # a wider row stride than the width
            buf = np.zeros((img.shape[0], img.shape[1] + 13), np.uint8)
            buf[:, :img.shape[1]] = img
            L.at_detect(td, buf.ctypes.data, img.shape[1], img.shape[0], buf.shape[1])
            d = td.contents
            raw = C.string_at(d.detections, d.ndetections * 168) if d.ndetections else b""
            dets = np.frombuffer(raw, atg.DET_DTYPE)
            assert det_rows(dets) == gold[path.name]
        # the stage times of the call (fast/detector.c:257-334): each stage ran
        tm = td.contents.timing
        stages = [tm.decimate, tm.threshold, tm.unionfind, tm.clusters, tm.fit_quads, tm.decode, tm.reconcile]
        assert all(t > 0 for t in stages), stages
        assert sum(stages) + tm.blur <= tm.total
        # under 8x8: nothing, as the reference
        L.at_detect(td, img.ctypes.data, 7, 7, img.shape[1])
        assert td.contents.ndetections == 0 and td.contents.nquads == 0
    finally:
        L.at_detector_destroy(td)


def test_device_pointer_input(atg, ctx):
    """A frame already in HBM (atg_detect with on_device) gives what the
    host-memory call gives."""
    img, p, fams = synth_case(atg, 1, 2)
    ctx.set_families(fams)
    prm = atg.params_from_dict(p)
    a = ctx.detect(img, prm)
    d = ctx.to_device(img)
    try:
        b = ctx.detect(None, prm, on_device_ptr=d)
    finally:
        ctx.free_device(d[0])
    assert same_dets(a["dets"], b["dets"]) and a["nquads"] == b["nquads"]


def test_upstream_api_golden_hex(tmp_path):
    """The upstream API (apriltag.h) from a C program, every golden case:
    apriltag_detector_detect gives the reference's detections."""
    import subprocess
    from test_abi import build_upstream_demo
    exe = build_upstream_demo(tmp_path / "upstream_demo")
    imgs = [str(p) for p in ref_images()]
    for name, args in CASES:
        print("case", name, flush=True)
        out = subprocess.run([str(exe)] + args.split() + imgs, check=True, capture_output=True, text=True,
                             timeout=120).stdout
        got = {}
        for line in out.splitlines():
            if line.startswith("#"):
                continue
            tok = line.split()
            img = tok[0].rsplit("/", 1)[-1]
            got.setdefault(img, []).append([int(tok[1]), int(tok[2])] + [float.fromhex(t) for t in tok[3:]])
        gold = parse_hex(GOLDEN / "ref_hex" / (name + ".hex"))
        for path in ref_images():
            assert got.get(path.name, []) == gold.get(path.name, []), (name, path.name)
        # td->tp: the upstream profile's stages, in order, stamped as they end
        tp = [l[len("# tp "):].split("|") for l in out.splitlines() if l.startswith("# tp ")]
        names = ["init", "decimate", "blur/sharp", "threshold", "unionfind", "make clusters",
                 "fit quads to clusters", "decode+refinement", "reconcile"]
        assert [n for n, _ in tp[:9]] == names
        us = [int(u) for _, u in tp[:9]]
        assert us == sorted(us) and us[-1] > 0
