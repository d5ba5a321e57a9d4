"""The oracle (oracle/oracle.c) pinned against the reference.

- the reference's own golden vectors (test/golden/*.hex over
  test/images/*.pgm, 18 parameter sets, copied to tests/golden/ref_hex);
- the reference build (oracle/_ref), stage by stage, where it was built;
- committed fixtures made with the reference build on seeded synthetic
  frames (tests/golden/make_synth_golden.py);
- properties the mathematics fixes: run-length reconstruction, the
  connected-component partition by brute-force flood fill, component sizes,
  the homography of exact corners.
"""
import numpy as np
import pytest

from atg_testlib import (DET_DTYPE, DIGEST_KEYS, GOLDEN, NAMED_CASES, RECONCILE_CASES, STAGE_DIGEST_CASES,
                         Oracle, Reference, batch64_frames, default_params, det_rows, digest, frame_hash,
                         golden_cases, have_ref, load_atg, named_case, parse_case_args, parse_hex, read_pnm,
                         ref_images, same_dets, SYNTH_CASES, synth_case, crowded_frame)

CASES = golden_cases()


@pytest.fixture(scope="module")
def orc():
    return Oracle()


@pytest.fixture(scope="module")
def imgs():
    return [(p.name, read_pnm(p)) for p in ref_images()]


@pytest.mark.parametrize("name,args", CASES, ids=[c[0] for c in CASES])
def test_oracle_matches_reference_golden_hex(orc, imgs, name, args):
    p, fam, ham = parse_case_args(args)
    gold = parse_hex(GOLDEN / "ref_hex" / (name + ".hex"))
    for iname, img in imgs:
        got = det_rows(orc.detect(img, p, [(fam, ham)])["dets"])
        assert got == gold.get(iname, []), (name, iname)


@pytest.mark.skipif(not have_ref(), reason="reference build (oracle/_ref) not present")
@pytest.mark.parametrize("name,args", CASES[:10], ids=[c[0] for c in CASES[:10]])
def test_oracle_stages_match_reference_build(orc, imgs, name, args):
    ref = Reference()
    p, fam, ham = parse_case_args(args)
    p["nthreads"] = 1
    for iname, img in imgs:
        o = orc.detect(img, p, [(fam, ham)], stages=True)
        r = ref.detect(img, p, [(fam, ham)], stages=True)
        for k in o["stages"]:
            a, b = o["stages"][k], r["stages"][k]
            if isinstance(a, np.ndarray):
                assert a.shape == b.shape and np.array_equal(a, b), (name, iname, k)
        assert o["nquads"] == r["nquads"]
        assert same_dets(o["dets"], r["dets"])


ALL_CASES = ["c%d_%d" % c for c in SYNTH_CASES] + list(NAMED_CASES)


@pytest.fixture(scope="module")
def fx():
    return np.load(GOLDEN / "synth_expected.npz")


@pytest.mark.parametrize("tag", ALL_CASES)
def test_oracle_matches_reference_fixtures(orc, fx, tag):
    """Detections, quad count and the count before reconcile equal the
    reference build's on every seeded case (configurations, reconcile
    frames, every family)."""
    atg = load_atg()
    img, p, fams = named_case(atg, tag)
    assert int(fx[tag + "_hash"]) == frame_hash(img), tag
    o = orc.detect(img, p, fams)
    assert o["dets"].tobytes() == fx[tag + "_dets"].tobytes(), tag
    assert o["nquads"] == int(fx[tag + "_nquads"]), tag
    assert o["nraw"] == int(fx[tag + "_nraw"]), tag


def test_reconcile_cases_remove_detections(fx):
    """The reconcile fixtures are frames on which the reference removes
    overlapping detections of one tag (fast/decode.c:860-891)."""
    removed = [int(fx[t + "_nraw"]) - len(fx[t + "_dets"]) for t in RECONCILE_CASES]
    assert all(r > 0 for r in removed), removed


@pytest.mark.parametrize("tag", STAGE_DIGEST_CASES)
def test_oracle_stage_digests_match_reference(orc, fx, tag):
    """Every stage array (quad image, threshold, runs, union-find roots and
    sizes, ok flags, cluster ids and order, points, quads) of production-
    size and heavy-clustering frames equals the reference build's dump."""
    atg = load_atg()
    img, p, fams = named_case(atg, tag)
    st = orc.detect(img, p, fams, stages=True)["stages"]
    for k in DIGEST_KEYS:
        assert digest(st[k]) == int(fx[tag + "_sd_" + k]), (tag, k)


def test_oracle_batch64_matches_reference(orc, fx):
    """The 64 uneven frames (16 each of 640x480 .. 2448x2048) of the
    batching workload: each frame's detections equal the reference's."""
    atg = load_atg()
    frames = batch64_frames(atg)
    p, fams = default_params(), [("tag36h11", 2)]
    for i, f in enumerate(frames):
        assert frame_hash(f) == int(fx["b64_hash"][i])
        o = orc.detect(f, p, fams)
        assert len(o["dets"]) == int(fx["b64_ndets"][i]), i
        assert o["nquads"] == int(fx["b64_nquads"][i]), i
        assert digest(o["dets"]) == int(fx["b64_dets"][i]), i


def reconcile_lists():
    z = np.load(GOLDEN / "reconcile_cases.npz")
    d, off, keep = z["dets"], z["offsets"], z["keep"]
    return [(d[off[i]:off[i + 1]], keep[off[i]:off[i + 1]]) for i in range(len(off) - 1)]


def test_oracle_reconcile_matches_reference_lists(orc):
    """The oracle's reconcile keeps what the reference's keeps on 400
    random lists (overlaps, nesting, shared corners, ties, both windings)."""
    removed = 0
    for k, (d, keep) in enumerate(reconcile_lists()):
        got = orc.reconcile(d)
        assert got.tobytes() == d[keep].tobytes(), k
        removed += len(d) - len(got)
    assert removed > 100


@pytest.mark.skipif(not have_ref(), reason="reference build (oracle/_ref) not present")
def test_reference_reconcile_reproduces_fixture():
    ref = Reference()
    for k, (d, keep) in enumerate(reconcile_lists()[:50]):
        assert ref.reconcile(d).tobytes() == d[keep].tobytes(), k


@pytest.mark.skipif(not have_ref(), reason="reference build (oracle/_ref) not present")
def test_oracle_crowded_frame_matches_reference():
    """A frame of 15530 quad candidates and 1250 detections: the oracle's
    stages and detections equal the reference build's."""
    img = crowded_frame(load_atg())
    p, fams = default_params(quad_decimate=1.0), [("tag16h5", 0)]
    o = Oracle().detect(img, p, fams, stages=True)
    r = Reference().detect(img, p, fams, stages=True)
    assert o["nquads"] == r["nquads"] > 4096 and len(r["dets"]) > 1024
    assert same_dets(o["dets"], r["dets"])
    for k in ("rep", "cluster_ids", "pts", "quads"):
        assert np.array_equal(o["stages"][k], r["stages"][k]), k


def _square(cx, cy, s, rot=0.0, cw=False):
    a = rot + np.arange(4) * np.pi / 2
    pts = np.stack([cx + s * np.cos(a), cy + s * np.sin(a)], 1)
    return (pts[::-1] if cw else pts).reshape(8)


def _dets(*rows):
    d = np.zeros(len(rows), DET_DTYPE)
    for i, (tid, fam, ham, margin, p) in enumerate(rows):
        d[i]["id"], d[i]["family"], d[i]["hamming"], d[i]["margin"], d[i]["p"] = tid, fam, ham, margin, p
    return d


def test_reconcile_hand_cases(orc):
    """Cases whose outcome the rule fixes (of two overlapping detections of
    one tag and family, the one with fewer corrected bits stays, then the
    one with the larger margin; detections of other tags or families, or
    apart, all stay), for the geometry that decides overlap: crossing
    edges, one quad inside the other, both windings, disjoint quads."""
    big, small, far = _square(50, 50, 20), _square(50, 50, 5, 0.3), _square(150, 50, 10)
    shifted = _square(60, 55, 20, 0.2, cw=True)
    # crossing edges: the lower hamming stays, wherever it is in the list
    for a, b in ((0, 1), (1, 0)):
        d = _dets((7, 0, a, 50.0, big), (7, 0, b, 50.0, shifted))
        got = orc.reconcile(d)
        assert len(got) == 1 and got[0]["hamming"] == 0
    # nested (no edge crosses): still one tag; equal hamming, larger margin stays
    d = _dets((7, 0, 1, 30.0, big), (7, 0, 1, 60.0, small))
    got = orc.reconcile(d)
    assert len(got) == 1 and got[0]["margin"] == 60.0
    d = _dets((7, 0, 1, 60.0, small), (7, 0, 1, 30.0, big))
    assert orc.reconcile(d)[0]["margin"] == 60.0
    # apart, other id, other family: all stay, in order
    d = _dets((7, 0, 0, 50.0, big), (7, 0, 0, 50.0, far), (8, 0, 0, 50.0, shifted), (7, 1, 0, 50.0, small))
    assert orc.reconcile(d).tobytes() == d.tobytes()
    # a chain of three over one tag: the best of them stays
    d = _dets((3, 0, 1, 40.0, big), (3, 0, 0, 20.0, shifted), (3, 0, 2, 90.0, small))
    got = orc.reconcile(d)
    assert len(got) == 1 and got[0]["hamming"] == 0


def _flood_partition(t):
    """Components of the thresholded image by flood fill over pixels, with
    the reference's connectivity (fast/unionfind.c:112-146): runs cover
    columns 0..w-2; white is 8-connected, black 4-connected, 127 joins
    nothing; a link between two rows needs the lower pixel in a column >= 1."""
    h, w = t.shape
    lab = -np.ones((h, w), np.int64)
    nxt = 0

    def links(y, x):
        v = t[y, x]
        for nx in (x - 1, x + 1):
            if 0 <= nx < w - 1 and t[y, nx] == v:
                yield y, nx
        for ny in (y - 1, y + 1):
            if not 0 <= ny < h:
                continue
            for nx in ((x - 1, x, x + 1) if v == 255 else (x,)):
                if not 0 <= nx < w - 1 or t[ny, nx] != v:
                    continue
                lower_x = x if y > ny else nx
                if lower_x >= 1:
                    yield ny, nx

    for y in range(h):
        for x in range(w - 1):
            if t[y, x] == 127 or lab[y, x] >= 0:
                continue
            lab[y, x] = nxt
            stack = [(y, x)]
            while stack:
                cy, cx = stack.pop()
                for ny, nx in links(cy, cx):
                    if lab[ny, nx] < 0:
                        lab[ny, nx] = nxt
                        stack.append((ny, nx))
            nxt += 1
    return lab


def test_runs_reconstruct_threshold_image(orc):
    atg = load_atg()
    img, _ = atg.synth_frame(1, 7)
    st = orc.detect(img, default_params(), [("tag36h11", 2)], stages=True)["stages"]
    t, runs, ro = st["t"], st["runs"], st["row_off"]
    h, w = t.shape
    rec = np.full((h, w - 1), 127, np.uint8)
    for y in range(h):
        for s, e, v in runs[ro[y]:ro[y + 1]]:
            rec[y, s:e + 1] = v
    assert np.array_equal(rec, t[:, :w - 1])


def test_partition_and_sizes_by_flood_fill(orc):
    atg = load_atg()
    img, _ = atg.synth_frame(1, 3)
    img = img[:200, :240].copy()
    st = orc.detect(img, default_params(), [("tag36h11", 2)], stages=True)["stages"]
    t, runs, ro, rep, csize = st["t"], st["runs"], st["row_off"], st["rep"], st["csize"]
    h, w = t.shape
    lab = _flood_partition(t)
    # node of each run pixel
    node = -np.ones((h, w - 1), np.int64)
    for y in range(h):
        for i in range(ro[y], ro[y + 1]):
            s, e, _ = runs[i]
            node[y, s:e + 1] = i
    m = node >= 0
    # same flood component <=> same representative
    pairs = {}
    for a, b in zip(lab[:, :w - 1][m], rep[node[m]]):
        pairs.setdefault(int(a), set()).add(int(b))
    assert all(len(v) == 1 for v in pairs.values())
    inv = {}
    for a, bs in pairs.items():
        inv.setdefault(next(iter(bs)), set()).add(a)
    # a component may also take in last-column virtual nodes, never two floods
    assert all(len(v) == 1 for v in inv.values())
    # component pixel counts: run pixels plus joined virtual nodes
    counts = np.bincount(rep[node[m]], minlength=len(rep))
    vn = np.arange(ro[h], len(rep))
    vcount = np.bincount(rep[vn], minlength=len(rep))
    tot = counts + vcount
    assert np.array_equal(csize[node[m]], tot[rep[node[m]]])


def test_homography_of_exact_corners():
    """The detection homography maps the tag corners (+-1, +-1) onto the
    detected corners (fast/detector.c:100-105: p[i] = H * corner i)."""
    orc = Oracle()
    img = read_pnm(ref_images()[0])
    dets = orc.detect(img, default_params(), [("tag36h11", 1)])["dets"]
    assert len(dets) > 0
    corners = [(-1, 1), (1, 1), (1, -1), (-1, -1)]
    for d in dets:
        H = d["H"].reshape(3, 3)
        for i, (x, y) in enumerate(corners):
            v = H @ np.array([x, y, 1.0])
            assert abs(v[0] / v[2] - d["p"][2 * i]) < 1e-9
            assert abs(v[1] / v[2] - d["p"][2 * i + 1]) < 1e-9
