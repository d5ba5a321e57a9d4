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

from atg_testlib import (GOLDEN, Oracle, Reference, default_params, det_rows, frame_hash, golden_cases,
                         have_ref, load_atg, parse_case_args, parse_hex, read_pnm, ref_images, same_dets,
                         SYNTH_CASES, synth_case)

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


def test_oracle_matches_synth_fixtures(orc):
    atg = load_atg()
    fx = np.load(GOLDEN / "synth_expected.npz")
    for cfg, seed in SYNTH_CASES:
        tag = "c%d_%d" % (cfg, seed)
        img, p, fams = synth_case(atg, cfg, seed)
        assert int(fx[tag + "_hash"]) == frame_hash(img), tag
        o = orc.detect(img, p, fams)
        assert o["dets"].tobytes() == fx[tag + "_dets"].tobytes(), tag
        assert o["nquads"] == int(fx[tag + "_nquads"]), tag


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
