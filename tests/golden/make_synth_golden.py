"""Write the fixtures made with the reference build (oracle/_ref, built from
the reference's own sources, single thread). Run in a container where
`make -C oracle ref` built it; the fixtures are data and travel with the
tests.

    python tests/golden/make_synth_golden.py

tests/golden/synth_expected.npz, per case of tests/atg_testlib.SYNTH_CASES
("c<config>_<seed>") and NAMED_CASES:
  <case>_dets     the detections (at_detection records)
  <case>_nquads   the quad-candidate count
  <case>_nraw     the detections before reconcile (fast/detector.c:325)
  <case>_hash     a digest of the frame (pins the seeded generator)
per case of STAGE_DIGEST_CASES, the digest of each stage array of the
reference's stage dump: <case>_sd_<array>;
for the uneven 64-frame batch (batch64_frames): b64_hash, b64_ndets,
b64_nquads, b64_dets (a digest of each frame's detections).

tests/golden/reconcile_cases.npz: random detection lists (ids, families,
hamming, margins and corners drawn so that overlaps, containment, exact
ties and both windings occur) and, for each, which detections the
reference's reconcile keeps (fast/decode.c:860-891).
"""
import sys
from pathlib import Path

import numpy as np

sys.path.insert(0, str(Path(__file__).resolve().parent.parent))
from atg_testlib import (DET_DTYPE, DIGEST_KEYS, GOLDEN, NAMED_CASES, STAGE_DIGEST_CASES, SYNTH_CASES,  # noqa: E402
                         Reference, batch64_frames, default_params, digest, frame_hash, load_atg, named_case)


def random_detection_lists(rng, nsets):
    """Detection lists for reconcile: quads around a few centres (so that
    many overlap), some nested, some sharing corners exactly, ids and
    families from small sets, margins and hamming with ties."""
    sets = []
    for k in range(nsets):
        n = int(rng.integers(2, 9))
        d = np.zeros(n, DET_DTYPE)
        centres = rng.uniform(0, 60, (3, 2)).round(1)
        for i in range(n):
            c = centres[rng.integers(0, 3)] + rng.normal(0, 4, 2).round(2)
            s = rng.choice([2.0, 6.0, 12.0, 20.0])
            th = rng.uniform(0, 2 * np.pi)
            pts = []
            for j in range(4):
                a = th + j * np.pi / 2 + rng.normal(0, 0.15)
                r = s * (1 + rng.normal(0, 0.1))
                pts.append((c[0] + r * np.cos(a), c[1] + r * np.sin(a)))
            pts = np.array(pts)
            if rng.random() < 0.2 and i > 0:
                pts = d["p"][i - 1].reshape(4, 2).copy()  # the same corners as the previous one
                if rng.random() < 0.5:
                    pts[rng.integers(0, 4)] += rng.choice([-1e-9, 1e-9, 0.5])
            if rng.random() < 0.3:
                pts = pts[::-1]  # the other winding
            if rng.random() < 0.3:
                pts = np.round(pts * 4) / 4  # collinear and touching edges
            d["p"][i] = pts.reshape(8)
            d["id"][i] = rng.integers(0, 3)
            d["family"][i] = rng.integers(0, 2) if rng.random() < 0.3 else 0
            d["hamming"][i] = rng.integers(0, 3)
            d["margin"][i] = rng.choice([10.0, 25.5, 25.5, 40.25])
            d["c"][i] = pts.mean(axis=0)
            d["H"][i] = 0.0  # reconcile never reads H
        sets.append(d)
    return sets


def main():
    atg = load_atg()
    ref = Reference()
    out = {}
    names = ["c%d_%d" % c for c in SYNTH_CASES] + list(NAMED_CASES)
    for tag in names:
        img, p, fams = named_case(atg, tag)
        r = ref.detect(img, p, fams, stages=tag in STAGE_DIGEST_CASES)
        raw = ref.raw_detections(img, p, fams)
        out[tag + "_dets"] = r["dets"]
        out[tag + "_nquads"] = np.int64(r["nquads"])
        out[tag + "_nraw"] = np.int64(len(raw))
        out[tag + "_hash"] = np.uint64(frame_hash(img))
        if r["stages"] is not None:
            for k in DIGEST_KEYS:
                out[tag + "_sd_" + k] = np.uint64(digest(r["stages"][k]))
        print(tag, img.shape, len(r["dets"]), "dets", len(raw), "before reconcile", r["nquads"], "quads",
              "(stage digests)" if r["stages"] is not None else "")
    frames = batch64_frames(atg)
    p, fams = default_params(), [("tag36h11", 2)]
    res = [ref.detect(f, p, fams) for f in frames]
    out["b64_hash"] = np.array([frame_hash(f) for f in frames], np.uint64)
    out["b64_ndets"] = np.array([len(r["dets"]) for r in res], np.int64)
    out["b64_nquads"] = np.array([r["nquads"] for r in res], np.int64)
    out["b64_dets"] = np.array([digest(r["dets"]) for r in res], np.uint64)
    print("batch64:", int(out["b64_ndets"].sum()), "detections")
    np.savez_compressed(GOLDEN / "synth_expected.npz", **out)

    sets = random_detection_lists(np.random.default_rng(860), 400)
    keep, total = [], 0
    for d in sets:
        r = ref.reconcile(d)
        # reconcile only removes: the kept ones are a subsequence of the input
        mask = np.zeros(len(d), bool)
        j = 0
        for i in range(len(d)):
            if j < len(r) and d[i].tobytes() == r[j].tobytes():
                mask[i] = True
                j += 1
        assert j == len(r)
        keep.append(mask)
        total += len(d) - len(r)
    flat = np.concatenate(sets)
    off = np.cumsum([0] + [len(d) for d in sets]).astype(np.int64)
    np.savez_compressed(GOLDEN / "reconcile_cases.npz", dets=flat, offsets=off, keep=np.concatenate(keep))
    print("reconcile: %d lists, %d detections, %d removed" % (len(sets), len(flat), total))


if __name__ == "__main__":
    main()
