"""Write the pose fixtures with the reference's own pose estimation
(oracle/_ref/libposeref.so: apriltag_pose.c, common/homography.c,
common/matd.c and common/svd22.c compiled unchanged). Run where
`make -C oracle ref` built it; the fixtures are data and travel with the
tests.

    python tests/golden/make_pose_golden.py

tests/golden/pose_cases.npz, per group g:
  <g>_in      the tags (atg_pose_in records: H, corners, tag size, camera)
  <g>_iters   the orthogonal-iteration steps the group ran with
  <g>_out     the reference's poses (atg_pose_out layout: hpose, the
              homography pose, both refined poses and errors, has2, best)
  <g>_R, <g>_t  the true pose (synthetic groups only)
groups:
  exact   tags seen by a pinhole camera at known poses (rotations up to
          70 degrees, 0.3-5 m), exact corners
  noisy   the same with 0.5 px of corner noise (a second minimum often)
  far     small tags 3-10 m away with 1 px of noise (ambiguous poses)
  dets    the detections of synthetic frames (the oracle's detector on
          640x480 and 1920x1080 frames, four families), cameras of the
          frame's size
  iters   detections and noisy tags at 0, 1, 5 and 100 steps
est_*: estimate_tag_pose itself (50 steps) on the dets group: est_err,
est_R, est_t.
h2p_*: homography_to_pose(H, fx, fy, cx, cy) on detection homographies
with the caller's own signs: h2p_H, h2p_cam (fx, fy, cx, cy), h2p_out.
"""
import sys
from pathlib import Path

import numpy as np

sys.path.insert(0, str(Path(__file__).resolve().parent.parent))
from atg_testlib import (GOLDEN, POSE_IN_DTYPE, Oracle, RefPose, default_params, load_atg, rotation,  # noqa: E402
                         tag_seen)


def synthetic(rng, n, noise, zlo, zhi, sizes):
    tags = np.zeros(n, POSE_IN_DTYPE)
    Rs, ts = np.zeros((n, 3, 3)), np.zeros((n, 3))
    for i in range(n):
        fx = rng.uniform(400, 1800)
        fy = fx * rng.uniform(0.95, 1.05)
        cx, cy = rng.uniform(300, 980), rng.uniform(200, 560)
        R = rotation(rng.normal(size=3), rng.uniform(0, 1.22))
        z = rng.uniform(zlo, zhi)
        t = np.array([rng.uniform(-0.4, 0.4) * z, rng.uniform(-0.25, 0.25) * z, z])
        size = float(rng.choice(sizes))
        H, p = tag_seen(R, t, size, fx, fy, cx, cy)
        p = p + rng.normal(0, noise, 8) if noise else p
        tags[i] = (H, p, size, fx, fy, cx, cy)
        Rs[i], ts[i] = R, t
    return tags, Rs, ts


def detection_tags(rng):
    atg = load_atg()
    orc = Oracle()
    cases = [(1, s, [("tag36h11", 2)]) for s in range(6)] + [(2, s, [("tag36h11", 2)]) for s in range(3)]
    cases += [(1, 10, [("tag16h5", 1)]), (1, 11, [("tagStandard41h12", 2)]), (1, 12, [("tag25h9", 2)])]
    rows = []
    for cfg, seed, fams in cases:
        img = atg.synth_frame(cfg, seed, family=atg.FAMILY_INDEX[fams[0][0]])[0]
        res = orc.detect(img, default_params(), fams)
        H, W = img.shape
        f = rng.uniform(0.6, 1.4) * W
        for d in res["dets"]:
            rows.append((d["H"], d["p"], float(rng.choice([0.08, 0.1, 0.162, 0.2])), f, f * rng.uniform(0.98, 1.02),
                         W / 2 + rng.uniform(-20, 20), H / 2 + rng.uniform(-20, 20)))
    tags = np.zeros(len(rows), POSE_IN_DTYPE)
    for i, r in enumerate(rows):
        tags[i] = r
    return tags


def main():
    rng = np.random.default_rng(20261204)
    ref = RefPose()
    out = {}

    def group(name, tags, iters, R=None, t=None):
        out[name + "_in"] = tags
        out[name + "_iters"] = np.array(iters)
        out[name + "_out"] = ref.poses(tags, iters)
        if R is not None:
            out[name + "_R"], out[name + "_t"] = R, t

    tags, R, t = synthetic(rng, 120, 0.0, 0.3, 5.0, [0.05, 0.1, 0.162, 0.3])
    group("exact", tags, 50, R, t)
    tags, R, t = synthetic(rng, 120, 0.5, 0.3, 5.0, [0.05, 0.1, 0.162, 0.3])
    group("noisy", tags, 50, R, t)
    tags, R, t = synthetic(rng, 80, 1.0, 3.0, 10.0, [0.05, 0.08])
    group("far", tags, 50, R, t)
    dets = detection_tags(rng)
    group("dets", dets, 50)
    mix = np.concatenate([dets[:20], out["noisy_in"][:20]])
    for k in (0, 1, 5, 100):
        group("iters%d" % k, mix, k)
    est = [ref.estimate_tag_pose(tg) for tg in dets]
    out["est_err"] = np.array([e for e, _, _ in est])
    out["est_R"] = np.array([R for _, R, _ in est])
    out["est_t"] = np.array([t for _, _, t in est])
    cams = np.column_stack([rng.uniform(400, 1800, len(dets)), rng.uniform(400, 1800, len(dets)),
                            rng.uniform(200, 900, len(dets)), rng.uniform(200, 600, len(dets))])
    cams[::3, 0] *= -1  # the caller's own signs, as the pose code's -fx
    out["h2p_H"] = dets["H"].copy()
    out["h2p_cam"] = cams
    out["h2p_out"] = np.array([ref.homography_to_pose(H, *c) for H, c in zip(dets["H"], cams)])
    np.savez_compressed(GOLDEN / "pose_cases.npz", **out)
    n2 = {g: int(out[g + "_out"]["has2"].sum()) for g in ("exact", "noisy", "far", "dets")}
    print("wrote", GOLDEN / "pose_cases.npz", "tags:", {g: len(out[g + "_in"]) for g in n2}, "second minima:", n2)


if __name__ == "__main__":
    main()
