"""The reference's known-answer tests on real photos
(test/test_detection.c and test/test_tag_pose_estimation.c: three photos
of test/data/, their true corners in test/data/*.txt to 1e-1 px).

tests/golden/photos/ holds the photos as the reference's own JPEG decoder
gives their pixels, and the true-corner files; photo_expected.npz the
reference build's detections (decimate 1, no edge refinement, tag36h11 at
2 bits: test_detection.c:90-93) and estimate_tag_pose of each (tag 0.16 m,
f 600 px, centre at the image centre: test_tag_pose_estimation.c:148-156);
both written by make_photo_golden.py."""
import re

import numpy as np
import pytest

from atg_testlib import GOLDEN, POSE_IN_DTYPE, Oracle, default_params, read_pnm, same_dets

PHOTOS = ["33369213973_9d9bb4cc96_c", "34085369442_304b6bafd9_c", "34139872896_defdb2f8d9_c"]
PARAMS = default_params(quad_decimate=1.0, refine_edges=False)
FAMILIES = [("tag36h11", 2)]
EXPECTED = np.load(GOLDEN / "photo_expected.npz")


def photo(name):
    return read_pnm(GOLDEN / "photos" / (name + ".pgm"))


def true_rows(name):
    rows = []
    for line in (GOLDEN / "photos" / (name + ".txt")).read_text().splitlines():
        if line.strip():
            nums = [float(v) for v in re.findall(r"-?\d+(?:\.\d+)?", line)]
            rows.append((int(nums[0]), np.array(nums[1:9])))
    return rows


def matches_truth(dets, name):
    """test_detection.c:40-140: detections sorted by id then corners, each
    equal to the true row (same id, corners within 1e-1)."""
    order = sorted(range(len(dets)), key=lambda i: (int(dets["id"][i]), *dets["p"][i]))
    rows = true_rows(name)
    if len(order) != len(rows):
        return False
    return all(int(dets["id"][i]) == rid and np.abs(dets["p"][i] - p).max() <= 1e-1
               for i, (rid, p) in zip(order, rows))


def pose_inputs(dets, shape):
    tags = np.zeros(len(dets), POSE_IN_DTYPE)
    tags["H"], tags["p"] = dets["H"], dets["p"]
    tags["tagsize"], tags["fx"], tags["fy"] = 0.16, 600.0, 600.0
    tags["cx"], tags["cy"] = shape[1] / 2.0, shape[0] / 2.0
    return tags


@pytest.mark.parametrize("name", PHOTOS)
def test_reference_detections_are_the_true_ones(name):
    assert matches_truth(EXPECTED[name + "_dets"], name)


@pytest.mark.parametrize("name", PHOTOS)
def test_oracle_on_photos(name):
    o = Oracle().detect(photo(name), PARAMS, FAMILIES)
    assert o["nquads"] == int(EXPECTED[name + "_nquads"])
    assert same_dets(o["dets"], EXPECTED[name + "_dets"])


@pytest.mark.gpu
@pytest.mark.parametrize("name", PHOTOS)
def test_gpu_on_photos(atg, name):
    """The GPU's detections on the photos equal the reference's bit for bit
    (so also the true corners), and the GPU pose of each equals the
    reference's estimate_tag_pose within 1e-9."""
    img = photo(name)
    ctx = atg.Context(0)
    try:
        ctx.set_families(FAMILIES)
        g = ctx.detect(img, atg.params_from_dict(PARAMS))
        assert g["nquads"] == int(EXPECTED[name + "_nquads"])
        assert same_dets(g["dets"], EXPECTED[name + "_dets"])
        assert matches_truth(g["dets"], name)
        poses = ctx.estimate_poses(pose_inputs(g["dets"], img.shape), 50)
    finally:
        ctx.close()
    R = np.where((poses["best"] == 1)[:, None], poses["R1"], poses["R2"]).reshape(-1, 3, 3)
    t = np.where((poses["best"] == 1)[:, None], poses["t1"], poses["t2"])
    e = np.where(poses["best"] == 1, poses["err1"], poses["err2"])
    assert np.abs(R - EXPECTED[name + "_R"]).max() < 1e-9
    assert (np.abs(t - EXPECTED[name + "_t"]).max(axis=1) / np.linalg.norm(EXPECTED[name + "_t"], axis=1)).max() < 1e-9
    want_e = EXPECTED[name + "_err"]
    assert (np.abs(e - want_e) <= 1e-9 * np.abs(want_e) + 1e-18).all()
    assert (e < 1.0).all()  # test_tag_pose_estimation.c:181's success bar
