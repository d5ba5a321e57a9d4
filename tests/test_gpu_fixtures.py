"""GPU parity against fixtures written by the reference build (oracle/_ref,
tests/golden/make_synth_golden.py), bit for bit:

- frames on which the reference's reconcile removes detections;
- every built-in family rendered and detected (640x480 and 1920x1080);
- the stage arrays of production-size frames (1920x1080, 2448x2048) and
  heavy-clustering 3840x2160 frames, by digest, on the default paths;
- the uneven 64-frame batch in one call, each frame against the
  reference's detections of it alone;
- the device's reconcile and id sort on 400 random detection lists.
"""
import numpy as np
import pytest

from atg_testlib import (DIGEST_KEYS, GOLDEN, NAMED_CASES, RECONCILE_CASES, STAGE_DIGEST_CASES, batch64_frames,
                         default_params, digest, frame_hash, named_case)

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def fx():
    return np.load(GOLDEN / "synth_expected.npz")


@pytest.mark.parametrize("tag", list(NAMED_CASES))
def test_named_cases_match_reference(atg, ctx, fx, tag):
    img, p, fams = named_case(atg, tag)
    assert frame_hash(img) == int(fx[tag + "_hash"])
    print("case", tag, img.shape, flush=True)
    ctx.set_families(fams)
    g = ctx.detect(img, atg.params_from_dict(p))
    assert g["nquads"] == int(fx[tag + "_nquads"])
    assert g["dets"].tobytes() == fx[tag + "_dets"].tobytes()
    if tag in RECONCILE_CASES:
        assert int(fx[tag + "_nraw"]) > len(g["dets"])


@pytest.mark.parametrize("tag", STAGE_DIGEST_CASES)
def test_stage_digests_match_reference(atg, ctx, fx, tag):
    img, p, fams = named_case(atg, tag)
    print("case", tag, img.shape, flush=True)
    ctx.set_families(fams)
    g = ctx.detect(img, atg.params_from_dict(p))
    st = ctx.stages(0)
    for k in DIGEST_KEYS:
        assert digest(st[k]) == int(fx[tag + "_sd_" + k]), k
    assert g["dets"].tobytes() == fx[tag + "_dets"].tobytes()


def test_batch64_matches_reference(atg, ctx, fx):
    """One call over the 64 uneven frames; each frame equals the reference."""
    frames = batch64_frames(atg)
    ctx.set_families([("tag36h11", 2)])
    print("case batch64: 16 each of 640x480 1280x720 1920x1080 2448x2048", flush=True)
    out = ctx.detect_batch(frames, atg.params_from_dict(default_params()))
    for i, (f, g) in enumerate(zip(frames, out)):
        assert frame_hash(f) == int(fx["b64_hash"][i])
        assert len(g["dets"]) == int(fx["b64_ndets"][i]), i
        assert g["nquads"] == int(fx["b64_nquads"][i]), i
        assert digest(g["dets"]) == int(fx["b64_dets"][i]), i


def test_device_finish_matches_reference_lists(ctx):
    """The device's reconcile then stable id sort: the reference's kept
    detections, stably sorted by id (fast/detector.c:164-179)."""
    z = np.load(GOLDEN / "reconcile_cases.npz")
    d, off, keep = z["dets"], z["offsets"], z["keep"]
    for i in range(len(off) - 1):
        lst, k = d[off[i]:off[i + 1]], keep[off[i]:off[i + 1]]
        want = lst[k]
        want = want[np.argsort(want["id"], kind="stable")]
        got = ctx.finish(lst)
        assert got.tobytes() == want.tobytes(), i
