"""Tag pose (the reference's apriltag_pose.c and homography_to_pose) on the
GPU against the reference's own pose code.

The fixtures (tests/golden/pose_cases.npz, written by make_pose_golden.py
with oracle/_ref/libposeref.so, the reference's apriltag_pose.c,
homography.c, matd.c and svd22.c compiled unchanged) hold tags seen at
known poses (exact and noisy corners, near and far) and the detections
of synthetic frames, with the reference's poses.

Tolerance (floating point whose operation order is not the reference's:
its SVD polar factors and LU inverses are formed by other methods here):
rotations and homography poses within 1e-9 absolute (entries of unit
size), translations within 1e-9 of their length, errors within 1e-9
relative plus 1e-18 m^2 (the exact-corner errors are rounding noise of
about 1e-28), and the same second-minimum and choice flags.
"""
import ctypes as C

import numpy as np
import pytest

from atg_testlib import GOLDEN, POSE_IN_DTYPE, POSE_OUT_DTYPE, RefPose, POSE_REF_SO, rotation, tag_seen

CASES = np.load(GOLDEN / "pose_cases.npz")
GROUPS = ["exact", "noisy", "far", "dets", "iters0", "iters1", "iters5", "iters100"]
TOL = 1e-9


def close_poses(got, want):
    """Problems of got against want (POSE_OUT_DTYPE arrays), [] if none."""
    bad = []
    if not np.array_equal(got["has2"], want["has2"]):
        bad.append("has2 %s" % np.nonzero(got["has2"] != want["has2"])[0][:10])
    if not np.array_equal(got["best"], want["best"]):
        bad.append("best %s" % np.nonzero(got["best"] != want["best"])[0][:10])
    two = (want["has2"] == 1) & (got["has2"] == 1)
    for f in ("hpose", "R0", "R1", "R2"):
        m = two if f == "R2" else np.ones(len(want), bool)
        d = np.abs(got[f][m] - want[f][m])
        if f == "hpose":  # its translation column is in the homography's units
            d = d / np.maximum(1.0, np.abs(want[f][m]))
        if d.size and d.max() > TOL:
            bad.append("%s off by %.3g" % (f, d.max()))
    for f in ("t0", "t1", "t2"):
        m = two if f == "t2" else np.ones(len(want), bool)
        n = np.maximum(np.linalg.norm(want[f][m], axis=1), 1e-12)
        d = np.abs(got[f][m] - want[f][m]).max(axis=1) / n if m.any() else np.zeros(0)
        if d.size and d.max() > TOL:
            bad.append("%s off by %.3g" % (f, d.max()))
    for f, m in (("err1", np.ones(len(want), bool)), ("err2", two)):
        a, b = got[f][m], want[f][m]
        inf = np.isinf(b)
        if not np.array_equal(np.isinf(a), inf):
            bad.append("%s infinities differ" % f)
        elif (np.abs(a[~inf] - b[~inf]) > TOL * np.abs(b[~inf]) + 1e-18).any():
            bad.append("%s off" % f)
    return bad


def rotation_ok(R):
    R = R.reshape(-1, 3, 3)
    return np.abs(R @ R.transpose(0, 2, 1) - np.eye(3)).max() < 1e-12 and np.abs(np.linalg.det(R) - 1).max() < 1e-12


# ------------------------------------------------------------------ CPU

@pytest.mark.skipif(not POSE_REF_SO.exists(), reason="the reference pose build (make -C oracle ref) is absent")
def test_fixture_is_the_reference_output():
    """Every group of the fixture is what the reference build gives now."""
    ref = RefPose()
    for g in GROUPS:
        again = ref.poses(CASES[g + "_in"], int(CASES[g + "_iters"]))
        want = CASES[g + "_out"]
        for f in POSE_OUT_DTYPE.names:
            assert np.array_equal(again[f], want[f]), (g, f)


def test_fixture_recovers_the_true_pose():
    """The reference's pose of a tag seen exactly is the true pose, its
    second minimum a different rotation, and every rotation proper."""
    out, R, t = CASES["exact_out"], CASES["exact_R"], CASES["exact_t"]
    assert (out["best"] == 1).all()
    assert np.abs(out["R1"].reshape(-1, 3, 3) - R).max() < 1e-9
    assert (np.abs(out["t1"] - t).max(axis=1) / np.linalg.norm(t, axis=1)).max() < 1e-9
    assert (out["err1"] < 1e-20).all()
    two = out["has2"] == 1
    assert two.sum() > 50 and (out["err2"][two] > out["err1"][two]).all()
    for g in GROUPS:
        o = CASES[g + "_out"]
        assert rotation_ok(o["R0"]) and rotation_ok(o["R1"]), g
        if (o["has2"] == 1).any():
            assert rotation_ok(o["R2"][o["has2"] == 1]), g


def test_fixture_noisy_poses_near_truth():
    """With corner noise the chosen pose's translation stays near the truth
    (its rotation may take the other minimum: far tags are ambiguous), and
    its error is the smaller one (estimate_tag_pose's choice)."""
    for g, lim in (("noisy", 0.06), ("far", 0.4)):
        out, t = CASES[g + "_out"], CASES[g + "_t"]
        tb = np.where((out["best"] == 1)[:, None], out["t1"], out["t2"])
        assert (np.linalg.norm(tb - t, axis=1) / np.linalg.norm(t, axis=1)).max() < lim, g
        eb = np.where(out["best"] == 1, out["err1"], out["err2"])
        assert (eb <= np.minimum(out["err1"], out["err2"])).all()


def test_matd_algebra(atg):
    """The matd_t helpers of the upstream API (host code) against numpy."""
    L = atg.lib()

    class Matd(C.Structure):
        _fields_ = [("nrows", C.c_uint), ("ncols", C.c_uint), ("data", C.POINTER(C.c_double))]

    P = C.POINTER(Matd)
    for name, res, args in (("matd_create_data", P, [C.c_int, C.c_int, C.c_void_p]), ("matd_multiply", P, [P, P]),
                            ("matd_inverse", P, [P]), ("matd_transpose", P, [P]), ("matd_det", C.c_double, [P]),
                            ("matd_crossproduct", P, [P, P]), ("matd_vec_normalize", P, [P]),
                            ("matd_destroy", None, [P])):
        getattr(L, name).restype = res
        getattr(L, name).argtypes = args
    rng = np.random.default_rng(3)

    def mk(a):
        a = np.ascontiguousarray(a, np.float64)
        return L.matd_create_data(a.shape[0], a.shape[1], a.ctypes.data)

    def val(m):
        return np.ctypeslib.as_array(m.contents.data, (m.contents.nrows, m.contents.ncols)).copy()

    for n in (1, 2, 3, 4, 6):
        a, b = rng.normal(size=(n, n)), rng.normal(size=(n, n))
        ma, mb = mk(a), mk(b)
        prod, inv, tr = L.matd_multiply(ma, mb), L.matd_inverse(ma), L.matd_transpose(ma)
        assert np.allclose(val(prod), a @ b, rtol=1e-13, atol=1e-13)
        assert np.allclose(val(tr), a.T)
        assert np.allclose(L.matd_det(ma), np.linalg.det(a), rtol=1e-12)
        if n > 1:
            assert np.allclose(val(inv), np.linalg.inv(a), rtol=1e-10, atol=1e-10)
        for m in (ma, mb, prod, inv, tr):
            L.matd_destroy(m)
    sing = mk(np.array([[1.0, 2.0], [2.0, 4.0]]))
    assert not L.matd_inverse(sing)  # NULL: exactly singular
    u, v = mk(rng.normal(size=(3, 1))), mk(rng.normal(size=(3, 1)))
    cr, nu = L.matd_crossproduct(u, v), L.matd_vec_normalize(u)
    assert np.allclose(val(cr).ravel(), np.cross(val(u).ravel(), val(v).ravel()))
    assert np.isclose(np.linalg.norm(val(nu)), 1.0)
    for m in (sing, u, v, cr, nu):
        L.matd_destroy(m)


# ------------------------------------------------------------------ GPU

@pytest.mark.gpu
@pytest.mark.parametrize("group", GROUPS)
def test_gpu_poses_match_reference(ctx, group):
    tags = CASES[group + "_in"]
    got = ctx.estimate_poses(tags, int(CASES[group + "_iters"]))
    assert close_poses(got, CASES[group + "_out"]) == [], group


@pytest.mark.gpu
def test_gpu_poses_batch_and_flags(ctx):
    """Many tags in one launch give each tag's own pose; the homography-only
    flag gives the same homography pose; no tags is a no-op."""
    tags = np.concatenate([CASES[g + "_in"] for g in ("exact", "noisy", "far", "dets")])
    big = np.tile(tags, 64)
    got = ctx.estimate_poses(big, 50)
    one = ctx.estimate_poses(tags, 50)
    for k in range(64):
        part = got[k * len(tags):(k + 1) * len(tags)]
        for f in POSE_OUT_DTYPE.names:
            assert np.array_equal(part[f], one[f]), f
    h = ctx.estimate_poses(tags, 50, flags=1)
    assert np.array_equal(h["hpose"], one["hpose"]) and np.array_equal(h["R0"], one["R0"])
    assert (h["has2"] == 0).all() and np.isinf(h["err1"]).all()
    assert len(ctx.estimate_poses(np.zeros(0, POSE_IN_DTYPE))) == 0


@pytest.mark.gpu
def test_gpu_exact_pose_is_the_truth(ctx):
    """Independently of the reference: tags rendered at known poses (the
    camera model of tag_seen) come back at those poses."""
    rng = np.random.default_rng(11)
    tags = np.zeros(200, POSE_IN_DTYPE)
    Rs, ts = [], []
    for i in range(len(tags)):
        R = rotation(rng.normal(size=3), rng.uniform(0, 1.0))
        t = np.array([rng.uniform(-0.5, 0.5), rng.uniform(-0.3, 0.3), rng.uniform(0.5, 4)])
        H, p = tag_seen(R, t, 0.15, 1000.0, 1000.0, 640.0, 360.0)
        tags[i] = (H, p, 0.15, 1000.0, 1000.0, 640.0, 360.0)
        Rs.append(R)
        ts.append(t)
    got = ctx.estimate_poses(tags, 50)
    assert (got["best"] == 1).all()
    assert np.abs(got["R1"].reshape(-1, 3, 3) - np.array(Rs)).max() < 1e-9
    assert np.abs(got["t1"] - np.array(ts)).max() < 1e-9
    assert rotation_ok(got["R0"]) and rotation_ok(got["R1"])


class _Matd(C.Structure):
    _fields_ = [("nrows", C.c_uint), ("ncols", C.c_uint), ("data", C.POINTER(C.c_double))]


class _Det(C.Structure):  # apriltag_detection_t (apriltag.h:215-250)
    _fields_ = [("family", C.c_void_p), ("id", C.c_int), ("hamming", C.c_int), ("decision_margin", C.c_float),
                ("H", C.POINTER(_Matd)), ("c", C.c_double * 2), ("p", (C.c_double * 2) * 4)]


class _Info(C.Structure):  # apriltag_detection_info_t (apriltag_pose.h:11-18)
    _fields_ = [("det", C.POINTER(_Det)), ("tagsize", C.c_double), ("fx", C.c_double), ("fy", C.c_double),
                ("cx", C.c_double), ("cy", C.c_double)]


class _Pose(C.Structure):  # apriltag_pose_t (apriltag_pose.h:25-28)
    _fields_ = [("R", C.POINTER(_Matd)), ("t", C.POINTER(_Matd))]


@pytest.mark.gpu
def test_gpu_upstream_pose_api(atg):
    """estimate_tag_pose, estimate_tag_pose_orthogonal_iteration,
    estimate_pose_for_tag_homography and homography_to_pose through the C
    ABI (apriltag_pose.h) against the reference's own calls."""
    L = atg.lib()
    PI, PP, PM = C.POINTER(_Info), C.POINTER(_Pose), C.POINTER(_Matd)
    L.estimate_tag_pose.restype = C.c_double
    L.estimate_tag_pose.argtypes = [PI, PP]
    L.estimate_tag_pose_orthogonal_iteration.argtypes = [PI, C.POINTER(C.c_double), PP, C.POINTER(C.c_double), PP,
                                                         C.c_int]
    L.estimate_pose_for_tag_homography.argtypes = [PI, PP]
    L.homography_to_pose.restype = PM
    L.homography_to_pose.argtypes = [PM, C.c_double, C.c_double, C.c_double, C.c_double]
    L.matd_create_data.restype = PM
    L.matd_create_data.argtypes = [C.c_int, C.c_int, C.c_void_p]
    L.matd_destroy.argtypes = [PM]

    def arr(m):
        return np.ctypeslib.as_array(m.contents.data, (m.contents.nrows * m.contents.ncols,)).copy()

    tags, want = CASES["dets_in"], CASES["dets_out"]
    for i in range(len(tags)):
        tg = tags[i]
        H = np.ascontiguousarray(tg["H"])
        det = _Det()
        det.H = L.matd_create_data(3, 3, H.ctypes.data)
        for k in range(4):
            det.p[k][0], det.p[k][1] = tg["p"][2 * k], tg["p"][2 * k + 1]
        info = _Info(C.pointer(det), tg["tagsize"], tg["fx"], tg["fy"], tg["cx"], tg["cy"])
        pose = _Pose()
        e = L.estimate_tag_pose(C.byref(info), C.byref(pose))
        assert abs(e - CASES["est_err"][i]) <= TOL * abs(CASES["est_err"][i]) + 1e-18
        assert np.abs(arr(pose.R) - CASES["est_R"][i].ravel()).max() < TOL
        assert np.abs(arr(pose.t) - CASES["est_t"][i]).max() < TOL * np.linalg.norm(CASES["est_t"][i])
        L.matd_destroy(pose.R)
        L.matd_destroy(pose.t)
        p1, p2 = _Pose(), _Pose()
        e1, e2 = C.c_double(), C.c_double()
        L.estimate_tag_pose_orthogonal_iteration(C.byref(info), C.byref(e1), C.byref(p1), C.byref(e2), C.byref(p2),
                                                 50)
        assert abs(e1.value - want["err1"][i]) <= TOL * abs(want["err1"][i]) + 1e-18
        assert np.abs(arr(p1.R) - want["R1"][i]).max() < TOL
        assert bool(p2.R) == bool(want["has2"][i])
        if p2.R:
            assert np.abs(arr(p2.R) - want["R2"][i]).max() < TOL
            L.matd_destroy(p2.R)
            L.matd_destroy(p2.t)
        else:
            assert np.isinf(e2.value) and not p2.t
        L.matd_destroy(p1.R)
        L.matd_destroy(p1.t)
        p0 = _Pose()
        L.estimate_pose_for_tag_homography(C.byref(info), C.byref(p0))
        assert np.abs(arr(p0.R) - want["R0"][i]).max() < TOL
        assert np.abs(arr(p0.t) - want["t0"][i]).max() < TOL * np.linalg.norm(want["t0"][i])
        L.matd_destroy(p0.R)
        L.matd_destroy(p0.t)
        L.matd_destroy(det.H)
    for H, cam, w in zip(CASES["h2p_H"], CASES["h2p_cam"], CASES["h2p_out"]):
        Hm = L.matd_create_data(3, 3, np.ascontiguousarray(H).ctypes.data)
        m = L.homography_to_pose(Hm, *cam)
        got = arr(m).reshape(4, 4)
        assert np.abs(got - w).max() / max(1.0, np.abs(w).max()) < TOL
        L.matd_destroy(m)
        L.matd_destroy(Hm)
