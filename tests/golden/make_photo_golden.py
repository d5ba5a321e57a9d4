"""Write the photo fixtures: the three photos of the reference's own tests
(test/data/*.jpg, NASA Swarmathon, CC BY-SA 2.0, see test/data/README.md
there) decoded by the reference's JPEG decoder (oracle/_ref/libjpegref.so:
pjpeg_to_u8_baseline, the pixels its tests detect in), the true corners
its tests hold (test/data/*.txt), and the reference build's detections
and poses on them. Run where `make -C oracle ref` built oracle/_ref.

    python tests/golden/make_photo_golden.py

tests/golden/photos/<name>.pgm   the decoded gray photo
tests/golden/photos/<name>.txt   the reference test's true detections
tests/golden/photo_expected.npz  per photo <name>: <name>_dets (the
    reference's detections at test/test_detection.c's settings: decimate
    1, no edge refinement, tag36h11 with 2 bits), <name>_nquads, and
    the pose of each detection by estimate_tag_pose at
    test/test_tag_pose_estimation.c's camera (tag 0.16 m, f 600 px,
    centre at the image centre): <name>_err, <name>_R, <name>_t.
"""
import ctypes as C
import shutil
import sys
from pathlib import Path

import numpy as np

sys.path.insert(0, str(Path(__file__).resolve().parent.parent))
from atg_testlib import GOLDEN, POSE_IN_DTYPE, REPO, Reference, RefPose, default_params  # noqa: E402

REF_DATA = Path("/root/reference/test/data")
PHOTOS = ["33369213973_9d9bb4cc96_c", "34085369442_304b6bafd9_c", "34139872896_defdb2f8d9_c"]
PHOTO_PARAMS = dict(quad_decimate=1.0, refine_edges=False)
PHOTO_FAMILIES = [("tag36h11", 2)]
PHOTO_CAMERA = dict(tagsize=0.16, f=600.0)


def decode(path):
    L = C.CDLL(str(REPO / "oracle" / "_ref" / "libjpegref.so"))
    L.ref_jpeg_gray.argtypes = [C.c_char_p, C.c_void_p, C.c_int64, C.POINTER(C.c_int), C.POINTER(C.c_int)]
    w, h = C.c_int(), C.c_int()
    assert L.ref_jpeg_gray(str(path).encode(), None, 0, C.byref(w), C.byref(h)) == 0
    img = np.zeros((h.value, w.value), np.uint8)
    assert L.ref_jpeg_gray(str(path).encode(), img.ctypes.data, img.size, C.byref(w), C.byref(h)) == 0
    return img


def write_pgm(path, img):
    with open(path, "wb") as f:
        f.write(b"P5\n%d %d\n255\n" % (img.shape[1], img.shape[0]))
        f.write(img.tobytes())


def main():
    out_dir = GOLDEN / "photos"
    out_dir.mkdir(exist_ok=True)
    ref, rp = Reference(), RefPose()
    out = {}
    for name in PHOTOS:
        img = decode(REF_DATA / (name + ".jpg"))
        write_pgm(out_dir / (name + ".pgm"), img)
        shutil.copyfile(REF_DATA / (name + ".txt"), out_dir / (name + ".txt"))
        r = ref.detect(img, default_params(**PHOTO_PARAMS), PHOTO_FAMILIES)
        out[name + "_dets"] = r["dets"]
        out[name + "_nquads"] = np.array(r["nquads"])
        H, W = img.shape
        errs, Rs, ts = [], [], []
        for d in r["dets"]:
            tag = np.zeros(1, POSE_IN_DTYPE)[0]
            tag["H"], tag["p"] = d["H"], d["p"]
            tag["tagsize"], tag["fx"], tag["fy"] = PHOTO_CAMERA["tagsize"], PHOTO_CAMERA["f"], PHOTO_CAMERA["f"]
            tag["cx"], tag["cy"] = W / 2.0, H / 2.0
            e, R, t = rp.estimate_tag_pose(tag)
            errs.append(e)
            Rs.append(R)
            ts.append(t)
        out[name + "_err"], out[name + "_R"], out[name + "_t"] = np.array(errs), np.array(Rs), np.array(ts)
        print(name, img.shape, "detections", len(r["dets"]), "nquads", r["nquads"])
    np.savez_compressed(GOLDEN / "photo_expected.npz", **out)


if __name__ == "__main__":
    main()
