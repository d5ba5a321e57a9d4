"""CPU checks of the product library: it loads, exports every entry point
the headers declare, its built-in families equal the packed tables, and
the synthetic generator is deterministic. No GPU call here."""
import ctypes as C
import re
import subprocess

import numpy as np
import pytest

from atg_testlib import PKG, REPO, frame_hash, load_atg, load_families

HEADERS = [REPO / "include" / "apriltag_fast.h", REPO / "include" / "atg.h", REPO / "include" / "apriltag.h",
           REPO / "include" / "common" / "matd.h", REPO / "include" / "common" / "image_u8.h",
           REPO / "include" / "apriltag_pose.h", REPO / "include" / "common" / "homography.h"] + \
    sorted((REPO / "include").glob("tag*.h"))

# sizeof of the upstream API's public structs in the reference (its
# apriltag.h and common/ headers, x86-64 gcc): code built against them
# must see the same layouts here.
UPSTREAM_SIZES = {"apriltag_detector_t": 160, "apriltag_detection_t": 112, "apriltag_family_t": 72,
                  "image_u8_t": 24, "zarray_t": 24, "matd_t": 16, "apriltag_detection_info_t": 48,
                  "apriltag_pose_t": 16}


def declared_functions():
    names = set()
    for h in HEADERS:
        text = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        for m in re.finditer(r"^[A-Za-z_][\w \*]*?\b(\w+)\s*\([^;{]*\)\s*;", text, re.M):
            if not m.group(0).startswith(("typedef", "static")):
                names.add(m.group(1))
    return sorted(names)


def test_library_exports_every_declared_function():
    lib = PKG / "libapriltag_amd.so"
    assert lib.exists(), "build the library first"
    out = subprocess.run(["nm", "-D", "--defined-only", str(lib)], capture_output=True, text=True).stdout
    exported = {l.split()[-1] for l in out.splitlines() if l.strip()}
    names = declared_functions()
    assert len(names) >= 25
    missing = [n for n in names if n not in exported]
    assert not missing, missing
    C.CDLL(str(lib))


def test_builtin_families_match_packed_tables():
    atg = load_atg()
    L = atg.lib()
    fams = load_families()
    for name, idx in atg.FAMILY_INDEX.items():
        f = L.atg_builtin_family(idx).contents
        ref = fams[name]
        assert f.name.decode() == name
        assert (f.ncodes, f.nbits, f.width_at_border, f.total_width, f.reversed_border) == \
            (ref["ncodes"], ref["nbits"], ref["wab"], ref["tw"], ref["reversed"])
        codes = np.ctypeslib.as_array(f.codes, (f.ncodes,))
        assert np.array_equal(codes, ref["codes"])
        assert list(f.bit_x[:f.nbits]) == list(ref["bit_x"])
        by_name = L.at_family_by_name(name.encode()) if hasattr(L, "at_family_by_name") else None
        assert by_name is not None


def test_family_table_facts():
    """Facts of the families that the configs use (tag36h11.c:624-704,
    tagStandard52h13.c:48751-48863)."""
    f = load_families()
    t = f["tag36h11"]
    assert (t["ncodes"], t["nbits"], t["h"], t["wab"], t["tw"], t["reversed"]) == (587, 36, 11, 8, 10, False)
    s = f["tagStandard52h13"]
    assert (s["ncodes"], s["nbits"], s["h"], s["wab"], s["tw"], s["reversed"]) == (48714, 52, 13, 6, 10, True)
    # a family's codes are pairwise at least h bits apart
    codes = t["codes"]
    x = np.bitwise_xor(codes[:, None], codes[None, :])
    pc = np.zeros(x.shape, np.int64)
    for b in range(36):
        pc += ((x >> np.uint64(b)) & np.uint64(1)).astype(np.int64)
    np.fill_diagonal(pc, 99)
    assert pc.min() >= 11


def test_default_params_match_reference():
    atg = load_atg()
    p = atg.default_params()
    assert p.quad_decimate == 2.0 and p.quad_sigma == 0.0 and p.refine_edges
    assert p.decode_sharpening == 0.25 and p.min_cluster_pixels == 24 and p.max_nmaxima == 10
    assert p.max_line_fit_mse == 10.0 and p.min_white_black_diff == 5 and not p.deglitch
    assert p.cos_critical_rad == np.float32(np.cos(10.0 * np.pi / 180.0))


def test_synth_is_deterministic():
    atg = load_atg()
    a, ta = atg.synth_frame(1, 11)
    b, tb = atg.synth_frame(1, 11)
    c, _ = atg.synth_frame(1, 12)
    assert a.tobytes() == b.tobytes() and ta == tb
    assert a.tobytes() != c.tobytes()
    assert len(ta) == 4 and all(f == 3 for f, _ in ta)


def test_pnm_roundtrip(tmp_path):
    atg = load_atg()
    L = atg.lib()
    L.at_write_pgm.argtypes = [C.c_char_p, C.c_void_p, C.c_int64, C.c_int64]
    L.at_write_pgm.restype = C.c_bool
    L.at_read_pnm.argtypes = [C.c_char_p, C.POINTER(C.c_int64), C.POINTER(C.c_int64), C.POINTER(C.c_void_p)]
    L.at_read_pnm.restype = C.c_bool
    img = (np.arange(37 * 23) % 251).astype(np.uint8).reshape(23, 37)
    path = str(tmp_path / "x.pgm").encode()
    assert L.at_write_pgm(path, img.ctypes.data, 37, 23)
    w, h, px = C.c_int64(), C.c_int64(), C.c_void_p()
    assert L.at_read_pnm(path, C.byref(w), C.byref(h), C.byref(px))
    got = np.frombuffer(C.string_at(px, w.value * h.value), np.uint8).reshape(h.value, w.value)
    assert np.array_equal(got, img)
    C.CDLL(None).free(px)


def build_upstream_demo(out):
    """tests/c/upstream_demo.c, built as a user's C program against
    include/ and the library."""
    lib_dir = PKG
    cmd = ["gcc", "-O1", "-Wall", "-Werror", "-I", str(REPO / "include"), str(REPO / "tests" / "c" / "upstream_demo.c"),
           "-L", str(lib_dir), "-lapriltag_amd", "-Wl,-rpath," + str(lib_dir), "-o", str(out)]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    return out


def test_upstream_api_builds_with_reference_layouts(tmp_path):
    exe = build_upstream_demo(tmp_path / "upstream_demo")
    out = subprocess.run([str(exe), "-s"], check=True, capture_output=True, text=True).stdout
    sizes = {l.split()[0]: int(l.split()[1]) for l in out.splitlines()}
    assert sizes == UPSTREAM_SIZES


class _ImageU8(C.Structure):
    _fields_ = [("width", C.c_int32), ("height", C.c_int32), ("stride", C.c_int32), ("buf", C.POINTER(C.c_uint8))]


def test_apriltag_to_image_round_trips_through_the_detector():
    """apriltag_to_image (upstream API, host code): the tag image of code
    idx of every family, scaled up and set on a white page, decodes to idx
    with hamming 0 in the oracle (which the reference's goldens pin)."""
    from atg_testlib import FAMILY_NAMES, Oracle, default_params
    atg = load_atg()
    L = atg.lib()
    L.apriltag_to_image.restype = C.POINTER(_ImageU8)
    L.apriltag_to_image.argtypes = [C.c_void_p, C.c_uint32]
    L.image_u8_destroy.argtypes = [C.POINTER(_ImageU8)]
    orc = Oracle()
    for name in FAMILY_NAMES:
        create = getattr(L, name + "_create")
        create.restype = C.c_void_p
        destroy = getattr(L, name + "_destroy")
        destroy.argtypes = [C.c_void_p]
        fam = create()
        try:
            for idx in (0, 3, 11):
                ptr = L.apriltag_to_image(fam, idx)
                im = ptr.contents
                h, w = im.height, im.width
                cells = np.ctypeslib.as_array(im.buf, (h, im.stride))[:, :w].copy()
                L.image_u8_destroy(ptr)
                page = np.full((h * 12 + 80, w * 12 + 80), 255, np.uint8)
                page[40:40 + h * 12, 40:40 + w * 12] = np.kron(cells, np.ones((12, 12), np.uint8))
                d = orc.detect(page, default_params(quad_decimate=1.0), [(name, 0)])["dets"]
                assert [(int(x["id"]), int(x["hamming"])) for x in d] == [(idx, 0)], (name, idx)
        finally:
            destroy(fam)
