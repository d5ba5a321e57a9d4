"""The allocation-free entry point (atg_detect_noalloc) through the C-ABI:
its detections and quad count equal atg_detect's, bit for bit, on two
calls in a row (the second on a smaller frame), from host and from device
memory; a frame beyond its static scratch is refused with an error."""
import ctypes as C

import numpy as np
import pytest

from atg_testlib import DET_DTYPE, default_params, same_dets, synth_case

pytestmark = pytest.mark.gpu

ATG_ERR_LIMIT = -3


def noalloc(atg, ctx, prm, img=None, dev=None, cap=4096):
    L = atg.lib()
    L.atg_detect_noalloc.restype = C.c_int
    L.atg_detect_noalloc.argtypes = [C.c_void_p, C.POINTER(atg.Params), C.c_void_p, C.c_int64, C.c_int64, C.c_int64,
                                     C.c_int, C.c_void_p, C.c_int64, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
    out = np.zeros(max(cap, 1), DET_DTYPE)
    nd, nq = C.c_int64(-1), C.c_int64(-1)
    if dev is not None:
        ptr, h, w, stride = dev
        on_device = 1
    else:
        ptr, h, w, stride = img.ctypes.data, img.shape[0], img.shape[1], img.strides[0]
        on_device = 0
    rc = L.atg_detect_noalloc(ctx.h, C.byref(prm), ptr, w, h, stride, on_device, out.ctypes.data, cap,
                              C.byref(nd), C.byref(nq))
    return rc, out[:min(max(nd.value, 0), cap)].copy(), nd.value, nq.value


def test_noalloc_equals_detect_two_calls(atg, ctx):
    ctx.set_families([("tag36h11", 2)])
    prm = atg.params_from_dict(default_params())
    frames = [synth_case(atg, 2, 0)[0], synth_case(atg, 1, 0)[0]]  # 1920x1080, then 640x480
    for on_device in (False, True):
        for img in frames:
            print("case noalloc", img.shape, "device" if on_device else "host", flush=True)
            want = ctx.detect(img, prm)
            dev = ctx.to_device(img) if on_device else None
            try:
                rc, dets, nd, nq = noalloc(atg, ctx, prm, img=img, dev=dev)
            finally:
                if dev:
                    ctx.free_device(dev[0])
            assert rc == 0, atg.lib().atg_last_error(ctx.h)
            assert nq == want["nquads"] and nd == len(want["dets"]) > 0
            assert same_dets(dets, want["dets"])


def test_noalloc_partial_output_and_limits(atg, ctx):
    ctx.set_families([("tag36h11", 2)])
    prm = atg.params_from_dict(default_params())
    img = synth_case(atg, 2, 1)[0]
    want = ctx.detect(img, prm)
    rc, dets, nd, nq = noalloc(atg, ctx, prm, img=img, cap=3)
    assert rc == 0 and nd == len(want["dets"]) and len(dets) == 3
    assert same_dets(dets, want["dets"][:3])
    # a 3840x2160 frame at decimate 1 needs far more than the scratch: refused, not run
    big = np.zeros((2160, 3840), np.uint8)
    rc, _, nd, _ = noalloc(atg, ctx, atg.params_from_dict(default_params(quad_decimate=1.0)), img=big)
    assert rc == ATG_ERR_LIMIT and nd == 0
    assert b"scratch" in atg.lib().atg_last_error(ctx.h)
    # the scratch keeps working after the refusal
    rc, dets, nd, nq = noalloc(atg, ctx, prm, img=img)
    assert rc == 0 and same_dets(dets, want["dets"]) and nq == want["nquads"]
