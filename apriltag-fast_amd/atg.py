"""Python binding of the MI355X AprilTag detector (ctypes over the C-ABI in
include/atg.h and include/apriltag_fast.h).

Loads the in-tree libapriltag_amd.so and raises if it is missing: there is
no CPU fallback.

    ctx = Context()
    ctx.set_families([("tag36h11", 2)])
    res = ctx.detect(img)                 # img: uint8 HxW numpy array
    res["dets"], res["nquads"]
    outs = ctx.detect_batch([img0, img1]) # frames of any sizes, one pass
"""
import ctypes as C
from pathlib import Path

import numpy as np

PKG = Path(__file__).resolve().parent
LIB_PATH = PKG / "libapriltag_amd.so"

FAMILY_INDEX = {
    "tag16h5": 0, "tag25h9": 1, "tag36h10": 2, "tag36h11": 3, "tagCircle21h7": 4,
    "tagCircle49h12": 5, "tagCustom48h12": 6, "tagStandard41h12": 7, "tagStandard52h13": 8,
}


class Family(C.Structure):
    _fields_ = [("name", C.c_char * 32), ("codes", C.POINTER(C.c_uint64)), ("ncodes", C.c_int64),
                ("nbits", C.c_int32), ("h", C.c_int32), ("width_at_border", C.c_int32),
                ("total_width", C.c_int32), ("reversed_border", C.c_bool),
                ("bit_x", C.c_int32 * 64), ("bit_y", C.c_int32 * 64)]


class Params(C.Structure):
    _fields_ = [("quad_decimate", C.c_float), ("quad_sigma", C.c_float), ("refine_edges", C.c_bool),
                ("decode_sharpening", C.c_double), ("min_cluster_pixels", C.c_int32),
                ("max_nmaxima", C.c_int32), ("cos_critical_rad", C.c_float),
                ("max_line_fit_mse", C.c_float), ("min_white_black_diff", C.c_int32),
                ("deglitch", C.c_bool)]


class Frame(C.Structure):
    _fields_ = [("pixels", C.c_void_p), ("width", C.c_int64), ("height", C.c_int64), ("stride", C.c_int64)]


class Result(C.Structure):
    _fields_ = [("dets", C.c_void_p), ("ndets", C.c_int64), ("nquads", C.c_int64)]


class Stages(C.Structure):
    _fields_ = [("qw", C.c_int64), ("qh", C.c_int64), ("qim", C.POINTER(C.c_uint8)),
                ("t", C.POINTER(C.c_uint8)), ("nruns", C.c_int64), ("runs", C.POINTER(C.c_uint32)),
                ("row_off", C.POINTER(C.c_uint32)), ("nnodes", C.c_int64),
                ("rep", C.POINTER(C.c_uint32)), ("csize", C.POINTER(C.c_uint32)),
                ("ok", C.POINTER(C.c_uint8)), ("nclusters", C.c_int64),
                ("cluster_ids", C.POINTER(C.c_uint64)), ("ccount", C.POINTER(C.c_uint32)),
                ("npts", C.c_int64), ("pts", C.POINTER(C.c_uint64)), ("nquads", C.c_int64),
                ("quads", C.POINTER(C.c_float))]


class SynthSpec(C.Structure):
    _fields_ = [("width", C.c_int32), ("height", C.c_int32), ("ntags_min", C.c_int32),
                ("ntags_max", C.c_int32), ("side_min", C.c_int32), ("side_max", C.c_int32),
                ("max_tilt", C.c_double), ("family_a", C.c_int32), ("family_b", C.c_int32),
                ("background", C.c_int32), ("bg_amp", C.c_int32), ("blur_min", C.c_double),
                ("blur_max", C.c_double), ("noise_min", C.c_double), ("noise_max", C.c_double),
                ("nbars", C.c_int32), ("nclutter", C.c_int32), ("ncheckers", C.c_int32)]


class SynthTag(C.Structure):
    _fields_ = [("family", C.c_int32), ("id", C.c_int32), ("corners", C.c_double * 8)]


class Tune(C.Structure):
    _fields_ = [("thr_tile_rows", C.c_int32), ("row_waves", C.c_int32), ("front_threads", C.c_int32),
                ("fit_blocks", C.c_int32), ("front_bands", C.c_int32)]

TUNE_FIELDS = [f for f, _ in Tune._fields_]


class ServerConfig(C.Structure):
    _fields_ = [("device", C.c_int32), ("budget_bytes", C.c_size_t), ("lanes", C.c_int32),
                ("max_batch", C.c_int32), ("slots_per_stream", C.c_int32)]


class Served(C.Structure):
    _fields_ = [("stream", C.c_int32), ("status", C.c_int32), ("tag", C.c_uint64), ("arrival_ns", C.c_int64),
                ("admit_ns", C.c_int64), ("done_ns", C.c_int64), ("nquads", C.c_int64), ("ndets", C.c_int64),
                ("dets", C.c_void_p), ("batch_frames", C.c_int32)]


class Arrival(C.Structure):
    _fields_ = [("stream", C.c_int32), ("at_ns", C.c_int64), ("pixels", C.c_void_p), ("stride", C.c_int64),
                ("tag", C.c_uint64)]


class ServerCounts(C.Structure):
    _fields_ = [("submitted", C.c_int64), ("dropped", C.c_int64), ("completed", C.c_int64),
                ("batches", C.c_int64), ("max_batch_seen", C.c_int64), ("lanes", C.c_int32),
                ("lane_arena_bytes", C.c_size_t), ("slot_bytes", C.c_size_t)]


# at_detection (include/apriltag_fast.h), 168 bytes
DET_DTYPE = np.dtype([("family", "<i4"), ("id", "<i4"), ("hamming", "<i4"), ("margin", "<f4"),
                      ("H", "<f8", 9), ("c", "<f8", 2), ("p", "<f8", 8)])

# atg_pose_in / atg_pose_out (include/atg.h)
POSE_IN_DTYPE = np.dtype([("H", "<f8", 9), ("p", "<f8", 8), ("tagsize", "<f8"), ("fx", "<f8"), ("fy", "<f8"),
                          ("cx", "<f8"), ("cy", "<f8")])
POSE_OUT_DTYPE = np.dtype([("hpose", "<f8", 12), ("R0", "<f8", 9), ("t0", "<f8", 3), ("R1", "<f8", 9),
                           ("t1", "<f8", 3), ("err1", "<f8"), ("R2", "<f8", 9), ("t2", "<f8", 3), ("err2", "<f8"),
                           ("has2", "<i4"), ("best", "<i4")])
POSE_HOMOGRAPHY_ONLY = 1

_lib = None


def lib():
    """The loaded library (raises if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise RuntimeError("libapriltag_amd.so is not built: run apriltag-fast_amd/build.py")
    L = C.CDLL(str(LIB_PATH))
    vp, i32, i64 = C.c_void_p, C.c_int32, C.c_int64
    L.atg_params_default.argtypes = [C.POINTER(Params)]
    L.atg_ctx_create.argtypes = [C.c_int, C.POINTER(vp)]
    L.atg_ctx_create.restype = C.c_int
    L.atg_ctx_destroy.argtypes = [vp]
    L.atg_last_error.argtypes = [vp]
    L.atg_last_error.restype = C.c_char_p
    L.atg_set_families.argtypes = [vp, C.POINTER(C.POINTER(Family)), C.POINTER(i32), i32]
    L.atg_set_families.restype = C.c_int
    L.atg_detect.argtypes = [vp, C.POINTER(Params), vp, i64, i64, i64, C.c_int, C.POINTER(Result)]
    L.atg_detect.restype = C.c_int
    L.atg_detect_batch.argtypes = [vp, C.POINTER(Params), C.POINTER(Frame), i32, C.c_int, C.POINTER(Result)]
    L.atg_detect_batch.restype = C.c_int
    L.atg_debug_stages.argtypes = [vp, i32, C.POINTER(Stages)]
    L.atg_debug_stages.restype = C.c_int
    L.atg_stages_free.argtypes = [C.POINTER(Stages)]
    L.atg_debug_counters.argtypes = [vp, i32, C.POINTER(C.c_uint32)]
    L.atg_debug_counters.restype = C.c_int
    L.atg_last_timing.argtypes = [vp, C.POINTER(C.c_double)]
    L.atg_last_timing.restype = C.c_int
    L.atg_set_profiling.argtypes = [vp, C.c_int]
    L.atg_set_profiling.restype = C.c_int
    L.atg_debug_copy.argtypes = [vp, vp, vp, i64, C.POINTER(C.c_double)]
    L.atg_debug_copy.restype = C.c_int
    L.atg_debug_force_scratch.argtypes = [vp, C.c_int]
    L.atg_debug_force_scratch.restype = C.c_int
    L.atg_debug_finish.argtypes = [vp, vp, C.POINTER(i64)]
    L.atg_debug_finish.restype = C.c_int
    L.atg_get_tune.argtypes = [vp, C.POINTER(Tune)]
    L.atg_get_tune.restype = C.c_int
    L.atg_set_tune.argtypes = [vp, C.POINTER(Tune)]
    L.atg_set_tune.restype = C.c_int
    L.atg_set_graphs.argtypes = [vp, C.c_int]
    L.atg_set_graphs.restype = C.c_int
    L.atg_autotune.argtypes = [vp, C.POINTER(Params), C.POINTER(Frame), i32, C.c_int, C.POINTER(Tune),
                               C.POINTER(C.c_double)]
    L.atg_autotune.restype = C.c_int
    L.atg_stage_times.argtypes = [vp, C.POINTER(C.c_char_p), C.POINTER(C.c_double), i32]
    L.atg_stage_times.restype = C.c_int
    L.atg_device_alloc.argtypes = [vp, C.c_size_t, C.POINTER(vp)]
    L.atg_device_alloc.restype = C.c_int
    L.atg_device_free.argtypes = [vp, vp]
    L.atg_device_free.restype = C.c_int
    L.atg_copy_to_device.argtypes = [vp, vp, vp, C.c_size_t]
    L.atg_copy_to_device.restype = C.c_int
    L.atg_builtin_family.argtypes = [i32]
    L.atg_builtin_family.restype = C.POINTER(Family)
    L.atg_synth_default.argtypes = [C.POINTER(SynthSpec), C.c_int]
    L.atg_synth_frame.argtypes = [C.POINTER(SynthSpec), C.c_uint64, vp, i64, C.POINTER(SynthTag), i32]
    L.atg_synth_frame.restype = C.c_int
    L.atg_synth_frame_drift.argtypes = [C.POINTER(SynthSpec), C.c_uint64, i64, vp, i64, C.POINTER(SynthTag), i32]
    L.atg_synth_frame_drift.restype = C.c_int
    L.atg_server_config_default.argtypes = [C.POINTER(ServerConfig)]
    L.atg_server_create.argtypes = [C.POINTER(ServerConfig), C.POINTER(Params), C.POINTER(C.POINTER(Family)),
                                    C.POINTER(i32), i32, C.POINTER(vp)]
    L.atg_server_create.restype = C.c_int
    L.atg_server_destroy.argtypes = [vp]
    L.atg_server_error.argtypes = [vp]
    L.atg_server_error.restype = C.c_char_p
    L.atg_server_now_ns.argtypes = [vp]
    L.atg_server_now_ns.restype = i64
    L.atg_server_add_stream.argtypes = [vp, i32, i32, C.POINTER(i32)]
    L.atg_server_add_stream.restype = C.c_int
    L.atg_server_submit.argtypes = [vp, i32, vp, i64, C.c_uint64]
    L.atg_server_submit.restype = C.c_int
    L.atg_server_feed.argtypes = [vp, C.POINTER(Arrival), i64]
    L.atg_server_feed.restype = C.c_int
    L.atg_server_poll.argtypes = [vp, C.POINTER(Served), i32]
    L.atg_server_poll.restype = C.c_int
    L.atg_server_drain.argtypes = [vp, i64]
    L.atg_server_drain.restype = C.c_int
    L.atg_server_stats.argtypes = [vp, C.POINTER(ServerCounts)]
    L.atg_server_stats.restype = C.c_int
    L.atg_host_alloc.argtypes = [C.c_size_t, C.POINTER(vp)]
    L.atg_host_alloc.restype = C.c_int
    L.atg_host_free.argtypes = [vp]
    L.atg_host_free.restype = C.c_int
    L.atg_debug_trig.argtypes = [vp, vp, vp, vp, i64, vp, vp, vp, C.c_int]
    L.atg_debug_trig.restype = C.c_int
    L.atg_estimate_poses.argtypes = [vp, vp, i64, i32, i32, vp]
    L.atg_estimate_poses.restype = C.c_int
    _lib = L
    return L


def default_params(**kw):
    p = Params()
    lib().atg_params_default(C.byref(p))
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def params_from_dict(d):
    """Parameters from a dict in the at_detector naming (see tests)."""
    keys = [f for f, _ in Params._fields_]
    return default_params(**{k: v for k, v in d.items() if k in keys})


def _dets(ptr, n):
    if n <= 0:
        return np.zeros(0, DET_DTYPE)
    return np.frombuffer(C.string_at(ptr, n * DET_DTYPE.itemsize), DET_DTYPE).copy()


class Context:
    """One device context (atg_ctx)."""

    def __init__(self, device=0):
        self.L = lib()
        self.h = C.c_void_p()
        rc = self.L.atg_ctx_create(device, C.byref(self.h))
        if rc != 0:
            raise RuntimeError("atg_ctx_create failed (%d)" % rc)
        self._fams = []

    def close(self):
        if self.h:
            self.L.atg_ctx_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc, what):
        if rc != 0:
            raise RuntimeError("%s failed (%d): %s" % (what, rc, self.L.atg_last_error(self.h).decode()))

    def set_families(self, families):
        """families: list of (name, bits_corrected)."""
        ptrs = (C.POINTER(Family) * max(1, len(families)))()
        bits = (C.c_int32 * max(1, len(families)))()
        for i, (name, b) in enumerate(families):
            ptrs[i] = self.L.atg_builtin_family(FAMILY_INDEX[name])
            bits[i] = b
        self._check(self.L.atg_set_families(self.h, ptrs, bits, len(families)), "atg_set_families")

    def detect(self, img, params=None, on_device_ptr=None):
        p = params if params is not None else default_params()
        r = Result()
        if on_device_ptr is not None:
            ptr, h, w, stride = on_device_ptr
            self._check(self.L.atg_detect(self.h, C.byref(p), ptr, w, h, stride, 1, C.byref(r)), "atg_detect")
        else:
            img = np.ascontiguousarray(img, np.uint8)
            self._check(self.L.atg_detect(self.h, C.byref(p), img.ctypes.data, img.shape[1], img.shape[0],
                                          img.strides[0], 0, C.byref(r)), "atg_detect")
        return dict(dets=_dets(r.dets, r.ndets), nquads=r.nquads)

    @staticmethod
    def _frames(imgs, on_device):
        n = len(imgs)
        fr = (Frame * n)()
        keep = []
        for i, im in enumerate(imgs):
            if on_device:
                ptr, h, w, stride = im
                fr[i] = Frame(ptr, w, h, stride)
            else:
                a = np.ascontiguousarray(im, np.uint8)
                keep.append(a)
                fr[i] = Frame(a.ctypes.data, a.shape[1], a.shape[0], a.strides[0])
        return fr, keep

    def detect_batch(self, imgs, params=None, on_device=False):
        """imgs: list of uint8 arrays, or with on_device a list of
        (device_ptr, height, width, stride)."""
        p = params if params is not None else default_params()
        n = len(imgs)
        fr, keep = self._frames(imgs, on_device)
        res = (Result * n)()
        self._check(self.L.atg_detect_batch(self.h, C.byref(p), fr, n, 1 if on_device else 0, res),
                    "atg_detect_batch")
        return [dict(dets=_dets(r.dets, r.ndets), nquads=r.nquads) for r in res]

    def chain_ms(self):
        v = C.c_double()
        self.L.atg_last_timing(self.h, C.byref(v))
        return v.value

    def to_device(self, img):
        """Copy a frame into a new device buffer: (ptr, height, width, stride)."""
        img = np.ascontiguousarray(img, np.uint8)
        p = C.c_void_p()
        self._check(self.L.atg_device_alloc(self.h, img.nbytes, C.byref(p)), "atg_device_alloc")
        self._check(self.L.atg_copy_to_device(self.h, p, img.ctypes.data, img.nbytes), "atg_copy_to_device")
        return (p.value, img.shape[0], img.shape[1], img.shape[1])

    def free_device(self, ptr):
        self._check(self.L.atg_device_free(self.h, ptr), "atg_device_free")

    def set_profiling(self, on=True):
        """on: True/1 device clock stamps (graphs stay on), 2 HIP events, False/0 off."""
        self._check(self.L.atg_set_profiling(self.h, int(on)), "atg_set_profiling")

    def get_tune(self):
        t = Tune()
        self._check(self.L.atg_get_tune(self.h, C.byref(t)), "atg_get_tune")
        return tuple(getattr(t, f) for f in TUNE_FIELDS)

    def set_tune(self, thr_tile_rows, row_waves, front_threads, fit_blocks, front_bands=0):
        """front_bands 0: the union-find front's row bands chosen by the call's size."""
        t = Tune(thr_tile_rows, row_waves, front_threads, fit_blocks, front_bands)
        self._check(self.L.atg_set_tune(self.h, C.byref(t)), "atg_set_tune")

    def set_graphs(self, on=True):
        self._check(self.L.atg_set_graphs(self.h, 1 if on else 0), "atg_set_graphs")

    def autotune(self, imgs, params=None, on_device=False):
        """Times the work splits on this batch; returns (tune tuple, ms)."""
        p = params if params is not None else default_params()
        fr, keep = self._frames(imgs, on_device)
        t = Tune()
        ms = C.c_double()
        self._check(self.L.atg_autotune(self.h, C.byref(p), fr, len(imgs), 1 if on_device else 0, C.byref(t),
                                        C.byref(ms)), "atg_autotune")
        return tuple(getattr(t, f) for f in TUNE_FIELDS), ms.value

    def copy_bandwidth(self, nbytes=1 << 29, reps=3):
        """A plain copy kernel over nbytes, reps times: (GB/s moved, ms each)."""
        a, b = C.c_void_p(), C.c_void_p()
        self._check(self.L.atg_device_alloc(self.h, nbytes, C.byref(a)), "atg_device_alloc")
        self._check(self.L.atg_device_alloc(self.h, nbytes, C.byref(b)), "atg_device_alloc")
        try:
            ms = C.c_double()
            times = []
            for _ in range(reps):
                self._check(self.L.atg_debug_copy(self.h, b, a, nbytes, C.byref(ms)), "atg_debug_copy")
                times.append(ms.value)
        finally:
            self.L.atg_device_free(self.h, a)
            self.L.atg_device_free(self.h, b)
        best = min(times)
        return 2 * nbytes / (best * 1e-3) / 1e9, times

    def finish(self, dets):
        """The device's reconcile and stable id sort on a DET_DTYPE array."""
        d = np.ascontiguousarray(dets, DET_DTYPE).copy()
        n = C.c_int64(len(d))
        self._check(self.L.atg_debug_finish(self.h, d.ctypes.data if len(d) else None, C.byref(n)),
                    "atg_debug_finish")
        return d[:n.value].copy()

    def force_scratch(self, on=True):
        """Keep the union-find replay state in HBM scratch (the wide-frame path)."""
        self._check(self.L.atg_debug_force_scratch(self.h, 1 if on else 0), "atg_debug_force_scratch")

    def estimate_poses(self, tags, n_iters=50, flags=0):
        """Poses of many tags in one launch: tags is a POSE_IN_DTYPE array;
        returns a POSE_OUT_DTYPE array."""
        t = np.ascontiguousarray(tags, POSE_IN_DTYPE)
        out = np.zeros(len(t), POSE_OUT_DTYPE)
        self._check(self.L.atg_estimate_poses(self.h, t.ctypes.data if len(t) else None, len(t), n_iters, flags,
                                              out.ctypes.data if len(t) else None), "atg_estimate_poses")
        return out

    def trig(self, y, x, th, slow=False):
        """The refinement's correctly rounded atan2f(y, x), cosf(th),
        sinf(th) on the device (equal-length float32 arrays)."""
        y, x, th = (np.ascontiguousarray(a, np.float32) for a in (y, x, th))
        n = len(y)
        assert len(x) == n and len(th) == n
        at, co, si = (np.zeros(n, np.float32) for _ in range(3))
        self._check(self.L.atg_debug_trig(self.h, y.ctypes.data, x.ctypes.data, th.ctypes.data, n, at.ctypes.data,
                                          co.ctypes.data, si.ctypes.data, 1 if slow else 0), "atg_debug_trig")
        return at, co, si

    def stage_times(self):
        """[(group.kernel, ms)] of the last call with profiling on."""
        names = (C.c_char_p * 64)()
        ms = (C.c_double * 64)()
        n = self.L.atg_stage_times(self.h, names, ms, 64)
        return [(names[i].decode(), ms[i]) for i in range(min(n, 64))]

    def counters(self, frame=0):
        out = (C.c_uint32 * 16)()
        self._check(self.L.atg_debug_counters(self.h, frame, out), "atg_debug_counters")
        return list(out)

    def stages(self, frame=0):
        s = Stages()
        self._check(self.L.atg_debug_stages(self.h, frame, C.byref(s)), "atg_debug_stages")

        def arr(ptr, n, dt):
            if n <= 0:
                return np.zeros(0, dt)
            return np.ctypeslib.as_array(ptr, (n,)).astype(dt, copy=True)
        qw, qh = s.qw, s.qh
        out = dict(qw=qw, qh=qh,
                   qim=arr(s.qim, qw * qh, np.uint8).reshape(qh, qw),
                   t=arr(s.t, qw * qh, np.uint8).reshape(qh, qw),
                   runs=arr(s.runs, 3 * s.nruns, np.uint32).reshape(-1, 3),
                   row_off=arr(s.row_off, qh + 1, np.uint32),
                   rep=arr(s.rep, s.nnodes, np.uint32), csize=arr(s.csize, s.nnodes, np.uint32),
                   ok=arr(s.ok, s.nnodes, np.uint8),
                   cluster_ids=arr(s.cluster_ids, s.nclusters, np.uint64),
                   ccount=arr(s.ccount, s.nclusters, np.uint32),
                   pts=arr(s.pts, s.npts, np.uint64),
                   quads=arr(s.quads, 9 * s.nquads, np.float32).reshape(-1, 9))
        self.L.atg_stages_free(C.byref(s))
        return out


def _family_arrays(families):
    L = lib()
    ptrs = (C.POINTER(Family) * max(1, len(families)))()
    bits = (C.c_int32 * max(1, len(families)))()
    for i, (name, b) in enumerate(families):
        ptrs[i] = L.atg_builtin_family(FAMILY_INDEX[name])
        bits[i] = b
    return ptrs, bits


class PinnedFrame:
    """A uint8 HxW frame in pinned host memory (DMA-able without staging)."""

    def __init__(self, h, w):
        self.ptr = C.c_void_p()
        if lib().atg_host_alloc(h * w, C.byref(self.ptr)) != 0:
            raise RuntimeError("atg_host_alloc failed")
        self.array = np.ctypeslib.as_array((C.c_uint8 * (h * w)).from_address(self.ptr.value)).reshape(h, w)

    def free(self):
        if self.ptr:
            lib().atg_host_free(self.ptr)
            self.ptr = C.c_void_p()


class Server:
    """The serving loop (atg_server): streams, arrivals, polled results."""

    def __init__(self, families, params=None, budget_bytes=16 << 30, lanes=3, max_batch=128, slots=0, device=0):
        self.L = lib()
        cfg = ServerConfig()
        self.L.atg_server_config_default(C.byref(cfg))
        cfg.device, cfg.budget_bytes, cfg.lanes, cfg.max_batch, cfg.slots_per_stream = (
            device, budget_bytes, lanes, max_batch, slots)
        self.params = params if params is not None else default_params()
        ptrs, bits = _family_arrays(families)
        self.h = C.c_void_p()
        rc = self.L.atg_server_create(C.byref(cfg), C.byref(self.params), ptrs, bits, len(families),
                                      C.byref(self.h))
        if rc != 0:
            raise RuntimeError("atg_server_create failed (%d)" % rc)
        # frames the server borrows, by the id passed to it as the frame's
        # tag: (caller's tag, array), released when the frame's record is polled
        self._keep = {}
        self._next_id = 0

    def _check(self, rc, what):
        if rc < 0:
            raise RuntimeError("%s: %s (%d)" % (what, self.L.atg_server_error(self.h).decode(), rc))
        return rc

    def _hold(self, img, tag):
        self._next_id += 1
        self._keep[self._next_id] = (tag, img)
        return self._next_id

    def add_stream(self, width, height):
        sid = C.c_int32()
        self._check(self.L.atg_server_add_stream(self.h, width, height, C.byref(sid)), "atg_server_add_stream")
        return sid.value

    def now_ns(self):
        return self.L.atg_server_now_ns(self.h)

    def submit(self, stream, img, tag=0):
        """img: a uint8 HxW array, held until its result is polled."""
        fid = self._hold(img, tag)
        rc = self.L.atg_server_submit(self.h, stream, img.ctypes.data, img.strides[0], fid)
        if rc != 0:
            del self._keep[fid]  # refused or dropped: the server holds nothing
        return rc

    def feed(self, arrivals):
        """arrivals: list of (stream, at_ns, img, tag); each img is held until
        its record (or, if dropped, its ATG_ERR_LIMIT record) is polled."""
        arr = (Arrival * max(1, len(arrivals)))()
        ids = [self._hold(img, tag) for (_, _, img, tag) in arrivals]
        for i, (sid, at, img, _) in enumerate(arrivals):
            arr[i] = Arrival(sid, int(at), img.ctypes.data, img.strides[0], ids[i])
        rc = self.L.atg_server_feed(self.h, arr, len(arrivals))
        if rc < 0:
            for f in ids:
                del self._keep[f]
        self._check(rc, "atg_server_feed")

    def poll(self, cap=1024):
        out = (Served * cap)()
        n = self._check(self.L.atg_server_poll(self.h, out, cap), "atg_server_poll")
        res = []
        for i in range(n):
            o = out[i]
            tag, _ = self._keep.pop(o.tag)
            res.append(dict(stream=o.stream, status=o.status, tag=tag, arrival_ns=o.arrival_ns,
                            admit_ns=o.admit_ns, done_ns=o.done_ns, nquads=o.nquads, batch_frames=o.batch_frames,
                            dets=_dets(o.dets, o.ndets)))
        return res

    def drain(self, timeout_ms=60000):
        self._check(self.L.atg_server_drain(self.h, timeout_ms), "atg_server_drain")

    def counts(self):
        c = ServerCounts()
        self._check(self.L.atg_server_stats(self.h, C.byref(c)), "atg_server_stats")
        return {f: getattr(c, f) for f, _ in ServerCounts._fields_}

    def close(self):
        if self.h:
            self.L.atg_server_destroy(self.h)
            self.h = C.c_void_p()
            self._keep = {}


def synth_spec(config, width=None, height=None, ntags=None):
    s = SynthSpec()
    lib().atg_synth_default(C.byref(s), config)
    if width:
        s.width = width
    if height:
        s.height = height
    if ntags is not None:
        s.ntags_min = s.ntags_max = ntags
    return s


def synth_stream_frame(spec, seed, step, out=None):
    """Frame `step` of the camera stream `seed` (tags drifting slowly) into
    out (a uint8 array, e.g. a PinnedFrame's) or a new array."""
    img = out if out is not None else np.zeros((spec.height, spec.width), np.uint8)
    n = lib().atg_synth_frame_drift(C.byref(spec), seed, step, img.ctypes.data, img.strides[0], None, 0)
    if n < 0:
        raise RuntimeError("atg_synth_frame_drift failed")
    return img


def synth_frame(config, seed, width=None, height=None, ntags=None, family=None):
    """A seeded synthetic frame (uint8 HxW) and its tags (family, id);
    family: the index of the built-in family its tags come from (the
    configuration's own by default)."""
    L = lib()
    s = SynthSpec()
    L.atg_synth_default(C.byref(s), config)
    if family is not None:
        s.family_a = family
    if width:
        s.width = width
    if height:
        s.height = height
    if ntags is not None:
        s.ntags_min = s.ntags_max = ntags
    img = np.zeros((s.height, s.width), np.uint8)
    truth = (SynthTag * 256)()
    n = L.atg_synth_frame(C.byref(s), seed, img.ctypes.data, s.width, truth, 256)
    if n < 0:
        raise RuntimeError("atg_synth_frame failed")
    return img, [(truth[i].family, truth[i].id) for i in range(min(n, 256))]
