"""Build the MI355X detection library in-tree.

    python apriltag-fast_amd/build.py

Compiles every HIP/C++ source of csrc/ for gfx950 with hipcc into
apriltag-fast_amd/libapriltag_amd.so. Device code is compiled without
floating-point contraction (the detector's results are bit-exact against
the reference, which is built without FMA), and without packed-FP32
instructions. Object files go to apriltag-fast_amd/_build/ and are rebuilt
when their source or a header changed.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
REPO = PKG.parent
CSRC = PKG / "csrc"
OUT = PKG / "libapriltag_amd.so"
BUILD = PKG / "_build"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("ATG_ARCH", "gfx950")

COMMON = ["-O3", "-fPIC", "-std=c++17", "-ffp-contract=off", "-fno-fast-math",
          "-I", str(REPO / "include"), "-I", str(CSRC),
          "-DATG_FAMILIES_BIN=\"%s\"" % (PKG / "families" / "families.bin")]
DEVICE = ["--offload-arch=" + ARCH, "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def sources():
    return sorted(list(CSRC.glob("*.hip")) + list(CSRC.glob("*.cpp")))


def headers_mtime():
    hs = list(CSRC.glob("*.h")) + list((REPO / "include").glob("*.h")) + [PKG / "families" / "families.bin"]
    return max(h.stat().st_mtime for h in hs)


def compile_one(src, force):
    obj = BUILD / (src.name + ".o")
    if not force and obj.exists() and obj.stat().st_mtime >= max(src.stat().st_mtime, headers_mtime()):
        return obj, None
    cmd = [HIPCC, "-x", "hip"] + DEVICE + COMMON + ["-c", str(src), "-o", str(obj)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        return obj, "%s\n%s%s" % (" ".join(cmd), r.stdout, r.stderr)
    return obj, None


def build(force=False, verbose=True):
    BUILD.mkdir(exist_ok=True)
    srcs = sources()
    with ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        res = list(ex.map(lambda s: compile_one(s, force), srcs))
    errs = [e for _, e in res if e]
    if errs:
        raise RuntimeError("hipcc failed:\n" + "\n".join(errs))
    objs = [str(o) for o, _ in res]
    cmd = [HIPCC, "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", str(OUT) + ".tmp"] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n%s\n%s%s" % (" ".join(cmd), r.stdout, r.stderr))
    os.replace(str(OUT) + ".tmp", OUT)
    if verbose:
        print("built", OUT)
    return OUT


if __name__ == "__main__":
    build(force="--force" in sys.argv)
