"""Build the MI355X detection library in-tree, and install it as a drop-in
for the reference's libapriltag.

    python apriltag-fast_amd/build.py [--force]
    python apriltag-fast_amd/build.py --install PREFIX

Compiles every HIP/C++ source of csrc/ for gfx950 with hipcc into
apriltag-fast_amd/libapriltag_amd.so. Device code is compiled without
floating-point contraction (the detector's results are bit-exact against
the reference, which is built without FMA), and without packed-FP32
instructions. Object files go to apriltag-fast_amd/_build/ and are rebuilt
when their source or a header changed.

--install lays the library out as the reference's CMake install does
(CMakeLists.txt:51-98): PREFIX/lib/libapriltag.so.3.4.5 (SONAME
libapriltag.so.3) with its libapriltag.so.3 and libapriltag.so links, the
headers under PREFIX/include/apriltag/ (common/ and fast/ beside them),
PREFIX/lib/pkgconfig/apriltag.pc, and the package files of
find_package(apriltag) under PREFIX/lib/apriltag/cmake/ with the imported
target apriltag::apriltag.
"""
import os
import shutil
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
# the upstream API version this library implements (the reference's
# project version, CMakeLists.txt:3) and its ABI's SONAME
VERSION = "3.4.5"
SONAME = "libapriltag.so.3"
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


def link(objs, out, soname):
    cmd = [HIPCC, "-shared", "-fPIC", "--offload-arch=" + ARCH, "-Wl,-soname," + soname, "-o", str(out) + ".tmp"]
    r = subprocess.run(cmd + [str(o) for o in objs], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n%s\n%s%s" % (" ".join(cmd), r.stdout, r.stderr))
    os.replace(str(out) + ".tmp", out)


def build(force=False, verbose=True):
    BUILD.mkdir(exist_ok=True)
    srcs = sources()
    with ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        res = list(ex.map(lambda s: compile_one(s, force), srcs))
    errs = [e for _, e in res if e]
    if errs:
        raise RuntimeError("hipcc failed:\n" + "\n".join(errs))
    objs = [str(o) for o, _ in res]
    link(objs, OUT, OUT.name)
    if verbose:
        print("built", OUT)
    return OUT


PC_TEMPLATE = """prefix={prefix}
exec_prefix=${{prefix}}
includedir=${{prefix}}/include
libdir=${{exec_prefix}}/lib

Name: apriltag
Description: AprilTag detection on the AMD MI355X (GPU), the upstream and fast C APIs
Version: {version}
Cflags: -I${{includedir}}
Libs: -L${{libdir}} -lapriltag
"""

CONFIG_TEMPLATE = """# find_package(apriltag): the imported target apriltag::apriltag
include(CMakeFindDependencyMacro)
find_dependency(Threads)
get_filename_component(_apriltag_prefix "${{CMAKE_CURRENT_LIST_DIR}}/../../.." ABSOLUTE)
if(NOT TARGET apriltag::apriltag)
  add_library(apriltag::apriltag SHARED IMPORTED)
  set_target_properties(apriltag::apriltag PROPERTIES
    IMPORTED_LOCATION "${{_apriltag_prefix}}/lib/libapriltag.so.{version}"
    IMPORTED_SONAME "{soname}"
    INTERFACE_INCLUDE_DIRECTORIES "${{_apriltag_prefix}}/include;${{_apriltag_prefix}}/include/apriltag"
    INTERFACE_LINK_LIBRARIES "Threads::Threads;m")
endif()
set(apriltag_FOUND TRUE)
"""

VERSION_TEMPLATE = """set(PACKAGE_VERSION "{version}")
# same major version: compatible (the reference's SameMajorVersion)
if(PACKAGE_FIND_VERSION_MAJOR STREQUAL "{major}" AND NOT PACKAGE_FIND_VERSION VERSION_GREATER PACKAGE_VERSION)
  set(PACKAGE_VERSION_COMPATIBLE TRUE)
  if(PACKAGE_FIND_VERSION STREQUAL PACKAGE_VERSION)
    set(PACKAGE_VERSION_EXACT TRUE)
  endif()
elseif(NOT PACKAGE_FIND_VERSION)
  set(PACKAGE_VERSION_COMPATIBLE TRUE)
else()
  set(PACKAGE_VERSION_COMPATIBLE FALSE)
endif()
"""


def install(prefix):
    """Install the built library under prefix (see the module docstring)."""
    prefix = Path(prefix).resolve()
    build()
    lib, inc = prefix / "lib", prefix / "include" / "apriltag"
    for d in (lib / "pkgconfig", lib / "apriltag" / "cmake", inc / "common", inc / "fast"):
        d.mkdir(parents=True, exist_ok=True)
    real = lib / ("libapriltag.so." + VERSION)
    # the same objects linked again under the upstream library's SONAME
    # (the in-tree library keeps its own name)
    link(sorted(BUILD.glob("*.o")), real, SONAME)
    for name, target in ((SONAME, real.name), ("libapriltag.so", SONAME)):
        ln = lib / name
        if ln.is_symlink() or ln.exists():
            ln.unlink()
        ln.symlink_to(target)
    src = REPO / "include"
    for h in sorted(src.glob("*.h")):
        shutil.copy2(h, inc / h.name)
    for h in sorted((src / "common").glob("*.h")):
        shutil.copy2(h, inc / "common" / h.name)
    shutil.copy2(src / "apriltag_fast.h", inc / "fast" / "apriltag_fast.h")  # the reference's fast/ path
    (lib / "pkgconfig" / "apriltag.pc").write_text(PC_TEMPLATE.format(prefix=prefix, version=VERSION))
    cm = lib / "apriltag" / "cmake"
    (cm / "apriltagConfig.cmake").write_text(CONFIG_TEMPLATE.format(version=VERSION, soname=SONAME))
    (cm / "apriltagConfigVersion.cmake").write_text(VERSION_TEMPLATE.format(version=VERSION,
                                                                            major=VERSION.split(".")[0]))
    return prefix


if __name__ == "__main__":
    if "--install" in sys.argv:
        build(force="--force" in sys.argv)
        print("installed under", install(sys.argv[sys.argv.index("--install") + 1]))
    else:
        build(force="--force" in sys.argv)
