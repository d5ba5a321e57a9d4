"""The installed library as a drop-in for the reference's libapriltag
(its CMake install, CMakeLists.txt:51-98): the SONAME and links, the
header tree, pkg-config and find_package(apriltag). CPU only: the programs
built here print struct sizes and use host functions, no GPU call."""
import os
import re
import subprocess
import sys
from pathlib import Path

import pytest

from atg_testlib import PKG, REPO

sys.path.insert(0, str(PKG))
import build as atg_build  # noqa: E402

REF = Path("/root/reference")

PROGRAM = r"""
#include <stdio.h>
#include <apriltag/apriltag.h>
#include <apriltag/apriltag_pose.h>
#include <apriltag/tag36h11.h>
#include <apriltag/common/matd.h>
int main(void) {
    apriltag_family_t *tf = tag36h11_create();
    matd_t *a = matd_identity(3);
    matd_t *b = matd_scale(a, 2.0);
    double d = matd_det(b);
    printf("%u %u %zu %zu %g\n", tf->ncodes, tf->nbits, sizeof(apriltag_detection_t), sizeof(apriltag_pose_t), d);
    matd_destroy(a);
    matd_destroy(b);
    tag36h11_destroy(tf);
    return 0;
}
"""
EXPECT = "587 36 112 16 8"


@pytest.fixture(scope="module")
def prefix(tmp_path_factory):
    return atg_build.install(tmp_path_factory.mktemp("prefix"))


def pc_flags(pc):
    """Cflags and Libs of a .pc file, its ${var} expanded."""
    env, fields = {}, {}
    for line in pc.read_text().splitlines():
        m = re.match(r"^(\w+)=(.*)$", line)
        if m:
            env[m.group(1)] = m.group(2)
        m = re.match(r"^(\w+):\s*(.*)$", line)
        if m:
            fields[m.group(1)] = m.group(2)

    def expand(s):
        for _ in range(5):
            s = re.sub(r"\$\{(\w+)\}", lambda m: env[m.group(1)], s)
        return s
    for k in list(env):
        env[k] = expand(env[k])
    return expand(fields["Cflags"]).split(), expand(fields["Libs"]).split(), fields["Version"]


def test_install_layout(prefix):
    lib = prefix / "lib"
    real = lib / "libapriltag.so.3.4.5"
    assert real.exists() and os.readlink(lib / "libapriltag.so.3") == real.name
    assert os.readlink(lib / "libapriltag.so") == "libapriltag.so.3"
    dyn = subprocess.run(["readelf", "-d", str(real)], capture_output=True, text=True, check=True).stdout
    assert "Library soname: [libapriltag.so.3]" in dyn
    inc = prefix / "include" / "apriltag"
    for h in ("apriltag.h", "apriltag_pose.h", "tag36h11.h", "common/matd.h", "common/zarray.h",
              "common/image_u8.h", "common/homography.h", "fast/apriltag_fast.h"):
        assert (inc / h).exists(), h


def test_pkg_config_program(prefix, tmp_path):
    cflags, libs, version = pc_flags(prefix / "lib" / "pkgconfig" / "apriltag.pc")
    assert version == "3.4.5"
    src = tmp_path / "main.c"
    src.write_text(PROGRAM)
    exe = tmp_path / "main"
    subprocess.run(["gcc", "-Wall", "-Werror", *cflags, str(src), *libs, "-Wl,-rpath," + str(prefix / "lib"), "-o",
                    str(exe)], check=True, capture_output=True, text=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.strip()
    assert out == EXPECT


def test_cmake_find_package(prefix, tmp_path):
    proj = tmp_path / "proj"
    proj.mkdir()
    (proj / "main.c").write_text(PROGRAM)
    (proj / "CMakeLists.txt").write_text(
        "cmake_minimum_required(VERSION 3.16)\nproject(demo C)\nfind_package(apriltag 3 REQUIRED)\n"
        "add_executable(demo main.c)\ntarget_link_libraries(demo apriltag::apriltag)\n")
    bld = tmp_path / "bld"
    subprocess.run(["cmake", "-S", str(proj), "-B", str(bld), "-DCMAKE_PREFIX_PATH=" + str(prefix)], check=True,
                   capture_output=True, text=True)
    subprocess.run(["cmake", "--build", str(bld)], check=True, capture_output=True, text=True)
    out = subprocess.run([str(bld / "demo")], check=True, capture_output=True, text=True).stdout.strip()
    assert out == EXPECT
    # another major version is refused (the reference's SameMajorVersion)
    (proj / "CMakeLists.txt").write_text(
        "cmake_minimum_required(VERSION 3.16)\nproject(demo C)\nfind_package(apriltag 4 REQUIRED)\n")
    r = subprocess.run(["cmake", "-S", str(proj), "-B", str(tmp_path / "bld4"), "-DCMAKE_PREFIX_PATH=" + str(prefix)],
                       capture_output=True, text=True)
    assert r.returncode != 0


@pytest.mark.skipif(not (REF / "apriltag.h").exists(), reason="the reference's headers are absent")
def test_program_built_against_reference_headers_links(prefix, tmp_path):
    """A user's program compiled against the REFERENCE's own headers links
    and runs against this library (the same symbols and layouts)."""
    exe = tmp_path / "upstream_demo"
    subprocess.run(["gcc", "-O1", "-I", str(REF), str(REPO / "tests" / "c" / "upstream_demo.c"), "-L",
                    str(prefix / "lib"), "-lapriltag", "-lm", "-lpthread", "-Wl,-rpath," + str(prefix / "lib"),
                    "-o", str(exe)], check=True, capture_output=True, text=True)
    out = subprocess.run([str(exe), "-s"], check=True, capture_output=True, text=True).stdout
    sizes = {l.split()[0]: int(l.split()[1]) for l in out.splitlines()}
    assert sizes["apriltag_detector_t"] == 160 and sizes["apriltag_detection_t"] == 112
    assert sizes["apriltag_detection_info_t"] == 48
