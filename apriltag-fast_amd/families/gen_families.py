"""Pack the AprilTag family tables into families.bin.

The nine tag families (code words and bit layouts) are data that a drop-in
detector must reproduce exactly: a tag id IS an index into its family's code
list. This script reads them from the reference's family files
(tag*.c, e.g. tag36h11.c:31 `codedata[]` and tag36h11.c:620-706
`tag36h11_create`) and writes one little-endian blob in this project's own
layout, which the library embeds at build time:

    "ATGF" u32 version=1 u32 nfamilies
    per family:
        char name[32]
        u32 ncodes, nbits, h
        i32 width_at_border, total_width
        u32 reversed_border
        i32 bit_x[nbits], bit_y[nbits]
        u64 codes[ncodes]

The families were generated by the AprilTag project (University of Michigan,
BSD 2-clause); see LICENSE.families in this folder.

Run (in a container that has the reference):
    python gen_families.py /path/to/reference
"""
import re
import struct
import sys
from pathlib import Path

FAMILIES = [
    "tag16h5", "tag25h9", "tag36h10", "tag36h11", "tagCircle21h7",
    "tagCircle49h12", "tagCustom48h12", "tagStandard41h12", "tagStandard52h13",
]


def parse_family(text, name):
    m = re.search(r"codedata\[(\d+)\]\s*=\s*\{(.*?)\};", text, re.S)
    ncodes_decl = int(m.group(1))
    codes = [int(c, 16) for c in re.findall(r"0x([0-9a-fA-F]+)UL", m.group(2))]
    assert len(codes) == ncodes_decl, name

    def field(key):
        mm = re.search(r"tf->%s\s*=\s*([^;]+);" % key, text)
        return mm.group(1).strip()

    nbits = int(field("nbits"))
    fields = {
        "h": int(field("h")),
        "ncodes": int(field("ncodes")),
        "wab": int(field("width_at_border")),
        "tw": int(field("total_width")),
        "rev": field("reversed_border") == "true",
    }
    assert fields["ncodes"] == len(codes)
    bx = [0] * nbits
    by = [0] * nbits
    for i, v in re.findall(r"tf->bit_x\[(\d+)\]\s*=\s*(-?\d+);", text):
        bx[int(i)] = int(v)
    for i, v in re.findall(r"tf->bit_y\[(\d+)\]\s*=\s*(-?\d+);", text):
        by[int(i)] = int(v)
    return nbits, fields, bx, by, codes


def main():
    ref = Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
    out = bytearray(b"ATGF" + struct.pack("<II", 1, len(FAMILIES)))
    for name in FAMILIES:
        nbits, f, bx, by, codes = parse_family((ref / (name + ".c")).read_text(), name)
        out += name.encode().ljust(32, b"\0")
        out += struct.pack("<IIIiiI", len(codes), nbits, f["h"], f["wab"], f["tw"], int(f["rev"]))
        out += struct.pack("<%di" % nbits, *bx)
        out += struct.pack("<%di" % nbits, *by)
        out += struct.pack("<%dQ" % len(codes), *codes)
    dst = Path(__file__).with_name("families.bin")
    dst.write_bytes(bytes(out))
    print("wrote", dst, len(out), "bytes")


if __name__ == "__main__":
    main()
