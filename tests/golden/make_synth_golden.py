"""Write tests/golden/synth_expected.npz: the reference build's detections
(oracle/_ref/libatref.so, single thread) on the seeded synthetic cases of
tests/atg_testlib.SYNTH_CASES, with a digest of each frame (which pins the
generator). Run in a container where `make -C oracle ref` built the
reference; the fixture is data and travels with the tests.

    python tests/golden/make_synth_golden.py
"""
import sys
from pathlib import Path

import numpy as np

sys.path.insert(0, str(Path(__file__).resolve().parent.parent))
from atg_testlib import GOLDEN, SYNTH_CASES, Reference, frame_hash, load_atg, synth_case  # noqa: E402


def main():
    atg = load_atg()
    ref = Reference()
    out = {}
    for cfg, seed in SYNTH_CASES:
        img, p, fams = synth_case(atg, cfg, seed)
        r = ref.detect(img, p, fams)
        tag = "c%d_%d" % (cfg, seed)
        out[tag + "_dets"] = r["dets"]
        out[tag + "_nquads"] = np.int64(r["nquads"])
        out[tag + "_hash"] = np.uint64(frame_hash(img))
        print(tag, img.shape, len(r["dets"]), "dets", r["nquads"], "quads")
    np.savez_compressed(GOLDEN / "synth_expected.npz", **out)


if __name__ == "__main__":
    main()
