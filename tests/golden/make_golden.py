"""Extract the reference-computed golden data used by the tests.

Run where the reference tree is present:
    python tests/golden/make_golden.py
Writes tests/golden/chig_frames.npz from the reference's own example run
(example/15.SpatialTransformation): the chignolin input (sequence, initial
positions, the ff_2.1 Rama maps stored in inputs/chig.up) and the 20 logged
frames of outputs/simple_test/chig.run.up (positions and the per-frame
potential and per-residue terms the reference logged).  Only data is copied.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "upside2-md_amd"))
import h5lite  # noqa: E402

EX = "/root/reference/example/15.SpatialTransformation"


def main():
    inp = h5lite.File(EX + "/inputs/chig.up")
    run = h5lite.File(EX + "/outputs/simple_test/chig.run.up")
    out = dict(
        sequence=np.array([s for s in inp.read("/input/sequence")], dtype="S3"),
        initial_pos=inp.read("/input/pos")[:, :, 0].astype(np.float32),
        rama_pot=inp.read("/input/potential/rama_map_pot/rama_pot"),
        frame_pos=run.read("/output/pos")[:, 0].astype(np.float32),
    )
    for k in ["potential", "rotamer_free_energy", "rotamer_potential", "rotamer_entropy", "hbond",
              "rama_map_potential", "sigmoid_coupling", "rotamer_1body_energy0", "rotamer_1body_energy1",
              "rotamer_1body_energy2", "kinetic", "time"]:
        out["frame_" + k] = run.read("/output/" + k)
    np.savez_compressed(os.path.join(HERE, "chig_frames.npz"), **out)
    print({k: v.shape for k, v in out.items()})


if __name__ == "__main__":
    main()
