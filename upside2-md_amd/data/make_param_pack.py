"""Extract the ff_2.1 parameter arrays the `.up` writer needs into one .npz.

Run where the reference tree is present:
    python upside2-md_amd/data/make_param_pack.py /root/reference/parameters
It reads (never executes) the reference's parameter data files:
    ff_2.1/sidechain.h5, ff_2.1/environment.h5, ff_2.1/hbond.h5,
    ff_2.1/bb_env.dat, ff_2.1/sheet, common/rama_reference.pkl
The pickle holds one (1,72,72) float64 array; its raw bytes are taken from the
opcode stream with pickletools (no unpickling, nothing from the file is run).
"""
import os
import pickletools
import sys

import numpy as np

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), ".."))
import h5lite  # noqa: E402


def rama_reference_from_pickle(path):
    data = open(path, "rb").read()
    shape, payload = None, None
    for op, arg, _ in pickletools.genops(data):
        if op.name in ("BINSTRING", "SHORT_BINSTRING", "BINBYTES", "SHORT_BINBYTES") and arg is not None:
            raw = arg.encode("latin1") if isinstance(arg, str) else arg
            if len(raw) == 72 * 72 * 8:
                payload = raw
    if payload is None:
        raise ValueError("payload of 72*72 float64 not found in %s" % path)
    shape = (72, 72)  # pickled state tuple: (version=1, shape=(72,72), dtype f8, ...)
    return np.frombuffer(payload, dtype="<f8").reshape(shape).copy()


def main(param_root, out):
    ff = os.path.join(param_root, "ff_2.1")
    pack = {}
    with h5lite.File(os.path.join(ff, "sidechain.h5")) as f:
        for k in ["restype_order", "bead_order", "coverage_interaction", "hydrophobe_interaction",
                  "hydrophobe_placement", "pair_interaction", "rotamer_center_fixed", "rotamer_prob",
                  "rotamer_start_stop_bead"]:
            pack["sidechain/" + k] = f.read("/" + k)
    with h5lite.File(os.path.join(ff, "environment.h5")) as f:
        for k in ["center", "coverage_param", "restype_order", "scale", "sharpness", "weights"]:
            pack["environment/" + k] = f.read("/" + k)
    with h5lite.File(os.path.join(ff, "hbond.h5")) as f:
        pack["hbond/parameter"] = f.read("/parameter")
    pack["bb_env"] = np.loadtxt(os.path.join(ff, "bb_env.dat"))
    pack["sheet"] = np.loadtxt(os.path.join(ff, "sheet"))
    pack["rama_reference"] = rama_reference_from_pickle(os.path.join(param_root, "common", "rama_reference.pkl"))
    np.savez_compressed(out, **pack)
    print("wrote", out, {k: v.shape for k, v in pack.items()})


if __name__ == "__main__":
    root = sys.argv[1] if len(sys.argv) > 1 else "/root/reference/parameters"
    main(root, os.path.join(os.path.dirname(os.path.abspath(__file__)), "ff21_params.npz"))
