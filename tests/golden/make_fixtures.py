"""Pack the reference's Fortran-derived convection fixtures into one compressed .npz.

Source: jcm/data/test/{iptop,qdif,cbmf,precnv,dfse,dfqa}.npy in the reference
(used by jcm/physics/speedy/convection_test.py:38-46,123-140).  These are data
files the reference's tests hold (inputs are produced by the recipe at
convection_test.py:30-37).  Run once in a container that has the reference:

    python tests/golden/make_fixtures.py /root/reference
"""
import pathlib
import sys

import numpy as np


def main(ref_root):
    src = pathlib.Path(ref_root) / "jcm" / "data" / "test"
    arrays = {}
    for name in ("iptop", "qdif", "cbmf", "precnv", "dfse", "dfqa"):
        arrays[name] = np.load(src / f"{name}.npy", allow_pickle=False)
    out = pathlib.Path(__file__).with_name("convection_f90.npz")
    np.savez_compressed(out, **arrays)
    print("wrote", out, {k: (v.shape, str(v.dtype)) for k, v in arrays.items()})


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
