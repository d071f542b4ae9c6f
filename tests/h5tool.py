"""HDF5 side of the run_sim_json tests, run by an interpreter that has h5py
(the image's /opt/conda/bin/python3.9); the test interpreter has none.

    h5tool.py write SRC.npz DST.h5   every array becomes a root dataset, as the
                                     front end's dump() writes the input file
                                     (ffip/geom.py:432-433, ffip/source.py:127-129)
    h5tool.py read SRC.h5 DST.npz    every dataset, keyed by its path without the
                                     leading "/", plus "__shape__/<path>" entries
"""
import sys

import h5py
import numpy as np


def write(src, dst):
    d = np.load(src)
    with h5py.File(dst, "w") as f:
        for k in d.files:
            f.create_dataset(k, data=d[k])


def read(src, dst):
    out = {}

    def visit(name, obj):
        if isinstance(obj, h5py.Dataset):
            a = obj[()]
            out[name] = np.asarray(a)
            out["__shape__/" + name] = np.asarray(obj.shape, dtype=np.int64)
            out["__dtype__/" + name] = np.asarray(str(obj.dtype))

    with h5py.File(src, "r") as f:
        f.visititems(visit)
    np.savez(dst, **out)


if __name__ == "__main__":
    {"write": write, "read": read}[sys.argv[1]](sys.argv[2], sys.argv[3])
