"""Copies the reference's committed test data into small fixtures.

Run once in a container where the reference tree is mounted:
    python tests/golden/make_golden.py /root/reference

Only data is copied: the example inputs (INI files, source/station tables,
proc000000_*.bin model files stored losslessly in .npz), the golden outputs
of the reference's integration tests (tests/fixtures/**, byte copies) and a
time-decimated subset of examples/adjoint/obs_traces (every 16th sample of
the reference's forward traces on model_true).  No reference source file is
copied.
"""
import os
import shutil
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def model_npz(src_dir, dst):
    arrs = {}
    for f in sorted(os.listdir(src_dir)):
        if f.endswith(".bin"):
            b = open(os.path.join(src_dir, f), "rb").read()
            arrs[f[:-4]] = np.frombuffer(b, np.uint8)   # exact bytes, header included
    np.savez_compressed(dst, **arrs)


def main(ref):
    ex_out = os.path.join(HERE, "examples")
    for ex in ("forward", "spin", "adjoint"):
        s = os.path.join(ref, "examples", ex)
        d = os.path.join(ex_out, ex)
        os.makedirs(d, exist_ok=True)
        for f in os.listdir(s):
            p = os.path.join(s, f)
            if f.endswith((".ini", ".dat")):
                shutil.copyfile(p, os.path.join(d, f))
            elif os.path.isdir(p) and f.startswith("model"):
                model_npz(p, os.path.join(d, f + ".npz"))
        fx = os.path.join(ref, "tests", "fixtures", ex)
        fd = os.path.join(HERE, "fixtures", ex)
        os.makedirs(fd, exist_ok=True)
        for f in os.listdir(fx):
            shutil.copyfile(os.path.join(fx, f), os.path.join(fd, f))
    obs_dir = os.path.join(ref, "examples", "adjoint", "obs_traces")
    sub = {}
    for i in range(12):
        a = np.load(os.path.join(obs_dir, f"uy_{i:06d}.npy"), allow_pickle=False)
        sub[f"uy_{i:06d}"] = np.ascontiguousarray(a[:, ::16])
    np.savez_compressed(os.path.join(HERE, "adjoint_obs_traces_every16.npz"), **sub)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
