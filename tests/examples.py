"""Helpers that materialise the reference's example inputs (kept as
fixtures under tests/golden/) into a scratch directory, and read the
reference's golden outputs."""
from __future__ import annotations

import configparser
import os
import shutil

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def materialise(example: str, dst: str) -> str:
    """Writes examples/<example>/ (INI, tables, model dirs) under dst and
    returns the example directory."""
    src = os.path.join(GOLDEN, "examples", example)
    d = os.path.join(dst, example)
    os.makedirs(d, exist_ok=True)
    for f in os.listdir(src):
        p = os.path.join(src, f)
        if f.endswith(".npz"):
            md = os.path.join(d, f[:-4])
            os.makedirs(md, exist_ok=True)
            with np.load(p, allow_pickle=False) as z:
                for k in z.files:
                    z[k].tofile(os.path.join(md, k + ".bin"))
        else:
            shutil.copyfile(p, os.path.join(d, f))
    return d


def fixture(example: str, name: str) -> np.ndarray:
    p = os.path.join(GOLDEN, "fixtures", example, name)
    if name.endswith(".npy"):
        return np.load(p, allow_pickle=False)
    return np.fromfile(p, "<f4")


def read_npy_payload(path: str) -> np.ndarray:
    """Payload of a v1 .npy, as the reference's test reads it (integration.rs:117-150)."""
    b = open(path, "rb").read()
    assert b[:6] == b"\x93NUMPY"
    hl = int.from_bytes(b[8:10], "little")
    return np.frombuffer(b[10 + hl:], "<f4")


def obs_traces_every16() -> dict:
    with np.load(os.path.join(GOLDEN, "adjoint_obs_traces_every16.npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def ini_values(path: str) -> dict:
    """[solver] values of an example INI as the oracle's load_problem wants them
    (booleans are true only for 'yes', config.rs:72-76)."""
    cp = configparser.ConfigParser()
    cp.read(path)
    s = cp["solver"]
    yes = lambda k, d: s.get(k, "yes" if d else "no").strip().lower() == "yes"
    return dict(nt=int(s["nt"]), dt=float(np.float32(float(s["dt"]))), sh=yes("sh", False),
                psv=yes("psv", True), spin=yes("spin", False), abs_left=yes("abs_left", False),
                abs_right=yes("abs_right", False), abs_bottom=yes("abs_bottom", False),
                abs_top=yes("abs_top", False), abs_width=int(s.get("abs_width", "20")),
                abs_alpha=float(np.float32(float(s.get("abs_alpha", "0.015")))),
                save_snapshot=int(s.get("save_snapshot", "0")),
                combine=yes("combine_sources", False),
                adj_interval=int(s.get("adj_interval", "10")),
                smooth=float(np.float32(float(s.get("smooth", "5.0")))),
                paths=dict(cp["path"]))
