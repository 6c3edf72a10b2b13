"""Builds libseiswg_hip.so (gfx950 kernels + C ABI + C++ host) and the
`seiswg` CLI in-tree with hipcc.

    python -m seiswg_amd.build          # or seiswg_amd.build.build()
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
OBJ = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libseiswg_hip.so")
CLI = os.path.join(PKG, "bin", "seiswg")
ARCH = os.environ.get("SEISWG_OFFLOAD_ARCH", "gfx950")

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
# -ffp-contract=off: keep the reference's operation order (no FMA formed).
# packed-fp32-ops off: no v_pk_*_f32 generated from scalar float code.
COMMON = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall",
          f"-I{os.path.join(ROOT, 'include')}"]
DEVICE = [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]

SOURCES = [
    ("kernels.hip", True),
    ("capi.cpp", True),
    (os.path.join("host", "seiswg_host.cpp"), False),
]


def _compile(src: str, device: bool) -> str:
    out = os.path.join(OBJ, os.path.basename(src) + ".o")
    cmd = [HIPCC] + COMMON
    if device:
        cmd += DEVICE + ["-x", "hip"]
    else:
        cmd += ["-x", "c++"]
    cmd += ["-c", os.path.join(CSRC, src), "-o", out]
    subprocess.run(cmd, check=True)
    return out


def build(verbose: bool = False) -> str:
    os.makedirs(OBJ, exist_ok=True)
    os.makedirs(os.path.dirname(CLI), exist_ok=True)
    with ThreadPoolExecutor(max_workers=len(SOURCES)) as ex:
        objs = list(ex.map(lambda s: _compile(*s), SOURCES))
    tmp = LIB + ".tmp"
    subprocess.run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp] + objs, check=True)
    os.replace(tmp, LIB)
    main_o = os.path.join(OBJ, "main.o")
    subprocess.run([HIPCC] + COMMON + ["-x", "c++", "-c", os.path.join(CSRC, "host", "main.cpp"), "-o", main_o],
                   check=True)
    tmp = CLI + ".tmp"
    subprocess.run([HIPCC, "-o", tmp, main_o, f"-L{PKG}", "-lseiswg_hip", "-Wl,-rpath,$ORIGIN/.."], check=True)
    os.replace(tmp, CLI)
    if verbose:
        print("built", LIB, CLI)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
