"""CPU tests of the product's host layer (no kernel is launched).

* libseiswg_hip.so loads and exports every symbol include/*.h declares;
* the host numerics equal the oracle's restatement bit for bit: Ricker table
  (solver.rs:350-356), grid inference (solver.rs:640-658), grid-index
  rounding (solver.rs:716-718);
* the CLI's failure classes that occur before any device work (main.rs:38-42,
  config.rs:67-70, 78-83, solver.rs:306-325, 614-623) end with a non-zero exit
  and the reference's message.
"""
import os
import re
import subprocess

import numpy as np
import pytest

import examples as E
from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for h in os.listdir(os.path.join(ROOT, "include")):
        if h.endswith(".h"):
            txt = open(os.path.join(ROOT, "include", h)).read()
            names |= set(re.findall(r"^SW_API\s+[\w\s\*]*?\b(\w+)\s*\(", txt, re.M))
    return names


def test_library_exports_every_declared_symbol(seiswg):
    names = declared_symbols()
    assert len(names) >= 20
    lib = seiswg.load()
    for n in names:
        assert hasattr(lib, n), n
    assert names == set(seiswg.EXPORTS)
    out = subprocess.run(["nm", "-D", "--defined-only", seiswg.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r" T (\w+)", out))
    assert names <= exported


def test_ricker_table_equals_oracle_bitwise(seiswg):
    rows = [[0, 0, 0, 2.0, 0.6, 0.0, 1e10], [0, 0, 0, 25.0, 0.05, 45.0, 1e10], [0, 0, 0, 8, 0.15, 90, 1e10],
            [0, 0, 0, 0.02, 0.0, 0.0, 1e10], [0, 0, 0, 50.0, 0.05, 0.0, 1e10]]
    for nt, dt in [(2000, 0.006), (3000, 4e-4), (4800, 0.06)]:
        params = np.array([[r[3], r[4], r[5], r[6]] for r in rows], np.float32)
        got = seiswg.ricker_table(params, nt, dt)
        want = O.ricker_table(4, rows, nt, dt).astype(np.float32)
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


@pytest.mark.parametrize("shape", [(200, 200, 100.0), (201, 201, 2400.0), (200, 200, 1.0), (1000, 600, 10.0),
                                   (500, 500, 5.0)])
def test_grid_inference_equals_oracle(seiswg, shape):
    nx, nz, h = shape
    i, j = np.meshgrid(np.arange(nx), np.arange(nz), indexing="ij")
    x, z = (i * h).astype(np.float32).ravel(), (j * h).astype(np.float32).ravel()
    got = seiswg.infer_grid(x, z)
    want = O.infer_grid(x, z)
    assert got[:4] == want and not got[4]


def test_grid_inference_exact_fallback_for_mis_factored_shapes(seiswg):
    nx, nz, h = 2000, 1000, 10.0
    i, j = np.meshgrid(np.arange(nx), np.arange(nz), indexing="ij")
    x, z = (i * h).astype(np.float32).ravel(), (j * h).astype(np.float32).ravel()
    assert O.infer_grid(x, z)[:2] == (2001, 1000)       # the reference's rule
    got = seiswg.infer_grid(x, z)
    assert got[:2] == (2000, 1000) and got[4] and got[2] == np.float32(10.0)


def test_grid_index_rounding_equals_oracle(seiswg):
    rng = np.random.default_rng(0)
    dx, dz, nz = np.float32(100.0), np.float32(37.5), 321
    xs = list(rng.uniform(-500, 20000, 200)) + [250.0, 350.0, 50.0, -49.0, -51.0, 0.0, 149.99999]
    zs = list(rng.uniform(-100, 12000, 200)) + [18.75, 56.25, 93.75, 0.0, -18.75, 1.0, 2.0]
    for x, z in zip(xs, zs):
        assert seiswg.grid_index(x, z, dx, dz, nz) == O.grid_index(x, z, dx, dz, nz), (x, z)


def run_cli(args, cwd=None):
    import seiswg_amd
    env = dict(os.environ, RUST_LOG="error")
    return subprocess.run([seiswg_amd.CLI_PATH] + args, capture_output=True, text=True, cwd=cwd, env=env,
                          timeout=60)


def write_ini(path, body):
    with open(path, "w") as f:
        f.write(body)
    return str(path)


BASE = """[workflow]
mode = forward
[solver]
sh = yes
nt = 10
dt = 0.01
[path]
output = out
model_init = model
sources = sources.dat
stations = stations.dat
"""


def test_cli_usage_without_argument(seiswg):
    r = run_cli([])
    assert r.returncode == 1 and "Usage:" in r.stderr


@pytest.mark.parametrize("body,msg", [
    (BASE.replace("mode = forward\n", ""), "[workflow] mode is required"),
    (BASE.replace("nt = 10", "nt = ten"), "[solver] nt must be an integer"),
    (BASE.replace("nt = 10", "nt = -3"), "[solver] nt must be an integer"),
    (BASE.replace("dt = 0.01\n", ""), "[solver] dt is required"),
    (BASE.replace("dt = 0.01", "dt = fast"), "[solver] dt must be a float"),
    (BASE.replace("sources = sources.dat\n", ""), "[path] sources is required"),
    (BASE.replace("model_init = model\n", ""), "No model path in config"),
    (BASE, "Cannot read model file"),
])
def test_cli_failure_classes(seiswg, tmp_path, body, msg):
    ini = write_ini(tmp_path / "config.ini", body)
    r = run_cli([ini])
    assert r.returncode != 0
    assert msg in r.stderr, r.stderr


def test_cli_missing_config_file(seiswg, tmp_path):
    r = run_cli([str(tmp_path / "nope.ini")])
    assert r.returncode != 0 and "Cannot open" in r.stderr


def test_cli_truncated_model_file(seiswg, tmp_path):
    d = E.materialise("forward", str(tmp_path))
    p = os.path.join(d, "model", "proc000000_vp.bin")
    b = open(p, "rb").read()
    open(p, "wb").write(b[:1000])
    r = run_cli([os.path.join(d, "config.ini")])
    assert r.returncode != 0 and "too short" in r.stderr


def test_ini_parsing_rules_match_config_rs(seiswg, tmp_path):
    """Comments, case-folded keys/sections, ':' delimiter and 'yes'-only
    booleans; the run then stops at the model files (no device needed)."""
    body = """; leading comment
[WORKFLOW]
MODE : Forward   # inline comment
[Solver]
SH = YES
nt=5
dt = 1e-3
abs_left = true
[path]
output = o
model_true = mdl
sources = s.dat
stations = r.dat
"""
    r = run_cli([write_ini(tmp_path / "c.ini", body)])
    assert r.returncode != 0 and "Cannot read model file" in r.stderr and "mdl" in r.stderr


def test_vfs_map_round_trip(seiswg):
    v = seiswg.Vfs({"b/x.bin": b"\x00\x01\x02", "a.txt": b"hello"})
    v["c"] = b""
    v["a.txt"] = b"again\x00with nul"
    assert v.keys() == ["a.txt", "b/x.bin", "c"]
    assert v["a.txt"] == b"again\x00with nul" and v["b/x.bin"] == b"\x00\x01\x02" and v["c"] == b""
    with pytest.raises(KeyError):
        v["missing"]
    v.close()


@pytest.mark.parametrize("body,msg", [
    (BASE.replace("mode = forward\n", ""), "[workflow] mode is required"),
    (BASE.replace("nt = 10", "nt = ten"), "[solver] nt must be an integer"),
    (BASE.replace("sources = sources.dat\n", ""), "[path] sources is required"),
    (BASE.replace("model_init = model\n", ""), "No model path in config"),
    (BASE, "Cannot read model file model/proc000000_rho.bin"),
])
def test_in_memory_run_failure_classes(seiswg, body, msg):
    """seiswg_run_simulation fails as the CLI does, before any device work,
    and looks for nothing on disk (no model files are given)."""
    with pytest.raises(seiswg.SeiswgError) as e:
        seiswg.run_simulation(body, "0 0\n", "0 0\n", {})
    assert msg in str(e.value)


def test_in_memory_run_reads_model_files_from_the_map(seiswg, tmp_path):
    """Model files come from the map under the path the config resolves them
    to, also when the config writes it with a leading "./": a truncated vp
    file in the map stops the run at the same check as the CLI's."""
    d = E.materialise("forward", str(tmp_path))
    files = {f"model/{n}": open(os.path.join(d, "model", n), "rb").read() for n in os.listdir(os.path.join(d, "model"))}
    files["model/proc000000_vp.bin"] = files["model/proc000000_vp.bin"][:1000]
    ini = open(os.path.join(d, "config.ini")).read()
    tables = [open(os.path.join(d, n)).read() for n in ("sources.dat", "stations.dat")]
    for text in (ini, ini.replace("model_init    = model", "model_init    = ./model")):
        with pytest.raises(seiswg.SeiswgError) as e:
            seiswg.run_simulation(text, *tables, files)
        assert "too short" in str(e.value) and "proc000000_vp.bin" in str(e.value)


def test_vfs_c_abi_edges(seiswg):
    """The map's C entry points on their edges: an absent key reads -1, an
    index outside the map gives NULL, a negative size and a NULL map are
    refused, and the in-memory run needs config text and an output map."""
    import ctypes as C
    lib = seiswg.load()
    v = lib.seiswg_vfs_new()
    try:
        p = C.c_void_p()
        assert lib.seiswg_vfs_get(v, b"none", C.byref(p)) == -1
        assert lib.seiswg_vfs_put(v, b"k", b"abc", 3) == 0
        assert lib.seiswg_vfs_get(v, b"k", None) == 3
        assert lib.seiswg_vfs_count(v) == 1 and lib.seiswg_vfs_key(v, 0) == b"k"
        assert lib.seiswg_vfs_key(v, 1) is None and lib.seiswg_vfs_key(v, -1) is None
        assert lib.seiswg_vfs_put(v, b"k", b"abc", -1) != 0
        assert lib.seiswg_vfs_put(None, b"k", b"abc", 3) != 0
        assert lib.seiswg_vfs_count(None) == 0
        assert lib.seiswg_run_simulation(None, None, None, None, None, v) != 0
        assert b"config text" in lib.sw_last_error()
        assert lib.seiswg_run_simulation(b"[workflow]\nmode = forward\n", None, None, None, None, None) != 0
    finally:
        lib.seiswg_vfs_free(v)
