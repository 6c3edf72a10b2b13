"""ctypes binding of libseiswg_hip.so (include/seiswg_hip.h, seiswg_host.h).

Thin plumbing for tests and benchmarks: every call goes through the C ABI
into the gfx950 kernels.  There is no CPU fallback: if the library is not
built, loading raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libseiswg_hip.so")
CLI_PATH = os.path.join(PKG, "bin", "seiswg")

# field ids (fd2d.wgsl:53-105)
F_VY, F_SXY, F_SZY, F_VX, F_VZ = 0, 2, 3, 7, 8
F_SXX, F_SZZ, F_SXZ = 11, 12, 13
F_VY_C, F_SYX_C, F_SYY_C, F_SYZ_C = 20, 22, 23, 24
F_LAM, F_MU, F_NU, F_J, F_LAM_C, F_MU_C, F_NU_C, F_RHO, F_BOUND = 30, 31, 32, 33, 34, 35, 36, 37, 38
F_K_LAM, F_K_MU, F_K_RHO = 39, 40, 41
OBS_X, OBS_Y, OBS_Z, OBS_YI, OBS_YC = range(5)
MODEL_VPVS, MODEL_LAME = 0, 1
DIV_EXACT, DIV_RECIPROCAL = 0, 1


class SeiswgError(RuntimeError):
    pass


class sw_grid(C.Structure):
    _fields_ = [("nx", C.c_int32), ("nz", C.c_int32), ("nt", C.c_int32),
                ("dx", C.c_float), ("dz", C.c_float), ("dt", C.c_float),
                ("sh", C.c_int32), ("psv", C.c_int32), ("spin", C.c_int32),
                ("abs_left", C.c_int32), ("abs_right", C.c_int32), ("abs_bottom", C.c_int32),
                ("abs_top", C.c_int32), ("abs_width", C.c_int32), ("abs_alpha", C.c_float)]


class sw_timing(C.Structure):
    _fields_ = [("loop_ms", C.c_double), ("steps", C.c_int64), ("step_launches", C.c_int64),
                ("cell_updates", C.c_int64)]


SNAP_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_int32, C.c_int32, C.c_int32, C.POINTER(C.c_float), C.c_int64)

# every symbol the headers declare (checked by the CPU test suite)
EXPORTS = ["sw_last_error", "sw_create", "sw_destroy", "sw_set_model", "sw_set_sources", "sw_set_receivers",
           "sw_forward", "sw_forward_batch", "sw_read_obs", "sw_read_field", "sw_write_field", "sw_prepare",
           "sw_advance", "sw_adjoint_plain", "sw_smooth", "sw_set_tiling", "sw_get_tiling", "sw_get_timing",
           "sw_set_steps_per_launch", "sw_set_tiling2", "sw_get_tiling2", "sw_release_cached_memory",
           "sw_bench_copy", "sw_gradient", "sw_autotune", "sw_multi_create", "sw_multi_destroy",
           "sw_multi_forward", "sw_multi_gradient", "sw_multi_timing", "sw_multi_set_merge_limit",
           "sw_multi_set_workers", "sw_set_finite_check", "sw_get_graph_state", "sw_smooth_noalloc",
           "sw_set_division",
           "seiswg_run", "seiswg_ricker_table", "seiswg_infer_grid", "seiswg_grid_index",
           "seiswg_vfs_new", "seiswg_vfs_free", "seiswg_vfs_put", "seiswg_vfs_get", "seiswg_vfs_count",
           "seiswg_vfs_key", "seiswg_run_simulation"]

_lib = None


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise SeiswgError(f"{LIB_PATH} is not built; run `python -m seiswg_amd.build`")
    L = C.CDLL(LIB_PATH)
    vp, ip, fp, dp = C.c_void_p, C.POINTER(C.c_int32), C.POINTER(C.c_float), C.POINTER(C.c_double)
    L.sw_last_error.restype = C.c_char_p
    L.sw_create.argtypes = [C.POINTER(sw_grid), C.c_int32, C.POINTER(C.c_void_p)]
    L.sw_destroy.argtypes = [vp]
    L.sw_destroy.restype = None
    L.sw_set_model.argtypes = [vp, C.c_int32, C.POINTER(fp)]
    L.sw_set_sources.argtypes = [vp, C.c_int32, ip, fp]
    L.sw_set_receivers.argtypes = [vp, C.c_int32, ip]
    L.sw_forward.argtypes = [vp, C.c_int32, C.c_int32, SNAP_FN, vp]
    L.sw_forward_batch.argtypes = [vp, C.c_int32, ip, C.c_int32, SNAP_FN, vp]
    L.sw_read_obs.argtypes = [vp, C.c_int32, C.c_int32, fp]
    L.sw_read_field.argtypes = [vp, C.c_int32, C.c_int32, fp]
    L.sw_write_field.argtypes = [vp, C.c_int32, C.c_int32, fp]
    L.sw_prepare.argtypes = [vp, C.c_int32, ip]
    L.sw_advance.argtypes = [vp, C.c_int32]
    L.sw_adjoint_plain.argtypes = [vp, C.c_int32, ip, fp, C.c_int32, C.c_float, dp]
    L.sw_smooth.argtypes = [vp, C.c_int32, C.c_float]
    L.sw_smooth_noalloc.argtypes = [vp, C.c_int32, C.c_float]
    L.sw_set_tiling.argtypes = [vp, C.c_int32, C.c_int32, C.c_int32]
    L.sw_set_steps_per_launch.argtypes = [vp, C.c_int32]
    L.sw_set_tiling2.argtypes = [vp, C.c_int32, C.c_int32, C.c_int32]
    L.sw_release_cached_memory.argtypes = [C.c_int32, C.POINTER(C.c_int64)]
    L.sw_get_tiling.argtypes = [vp, ip, ip, ip]
    L.sw_get_timing.argtypes = [vp, C.POINTER(sw_timing)]
    L.sw_multi_create.argtypes = [C.POINTER(C.c_void_p), C.c_int32, C.POINTER(C.c_void_p)]
    L.sw_multi_destroy.argtypes = [vp]
    L.sw_multi_destroy.restype = None
    L.sw_multi_forward.argtypes = [vp, ip, C.POINTER(ip)]
    L.sw_multi_gradient.argtypes = [vp, ip, C.POINTER(ip), C.POINTER(fp), C.c_int32, C.POINTER(dp)]
    L.sw_multi_timing.argtypes = [vp, dp, C.POINTER(C.c_int64)]
    L.sw_multi_set_merge_limit.argtypes = [vp, C.c_int64]
    L.sw_multi_set_workers.argtypes = [vp, C.c_int32]
    L.sw_autotune.argtypes = [vp, C.c_int32, C.c_int32, ip, ip, ip, dp]
    L.sw_gradient.argtypes = [vp, C.c_int32, ip, fp, C.c_int32, C.c_int32, dp]
    L.sw_bench_copy.argtypes = [C.c_int32, C.c_int64, C.c_int32, dp]
    L.sw_set_finite_check.argtypes = [vp, C.c_int32]
    L.sw_set_division.argtypes = [vp, C.c_int32]
    L.sw_get_graph_state.argtypes = [vp, ip, ip, C.c_char_p, C.c_int32]
    L.seiswg_run.argtypes = [C.c_char_p]
    L.seiswg_ricker_table.argtypes = [C.c_int32, C.c_int32, C.c_float, fp, fp]
    L.seiswg_infer_grid.argtypes = [C.c_int64, fp, fp, ip, ip, fp, fp, ip]
    L.seiswg_grid_index.argtypes = [C.c_float, C.c_float, C.c_float, C.c_float, C.c_uint32]
    L.seiswg_grid_index.restype = C.c_uint32
    L.seiswg_vfs_new.argtypes = []
    L.seiswg_vfs_new.restype = vp
    L.seiswg_vfs_free.argtypes = [vp]
    L.seiswg_vfs_free.restype = None
    L.seiswg_vfs_put.argtypes = [vp, C.c_char_p, C.c_char_p, C.c_int64]
    L.seiswg_vfs_get.argtypes = [vp, C.c_char_p, C.POINTER(C.c_void_p)]
    L.seiswg_vfs_get.restype = C.c_int64
    L.seiswg_vfs_count.argtypes = [vp]
    L.seiswg_vfs_key.argtypes = [vp, C.c_int32]
    L.seiswg_vfs_key.restype = C.c_char_p
    L.seiswg_run_simulation.argtypes = [C.c_char_p, C.c_char_p, C.c_char_p, C.c_char_p, vp, vp]
    for n in EXPORTS:
        if n not in ("sw_destroy", "sw_last_error", "seiswg_grid_index", "sw_multi_destroy", "seiswg_vfs_new",
                     "seiswg_vfs_free", "seiswg_vfs_get", "seiswg_vfs_key"):
            getattr(L, n).restype = C.c_int
    _lib = L
    return L


def _check(rc: int):
    if rc != 0:
        raise SeiswgError(load().sw_last_error().decode(errors="replace"))


def _f32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float32)


def _fptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _iptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_int32))


# ------------------------------------------------------------ host helpers
def ricker_table(src_params, nt: int, dt: float) -> np.ndarray:
    """[3][nsrc][nt] from rows (f0, t0, ang_deg, amp)."""
    p = _f32(np.asarray(src_params, np.float32).reshape(-1, 4))
    out = np.zeros((3, p.shape[0], nt), np.float32)
    _check(load().seiswg_ricker_table(p.shape[0], nt, C.c_float(dt), _fptr(p), _fptr(out)))
    return out


def infer_grid(x, z):
    x, z = _f32(x), _f32(z)
    nx, nz, fb = C.c_int32(), C.c_int32(), C.c_int32()
    dx, dz = C.c_float(), C.c_float()
    rc = load().seiswg_infer_grid(len(x), _fptr(x), _fptr(z), C.byref(nx), C.byref(nz), C.byref(dx), C.byref(dz),
                                  C.byref(fb))
    if rc != 0:
        raise SeiswgError(f"grid inference failed ({rc})")
    return nx.value, nz.value, np.float32(dx.value), np.float32(dz.value), bool(fb.value)


def grid_index(x: float, z: float, dx: float, dz: float, nz: int) -> int:
    return int(load().seiswg_grid_index(C.c_float(x), C.c_float(z), C.c_float(dx), C.c_float(dz), nz))


def bench_copy(nbytes: int, iters: int = 20, device: int = 0) -> float:
    """Average ms of a plain copy kernel moving nbytes in + nbytes out."""
    ms = C.c_double()
    _check(load().sw_bench_copy(device, nbytes, iters, C.byref(ms)))
    return ms.value


def release_cached_memory(device: int = -1) -> int:
    """Return the library's cached device memory to the runtime; bytes freed."""
    n = C.c_int64()
    _check(load().sw_release_cached_memory(device, C.byref(n)))
    return n.value


def run_config(path: str) -> int:
    """In-process `seiswg <config.ini>`."""
    return load().seiswg_run(path.encode())


class Vfs:
    """A path -> bytes map of the library (sw_vfs)."""

    def __init__(self, files: dict | None = None):
        L = load()
        self._L = L
        self.h = L.seiswg_vfs_new()
        if not self.h:
            raise SeiswgError("seiswg_vfs_new failed")
        for k, v in (files or {}).items():
            self[k] = v

    def __setitem__(self, key: str, data: bytes):
        data = bytes(data)
        _check(self._L.seiswg_vfs_put(self.h, key.encode(), data, len(data)))

    def __getitem__(self, key: str) -> bytes:
        p = C.c_void_p()
        n = self._L.seiswg_vfs_get(self.h, key.encode(), C.byref(p))
        if n < 0:
            raise KeyError(key)
        return C.string_at(p, n) if n else b""

    def keys(self) -> list:
        return [self._L.seiswg_vfs_key(self.h, i).decode() for i in range(self._L.seiswg_vfs_count(self.h))]

    def to_dict(self) -> dict:
        return {k: self[k] for k in self.keys()}

    def close(self):
        if self.h:
            self._L.seiswg_vfs_free(self.h)
            self.h = None

    def __del__(self):
        self.close()


def run_simulation(config_ini: str, sources_dat: str | None, stations_dat: str | None, files: dict,
                   config_true_ini: str | None = None) -> dict:
    """In-memory workflow (lib.rs:74-180): model files (and observed traces)
    by path in `files`, outputs returned as a dict path -> bytes."""
    L = load()
    fin, fout = Vfs(files), Vfs()
    try:
        enc = lambda t: None if t is None else t.encode()
        _check(L.seiswg_run_simulation(config_ini.encode(), enc(sources_dat), enc(stations_dat),
                                       enc(config_true_ini), fin.h, fout.h))
        return fout.to_dict()
    finally:
        fin.close()
        fout.close()


# ------------------------------------------------------------ device context
class Context:
    """One device problem (sw_ctx)."""

    def __init__(self, nx, nz, nt, dx, dz, dt, sh=False, psv=True, spin=False, abs_left=False,
                 abs_right=False, abs_bottom=False, abs_top=False, abs_width=20, abs_alpha=0.015, device=0):
        L = load()
        self.g = sw_grid(nx, nz, nt, dx, dz, dt, int(sh), int(psv), int(spin), int(abs_left), int(abs_right),
                         int(abs_bottom), int(abs_top), abs_width, abs_alpha)
        h = C.c_void_p()
        _check(L.sw_create(C.byref(self.g), device, C.byref(h)))
        self.h = h
        self.L = L
        self.npt = nx * nz
        self.nrec = 0
        self.nshots = 0
        self._keep = []

    @classmethod
    def from_problem(cls, p, device=0):
        """From an oracle-style Problem description (same fields)."""
        c = cls(p.nx, p.nz, p.nt, p.dx, p.dz, p.dt, p.sh, p.psv, p.spin, p.abs_left, p.abs_right, p.abs_bottom,
                p.abs_top, p.abs_width, p.abs_alpha, device)
        m = p.model
        if p.spin:
            c.set_model(MODEL_LAME, [m[n] for n in ("rho", "lambda", "mu", "nu", "j", "lambda_c", "mu_c", "nu_c")])
        else:
            c.set_model(MODEL_VPVS, [m["rho"], m["vp"], m["vs"]])
        return c

    def close(self):
        if getattr(self, "h", None):
            self.L.sw_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_model(self, kind: int, fields):
        arrs = [_f32(np.asarray(f).reshape(-1)) for f in fields]
        for a in arrs:
            assert a.size == self.npt
        ptrs = (C.POINTER(C.c_float) * len(arrs))(*[_fptr(a) for a in arrs])
        _check(self.L.sw_set_model(self.h, kind, ptrs))

    def set_sources(self, grid_index, stf):
        idx = np.ascontiguousarray(np.asarray(grid_index, np.int64).astype(np.uint32).view(np.int32))
        stf = _f32(stf)
        _check(self.L.sw_set_sources(self.h, len(idx), _iptr(idx), _fptr(stf)))

    def set_receivers(self, grid_index):
        idx = np.ascontiguousarray(np.asarray(grid_index, np.int64).astype(np.uint32).view(np.int32))
        self.nrec = len(idx)
        _check(self.L.sw_set_receivers(self.h, len(idx), _iptr(idx)))

    def set_tiling(self, threads: int, seglen: int, shots_per_block: int = 0):
        _check(self.L.sw_set_tiling(self.h, threads, seglen, shots_per_block))

    def set_steps_per_launch(self, steps: int):
        """Forward steps per launch (2: launches of two steps, the default;
        3: P-SV alone in launches of three steps; 1: one step per launch).
        Same results bit for bit."""
        _check(self.L.sw_set_steps_per_launch(self.h, steps))

    def set_tiling2(self, threads: int, seglen: int, shots_per_block: int = 0):
        """Tile shape of the two-step P-SV launches."""
        _check(self.L.sw_set_tiling2(self.h, threads, seglen, shots_per_block))

    def tiling2(self):
        """(threads, seglen, shots per block) of the two-step launches of the
        last prepared batch; zeros when it runs one step per launch."""
        a, b, g = C.c_int32(), C.c_int32(), C.c_int32()
        _check(self.L.sw_get_tiling2(self.h, C.byref(a), C.byref(b), C.byref(g)))
        return a.value, b.value, g.value

    def autotune(self, nshots: int, steps: int = 10):
        """Fastest tile shape for nshots shots: (threads, seglen, shots per block, ms per step)."""
        a, b, g, ms = C.c_int32(), C.c_int32(), C.c_int32(), C.c_double()
        _check(self.L.sw_autotune(self.h, nshots, steps, C.byref(a), C.byref(b), C.byref(g), C.byref(ms)))
        return a.value, b.value, g.value, ms.value

    def tiling(self):
        """(threads, seglen, shots per block) of the last prepared batch."""
        a, b, g = C.c_int32(), C.c_int32(), C.c_int32()
        _check(self.L.sw_get_tiling(self.h, C.byref(a), C.byref(b), C.byref(g)))
        return a.value, b.value, g.value

    def forward(self, isrc: int = -1, snap: int = 0, snapshots: dict | None = None):
        return self.forward_batch([isrc], snap, snapshots)

    def forward_batch(self, isrc_list, snap: int = 0, snapshots: dict | None = None):
        """snapshots (optional dict) receives {(shot, it, field): array}."""
        isrc = np.ascontiguousarray(isrc_list, np.int32)

        def cb(user, shot, it, field, data, npt):
            if snapshots is not None:
                snapshots[(shot, it, field)] = np.ctypeslib.as_array(data, shape=(npt,)).copy()
            return 0

        fn = SNAP_FN(cb)
        _check(self.L.sw_forward_batch(self.h, len(isrc), _iptr(isrc), snap, fn, None))
        self.nshots = len(isrc)

    def prepare(self, isrc_list):
        isrc = np.ascontiguousarray(isrc_list, np.int32)
        _check(self.L.sw_prepare(self.h, len(isrc), _iptr(isrc)))
        self.nshots = len(isrc)

    def advance(self, nsteps: int):
        _check(self.L.sw_advance(self.h, nsteps))

    def obs(self, shot: int = 0, comp: int = OBS_Y) -> np.ndarray:
        out = np.zeros((self.nrec, self.g.nt), np.float32)
        _check(self.L.sw_read_obs(self.h, shot, comp, _fptr(out)))
        return out

    def field(self, fid: int, shot: int = 0) -> np.ndarray:
        out = np.zeros(self.npt, np.float32)
        _check(self.L.sw_read_field(self.h, shot, fid, _fptr(out)))
        return out

    def write_field(self, fid: int, data, shot: int = 0):
        a = _f32(np.asarray(data).reshape(-1))
        assert a.size == self.npt
        _check(self.L.sw_write_field(self.h, shot, fid, _fptr(a)))

    def adjoint_plain(self, isrc_list, obs_ref, adj_interval: int, smooth: float):
        isrc = np.ascontiguousarray(isrc_list, np.int32)
        ref = _f32(obs_ref)
        mis = np.zeros(max(len(isrc), 1), np.float64)
        _check(self.L.sw_adjoint_plain(self.h, len(isrc), _iptr(isrc), _fptr(ref), adj_interval, C.c_float(smooth),
                                       mis.ctypes.data_as(C.POINTER(C.c_double))))
        return self.field(F_K_MU), mis[:len(isrc)]

    def gradient(self, isrc_list, obs_ref, ckpt_every: int = 50, shots_per_launch: int = 0, smooth: float = 0.0):
        """Adjoint gradient; returns (misfit per shot, K_rho, K_mu, K_lambda),
        the kernels Gaussian-smoothed with sigma = smooth cells if smooth > 0."""
        isrc = np.ascontiguousarray(isrc_list, np.int32)
        ref = _f32(obs_ref)
        mis = np.zeros(len(isrc), np.float64)
        _check(self.L.sw_gradient(self.h, len(isrc), _iptr(isrc), _fptr(ref), ckpt_every, shots_per_launch,
                                  mis.ctypes.data_as(C.POINTER(C.c_double))))
        if smooth > 0:
            for f in (F_K_RHO, F_K_MU, F_K_LAM):
                self.smooth(f, smooth)
        return mis, self.field(F_K_RHO), self.field(F_K_MU), self.field(F_K_LAM)

    def smooth(self, fid: int, sigma: float):
        _check(self.L.sw_smooth(self.h, fid, C.c_float(sigma)))

    def smooth_noalloc(self, fid: int, sigma: float):
        """sw_smooth through the library's fixed device scratch (no allocation)."""
        _check(self.L.sw_smooth_noalloc(self.h, fid, C.c_float(sigma)))

    def timing(self) -> sw_timing:
        t = sw_timing()
        _check(self.L.sw_get_timing(self.h, C.byref(t)))
        return t

    def set_division(self, reciprocal: bool):
        """Forward-step divisions: exact (default; bit-identical to the FP32
        restatement) or by multiplication with the rounded reciprocal."""
        _check(self.L.sw_set_division(self.h, int(bool(reciprocal))))

    def set_finite_check(self, on: bool):
        """Blow-up detection (on by default): forward runs, gradients and the
        plain adjoint raise when the wavefield, misfits or kernels are not finite."""
        _check(self.L.sw_set_finite_check(self.h, int(bool(on))))

    def graph_state(self):
        """(state, captured blocks, message): state 1 graphs in use, 0 off,
        -1 off after a failed capture (message = the HIP error)."""
        st, n = C.c_int32(), C.c_int32()
        buf = C.create_string_buffer(256)
        _check(self.L.sw_get_graph_state(self.h, C.byref(st), C.byref(n), buf, 256))
        return st.value, n.value, buf.value.decode(errors="replace")


class Multi:
    """Several Contexts advanced together (sw_multi)."""

    def __init__(self, contexts):
        L = load()
        self.ctxs = list(contexts)
        arr = (C.c_void_p * len(self.ctxs))(*[c.h.value for c in self.ctxs])
        h = C.c_void_p()
        _check(L.sw_multi_create(arr, len(self.ctxs), C.byref(h)))
        self.h, self.L = h, L

    def close(self):
        if getattr(self, "h", None):
            self.L.sw_multi_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _isrc(self, isrc_lists):
        arrs = [np.ascontiguousarray(x, np.int32) for x in isrc_lists]
        n = np.ascontiguousarray([len(a) for a in arrs], np.int32)
        ptrs = (C.POINTER(C.c_int32) * len(arrs))(*[_iptr(a) for a in arrs])
        return arrs, n, ptrs

    def forward(self, isrc_lists):
        arrs, n, ptrs = self._isrc(isrc_lists)
        _check(self.L.sw_multi_forward(self.h, _iptr(n), ptrs))
        for c, a in zip(self.ctxs, arrs):
            c.nshots = len(a)

    def set_merge_limit(self, shot_cells: int):
        """Gradients: problems with at least this many shot-cells per step run
        on their own launches, smaller ones in merged launches (default 0:
        all on their own launches)."""
        _check(self.L.sw_multi_set_merge_limit(self.h, int(shot_cells)))

    def set_workers(self, workers: int):
        """Problems on their own launches run this many GPU-filling problems'
        worth at a time (smaller problems count for their share)."""
        _check(self.L.sw_multi_set_workers(self.h, workers))

    def gradient(self, isrc_lists, obs_refs, ckpt_every: int = 50):
        """Returns per problem (misfit per shot, K_rho, K_mu, K_lambda)."""
        arrs, n, ptrs = self._isrc(isrc_lists)
        refs = [_f32(r) for r in obs_refs]
        rp = (C.POINTER(C.c_float) * len(refs))(*[_fptr(r) for r in refs])
        mis = [np.zeros(len(a), np.float64) for a in arrs]
        mp = (C.POINTER(C.c_double) * len(mis))(*[m.ctypes.data_as(C.POINTER(C.c_double)) for m in mis])
        _check(self.L.sw_multi_gradient(self.h, _iptr(n), ptrs, rp, ckpt_every, mp))
        return [(m, c.field(F_K_RHO), c.field(F_K_MU), c.field(F_K_LAM)) for m, c in zip(mis, self.ctxs)]

    def timing(self):
        ms, n = C.c_double(), C.c_int64()
        _check(self.L.sw_multi_timing(self.h, C.byref(ms), C.byref(n)))
        return ms.value, n.value
