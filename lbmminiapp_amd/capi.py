"""ctypes binding of the C ABI declared in include/lbm_capi.h.

The HIP library must be built (``python -m lbmminiapp_amd.build``); there is no
CPU fallback: every class here raises if liblbm_hip.so is missing or a call
fails.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "liblbm_hip.so")

# enums (include/lbm_capi.h)
OK, ERR_ARG, ERR_HIP, ERR_OOM, ERR_UNSUPPORTED, ERR_STATE, ERR_NUMERIC = range(7)
PATTERN_AB, PATTERN_AA = 0, 1
LAYOUT_SOA, LAYOUT_AOS = 0, 1
STORE_F64, STORE_F32S, STORE_F16S, STORE_F32 = 0, 1, 2, 3

STORAGE_NAMES = {"f64": STORE_F64, "f32s": STORE_F32S, "f16s": STORE_F16S, "f32": STORE_F32}
LAYOUT_NAMES = {"soa": LAYOUT_SOA, "aos": LAYOUT_AOS}

EXPORTED = [
    "lbm_last_error", "lbm_version", "lbm_device_count", "lbm_set_device", "lbm_device_alloc",
    "lbm_device_free", "lbm_memcpy_h2d", "lbm_memcpy_d2h", "lbm_device_synchronize",
    "lbm_sparse_create", "lbm_sparse_destroy", "lbm_sparse_aa_even", "lbm_sparse_aa_odd",
    "lbm_sparse_ab", "lbm_sparse_init_rest", "lbm_sparse_mass", "lbm_sparse_mass_noalloc",
    "lbm_sparse_synchronize", "lbm_sparse_set_stream",
    "lbm_dense_create", "lbm_dense_destroy", "lbm_dense_get_info", "lbm_dense_set_state",
    "lbm_dense_get_state", "lbm_dense_init_equilibrium", "lbm_dense_init_tgv", "lbm_dense_step",
    "lbm_dense_run_timed", "lbm_dense_moments", "lbm_dense_mass", "lbm_dense_synchronize",
    "lbm_dense_set_stream", "lbm_dense_graph_create", "lbm_dense_graph_launch",
    "lbm_dense_graph_destroy",
    "lbm_geometry_cylinder", "lbm_geometry_channel", "lbm_geometry_add_spheres",
    "lbm_field_smooth_random", "lbm_copy_bandwidth",
    "lbm_adjoint_create", "lbm_adjoint_destroy", "lbm_adjoint_set_initial_state",
    "lbm_adjoint_forward", "lbm_adjoint_backward", "lbm_adjoint_get_lambda0",
    "lbm_adjoint_get_final_state", "lbm_adjoint_fd", "lbm_adjoint_fd_direction",
    "lbm_adjoint_device_bytes",
]


class LbmError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"lbm status {status}: {msg}")
        self.status = status


class _SparseConfig(ctypes.Structure):
    _fields_ = [("n_fluid", ctypes.c_int64), ("q", ctypes.c_int32), ("layout", ctypes.c_int32),
                ("storage", ctypes.c_int32), ("adjacency_mode", ctypes.c_int32),
                ("omega", ctypes.c_double), ("force", ctypes.c_double * 3),
                ("adjacency", ctypes.c_void_p)]


class _DenseConfig(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("nz", ctypes.c_int32),
                ("q", ctypes.c_int32), ("pattern", ctypes.c_int32), ("pull", ctypes.c_int32),
                ("layout", ctypes.c_int32), ("storage", ctypes.c_int32),
                ("periodic", ctypes.c_int32 * 3), ("omega", ctypes.c_double),
                ("force", ctypes.c_double * 3), ("solid", ctypes.c_void_p),
                ("tau", ctypes.c_void_p), ("gamma", ctypes.c_void_p),
                ("f16_scale", ctypes.c_double), ("density_jump_x", ctypes.c_double)]


class _AdjointConfig(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("nz", ctypes.c_int32),
                ("q", ctypes.c_int32), ("storage", ctypes.c_int32),
                ("periodic", ctypes.c_int32 * 3), ("omega", ctypes.c_double),
                ("force", ctypes.c_double * 3), ("solid", ctypes.c_void_p),
                ("tau", ctypes.c_void_p), ("gamma", ctypes.c_void_p),
                ("objective", ctypes.c_int32), ("outlet_x", ctypes.c_int32),
                ("checkpoint_every", ctypes.c_int32), ("checkpoint_inner", ctypes.c_int32 * 2),
                ("density_jump_x", ctypes.c_double)]


class _DenseInfo(ctypes.Structure):
    _fields_ = [("ncell", ctypes.c_int64), ("n_fluid", ctypes.c_int64),
                ("time_step", ctypes.c_int64), ("bytes_per_lup", ctypes.c_double),
                ("device_bytes", ctypes.c_size_t)]


_lib = None


def load(path: str = LIB_PATH):
    """Loads liblbm_hip.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"{path} not found: build it with `python -m lbmminiapp_amd.build`")
    L = ctypes.CDLL(path)
    P, I32, I64, D = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.c_double
    L.lbm_last_error.restype = ctypes.c_char_p
    L.lbm_version.restype = ctypes.c_char_p
    sig = {
        "lbm_device_count": [P], "lbm_set_device": [ctypes.c_int],
        "lbm_device_alloc": [ctypes.c_size_t, P], "lbm_device_free": [P],
        "lbm_memcpy_h2d": [P, P, ctypes.c_size_t], "lbm_memcpy_d2h": [P, P, ctypes.c_size_t],
        "lbm_device_synchronize": [],
        "lbm_sparse_create": [P, P], "lbm_sparse_destroy": [P],
        "lbm_sparse_aa_even": [P, P, I64, I64], "lbm_sparse_aa_odd": [P, P, I64, I64],
        "lbm_sparse_ab": [P, P, P, I64, I64, I32], "lbm_sparse_init_rest": [P, P, I32],
        "lbm_sparse_mass": [P, P, I64, I64, P], "lbm_sparse_mass_noalloc": [P, P, I64, I64, P],
        "lbm_sparse_synchronize": [P], "lbm_sparse_set_stream": [P, P],
        "lbm_dense_create": [P, P], "lbm_dense_destroy": [P], "lbm_dense_get_info": [P, P],
        "lbm_dense_set_state": [P, P], "lbm_dense_get_state": [P, P],
        "lbm_dense_init_equilibrium": [P, P, P], "lbm_dense_init_tgv": [P, D, D, ctypes.c_uint64],
        "lbm_dense_step": [P, I32], "lbm_dense_run_timed": [P, I32, P],
        "lbm_dense_moments": [P, P, P], "lbm_dense_mass": [P, P], "lbm_dense_synchronize": [P],
        "lbm_dense_set_stream": [P, P], "lbm_dense_graph_create": [P, I32, P],
        "lbm_dense_graph_launch": [P, I32], "lbm_dense_graph_destroy": [P],
        "lbm_copy_bandwidth": [ctypes.c_size_t, I32, P],
        "lbm_geometry_cylinder": [D, P, P], "lbm_geometry_channel": [I32, I32, I32, P],
        "lbm_geometry_add_spheres": [I32, I32, I32, P, D, D, D, ctypes.c_uint64, P, P],
        "lbm_field_smooth_random": [I32, I32, I32, D, ctypes.c_uint64, P],
        "lbm_adjoint_create": [P, P], "lbm_adjoint_destroy": [P],
        "lbm_adjoint_set_initial_state": [P, P], "lbm_adjoint_forward": [P, I32, P],
        "lbm_adjoint_backward": [P, P, P, P], "lbm_adjoint_get_lambda0": [P, P],
        "lbm_adjoint_get_final_state": [P, P], "lbm_adjoint_fd": [P, I32, I64, D, I32, I32, P],
        "lbm_adjoint_fd_direction": [P, I32, P, D, I32, I32, P],
        "lbm_adjoint_device_bytes": [P, P],
    }
    for name, args in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = ctypes.c_int
    _lib = L
    return L


def check(status: int) -> None:
    if status != OK:
        msg = load().lbm_last_error().decode(errors="replace")
        raise LbmError(status, msg)


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def device_count() -> int:
    n = ctypes.c_int(0)
    check(load().lbm_device_count(ctypes.byref(n)))
    return n.value


def copy_bandwidth(nbytes: int, reps: int = 10) -> float:
    """Achieved copy bandwidth in GB/s (read + write bytes / time)."""
    ms = ctypes.c_double(0)
    check(load().lbm_copy_bandwidth(nbytes, reps, ctypes.byref(ms)))
    return 2.0 * nbytes * reps / (ms.value * 1e-3) / 1e9


# --------------------------------------------------------------------------- geometry
def cylinder(simulation_size: float):
    """Reference cylinder mask; returns (dims (nx, ny, nz), solid uint8[ncell])."""
    dims = np.zeros(3, np.int32)
    check(load().lbm_geometry_cylinder(float(simulation_size), _p(dims), None))
    solid = np.zeros(int(np.prod(dims.astype(np.int64))), np.uint8)
    check(load().lbm_geometry_cylinder(float(simulation_size), _p(dims), _p(solid)))
    return tuple(int(d) for d in dims), solid


def channel(nx, ny, nz):
    solid = np.zeros(nx * ny * nz, np.uint8)
    check(load().lbm_geometry_channel(nx, ny, nz, _p(solid)))
    return solid


def add_spheres(solid, nx, ny, nz, rmin, rmax, fraction, seed, periodic=(1, 1, 1)):
    """Adds seeded random spheres to `solid` in place; returns the number drawn."""
    per = np.asarray(periodic, np.int32)
    count = ctypes.c_int64(0)
    check(load().lbm_geometry_add_spheres(nx, ny, nz, _p(per), float(rmin), float(rmax),
                                          float(fraction), int(seed), _p(solid),
                                          ctypes.byref(count)))
    return count.value


def smooth_random_field(nx, ny, nz, sigma, seed):
    out = np.empty(nx * ny * nz, np.float64)
    check(load().lbm_field_smooth_random(nx, ny, nz, float(sigma), int(seed), _p(out)))
    return out


MASK_MAGIC = b"LBMMASK1"


def write_mask(path, solid, dims):
    """Writes a solid mask file: the 8 bytes "LBMMASK1", int32 nx, ny, nz
    (little-endian), then one byte per cell, x fastest (cell x + nx*(y + ny*z)),
    0 fluid and 1 solid -- the format of the driver's `maskfile` / `writemask`."""
    nx, ny, nz = (int(d) for d in dims)
    solid = np.ascontiguousarray(solid).reshape(-1)
    if solid.size != nx * ny * nz or min(nx, ny, nz) <= 0:
        raise ValueError("mask size does not match its dimensions")
    with open(path, "wb") as fh:
        fh.write(MASK_MAGIC)
        fh.write(np.asarray([nx, ny, nz], "<i4").tobytes())
        fh.write((solid != 0).astype(np.uint8).tobytes())


def read_mask(path):
    """Reads a solid mask file (write_mask); returns (dims (nx, ny, nz), solid uint8[ncell])."""
    with open(path, "rb") as fh:
        raw = fh.read()
    if raw[:8] != MASK_MAGIC or len(raw) < 20:
        raise ValueError(f"{path} is not a mask file (LBMMASK1)")
    nx, ny, nz = (int(d) for d in np.frombuffer(raw[8:20], "<i4"))
    if min(nx, ny, nz) <= 0:
        raise ValueError(f"{path}: bad dimensions")
    n = nx * ny * nz
    if len(raw) < 20 + n:
        raise ValueError(f"{path} is truncated")
    solid = (np.frombuffer(raw[20:20 + n], np.uint8) != 0).astype(np.uint8)
    return (nx, ny, nz), solid


# --------------------------------------------------------------------------- device buffers
class DeviceBuffer:
    """A device allocation owned by the library (lbm_device_alloc)."""

    def __init__(self, nbytes: int):
        self.nbytes = int(nbytes)
        self.ptr = ctypes.c_void_p()
        check(load().lbm_device_alloc(self.nbytes, ctypes.byref(self.ptr)))

    @classmethod
    def from_array(cls, a: np.ndarray):
        a = np.ascontiguousarray(a)
        b = cls(a.nbytes)
        check(load().lbm_memcpy_h2d(b.ptr, _p(a), a.nbytes))
        return b

    def to_array(self, dtype, count=None) -> np.ndarray:
        dtype = np.dtype(dtype)
        count = self.nbytes // dtype.itemsize if count is None else count
        out = np.empty(count, dtype)
        check(load().lbm_memcpy_d2h(_p(out), self.ptr, out.nbytes))
        return out

    def upload(self, a: np.ndarray):
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        check(load().lbm_memcpy_h2d(self.ptr, _p(a), a.nbytes))

    def free(self):
        if self.ptr:
            check(load().lbm_device_free(self.ptr))
            self.ptr = ctypes.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


# --------------------------------------------------------------------------- sparse
class SparseKernels:
    """Drop-in kernels on the reference's fluid list + adjacency (lbm_sparse_*)."""

    def __init__(self, n_fluid, q, adjacency, omega, force=(1e-6, 0.0, 0.0), layout="soa",
                 storage="f64", adjacency_mode=1):
        self.q = q
        self.n = int(n_fluid)
        self.storage = STORAGE_NAMES[storage] if isinstance(storage, str) else storage
        self.dtype = np.float64 if self.storage == STORE_F64 else np.float32
        self._adj = np.ascontiguousarray(adjacency, np.int32)
        cfg = _SparseConfig(self.n, q, LAYOUT_NAMES[layout] if isinstance(layout, str) else layout,
                            self.storage, adjacency_mode, float(omega),
                            (ctypes.c_double * 3)(*force), _p(self._adj))
        self.h = ctypes.c_void_p()
        check(load().lbm_sparse_create(ctypes.byref(cfg), ctypes.byref(self.h)))

    def aa_even(self, buf: DeviceBuffer, start=0, count=None):
        check(load().lbm_sparse_aa_even(self.h, buf.ptr, start, self.n if count is None else count))

    def aa_odd(self, buf: DeviceBuffer, start=0, count=None):
        check(load().lbm_sparse_aa_odd(self.h, buf.ptr, start, self.n if count is None else count))

    def ab(self, src: DeviceBuffer, tgt: DeviceBuffer, start=0, count=None, pull=False):
        check(load().lbm_sparse_ab(self.h, src.ptr, tgt.ptr, start,
                                   self.n if count is None else count, int(pull)))

    def init_rest(self, buf: DeviceBuffer, aa: bool):
        check(load().lbm_sparse_init_rest(self.h, buf.ptr, int(aa)))

    def mass(self, buf: DeviceBuffer, start=0, count=None) -> float:
        """FP64 sum of the Q stored values of the nodes [start, start + count)."""
        m = ctypes.c_double(0)
        count = self.n - start if count is None else count
        check(load().lbm_sparse_mass(self.h, buf.ptr, int(start), int(count), ctypes.byref(m)))
        return m.value

    def mass_noalloc(self, buf: DeviceBuffer, out: DeviceBuffer, start=0, count=None):
        """lbm_sparse_mass_noalloc: enqueues the FP64 sum into the device double
        `out` (no allocation, no wait)."""
        count = self.n - start if count is None else count
        check(load().lbm_sparse_mass_noalloc(self.h, buf.ptr, int(start), int(count), out.ptr))

    def synchronize(self):
        check(load().lbm_sparse_synchronize(self.h))

    def set_stream(self, stream=None):
        """Launch on `stream` (a hipStream_t as an int, created by the HIP
        runtime this library is bound to; None: the null stream).  Buffer
        copies stay blocking on the null stream: synchronize() before reading
        results that were computed on another stream."""
        check(load().lbm_sparse_set_stream(self.h, ctypes.c_void_p(stream or None)))

    def close(self):
        if self.h:
            check(load().lbm_sparse_destroy(self.h))
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# --------------------------------------------------------------------------- dense
class DenseLattice:
    """Structured-grid engine (lbm_dense_*)."""

    def __init__(self, nx, ny, nz, q=19, pattern="aa", pull=False, layout="soa", storage="f64",
                 periodic=(1, 1, 1), omega=1.0, force=(0.0, 0.0, 0.0), solid=None, tau=None,
                 gamma=None, f16_scale=0.0, density_jump_x=0.0):
        self.shape = (int(nx), int(ny), int(nz))
        self.q = q
        self.ncell = int(nx) * int(ny) * int(nz)
        pat = {"aa": PATTERN_AA, "ab": PATTERN_AB}[pattern] if isinstance(pattern, str) else pattern
        self._solid = None if solid is None else np.ascontiguousarray(solid, np.uint8).reshape(-1)
        self._tau = None if tau is None else np.ascontiguousarray(tau, np.float64).reshape(-1)
        self._gamma = None if gamma is None else np.ascontiguousarray(gamma, np.float64).reshape(-1)
        for a in (self._solid, self._tau, self._gamma):
            if a is not None and a.size != self.ncell:
                raise ValueError("per-cell arrays must have nx*ny*nz entries")
        cfg = _DenseConfig(nx, ny, nz, q, pat, int(pull),
                           LAYOUT_NAMES[layout] if isinstance(layout, str) else layout,
                           STORAGE_NAMES[storage] if isinstance(storage, str) else storage,
                           (ctypes.c_int32 * 3)(*[int(p) for p in periodic]), float(omega),
                           (ctypes.c_double * 3)(*force), _p(self._solid), _p(self._tau),
                           _p(self._gamma), float(f16_scale), float(density_jump_x))
        self.h = ctypes.c_void_p()
        check(load().lbm_dense_create(ctypes.byref(cfg), ctypes.byref(self.h)))

    def info(self) -> dict:
        i = _DenseInfo()
        check(load().lbm_dense_get_info(self.h, ctypes.byref(i)))
        return {"ncell": i.ncell, "n_fluid": i.n_fluid, "time_step": i.time_step,
                "bytes_per_lup": i.bytes_per_lup, "device_bytes": i.device_bytes}

    def set_state(self, f: np.ndarray):
        f = np.ascontiguousarray(f, np.float64).reshape(-1)
        assert f.size == self.q * self.ncell
        check(load().lbm_dense_set_state(self.h, _p(f)))

    def get_state(self) -> np.ndarray:
        out = np.empty(self.q * self.ncell, np.float64)
        check(load().lbm_dense_get_state(self.h, _p(out)))
        return out.reshape(self.q, self.ncell)

    def init_equilibrium(self, u: np.ndarray, rho: np.ndarray | None = None):
        u = np.ascontiguousarray(u, np.float64).reshape(-1)
        assert u.size == 3 * self.ncell
        r = None if rho is None else np.ascontiguousarray(rho, np.float64).reshape(-1)
        check(load().lbm_dense_init_equilibrium(self.h, _p(r), _p(u)))

    def init_tgv(self, u0=0.05, noise=1e-4, seed=1):
        check(load().lbm_dense_init_tgv(self.h, float(u0), float(noise), int(seed)))

    def step(self, n=1):
        check(load().lbm_dense_step(self.h, int(n)))

    def run_timed(self, n) -> float:
        ms = ctypes.c_double(0)
        check(load().lbm_dense_run_timed(self.h, int(n), ctypes.byref(ms)))
        return ms.value

    def moments(self):
        rho = np.empty(self.ncell)
        u = np.empty(3 * self.ncell)
        check(load().lbm_dense_moments(self.h, _p(rho), _p(u)))
        return rho, u.reshape(self.ncell, 3)

    def mass(self) -> float:
        m = ctypes.c_double(0)
        check(load().lbm_dense_mass(self.h, ctypes.byref(m)))
        return m.value

    def synchronize(self):
        check(load().lbm_dense_synchronize(self.h))

    def set_stream(self, stream=None):
        """Enqueue later calls on `stream` (a hipStream_t as an int, created by
        the HIP runtime this library is bound to; None: the handle's own stream)."""
        check(load().lbm_dense_set_stream(self.h, ctypes.c_void_p(stream or None)))

    def graph(self, nsteps) -> "DenseGraph":
        """Capture `nsteps` (even) time steps as a HIP graph (nothing runs)."""
        return DenseGraph(self, nsteps)

    def close(self):
        if self.h:
            check(load().lbm_dense_destroy(self.h))
            self.h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DenseGraph:
    """An even number of time steps captured as a HIP graph (lbm_dense_graph_*).
    Keeps its lattice alive; close it before the lattice."""

    def __init__(self, lattice, nsteps):
        self.lattice = lattice
        self.nsteps = int(nsteps)
        self.g = ctypes.c_void_p()
        check(load().lbm_dense_graph_create(lattice.h, self.nsteps, ctypes.byref(self.g)))

    def launch(self, n=1):
        """Replay n times on the lattice's current stream (asynchronous)."""
        check(load().lbm_dense_graph_launch(self.g, int(n)))

    def close(self):
        if self.g:
            check(load().lbm_dense_graph_destroy(self.g))
            self.g = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# --------------------------------------------------------------------------- adjoint
class Adjoint:
    """Discrete adjoint forward/backward pair (lbm_adjoint_*)."""

    OBJECTIVES = {"kinetic_energy": 0, "outlet_flux": 1}

    def __init__(self, nx, ny, nz, q=19, storage="f64", periodic=(1, 1, 1), omega=1.0,
                 force=(0.0, 0.0, 0.0), solid=None, tau=None, gamma=None,
                 objective="kinetic_energy", outlet_x=0, checkpoint_every=50,
                 checkpoint_inner=(0, 0), density_jump_x=0.0):
        self.shape = (int(nx), int(ny), int(nz))
        self.q = q
        self.ncell = int(nx) * int(ny) * int(nz)
        self._solid = None if solid is None else np.ascontiguousarray(solid, np.uint8).reshape(-1)
        self._tau = None if tau is None else np.ascontiguousarray(tau, np.float64).reshape(-1)
        self._gamma = None if gamma is None else np.ascontiguousarray(gamma, np.float64).reshape(-1)
        obj = self.OBJECTIVES[objective] if isinstance(objective, str) else int(objective)
        cfg = _AdjointConfig(nx, ny, nz, q, STORAGE_NAMES[storage] if isinstance(storage, str) else storage,
                             (ctypes.c_int32 * 3)(*[int(p) for p in periodic]), float(omega),
                             (ctypes.c_double * 3)(*force), _p(self._solid), _p(self._tau),
                             _p(self._gamma), obj, int(outlet_x), int(checkpoint_every),
                             (ctypes.c_int32 * 2)(*[int(k) for k in checkpoint_inner]),
                             float(density_jump_x))
        self.h = ctypes.c_void_p()
        check(load().lbm_adjoint_create(ctypes.byref(cfg), ctypes.byref(self.h)))

    def set_initial_state(self, f):
        f = np.ascontiguousarray(f, np.float64).reshape(-1)
        assert f.size == self.q * self.ncell
        check(load().lbm_adjoint_set_initial_state(self.h, _p(f)))

    def forward(self, nsteps) -> float:
        j = ctypes.c_double(0)
        check(load().lbm_adjoint_forward(self.h, int(nsteps), ctypes.byref(j)))
        return j.value

    def backward(self, gamma=True, tau=True):
        gg = np.empty(self.ncell) if (gamma and self._gamma is not None) else None
        gt = np.empty(self.ncell) if (tau and self._tau is not None) else None
        ms = ctypes.c_double(0)
        check(load().lbm_adjoint_backward(self.h, _p(gg), _p(gt), ctypes.byref(ms)))
        return gg, gt, ms.value

    def lambda0(self):
        out = np.empty(self.q * self.ncell)
        check(load().lbm_adjoint_get_lambda0(self.h, _p(out)))
        return out.reshape(self.q, self.ncell)

    def final_state(self):
        out = np.empty(self.q * self.ncell)
        check(load().lbm_adjoint_get_final_state(self.h, _p(out)))
        return out.reshape(self.q, self.ncell)

    def finite_difference(self, param, cell, step, nsteps, order=6) -> float:
        d = ctypes.c_double(0)
        p = {"gamma": 0, "tau": 1}[param] if isinstance(param, str) else int(param)
        check(load().lbm_adjoint_fd(self.h, p, int(cell), float(step), int(nsteps), int(order),
                                    ctypes.byref(d)))
        return d.value

    def finite_difference_direction(self, param, direction, step, nsteps, order=6) -> float:
        """d/de J(p + e v) at e = 0 for the whole-field direction v (ncell values)."""
        d = ctypes.c_double(0)
        v = np.ascontiguousarray(direction, np.float64).reshape(-1)
        assert v.size == self.ncell
        p = {"gamma": 0, "tau": 1}[param] if isinstance(param, str) else int(param)
        check(load().lbm_adjoint_fd_direction(self.h, p, _p(v), float(step), int(nsteps),
                                              int(order), ctypes.byref(d)))
        return d.value

    def device_bytes(self) -> int:
        b = ctypes.c_size_t(0)
        check(load().lbm_adjoint_device_bytes(self.h, ctypes.byref(b)))
        return b.value

    def close(self):
        if self.h:
            check(load().lbm_adjoint_destroy(self.h))
            self.h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
