"""The C-ABI library without a GPU: it loads, exports every entry point that
include/lbm_capi.h declares, and rejects invalid arguments with status codes
before any device call (no compute runs here)."""
import ctypes
import os
import re

import numpy as np
import pytest

from lbmminiapp_amd import capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "lbm_capi.h")


def header_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(lbm_[a-z0-9_]+)\s*\(", text)))


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(capi.LIB_PATH):
        from lbmminiapp_amd import build
        build.build()
    return capi.load()


def test_header_declares_expected_entry_points():
    names = header_functions()
    assert set(capi.EXPORTED) == set(names), set(capi.EXPORTED) ^ set(names)


def test_library_exports_every_declared_symbol(lib):
    missing = [n for n in header_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_version_and_error_string(lib):
    assert b"gfx950" in lib.lbm_version()
    assert isinstance(lib.lbm_last_error(), bytes)


def test_dense_create_rejects_bad_arguments(lib):
    h = ctypes.c_void_p()
    base = dict(nx=8, ny=8, nz=8, q=19, pattern=1, pull=0, layout=0, storage=0,
                periodic=(ctypes.c_int32 * 3)(1, 1, 1), omega=1.0,
                force=(ctypes.c_double * 3)(0, 0, 0), solid=None, tau=None, gamma=None,
                f16_scale=0.0)
    for bad in (dict(q=20), dict(nx=0), dict(pattern=7), dict(layout=3), dict(storage=9),
                dict(omega=2.5), dict(omega=float("nan")), dict(ny=70000)):
        cfg = capi._DenseConfig(**{**base, **bad})
        assert lib.lbm_dense_create(ctypes.byref(cfg), ctypes.byref(h)) == capi.ERR_ARG, bad
        assert lib.lbm_last_error()
    assert lib.lbm_dense_create(None, ctypes.byref(h)) == capi.ERR_ARG


def test_sparse_create_rejects_bad_adjacency(lib):
    import numpy as np
    adj = np.zeros(19 * 4, np.int32)
    adj[5] = 19 * 4  # out of range
    cfg = capi._SparseConfig(4, 19, 0, capi.STORE_F64, 1, 1.0, (ctypes.c_double * 3)(1e-6, 0, 0),
                             adj.ctypes.data_as(ctypes.c_void_p))
    h = ctypes.c_void_p()
    assert lib.lbm_sparse_create(ctypes.byref(cfg), ctypes.byref(h)) == capi.ERR_ARG
    assert b"adjacency" in lib.lbm_last_error()
    cfg.q = 21
    assert lib.lbm_sparse_create(ctypes.byref(cfg), ctypes.byref(h)) == capi.ERR_ARG
    cfg.q = 19
    cfg.storage = capi.STORE_F16S
    assert lib.lbm_sparse_create(ctypes.byref(cfg), ctypes.byref(h)) == capi.ERR_ARG


def test_null_handles_are_rejected(lib):
    assert lib.lbm_dense_step(None, 1) == capi.ERR_ARG
    assert lib.lbm_dense_mass(None, None) == capi.ERR_ARG
    assert lib.lbm_sparse_aa_even(None, None, 0, 1) == capi.ERR_ARG
    assert lib.lbm_copy_bandwidth(15, 1, None) == capi.ERR_ARG


def test_adjoint_create_rejects_bad_checkpoint_plans(lib):
    h = ctypes.c_void_p()
    base = dict(nx=8, ny=8, nz=8, q=19, storage=capi.STORE_F64,
                periodic=(ctypes.c_int32 * 3)(1, 1, 1), omega=1.0,
                force=(ctypes.c_double * 3)(0, 0, 0), solid=None, tau=None, gamma=None,
                objective=0, outlet_x=0, checkpoint_every=10)
    for bad in (dict(checkpoint_every=0), dict(checkpoint_inner=(ctypes.c_int32 * 2)(11, 0)),
                dict(checkpoint_inner=(ctypes.c_int32 * 2)(0, 3)),
                dict(checkpoint_inner=(ctypes.c_int32 * 2)(4, 5)),
                dict(checkpoint_inner=(ctypes.c_int32 * 2)(-1, 0)),
                dict(objective=1, outlet_x=8), dict(storage=capi.STORE_F16S)):
        cfg = capi._AdjointConfig(**{**base, **bad})
        assert lib.lbm_adjoint_create(ctypes.byref(cfg), ctypes.byref(h)) == capi.ERR_ARG, bad
        assert lib.lbm_last_error()


def test_mask_file_round_trip_and_errors(tmp_path):
    """Solid mask files (the driver's `maskfile` / `writemask` format): written
    and read back equal, nonzero bytes read as solid, and malformed files
    (wrong magic, bad dimensions, truncated data) are rejected."""
    from lbmminiapp_amd import capi

    dims = (7, 5, 3)
    solid = (np.arange(7 * 5 * 3) % 4 == 0).astype(np.uint8) * 3
    path = tmp_path / "m.bin"
    capi.write_mask(path, solid, dims)
    raw = path.read_bytes()
    assert raw[:8] == b"LBMMASK1" and len(raw) == 20 + 105
    assert list(np.frombuffer(raw[8:20], "<i4")) == [7, 5, 3]
    got_dims, got = capi.read_mask(path)
    assert got_dims == dims and np.array_equal(got, (solid != 0).astype(np.uint8))
    # layout: cell x + nx*(y + ny*z)
    assert got[2 + 7 * (1 + 5 * 2)] == (solid[2 + 7 * (1 + 5 * 2)] != 0)
    bad = tmp_path / "bad.bin"
    bad.write_bytes(b"LBMMASK2" + raw[8:])
    with pytest.raises(ValueError):
        capi.read_mask(bad)
    bad.write_bytes(raw[:8] + np.asarray([7, 0, 3], "<i4").tobytes() + raw[20:])
    with pytest.raises(ValueError):
        capi.read_mask(bad)
    bad.write_bytes(raw[:-1])
    with pytest.raises(ValueError):
        capi.read_mask(bad)
    with pytest.raises(ValueError):
        capi.write_mask(bad, solid[:-1], dims)


def test_header_is_plain_c(tmp_path):
    """include/lbm_capi.h compiles as strict C99 (the ABI is for C, cgo and
    ctypes callers), and a C program links against the library and calls it:
    the version string, and an argument error reported through
    lbm_last_error without any device call."""
    import shutil
    import subprocess

    gcc = shutil.which("gcc")
    if gcc is None:
        pytest.skip("no C compiler")
    src = tmp_path / "use_capi.c"
    src.write_text(r'''
#include <stdio.h>
#include <string.h>
#include "lbm_capi.h"
int main(void) {
  lbm_dense* h = NULL;
  lbm_status st;
  if (!lbm_version() || !lbm_version()[0]) return 2;
  st = lbm_dense_create(NULL, &h);
  if (st != LBM_ERR_ARG || h != NULL || strlen(lbm_last_error()) == 0) return 3;
  printf("%s\n", lbm_version());
  return 0;
}
''')
    exe = tmp_path / "use_capi"
    lib_dir = os.path.join(ROOT, "lbmminiapp_amd")
    r = subprocess.run([gcc, "-std=c99", "-Wall", "-Wextra", "-pedantic", "-Werror",
                        "-I" + os.path.join(ROOT, "include"), str(src), "-o", str(exe),
                        "-L" + lib_dir, "-llbm_hip", "-Wl,-rpath," + lib_dir],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, (r.returncode, r.stdout, r.stderr)
    assert r.stdout.strip()
