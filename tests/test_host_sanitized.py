"""The C++ host layer and the deck driver built with AddressSanitizer and
UndefinedBehaviorSanitizer (g++, host code only: no device code is built
here), run through the CPU scenarios of test_host_geometry: the reference's
cylinder decks (fluid order, ranges and adjacency against the reference
fixtures), deck parsing errors and warnings, the rejected MR pattern, and mask
files (read, written back, bad magic).  Any sanitizer report fails the run
(exit code 86 / 87, leak checking on)."""
import os
import subprocess

import pytest

import test_host_geometry as H

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "lbmminiapp_amd")
HOST = os.path.join(PKG, "host")


@pytest.fixture(scope="module")
def driver(tmp_path_factory):
    lib = os.path.join(PKG, "liblbm_hip.so")
    if not os.path.exists(lib):
        from lbmminiapp_amd import build
        build.build()
    out = tmp_path_factory.mktemp("sanitized") / "lbm-proxy-app"
    srcs = sorted(os.path.join(HOST, f) for f in os.listdir(HOST) if f.endswith(".cpp"))
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer",
           "-fsanitize=address,undefined", "-fno-sanitize-recover=undefined",
           "-I" + os.path.join(ROOT, "include"), "-I" + HOST, *srcs, "-o", str(out),
           "-L" + PKG, "-llbm_hip", "-Wl,-rpath," + PKG]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    saved = {k: os.environ.get(k) for k in ("ASAN_OPTIONS", "UBSAN_OPTIONS")}
    os.environ["ASAN_OPTIONS"] = "exitcode=86:detect_leaks=1:abort_on_error=0"
    os.environ["UBSAN_OPTIONS"] = "exitcode=87:print_stacktrace=1:halt_on_error=1"
    yield str(out)
    for k, v in saved.items():
        if v is None:
            os.environ.pop(k, None)
        else:
            os.environ[k] = v


def test_sanitized_cylinder_fixtures(driver, tmp_path):
    assert H.GOLDEN
    for i, path in enumerate(H.GOLDEN):
        d = tmp_path / str(i)
        d.mkdir()
        H.test_cylinder_geometry_matches_reference(driver, path, d)


def test_sanitized_reference_deck(driver, tmp_path):
    H.test_reference_deck_geometry_s2(driver, tmp_path)


def test_sanitized_deck_errors(driver, tmp_path):
    H.test_deck_errors(driver, tmp_path)
    H.test_deck_rejects_unimplemented_pattern(driver, tmp_path)


def test_sanitized_mask_files(driver, tmp_path):
    H.test_mask_file_geometry(driver, tmp_path)


def test_sanitizer_is_live(driver, tmp_path):
    """The build really carries the sanitizers: the binary links the ASan and
    UBSan runtimes."""
    r = subprocess.run(["ldd", driver], capture_output=True, text=True)
    assert "libasan" in r.stdout and "libubsan" in r.stdout
