"""The C ABI library (no GPU needed): it loads, exports exactly the symbols
include/cp3_hip.h declares, the Python view declares them all, and without a
device every entry point fails with a status code instead of crashing."""
import ctypes as C
import os
import re

import pytest

from conftest import ROOT


def header_symbols():
    src = open(os.path.join(ROOT, "include", "cp3_hip.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(cp3_\w+)\s*\(", src)))


def test_header_symbols_exported():
    from cudaprob3_amd import _native
    lib = _native.load()
    missing = [s for s in header_symbols() if not hasattr(lib, s)]
    assert not missing, missing
    assert len(header_symbols()) >= 50


def test_python_view_declares_every_symbol():
    from cudaprob3_amd import _native
    assert sorted(_native.declared_symbols()) == header_symbols()


def test_exports_are_only_the_c_abi():
    import subprocess
    out = subprocess.run(["nm", "-D", "--defined-only",
                          os.path.join(ROOT, "cudaprob3_amd", "libcp3hip.so")],
                         capture_output=True, text=True).stdout
    cp3 = sorted(l.split()[-1] for l in out.splitlines()
                 if l.split() and l.split()[-1].startswith("cp3_"))
    assert cp3 == header_symbols()


def no_gpu():
    from cudaprob3_amd import _native
    n = C.c_int(0)
    rc = _native.load().cp3_device_count(C.byref(n))
    return rc != 0 or n.value == 0


def test_version_and_errors_without_device():
    from cudaprob3_amd import _native as N
    lib = N.load()
    assert lib.cp3_version() >= 10000
    if not no_gpu():
        pytest.skip("a GPU is present; the no-device path is not reachable")
    h = C.c_void_p()
    rc = lib.cp3_atmos_create(0, 10, 10, 0, C.byref(h))
    assert rc == N.CP3_ERR_DEVICE
    assert not h.value
    assert lib.cp3_last_error()
    rc = lib.cp3_batch_create(0, 2, C.byref(h))
    assert rc in (N.CP3_ERR_DEVICE, N.CP3_ERR_INVALID)


def test_null_handles_rejected():
    from cudaprob3_amd import _native as N
    lib = N.load()
    assert lib.cp3_atmos_calculate(None, 0) == N.CP3_ERR_INVALID
    assert b"null" in lib.cp3_last_error()
    assert lib.cp3_beam_set_density(None, 2.6) == N.CP3_ERR_INVALID
    assert lib.cp3_batch_run(None) == N.CP3_ERR_INVALID
    assert lib.cp3_atmos_destroy(None) == N.CP3_OK


def test_python_view_fails_loudly_without_library(tmp_path):
    from cudaprob3_amd import _native as N
    with pytest.raises(ImportError, match="no CPU fallback"):
        N._lib_backup = N._lib
        try:
            N._lib = None
            N.load(str(tmp_path / "missing.so"))
        finally:
            N._lib = N._lib_backup
