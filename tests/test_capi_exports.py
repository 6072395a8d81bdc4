"""The HIP library loads on the CPU host and exports every entry point that
include/picopic.h declares (no compute calls: there is no GPU here)."""
import ctypes
import os
import re

import pytest

from picopic_amd import capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pcp\d*_\w+)\s*\(", text)))


def test_library_built():
    assert os.path.exists(capi.LIB_PATH), "run python -m picopic_amd.build"


def test_every_declared_symbol_is_exported():
    lib = ctypes.CDLL(capi.LIB_PATH)
    names = declared("picopic.h")
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert set(capi.EXPORTS) <= set(names)


def test_error_message_entry_point():
    lib = capi.load()
    assert isinstance(lib.pcp_last_error(), bytes)


def test_create_rejects_bad_config_without_gpu():
    # argument validation happens before any device call
    lib = capi.load()
    cfg = capi.Config()
    cfg.grid_r, cfg.grid_z, cfg.domains_r, cfg.domains_z = 10, 10, 3, 1
    cfg.n_species, cfg.dt = 1, 1e-12
    h = ctypes.c_void_p()
    rc = lib.pcp_create(ctypes.byref(cfg), ctypes.byref(h))
    assert rc == -1
    assert b"divide" in lib.pcp_last_error()


def test_cpp_driver_links_against_the_library():
    """The SMB-style C++ driver over include/picopic_solvers.hpp is built and
    resolves libpicopic_hip.so from its own directory."""
    import subprocess
    drv = os.path.join(ROOT, "picopic_amd", "_lib", "smb_driver")
    assert os.path.exists(drv), "run python -m picopic_amd.build"
    out = subprocess.run(["ldd", drv], capture_output=True, text=True).stdout
    line = [l for l in out.splitlines() if "libpicopic_hip.so" in l]
    assert line and "not found" not in line[0], out
