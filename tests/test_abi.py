"""The C ABI without a GPU: the library loads, exports every symbol the public
header declares, and reports a missing device as an error (no compute)."""

import ctypes as C
import os

import pytest

import splattar_amd as sa
from splattar_amd import _lib as L


def test_library_exports_every_declared_symbol():
    lib = sa.lib()
    syms = sa.header_symbols()
    assert len(syms) >= 60
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, missing


def test_prototypes_cover_the_header():
    # Every declared function has a ctypes prototype in the binding.
    assert set(sa.header_symbols()) == set(L._PROTOS)


def test_abi_version_and_errors():
    lib = sa.lib()
    assert lib.gs_abi_version() == 1
    assert b"gfx950" in lib.gs_build_info()
    # Null-argument misuse is an error status with a message, not a crash.
    assert lib.gs_ctx_create(0, None) == 1
    assert b"null" in lib.gs_last_error()


@pytest.mark.skipif(os.environ.get("HIP_VISIBLE_DEVICES", None) is None and os.path.exists("/dev/kfd"),
                    reason="a GPU may be present; the no-device path is checked only without one")
def test_context_without_device_reports_no_device():
    h = C.c_void_p()
    st = sa.lib().gs_ctx_create(0, C.byref(h))
    assert st == 6 and not h.value


def test_header_is_plain_c():
    import re
    text = re.sub(r"/\*.*?\*/", "", open(L.HEADER_PATH).read(), flags=re.S)
    for banned in ("torch", "hip_runtime", "std::", "class ", "template"):
        assert banned not in text


def test_rust_binding_declares_every_entry_point():
    """The `-sys` crate a maintainer adds beside gs-wgpu declares one
    `pub fn` per function of the header, so the documented binding and the
    library cannot drift apart."""
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    text = open(os.path.join(root, "INTEGRATION.md")).read()
    missing = [s for s in sa.header_symbols() if f"pub fn {s}(" not in text]
    assert not missing, missing
