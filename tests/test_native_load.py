"""The C-ABI libraries load without a GPU and export every symbol that
include/*.h declares (no compute calls here)."""
import ctypes as C

import pytest

from tetragpu_amd import _native

HEADER_LIB = {
    "tetra_c.h": "libtetra_hip.so",
    "tetra_meshgen.h": "libtetra_host.so",
}


def test_every_header_maps_to_a_library():
    decl = _native.declared_symbols()
    assert set(decl) == set(HEADER_LIB), f"unmapped headers: {set(decl) ^ set(HEADER_LIB)}"


@pytest.mark.parametrize("header", sorted(HEADER_LIB))
def test_header_symbols_exported(header):
    names = _native.declared_symbols()[header]
    assert names, header
    _native.hip()
    lib = C.CDLL(str(_native.LIB_DIR / HEADER_LIB[header]))
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, f"{header}: not exported: {missing}"


def test_no_device_is_reported_cleanly():
    import tetragpu_amd as t

    if t.device_count() > 0:
        pytest.skip("a device is visible")
    with pytest.raises(_native.NativeError) as e:
        t.init(0)
    assert e.value.status == -7
