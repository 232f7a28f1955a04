"""The C-ABI library loads and exports every entry point include/fastlpr_amd.h declares
(no GPU work is launched)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "fastlpr_amd.h")


def _declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:int|int64_t|double)\s+(flpr_\w+)\s*\(", text, flags=re.M)))


def test_header_declares_entry_points():
    names = _declared()
    assert "flpr_spread" in names and "flpr_conv" in names and "flpr_solve" in names
    from fastlpr_amd import _lib
    assert set(names) == set(_lib.EXPORTS)


def test_library_exports_every_declared_symbol():
    from fastlpr_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libfastlpr_amd.so not built (run __graft_entry__.build())")
    lib = _lib.load()
    for name in _declared():
        assert hasattr(lib, name), name
        assert isinstance(getattr(lib, name), ctypes._CFuncPtr)


def test_geometry_struct_matches_header_layout():
    """FlprGeom field order/size in ctypes equals the C struct (22 int32 then 13x3 doubles)."""
    from fastlpr_amd._lib import FlprGeom
    assert ctypes.sizeof(FlprGeom) == 22 * 4 + 13 * 3 * 8
    text = open(HEADER).read()
    body = text[text.index("typedef struct FlprGeom"):text.index("} FlprGeom;")]
    fields = re.findall(r"(?:int32_t|double)\s+(\w+)\s*[\[;]", body)
    assert fields == [f[0] for f in FlprGeom._fields_]


def test_key_roundtrip_without_gpu():
    """flpr_key_to_double is host code: the order-preserving key round-trips."""
    from fastlpr_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("library not built")
    lib = _lib.load()
    import struct
    for v in (0.0, 1.5, -2.25, 1e-300, 3e300):
        bits = struct.unpack("<Q", struct.pack("<d", v))[0]
        key = (~bits & (2**64 - 1)) if bits >> 63 else (bits | (1 << 63))
        assert lib.flpr_key_to_double(key) == v


def test_conv_strip_length_without_gpu():
    """flpr_conv_strip_len is host code: each of the eight shifted copies of a
    tap row holds the reversed window positions nmax + 62 + 95 (the windows a
    float32-form tile can read), rounded up to whole 16-byte units."""
    from fastlpr_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("library not built")
    lib = _lib.load()
    assert lib.flpr_conv_strip_len(0) == -1
    for nmax in (1, 40, 255, 256, 1024):
        v = lib.flpr_conv_strip_len(nmax)
        assert v % 8 == 0 and nmax + 157 <= v < nmax + 157 + 8
