"""The C ABI: layouts, exported symbols, and host-side validation (no GPU)."""
import ctypes
import re
from pathlib import Path

import numpy as np

from stargaze_amd import _native, layouts

ROOT = Path(__file__).resolve().parent.parent
HEADER = ROOT / "include" / "stargaze_amd.h"


def declared_functions():
    text = HEADER.read_text()
    return sorted(set(re.findall(r"^\s*(?:int|uint32_t|uint64_t|const char \*)\s*(sg_\w+)\s*\(", text, re.M)))


def test_header_declares_what_the_binding_binds():
    assert declared_functions() == sorted(_native.EXPORTS)


def test_library_loads_and_exports_every_symbol():
    lib = ctypes.CDLL(str(_native.LIBRARY))
    for name in declared_functions():
        assert hasattr(lib, name), name


def test_reference_layout_sizes():
    """src/pathtracer_tests.rs:63-97 checks CPU/GPU struct sizes; same sizes here."""
    assert layouts.GLOBALS.itemsize == 176
    assert layouts.SPHERE.itemsize == 160
    assert layouts.PRIMITIVE.itemsize == 48
    assert layouts.RAY_STAR.itemsize == 32
    assert layouts.SETTINGS.itemsize == 240
    assert layouts.CAMERA_AIR.itemsize == 16


def test_field_offsets_match_the_wgsl_order():
    g = layouts.GLOBALS
    assert g.fields["viewport"][1] == 48 and g.fields["ground_counts"][1] == 160
    s = layouts.SPHERE
    assert s.fields["radius"][1] == 12 and s.fields["emissive"][1] == 28 and s.fields["surface_params"][1] == 144
    st = layouts.SETTINGS
    assert st.fields["counts"][1] == 176 and st.fields["lights"][1] == 224


def test_errors_without_context_are_reported():
    lib = _native.load()
    assert lib.sg_destroy(None) == 0
    assert lib.sg_resize(None, 1, 1) == -1
    assert lib.sg_samples(None) == 0
    out = ctypes.c_void_p()
    assert lib.sg_create(ctypes.byref(out), 0, 0, None, 0, None, 0, None) == -1
