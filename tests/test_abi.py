"""The C-ABI library loads and exports every symbol include/fesom_fct.h declares (CPU
only: no compute call is made without a GPU)."""

import ctypes
import re

import numpy as np
from pathlib import Path

import pytest

from fesom_jax_amd import _lib

HEADER = Path(__file__).resolve().parents[1] / "include" / "fesom_fct.h"


def _declared():
    text = HEADER.read_text()
    return sorted(set(re.findall(r"^(?:int|const char\*)\s+(fct_\w+)\(", text, re.M)))


def test_header_and_binding_agree():
    assert _declared() == sorted(_lib.EXPORTED_SYMBOLS)


def test_library_exports_all_symbols():
    if not _lib.LIB_PATH.exists():
        pytest.skip("library not built")
    lib = _lib.load(require_torch=False)
    for name in _declared():
        assert hasattr(lib, name), name
    assert lib.fct_abi_version() == _lib.FCT_ABI_VERSION
    assert lib.fct_last_error() == b""


def test_struct_layouts_match_header():
    # field counts mirror the C structs; a missing/extra field shifts every pointer
    text = HEADER.read_text()
    body = text[text.index("typedef struct fct_mesh_desc"):text.index("} fct_mesh_desc;")]
    n_ptr = len(re.findall(r"\*\s*\w+;", body)) + len(re.findall(r"\*\s*\w+,", body))
    n_int = 5  # nod2D, elem2D, edge2D, nl, max_tracers
    assert len(_lib.FctMeshDesc._fields_) == n_ptr + n_int
    assert ctypes.sizeof(_lib.FctMeshDesc) == 4 * 4 + 8 * n_ptr + 8


def test_plan_create_rejects_bad_mesh_without_gpu():
    """Argument validation happens before any device call."""
    if not _lib.LIB_PATH.exists():
        pytest.skip("library not built")
    lib = _lib.load(require_torch=False)
    out = ctypes.c_void_p()
    rc = lib.fct_plan_create(None, ctypes.byref(out))
    assert rc == _lib.FCT_E_INVALID
    assert b"NULL" in lib.fct_last_error()


def test_enum_constants_match_header():
    """Every FCT_TUNE_*, FCT_PATH_* and FCT_E_* value in the header has the same value
    in the Python binding."""
    text = HEADER.read_text()
    pairs = dict(re.findall(r"\b(FCT_(?:TUNE|PATH|E)_\w+)\s*=\s*(-?\d+)", text))
    pairs.update(re.findall(r"#define\s+(FCT_PATH_BIT_\w+)\s+(\d+)", text))
    pairs.update(re.findall(r"#define\s+(FCT_OK|FCT_E_\w+)\s+\(?(-?\d+)\)?", text))
    assert "FCT_E_UNSUPPORTED" in pairs and "FCT_PATH_BIT_FUSED" in pairs
    assert any(k.startswith("FCT_TUNE_") for k in pairs)
    for name, value in pairs.items():
        assert hasattr(_lib, name), name
        assert getattr(_lib, name) == int(value), name


def test_stage_ids_match_header():
    text = HEADER.read_text()
    ids = dict(re.findall(r"\bFCT_STAGE_(\w+)\s*=\s*(\d+)", text))
    assert ids and {k: int(v) for k, v in ids.items()} == _lib.FCT_STAGE


def test_mesh_desc_refuses_mismatched_shapes(pi_mesh):
    """A field whose shape disagrees with the mesh counts is refused before the library
    could read past its end."""
    import dataclasses
    bad = dataclasses.replace(pi_mesh, edge_up_dn_tri=pi_mesh.edge_up_dn_tri[:-3])
    with pytest.raises(ValueError, match="edge_up_dn_tri"):
        _lib.mesh_desc_from(bad)
    bad = dataclasses.replace(pi_mesh, area=pi_mesh.area[:, :-1])
    with pytest.raises(ValueError, match="area"):
        _lib.mesh_desc_from(bad)


@pytest.mark.parametrize("field, value, word", [
    ("ulevels", 0, "element"), ("nlevels", 10_000, "element"),
    ("ulevels_nod2D_max", 10_000, "node"), ("nlevels_nod2D_min", 0, "node")])
def test_plan_create_checks_level_ranges_without_gpu(pi_mesh, field, value, word):
    """Level bounds outside [1, nl] are refused by the plan's validation, which runs
    before any device call."""
    if not _lib.LIB_PATH.exists():
        pytest.skip("library not built")
    lib = _lib.load(require_torch=False)
    arr = np.array(getattr(pi_mesh, field), copy=True)
    arr[5] = value
    import dataclasses
    desc, keep = _lib.mesh_desc_from(dataclasses.replace(pi_mesh, **{field: arr}))
    out = ctypes.c_void_p()
    rc = lib.fct_plan_create(ctypes.byref(desc), ctypes.byref(out))
    assert rc == _lib.FCT_E_INVALID
    assert word.encode() in lib.fct_last_error()
