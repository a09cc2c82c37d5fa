"""The C ABI: the built library loads and exports every entry point that
include/fem355.h declares, and the ctypes table of the package covers exactly
that set.  No compute call is made (CPU only)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "fem355.h")
LIB = os.path.join(ROOT, "fem_taichi_amd", "libfem355.so")


def declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(fem_\w+)\s*\(", text, flags=re.M)))


def test_header_parses():
    names = declared()
    assert "fem_hex8_assemble" in names and "fem_bsr_spmv" in names and "fem_last_error" in names
    assert len(names) >= 30


@pytest.mark.skipif(not os.path.exists(LIB), reason="libfem355.so not built")
def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(LIB)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing
    lib.fem_abi_version.restype = ctypes.c_int
    assert lib.fem_abi_version() >= 1


def test_ctypes_table_matches_header():
    from fem_taichi_amd import _lib
    assert sorted(_lib.SIGNATURES) == declared()


def test_fem_mesh_struct_matches_header():
    """Field names and order of the ctypes mirror equal struct fem_mesh's."""
    from fem_taichi_amd import _lib
    text = open(HEADER).read()
    body = re.search(r"typedef struct fem_mesh \{(.*?)\} fem_mesh;", text, flags=re.S).group(1)
    body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
    fields = re.findall(r"(\w+)\s*(?:\[\d+\])?\s*;", body)
    assert fields == [f[0] for f in _lib.FemMesh._fields_]


def test_grid_transfer_struct_matches_header():
    """Field names and order of the ctypes mirror equal struct fem_grid_transfer's."""
    from fem_taichi_amd import _lib
    text = open(HEADER).read()
    body = re.search(r"typedef struct fem_grid_transfer \{(.*?)\} fem_grid_transfer;", text, flags=re.S).group(1)
    body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
    fields = re.findall(r"(\w+)\s*(?:\[\d+\])?\s*;", body)
    assert fields == [f[0] for f in _lib.GridTransfer._fields_]


@pytest.mark.skipif(not os.path.exists(LIB), reason="libfem355.so not built")
def test_assembly_rejects_patterns_beyond_int32_values():
    """The assembly's CSR/BSR value index 9*s is int32: a pattern of 4x the
    500x60x75 blade (61.4M node-pair blocks there, 9*nnzb = 552.6M) is refused
    with FEM_E_LIMIT before any device work (argument check only)."""
    from fem_taichi_amd import _lib
    m = _lib.FemMesh()
    m.num_node, m.num_elem = 4 * 2322636, 4 * 2250000
    m.nnzb = 4 * 61399906
    rc = _lib.lib().fem_hex8_assemble(ctypes.byref(m), None, 2e5, 0.3, 7.8e-9, ctypes.c_void_p(64), 0, 0, None, None)
    assert rc == _lib.FEM_E_LIMIT
    assert "int32" in _lib.last_error()
    m.nnzb = 61399906                      # the blade itself passes that check (then fails on the NULL pattern)
    rc = _lib.lib().fem_hex8_assemble(ctypes.byref(m), None, 2e5, 0.3, 7.8e-9, ctypes.c_void_p(64), 0, 0, None, None)
    assert rc == _lib.FEM_E_INVALID
