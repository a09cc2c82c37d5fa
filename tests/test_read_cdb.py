"""The .cdb / .lis readers (READ_physics.py:63-137) and the mesh2 fixture.  CPU only."""
import os

import numpy as np
import pytest
import scipy.sparse as sp

from fem_taichi_amd import mesh
from fem_taichi_amd.read_cdb import read_cdb, read_lis, read_model
from oracle import fem_oracle as O
from fixtures import load_mesh2

REF = "/root/reference"


def write_cdb(path, g):
    """An ANSYS-format .cdb of a StructuredGrid, the way ANSA writes it."""
    with open(path, "w") as f:
        f.write("/PREP7\nNBLOCK,6,SOLID\n(3i9,6e20.13)\n")
        for i in range(g.Node.shape[1]):
            x, y, z = g.Node[-3:, i]
            f.write(f"{int(g.Node[0, i]):9d}{0:9d}{0:9d}{x:20.13E}{y:20.13E}{z:20.13E}\n")
        f.write("N,R5.3,LOC,-1\nET,1,185\nEBLOCK,19,SOLID\n(19i8)\n")
        for e in range(g.Elem.shape[1]):
            nodes = "".join(f"{int(v) + 1:8d}" for v in g.Elem[-8:, e])
            f.write(f"{2:8d}{1:8d}{0:8d}{0:8d}{'':8s}{0:8d}{0:8d}{0:8d}{8:8d}{0:8d}{e + 1:8d}{nodes}\n")
        f.write("      -1\nFINISH\n")


def test_read_cdb_round_trip(tmp_path):
    g = mesh.structured_block(3, 2, 4, 10.0, 0.1, 3)
    g.Node[-3:, 5] *= -1.0          # negative values touching the previous column
    p = tmp_path / "t.cdb"
    write_cdb(p, g)
    node, elem = read_cdb(str(p))
    assert node.shape == (6, g.Node.shape[1]) and elem.shape == (18, g.Elem.shape[1])
    np.testing.assert_array_equal(node[0], g.Node[0])
    np.testing.assert_allclose(node[3:6], g.Node[-3:], rtol=1e-12, atol=1e-12)
    np.testing.assert_array_equal(elem[-8:], g.Elem[-8:])
    np.testing.assert_array_equal(elem[9], np.arange(g.Elem.shape[1]))   # element IDs - 1


def test_read_lis(tmp_path):
    p = tmp_path / "m_bot.lis"
    p.write_text(" LIST ALL SELECTED COMPONENTS\n\n    NAME   TYPE\n  NODE LIST\n   12   7\n 3\n")
    np.testing.assert_array_equal(read_lis(str(p)), [11, 6, 2])


@pytest.mark.skipif(not os.path.exists(os.path.join(REF, "Mesh", "2.cdb")), reason="reference files absent")
def test_fixture_equals_reference_mesh():
    node, elem, bounds, forces = read_model("2", os.path.join(REF, "Mesh"), os.path.join(REF, "Boundary"))
    n2, e2, b2, f2 = load_mesh2()
    np.testing.assert_array_equal(node[0], n2[0])
    np.testing.assert_array_equal(node[3:6], n2[3:6])
    np.testing.assert_array_equal(elem[-8:], e2)
    for (a, _), (b, _) in zip(bounds, b2):
        np.testing.assert_array_equal(a, b)
    assert forces == f2 == []


def test_mesh2_pattern_brute_force():
    """Gapped IDs (max 10731 for 9468 nodes): the oracle's pattern equals the
    brute-force neighbour sets and is symmetric with the diagonal present."""
    Node, Elem, _, _ = load_mesh2()
    pp = O.pre_process(Node, Elem)
    assert pp["Num_MaxN"] == 10731 and pp["Num_Node"] == 9468
    en = O.element_nodes(Elem, pp["Nind"])
    N = Node.shape[1]
    rows = np.repeat(np.arange(len(en)), 8)
    A = sp.csr_matrix((np.ones(en.size), (rows, en.reshape(-1))), shape=(len(en), N))
    nn = (A.T @ A).tocsr()
    nn.sort_indices()
    np.testing.assert_array_equal(pp["SmNode"], nn.indptr)
    np.testing.assert_array_equal(pp["ReNode"], nn.indices)
    for i in (0, 77, 4000, N - 1):
        row = set(pp["ReNode"][pp["SmNode"][i]:pp["SmNode"][i + 1]])
        brute = set(en[np.any(en == i, axis=1)].reshape(-1))
        assert row == brute and i in row
