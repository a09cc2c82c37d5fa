"""Post.Output_field_BIN / Output_field_ASC (POST_process.py:32-69, 101-135)
on the drop-in Post: the files land in Results/<name>/, the binary file
reads back to the reference's variable list with the displacements, the
stresses of Get_Stress and the mesh's 1-based connectivity minus one."""
import numpy as np
import pytest

from fem_taichi_amd import mesh

pytestmark = pytest.mark.gpu


def test_post_writes_reference_fields(gpu, tmp_path, monkeypatch):
    from fem_taichi_amd import Model, Physics, Post, Static_solver, tecplot
    monkeypatch.chdir(tmp_path)
    g = mesh.structured_block(3, 2, 4, 1.0, 0.05, 0)
    phys = Physics.from_grid(g)
    model = Model(g.name, phys, g.Materials, Omegas=(0, 0, 0))
    U = Static_solver(model, 0).Solve_cpu(tol=1e-10)
    Disp = U.reshape(phys.Num_Node, 3, 1)
    post = Post("part", phys, g.Materials, Disp)
    S = post.Get_Stress()[:, :, None]
    Fields = np.concatenate((Disp, S), axis=1)
    names = ['u', 'v', 'w', 's-xx', 's-yy', 's-zz', 's-xy', 's-yz', 's-xz', '1-prin', '2-prin', '3-prin',
             'von-mises']
    out = post.Output_field_BIN("statics", Fields, names)
    assert out.endswith("statics.plt") and (tmp_path / "Results" / "part" / "statics.plt").exists()
    d = tecplot.read_plt(out)
    assert d["var_names"] == ["x", "y", "z"] + [n + "1" for n in names]
    assert np.array_equal(d["values"][0:3].T, g.Node[1:4].T)
    assert np.array_equal(d["values"][3:6].T, Disp[:, :, 0])
    assert np.array_equal(d["values"][6:].T, S[:, :, 0])
    assert np.array_equal(d["elem"], g.Elem.T)                       # Nind[Elem] + 1 - 1, contiguous IDs
    asc = post.Output_field_ASC("statics", Fields, names)
    lines = open(asc).read().splitlines()
    assert lines[2] == f"ZONE  I={phys.Num_Node}, J={phys.Num_Elem}, F=FEPOINT, ET=BRICK"
    assert len(lines) == 3 + phys.Num_Node + phys.Num_Elem
