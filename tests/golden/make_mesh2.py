"""Writes tests/golden/mesh2_cdb.npz from the reference's own mesh files.

Input (read as data, in this container only): Mesh/2.cdb and
Boundary/2_bot_in.lis, Boundary/2_bot_out.lis of the reference, parsed with
fem_taichi_amd.read_cdb.  The fixture keeps what Physics needs: node IDs and
coordinates (Node rows 0, 3, 4, 5), the element node IDs - 1 (Elem's last 8
rows) and the two boundary sets (node IDs - 1).

    python tests/golden/make_mesh2.py [/root/reference]
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from fem_taichi_amd.read_cdb import read_model  # noqa: E402


def main(ref="/root/reference"):
    node, elem, bounds, forces = read_model("2", os.path.join(ref, "Mesh"), os.path.join(ref, "Boundary"))
    assert not forces
    np.savez_compressed(os.path.join(HERE, "mesh2_cdb.npz"),
                        node_id=node[0].astype(np.int32), xyz=node[3:6].T.copy(),
                        elem_nodes=elem[-8:].T.astype(np.int32).copy(),
                        bound_in=bounds[0][0].astype(np.int32), bound_out=bounds[1][0].astype(np.int32))
    print(f"mesh2_cdb.npz: {node.shape[1]} nodes, {elem.shape[1]} elements, "
          f"bounds {[len(b[0]) for b in bounds]}")


if __name__ == "__main__":
    main(*sys.argv[1:])
