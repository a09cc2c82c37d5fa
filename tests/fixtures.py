"""Loaders of the committed fixtures under tests/golden."""
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
STEEL = {"EX": 2.0e5, "PO": 0.3, "DE": 7.8e-9}


def load_mesh2():
    """The reference's Mesh/2.cdb (ANSYS hex mesh, gapped node IDs) with its
    two root boundary sets, as (Node (6, N), Elem (8, E), Bounds, Forces) in
    the reference's Read layout."""
    z = np.load(os.path.join(GOLDEN, "mesh2_cdb.npz"))
    n = len(z["node_id"])
    node = np.zeros((6, n))
    node[0] = z["node_id"]
    node[3:6] = z["xyz"].T
    elem = z["elem_nodes"].T.astype(np.int64)
    bounds = [[z["bound_in"].astype(np.int64), [0, 0, 0]], [z["bound_out"].astype(np.int64), [0, 0, 0]]]
    return node, elem, bounds, []
