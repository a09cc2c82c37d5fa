"""Readers of the reference's mesh and boundary files (READ_physics.py:63-137).

``read_cdb`` reads an ANSYS ``.cdb`` (NBLOCK ``(3i9,6e20.13)`` node lines
between the ``NBLOCK`` header and ``N,R5``; EBLOCK ``(19i8)`` element lines up
to the ``-1`` terminator) into the reference's arrays:

* ``Node`` (6, N) float64: rows [id, 0, 0, x, y, z] (READ_physics.py:79-83, 96);
* ``Elem`` (18, E) int64 = tokens - 1, whose last 8 rows are node ID - 1
  (READ_physics.py:86-89, 95).

``read_lis`` reads an ANSYS component listing (``.lis``): every line without
NAME / NODE / LIST holds node IDs (READ_physics.py:105-114).  ``read_model``
mirrors ``Read(name, bounds_file, forces_file)``: the mesh ``<name>.cdb``,
boundary sets ``<name>_<prefix>*.<ext>`` -> ``Bounds`` = [[ids - 1, [0, 0, 0]],
...] (all DOFs fixed, READ_physics.py:114) and likewise ``Forces``.

Host-side file parsing, like the reference's (pandas there, plain Python
here); the result feeds ``Physics``.
"""
from __future__ import annotations

import glob
import os

import numpy as np


def _node_values(line):
    # (3i9, 6e20.13): fixed columns; values can touch ("...E+01-5.12...")
    vals = [line[0:9], line[9:18], line[18:27]]
    vals += [line[27 + 20 * k: 47 + 20 * k] for k in range(3)]
    return [float(v) for v in vals]


def read_cdb(path):
    """(Node (6, N) float64, Elem (18, E) int64) of an ANSYS .cdb file."""
    with open(path, "r", errors="replace") as f:
        lines = f.read().splitlines()
    n_sta = next(i for i, s in enumerate(lines) if "NBLOCK" in s) + 2
    n_end = next(i for i, s in enumerate(lines) if i >= n_sta and "N,R5" in s)
    e_sta = next(i for i, s in enumerate(lines) if "EBLOCK" in s) + 2
    e_end = next(i for i, s in enumerate(lines) if i >= e_sta and s.strip().startswith("-1"))
    node = np.array([_node_values(s) for s in lines[n_sta:n_end] if s.strip()], dtype=np.float64)
    toks = " ".join(lines[e_sta:e_end]).split()
    if len(toks) % 18:
        raise ValueError(f"{path}: EBLOCK does not hold 18 values per element")
    elem = np.array(toks, dtype=np.int64).reshape(-1, 18)
    return node.T.copy(), elem.T - 1


def read_lis(path):
    """Node IDs - 1 listed in an ANSYS component .lis file."""
    ids = []
    with open(path, "r", errors="replace") as f:
        for s in f:
            if any(key in s for key in ("NAME", "NODE", "LIST")):
                continue
            ids.extend(s.split())
    return np.array(ids, dtype=np.int64) - 1


def _sets(directory, name, prefix, ext):
    out = []
    for path in sorted(glob.glob(os.path.join(directory, "*." + ext))):
        if f"{name}_{prefix}" in os.path.basename(path):
            out.append([read_lis(path), [0, 0, 0]])
    return out


def read_model(name, mesh_dir="Mesh", bounds_dir="Boundary", bounds_file=("bot", "lis"),
               forces_file=("top", "lis")):
    """(Node, Elem, Bounds, Forces) as ``Read(name, ...)`` stages them."""
    node, elem = read_cdb(os.path.join(mesh_dir, f"{name}.cdb"))
    bounds = _sets(bounds_dir, name, *bounds_file)
    forces = _sets(bounds_dir, name, *forces_file)
    return node, elem, bounds, forces
