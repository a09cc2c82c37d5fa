"""Synthetic structured hex8 meshes for the benchmark and test workloads.

A structured grid of ``nx x ny x nz`` elements has ``(nx+1)(ny+1)(nz+1)``
nodes numbered x-fastest: ``n(i, j, k) = i + (nx+1) * (j + (ny+1) * k)``,
with 1-based contiguous node IDs ``n + 1``.  Element connectivity follows the
ANSYS I..P order the reference's shape functions assume
(FEM_model_csr.py:357-364): bottom face counter-clockwise seen from +z
starting at the (+x, +y) corner, then the top face in the same order, so
``det J > 0`` for an undistorted element.

Arrays follow the reference's ``Read`` duck type (READ_physics.py:44-61,
96-97, 114, 134): ``Node`` is (4, N) = [ID, x, y, z]; ``Elem`` is (8, E) =
node ID - 1; ``Bounds`` and ``Forces`` are lists of ``[ids - 1, [c0, c1, c2]]``.
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np


@dataclasses.dataclass
class StructuredGrid:
    """A structured hex8 mesh plus its boundary sets and material field."""

    nx: int
    ny: int
    nz: int
    Node: np.ndarray              # (4, N) [ID, x, y, z]
    Elem: np.ndarray              # (8, E) node ID - 1
    Bounds: list
    Forces: list
    EX: object                    # scalar or (E,) per-element Young's modulus
    PO: float = 0.3
    DE: float = 7.8e-9
    Omegas: tuple = (0.0, 0.0, 0.0)
    name: str = "grid"

    @property
    def Num_Node(self) -> int:
        return self.Node.shape[-1]

    @property
    def Num_Elem(self) -> int:
        return self.Elem.shape[-1]

    @property
    def node_dims(self):
        return (self.nx + 1, self.ny + 1, self.nz + 1)

    @property
    def Materials(self) -> dict:
        return {'DE': self.DE, 'EX': self.EX, 'PO': self.PO}


def grid_connectivity(nx: int, ny: int, nz: int) -> np.ndarray:
    """(E, 8) node indices of a structured grid, elements x-fastest."""
    ex, ey, ez = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing='ij')
    ex = ex.transpose(2, 1, 0).ravel()
    ey = ey.transpose(2, 1, 0).ravel()
    ez = ez.transpose(2, 1, 0).ravel()
    sx, sy = nx + 1, (nx + 1) * (ny + 1)

    def n(i, j, k):
        return i + sx * j + sy * k

    # corner order (+,+,-), (-,+,-), (-,-,-), (+,-,-), then the same with +z.
    cols = [n(ex + 1, ey + 1, ez), n(ex, ey + 1, ez), n(ex, ey, ez), n(ex + 1, ey, ez),
            n(ex + 1, ey + 1, ez + 1), n(ex, ey + 1, ez + 1), n(ex, ey, ez + 1), n(ex + 1, ey, ez + 1)]
    return np.stack(cols, axis=1).astype(np.int64)


def grid_coords(nx, ny, nz, h=1.0):
    i, j, k = np.meshgrid(np.arange(nx + 1), np.arange(ny + 1), np.arange(nz + 1), indexing='ij')
    xyz = np.stack([i.transpose(2, 1, 0).ravel(), j.transpose(2, 1, 0).ravel(),
                    k.transpose(2, 1, 0).ravel()], axis=1).astype(np.float64)
    return xyz * h


def is_grid_mesh(elem_idx, num_node, dims) -> bool:
    """True when (E, 8) node indices over ``num_node`` nodes are exactly the
    grid of ``dims`` = (nx, ny, nz) elements in the numbering of
    grid_connectivity (node ids x-fastest, elements x-fastest, the I..P
    corner order)."""
    nx, ny, nz = (int(d) for d in dims)
    if min(nx, ny, nz) < 1 or num_node != (nx + 1) * (ny + 1) * (nz + 1):
        return False
    elem_idx = np.asarray(elem_idx)
    if elem_idx.shape != (nx * ny * nz, 8):
        return False
    return bool(np.array_equal(elem_idx, grid_connectivity(nx, ny, nz)))


def structured_block(nx, ny, nz, h=1.0, jitter=0.0, seed=0, E=200000.0, E_sigma=0.0,
                     nu=0.3, rho=7.8e-9, fix_face='z0', load_face='zmax', total_load=1000.0,
                     twist_deg=0.0, omegas=(0.0, 0.0, 0.0), name='grid') -> StructuredGrid:
    """Build a jittered (optionally twisted) structured hex8 block.

    Coordinates are ``(i, j, k) * h + U(-jitter*h, +jitter*h)^3`` drawn from
    ``np.random.default_rng(seed)``; a per-element modulus
    ``E * exp(E_sigma * z - E_sigma**2 / 2)``, ``z ~ N(0, 1)`` (arithmetic mean
    ``E``) is drawn afterwards from the same generator when ``E_sigma > 0``.
    ``twist_deg`` rotates (y, z) about the long x axis through the section
    centre by ``twist_deg * i / nx``.
    """
    rng = np.random.default_rng(seed)
    xyz = grid_coords(nx, ny, nz, h)
    N = xyz.shape[0]
    if jitter:
        xyz = xyz + rng.uniform(-jitter * h, jitter * h, size=(N, 3))
    if twist_deg:
        ii = np.arange(N) % (nx + 1)
        th = np.deg2rad(twist_deg) * ii / nx
        yc, zc = ny * h / 2.0, nz * h / 2.0
        y, z = xyz[:, 1] - yc, xyz[:, 2] - zc
        c, s = np.cos(th), np.sin(th)
        xyz[:, 1] = yc + c * y - s * z
        xyz[:, 2] = zc + s * y + c * z
    en = grid_connectivity(nx, ny, nz)
    Ecount = en.shape[0]
    if E_sigma:
        EX = E * np.exp(E_sigma * rng.standard_normal(Ecount) - 0.5 * E_sigma ** 2)
    else:
        EX = float(E)
    ids = np.arange(1, N + 1, dtype=np.float64)
    Node = np.concatenate([ids[None, :], xyz.T], axis=0)
    Elem = en.T.copy()                      # node index == ID - 1 here
    node_ijk = np.stack([np.arange(N) % (nx + 1), (np.arange(N) // (nx + 1)) % (ny + 1),
                         np.arange(N) // ((nx + 1) * (ny + 1))], axis=1)
    faces = {'z0': node_ijk[:, 2] == 0, 'zmax': node_ijk[:, 2] == nz,
             'x0': node_ijk[:, 0] == 0, 'xmax': node_ijk[:, 0] == nx}
    Bounds, Forces = [], []
    if fix_face:
        Bounds.append([np.nonzero(faces[fix_face])[0].astype(np.int64), [0, 0, 0]])
    if load_face and total_load:
        sel = np.nonzero(faces[load_face])[0].astype(np.int64)
        Forces.append([sel, [0.0, 0.0, -float(total_load) / len(sel)]])
    return StructuredGrid(nx, ny, nz, Node, Elem, Bounds, Forces, EX, nu, rho,
                          tuple(float(w) for w in omegas), name)


# Workloads named by what they are (element grid, jitter, seed, material, loads).
def cantilever_20x20x50(**kw):
    """Cantilever block 20x20x50 elements, 1 mm, jitter 5 % (seed 0), z=0 fixed,
    -1000 N spread over the z=max face, no spin."""
    return structured_block(20, 20, 50, 1.0, 0.05, 0, name='cantilever_20x20x50', **kw)


def slab_60x60x55(**kw):
    """Slab 60x60x55 elements, jitter 10 % (seed 1), lognormal E (sigma 0.3)."""
    return structured_block(60, 60, 55, 1.0, 0.10, 1, E_sigma=0.3, name='slab_60x60x55', **kw)


def twisted_blade(nx=400, ny=50, nz=50, seed=2, **kw):
    """Rotating twisted slab (long axis x, 60 deg sweep), root face x=0 fixed,
    omega = (0, 0, 100) rad/s centrifugal load, no nodal load."""
    return structured_block(nx, ny, nz, 1.0, 0.05, seed, fix_face='x0', load_face=None,
                            total_load=0.0, twist_deg=60.0, omegas=(0.0, 0.0, 100.0),
                            name=f'twisted_blade_{nx}x{ny}x{nz}', **kw)


def training_slab(seed, nx=60, ny=60, nz=55, **kw):
    """One learned-preconditioner training slab: lognormal E (sigma 0.5), jitter 10 %."""
    return structured_block(nx, ny, nz, 1.0, 0.10, seed, E_sigma=0.5, name=f'train_slab_{seed}', **kw)


def permute_ids(grid: StructuredGrid, seed=0, gap=3):
    """Return (Node, Elem) with shuffled, gapped 1-based node IDs and shuffled
    node columns -- the situation of a real .cdb mesh (non-contiguous IDs)."""
    rng = np.random.default_rng(seed)
    N = grid.Num_Node
    new_ids = np.sort(rng.choice(N * gap, size=N, replace=False)) + 1
    rng.shuffle(new_ids)
    old_to_new = new_ids                                  # old index -> new ID
    perm = rng.permutation(N)                             # column order of Node
    Node = np.concatenate([old_to_new[perm][None, :].astype(np.float64), grid.Node[1:, perm]], axis=0)
    Elem = (old_to_new[grid.Elem.astype(np.int64)] - 1).astype(np.int64)
    return Node, Elem


def hex_volume_estimate(nx, ny, nz, h=1.0):
    return nx * ny * nz * h ** 3


def nnzb_structured(nx, ny, nz):
    """Closed form node-pair block count of a structured grid."""
    return (3 * (nx + 1) - 2) * (3 * (ny + 1) - 2) * (3 * (nz + 1) - 2)


def isclose(a, b):
    return math.isclose(a, b, rel_tol=1e-12)
