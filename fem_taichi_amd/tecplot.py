"""Tecplot output of nodal fields on a hex8 mesh, written from scratch.

The reference writes its results through Tecplot's closed TecIO library
(POST_process.py:32-69 calling Utilize/Dll1/dllmain.cpp:28-115, TECINI142 /
TECZNE142 / TECDAT142 / TECNODE142) or as ASCII text (POST_process.py:101-135).
Its Readme (Readme.md:186, 195-198) asks for binary ``.plt`` output.  This
module writes the binary data file format version 112 (``#!TDV112``) as
Tecplot's published data-format guide describes it, directly with numpy: no
TecIO, no compiled code.

What the reference's DLL writes, and what is written here: one FEBRICK zone
("Simple Zone", solution time 360.0, static strand, no parent), all
variables nodal, double precision, block order (each variable's N values in
turn), then the 8 node indices of every element (zero-based in the file;
the DLL hands TecIO 1-based indices, which TecIO stores zero-based).

Binary layout (all little-endian; INT32 strings are one INT32 per character
followed by a 0):

  header   "#!TDV112", INT32 1, INT32 file type (0 = full), title,
           INT32 NumVar, NumVar names;
  zone     FLOAT32 299.0, zone name, INT32 parent zone (-1), INT32 strand
           (-1 = static), FLOAT64 solution time, INT32 -1 (unused),
           INT32 zone type (5 = FEBRICK), INT32 0 (all variables nodal),
           INT32 0 (no raw face neighbours), INT32 0 (no user face
           neighbours), INT32 NumPts, INT32 NumElements, INT32 x 3 cell
           dimensions (0), INT32 0 (no auxiliary data);
  end      FLOAT32 357.0 (end of header);
  data     FLOAT32 299.0, INT32 x NumVar data format (2 = double), INT32 0
           (no passive variables), INT32 0 (no variable sharing), INT32 -1
           (no connectivity sharing), FLOAT64 (min, max) per variable, the
           variables' values (FLOAT64, block), INT32 x 8 E connectivity.
"""
from __future__ import annotations

import os

import numpy as np

ZONE_MARKER = 299.0
EOH_MARKER = 357.0
FEBRICK = 5
SOLUTION_TIME = 360.0          # dllmain.cpp: solTime


def _istr(s: str) -> np.ndarray:
    """A string as the format's INT32 characters plus the terminating 0."""
    return np.array([ord(c) for c in s] + [0], dtype="<i4")


def write_plt(path, title, var_names, values, elem, zone_name="Simple Zone"):
    """Write one FEBRICK zone.  values (V, N) float64, variable-major;
    elem (E, 8) zero-based node indices."""
    values = np.ascontiguousarray(values, dtype="<f8")
    elem = np.ascontiguousarray(elem, dtype="<i4")
    V, N = values.shape
    if len(var_names) != V:
        raise ValueError(f"{len(var_names)} variable names for {V} variables")
    if elem.ndim != 2 or elem.shape[1] != 8:
        raise ValueError("elem must be (E, 8)")
    if elem.size and (elem.min() < 0 or elem.max() >= N):
        raise ValueError("connectivity refers to a node outside the zone")
    E = elem.shape[0]
    with open(path, "wb") as f:
        f.write(b"#!TDV112")
        np.array([1, 0], dtype="<i4").tofile(f)                      # byte order, FULL
        _istr(title).tofile(f)
        np.array([V], dtype="<i4").tofile(f)
        for name in var_names:
            _istr(name).tofile(f)
        np.array([ZONE_MARKER], dtype="<f4").tofile(f)
        _istr(zone_name).tofile(f)
        np.array([-1, -1], dtype="<i4").tofile(f)                    # parent zone, static strand
        np.array([SOLUTION_TIME], dtype="<f8").tofile(f)
        np.array([-1, FEBRICK, 0, 0, 0, N, E, 0, 0, 0, 0], dtype="<i4").tofile(f)
        np.array([EOH_MARKER], dtype="<f4").tofile(f)
        np.array([ZONE_MARKER], dtype="<f4").tofile(f)
        np.full(V, 2, dtype="<i4").tofile(f)                         # double
        np.array([0, 0, -1], dtype="<i4").tofile(f)                  # passive, sharing, connectivity sharing
        mm = np.empty((V, 2), dtype="<f8")
        if N:
            mm[:, 0], mm[:, 1] = values.min(axis=1), values.max(axis=1)
        else:
            mm[:] = 0.0
        mm.tofile(f)
        values.tofile(f)
        elem.tofile(f)


def read_plt(path):
    """Parse a file written by write_plt (one FE zone of double variables):
    returns dict(title, var_names, zone_name, zone_type, solution_time,
    values (V, N), minmax (V, 2), elem (E, 8))."""
    buf = open(path, "rb").read()
    pos = [0]

    def take(dtype, count=1):
        dt = np.dtype(dtype)
        a = np.frombuffer(buf, dtype=dt, count=count, offset=pos[0])
        pos[0] += dt.itemsize * count
        return a

    def istr():
        chars = []
        while True:
            c = int(take("<i4")[0])
            if c == 0:
                return "".join(chars)
            chars.append(chr(c))

    if buf[:8] != b"#!TDV112":
        raise ValueError("not a version 112 Tecplot binary file")
    pos[0] = 8
    if int(take("<i4")[0]) != 1:
        raise ValueError("byte order mark is not 1")
    file_type = int(take("<i4")[0])
    title = istr()
    V = int(take("<i4")[0])
    names = [istr() for _ in range(V)]
    if float(take("<f4")[0]) != ZONE_MARKER:
        raise ValueError("zone marker expected")
    zone_name = istr()
    parent, strand = (int(v) for v in take("<i4", 2))
    sol_time = float(take("<f8")[0])
    unused, ztype, varloc, raw_nb, misc_nb, N, E, c1, c2, c3, aux = (int(v) for v in take("<i4", 11))
    if float(take("<f4")[0]) != EOH_MARKER:
        raise ValueError("end-of-header marker expected")
    if float(take("<f4")[0]) != ZONE_MARKER:
        raise ValueError("data-section zone marker expected")
    fmts = take("<i4", V)
    if np.any(fmts != 2):
        raise ValueError("only double variables are read")
    passive, sharing, conn_share = (int(v) for v in take("<i4", 3))
    minmax = take("<f8", 2 * V).reshape(V, 2)
    values = take("<f8", V * N).reshape(V, N)
    elem = take("<i4", 8 * E).reshape(E, 8)
    if pos[0] != len(buf):
        raise ValueError(f"{len(buf) - pos[0]} trailing bytes")
    return dict(file_type=file_type, title=title, var_names=names, zone_name=zone_name, parent=parent,
                strand=strand, solution_time=sol_time, unused=unused, zone_type=ztype, var_location=varloc,
                raw_face_neighbours=raw_nb, misc_face_neighbours=misc_nb, cell_dims=(c1, c2, c3), aux=aux,
                passive=passive, sharing=sharing, connectivity_sharing=conn_share, values=values.copy(),
                minmax=minmax.copy(), elem=elem.copy())


def field_table(node_xyz, fields, field_names):
    """The reference's variable list and values (POST_process.py:37-46) for
    node_xyz (N, 3) and fields (N, F, S): x, y, z, then for each step i every
    field name with the suffix i + 1.  Returns (names, values (3 + F S, N))."""
    fields = np.asarray(fields, dtype=np.float64)
    if fields.ndim != 3:
        raise ValueError("fields must be (N, F, S)")
    N, F, S = fields.shape
    if len(field_names) != F:
        raise ValueError(f"{len(field_names)} names for {F} fields")
    names = ["x", "y", "z"] + [f"{nm}{i + 1}" for i in range(S) for nm in field_names]
    xyz = np.asarray(node_xyz, dtype=np.float64)
    if xyz.shape != (N, 3):
        raise ValueError(f"node_xyz must be ({N}, 3)")
    data = np.concatenate([xyz, np.transpose(fields, (0, 2, 1)).reshape(N, F * S)], axis=1)
    return names, np.ascontiguousarray(data.T)


def write_dat(path, title, node_xyz, fields, field_names, elem1):
    """The reference's ASCII form (POST_process.py:101-135): a TITLE line, the
    VARIABLES line as it writes it, a FEPOINT BRICK zone line, the node rows
    (%10.5e, space separated) and the 1-based element rows.

    POST_process.py:125 writes ``"%s " % ',"' + name + str(step + 1) + '"'``:
    ``%`` binds tighter than ``+``, so each field comes out as ``," u1"`` (the
    space after the comma-quote, none after the name), and that is what is
    written here."""
    fields = np.asarray(fields, dtype=np.float64)
    N, F, S = fields.shape
    data = np.concatenate([np.asarray(node_xyz, dtype=np.float64),
                           np.transpose(fields, (0, 2, 1)).reshape(N, F * S)], axis=1)
    with open(path, "w") as f:
        f.write('TITLE = "%s.data"\n' % title)
        f.write('VARIABLES = "X [ m ]","Y [ m ]","Z [m]"')
        for step in range(S):
            for name in field_names:
                f.write(('%s ' % ',"') + str(name) + str(step + 1) + '"')
        f.write("\n")
        f.write("ZONE  I=%d, J=%d, F=FEPOINT, ET=BRICK\n" % (N, elem1.shape[0]))
        np.savetxt(f, data, fmt="%10.5e", delimiter=" ")
        np.savetxt(f, np.asarray(elem1, dtype=np.int64), fmt="%d", delimiter=" ")


def results_dir(name):
    """Results/<name>/ (POST_process.py:20-22), created on demand."""
    path = os.path.join("Results", name)
    os.makedirs(path, exist_ok=True)
    return path
