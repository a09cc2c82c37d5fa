"""Tecplot output (CPU): the binary writer against a byte string assembled
here field by field from the published version-112 layout, a write/read
round trip, the reference's variable naming, and the ASCII layout of
POST_process.py:101-135.  No Tecplot reader exists in this environment, so
the binary format is pinned only against the published layout as restated
(parity unpinned beyond that restatement)."""
import struct

import numpy as np

from fem_taichi_amd import tecplot as T


def _istr(s):
    return b"".join(struct.pack("<i", ord(c)) for c in s) + struct.pack("<i", 0)


def test_one_brick_bytes(tmp_path):
    vals = np.arange(2 * 8, dtype=np.float64).reshape(2, 8) * 0.5 - 1.0
    elem = np.array([[1, 0, 3, 2, 5, 4, 7, 6]])
    p = tmp_path / "one.plt"
    T.write_plt(p, "Solution", ["x", "u1"], vals, elem)
    exp = b"#!TDV112" + struct.pack("<ii", 1, 0) + _istr("Solution") + struct.pack("<i", 2)
    exp += _istr("x") + _istr("u1")
    exp += struct.pack("<f", 299.0) + _istr("Simple Zone") + struct.pack("<ii", -1, -1) + struct.pack("<d", 360.0)
    exp += struct.pack("<11i", -1, 5, 0, 0, 0, 8, 1, 0, 0, 0, 0) + struct.pack("<f", 357.0)
    exp += struct.pack("<f", 299.0) + struct.pack("<ii", 2, 2) + struct.pack("<iii", 0, 0, -1)
    exp += struct.pack("<4d", vals[0].min(), vals[0].max(), vals[1].min(), vals[1].max())
    exp += vals.astype("<f8").tobytes() + elem.astype("<i4").tobytes()
    assert p.read_bytes() == exp


def test_round_trip_and_reference_names(tmp_path):
    rng = np.random.default_rng(0)
    N, E, S = 27, 8, 2
    xyz = rng.standard_normal((N, 3))
    fields = rng.standard_normal((N, 3, S))
    names, values = T.field_table(xyz, fields, ["u", "v", "w"])
    assert names == ["x", "y", "z", "u1", "v1", "w1", "u2", "v2", "w2"]
    assert np.array_equal(values[3:6].T, fields[:, :, 0]) and np.array_equal(values[6:9].T, fields[:, :, 1])
    elem = rng.integers(0, N, size=(E, 8))
    p = tmp_path / "rt.plt"
    T.write_plt(p, "Solution", names, values, elem)
    d = T.read_plt(p)
    assert d["var_names"] == names and d["zone_type"] == T.FEBRICK and d["solution_time"] == 360.0
    assert np.array_equal(d["values"], values) and np.array_equal(d["elem"], elem)
    assert np.array_equal(d["minmax"][:, 0], values.min(axis=1))


def test_ascii_layout(tmp_path):
    xyz = np.array([[0.0, 1.0, 2.0], [3.0, 4.0, 5.0]])
    fields = np.array([[[1.5], [-2.0]], [[0.25], [8.0]]])          # (N=2, F=2, S=1)
    elem1 = np.array([[1, 2, 1, 2, 1, 2, 1, 2]])
    p = tmp_path / "a.dat"
    T.write_dat(p, "part", xyz, fields, ["u", "v"], elem1)
    lines = p.read_text().splitlines()
    assert lines[0] == 'TITLE = "part.data"'
    # POST_process.py:125, "%s " % ',"' + name + str(step + 1) + '"', evaluated by
    # hand: ("%s " % ',"') is ',"' + ' ', then the name, the step and a quote
    assert lines[1] == 'VARIABLES = "X [ m ]","Y [ m ]","Z [m]"," u1"," v1"'
    assert lines[2] == "ZONE  I=2, J=1, F=FEPOINT, ET=BRICK"
    assert lines[3] == "0.00000e+00 1.00000e+00 2.00000e+00 1.50000e+00 -2.00000e+00"
    assert lines[5] == "1 2 1 2 1 2 1 2"
