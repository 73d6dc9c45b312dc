"""OpenDX grid files (the APBS potential format) as the reference reads them
(pystarc/forces/electrostatic/grid_force.py:110-143): header, values only
inside the array object, z fastest.  The 3x3x3 file below holds the value
i + j + k at node (i, j, k), so a wrong axis order or a dropped line fails
the node-by-node check; the reference's own test of the reader uses a file of
the same shape (tests/test_pystarc.py:971-991)."""
import numpy as np
import pytest

from pystarc_amd.forces.electrostatic.grid_force import DXGrid

SUM_GRID = """# APBS-style header
object 1 class gridpositions counts 3 3 3
origin 0.000 0.000 0.000
delta 1.000 0.000 0.000
delta 0.000 1.000 0.000
delta 0.000 0.000 1.000
object 2 class gridconnections counts 3 3 3
object 3 class array type double rank 0 items 27 data follows
0.0 1.0 2.0 1.0 2.0 3.0 2.0 3.0 4.0
1.0 2.0 3.0 2.0 3.0 4.0 3.0 4.0 5.0
2.0 3.0 4.0 3.0 4.0 5.0 4.0 5.0 6.0
attribute "dep" string "positions"
object "regular positions regular connections" class field
component "positions" value 1
component "connections" value 2
component "data" value 3
"""


def test_reads_node_values_in_c_order(tmp_path):
    p = tmp_path / "sum.dx"
    p.write_text(SUM_GRID)
    g = DXGrid.from_file(p)
    assert g.data.shape == (3, 3, 3)
    i, j, k = np.meshgrid(np.arange(3), np.arange(3), np.arange(3), indexing="ij")
    assert np.array_equal(g.data, (i + j + k).astype(float))
    assert np.array_equal(g.origin, [0, 0, 0]) and np.array_equal(g.delta, np.eye(3))
    assert np.array_equal(g.spacing, [1, 1, 1])


def test_anisotropic_header(tmp_path):
    text = SUM_GRID.replace("origin 0.000 0.000 0.000", "origin -1.5 2.25 -30.0") \
                   .replace("delta 0.000 1.000 0.000", "delta 0.000 0.500 0.000")
    p = tmp_path / "a.dx"
    p.write_text(text)
    g = DXGrid.from_file(p)
    assert np.array_equal(g.origin, [-1.5, 2.25, -30.0])
    assert np.array_equal(g.spacing, [1.0, 0.5, 1.0])


def test_round_trip_exact(tmp_path):
    rng = np.random.default_rng(5)
    d = rng.normal(size=(4, 5, 7)) * 1e3      # 140 values: the last line holds 2
    g = DXGrid([1.0, -2.0, 0.5], np.diag([0.5, 0.5, 0.5]), d)
    p = tmp_path / "r.dx"
    g.to_file(p, fmt="%.17g")
    h = DXGrid.from_file(p)
    assert np.array_equal(h.data, d)
    assert np.array_equal(h.origin, g.origin) and np.array_equal(h.delta, g.delta)
    # the default 7-digit format keeps the values to 7 significant digits
    g.to_file(p)
    assert np.allclose(DXGrid.from_file(p).data, d, rtol=1e-6, atol=0)


@pytest.mark.parametrize("edit", ["drop_line", "extra_value"])
def test_value_count_mismatch_is_an_error(tmp_path, edit):
    lines = SUM_GRID.splitlines()
    k = lines.index("0.0 1.0 2.0 1.0 2.0 3.0 2.0 3.0 4.0")
    if edit == "drop_line":
        del lines[k]
    else:
        lines[k] += " 9.0"
    p = tmp_path / "bad.dx"
    p.write_text("\n".join(lines) + "\n")
    with pytest.raises(ValueError):
        DXGrid.from_file(p)


def test_values_after_the_array_object_are_not_data(tmp_path):
    # numbers on lines after the array ends (field components) are ignored
    p = tmp_path / "t.dx"
    p.write_text(SUM_GRID + "object 5 class array items 2 data follows\n1.0 2.0\n")
    assert DXGrid.from_file(p).data.shape == (3, 3, 3)


def test_not_a_grid_file(tmp_path):
    p = tmp_path / "n.dx"
    p.write_text("# nothing here\n1 2 3\n")
    with pytest.raises(ValueError):
        DXGrid.from_file(p)
