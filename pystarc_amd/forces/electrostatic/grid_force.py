"""OpenDX scalar grids (the APBS output format), as
pystarc/forces/electrostatic/grid_force.py:90-145 reads them: ``counts``,
``origin``, three ``delta`` rows, then values with z varying fastest, held as
``data`` (nx, ny, nz) in C order."""
from __future__ import annotations

from pathlib import Path

import numpy as np


class DXGrid:
    def __init__(self, origin, delta, data):
        self.origin = np.asarray(origin, dtype=float)
        self.delta = np.asarray(delta, dtype=float)
        self.data = np.asarray(data)

    @property
    def spacing(self):
        return np.array([self.delta[i, i] for i in range(3)])

    @classmethod
    def from_file(cls, path):
        """Header lines anywhere before the data (``object 1 ... counts``,
        ``origin``, three ``delta`` rows); values only inside the array object
        (after the ``object 3`` line, up to the next ``object`` or
        ``attribute`` line), as grid_force.py:110-143 reads them.  A value
        count other than nx ny nz is an error, as the reference's reshape
        makes it."""
        counts = origin = None
        deltas = []
        values = []
        in_data = False
        with open(Path(path), "r") as fh:
            for line in fh:
                s = line.strip()
                if not s or s.startswith("#"):
                    continue
                if in_data:
                    if s.startswith("object") or s.startswith("attribute"):
                        break
                    values.extend(float(v) for v in s.split())
                elif s.startswith("object 1"):
                    counts = tuple(int(v) for v in s.split()[-3:])
                elif s.startswith("origin"):
                    origin = [float(v) for v in s.split()[1:4]]
                elif s.startswith("delta"):
                    deltas.append([float(v) for v in s.split()[1:4]])
                elif s.startswith("object 3"):
                    in_data = True
        if counts is None or origin is None or len(deltas) != 3:
            raise ValueError(f"{path}: not an OpenDX gridpositions file")
        n = counts[0] * counts[1] * counts[2]
        if len(values) != n:
            raise ValueError(f"{path}: {len(values)} values for {n} grid points")
        data = np.asarray(values, dtype=np.float64).reshape(counts)
        return cls(origin, np.array(deltas), data)

    def to_file(self, path, fmt="%.6e"):
        nx, ny, nz = self.data.shape
        flat = np.asarray(self.data, dtype=np.float64).ravel()
        with open(Path(path), "w") as fh:
            fh.write(f"object 1 class gridpositions counts {nx} {ny} {nz}\n")
            # geometry as shortest round-trip decimals: a grid read back sits
            # exactly where it was written
            fh.write("origin " + " ".join(repr(float(v)) for v in self.origin) + "\n")
            for row in self.delta:
                fh.write("delta " + " ".join(repr(float(v)) for v in row) + "\n")
            fh.write(f"object 2 class gridconnections counts {nx} {ny} {nz}\n")
            fh.write(f"object 3 class array type double rank 0 items {flat.size} data follows\n")
            full = flat.size - flat.size % 3
            if full:
                np.savetxt(fh, flat[:full].reshape(-1, 3), fmt=fmt)
            if flat.size % 3:
                fh.write(" ".join(fmt % v for v in flat[full:]) + "\n")
            fh.write('attribute "dep" string "positions"\n')
            fh.write('object "regular positions regular connections" class field\n')
            fh.write('component "positions" value 1\ncomponent "connections" value 2\n')
            fh.write('component "data" value 3\n')
