"""PQR reading and writing for the batch path's molecules.

Restates pystarc/structures/pqr_io.py:64-199: ATOM / HETATM records parsed by
the fixed PDB columns (serial 6-11, name 12-16, resname 16-21, chain 21,
resid 22-26, x/y/z 30-54, then charge and radius, optional element), with a
whitespace fallback that detects whether a chain column is present;
REMARK / END / blank lines are skipped.  ``parse_pqr`` returns a Molecule
whose atoms keep the file order (the order rxns.xml indices refer to).
"""
from __future__ import annotations

from dataclasses import dataclass
from pathlib import Path
from typing import List, Optional

from .molecules import Atom, Molecule


@dataclass
class PQRRecord:
    record_type: str
    serial: int
    name: str
    resname: str
    chain: str
    resid: int
    x: float
    y: float
    z: float
    charge: float
    radius: float
    element: str = ""


def _is_int(s: str) -> bool:
    try:
        int(s)
        return True
    except ValueError:
        return False


def _columns(line: str, rtype: str) -> Optional[PQRRecord]:
    if len(line) < 54:
        return None
    try:
        tail = line[54:].split()
        if len(tail) < 2:
            return None
        return PQRRecord(rtype, int(line[6:11]), line[12:16].strip(), line[16:21].strip(),
                         line[21:22].strip(), int(line[22:26]), float(line[30:38]),
                         float(line[38:46]), float(line[46:54]), float(tail[0]), float(tail[1]),
                         tail[2] if len(tail) >= 3 else "")
    except (ValueError, IndexError):
        return None


def _fields(line: str, rtype: str) -> Optional[PQRRecord]:
    parts = line.split()
    if len(parts) < 10:
        return None
    try:
        if _is_int(parts[4]):
            chain, resid, off = "", int(parts[4]), 5
        elif len(parts) >= 11 and _is_int(parts[5]):
            chain, resid, off = parts[4], int(parts[5]), 6
        else:
            return None
        return PQRRecord(rtype, int(parts[1]) if _is_int(parts[1]) else 0, parts[2], parts[3],
                         chain, resid, float(parts[off]), float(parts[off + 1]),
                         float(parts[off + 2]), float(parts[off + 3]), float(parts[off + 4]),
                         parts[off + 5] if len(parts) > off + 5 else "")
    except (ValueError, IndexError):
        return None


def parse_pqr_records(path) -> List[PQRRecord]:
    out: List[PQRRecord] = []
    with open(Path(path)) as fh:
        for raw in fh:
            line = raw.rstrip("\n").rstrip("\r")
            s = line.strip()
            if not s or s.startswith("REMARK") or s.startswith("END"):
                continue
            rtype = line[:6].strip()
            if rtype not in ("ATOM", "HETATM"):
                continue
            rec = _columns(line, rtype) or _fields(line, rtype)
            if rec is not None:
                out.append(rec)
    return out


def parse_pqr(path) -> Molecule:
    path = Path(path)
    mol = Molecule(name=path.stem)
    for i, r in enumerate(parse_pqr_records(path)):
        mol.atoms.append(Atom(index=i, name=r.name, residue_name=r.resname, residue_index=r.resid,
                              chain=r.chain or "A", x=r.x, y=r.y, z=r.z, charge=r.charge,
                              radius=r.radius))
    return mol


def write_pqr(mol: Molecule, path) -> None:
    """Fixed-column records that parse_pqr_records reads back exactly (3 decimals)."""
    with open(Path(path), "w") as fh:
        for i, a in enumerate(mol.atoms):
            fh.write(f"ATOM  {i + 1:5d} {a.name[:4]:<4s} {a.residue_name[:4]:<4s}{a.chain[:1]:1s}"
                     f"{a.residue_index:4d}    {a.x:8.3f}{a.y:8.3f}{a.z:8.3f} {a.charge:8.4f} "
                     f"{a.radius:7.4f}\n")
        fh.write("END\n")
