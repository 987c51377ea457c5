"""Oracle restatement of the reference's molecule input path (pure Python).

TEST INFRASTRUCTURE ONLY: used by tests/ to check the engine's own C++
readers and type table; never imported by cosims_amd.

Restates:
  molecule2::assignDefaultFF   src/molecule2.cpp:43-129   (He-atom LJ table)
  molecule2::importFile        src/molecule2.cpp:132-192  (atoms.dat overrides)
  molecule2::importMFJ         src/molecule2.cpp:442-647
  molecule2::importPDB         src/molecule2.cpp:649-768
  setMolProperties type lookup src/molecule2.cpp:782-839
"""
from __future__ import annotations


def default_table():
    """[name, symbol, mass, sigma, epsilon, intMass] rows, reference order."""
    return [
        ["Hydrogen", "H", 1.0079, 2.38, 0.01498936, 1],
        ["Carbon", "C", 12.010, 3.043, 0.03090114, 12],
        ["Nitrogen", "N", 14.006, 3.043, 0.03090114, 14],
        ["Oxygen", "O", 15.999, 3.043, 0.03090114, 16],
        ["Sodium", "Na", 22.98977, 3.043, 0.03090114, 23],
        ["Silicon", "Si", 28.086, 3.5, 0.03113175, 28],
        ["Phosphorus", "P", 30.97396, 3.5, 0.03113175, 31],
        ["Sulfur", "S", 32.07, 3.5, 0.03113175, 32],
        ["Iron", "Fe", 55.84, 3.5, 0.03113175, 56],
    ]


def _atof(s: str) -> float:
    # C atof: longest numeric prefix, 0.0 when none
    import re
    m = re.match(r"\s*[+-]?(\d+\.?\d*([eE][+-]?\d+)?|\.\d+([eE][+-]?\d+)?)", s)
    return float(m.group(0)) if m else 0.0


def _atoi(s: str) -> int:
    import re
    m = re.match(r"\s*[+-]?\d+", s)
    return int(m.group(0)) if m else 0


def apply_atoms_file(table, path):
    """importFile: whitespace tokens in groups of six; '#'-names skipped."""
    toks = open(path).read().split()
    for i in range(0, len(toks) - 5, 6):
        name, symbol, mass, sigma, eps, imass = toks[i:i + 6]
        name = name.lower()
        if name[0] == "#":
            continue
        idx = len(table)
        for j, row in enumerate(table):
            if symbol == row[1] or row[5] == _atoi(imass):
                idx = j
                break
        if idx == len(table):
            table.append([None] * 6)
        table[idx] = [name, symbol, _atof(mass), _atof(sigma), _atof(eps), _atoi(imass)]
    return table


def read_mfj(path, table):
    """importMFJ: returns dict(symbols, xyz, charge, charged, total_charge)."""
    lines = open(path).read().split("\n")
    hdr = ["".join(l.split()) for l in lines[:6]]
    unit = hdr[3].lower()
    ctype = hdr[4]
    charged, uniform, calc, total = False, False, False, 0.0
    if ctype == "" or ctype == "none":
        charged = False
    elif ctype == "equal":
        charged, uniform, total = True, True, 1.0
    elif ctype.isdigit():
        charged, uniform, total = True, True, float(int(ctype))
    elif ctype == "calc":
        charged, calc = True, True
    scale = 1.0
    if hdr[5] and all(c.isdigit() or c == "." for c in hdr[5]):
        scale = _atof(hdr[5])
    conv = 0.529177210 if unit == "au" else 1.0
    natoms = _atoi(hdr[2])
    toks = " ".join(lines[6:]).split()
    xyz, q, sym = [], [], []
    for i in range(0, len(toks) - 4, 5):
        x, y, z, m, c = toks[i:i + 5]
        xyz.append([_atof(x) * scale * conv, _atof(y) * scale * conv, _atof(z) * scale * conv])
        if charged:
            q.append(total / natoms if uniform else _atof(c))
        else:
            q.append(0.0)
        im = _atoi(m)
        for row in table:
            if row[5] == im:
                sym.append(row[1])
                break
        else:
            sym.append(None)
    tot = 0.0
    for v in q:
        tot += v
    if calc:
        total = tot
    return dict(symbols=sym, xyz=xyz, charge=q, charged=charged, total_charge=total)


def read_pdb(path):
    """importPDB: ATOM/HETATM records, columns 31-54, element 77-78."""
    xyz, sym = [], []
    for raw in open(path).read().split("\n"):
        line = raw[:799]
        if line[:4] == "ATOM" or line[:6] == "HETATM":
            namelen = 1
            if len(line) >= 78:
                at = line[76:78]
                if at[0] == " " and at[1] == " ":
                    at = line[12:17]
                else:
                    namelen = 2
            else:
                at = line[12:17]
            t = ""
            for ch in at:
                if ch.isalpha():
                    t += ch
                    if len(t) == namelen:
                        break
            sym.append(t)
            xyz.append([_atof(line[30:38]), _atof(line[38:46]), _atof(line[46:54])])
    return dict(symbols=sym, xyz=xyz, charge=None, charged=False, total_charge=0.0)


def resolve(table, symbols):
    """setMolProperties lookup by symbol: (mass, sigma, eps) per atom."""
    out = []
    for s in symbols:
        for row in table:
            if row[1] == s:
                out.append((row[2], row[3], row[4]))
                break
        else:
            raise KeyError(f"atom type {s!r} not in force field")
    return out
