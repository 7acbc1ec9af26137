"""PDB text I/O (src/pdb.rs): the writers against the reference's format
string restated with Python's formatting (round-half-even on the exact
value, as Rust's {:8.3}), the reader against the reference's column rules."""
import numpy as np
import pytest

from dnab.structure import (ResidueCloud, aa_from_sequence, parse_pdb, protein_from_ca_trace,
                            write_pdb, write_pdb_allatom, write_pdb_multi)
from oracle import oracle as O

THREE = ["ALA", "ARG", "ASN", "ASP", "CYS", "GLN", "GLU", "GLY", "HIS", "ILE", "LEU", "LYS", "MET",
         "PHE", "PRO", "SER", "THR", "TRP", "TYR", "VAL"]


def rec(serial, name, aa, chain, resseq, x, y, z):
    # "ATOM  {:5} {:<4} {:3} {:1}{:4}    {:8.3}{:8.3}{:8.3}  1.00  0.00           C\n"
    return (f"ATOM  {min(serial, 99999):5d} {name:<4} {THREE[aa]:<3} {chain}{min(resseq, 9999):4d}"
            f"    {float(x):8.3f}{float(y):8.3f}{float(z):8.3f}  1.00  0.00           C\n")


def names(aa):
    return O.reference_tables()["atom_names"][aa]


def test_allatom_and_multi_records():
    # write_pdb_allatom (src/pdb.rs:121-155), write_pdb_multi (157-191)
    h = protein_from_ca_trace([0.0, 3.8, 7.6], [1.25, 0.0, -2.5], [0.0, 0.5, 1.0],
                              aa_from_sequence("WGC"))
    l = protein_from_ca_trace([10.0], [0.0], [0.0], aa_from_sequence("K"))
    def lines(p, chain, serial0):
        out, s = [], serial0
        for r, a in enumerate(p.amino_acid):
            b, e = int(p.res_start[r]), int(p.res_start[r + 1])
            for k in range(b, e):
                s += 1
                nm = names(a)[k - b]
                nm = (" " + nm + "   ")[:4] if len(nm) < 4 else nm   # the tables' 4-column names
                out.append(rec(s, nm, a, chain, r + 1, p.atoms.x[k], p.atoms.y[k], p.atoms.z[k]))
        return out, s
    a, s = lines(h, "H", 0)
    assert write_pdb_allatom(h, "H") == "".join(a) + "END\n"
    b, _ = lines(l, "L", s)
    assert write_pdb_multi([(h, "H"), (l, "L")]) == "".join(a) + "TER\n" + "".join(b) + "TER\nEND\n"


def test_ca_records_caps_and_specials():
    # write_pdb (src/pdb.rs:94-119): serial and residue number capped at 99999 / 9999
    n = 10002
    x = np.linspace(-5, 5, n).astype(np.float32)
    x[3] = np.float32(np.nan)
    x[4] = np.float32(np.inf)
    x[5] = np.float32(-0.0)
    x[6] = np.float32(0.0625)            # a tie at the third decimal: round half to even
    aa = (np.arange(n) % 20).astype(np.uint8)
    c = ResidueCloud.from_points(x, np.zeros(n, np.float32), np.full(n, 1.5, np.float32), aa)
    t = write_pdb(c, "A").splitlines(keepends=True)
    assert t[-1] == "END\n" and len(t) == n + 1
    assert t[0] == rec(1, " CA ", 0, "A", 1, x[0], 0.0, 1.5)
    assert t[3][30:38] == "     NaN" and t[4][30:38] == "     inf" and t[5][30:38] == "  -0.000"
    assert t[6][30:38] == "   0.062"
    assert t[n - 1] == rec(n, " CA ", int(aa[n - 1]), "A", n, x[n - 1], 0.0, 1.5)
    assert t[n - 1][22:26] == "9999"


def test_parse_pdb_reads_calphas_by_column():
    # parse_pdb (src/pdb.rs:19-57)
    lines = [
        "HEADER    TEST",
        rec(1, " N  ", 0, "A", 1, 1.0, 2.0, 3.0).rstrip("\n"),
        rec(2, " CA ", 0, "A", 1, 1.5, -2.25, 3.125).rstrip("\n"),
        "HETATM    3  CA  HID A   2      +1.500   1e1     -0.5  1.00  0.00           C\r",
        "ATOM      4  CA  XYZ A   3    " + " abc    " + "   2.000",  # bad x; the line ends before z
        "ATOM      5  CA  CYS A",                                  # no coordinates at all
        "ATOMS     6  CA  ALA A   4       9.000   9.000   9.000",  # still an ATOM prefix
        "REMARK    7  CA  ALA A   4       9.000   9.000   9.000",
        "ATOM      8  CB  ALA A   4       9.000   9.000   9.000",
    ]
    c = parse_pdb("\n".join(lines) + "\n")
    assert list(c.amino_acid) == [0, 8, 7, 4, 0]
    assert np.array_equal(c.x, np.array([1.5, 1.5, 0.0, 0.0, 9.0], np.float32))
    assert np.array_equal(c.y, np.array([-2.25, 10.0, 2.0, 0.0, 9.0], np.float32))
    assert np.array_equal(c.z, np.array([3.125, -0.5, 0.0, 0.0, 9.0], np.float32))
    with pytest.raises(ValueError):
        parse_pdb("HEADER only\nATOM      1  N   ALA A   1       1.000   2.000   3.000\n")


def test_pdb_round_trip():
    p = protein_from_ca_trace(np.arange(12, dtype=np.float32) * 3.8, np.full(12, 1.0, np.float32),
                              np.linspace(-3, 3, 12).astype(np.float32),
                              aa_from_sequence("ACDEFGHIKLMN"))
    c = parse_pdb(write_pdb_allatom(p, "B"))
    assert np.array_equal(c.amino_acid, p.amino_acid)
    ca = p.ca_pos()
    assert np.abs(np.stack([c.x, c.y, c.z], 1) - ca).max() <= 5e-4
