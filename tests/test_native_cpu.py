"""CPU tests of libdnab's host side: the library loads, every symbol declared
in include/dnab.h is exported, the structure builders follow the reference
(atom counts, NERF geometry, force-field tables), and context creation
without a GPU reports 'no device' (the reference's try_init -> None)."""
import json
import re
from pathlib import Path

import numpy as np
import pytest

import dnab
from dnab import _native as N
from dnab.structure import aa_from_sequence, parse_peptide, protein_from_ca_trace
from dnab import workloads as W

ROOT = Path(__file__).resolve().parent.parent


def test_exports_every_declared_symbol():
    header = (ROOT / "include" / "dnab.h").read_text()
    declared = set(re.findall(r"^\s*(?:int|void|const char\*)\s+(dnab_\w+)\s*\(", header, re.M))
    assert len(declared) > 40
    missing = [s for s in declared if not hasattr(N.lib, s)]
    assert not missing, missing
    assert declared == set(N.EXPORTED), set(N.EXPORTED) ^ declared


def test_status_strings():
    assert N.lib.dnab_status_string(0) == b"ok"
    assert N.lib.dnab_status_string(N.DNAB_ERR_NO_DEVICE) == b"no usable GPU"


def test_residue_atom_counts():
    # src/allatom.rs:560-578: ALA 5, GLY 4; the 20 topologies hold 167 atoms.
    counts = [N.lib.dnab_residue_atom_count(a) for a in range(20)]
    assert counts[0] == 5 and counts[7] == 4
    assert sum(counts) == 167
    assert N.lib.dnab_residue_atom_count(20) == -1
    p = protein_from_ca_trace([10.0], [0.0], [0.0], aa_from_sequence("A"))
    assert p.n_atoms() == 5
    p = protein_from_ca_trace([0.0], [0.0], [0.0], aa_from_sequence("G"))
    assert p.n_atoms() == 4


def test_nerf_bond_length():
    # src/rotamer.rs:477-487
    import ctypes as C
    a = (C.c_float * 3)(0.0, 0.0, 0.0)
    b = (C.c_float * 3)(1.5, 0.0, 0.0)
    c = (C.c_float * 3)(1.5, 1.4, 0.0)
    d = (C.c_float * 3)()
    N.check(N.lib.dnab_place_atom(a, b, c, 1.52, 112.0, -60.0, d))
    dist = np.sqrt((d[0] - 1.5) ** 2 + (d[1] - 1.4) ** 2 + d[2] ** 2)
    assert abs(dist - 1.52) < 1e-4
    # The reference's place_atom (src/rotamer.rs:209-230) measures along the
    # C->B unit vector with -cos(angle), so the placed atom has bond angle
    # B-C-D = 180 - angle and dihedral A-B-C-D = 180 - dihedral. The builder
    # reproduces the reference's geometry, so pin exactly that.
    B, Cc, A_ = np.array(b[:]), np.array(c[:]), np.array(a[:])
    for dih_in in (-60.0, 0.0, 60.0, 120.0, 175.0):
        N.check(N.lib.dnab_place_atom(a, b, c, 1.52, 112.0, dih_in, d))
        D = np.array(d[:])
        u, v = B - Cc, D - Cc
        ang = np.degrees(np.arccos(u @ v / np.linalg.norm(u) / np.linalg.norm(v)))
        assert abs(ang - (180.0 - 112.0)) < 1e-3
        b0, b1, b2 = B - A_, Cc - B, D - Cc
        n1, n2 = np.cross(b0, b1), np.cross(b1, b2)
        m1 = np.cross(n1, b1 / np.linalg.norm(b1))
        dih = np.degrees(np.arctan2(m1 @ n2, n1 @ n2))
        assert abs(((dih - (180.0 - dih_in)) + 180.0) % 360.0 - 180.0) < 1e-3


def test_side_chain_bond_lengths():
    # every residue built from a trace: backbone at the ideal offsets
    # (src/allatom.rs:502-517) and side-chain bonds at their table lengths.
    seq = "ARNDCQEGHILKMFPSTWYV"
    xs = np.arange(20, dtype=np.float32) * 10.0
    p = protein_from_ca_trace(xs, np.zeros(20), np.zeros(20), aa_from_sequence(seq))
    assert p.n_atoms() == 167
    for r in range(20):
        s = int(p.res_start[r])
        ca = np.array([p.atoms.x[s + 1], p.atoms.y[s + 1], p.atoms.z[s + 1]])
        n = np.array([p.atoms.x[s], p.atoms.y[s], p.atoms.z[s]])
        assert np.allclose(ca - n, [1.46, 0, 0], atol=1e-5)
        assert p.ca_atom_idx[r] == s + 1


def test_ff_tables_match_reference_fixture():
    fx = json.loads((ROOT / "tests" / "golden" / "ff_tables.json").read_text())
    import ctypes as C
    for aa, atoms in enumerate(fx["residues"]):
        assert N.lib.dnab_residue_atom_count(aa) == len(atoms)
        for slot, (atype, charge, hyd) in enumerate(atoms):
            t, q, h, rmh, eps = C.c_uint8(), C.c_float(), C.c_uint8(), C.c_float(), C.c_float()
            N.check(N.lib.dnab_ff_table(aa, slot, C.byref(t), C.byref(q), C.byref(h), C.byref(rmh),
                                        C.byref(eps)))
            assert t.value == atype and h.value == hyd
            assert q.value == np.float32(charge)
            assert rmh.value == np.float32(fx["amber_nb"][atype][0])
            assert eps.value == np.float32(fx["amber_nb"][atype][1])
    for t in range(16):
        dg, vol, lam = C.c_float(), C.c_float(), C.c_float()
        don, acc = C.c_uint8(), C.c_uint8()
        N.check(N.lib.dnab_ff_type_params(t, C.byref(dg), C.byref(vol), C.byref(lam), C.byref(don),
                                          C.byref(acc)))
        assert [dg.value, vol.value, lam.value] == [np.float32(v) for v in fx["eef1"][t]]
        assert don.value == fx["hbond_donor"][t] and acc.value == fx["hbond_acceptor"][t]
    assert N.lib.dnab_ff_type_params(16, None, None, None, None, None) != 0
    # the disulfide term reads the Cys SG atom at topology slot 5
    cys = fx["residue_order"].index("CYS")
    assert fx["atom_names"][cys].index("SG") == 5


def test_parse_peptide_layout():
    x, y, z, aa = parse_peptide("ACDE")
    assert list(x) == [np.float32(-7.6), np.float32(-3.8), 0.0, np.float32(3.8)]
    with pytest.raises(ValueError):
        parse_peptide("AXB")


def test_workload_shapes():
    w = W.config1()
    assert w.antigens[0].n_residues() == 16 and w.pop.n_cand == 64
    assert all(np.diff(w.pop.cand_res_off) == 12)
    w = W.config2(n_cand=16)
    assert 2200 < w.antigens[0].n_atoms() < 2900          # ~8.35 atoms/residue
    assert all(np.diff(w.pop.cand_res_off) == 20)


def test_no_device_reports_none(monkeypatch):
    if dnab.visible_gpu():
        pytest.skip("a GPU is visible")
    assert dnab.GpuContext.try_init() is None


def test_population_desc_layout():
    w = W.config2(n_cand=8)
    off = w.pop.cand_atom_off()
    assert off[0] == 0 and off[-1] == len(w.pop.atoms)
    assert (w.pop.res_ca_atom >= w.pop.res_atom_off[:-1]).all()
    assert (w.pop.res_ca_atom < w.pop.res_atom_off[1:]).all()


def test_shell_population_matches_per_binder_path():
    # the threaded batch generator builds the same candidates as one call per binder
    import numpy as np
    from dnab import workloads as W
    from dnab.structure import population_from_proteins
    ctr = np.array([1.0, -2.0, 0.5], np.float32)
    seed = (2 << 32) | (7 << 20)
    batch = W.shell_population(37, 9, ctr, 28.0, 25.0, seed, nthreads=5)
    ref = population_from_proteins([W.shell_binder(9, ctr, 28.0, 25.0, seed + k) for k in range(37)])
    for name in ("cand_res_off", "res_atom_off", "res_ca_atom", "res_amino_acid"):
        assert np.array_equal(getattr(batch, name), getattr(ref, name)), name
    for name in ("x", "y", "z", "charge", "r_min_half", "epsilon", "hydrophobic"):
        assert np.array_equal(getattr(batch.atoms, name), getattr(ref.atoms, name)), name
