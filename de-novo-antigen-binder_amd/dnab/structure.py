"""Host-side structures (numpy) mirroring the reference's data model.

AtomCloud      src/allatom.rs:16-31 (flat SoA of heavy atoms)
AtomProtein    src/allatom.rs:92-101 (residue bookkeeping over an AtomCloud)
ResidueCloud   src/atom.rs:185-195 (one point per residue, residue-level params)

Atom placement is done by libdnab (dnab_build_protein), which restates
protein_from_ca_trace (src/allatom.rs:523-541).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _native as N

AMINO_ONE_LETTER = "ARNDCQEGHILKMFPSTWYV"  # AminoAcid order, src/atom.rs:4-25
AA_COUNT = 20
MAX_ATOMS_PER_RESIDUE = 14


def aa_from_sequence(seq: str) -> np.ndarray:
    """One-letter codes to AminoAcid indices (src/atom.rs:180-206)."""
    out = np.empty(len(seq), np.uint8)
    for i, ch in enumerate(seq.upper()):
        k = AMINO_ONE_LETTER.find(ch)
        if k < 0:
            raise ValueError(f"unknown one-letter code {ch!r} in peptide")
        out[i] = k
    return out


def sequence_of(aa: np.ndarray) -> str:
    return "".join(AMINO_ONE_LETTER[int(a)] for a in aa)


@dataclass
class AtomCloud:
    x: np.ndarray
    y: np.ndarray
    z: np.ndarray
    charge: np.ndarray
    r_min_half: np.ndarray
    epsilon: np.ndarray
    hydrophobic: np.ndarray
    atom_type: np.ndarray | None = None
    residue_idx: np.ndarray | None = None
    _keep: list = field(default_factory=list, repr=False)

    def __len__(self) -> int:
        return int(self.x.shape[0])

    def __post_init__(self):
        for n in ("x", "y", "z", "charge", "r_min_half", "epsilon"):
            setattr(self, n, np.ascontiguousarray(getattr(self, n), dtype=np.float32))
        self.hydrophobic = np.ascontiguousarray(self.hydrophobic, dtype=np.uint8)
        if self.atom_type is not None:
            self.atom_type = np.ascontiguousarray(self.atom_type, dtype=np.uint8)
        if self.residue_idx is not None:
            self.residue_idx = np.ascontiguousarray(self.residue_idx, dtype=np.uint32)

    def c(self) -> N.AtomCloudC:
        return N.AtomCloudC(len(self), N.ptr(self.x, C.c_float), N.ptr(self.y, C.c_float),
                            N.ptr(self.z, C.c_float), N.ptr(self.charge, C.c_float),
                            N.ptr(self.r_min_half, C.c_float), N.ptr(self.epsilon, C.c_float),
                            N.ptr(self.hydrophobic, C.c_uint8),
                            N.ptr(self.atom_type, C.c_uint8),
                            N.ptr(self.residue_idx, C.c_uint32))

    def subset(self, idx) -> "AtomCloud":
        def s(a):
            return None if a is None else a[idx]
        return AtomCloud(self.x[idx], self.y[idx], self.z[idx], self.charge[idx],
                         self.r_min_half[idx], self.epsilon[idx], self.hydrophobic[idx],
                         s(self.atom_type), s(self.residue_idx))

    def gpu_atoms(self) -> np.ndarray:
        """Pack as the 32-byte GpuAtom AoS (src/gpu.rs:18-29, src/diffusion.rs:1029-1040)."""
        a = np.zeros((len(self), 8), np.float32)
        a[:, 0], a[:, 1], a[:, 2] = self.x, self.y, self.z
        a[:, 3], a[:, 4], a[:, 5] = self.charge, self.r_min_half, self.epsilon
        a[:, 6] = self.hydrophobic.astype(np.float32)
        return a

    @staticmethod
    def concat(clouds: list["AtomCloud"]) -> "AtomCloud":
        def cat(name):
            parts = [getattr(c, name) for c in clouds]
            if any(p is None for p in parts):
                return None
            return np.concatenate(parts) if parts else np.zeros(0)
        return AtomCloud(cat("x"), cat("y"), cat("z"), cat("charge"), cat("r_min_half"),
                         cat("epsilon"), cat("hydrophobic"), cat("atom_type"), None)


@dataclass
class ResidueCloud:
    x: np.ndarray
    y: np.ndarray
    z: np.ndarray
    charge: np.ndarray
    epsilon: np.ndarray
    sigma: np.ndarray
    hydrophobic: np.ndarray
    amino_acid: np.ndarray

    def __len__(self) -> int:
        return int(self.x.shape[0])

    def __post_init__(self):
        for n in ("x", "y", "z", "charge", "epsilon", "sigma"):
            setattr(self, n, np.ascontiguousarray(getattr(self, n), dtype=np.float32))
        self.hydrophobic = np.ascontiguousarray(self.hydrophobic, dtype=np.uint8)
        self.amino_acid = np.ascontiguousarray(self.amino_acid, dtype=np.uint8)

    @staticmethod
    def from_points(x, y, z, aa) -> "ResidueCloud":
        """ResidueCloud::push per point (src/atom.rs:226-235)."""
        aa = np.asarray(aa, np.uint8)
        q = np.empty(len(aa), np.float32)
        e = np.empty(len(aa), np.float32)
        s = np.empty(len(aa), np.float32)
        h = np.empty(len(aa), np.uint8)
        for i, a in enumerate(aa):
            qq, ee, ss, hh = C.c_float(), C.c_float(), C.c_float(), C.c_uint8()
            N.check(N.lib.dnab_residue_cg_params(int(a), C.byref(qq), C.byref(ee), C.byref(ss),
                                                 C.byref(hh)))
            q[i], e[i], s[i], h[i] = qq.value, ee.value, ss.value, hh.value
        return ResidueCloud(np.asarray(x, np.float32), np.asarray(y, np.float32),
                            np.asarray(z, np.float32), q, e, s, h, aa)

    def c(self) -> N.ResidueCloudC:
        return N.ResidueCloudC(len(self), N.ptr(self.x, C.c_float), N.ptr(self.y, C.c_float),
                               N.ptr(self.z, C.c_float), N.ptr(self.charge, C.c_float),
                               N.ptr(self.epsilon, C.c_float), N.ptr(self.sigma, C.c_float),
                               N.ptr(self.hydrophobic, C.c_uint8))


@dataclass
class AtomProtein:
    atoms: AtomCloud
    res_start: np.ndarray   # n_res + 1
    amino_acid: np.ndarray  # n_res
    ca_atom_idx: np.ndarray  # n_res

    def n_residues(self) -> int:
        return int(self.amino_acid.shape[0])

    def n_atoms(self) -> int:
        return len(self.atoms)

    def ca_pos(self) -> np.ndarray:
        i = self.ca_atom_idx
        return np.stack([self.atoms.x[i], self.atoms.y[i], self.atoms.z[i]], axis=1)

    def ca_center_of_mass(self) -> np.ndarray:
        """f32 sequential sums / n, src/allatom.rs:183-195."""
        p = self.ca_pos()
        if len(p) == 0:
            return np.zeros(3, np.float32)
        acc = np.zeros(3, np.float32)
        for row in p:
            acc = (acc + row).astype(np.float32)
        return (acc / np.float32(len(p))).astype(np.float32)

    def ca_residue_cloud(self) -> ResidueCloud:
        """Antigen Calpha cloud of run_allatom, src/diffusion.rs:994-1002."""
        p = self.ca_pos()
        return ResidueCloud.from_points(p[:, 0], p[:, 1], p[:, 2], self.amino_acid)

    def sequence(self) -> str:
        return sequence_of(self.amino_acid)


def protein_from_ca_trace(x, y, z, amino_acid, rotamer_idx=None) -> AtomProtein:
    """protein_from_ca_trace (src/allatom.rs:523-541) via dnab_build_protein."""
    x = np.ascontiguousarray(x, np.float32)
    y = np.ascontiguousarray(y, np.float32)
    z = np.ascontiguousarray(z, np.float32)
    aa = np.ascontiguousarray(amino_acid, np.uint8)
    rot = None if rotamer_idx is None else np.ascontiguousarray(rotamer_idx, np.uint8)
    n = len(aa)
    cap = max(1, MAX_ATOMS_PER_RESIDUE * n)
    f = {k: np.zeros(cap, np.float32) for k in ("x", "y", "z", "charge", "rmh", "eps")}
    hyd = np.zeros(cap, np.uint8)
    at = np.zeros(cap, np.uint8)
    ri = np.zeros(cap, np.uint32)
    rs = np.zeros(n + 1, np.uint32)
    ca = np.zeros(max(n, 1), np.uint32)
    aa_out = np.zeros(max(n, 1), np.uint8)
    out = N.ProteinOutC(cap, N.ptr(f["x"], C.c_float), N.ptr(f["y"], C.c_float),
                        N.ptr(f["z"], C.c_float), N.ptr(f["charge"], C.c_float),
                        N.ptr(f["rmh"], C.c_float), N.ptr(f["eps"], C.c_float),
                        N.ptr(hyd, C.c_uint8), N.ptr(at, C.c_uint8), N.ptr(ri, C.c_uint32),
                        N.ptr(rs, C.c_uint32), N.ptr(ca, C.c_uint32), N.ptr(aa_out, C.c_uint8))
    na = C.c_size_t(0)
    N.check(N.lib.dnab_build_protein(N.ptr(x, C.c_float), N.ptr(y, C.c_float),
                                     N.ptr(z, C.c_float), N.ptr(aa, C.c_uint8),
                                     N.ptr(rot, C.c_uint8), n, C.byref(out), C.byref(na)))
    m = na.value
    cloud = AtomCloud(f["x"][:m].copy(), f["y"][:m].copy(), f["z"][:m].copy(),
                      f["charge"][:m].copy(), f["rmh"][:m].copy(), f["eps"][:m].copy(),
                      hyd[:m].copy(), at[:m].copy(), ri[:m].copy())
    return AtomProtein(cloud, rs, aa_out[:n].copy(), ca[:n].copy())


def parse_peptide(seq: str) -> tuple[np.ndarray, np.ndarray, np.ndarray, np.ndarray]:
    """Straight Calpha line along x, src/pdb.rs:64-83: x_i = (i - n/2) * 3.8."""
    aa = aa_from_sequence(seq)
    if len(aa) == 0:
        raise ValueError("empty peptide")
    n = np.float32(len(aa))
    x = np.array([(np.float32(i) - n * np.float32(0.5)) * np.float32(3.8) for i in range(len(aa))],
                 np.float32)
    zero = np.zeros(len(aa), np.float32)
    return x, zero.copy(), zero.copy(), aa


@dataclass
class PopulationHost:
    """Ragged batch of candidates bound to antigens (no reference counterpart;
    src/gpu.rs:166-174 needs one atom count for all candidates)."""
    cand_antigen: np.ndarray
    cand_res_off: np.ndarray
    res_atom_off: np.ndarray
    res_ca_atom: np.ndarray
    res_amino_acid: np.ndarray
    atoms: AtomCloud

    @property
    def n_cand(self) -> int:
        return int(self.cand_antigen.shape[0])

    @property
    def n_res(self) -> int:
        return int(self.res_ca_atom.shape[0])

    def cand_atom_off(self) -> np.ndarray:
        return self.res_atom_off[self.cand_res_off].astype(np.uint32)

    def c(self) -> N.PopulationDescC:
        return N.PopulationDescC(self.n_cand, self.n_res, N.ptr(self.cand_antigen, C.c_uint32),
                                 N.ptr(self.cand_res_off, C.c_uint32),
                                 N.ptr(self.res_atom_off, C.c_uint32),
                                 N.ptr(self.res_ca_atom, C.c_uint32),
                                 N.ptr(self.res_amino_acid, C.c_uint8), self.atoms.c())

    def candidate_cloud(self, c: int) -> AtomCloud:
        off = self.cand_atom_off()
        return self.atoms.subset(slice(int(off[c]), int(off[c + 1])))


def population_from_proteins(proteins: list[AtomProtein], antigen_ids=None) -> PopulationHost:
    n = len(proteins)
    ag = np.zeros(n, np.uint32) if antigen_ids is None else np.asarray(antigen_ids, np.uint32)
    cand_res_off = np.zeros(n + 1, np.uint32)
    res_atom_off = [np.zeros(1, np.uint32)]
    res_ca, res_aa, clouds = [], [], []
    atom_base = 0
    for c, p in enumerate(proteins):
        cand_res_off[c + 1] = cand_res_off[c] + p.n_residues()
        res_atom_off.append(p.res_start[1:].astype(np.uint32) + np.uint32(atom_base))
        res_ca.append(p.ca_atom_idx.astype(np.uint32) + np.uint32(atom_base))
        res_aa.append(p.amino_acid)
        clouds.append(p.atoms)
        atom_base += p.n_atoms()
    return PopulationHost(
        ag, cand_res_off, np.ascontiguousarray(np.concatenate(res_atom_off), np.uint32),
        np.ascontiguousarray(np.concatenate(res_ca) if res_ca else np.zeros(0), np.uint32),
        np.ascontiguousarray(np.concatenate(res_aa) if res_aa else np.zeros(0), np.uint8),
        AtomCloud.concat(clouds))


# ---- PDB text I/O (src/pdb.rs) ----

def parse_pdb(text: str | bytes) -> ResidueCloud:
    """parse_pdb (src/pdb.rs:19-57) on a PDB file's text: the CA atoms."""
    data = text.encode() if isinstance(text, str) else bytes(text)
    n = C.c_size_t(0)
    st = N.lib.dnab_parse_pdb(data, len(data), None, None, None, None, 0, C.byref(n))
    if st not in (N.DNAB_OK, N.DNAB_ERR_CAPACITY):
        raise ValueError("no CA atom in the PDB text")
    m = n.value
    x, y, z = (np.zeros(m, np.float32) for _ in range(3))
    aa = np.zeros(m, np.uint8)
    N.check(N.lib.dnab_parse_pdb(data, len(data), N.ptr(x, C.c_float), N.ptr(y, C.c_float),
                                 N.ptr(z, C.c_float), N.ptr(aa, C.c_uint8), m, C.byref(n)))
    return ResidueCloud.from_points(x, y, z, aa)


def _text(fn):
    n = C.c_size_t(0)
    N.check(fn(None, 0, C.byref(n)))
    buf = C.create_string_buffer(n.value + 1)
    N.check(fn(buf, n.value + 1, C.byref(n)))
    return buf.value.decode()


def write_pdb(cloud: ResidueCloud, chain_id: str = "B") -> str:
    """write_pdb (src/pdb.rs:94-119): CA-only ATOM records."""
    return _text(lambda o, cap, n: N.lib.dnab_write_pdb_ca(
        N.ptr(cloud.x, C.c_float), N.ptr(cloud.y, C.c_float), N.ptr(cloud.z, C.c_float),
        N.ptr(cloud.amino_acid, C.c_uint8), len(cloud), chain_id.encode(), o, cap, n))


def write_pdb_multi(chains: list[tuple[AtomProtein, str]], ter: bool = True) -> str:
    """write_pdb_multi (src/pdb.rs:157-191); ter=False with one chain is
    write_pdb_allatom (src/pdb.rs:121-155)."""
    arr = (N.PdbChainC * max(len(chains), 1))()
    keep = []
    for k, (p, cid) in enumerate(chains):
        rs = np.ascontiguousarray(p.res_start, np.uint32)
        aa = np.ascontiguousarray(p.amino_acid, np.uint8)
        keep.append((rs, aa))
        arr[k] = N.PdbChainC(N.ptr(p.atoms.x, C.c_float), N.ptr(p.atoms.y, C.c_float),
                             N.ptr(p.atoms.z, C.c_float), N.ptr(rs, C.c_uint32),
                             N.ptr(aa, C.c_uint8), len(aa), cid.encode())
    return _text(lambda o, cap, n: N.lib.dnab_write_pdb_chains(arr, len(chains), 1 if ter else 0,
                                                              o, cap, n))


def write_pdb_allatom(p: AtomProtein, chain_id: str = "B") -> str:
    return write_pdb_multi([(p, chain_id)], ter=False)
