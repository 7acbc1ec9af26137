"""ctypes loader for the CPU oracle (liboracle.so, built from dnab_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product. Each function states the
reference lines it restates (see dnab_oracle.c).

Cloud arguments are any objects with numpy attributes x, y, z, charge,
r_min_half, epsilon, hydrophobic (atom clouds) or x, y, z, charge, epsilon,
sigma, hydrophobic (residue clouds).
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle.so"

f32p = C.POINTER(C.c_float)
f64p = C.POINTER(C.c_double)
u8p = C.POINTER(C.c_uint8)
u32p = C.POINTER(C.c_uint32)
u64p = C.POINTER(C.c_uint64)


class OrCloud(C.Structure):
    _fields_ = [("n", C.c_size_t), ("x", f32p), ("y", f32p), ("z", f32p), ("q", f32p),
                ("rmh", f32p), ("eps", f32p), ("hyd", u8p)]


class OrRCloud(C.Structure):
    _fields_ = [("n", C.c_size_t), ("x", f32p), ("y", f32p), ("z", f32p), ("q", f32p),
                ("eps", f32p), ("sig", f32p), ("hyd", u8p)]


class OrExt(C.Structure):
    """Per-atom EEF1 parameters and H-bond classes."""
    _fields_ = [("dg", f32p), ("vol", f32p), ("lam", f32p), ("don", u8p), ("acc", u8p)]


class OrStepParams(C.Structure):
    _fields_ = [("step", C.c_float), ("max_disp", C.c_float), ("restraint_k", C.c_float),
                ("restraint_radius", C.c_float), ("sigma", C.c_float), ("with_noise", C.c_int),
                ("force_model", C.c_int), ("iter", C.c_uint32), ("seed_lo", C.c_uint32),
                ("seed_hi", C.c_uint32), ("center", C.c_float * 3)]


def _load():
    if not LIB.exists():
        raise ImportError(f"{LIB} not built: run `make -C {HERE}`")
    lib = C.CDLL(str(LIB))
    protos = {
        "or_cell_of": (C.c_int32, [C.c_float, C.c_float]),
        "or_inv_cell": (C.c_float, [C.c_float]),
        "or_grid_query": (C.c_int, [f32p, f32p, f32p, C.c_size_t, C.c_float, f32p, f32p, f32p,
                                    C.c_size_t, u32p, u32p, C.c_size_t]),
        "or_energy_f32": (None, [C.POINTER(OrCloud), C.POINTER(OrCloud), C.c_float, f32p]),
        "or_energy_f64": (None, [C.POINTER(OrCloud), C.POINTER(OrCloud), C.c_float, f64p, f64p,
                                 f64p, f64p, u64p]),
        "or_population_f64": (None, [C.POINTER(OrCloud), C.POINTER(OrCloud), u32p, C.c_size_t,
                                     C.c_float, f64p, f64p, f64p, f64p, f64p, f64p, f64p, u64p,
                                     C.c_int]),
        "or_baseline_energy_brute": (None, [C.POINTER(OrCloud), C.POINTER(OrCloud), u32p,
                                            C.c_size_t, C.c_float, f32p, C.c_int]),
        "or_baseline_grid_ef": (C.c_int, [C.POINTER(OrCloud), C.POINTER(OrCloud), u32p,
                                          C.c_size_t, C.c_float, f32p, f32p, f32p, f32p,
                                          C.c_int]),
        "or_cg_forces": (None, [C.POINTER(OrRCloud), C.POINTER(OrRCloud), f32p, f32p, f32p]),
        "or_cg_forces_with_grid": (C.c_int, [C.POINTER(OrRCloud), C.c_float,
                                             C.POINTER(OrRCloud), f32p, f32p, f32p]),
        "or_hbond_energy_f32": (C.c_float, [C.c_float]),
        "or_hbond_energy_f64": (C.c_double, [C.c_double]),
        "or_pair_solvation_f32": (C.c_float, [C.c_float] * 7),
        "or_pair_solvation_f64": (C.c_double, [C.c_double] * 7),
        "or_energy_full_f32": (None, [C.POINTER(OrCloud), C.POINTER(OrCloud), C.POINTER(OrExt),
                                      C.POINTER(OrExt), C.c_float, f32p]),
        "or_energy_full_f64": (None, [C.POINTER(OrCloud), C.POINTER(OrCloud), C.POINTER(OrExt),
                                      C.POINTER(OrExt), C.c_float, f64p, f64p]),
        "or_disulfide_f32": (C.c_float, [f32p, f32p, f32p, C.c_size_t]),
        "or_disulfide_f64": (C.c_double, [f32p, f32p, f32p, C.c_size_t]),
        "or_philox4x32_10": (None, [u32p, u32p, u32p]),
        "or_sigma": (C.c_float, [C.c_uint32, C.c_uint32, C.c_float, C.c_float, C.c_float]),
        "or_temperature": (C.c_float, [C.c_uint32, C.c_uint32, C.c_float, C.c_float]),
        "or_refine_run": (C.c_int, [C.c_void_p, C.POINTER(OrRCloud), u8p, C.c_size_t, u32p, u8p,
                                    u8p, f32p, f32p, f32p, C.c_uint32, C.c_uint32, C.c_void_p,
                                    C.c_int]),
        "or_langevin_step": (None, [f32p, f32p, f32p, C.c_size_t, u32p, u32p, f32p, f32p, f32p,
                                    C.POINTER(OrRCloud), f32p, f32p, f32p, u8p, C.c_uint32,
                                    C.POINTER(OrStepParams)]),
    }
    for name, (res, args) in protos.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


lib = _load()


def _p(a, ct):
    if a is None:
        return C.cast(None, C.POINTER(ct))
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.POINTER(ct))


def _f32(a):
    return np.ascontiguousarray(a, np.float32)


def cloud(c) -> tuple[OrCloud, list]:
    keep = [_f32(c.x), _f32(c.y), _f32(c.z), _f32(c.charge), _f32(c.r_min_half),
            _f32(c.epsilon), np.ascontiguousarray(c.hydrophobic, np.uint8)]
    oc = OrCloud(len(keep[0]), _p(keep[0], C.c_float), _p(keep[1], C.c_float),
                 _p(keep[2], C.c_float), _p(keep[3], C.c_float), _p(keep[4], C.c_float),
                 _p(keep[5], C.c_float), _p(keep[6], C.c_uint8))
    return oc, keep


def rcloud(c) -> tuple[OrRCloud, list]:
    keep = [_f32(c.x), _f32(c.y), _f32(c.z), _f32(c.charge), _f32(c.epsilon), _f32(c.sigma),
            np.ascontiguousarray(c.hydrophobic, np.uint8)]
    oc = OrRCloud(len(keep[0]), *[_p(k, C.c_float) for k in keep[:6]], _p(keep[6], C.c_uint8))
    return oc, keep


def cell_of(x: float, cell: float) -> int:
    return int(lib.or_cell_of(x, lib.or_inv_cell(cell)))


def grid_query(x, y, z, cell, px, py, pz):
    """Neighbour lists by the definition of src/spatial.rs:31-58."""
    x, y, z, px, py, pz = (_f32(np.atleast_1d(v)) for v in (x, y, z, px, py, pz))
    off = np.zeros(len(px) + 1, np.uint32)
    lib.or_grid_query(_p(x, C.c_float), _p(y, C.c_float), _p(z, C.c_float), len(x), cell,
                      _p(px, C.c_float), _p(py, C.c_float), _p(pz, C.c_float), len(px),
                      _p(off, C.c_uint32), _p(None, C.c_uint32), 0)
    idx = np.zeros(max(int(off[-1]), 1), np.uint32)
    lib.or_grid_query(_p(x, C.c_float), _p(y, C.c_float), _p(z, C.c_float), len(x), cell,
                      _p(px, C.c_float), _p(py, C.c_float), _p(pz, C.c_float), len(px),
                      _p(off, C.c_uint32), _p(idx, C.c_uint32), len(idx))
    return off, idx[:int(off[-1])]


def energy_f32(ag, ab, cutoff=10.0):
    """(lj, coulomb, hydrophobic) in f32, reference order (src/energy.rs:269-301)."""
    a, ka = cloud(ag)
    b, kb = cloud(ab)
    out = np.zeros(3, np.float32)
    lib.or_energy_f32(C.byref(a), C.byref(b), cutoff, _p(out, C.c_float))
    return out


def energy_f64(ag, ab, cutoff=10.0):
    """FP64 (lj, coulomb, hydrophobic), per-atom FP64 forces, pair count."""
    a, ka = cloud(ag)
    b, kb = cloud(ab)
    n = len(kb[0])
    out = np.zeros(3, np.float64)
    fx, fy, fz = (np.zeros(max(n, 1), np.float64) for _ in range(3))
    np_ = C.c_uint64()
    lib.or_energy_f64(C.byref(a), C.byref(b), cutoff, _p(out, C.c_double), _p(fx, C.c_double),
                      _p(fy, C.c_double), _p(fz, C.c_double), C.byref(np_))
    return out, (fx[:n], fy[:n], fz[:n]), np_.value


def population_f64(ag, atoms, cand_atom_off, cutoff=10.0, forces=True, nthreads=None):
    """Per candidate FP64 energy + terms, per atom FP64 forces, per candidate pairs."""
    nthreads = nthreads or os.cpu_count() or 1
    a, ka = cloud(ag)
    b, kb = cloud(atoms)
    off = np.ascontiguousarray(cand_atom_off, np.uint32)
    nc = len(off) - 1
    na = len(kb[0])
    e, lj, co, hy = (np.zeros(max(nc, 1), np.float64) for _ in range(4))
    pairs = np.zeros(max(nc, 1), np.uint64)
    f = [np.zeros(max(na, 1), np.float64) for _ in range(3)] if forces else [None] * 3
    lib.or_population_f64(C.byref(a), C.byref(b), _p(off, C.c_uint32), nc, cutoff,
                          _p(e, C.c_double), _p(lj, C.c_double), _p(co, C.c_double),
                          _p(hy, C.c_double), *[_p(v, C.c_double) for v in f],
                          _p(pairs, C.c_uint64), int(nthreads))
    res = {"energy": e[:nc], "lj": lj[:nc], "coulomb": co[:nc], "hydrophobic": hy[:nc],
           "pairs": pairs[:nc]}
    if forces:
        res["forces"] = tuple(v[:na] for v in f)
    return res


# ---- H-bond, EEF1 solvation, disulfide (the reference's other energy terms) ----

def reference_tables():
    """The reference's force-field tables as recorded in the committed fixture
    tests/golden/ff_tables.json (EEF1 rows and H-bond classes per AtomType)."""
    import json
    return json.loads((HERE.parent / "tests" / "golden" / "ff_tables.json").read_text())


def ext(types) -> tuple[OrExt, list]:
    """Per-atom EEF1 parameters and H-bond classes from atom types."""
    tab = reference_tables()
    t = np.asarray(types, np.int64)
    eef = np.asarray(tab["eef1"], np.float32).reshape(-1, 3)
    don = np.asarray(tab["hbond_donor"], np.uint8)
    acc = np.asarray(tab["hbond_acceptor"], np.uint8)
    keep = [np.ascontiguousarray(eef[t, 0]), np.ascontiguousarray(eef[t, 1]),
            np.ascontiguousarray(eef[t, 2]), np.ascontiguousarray(don[t]),
            np.ascontiguousarray(acc[t])]
    oe = OrExt(*[_p(k, C.c_float) for k in keep[:3]], _p(keep[3], C.c_uint8),
               _p(keep[4], C.c_uint8))
    return oe, keep


def hbond_energy_f32(r_sq):
    return float(np.float32(lib.or_hbond_energy_f32(float(r_sq))))


def hbond_energy_f64(r_sq):
    return lib.or_hbond_energy_f64(float(r_sq))


def pair_solvation_f32(pi, pj, r_sq):
    """pi, pj: (dg_ref, vol, lambda) of the two atoms."""
    return float(np.float32(lib.or_pair_solvation_f32(*map(float, pi), *map(float, pj),
                                                      float(r_sq))))


def pair_solvation_f64(pi, pj, r_sq):
    return lib.or_pair_solvation_f64(*map(float, pi), *map(float, pj), float(r_sq))


def energy_full_f32(ag, ab, cutoff=10.0):
    """(lj, coulomb, hydrophobic, hbond, solvation) in f32, the reference's order."""
    a, ka = cloud(ag)
    b, kb = cloud(ab)
    ea, kea = ext(ag.atom_type)
    eb, keb = ext(ab.atom_type)
    out = np.zeros(5, np.float32)
    lib.or_energy_full_f32(C.byref(a), C.byref(b), C.byref(ea), C.byref(eb), cutoff,
                           _p(out, C.c_float))
    return out


def energy_full_f64(ag, ab, cutoff=10.0):
    """FP64 (lj, coulomb, hydrophobic, hbond, solvation) and the per-atom (n, 5) terms."""
    a, ka = cloud(ag)
    b, kb = cloud(ab)
    ea, kea = ext(ag.atom_type)
    eb, keb = ext(ab.atom_type)
    n = len(kb[0])
    out = np.zeros(5, np.float64)
    per = np.zeros((max(n, 1), 5), np.float64)
    lib.or_energy_full_f64(C.byref(a), C.byref(b), C.byref(ea), C.byref(eb), cutoff,
                           _p(out, C.c_double), _p(per, C.c_double))
    return out, per[:n]


def disulfide_f32(x, y, z):
    x, y, z = _f32(x), _f32(y), _f32(z)
    return float(np.float32(lib.or_disulfide_f32(_p(x, C.c_float), _p(y, C.c_float),
                                                 _p(z, C.c_float), len(x))))


def disulfide_f64(x, y, z):
    x, y, z = _f32(x), _f32(y), _f32(z)
    return lib.or_disulfide_f64(_p(x, C.c_float), _p(y, C.c_float), _p(z, C.c_float), len(x))


CYS, CYS_SG_SLOT = 4, 5   # AminoAcid::Cys; SG's slot in the Cys topology (src/amber.rs)


def sg_positions(x, y, z, res_atom_off, res_amino_acid, r0, r1):
    """SG atoms of the Cys residues r0..r1 in residue order (sg_pos,
    src/allatom.rs:139-150)."""
    sel = [res_atom_off[r] + CYS_SG_SLOT for r in range(r0, r1)
           if res_amino_acid[r] == CYS and res_atom_off[r + 1] - res_atom_off[r] > CYS_SG_SLOT]
    sel = np.asarray(sel, np.int64)
    return np.asarray(x)[sel], np.asarray(y)[sel], np.asarray(z)[sel]


def baseline_energy_brute(ag, atoms, cand_atom_off, cutoff, nthreads):
    a, ka = cloud(ag)
    b, kb = cloud(atoms)
    off = np.ascontiguousarray(cand_atom_off, np.uint32)
    e = np.zeros(max(len(off) - 1, 1), np.float32)
    lib.or_baseline_energy_brute(C.byref(a), C.byref(b), _p(off, C.c_uint32), len(off) - 1,
                                 cutoff, _p(e, C.c_float), int(nthreads))
    return e[:len(off) - 1]


def baseline_grid_ef(ag, atoms, cand_atom_off, cutoff, nthreads, forces=False):
    a, ka = cloud(ag)
    b, kb = cloud(atoms)
    off = np.ascontiguousarray(cand_atom_off, np.uint32)
    nc = len(off) - 1
    e = np.zeros(max(nc, 1), np.float32)
    na = len(kb[0])
    f = [np.zeros(max(na, 1), np.float32) for _ in range(3)] if forces else [None] * 3
    rc = lib.or_baseline_grid_ef(C.byref(a), C.byref(b), _p(off, C.c_uint32), nc, cutoff,
                                 _p(e, C.c_float), *[_p(v, C.c_float) for v in f], int(nthreads))
    assert rc == 0
    return (e[:nc], tuple(v[:na] for v in f)) if forces else e[:nc]


def cg_forces(ag, ab):
    a, ka = rcloud(ag)
    b, kb = rcloud(ab)
    n = len(kb[0])
    f = [np.zeros(max(n, 1), np.float32) for _ in range(3)]
    lib.or_cg_forces(C.byref(a), C.byref(b), *[_p(v, C.c_float) for v in f])
    return tuple(v[:n] for v in f)


def cg_forces_with_grid(ag, ab, cell=10.0):
    a, ka = rcloud(ag)
    b, kb = rcloud(ab)
    n = len(kb[0])
    f = [np.zeros(max(n, 1), np.float32) for _ in range(3)]
    lib.or_cg_forces_with_grid(C.byref(a), cell, C.byref(b), *[_p(v, C.c_float) for v in f])
    return tuple(v[:n] for v in f)


def philox(ctr, key):
    c = np.ascontiguousarray(ctr, np.uint32)
    k = np.ascontiguousarray(key, np.uint32)
    o = np.zeros(4, np.uint32)
    lib.or_philox4x32_10(_p(c, C.c_uint32), _p(k, C.c_uint32), _p(o, C.c_uint32))
    return o


def temperature(it, total, t_hot=3.0, t_cold=0.02):
    return float(lib.or_temperature(it, total, t_hot, t_cold))


def sigma(it, total, t_hot=3.0, t_cold=0.02, noise_base=0.25):
    return float(lib.or_sigma(it, total, t_hot, t_cold, noise_base))


def langevin_step(x, y, z, res_off, res_ca, params: dict, forces=None, ag_ca=None, res_cg=None,
                  res_id0=0):
    """In-place step on float32 arrays x, y, z (src/diffusion.rs:300-339)."""
    p = OrStepParams()
    for k, v in params.items():
        if k == "center":
            for i in range(3):
                p.center[i] = float(v[i])
        else:
            setattr(p, k, v)
    res_off = np.ascontiguousarray(res_off, np.uint32)
    res_ca = np.ascontiguousarray(res_ca, np.uint32)
    n_res = len(res_ca)
    fx = fy = fz = None
    if forces is not None:
        fx, fy, fz = (_f32(v) for v in forces)
    keep = []
    agp = C.cast(None, C.POINTER(OrRCloud))
    rq = re = rs = None
    rh = None
    if ag_ca is not None:
        oc, k1 = rcloud(ag_ca)
        keep.append((oc, k1))
        agp = C.pointer(oc)
        rq, re, rs = (_f32(v) for v in res_cg[:3])
        rh = np.ascontiguousarray(res_cg[3], np.uint8)
    lib.or_langevin_step(_p(x, C.c_float), _p(y, C.c_float), _p(z, C.c_float), n_res,
                         _p(res_off, C.c_uint32), _p(res_ca, C.c_uint32), _p(fx, C.c_float),
                         _p(fy, C.c_float), _p(fz, C.c_float), agp, _p(rq, C.c_float),
                         _p(re, C.c_float), _p(rs, C.c_float), _p(rh, C.c_uint8), res_id0,
                         C.byref(p))


# ---- affinity maturation: the greedy scan restated literally ----

_ATOM_FIELDS = ("x", "y", "z", "charge", "r_min_half", "epsilon", "hydrophobic", "atom_type")


def concat_residues(residues):
    """One atom cloud of per-residue clouds, in residue order."""
    from types import SimpleNamespace
    return SimpleNamespace(**{k: np.concatenate([np.asarray(getattr(r, k)) for r in residues])
                              for k in _ATOM_FIELDS})


def allatom_energy_f64(ag, residues, aa, cutoff=10.0):
    """allatom_energy (src/diffusion.rs:281-283) in FP64: the five pair terms
    in total() order plus the disulfide wells over the SG atom (slot 5) of each
    Cys residue, in residue order."""
    t, _ = energy_full_f64(ag, concat_residues(residues), cutoff)
    sg = [(r.x[CYS_SG_SLOT], r.y[CYS_SG_SLOT], r.z[CYS_SG_SLOT])
          for r, a in zip(residues, aa) if a == CYS and len(r.x) > CYS_SG_SLOT]
    xs, ys, zs = (np.array([p[k] for p in sg], np.float32) for k in range(3))
    return ((((t[0] + t[1]) + t[2]) + t[3]) + t[4]) + disulfide_f64(xs, ys, zs)


def affinity_maturation_f64(ag, residues, aa, rot, ca, mutable, build_residue, cutoff=10.0,
                            trace=None):
    """affinity_maturation (src/diffusion.rs:924-961) with every trial scored
    by a full FP64 allatom_energy. residues (per-residue atom clouds), aa and
    rot (rotamer index per residue) are lists updated in place; ca[r] is
    residue r's Calpha; build_residue(aa, rot, ca) gives the atoms
    mutate_residue places (src/allatom.rs:299-384). Returns the best score;
    trace (a list) receives (position, type, score, best) of every trial."""
    best = allatom_energy_f64(ag, residues, aa, cutoff)
    for r in mutable:
        old, old_rot = aa[r], rot[r]
        for cand in range(20):                     # ALL_AA order, src/atom.rs:29-50
            if cand == old:
                continue
            residues[r], aa[r], rot[r] = build_residue(cand, 0, ca[r]), cand, 0
            e = allatom_energy_f64(ag, residues, aa, cutoff)
            if trace is not None:
                trace.append((r, cand, e, best))
            if e < best:                           # first strict improvement is kept
                best = e
                break
            residues[r], aa[r], rot[r] = build_residue(old, old_rot, ca[r]), old, old_rot
    return best


# ---- refinement (allatom_diffusion_step), the C port ----

class OrRfTables(C.Structure):
    _fields_ = [("natoms", u8p), ("ndef", u8p), ("defs", f32p), ("nrot", u8p), ("chi", f32p),
                ("ca_par", f32p), ("cg_par", f32p), ("hb_don", u8p), ("hb_acc", u8p)]


class OrRfParams(C.Structure):
    _fields_ = [("step", C.c_float), ("max_disp", C.c_float), ("restraint_k", C.c_float),
                ("restraint_radius", C.c_float), ("with_noise", C.c_int), ("seed_lo", C.c_uint32),
                ("seed_hi", C.c_uint32), ("mut_p", C.c_float), ("rot_p", C.c_float),
                ("bb_p", C.c_float), ("max_torsion", C.c_float), ("iface_sq", C.c_float),
                ("moves", C.c_uint32), ("center", C.c_float * 3)]


def refine_tables(cg_par):
    """Tables of the refinement port from the committed fixture (residue
    topologies, rotamers, side-chain placements, H-bond classes) and the
    residue-level parameters cg_par (20 x 4: charge, epsilon, sigma,
    hydrophobic) the caller supplies."""
    tab = reference_tables()
    natoms = np.array([len(r) for r in tab["residues"]], np.uint8)
    ndef = np.array([len(d) for d in tab["sidechain"]], np.uint8)
    defs = np.zeros((20, 12, 7), np.float32)
    for a, d in enumerate(tab["sidechain"]):
        if d:
            defs[a, :len(d)] = np.asarray(d, np.float32)
    nrot = np.array([len(r) for r in tab["rotamers"]], np.uint8)
    chi = np.zeros((20, 5, 4), np.float32)
    for a, r in enumerate(tab["rotamers"]):
        if r:
            chi[a, :len(r)] = np.asarray(r, np.float32)
    nb = np.asarray(tab["amber_nb"], np.float32)
    ca = np.zeros((20, 5), np.float32)
    for a, atoms in enumerate(tab["residues"]):
        t, q, h = atoms[1]
        ca[a] = (nb[t, 0], nb[t, 1], q, h, t)
    keep = [natoms, ndef, np.ascontiguousarray(defs), nrot, np.ascontiguousarray(chi),
            np.ascontiguousarray(ca), np.ascontiguousarray(cg_par, np.float32),
            np.asarray(tab["hbond_donor"], np.uint8), np.asarray(tab["hbond_acceptor"], np.uint8)]
    t = OrRfTables(_p(keep[0], C.c_uint8), _p(keep[1], C.c_uint8), _p(keep[2], C.c_float),
                   _p(keep[3], C.c_uint8), _p(keep[4], C.c_float), _p(keep[5], C.c_float),
                   _p(keep[6], C.c_float), _p(keep[7], C.c_uint8), _p(keep[8], C.c_uint8))
    return t, keep


def refine_run(tables, ag_ca, ag_aa, cand_res_off, res_aa, res_rot, slots, temp, sigma, iter0,
               params: dict, nthreads=1):
    """The C port of allatom_diffusion_step over len(temp) iterations, in
    place on res_aa, res_rot (uint8) and slots (n_res x 14 x 3 float32)."""
    t, _keep = tables
    oc, k1 = rcloud(ag_ca)
    p = OrRfParams()
    for k, v in params.items():
        if k == "center":
            for i in range(3):
                p.center[i] = float(v[i])
        else:
            setattr(p, k, v)
    ag_aa = np.ascontiguousarray(ag_aa, np.uint8)
    off = np.ascontiguousarray(cand_res_off, np.uint32)
    temp, sigma = _f32(temp), _f32(sigma)
    rc = lib.or_refine_run(C.addressof(t), C.byref(oc), _p(ag_aa, C.c_uint8), len(off) - 1,
                           _p(off, C.c_uint32), _p(res_aa, C.c_uint8), _p(res_rot, C.c_uint8),
                           _p(slots, C.c_float), _p(temp, C.c_float), _p(sigma, C.c_float),
                           len(temp), iter0, C.addressof(p), int(nthreads))
    assert rc == 0
