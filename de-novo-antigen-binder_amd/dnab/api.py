"""Python handles over the device-side C ABI.

GpuContext mirrors the reference's GpuContext (src/gpu.rs:46-57):
``GpuContext.try_init()`` returns None when no usable GPU exists
(src/gpu.rs:55-57, its async body src/gpu.rs:59-155 returns None at 65-71), and ``score_batch`` keeps the reference's signature
(src/gpu.rs:166-171). The energy/force methods mirror src/energy.rs.
"""
from __future__ import annotations

import ctypes as C
import weakref
from dataclasses import dataclass

import numpy as np

from . import _native as N
from .structure import (MAX_ATOMS_PER_RESIDUE, AtomCloud, AtomProtein, PopulationHost,
                        ResidueCloud, population_from_proteins)

vp = C.c_void_p


@dataclass
class AtomEnergyBreakdown:
    """src/energy.rs:238-250 (hbond and solvation are not computed: 0)."""
    lj: float
    coulomb: float
    hydrophobic: float
    hbond: float = 0.0
    solvation: float = 0.0

    def total(self) -> float:
        return self.lj + self.coulomb + self.hydrophobic + self.hbond + self.solvation


class GpuContext:
    """Owns a libdnab context. Objects made on it (grids, antigens,
    populations, networks, datasets) register here and are destroyed before
    the context, since their C handles refer to its stream."""

    def __init__(self, handle: int):
        self._h = vp(handle)
        self._children = weakref.WeakSet()

    def _adopt(self, obj):
        self._children.add(obj)

    @classmethod
    def try_init(cls, device: int = 0) -> "GpuContext | None":
        h = vp()
        st = N.lib.dnab_ctx_create(device, C.byref(h))
        if st == N.DNAB_ERR_NO_DEVICE:
            return None
        N.check(st)
        return cls(h.value)

    @property
    def handle(self):
        return self._h

    def close(self):
        if self._h:
            for obj in list(self._children):
                obj.close()
            N.lib.dnab_ctx_destroy(self._h)
            self._h = vp()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def check(self, status: int):
        N.check(status, self._h)

    def synchronize(self):
        self.check(N.lib.dnab_ctx_synchronize(self._h))

    def device_info(self):
        cu, ws = C.c_int(), C.c_int()
        arch = C.create_string_buffer(64)
        self.check(N.lib.dnab_ctx_device_info(self._h, C.byref(cu), C.byref(ws), arch, 64))
        return {"cu_count": cu.value, "wave_size": ws.value, "arch": arch.value.decode()}

    def set_timing(self, enable: bool):
        self.check(N.lib.dnab_ctx_set_timing(self._h, 1 if enable else 0))

    def kernel_time(self, name: str) -> tuple[float, int]:
        ms, n = C.c_double(), C.c_uint64()
        self.check(N.lib.dnab_ctx_kernel_time(self._h, name.encode(), C.byref(ms), C.byref(n)))
        return ms.value, n.value

    # ---- src/gpu.rs:166-267 ------------------------------------------------
    def score_batch(self, antigen: AtomCloud, candidates: np.ndarray, n_ab: int) -> np.ndarray:
        """candidates: (n_atoms, 8) float32 GpuAtom rows, candidate-major."""
        cand = np.ascontiguousarray(candidates, np.float32).reshape(-1, 8)
        n_atoms = cand.shape[0]
        n_cand = n_atoms // n_ab if n_ab else 0
        out = np.zeros(max(n_cand, 1), np.float32)
        got = C.c_size_t(0)
        agc = antigen.c()
        self.check(N.lib.dnab_score_batch(self._h, C.byref(agc),
                                          cand.ctypes.data_as(C.POINTER(N.GpuAtom)), n_atoms,
                                          n_ab, N.ptr(out, C.c_float), len(out), C.byref(got)))
        return out[:got.value]

    # ---- src/energy.rs ---------------------------------------------------------
    def interaction_energy_atoms_breakdown(self, antigen: AtomCloud, antibody: AtomCloud,
                                           cutoff: float = 10.0) -> AtomEnergyBreakdown:
        b = N.EnergyBreakdownC()
        a, c = antigen.c(), antibody.c()
        self.check(N.lib.dnab_interaction_energy_atoms_breakdown(self._h, C.byref(a), C.byref(c),
                                                                 cutoff, C.byref(b)))
        return AtomEnergyBreakdown(b.lj, b.coulomb, b.hydrophobic, b.hbond, b.solvation)

    def interaction_energy_atoms(self, antigen: AtomCloud, antibody: AtomCloud,
                                 cutoff: float = 10.0) -> float:
        e = C.c_double()
        a, c = antigen.c(), antibody.c()
        self.check(N.lib.dnab_interaction_energy_atoms(self._h, C.byref(a), C.byref(c), cutoff,
                                                       C.byref(e)))
        return e.value

    def allatom_energy(self, antigen, antibody: AtomProtein, cutoff: float = 10.0,
                       with_disulfide: bool = False):
        """allatom_energy (src/diffusion.rs:281-283): interaction_energy_atoms
        plus the antibody's disulfide_energy. antigen: AtomProtein or AtomCloud."""
        ag = antigen.atoms if isinstance(antigen, AtomProtein) else antigen
        a, c = ag.c(), antibody.atoms.c()
        rs = np.ascontiguousarray(antibody.res_start, np.uint32)
        ca = np.ascontiguousarray(antibody.ca_atom_idx, np.uint32)
        aa = np.ascontiguousarray(antibody.amino_acid, np.uint8)
        e, ds = C.c_double(), C.c_double()
        self.check(N.lib.dnab_allatom_energy(self._h, C.byref(a), C.byref(c), N.ptr(rs, C.c_uint32),
                                             N.ptr(ca, C.c_uint32), N.ptr(aa, C.c_uint8),
                                             len(aa), cutoff, C.byref(e), C.byref(ds)))
        return (e.value, ds.value) if with_disulfide else e.value

    def affinity_maturation(self, antigen, antibodies: list[AtomProtein], mutable,
                            rotamers=None, cutoff: float = 10.0):
        """affinity_maturation_allatom (src/diffusion.rs:955-961) for every
        antibody in one call. mutable[a]: ascending residue indices of antibody
        a; rotamers[a] (optional): each residue's rotamer index. Returns the
        final proteins, the scores (the reference's return value), and per
        antibody the rotamer index and rebuilt flag of every residue."""
        antibodies = list(antibodies)
        call = MaturationCall(self, antigen, population_from_proteins(antibodies), mutable,
                              rotamers, cutoff)
        call.run()
        return call.results()

    def refine(self, antigen: "DeviceAntigen", antibodies: list[AtomProtein], lp=None, mc=None,
               rotamers=None, energies: bool = True):
        """allatom_diffusion_step (src/diffusion.rs:290-416) lp.n_steps times
        for every antibody; returns the refined proteins, their rotamer
        indices and (energies) their allatom_energy."""
        call = RefineCall(self, antigen, population_from_proteins(list(antibodies)),
                          lp if lp is not None else langevin_params(), mc, rotamers, energies)
        call.run()
        return call.results()

    def atom_forces(self, antigen: AtomCloud, antibody: AtomCloud, cutoff: float = 10.0):
        n = len(antibody)
        fx, fy, fz = (np.zeros(max(n, 1), np.float32) for _ in range(3))
        e = C.c_double()
        a, c = antigen.c(), antibody.c()
        self.check(N.lib.dnab_atom_forces(self._h, C.byref(a), C.byref(c), cutoff,
                                          N.ptr(fx, C.c_float), N.ptr(fy, C.c_float),
                                          N.ptr(fz, C.c_float), C.byref(e)))
        return fx[:n], fy[:n], fz[:n], e.value

    def compute_forces_with_grid(self, antigen: ResidueCloud, grid: "SpatialHashGrid",
                                 antibody: ResidueCloud):
        n = len(antibody)
        fx, fy, fz = (np.zeros(max(n, 1), np.float32) for _ in range(3))
        a, b = antigen.c(), antibody.c()
        self.check(N.lib.dnab_cg_forces_with_grid(self._h, C.byref(a), grid.handle, C.byref(b),
                                                  N.ptr(fx, C.c_float), N.ptr(fy, C.c_float),
                                                  N.ptr(fz, C.c_float)))
        return fx[:n], fy[:n], fz[:n]

    def compute_forces(self, antigen: ResidueCloud, antibody: ResidueCloud):
        n = len(antibody)
        fx, fy, fz = (np.zeros(max(n, 1), np.float32) for _ in range(3))
        a, b = antigen.c(), antibody.c()
        self.check(N.lib.dnab_cg_forces(self._h, C.byref(a), C.byref(b), N.ptr(fx, C.c_float),
                                        N.ptr(fy, C.c_float), N.ptr(fz, C.c_float)))
        return fx[:n], fy[:n], fz[:n]


class SpatialHashGrid:
    """src/spatial.rs:12-68 on the device."""

    def __init__(self, ctx: GpuContext, cell_size: float):
        self.ctx = ctx
        ctx._adopt(self)
        h = vp()
        ctx.check(N.lib.dnab_grid_create(ctx.handle, cell_size, C.byref(h)))
        self._h = h

    @property
    def handle(self):
        return self._h

    def build(self, x, y, z):
        x, y, z = (np.ascontiguousarray(v, np.float32) for v in (x, y, z))
        self.ctx.check(N.lib.dnab_grid_build(self.ctx.handle, self._h, N.ptr(x, C.c_float),
                                             N.ptr(y, C.c_float), N.ptr(z, C.c_float), len(x)))

    def query_neighbors(self, px, py, pz):
        """Batched query: returns (offsets, indices) in the reference's order."""
        px, py, pz = (np.ascontiguousarray(np.atleast_1d(v), np.float32) for v in (px, py, pz))
        n = len(px)
        off = np.zeros(n + 1, np.uint32)
        total = C.c_size_t(0)
        st = N.lib.dnab_grid_query(self.ctx.handle, self._h, N.ptr(px, C.c_float),
                                   N.ptr(py, C.c_float), N.ptr(pz, C.c_float), n,
                                   N.ptr(off, C.c_uint32), N.ptr(None, C.c_uint32), 0,
                                   C.byref(total))
        if st not in (N.DNAB_OK, N.DNAB_ERR_CAPACITY):
            self.ctx.check(st)
        idx = np.zeros(max(total.value, 1), np.uint32)
        self.ctx.check(N.lib.dnab_grid_query(self.ctx.handle, self._h, N.ptr(px, C.c_float),
                                             N.ptr(py, C.c_float), N.ptr(pz, C.c_float), n,
                                             N.ptr(off, C.c_uint32), N.ptr(idx, C.c_uint32),
                                             len(idx), C.byref(total)))
        return off, idx[:total.value]

    def cell_of(self, x, y, z) -> np.ndarray:
        x, y, z = (np.ascontiguousarray(np.atleast_1d(v), np.float32) for v in (x, y, z))
        keys = np.zeros((max(len(x), 1), 3), np.int32)
        self.ctx.check(N.lib.dnab_grid_cell_of(self.ctx.handle, self._h, N.ptr(x, C.c_float),
                                               N.ptr(y, C.c_float), N.ptr(z, C.c_float), len(x),
                                               N.ptr(keys, C.c_int32)))
        return keys[:len(x)]

    def info(self):
        n, ne = C.c_size_t(), C.c_size_t()
        kmin, kmax = (C.c_int32 * 3)(), (C.c_int32 * 3)()
        self.ctx.check(N.lib.dnab_grid_info(self._h, C.byref(n), C.byref(ne), kmin, kmax))
        return {"n_points": n.value, "n_cells_nonempty": ne.value, "key_min": list(kmin),
                "key_max": list(kmax)}

    def close(self):
        if self._h:
            N.lib.dnab_grid_destroy(self._h)
            self._h = vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeviceAntigen:
    """A device-resident antigen (Morton-ordered cell-major SoA in HBM)."""

    def __init__(self, ctx: GpuContext, protein: AtomProtein | AtomCloud, cutoff: float = 10.0,
                 cell_size: float | None = None, center=None, with_ca: bool = True,
                 restraint_radius: float = 0.0):
        self.ctx = ctx
        ctx._adopt(self)
        if isinstance(protein, AtomProtein):
            cloud = protein.atoms
            ca = protein.ca_residue_cloud() if with_ca else None
            ctr = protein.ca_center_of_mass() if center is None else center
        else:
            cloud, ca = protein, None
            ctr = np.zeros(3, np.float32) if center is None else center
        self.cloud, self.ca = cloud, ca
        self.center = np.asarray(ctr, np.float32)
        d = N.AntigenDescC()
        d.atoms = cloud.c()
        d.ca = ca.c() if ca is not None else N.ResidueCloudC()
        for k in range(3):
            d.center[k] = float(self.center[k])
        if ca is not None:
            d.ca_amino_acid = N.ptr(ca.amino_acid, C.c_uint8)
        d.cell_size = float(cutoff if cell_size is None else cell_size)
        d.cutoff = float(cutoff)
        d.restraint_radius = float(restraint_radius)
        self.restraint_radius = float(restraint_radius)
        self.cutoff = float(cutoff)
        h = vp()
        ctx.check(N.lib.dnab_antigen_create(ctx.handle, C.byref(d), C.byref(h)))
        self._h = h

    @property
    def handle(self):
        return self._h

    def info(self):
        n, ne = C.c_size_t(), C.c_size_t()
        rc, cs = C.c_float(), C.c_float()
        self.ctx.check(N.lib.dnab_antigen_info(self._h, C.byref(n), C.byref(ne), C.byref(rc),
                                               C.byref(cs)))
        return {"n_atoms": n.value, "n_cells_nonempty": ne.value, "cutoff": rc.value,
                "cell_size": cs.value}

    def build_maps(self, spacing: float = 0.0):
        """(Re)build the prescreen's antigen maps (dnab_antigen_build_maps);
        spacing <= 0 selects the default 0.625 A."""
        self.ctx.check(N.lib.dnab_antigen_build_maps(self.ctx.handle, self._h, spacing))

    def map_info(self):
        h = C.c_float()
        o = np.zeros(3, np.float32)
        d = np.zeros(3, np.int32)
        b = C.c_size_t()
        self.ctx.check(N.lib.dnab_antigen_map_info(self._h, C.byref(h), N.ptr(o, C.c_float),
                                                   N.ptr(d, C.c_int32), C.byref(b)))
        return {"spacing": h.value, "origin": o, "dims": d, "bytes": b.value}

    def map_points(self, ijk) -> np.ndarray:
        """Map values at grid points ijk [n, 3]: [n, 18] (7 LJ classes,
        potential, hydrophobic, 8 radial densities, the occupancy bit of the
        cell whose lower corner the point is)."""
        ijk = np.ascontiguousarray(ijk, np.int32).reshape(-1, 3)
        out = np.zeros((max(len(ijk), 1), N.MAP_VALUES), np.float32)
        self.ctx.check(N.lib.dnab_antigen_map_points(self.ctx.handle, self._h,
                                                     N.ptr(ijk, C.c_int32), len(ijk),
                                                     N.ptr(out, C.c_float)))
        return out[:len(ijk)]

    def close(self):
        if self._h:
            N.lib.dnab_antigen_destroy(self._h)
            self._h = vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class _ProteinOut:
    """Caller-allocated dnab_protein_out buffers for n_res residues, and their
    split into one AtomProtein per candidate of a population description."""

    def __init__(self, n_res: int):
        cap = max(1, MAX_ATOMS_PER_RESIDUE * n_res)
        self.f = {k: np.zeros(cap, np.float32) for k in ("x", "y", "z", "charge", "rmh", "eps")}
        self.hyd, self.at = np.zeros(cap, np.uint8), np.zeros(cap, np.uint8)
        self.ri = np.zeros(cap, np.uint32)
        self.rs = np.zeros(n_res + 1, np.uint32)
        self.ca = np.zeros(max(n_res, 1), np.uint32)
        self.aa = np.zeros(max(n_res, 1), np.uint8)
        f = self.f
        self.c = N.ProteinOutC(cap, N.ptr(f["x"], C.c_float), N.ptr(f["y"], C.c_float),
                               N.ptr(f["z"], C.c_float), N.ptr(f["charge"], C.c_float),
                               N.ptr(f["rmh"], C.c_float), N.ptr(f["eps"], C.c_float),
                               N.ptr(self.hyd, C.c_uint8), N.ptr(self.at, C.c_uint8),
                               N.ptr(self.ri, C.c_uint32), N.ptr(self.rs, C.c_uint32),
                               N.ptr(self.ca, C.c_uint32), N.ptr(self.aa, C.c_uint8))

    def proteins(self, cand_res_off) -> list[AtomProtein]:
        f, rs = self.f, self.rs
        out = []
        for a in range(len(cand_res_off) - 1):
            g0, g1 = int(cand_res_off[a]), int(cand_res_off[a + 1])
            a0, a1 = int(rs[g0]), int(rs[g1])
            cloud = AtomCloud(f["x"][a0:a1].copy(), f["y"][a0:a1].copy(), f["z"][a0:a1].copy(),
                              f["charge"][a0:a1].copy(), f["rmh"][a0:a1].copy(),
                              f["eps"][a0:a1].copy(), self.hyd[a0:a1].copy(),
                              self.at[a0:a1].copy(), self.ri[a0:a1].copy())
            out.append(AtomProtein(cloud, (rs[g0:g1 + 1] - a0).astype(np.uint32),
                                   self.aa[g0:g1].copy(), (self.ca[g0:g1] - a0).astype(np.uint32)))
        return out


def mc_params(**kw) -> N.McParamsC:
    """Monte-Carlo move rates (the reference's by default)."""
    p = N.McParamsC()
    N.check(N.lib.dnab_mc_default_params(C.byref(p)))
    for k, v in kw.items():
        if not hasattr(p, k):
            raise AttributeError(f"unknown Monte-Carlo parameter {k}")
        setattr(p, k, v)
    return p


def run_allatom(ctx: "GpuContext", antigen: "DeviceAntigen", **kw):
    """run_allatom (src/diffusion.rs:983-1107) in one call (dnab_run_allatom):
    keyword arguments set dnab_run_allatom_params fields (ab_length,
    population, cpu_iterations, top_n, n_broad, broad_steps, broad_moves,
    seed, ring_radius). Returns [(AtomProtein, allatom_energy, sequence)]
    in ascending energy, as the reference's Vec<AllAtomResult>."""
    p = N.RunAllatomParamsC()
    N.check(N.lib.dnab_run_allatom_default_params(C.byref(p)))
    for k, v in kw.items():
        if k.startswith("_") or not hasattr(p, k):
            raise AttributeError(f"unknown run_allatom parameter {k}")
        setattr(p, k, v)
    ab_len = p.ab_length or (len(antigen.ca) if antigen.ca is not None else 0)
    pool = p.population if p.n_broad == 0 else min(p.population, p.n_broad)
    n_keep = min(max(p.top_n, 1), pool)
    buf = _ProteinOut(n_keep * ab_len)
    e = np.zeros(max(n_keep, 1), np.float64)
    n, na = C.c_uint32(0), C.c_size_t(0)
    ctx.check(N.lib.dnab_run_allatom(ctx.handle, antigen.handle, C.byref(p), C.byref(buf.c),
                                     N.ptr(e, C.c_double), C.byref(n), C.byref(na)))
    off = np.arange(n.value + 1, dtype=np.uint32) * ab_len
    prots = buf.proteins(off)
    return [(pr, float(e[k]), pr.sequence()) for k, pr in enumerate(prots)]


class RefineCall:
    """One dnab_refine call (allatom_diffusion_step for every candidate, all
    iterations in one launch) with inputs packed and outputs allocated up
    front, so that run() is the C call alone."""

    def __init__(self, ctx: "GpuContext", antigen: "DeviceAntigen", host: PopulationHost,
                 lp: N.LangevinParamsC, mc: N.McParamsC | None = None, rotamers=None,
                 energies: bool = True):
        self.ctx, self.host, self.antigen, self.lp = ctx, host, antigen, lp
        self.mc = mc if mc is not None else mc_params()
        n_res = host.n_res
        self.rot = None if rotamers is None else np.ascontiguousarray(
            np.concatenate([np.asarray(r, np.uint8) for r in rotamers]), np.uint8)
        self.buf = _ProteinOut(n_res)
        self.rot_out = np.zeros(max(n_res, 1), np.uint8)
        self.energy = np.zeros(max(host.n_cand, 1), np.float64) if energies else None
        self.na = C.c_size_t(0)
        self.desc = host.c()

    def run(self):
        self.ctx.check(N.lib.dnab_refine(
            self.ctx.handle, self.antigen.handle, C.byref(self.desc), N.ptr(self.rot, C.c_uint8),
            C.byref(self.lp), C.byref(self.mc), C.byref(self.buf.c),
            N.ptr(self.rot_out, C.c_uint8), N.ptr(self.energy, C.c_double), C.byref(self.na)))

    def results(self):
        off = self.host.cand_res_off
        rots = [self.rot_out[int(off[a]):int(off[a + 1])].copy() for a in range(self.host.n_cand)]
        e = None if self.energy is None else self.energy[:self.host.n_cand].copy()
        return self.buf.proteins(off), rots, e


class MaturationCall:
    """One dnab_affinity_maturation call with its inputs packed and its output
    buffers allocated up front, so that run() is the C call alone."""

    def __init__(self, ctx: "GpuContext", antigen, host: PopulationHost, mutable,
                 rotamers=None, cutoff: float = 10.0):
        """antigen: a DeviceAntigen (its cutoff applies), or an AtomProtein /
        AtomCloud placed on the device here with `cutoff`."""
        self.ctx, self.host = ctx, host
        if not isinstance(antigen, DeviceAntigen):
            antigen = DeviceAntigen(ctx, antigen, cutoff=cutoff, with_ca=False)
        self.antigen = antigen
        n_ab, n_res = host.n_cand, host.n_res
        self.moff = np.zeros(n_ab + 1, np.uint32)
        for a, m in enumerate(mutable):
            self.moff[a + 1] = self.moff[a] + len(m)
        self.mres = np.ascontiguousarray(
            np.concatenate([np.asarray(m, np.int64) for m in mutable]) if n_ab else np.zeros(0),
            np.uint32)
        self.rot = None if rotamers is None else np.ascontiguousarray(
            np.concatenate([np.asarray(r, np.uint8) for r in rotamers]), np.uint8)
        self.buf = _ProteinOut(n_res)
        self.rot_out = np.zeros(max(n_res, 1), np.uint8)
        self.rebuilt = np.zeros(max(n_res, 1), np.uint8)
        self.energy = np.zeros(max(n_ab, 1), np.float64)
        self.na = C.c_size_t(0)
        self.desc = host.c()

    def run(self):
        self.ctx.check(N.lib.dnab_affinity_maturation(
            self.ctx.handle, self.antigen.handle, C.byref(self.desc),
            N.ptr(self.rot, C.c_uint8), N.ptr(self.moff, C.c_uint32),
            N.ptr(self.mres, C.c_uint32), C.byref(self.buf.c), N.ptr(self.rot_out, C.c_uint8),
            N.ptr(self.rebuilt, C.c_uint8), N.ptr(self.energy, C.c_double), C.byref(self.na)))

    def results(self):
        host = self.host
        off = host.cand_res_off
        rots = [self.rot_out[int(off[a]):int(off[a + 1])].copy() for a in range(host.n_cand)]
        flags = [self.rebuilt[int(off[a]):int(off[a + 1])].astype(bool) for a in range(host.n_cand)]
        return self.buf.proteins(off), self.energy[:host.n_cand].copy(), rots, flags


def langevin_params(**kw) -> N.LangevinParamsC:
    p = N.LangevinParamsC()
    N.check(N.lib.dnab_langevin_default_params(C.byref(p)))
    for k, v in kw.items():
        if not hasattr(p, k):
            raise AttributeError(f"unknown langevin parameter {k}")
        setattr(p, k, v)
    return p


class DevicePopulation:
    """A ragged batch of candidates in HBM, each bound to one antigen."""

    def __init__(self, ctx: GpuContext, antigens: list[DeviceAntigen], host: PopulationHost):
        self.ctx = ctx
        ctx._adopt(self)
        self.antigens = list(antigens)
        self.host = host
        arr = (vp * len(antigens))(*[a.handle.value for a in antigens])
        desc = host.c()
        h = vp()
        ctx.check(N.lib.dnab_population_create(ctx.handle, arr, len(antigens), C.byref(desc),
                                               C.byref(h)))
        self._h = h
        self.n_cand, self.n_res, self.n_atoms = host.n_cand, host.n_res, len(host.atoms)

    @property
    def handle(self):
        return self._h

    def evaluate(self, forces: bool = True, breakdown: bool = False, fp64: bool = False):
        flags = (N.DNAB_EVAL_FORCES if forces else 0) | (N.DNAB_EVAL_BREAKDOWN if breakdown else 0)
        flags |= N.DNAB_EVAL_FP64 if fp64 else 0
        self.ctx.check(N.lib.dnab_population_evaluate(self.ctx.handle, self._h, flags))

    def energies(self):
        e = np.zeros(max(self.n_cand, 1), np.float64)
        err = np.zeros(max(self.n_cand, 1), np.float32)
        self.ctx.check(N.lib.dnab_population_read_energies(self.ctx.handle, self._h,
                                                           N.ptr(e, C.c_double),
                                                           N.ptr(err, C.c_float)))
        return e[:self.n_cand], err[:self.n_cand]

    def status(self):
        """Per candidate, DNAB_STATUS_* bits of the last evaluation: 1 NaN
        energy (a non-finite coordinate), 2 a saturated sum (reported as inf)."""
        out = np.zeros(max(self.n_cand, 1), np.uint8)
        self.ctx.check(N.lib.dnab_population_read_status(self.ctx.handle, self._h,
                                                         N.ptr(out, C.c_uint8)))
        return out[:self.n_cand]

    def terms(self):
        """(n_cand, 6) of the last breakdown evaluation: LJ, Coulomb,
        hydrophobic, H-bond, solvation, disulfide."""
        out = np.zeros((max(self.n_cand, 1), 6), np.float64)
        self.ctx.check(N.lib.dnab_population_read_terms(self.ctx.handle, self._h,
                                                        N.ptr(out, C.c_double)))
        return out[:self.n_cand]

    def breakdown(self):
        out = [np.zeros(max(self.n_cand, 1), np.float64) for _ in range(3)]
        self.ctx.check(N.lib.dnab_population_read_breakdown(
            self.ctx.handle, self._h, *[N.ptr(o, C.c_double) for o in out]))
        return [o[:self.n_cand] for o in out]

    def forces(self):
        out = [np.zeros(max(self.n_atoms, 1), np.float32) for _ in range(3)]
        self.ctx.check(N.lib.dnab_population_read_forces(
            self.ctx.handle, self._h, *[N.ptr(o, C.c_float) for o in out]))
        return [o[:self.n_atoms] for o in out]

    def forces_f64(self):
        out = [np.zeros(max(self.n_atoms, 1), np.float64) for _ in range(3)]
        self.ctx.check(N.lib.dnab_population_read_forces_f64(
            self.ctx.handle, self._h, *[N.ptr(o, C.c_double) for o in out]))
        return [o[:self.n_atoms] for o in out]

    def coords(self):
        out = [np.zeros(max(self.n_atoms, 1), np.float32) for _ in range(3)]
        self.ctx.check(N.lib.dnab_population_read_coords(
            self.ctx.handle, self._h, *[N.ptr(o, C.c_float) for o in out]))
        return [o[:self.n_atoms] for o in out]

    def set_coords(self, x, y, z):
        x, y, z = (np.ascontiguousarray(v, np.float32) for v in (x, y, z))
        self.ctx.check(N.lib.dnab_population_write_coords(self.ctx.handle, self._h,
                                                          N.ptr(x, C.c_float),
                                                          N.ptr(y, C.c_float),
                                                          N.ptr(z, C.c_float)))

    def tuning(self) -> dict:
        t = N.PairTuningC()
        N.check(N.lib.dnab_population_get_tuning(self._h, C.byref(t)))
        return {"direct_survivors": t.direct_survivors, "direct_chunks": t.direct_chunks,
                "rest_waves": t.rest_waves}

    def set_tuning(self, direct_survivors: int, direct_chunks: int, rest_waves: int):
        t = N.PairTuningC(direct_survivors, direct_chunks, rest_waves)
        self.ctx.check(N.lib.dnab_population_set_tuning(self.ctx.handle, self._h, C.byref(t)))

    def trajectory_tuning(self) -> dict:
        t = N.TrajTuningC()
        N.check(N.lib.dnab_population_get_trajectory_tuning(self._h, C.byref(t)))
        return {"waves": t.waves, "cand_per_group": t.cand_per_group, "balance": t.balance,
                "cells_per_item": t.cells_per_item, "chunks_per_item": t.chunks_per_item,
                "rebalance_steps": t.rebalance_steps, "engine": t.engine}

    def set_trajectory_tuning(self, waves: int, cand_per_group: int, balance: int = 1,
                              cells_per_item: int = 2, chunks_per_item: int = 2,
                              rebalance_steps: int = 0, engine: int = 0):
        t = N.TrajTuningC(waves, cand_per_group, balance, cells_per_item, chunks_per_item,
                          rebalance_steps, engine)
        self.ctx.check(N.lib.dnab_population_set_trajectory_tuning(self.ctx.handle, self._h,
                                                                   C.byref(t)))

    def autotune_trajectory(self, params: N.LangevinParamsC, trial_steps: int = 20) -> dict:
        """Pick the fastest fused-loop shape for these loop parameters and
        keep it (the coordinates are restored; results do not depend on it)."""
        t = N.TrajTuningC()
        self.ctx.check(N.lib.dnab_population_autotune_trajectory(
            self.ctx.handle, self._h, C.byref(params), trial_steps, C.byref(t)))
        return self.trajectory_tuning()

    def autotune(self, reps: int = 3) -> tuple[dict, float]:
        """Pick the fastest pair-kernel work split for the current state."""
        t = N.PairTuningC()
        ms = C.c_double()
        self.ctx.check(N.lib.dnab_population_autotune(self.ctx.handle, self._h, reps, C.byref(t),
                                                      C.byref(ms)))
        return ({"direct_survivors": t.direct_survivors, "direct_chunks": t.direct_chunks,
                 "rest_waves": t.rest_waves}, ms.value)

    def pair_count(self):
        a, b = C.c_uint64(), C.c_uint64()
        self.ctx.check(N.lib.dnab_population_pair_count(self.ctx.handle, self._h, C.byref(a),
                                                        C.byref(b)))
        return a.value, b.value

    def pair_stats(self, reset: bool = False):
        a, b = C.c_uint64(), C.c_uint64()
        self.ctx.check(N.lib.dnab_population_read_pair_stats(self.ctx.handle, self._h,
                                                             C.byref(a), C.byref(b)))
        if reset:
            self.ctx.check(N.lib.dnab_population_reset_pair_stats(self.ctx.handle, self._h))
        return a.value, b.value

    def reset_pair_stats(self):
        self.ctx.check(N.lib.dnab_population_reset_pair_stats(self.ctx.handle, self._h))

    def step(self, params: N.LangevinParamsC, iteration: int):
        self.ctx.check(N.lib.dnab_langevin_step(self.ctx.handle, self._h, C.byref(params),
                                                iteration))

    def run(self, params: N.LangevinParamsC):
        k = max(1, params.top_k)
        idx = np.zeros(k, np.uint32)
        e = np.zeros(k, np.float64)
        n = C.c_uint32()
        self.ctx.check(N.lib.dnab_langevin_run(self.ctx.handle, self._h, C.byref(params),
                                               N.ptr(idx, C.c_uint32), N.ptr(e, C.c_double),
                                               C.byref(n)))
        return idx[:n.value], e[:n.value]

    def scores(self):
        """Score history of the last run: (idx, energy), each (n_scores, k);
        row j is the top k of the j-th scoring, the last row the final one."""
        n, k = C.c_uint32(), C.c_uint32()
        self.ctx.check(N.lib.dnab_population_read_scores(self.ctx.handle, self._h, C.byref(n),
                                                         C.byref(k), None, None))
        idx = np.zeros((max(n.value, 1), max(k.value, 1)), np.uint32)
        e = np.zeros_like(idx, dtype=np.float64)
        self.ctx.check(N.lib.dnab_population_read_scores(self.ctx.handle, self._h, None, None,
                                                         N.ptr(idx, C.c_uint32),
                                                         N.ptr(e, C.c_double)))
        return idx[:n.value, :k.value], e[:n.value, :k.value]

    def close(self):
        if self._h:
            N.lib.dnab_population_destroy(self._h)
            self._h = vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Mlp:
    """The prescreen MLP (32 -> 256 -> 256 -> 256 -> 1, SiLU, bf16 weights)."""

    def __init__(self, ctx: GpuContext, seed: int = 0):
        self.ctx = ctx
        ctx._adopt(self)
        h = vp()
        ctx.check(N.lib.dnab_mlp_create(ctx.handle, seed, C.byref(h)))
        self._h = h
        n = C.c_size_t()
        N.check(N.lib.dnab_mlp_param_count(C.byref(n)))
        self.n_params = n.value

    @property
    def handle(self):
        return self._h

    def get_params(self) -> np.ndarray:
        p = np.zeros(self.n_params, np.float32)
        self.ctx.check(N.lib.dnab_mlp_get_params(self.ctx.handle, self._h, N.ptr(p, C.c_float),
                                                 self.n_params))
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        self.ctx.check(N.lib.dnab_mlp_set_params(self.ctx.handle, self._h, N.ptr(p, C.c_float),
                                                 self.n_params))

    def loss_grad(self, x_bf16: np.ndarray, t: np.ndarray):
        x = np.ascontiguousarray(x_bf16, np.uint16)
        t = np.ascontiguousarray(t, np.float32)
        B = len(t)
        g = np.zeros(self.n_params, np.float32)
        pred = np.zeros(B, np.float32)
        loss = C.c_float()
        self.ctx.check(N.lib.dnab_mlp_loss_grad(self.ctx.handle, self._h, N.ptr(x, C.c_uint16),
                                                N.ptr(t, C.c_float), B, N.ptr(g, C.c_float),
                                                C.byref(loss), N.ptr(pred, C.c_float)))
        return loss.value, g, pred

    def train(self, ds: "Dataset", **kw):
        p = N.TrainParamsC()
        N.check(N.lib.dnab_train_default_params(C.byref(p)))
        for k, v in kw.items():
            setattr(p, k, v)
        hist = np.zeros(max(p.n_steps, 1), np.float32)
        self.ctx.check(N.lib.dnab_mlp_train(self.ctx.handle, self._h, ds.handle, C.byref(p),
                                            N.ptr(hist, C.c_float)))
        return hist[:p.n_steps]

    def predict(self, pop: DevicePopulation) -> np.ndarray:
        out = np.zeros(max(pop.n_cand, 1), np.float32)
        self.ctx.check(N.lib.dnab_mlp_predict(self.ctx.handle, self._h, pop.handle,
                                              N.ptr(out, C.c_float)))
        return out[:pop.n_cand]

    def prescreen(self, pop: DevicePopulation, n_keep: int, top_k: int):
        idx = np.zeros(top_k, np.uint32)
        e = np.zeros(top_k, np.float64)
        n = C.c_uint32()
        self.ctx.check(N.lib.dnab_prescreen(self.ctx.handle, self._h, pop.handle, n_keep, top_k,
                                            N.ptr(idx, C.c_uint32), N.ptr(e, C.c_double),
                                            C.byref(n)))
        return idx[:n.value], e[:n.value]

    def close(self):
        if self._h:
            N.lib.dnab_mlp_destroy(self._h)
            self._h = vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Dataset:
    """Device-resident (features, target) rows for training the prescreen."""

    def __init__(self, ctx: GpuContext, capacity: int, energy_scale: float = 10.0,
                 energy_cap: float = 0.0):
        self.ctx = ctx
        ctx._adopt(self)
        self.energy_scale = energy_scale
        self.energy_cap = energy_cap
        h = vp()
        ctx.check(N.lib.dnab_dataset_create(ctx.handle, capacity, energy_scale, energy_cap,
                                            C.byref(h)))
        self._h = h

    @property
    def handle(self):
        return self._h

    def add(self, pop: DevicePopulation):
        self.ctx.check(N.lib.dnab_dataset_add_population(self.ctx.handle, self._h, pop.handle))

    def __len__(self):
        n = C.c_size_t()
        N.check(N.lib.dnab_dataset_size(self._h, C.byref(n)))
        return n.value

    def close(self):
        if self._h:
            N.lib.dnab_dataset_destroy(self._h)
            self._h = vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def precision_study(pop: "DevicePopulation", n_atoms_per_cand=None) -> dict:
    """FP32 path (per-chunk TwoSum, exact fixed-point across chunks, FP64
    totals) against the all-FP64 kernel on the same pair set: per-candidate
    energy error, per-atom force RMS error, and how often the a-priori error
    estimate covers the measured error."""
    pop.evaluate(forces=True)
    e32, est = pop.energies()
    f32 = np.stack(pop.forces(), 1).astype(np.float64)
    pop.evaluate(fp64=True)
    e64, _ = pop.energies()
    f64 = np.stack(pop.forces_f64(), 1)
    de = np.abs(e32 - e64)
    rel = de / np.maximum(np.abs(e64), 1e-30)
    df = f32 - f64
    tol_e = 1e-3 + 1e-4 * np.abs(e64)
    tol_f = 1e-3 + 1e-4 * np.abs(f64)
    return {"candidates": int(len(e32)), "atoms": int(len(f32)),
            "energy_abs_err_max": float(de.max()), "energy_abs_err_median": float(np.median(de)),
            "energy_rel_err_max": float(rel.max()),
            "force_rms_err": float(np.sqrt(np.mean(df ** 2))),
            "force_rms": float(np.sqrt(np.mean(f64 ** 2))),
            "force_abs_err_max": float(np.abs(df).max()),
            "energy_within_tol": float(np.mean(de <= tol_e)),
            "force_within_tol": float(np.mean(np.abs(df) <= tol_f)),
            "estimate_covers": float(np.mean(de <= est + 1e-12)),
            "energy_range": [float(e64.min()), float(np.median(e64)), float(e64.max())]}


def population_features(pop: DevicePopulation) -> np.ndarray:
    out = np.zeros((max(pop.n_cand, 1), N.FEATURE_DIM), np.uint16)
    pop.ctx.check(N.lib.dnab_population_features(pop.ctx.handle, pop.handle,
                                                 N.ptr(out, C.c_uint16)))
    return out[:pop.n_cand]


def select_smallest(ctx: GpuContext, scores, m: int) -> np.ndarray:
    """Indices of the min(m, n) smallest scores (ties by index, NaN last), in
    ascending index order (dnab_select_smallest)."""
    s = np.ascontiguousarray(scores, np.float32)
    out = np.zeros(max(m, 1), np.uint32)
    n = C.c_uint32()
    ctx.check(N.lib.dnab_select_smallest(ctx.handle, N.ptr(s, C.c_float), len(s), m,
                                         N.ptr(out, C.c_uint32), C.byref(n)))
    return out[:n.value]


def map_energy(pop: DevicePopulation) -> np.ndarray:
    """Per candidate map energy (dnab_population_map_energy): the ranking
    score of the screen without the network."""
    out = np.zeros(max(pop.n_cand, 1), np.float32)
    pop.ctx.check(N.lib.dnab_population_map_energy(pop.ctx.handle, pop.handle,
                                                   N.ptr(out, C.c_float)))
    return out[:pop.n_cand]


def exact_topk(pop: DevicePopulation, top_k: int):
    idx = np.zeros(top_k, np.uint32)
    e = np.zeros(top_k, np.float64)
    n = C.c_uint32()
    pop.ctx.check(N.lib.dnab_exact_topk(pop.ctx.handle, pop.handle, top_k,
                                        N.ptr(idx, C.c_uint32), N.ptr(e, C.c_double),
                                        C.byref(n)))
    return idx[:n.value], e[:n.value]
