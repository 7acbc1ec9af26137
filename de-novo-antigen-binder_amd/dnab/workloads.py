"""Synthetic inputs for the five measured shapes.

c1  16-residue peptide antigen (parse_peptide, src/pdb.rs:64-83), 64 binders
    of 12 residues in the reference ring layout (src/diffusion.rs:264-277),
    cutoff 10, grid cell 8.
c2  300-residue globule antigen (Calpha random walk, 3.8 A steps, inside 25 A),
    1,024 binders of 20 residues starting on a shell at the antigen surface
    (globule radius + 3 A, walks kept outside the globule radius), restraint
    radius = shell radius, cutoff 10, cell 10.
c3  8 globule antigens of 100..1,000 residues (radius 25 (N/300)^(1/3)),
    4,096 binders of U{8..40} residues, 512 per antigen, cutoff 12, cell 12.
c4  5,000-residue globule, 1,024 binders of 30 residues, cutoff 12, cell 12.
c5  prescreen pool: binders of 20 residues against c2-like antigens.

The antigens and binders have no reference counterpart beyond the two
layouts named above; they are seeded and built with dnab_build_protein.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _native as N
from .structure import (AtomCloud, AtomProtein, PopulationHost, parse_peptide,
                        population_from_proteins, protein_from_ca_trace)


def globule_antigen(n_res: int, radius: float, seed: int) -> AtomProtein:
    x, y, z = (np.zeros(n_res, np.float32) for _ in range(3))
    aa = np.zeros(n_res, np.uint8)
    rot = np.zeros(n_res, np.uint8)
    N.check(N.lib.dnab_synth_globule_trace(n_res, radius, seed, N.ptr(x, C.c_float),
                                           N.ptr(y, C.c_float), N.ptr(z, C.c_float),
                                           N.ptr(aa, C.c_uint8), N.ptr(rot, C.c_uint8)))
    return protein_from_ca_trace(x, y, z, aa, rot)


def peptide_antigen(seq: str) -> AtomProtein:
    x, y, z, aa = parse_peptide(seq)
    return protein_from_ca_trace(x, y, z, aa)


def random_sequence(n: int, seed: int) -> str:
    rng = np.random.default_rng(seed)
    return "".join("ARNDCQEGHILKMFPSTWYV"[i] for i in rng.integers(0, 20, n))


def ring_binder(n_res: int, center, seed: int, ring_radius: float = 20.0) -> AtomProtein:
    x, y, z = (np.zeros(n_res, np.float32) for _ in range(3))
    aa = np.zeros(n_res, np.uint8)
    ctr = np.ascontiguousarray(center, np.float32)
    N.check(N.lib.dnab_synth_binder_ring(n_res, N.ptr(ctr, C.c_float), ring_radius, seed,
                                         N.ptr(x, C.c_float), N.ptr(y, C.c_float),
                                         N.ptr(z, C.c_float), N.ptr(aa, C.c_uint8)))
    return protein_from_ca_trace(x, y, z, aa)


def shell_binder(n_res: int, center, shell_radius: float, min_radius: float,
                 seed: int) -> AtomProtein:
    x, y, z = (np.zeros(n_res, np.float32) for _ in range(3))
    aa = np.zeros(n_res, np.uint8)
    ctr = np.ascontiguousarray(center, np.float32)
    N.check(N.lib.dnab_synth_binder_shell(n_res, N.ptr(ctr, C.c_float), shell_radius,
                                          min_radius, seed, N.ptr(x, C.c_float),
                                          N.ptr(y, C.c_float), N.ptr(z, C.c_float),
                                          N.ptr(aa, C.c_uint8)))
    return protein_from_ca_trace(x, y, z, aa)


def shell_population(n_cand: int, n_res: int, center, shell_radius: float, min_radius: float,
                     seed_base: int, antigen_id: int = 0, nthreads: int = 16) -> PopulationHost:
    """n_cand shell binders at once (dnab_synth_shell_population): the same
    candidates as shell_binder(n_res, center, shell_radius, min_radius,
    seed_base + k) for k < n_cand, built by host threads."""
    ctr = np.ascontiguousarray(center, np.float32)
    na = C.c_size_t(0)
    N.check(N.lib.dnab_synth_shell_population(n_cand, n_res, N.ptr(ctr, C.c_float), shell_radius,
                                              min_radius, seed_base, nthreads, None, None,
                                              C.byref(na)))
    m = na.value
    nr = n_cand * n_res
    f = {k: np.zeros(max(m, 1), np.float32) for k in ("x", "y", "z", "charge", "rmh", "eps")}
    hyd = np.zeros(max(m, 1), np.uint8)
    at = np.zeros(max(m, 1), np.uint8)
    ri = np.zeros(max(m, 1), np.uint32)
    rs = np.zeros(nr + 1, np.uint32)
    ca = np.zeros(max(nr, 1), np.uint32)
    aa = np.zeros(max(nr, 1), np.uint8)
    off = np.zeros(n_cand + 1, np.uint32)
    out = N.ProteinOutC(m, N.ptr(f["x"], C.c_float), N.ptr(f["y"], C.c_float),
                        N.ptr(f["z"], C.c_float), N.ptr(f["charge"], C.c_float),
                        N.ptr(f["rmh"], C.c_float), N.ptr(f["eps"], C.c_float),
                        N.ptr(hyd, C.c_uint8), N.ptr(at, C.c_uint8), N.ptr(ri, C.c_uint32),
                        N.ptr(rs, C.c_uint32), N.ptr(ca, C.c_uint32), N.ptr(aa, C.c_uint8))
    N.check(N.lib.dnab_synth_shell_population(n_cand, n_res, N.ptr(ctr, C.c_float), shell_radius,
                                              min_radius, seed_base, nthreads, C.byref(out),
                                              N.ptr(off, C.c_uint32), C.byref(na)))
    cloud = AtomCloud(f["x"][:m], f["y"][:m], f["z"][:m], f["charge"][:m], f["rmh"][:m],
                      f["eps"][:m], hyd[:m], at[:m], ri[:m])
    return PopulationHost(np.full(n_cand, antigen_id, np.uint32), off, rs, ca[:nr], aa[:nr], cloud)


# Binders start this far outside the globule radius: at the antigen surface,
# where their atoms have in-cutoff antigen neighbours (at +8 A only ~7% of
# 20-residue binders touch a 300-residue antigen).
SHELL_GAP = 3.0


def globule_radius(n_res: int) -> float:
    return 25.0 * (n_res / 300.0) ** (1.0 / 3.0)


@dataclass
class Workload:
    name: str
    antigens: list[AtomProtein]
    pop: PopulationHost
    cutoff: float
    cell_size: float
    restraint_radius: list[float]  # per antigen
    n_steps: int
    score_every: int
    top_k: int
    with_noise: int
    description: str


def config1(n_cand: int = 64, n_res: int = 12, seed: int = 1) -> Workload:
    seq = random_sequence(16, 1000 + seed)
    ag = peptide_antigen(seq)
    ctr = ag.ca_center_of_mass()
    binders = [ring_binder(n_res, ctr, (1 << 32) | (seed << 20) | k) for k in range(n_cand)]
    return Workload("c1", [ag], population_from_proteins(binders), 10.0, 8.0, [20.0], 600, 0,
                    n_cand, 1, f"16-res peptide {seq} vs {n_cand} x {n_res}-res ring binders")


def config2(n_cand: int = 1024, n_res: int = 20, seed: int = 2) -> Workload:
    r = 25.0
    ag = globule_antigen(300, r, 2000 + seed)
    ctr = ag.ca_center_of_mass()
    pop = shell_population(n_cand, n_res, ctr, r + SHELL_GAP, r, (2 << 32) | (seed << 20))
    return Workload("c2", [ag], pop, 10.0, 10.0, [r + SHELL_GAP],
                    200, 10,
                    64, 0, f"300-res globule vs {n_cand} x {n_res}-res shell binders")


def config3(n_cand: int = 4096, seed: int = 3) -> Workload:
    ags = [globule_antigen(100 * (k + 1), globule_radius(100 * (k + 1)), 3000 + 16 * seed + k)
           for k in range(8)]
    rng = np.random.default_rng(3000 + seed)
    lens = rng.integers(8, 41, n_cand)
    binders, ids = [], []
    for c in range(n_cand):
        a = c % 8
        r = globule_radius(100 * (a + 1))
        binders.append(shell_binder(int(lens[c]), ags[a].ca_center_of_mass(), r + SHELL_GAP, r,
                                    (3 << 32) | (seed << 20) | c))
        ids.append(a)
    radii = [globule_radius(100 * (a + 1)) + SHELL_GAP for a in range(8)]
    return Workload("c3", ags, population_from_proteins(binders, ids), 12.0, 12.0, radii, 200, 10,
                    64, 0, f"8 globules 100..1000 res vs {n_cand} binders of 8..40 res")


def config4(n_cand: int = 1024, n_res: int = 30, seed: int = 4) -> Workload:
    r = globule_radius(5000)
    ag = globule_antigen(5000, r, 4000 + seed)
    ctr = ag.ca_center_of_mass()
    pop = shell_population(n_cand, n_res, ctr, r + SHELL_GAP, r, (4 << 32) | (seed << 20))
    return Workload("c4", [ag], pop, 12.0, 12.0, [r + SHELL_GAP],
                    200, 10,
                    64, 0, f"5000-res globule vs {n_cand} x {n_res}-res binders")


def device_objects(ctx, w: Workload):
    """Upload a workload: one DeviceAntigen per antigen, one DevicePopulation."""
    from .api import DeviceAntigen, DevicePopulation
    ags = [DeviceAntigen(ctx, a, cutoff=w.cutoff, cell_size=w.cell_size,
                         restraint_radius=w.restraint_radius[k])
           for k, a in enumerate(w.antigens)]
    return ags, DevicePopulation(ctx, ags, w.pop)
