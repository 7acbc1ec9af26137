"""CPU restatement of allatom_diffusion_step (src/diffusion.rs:290-416): the
gradient step with residue-level forces, then the Metropolis rotamer /
mutation moves and the phi/psi torsion moves, on one candidate.

TEST INFRASTRUCTURE ONLY: imported by tests/, never by the product.

The candidate is a list of per-residue (n, 3) float32 coordinate arrays in
topology order, with residue types and rotamer indices, as the reference's
AtomProtein holds it (src/allatom.rs:92-101); a mutation replaces one entry
(mutate_residue's splice, src/allatom.rs:299-384). Geometry follows the
reference's f32 expressions operation by operation (numpy float32, IEEE per
operation); residue building is passed in (the structure builder's
protein_from_ca_trace and rebuild_side_chain). The energies of the Metropolis
tests are FP64 sums of the reference's f32 terms; random draws are
Philox4x32-10 by (iteration, residue, kind), the scheme the device uses in
place of the reference's SmallRng.
"""
from __future__ import annotations

import math

import numpy as np

from . import oracle as O

f32 = np.float32
TAG_NOISE, TAG_MC, TAG_BB = 0x6e6f6973, 0x6d636d76, 0x62626d76
GLY, CYS, SG_SLOT, N_AA = 7, 4, 5, 20


def uniforms(it, r, tag, seed):
    """Four uniforms in [0, 1) from the top 24 bits of a Philox4x32-10 block."""
    v = O.philox([it, r, tag, 0], [seed & 0xFFFFFFFF, seed >> 32])
    return [f32(int(x) >> 8) * f32(1.0 / 16777216.0) for x in v]


def deg2rad(d):
    return f32(d) * f32(f32(math.pi) / f32(180.0))


def ca_params():
    """Per residue type, the Calpha atom's (r_min_half, epsilon, charge,
    hydrophobic, atom type) from the tables (topology slot 1)."""
    tab = O.reference_tables()
    out = []
    for atoms in tab["residues"]:
        t, q, h = atoms[1]
        rmh, eps = tab["amber_nb"][t]
        out.append((f32(rmh), f32(eps), f32(q), int(h), int(t)))
    return out


def contribution(ca, ag_x, ag_y, ag_z, ag_aa, aa, cap):
    """residue_contribution_allatom (src/diffusion.rs:422-456): the Calpha
    of a residue of type aa at ca against every antigen Calpha."""
    brm, be, bq, bh, bt = cap[aa]
    tab = O.reference_tables()
    don, acc = tab["hbond_donor"], tab["hbond_acceptor"]
    e = 0.0
    for i in range(len(ag_x)):
        dx, dy, dz = f32(ca[0] - ag_x[i]), f32(ca[1] - ag_y[i]), f32(ca[2] - ag_z[i])
        r_sq = f32(f32(dx * dx + dy * dy) + dz * dz)
        if r_sq > f32(100.0) or r_sq < f32(0.25):
            continue
        arm, ae, aq, ah, at = cap[ag_aa[i]]
        r_ij = f32(brm + arm)
        eps = f32(np.sqrt(f32(be * ae)))
        r2 = f32(f32(r_ij * r_ij) / r_sq)
        r6 = f32(f32(r2 * r2) * r2)
        e += float(f32(eps * f32(f32(r6 * r6) - f32(f32(2.0) * r6))))
        if bq != 0.0 and aq != 0.0:
            e += float(f32(f32(f32(f32(332.0) * bq) * aq) / f32(np.sqrt(r_sq))))
        if bh == 1 and ah == 1 and r_sq < f32(36.0):
            e -= 0.5
        if (don[bt] and acc[at]) or (acc[bt] and don[at]):
            e += O.hbond_energy_f64(float(r_sq))
    return e


def disulfide(res, aa):
    sg = [r[SG_SLOT] for r, a in zip(res, aa) if a == CYS and len(r) > SG_SLOT]
    if not sg:
        return 0.0
    p = np.array(sg, np.float32)
    return O.disulfide_f64(p[:, 0], p[:, 1], p[:, 2])


def normalize(v):
    l = f32(np.sqrt(f32(f32(v[0] * v[0] + v[1] * v[1]) + v[2] * v[2])))
    if l < f32(1e-8):
        l = f32(1e-8)
    return np.array([v[0] / l, v[1] / l, v[2] / l], np.float32)


def rodrigues(p, origin, axis, ca, sa):
    """rodrigues (src/allatom.rs:467-479), f32 operation by operation."""
    v = (p - origin).astype(np.float32)
    dot = f32(f32(v[0] * axis[0] + v[1] * axis[1]) + v[2] * axis[2])
    cx = np.array([f32(axis[1] * v[2] - axis[2] * v[1]), f32(axis[2] * v[0] - axis[0] * v[2]),
                   f32(axis[0] * v[1] - axis[1] * v[0])], np.float32)
    omc = f32(f32(1.0) - ca)
    r = np.array([f32(f32(v[k] * ca + cx[k] * sa) + f32(f32(axis[k] * dot) * omc))
                  for k in range(3)], np.float32)
    return (r + origin).astype(np.float32)


def torsion(res, r, rad, phi):
    """perturb_phi / perturb_psi (src/allatom.rs:392-456)."""
    if phi:
        p0, p1 = res[r][0].copy(), res[r][1].copy()
    else:
        p0, p1 = res[r][1].copy(), res[r][2].copy()
    axis = normalize((p1 - p0).astype(np.float32))
    cs, sn = f32(math.cos(float(rad))), f32(math.sin(float(rad)))
    rr = res[r]
    for k in (range(2, len(rr)) if phi else [3]):
        rr[k] = rodrigues(rr[k], p1, axis, cs, sn)
    for j in range(r + 1, len(res)):
        for k in range(len(res[j])):
            res[j][k] = rodrigues(res[j][k], p1, axis, cs, sn)


def step(state, it, temp, sigma, ag, mc, lp, build_residue, rebuild_side_chain, cg_of):
    """One allatom_diffusion_step (src/diffusion.rs:290-416) on state, a dict
    of res (per-residue float32 (n, 3) arrays), aa, rot (lists) and g0 (the
    candidate's first global residue index, which keys its random draws).
    ag: dict of ca (antigen Calpha ResidueCloud), aa (their types), center,
    r0. lp: dict of step, max_disp, restraint_k, with_noise, seed. cg_of(aa):
    residue-level (charge, epsilon, sigma, hydrophobic)."""
    res, aa, rot, g0 = state["res"], state["aa"], state["rot"], state["g0"]
    n = len(res)
    seed = lp["seed"]
    # 1. forces on the Calphas, restraint, clamp, noise, set_ca_pos
    xyz = np.concatenate(res).astype(np.float32)
    off = np.concatenate([[0], np.cumsum([len(r) for r in res])]).astype(np.uint32)
    x, y, z = (np.ascontiguousarray(xyz[:, k]) for k in range(3))
    cg = [cg_of(a) for a in aa]
    res_cg = (np.array([c[0] for c in cg], np.float32), np.array([c[1] for c in cg], np.float32),
              np.array([c[2] for c in cg], np.float32), np.array([c[3] for c in cg], np.uint8))
    O.langevin_step(x, y, z, off, off[:-1] + 1,
                    dict(step=lp["step"], max_disp=lp["max_disp"], restraint_k=lp["restraint_k"],
                         restraint_radius=ag["r0"], sigma=sigma, with_noise=lp["with_noise"],
                         force_model=1, iter=it, seed_lo=seed & 0xFFFFFFFF, seed_hi=seed >> 32,
                         center=ag["center"]),
                    ag_ca=ag["ca"], res_cg=res_cg, res_id0=g0)
    for r in range(n):
        b, e = int(off[r]), int(off[r + 1])
        res[r] = np.stack([x[b:e], y[b:e], z[b:e]], 1).astype(np.float32)
    cap = ca_params()
    agx, agy, agz = ag["ca"].x, ag["ca"].y, ag["ca"].z
    # 2. interface mask
    iface = []
    for r in range(n):
        c = res[r][1]
        dx, dy, dz = (c[0] - agx).astype(np.float32), (c[1] - agy).astype(np.float32), (c[2] - agz).astype(np.float32)
        d2 = ((dx * dx + dy * dy).astype(np.float32) + dz * dz).astype(np.float32)
        iface.append(bool((d2 < f32(mc["iface_cutoff_sq"])).any()))

    def energy(r):
        return contribution(res[r][1], agx, agy, agz, ag["aa"], aa[r], cap)

    def accept(delta, u):
        return delta <= 0.0 or float(u) < math.exp(-delta / float(temp))

    def mutate(r, new_aa, new_rot):
        if aa[r] == new_aa:
            rot[r] = new_rot
            rebuild_side_chain(new_aa, new_rot, res[r])
            return
        res[r] = build_residue(new_aa, new_rot, res[r][1].copy())
        aa[r], rot[r] = new_aa, new_rot

    # 3. rotamer moves and mutations
    if mc["moves"] & 1:
        for r in range(n):
            u = uniforms(it, g0 + r, TAG_MC, seed)
            eff_mut = f32(mc["mutation_p"]) * f32(3.0 if iface[r] else 0.4)
            eff_rot = f32(mc["rotamer_move_p"]) * f32(2.0) if iface[r] else f32(mc["rotamer_move_p"])
            if not u[0] < eff_mut:
                continue
            old_aa, old_rot = aa[r], rot[r]
            if u[1] < eff_rot:
                nrot = 0 if old_aa in (0, GLY) else 5
                if nrot == 0:
                    continue
                new_rot = min(int(u[2] * f32(nrot)), nrot - 1)
                old_e = energy(r) + disulfide(res, aa)
                rot[r] = new_rot
                rebuild_side_chain(old_aa, new_rot, res[r])
                new_e = energy(r) + disulfide(res, aa)
                if not accept(new_e - old_e, u[3]):
                    rot[r] = old_rot
                    rebuild_side_chain(old_aa, old_rot, res[r])
            else:
                new_aa = min(int(u[2] * f32(N_AA)), N_AA - 1)
                old_e = energy(r) + disulfide(res, aa)
                mutate(r, new_aa, 0)
                new_e = energy(r) + disulfide(res, aa)
                if not accept(new_e - old_e, u[3]):
                    mutate(r, old_aa, old_rot)
    # 4. torsion moves
    if mc["moves"] & 2:
        for r in range(n):
            u = uniforms(it, g0 + r, TAG_BB, seed)
            if not u[0] < f32(mc["backbone_move_p"]):
                continue
            mt = f32(mc["max_torsion_deg"])
            rad = deg2rad(f32(-mt + f32(f32(mt + mt) * u[1])))
            phi = bool(u[2] < f32(0.5))
            old_e = sum(energy(i) for i in range(r, n)) + disulfide(res, aa)
            torsion(res, r, rad, phi)
            new_e = sum(energy(i) for i in range(r, n)) + disulfide(res, aa)
            if not accept(new_e - old_e, u[3]):
                torsion(res, r, f32(-rad), phi)
