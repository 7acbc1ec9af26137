"""CPU tests that pin the refinement oracle (oracle/refine_oracle.py): the
side-chain rebuild against the structure builder, the Calpha-level residue
energy against the pinned pair-term oracle, the torsion moves against the
geometry of a rotation, and the step's parts against the reference's rules."""
import math

import numpy as np

from oracle import oracle as O
from oracle import refine_oracle as R
from dnab import _native as N
from dnab.structure import AtomCloud, ResidueCloud, aa_from_sequence, protein_from_ca_trace
import ctypes as C


def rebuild(aa, rot, xyz):
    buf = np.ascontiguousarray(xyz, np.float32).reshape(-1)
    assert N.lib.dnab_rebuild_side_chain(int(aa), int(rot), N.ptr(buf, C.c_float), len(xyz)) == 0
    xyz[:] = buf.reshape(-1, 3)


def build(aa, rot, ca):
    p = protein_from_ca_trace([ca[0]], [ca[1]], [ca[2]], np.array([aa], np.uint8),
                              np.array([rot], np.uint8))
    return np.stack([p.atoms.x, p.atoms.y, p.atoms.z], 1).astype(np.float32)


def test_rebuild_side_chain_reproduces_the_builder():
    # rebuild_side_chain (src/allatom.rs:244-271) on an ideal backbone gives
    # the residue protein_from_ca_trace builds with the same rotamer
    ca = np.array([1.5, -2.25, 3.0], np.float32)
    for aa in range(20):
        for rot in range(5 if aa not in (0, 7) else 1):
            ref = build(aa, rot, ca)
            xyz = ref.copy()
            nb = 4 if aa == 7 else 5
            xyz[nb:] = 0.0
            rebuild(aa, rot, xyz)
            assert np.array_equal(xyz.view(np.uint32), ref.view(np.uint32)), (aa, rot)
    bad = np.zeros((5, 3), np.float32)
    assert N.lib.dnab_rebuild_side_chain(0, 1, N.ptr(bad, C.c_float), 5) != 0   # Ala has no rotamers
    assert N.lib.dnab_rebuild_side_chain(1, 0, N.ptr(bad, C.c_float), 5) != 0   # Arg has 11 atoms


def test_contribution_is_the_calpha_pair_terms():
    # residue_contribution_allatom (src/diffusion.rs:422-456) = the LJ +
    # Coulomb + hydrophobic + H-bond terms of interaction_energy_atoms_breakdown
    # (src/energy.rs:261-314) between the residue's Calpha atom and the
    # antigen's Calpha atoms
    ag = protein_from_ca_trace(np.arange(12, dtype=np.float32) * 3.8, np.zeros(12), np.zeros(12),
                               aa_from_sequence("DKLEGSCRWYVA"))
    ca_idx = ag.ca_atom_idx
    ag_ca_atoms = ag.atoms.subset(ca_idx)
    cap = R.ca_params()
    rng = np.random.default_rng(3)
    for aa in range(20):
        p = np.array([rng.uniform(0, 40), rng.uniform(2, 8), rng.uniform(-3, 3)], np.float32)
        one = protein_from_ca_trace([p[0]], [p[1]], [p[2]], np.array([aa], np.uint8))
        ca_atom = one.atoms.subset(one.ca_atom_idx)
        t, _ = O.energy_full_f64(ag_ca_atoms, ca_atom, 10.0)
        want = t[0] + t[1] + t[2] + t[3]
        got = R.contribution(p, ag.atoms.x[ca_idx], ag.atoms.y[ca_idx], ag.atoms.z[ca_idx],
                             ag.amino_acid, aa, cap)
        assert abs(got - want) <= 1e-5 * (1 + abs(want)), (aa, got, want)


def _chain():
    return protein_from_ca_trace(np.arange(5, dtype=np.float32) * 3.8, np.zeros(5), np.zeros(5),
                                 aa_from_sequence("LKDWS"))


def _res(p):
    rs = p.res_start
    xyz = np.stack([p.atoms.x, p.atoms.y, p.atoms.z], 1).astype(np.float32)
    return [xyz[int(rs[r]):int(rs[r + 1])].copy() for r in range(len(p.amino_acid))]


def test_torsion_moves_rotate_the_downstream_chain():
    # perturb_phi / perturb_psi (src/allatom.rs:392-456)
    for phi in (True, False):
        res = _res(_chain())
        before = [r.copy() for r in res]
        rad = R.deg2rad(np.float32(12.5))
        R.torsion(res, 2, rad, phi)
        # the axis atoms stay; residues before r stay
        keep = (0, 1) if phi else (1, 2)
        for k in keep:
            assert np.array_equal(res[2][k], before[2][k])
        for r in (0, 1):
            assert np.array_equal(res[r], before[r])
        # psi moves only O of residue r; phi every atom after CA off the axis
        # (the ideal backbone's C lies on the N-CA line)
        moved = [k for k in range(len(res[2])) if not np.array_equal(res[2][k], before[2][k])]
        a0, a1 = (before[2][0], before[2][1]) if phi else (before[2][1], before[2][2])
        ax = (a1 - a0).astype(np.float64) / np.linalg.norm((a1 - a0).astype(np.float64))

        def off_axis(pnt):
            v = pnt.astype(np.float64) - a1
            return np.linalg.norm(v - ax * (v @ ax)) > 1e-3
        want = [3] if not phi else [k for k in range(2, len(res[2])) if off_axis(before[2][k])]
        assert moved == want
        # a rigid rotation: distances among the moved atoms and to the axis kept
        pts0 = np.concatenate(before[3:]).astype(np.float64)
        pts1 = np.concatenate(res[3:]).astype(np.float64)
        d0 = np.linalg.norm(pts0[:, None] - pts0[None], axis=2)
        d1 = np.linalg.norm(pts1[:, None] - pts1[None], axis=2)
        assert np.abs(d0 - d1).max() < 1e-4
        # by the requested angle about the axis
        v0 = pts0[-1] - a1
        v1 = pts1[-1] - a1
        w0 = v0 - ax * (v0 @ ax)
        w1 = v1 - ax * (v1 @ ax)
        ang = math.atan2(np.cross(w0, w1) @ ax, w0 @ w1)
        assert abs(ang - float(rad)) < 1e-4
        # and back
        R.torsion(res, 2, np.float32(-rad), phi)
        for r in range(5):
            assert np.abs(res[r] - before[r]).max() < 1e-4


def test_uniforms_follow_philox_and_the_24_bit_rule():
    u = R.uniforms(3, 17, R.TAG_MC, 42)
    v = O.philox([3, 17, R.TAG_MC, 0], [42, 0])
    assert [float(x) for x in u] == [float((int(w) >> 8) / 16777216.0) for w in v]
    assert all(0.0 <= float(x) < 1.0 for x in u)


def _setup(seq="DKLEGSCRWYVAQNH", nb=3, seed=5):
    ag = protein_from_ca_trace(np.arange(len(seq), dtype=np.float32) * 3.8,
                               np.zeros(len(seq)), np.zeros(len(seq)), aa_from_sequence(seq))
    rng = np.random.default_rng(seed)
    ab = protein_from_ca_trace(np.arange(nb, dtype=np.float32) * 3.8 + 10.0,
                               np.full(nb, 6.5, np.float32), np.zeros(nb, np.float32),
                               rng.integers(0, 20, nb).astype(np.uint8))
    ca = ag.ca_residue_cloud()
    agd = dict(ca=ca, aa=[int(a) for a in ag.amino_acid], center=ag.ca_center_of_mass(), r0=20.0)
    state = dict(res=_res(ab), aa=[int(a) for a in ab.amino_acid], rot=[0] * nb, g0=0)
    return agd, state


def _cg(aa):
    q, e, s, h = C.c_float(), C.c_float(), C.c_float(), C.c_uint8()
    N.lib.dnab_residue_cg_params(int(aa), C.byref(q), C.byref(e), C.byref(s), C.byref(h))
    return q.value, e.value, s.value, h.value


LP = dict(step=0.08, max_disp=2.0, restraint_k=0.02, with_noise=1, seed=11)


def _rebuild_cb(aa, rot, xyz):
    rebuild(aa, rot, xyz)


def test_step_without_moves_is_the_gradient_step():
    agd, state = _setup()
    mc = dict(mutation_p=0.0, rotamer_move_p=0.12, backbone_move_p=0.0, max_torsion_deg=15.0,
              iface_cutoff_sq=64.0, moves=3)
    res0 = [r.copy() for r in state["res"]]
    R.step(state, 4, np.float32(1.0), np.float32(0.25), agd, mc, LP, build, _rebuild_cb, _cg)
    # every residue moved rigidly (the same displacement for all its atoms)
    for r0, r1 in zip(res0, state["res"]):
        d = (r1 - r0).astype(np.float64)
        assert np.abs(d - d[1]).max() < 1e-5
    assert state["aa"] == [int(a) for a in state["aa"]]


def test_metropolis_moves_happen_and_keep_residues_whole():
    agd, state = _setup(nb=6)
    mc = dict(mutation_p=0.9, rotamer_move_p=0.3, backbone_move_p=0.5, max_torsion_deg=15.0,
              iface_cutoff_sq=400.0, moves=3)
    aa0 = list(state["aa"])
    for it in range(6):
        R.step(state, it, np.float32(3.0), np.float32(0.0), agd, mc, dict(LP, with_noise=0), build,
               _rebuild_cb, _cg)
    assert state["aa"] != aa0                       # some mutation accepted
    tab = O.reference_tables()
    for r, a in zip(state["res"], state["aa"]):
        assert len(r) == len(tab["residues"][a])   # mutate_residue keeps topology counts
        # bond N-CA of an ideal or rotated backbone keeps its length
        assert abs(np.linalg.norm(r[1].astype(np.float64) - r[0]) - 1.46) < 1e-3


def _cg_table():
    return np.array([_cg(a) for a in range(20)], np.float32)


def _slots_of(res):
    s = np.zeros((len(res), 14, 3), np.float32)
    for r, xyz in enumerate(res):
        s[r, :len(xyz)] = xyz
    return s


def test_c_port_matches_the_literal_restatement():
    # the C port of the step (the CPU baseline) against refine_oracle.step:
    # same draws, same f32 geometry, same decisions
    tables = O.refine_tables(_cg_table())
    for moves, noise in ((3, 1), (1, 0), (2, 1)):
        agd, state = _setup(nb=6, seed=moves * 7 + noise)
        mc = dict(mutation_p=0.5, rotamer_move_p=0.4, backbone_move_p=0.4, max_torsion_deg=15.0,
                  iface_cutoff_sq=400.0, moves=moves)
        lp = dict(LP, with_noise=noise)
        slots = _slots_of(state["res"])
        aa = np.array(state["aa"], np.uint8)
        rot = np.array(state["rot"], np.uint8)
        steps, total = 5, 600
        temps = [O.temperature(it, total) for it in range(steps)]
        sigs = [O.sigma(it, total) if noise else 0.0 for it in range(steps)]
        for it in range(steps):
            R.step(state, it, np.float32(temps[it]), np.float32(sigs[it]), agd, mc, lp, build,
                   _rebuild_cb, _cg)
        O.refine_run(tables, agd["ca"], agd["aa"], [0, len(aa)], aa, rot, slots, temps, sigs, 0,
                     dict(step=0.08, max_disp=2.0, restraint_k=0.02, restraint_radius=20.0,
                          with_noise=noise, seed_lo=11, seed_hi=0, mut_p=0.5, rot_p=0.4,
                          bb_p=0.4, max_torsion=15.0, iface_sq=400.0, moves=moves,
                          center=agd["center"]))
        assert list(aa) == state["aa"] and list(rot) == state["rot"], moves
        for r, xyz in enumerate(state["res"]):
            assert np.array_equal(slots[r, :len(xyz)].view(np.uint32), xyz.view(np.uint32)), (moves, r)
