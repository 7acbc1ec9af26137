"""All-atom refinement with Monte-Carlo moves (dnab_refine) on the GPU vs the
oracle's restatement of allatom_diffusion_step (oracle/refine_oracle.py) fed
the same Philox draws.

Bars, as written below: amino-acid sequences and rotamer indices exact (every
Metropolis decision the same); coordinates within 1e-4 A (the geometry is the
same f32 operations on both sides; the only admitted difference is a last-bit
one in a torsion move's cosine/sine, taken from the FP64 libm of each side);
allatom_energy of the refined candidates |x - x64| <= 1e-3 + 1e-4 |x64|.
"""
import ctypes as C

import numpy as np
import pytest

from oracle import oracle as O
from oracle import refine_oracle as R
from dnab import DeviceAntigen, _native as N
from dnab import workloads as W
from dnab.api import langevin_params, mc_params
from dnab.structure import protein_from_ca_trace

pytestmark = pytest.mark.gpu


def tol(x, x64):
    return np.abs(np.asarray(x, np.float64) - x64) <= 1e-3 + 1e-4 * np.abs(x64)


def _cg(aa):
    q, e, s, h = C.c_float(), C.c_float(), C.c_float(), C.c_uint8()
    N.lib.dnab_residue_cg_params(int(aa), C.byref(q), C.byref(e), C.byref(s), C.byref(h))
    return q.value, e.value, s.value, h.value


def _build(aa, rot, ca):
    p = protein_from_ca_trace([ca[0]], [ca[1]], [ca[2]], np.array([aa], np.uint8),
                              np.array([rot], np.uint8))
    return np.stack([p.atoms.x, p.atoms.y, p.atoms.z], 1).astype(np.float32)


def _rebuild(aa, rot, xyz):
    buf = np.ascontiguousarray(xyz, np.float32).reshape(-1)
    assert N.lib.dnab_rebuild_side_chain(int(aa), int(rot), N.ptr(buf, C.c_float), len(xyz)) == 0
    xyz[:] = buf.reshape(-1, 3)


def _workload(n_cand=4, n_res=10, seed=3, n_ag=60, inside=False):
    ag = W.globule_antigen(n_ag, W.globule_radius(n_ag), 900 + seed)
    ctr = ag.ca_center_of_mass()
    r = W.globule_radius(n_ag)
    rng = np.random.default_rng(seed)
    abs_, rots = [], []
    for c in range(n_cand):
        shell, rmin = (0.5 * r, 0.0) if inside else (r + 6.0, r)
        sb = W.shell_binder(n_res, ctr, shell, rmin, 70 + seed * 100 + c)
        ca = sb.ca_pos()
        aa = sb.amino_acid
        rot = np.where((aa == 0) | (aa == 7), 0, rng.integers(0, 5, n_res)).astype(np.uint8)
        abs_.append(protein_from_ca_trace(ca[:, 0], ca[:, 1], ca[:, 2], aa, rot))
        rots.append(rot)
    return ag, abs_, rots


def _oracle_run(ag, ab, rot, g0, lp, mc):
    rs = ab.res_start
    xyz = np.stack([ab.atoms.x, ab.atoms.y, ab.atoms.z], 1).astype(np.float32)
    state = dict(res=[xyz[int(rs[r]):int(rs[r + 1])].copy() for r in range(len(ab.amino_acid))],
                 aa=[int(a) for a in ab.amino_acid], rot=[int(k) for k in rot], g0=g0)
    agd = dict(ca=ag.ca_residue_cloud(), aa=[int(a) for a in ag.amino_acid],
               center=ag.ca_center_of_mass(), r0=float(lp.restraint_radius))
    lpd = dict(step=lp.step, max_disp=lp.max_disp, restraint_k=lp.restraint_k,
               with_noise=lp.with_noise, seed=lp.seed)
    mcd = dict(mutation_p=mc.mutation_p, rotamer_move_p=mc.rotamer_move_p,
               backbone_move_p=mc.backbone_move_p, max_torsion_deg=mc.max_torsion_deg,
               iface_cutoff_sq=mc.iface_cutoff_sq, moves=mc.moves)
    for s in range(lp.n_steps):
        it = lp.schedule_offset + s
        temp = np.float32(O.temperature(it, lp.schedule_total, lp.t_hot, lp.t_cold))
        sig = np.float32(O.sigma(it, lp.schedule_total, lp.t_hot, lp.t_cold, lp.noise_base)
                         if lp.with_noise else 0.0)
        R.step(state, it, temp, sig, agd, mcd, lpd, _build, _rebuild, _cg)
    return state


@pytest.mark.parametrize("noise,moves,waves", [(1, 3, 0), (0, 3, 1), (1, 1, 4), (1, 2, 1),
                                               (1, 3, 1), (1, 3, 4), (1, 3, 2)],
                         ids=["noise-all", "quiet-all-1wave", "noise-sidechain-4waves",
                              "noise-torsion-1wave", "noise-all-1wave", "noise-all-4waves",
                              "noise-all-2waves"])
def test_refine_matches_the_oracle(ctx, noise, moves, waves):
    """waves: per candidate, 0 = the library's choice (8 for this batch)."""
    ag, abs_, rots = _workload(seed=3 + moves)
    agd = DeviceAntigen(ctx, ag, cutoff=10.0)
    lp = langevin_params(n_steps=6, with_noise=noise, schedule_offset=100, schedule_total=800,
                         seed=0x1234_5678_9ABC)
    mc = mc_params(mutation_p=0.3, rotamer_move_p=0.35, backbone_move_p=0.25, moves=moves,
                   waves=waves)
    prots, rot_out, e = ctx.refine(agd, abs_, lp, mc, rots)
    g0 = 0
    n_exact = n_atoms = n_mut = 0
    for c, (ab, rot) in enumerate(zip(abs_, rots)):
        st = _oracle_run(ag, ab, rot, g0, lp, mc)
        g0 += len(ab.amino_acid)
        p = prots[c]
        assert list(p.amino_acid) == st["aa"], (c, list(p.amino_acid), st["aa"])
        assert list(rot_out[c]) == st["rot"], (c, list(rot_out[c]), st["rot"])
        want = np.concatenate(st["res"]).astype(np.float32)
        got = np.stack([p.atoms.x, p.atoms.y, p.atoms.z], 1)
        assert got.shape == want.shape
        d = np.abs(got.astype(np.float64) - want)
        assert d.max() <= 1e-4, (c, d.max())
        n_exact += int((got == want).all(axis=1).sum())
        n_atoms += len(got)
        n_mut += sum(int(a != b) for a, b in zip(st["aa"], ab.amino_acid))
        res = [np.ascontiguousarray(r) for r in st["res"]]
        cloud = O.concat_residues([_cloud(r, a) for r, a in zip(res, st["aa"])])
        t64, _ = O.energy_full_f64(ag.atoms, cloud, 10.0)
        sg = [r[R.SG_SLOT] for r, a in zip(res, st["aa"]) if a == R.CYS and len(r) > R.SG_SLOT]
        ds = O.disulfide_f64(*(np.array([q[k] for q in sg], np.float32) for k in range(3)))
        assert tol(e[c], t64.sum() + ds), (c, e[c], t64.sum() + ds)
    print(f"atoms bit-identical {n_exact}/{n_atoms}, mutations kept {n_mut}")
    if moves & 1:
        assert n_mut > 0


@pytest.mark.parametrize("waves", [1, 8])
def test_refine_long_binder_matches_the_oracle(ctx, waves):
    """A 150-residue binder threaded through a 300-residue antigen: more than
    64 residues per wave (chunked force pass and interface mask at one wave)
    and 27-cell pair lists longer than a wave's list (batches of residues)."""
    ag, abs_, rots = _workload(n_cand=1, n_res=150, seed=11, n_ag=300, inside=True)
    agd = DeviceAntigen(ctx, ag, cutoff=10.0)
    lp = langevin_params(n_steps=3, with_noise=1, schedule_offset=20, schedule_total=800, seed=99)
    mc = mc_params(mutation_p=0.2, rotamer_move_p=0.3, backbone_move_p=0.01, waves=waves)
    prots, rot_out, _ = ctx.refine(agd, abs_, lp, mc, rots)
    st = _oracle_run(ag, abs_[0], rots[0], 0, lp, mc)
    assert list(prots[0].amino_acid) == st["aa"] and list(rot_out[0]) == st["rot"]
    want = np.concatenate(st["res"]).astype(np.float64)
    got = np.stack([prots[0].atoms.x, prots[0].atoms.y, prots[0].atoms.z], 1).astype(np.float64)
    assert got.shape == want.shape and np.abs(got - want).max() <= 1e-4


def test_refine_literal_moves_match_the_oracle(ctx):
    """A non-finite input coordinate sends every move down the literal path
    (rebuild, score, undo) instead of deciding on the residue types first; a
    NaN on a side-chain atom of a non-Cys residue leaves every energy finite,
    so both paths must end where the oracle does."""
    ag, abs_, rots = _workload(n_cand=3, seed=6)
    ab = abs_[1]
    r = next(i for i, a in enumerate(ab.amino_acid) if a not in (0, 4, 7))
    ab.atoms.x[int(ab.res_start[r]) + 5] = np.nan
    agd = DeviceAntigen(ctx, ag, cutoff=10.0)
    lp = langevin_params(n_steps=6, with_noise=1, schedule_offset=50, schedule_total=800, seed=4242)
    mc = mc_params(mutation_p=0.35, rotamer_move_p=0.4, backbone_move_p=0.25)
    prots, rot_out, _ = ctx.refine(agd, abs_, lp, mc, rots)
    g0 = 0
    for c, (ab, rot) in enumerate(zip(abs_, rots)):
        st = _oracle_run(ag, ab, rot, g0, lp, mc)
        g0 += len(ab.amino_acid)
        p = prots[c]
        assert list(p.amino_acid) == st["aa"] and list(rot_out[c]) == st["rot"], c
        want = np.concatenate(st["res"]).astype(np.float64)
        got = np.stack([p.atoms.x, p.atoms.y, p.atoms.z], 1).astype(np.float64)
        assert np.array_equal(np.isnan(got), np.isnan(want)), c
        ok = ~np.isnan(want)
        assert np.abs(got[ok] - want[ok]).max() <= 1e-4, c


def _cloud(xyz, aa):
    tab = O.reference_tables()
    from types import SimpleNamespace
    t = np.array([a[0] for a in tab["residues"][aa]], np.uint8)
    nb = np.array(tab["amber_nb"], np.float32)
    return SimpleNamespace(x=xyz[:, 0].copy(), y=xyz[:, 1].copy(), z=xyz[:, 2].copy(),
                           charge=np.array([a[1] for a in tab["residues"][aa]], np.float32),
                           r_min_half=nb[t, 0].copy(), epsilon=nb[t, 1].copy(),
                           hydrophobic=np.array([a[2] for a in tab["residues"][aa]], np.uint8),
                           atom_type=t)


def test_refine_repeats_and_seeds_differ(ctx):
    ag, abs_, rots = _workload(n_cand=3, seed=9)
    agd = DeviceAntigen(ctx, ag, cutoff=10.0)
    lp = langevin_params(n_steps=5, with_noise=1, schedule_total=600, seed=77)
    mc = mc_params(mutation_p=0.3, backbone_move_p=0.2)
    a = ctx.refine(agd, abs_, lp, mc, rots)
    b = ctx.refine(agd, abs_, lp, mc, rots)
    for p, q in zip(a[0], b[0]):
        assert np.array_equal(p.atoms.x, q.atoms.x) and list(p.amino_acid) == list(q.amino_acid)
    assert np.array_equal(a[2], b[2])
    # one wave or four per candidate: the same moves and coordinates
    for w in (1, 2, 4, 8):
        d = ctx.refine(agd, abs_, lp, mc_params(mutation_p=0.3, backbone_move_p=0.2, waves=w), rots)
        for p, q in zip(a[0], d[0]):
            assert np.array_equal(p.atoms.x, q.atoms.x) and np.array_equal(p.atoms.z, q.atoms.z)
            assert list(p.amino_acid) == list(q.amino_acid)
    lp2 = langevin_params(n_steps=5, with_noise=1, schedule_total=600, seed=78)
    c = ctx.refine(agd, abs_, lp2, mc, rots)
    assert any(not np.array_equal(p.atoms.x, q.atoms.x) for p, q in zip(a[0], c[0]))


def test_refine_rejects_bad_inputs(ctx):
    ag, abs_, rots = _workload(n_cand=1, seed=2)
    agd = DeviceAntigen(ctx, ag, cutoff=10.0)
    with pytest.raises(Exception):          # antigen without its Calpha cloud
        ctx.refine(DeviceAntigen(ctx, ag, cutoff=10.0, with_ca=False), abs_)
    with pytest.raises(Exception):          # rotamer index out of range
        ctx.refine(agd, abs_, rotamers=[np.full(len(abs_[0].amino_acid), 7, np.uint8)])
    with pytest.raises(Exception):          # waves per candidate other than 0, 1, 2, 4, 8
        ctx.refine(agd, abs_, langevin_params(n_steps=2), mc_params(waves=3))
    prots, _, e = ctx.refine(agd, [], langevin_params(n_steps=2))
    assert prots == [] and len(e) == 0
