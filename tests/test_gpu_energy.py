"""GPU all-atom energies and forces vs the oracle.

Tolerance (written here, applied per candidate energy and per atom-force
component): |x - x64| <= 1e-3 + 1e-4 |x64| (kcal/mol, kcal/mol/A), x64 the
oracle's FP64 value on the same f32 inputs and the same f32 pair predicate.
The FP64 GPU kernel must agree with the CPU FP64 oracle to 1e-9 relative.
Pair counts and hydrophobic terms are integer-exact.
"""
import numpy as np
import pytest

from oracle import oracle as O
from dnab import DeviceAntigen, DevicePopulation
from dnab import workloads as W
from dnab.structure import AtomCloud, population_from_proteins, protein_from_ca_trace, \
    aa_from_sequence

pytestmark = pytest.mark.gpu


def tol(x, x64):
    return np.abs(np.asarray(x, np.float64) - x64) <= 1e-3 + 1e-4 * np.abs(x64)


def cloud(xyz, q, rmh, eps, hyd):
    xyz = np.asarray(xyz, np.float32).reshape(-1, 3)
    n = len(xyz)
    f = lambda v: np.broadcast_to(np.asarray(v, np.float32), (n,)).copy()
    return AtomCloud(xyz[:, 0], xyz[:, 1], xyz[:, 2], f(q), f(rmh), f(eps),
                     np.broadcast_to(np.asarray(hyd, np.uint8), (n,)).copy())


def test_two_atom_closed_forms(ctx):
    R = float(np.float32(1.908) + np.float32(1.824))
    ag = cloud([0, 0, 0], 0.8, 1.908, 0.1094, 1)
    ab = cloud([R, 0, 0], -0.5, 1.824, 0.17, 1)
    b = ctx.interaction_energy_atoms_breakdown(ag, ab)
    eps = np.sqrt(np.float64(np.float32(0.1094)) * np.float64(np.float32(0.17)))
    assert abs(b.lj - (-eps)) < 1e-6
    assert abs(b.coulomb - 332.0 * float(np.float32(0.8)) * float(np.float32(-0.5)) / R) < 1e-5
    assert b.hydrophobic == -0.5 and b.hbond == 0.0 and b.solvation == 0.0
    fx, fy, fz, e = ctx.atom_forces(ag, ab)
    assert fx[0] < 0          # attraction dominates at the LJ minimum
    # cutoff edge: r^2 == 100 is in, just beyond is out; r^2 < 0.25 is out
    ab10 = cloud([10.0, 0, 0], -0.5, 1.824, 0.17, 0)
    assert ctx.interaction_energy_atoms_breakdown(ag, ab10).coulomb != 0.0
    ab_out = cloud([float(np.nextafter(np.float32(10), np.float32(11))), 0, 0], -0.5, 1.824, 0.17, 0)
    assert ctx.interaction_energy_atoms_breakdown(ag, ab_out).coulomb == 0.0
    ab_in = cloud([0.49, 0, 0], -0.5, 1.824, 0.17, 0)
    assert ctx.interaction_energy_atoms(ag, ab_in) == 0.0


def test_close_pairs_in_the_pair_loop(ctx):
    # The pair loop skips a staged antigen atom by the cutoff comparison alone
    # and drops pairs with r^2 < 0.25 where it evaluates them. Binder atoms:
    # 0.49 A from antigen atom 0 (no pair) beside atoms that pair with it, one
    # 0.5 A away (r^2 = 0.25: a pair), and one whose only antigen atom within
    # the cutoff is 0.49 A away (antigen atom 1 is evaluated for it alone).
    ag = cloud([[0, 0, 0], [30.0, 0, 0]], [0.8, -0.3], 1.908, 0.1094, [1, 0])
    ab = cloud([[0.49, 0, 0], [0.5, 0, 0], [3.0, 0, 0], [30.49, 0, 0]], -0.5, 1.824, 0.17, 1)
    fx, fy, fz, e = ctx.atom_forces(ag, ab)
    (lj, co, hy), f64, npairs = O.energy_f64(ag, ab)
    assert npairs == 2
    assert fx[0] == 0.0 and fy[0] == 0.0 and fz[0] == 0.0
    assert fx[3] == 0.0 and fy[3] == 0.0 and fz[3] == 0.0
    assert fx[1] != 0.0 and fx[2] != 0.0
    assert tol(e, lj + co + hy)
    for got, ref in zip((fx, fy, fz), f64):
        assert tol(got, ref).all(), (got, ref)
    # the close atoms add exact zeros: the same bits without them
    ab2 = cloud([[0.5, 0, 0], [3.0, 0, 0]], -0.5, 1.824, 0.17, 1)
    gx, gy, gz, e2 = ctx.atom_forces(ag, ab2)
    assert np.array_equal(np.array([fx[1], fx[2]]).view(np.uint32), gx.view(np.uint32))
    assert np.array_equal(np.array([fy[1], fy[2]]).view(np.uint32), gy.view(np.uint32))
    assert np.float64(e).view(np.uint64) == np.float64(e2).view(np.uint64)


def test_atom_breakdown_sums_to_total(ctx):
    # src/energy.rs:496-507
    ag = protein_from_ca_trace([0.0], [0.0], [0.0], aa_from_sequence("D"))
    ab = protein_from_ca_trace([3.0], [0.0], [0.0], aa_from_sequence("K"))
    b = ctx.interaction_energy_atoms_breakdown(ag.atoms, ab.atoms)
    total = ctx.interaction_energy_atoms(ag.atoms, ab.atoms)
    assert abs(b.total() - total) < 1e-4
    t64, _ = O.energy_full_f64(ag.atoms, ab.atoms)
    assert tol(total, t64.sum())
    got = [b.lj, b.coulomb, b.hydrophobic, b.hbond, b.solvation]
    assert tol(got, t64).all(), (got, t64)
    assert b.solvation != 0.0          # typed atoms: EEF1 is live


def typed_cloud(xyz, types, q=0.0, rmh=1.9, eps=0.1, hyd=0):
    c = cloud(xyz, q, rmh, eps, hyd)
    c.atom_type = np.ascontiguousarray(types, np.uint8)
    return c


def test_hbond_and_solvation_two_atoms(ctx):
    # backbone N (AtomType 5, donor) and O (10, acceptor) at the well's centre
    # (src/energy.rs:477-491): -1.5 kcal/mol; EEF1 (src/solvation.rs:52-84)
    # from the tables of the two types
    tab = O.reference_tables()
    eef = np.asarray(tab["eef1"], np.float32).reshape(-1, 3)
    r = np.float32(2.9)
    r2 = float(r * r)
    ag, ab = typed_cloud([0, 0, 0], [5]), typed_cloud([r, 0, 0], [10])
    b = ctx.interaction_energy_atoms_breakdown(ag, ab)
    assert abs(b.hbond - O.hbond_energy_f64(r2)) < 1e-6 and abs(b.hbond + 1.5) < 1e-5
    sv64 = O.pair_solvation_f64(eef[5], eef[10], r2)
    assert abs(b.solvation - sv64) < 1e-6 * (1 + abs(sv64)) and sv64 != 0.0
    # donor-donor (types 5 and 6): no H-bond, solvation stays (type 5 has zero
    # volume, so only type 6's volume enters)
    b2 = ctx.interaction_energy_atoms_breakdown(ag, typed_cloud([r, 0, 0], [6]))
    assert b2.hbond == 0.0 and b2.solvation != 0.0
    sv2 = O.pair_solvation_f64(eef[5], eef[6], r2)
    assert abs(b2.solvation - sv2) < 1e-6 * (1 + abs(sv2))
    # acceptor on the antigen side, donor on the binder: the term is symmetric
    b3 = ctx.interaction_energy_atoms_breakdown(typed_cloud([0, 0, 0], [10]), ab.subset([0]))
    assert b3.hbond == 0.0
    b4 = ctx.interaction_energy_atoms_breakdown(typed_cloud([0, 0, 0], [10]),
                                                typed_cloud([r, 0, 0], [5]))
    assert abs(b4.hbond - b.hbond) < 1e-7
    # the H-bond cutoff is r^2 <= 25, solvation's r^2 < 81; both need both types
    far = np.float32(5.0)
    b5 = ctx.interaction_energy_atoms_breakdown(ag, typed_cloud([far, 0, 0], [10]))
    assert b5.hbond != 0.0
    b6 = ctx.interaction_energy_atoms_breakdown(
        ag, typed_cloud([np.nextafter(far, np.float32(6)), 0, 0], [10]))
    assert b6.hbond == 0.0
    b7 = ctx.interaction_energy_atoms_breakdown(ag, typed_cloud([9.0, 0, 0], [10]))
    assert b7.solvation == 0.0
    nine = np.nextafter(np.float32(9), np.float32(0))
    b9 = ctx.interaction_energy_atoms_breakdown(ag, typed_cloud([nine, 0, 0], [10]))
    assert b9.solvation != 0.0
    untyped = cloud([r, 0, 0], 0.0, 1.9, 0.1, 0)
    b8 = ctx.interaction_energy_atoms_breakdown(ag, untyped)
    assert b8.hbond == 0.0 and b8.solvation == 0.0


@pytest.mark.parametrize("make", [lambda: W.config1(), lambda: W.config2(n_cand=96),
                                  lambda: W.config4(n_cand=12)],
                         ids=["c1", "c2-96", "c4-12"])
def test_all_terms_parity(ctx, make):
    # allatom_energy (src/diffusion.rs:281-283) per candidate: the five pair
    # terms vs the FP64 oracle and the binder's disulfide term vs its oracle
    w = make()
    ags, pop = W.device_objects(ctx, w)
    pop.evaluate(forces=False, breakdown=True)
    t = pop.terms()
    e, _ = pop.energies()
    p = w.pop
    off = p.cand_atom_off()
    ag = w.antigens[0].atoms
    n_hb = 0
    tot64 = np.zeros(p.n_cand)
    for c in range(p.n_cand):
        t64, _ = O.energy_full_f64(ag, p.candidate_cloud(c), w.cutoff)
        tot64[c] = t64.sum()
        ok = tol(t[c, :5], t64)
        assert ok.all(), (c, t[c, :5], t64)
        r0, r1 = int(p.cand_res_off[c]), int(p.cand_res_off[c + 1])
        sg = O.sg_positions(p.atoms.x, p.atoms.y, p.atoms.z, p.res_atom_off,
                            p.res_amino_acid, r0, r1)
        ds = O.disulfide_f64(*sg)
        assert abs(t[c, 5] - ds) <= 1e-5 * (1 + abs(ds)), (c, t[c, 5], ds)
        n_hb += t64[3] != 0.0
    assert tol(e, tot64).all()           # the candidate energy is the five terms' total
    assert n_hb > 0                      # the workload exercises the H-bond term
    # a forces evaluation keeps LJ + Coulomb + hydrophobic, the terms its forces derive from
    pop.evaluate(forces=True)
    e3, _ = pop.energies()
    r = O.population_f64(ag, p.atoms, off, w.cutoff, forces=False)
    assert tol(e3, r["energy"]).all()


def test_allatom_energy_entry_point(ctx):
    # allatom_energy (src/diffusion.rs:281-283) through its own entry point
    # equals the population's six terms for the same candidate
    from dnab.structure import AtomProtein
    w = W.config2(n_cand=8)
    ags, pop = W.device_objects(ctx, w)
    pop.evaluate(forces=False, breakdown=True)
    t = pop.terms()
    p = w.pop
    off = p.cand_atom_off()
    ag = w.antigens[0]
    for c in range(p.n_cand):
        r0, r1 = int(p.cand_res_off[c]), int(p.cand_res_off[c + 1])
        prot = AtomProtein(p.candidate_cloud(c), (p.res_atom_off[r0:r1 + 1] - off[c]).astype(np.uint32),
                           p.res_amino_acid[r0:r1].copy(),
                           (p.res_ca_atom[r0:r1] - off[c]).astype(np.uint32))
        e, ds = ctx.allatom_energy(ag, prot, w.cutoff, with_disulfide=True)
        assert ds == t[c, 5]
        assert abs(e - t[c].sum()) <= 1e-9 * (1 + np.abs(t[c]).sum()), (c, e, t[c])
        inter = ctx.interaction_energy_atoms(ag.atoms, prot.atoms, w.cutoff)
        assert abs(e - (inter + ds)) <= 1e-9 * (1 + abs(e))


def test_disulfide_closed_form(ctx):
    # two Cys whose SG atoms sit 2.05 A apart: the well's bottom, -4 kcal/mol
    # (src/allatom.rs:155-180); beyond 4 A nothing; a third SG adds its pairs
    prot = protein_from_ca_trace(np.arange(4, dtype=np.float32) * 3.8, np.zeros(4), np.zeros(4),
                                 aa_from_sequence("ACCA"))
    host = population_from_proteins([prot, prot])
    ags = [DeviceAntigen(ctx, W.globule_antigen(20, 8.0, 0), cutoff=10.0)]
    pop = DevicePopulation(ctx, ags, host)
    x, y, z = (v.copy() for v in pop.coords())
    sg = [int(host.res_atom_off[r]) + O.CYS_SG_SLOT for r in (1, 2)]
    base = int(host.res_atom_off[4])
    d0 = np.float32(2.05)
    for k, d in ((0, d0), (1, np.float32(4.5))):
        a, b = sg[0] + k * base, sg[1] + k * base
        y[b], z[b] = y[a], z[a]
        x[b] = x[a] + d
    pop.set_coords(x, y, z)
    pop.evaluate(forces=False, breakdown=True)
    t = pop.terms()
    assert abs(t[0, 5] + 4.0) < 1e-5, t[:, 5]
    assert t[1, 5] == 0.0
    for c in range(2):
        xs = x[[sg[0] + c * base, sg[1] + c * base]]
        ys = y[[sg[0] + c * base, sg[1] + c * base]]
        zs = z[[sg[0] + c * base, sg[1] + c * base]]
        assert abs(t[c, 5] - O.disulfide_f64(xs, ys, zs)) < 1e-6


def eval_vs_oracle(ctx, w, n_check=None):
    ags, pop = W.device_objects(ctx, w)
    pop.evaluate(forces=True)
    e, err = pop.energies()
    fx, fy, fz = pop.forces()
    off = w.pop.cand_atom_off()
    assert len(w.antigens) == 1
    r = O.population_f64(w.antigens[0].atoms, w.pop.atoms, off, w.cutoff)
    return ags, pop, e, err, (fx, fy, fz), r


@pytest.mark.parametrize("make", [lambda: W.config1(), lambda: W.config2(n_cand=96)],
                         ids=["c1", "c2-96"])
def test_energy_force_parity(ctx, make):
    w = make()
    ags, pop, e, err, f, r = eval_vs_oracle(ctx, w)
    ok = tol(e, r["energy"])
    assert ok.all(), (np.nonzero(~ok)[0][:5], e[~ok][:5], r["energy"][~ok][:5])
    for a, b in zip(f, r["forces"]):
        okf = tol(a, b)
        assert okf.all(), (np.nonzero(~okf)[0][:5], a[~okf][:5], b[~okf][:5])
    n_in, n_tested = pop.pair_count()
    assert n_in == int(r["pairs"].sum())
    assert n_tested >= n_in
    print(f"pairs in cutoff {n_in}, tested {n_tested}, "
          f"max |E-E64| {np.abs(e - r['energy']).max():.3g}")


def test_breakdown_and_fp64_kernel(ctx):
    w = W.config2(n_cand=64)
    ags, pop = W.device_objects(ctx, w)
    r = O.population_f64(w.antigens[0].atoms, w.pop.atoms, w.pop.cand_atom_off(), w.cutoff)
    pop.evaluate(forces=False, breakdown=True)
    lj, co, hy = pop.breakdown()
    assert tol(lj, r["lj"]).all() and tol(co, r["coulomb"]).all()
    assert np.array_equal(hy, r["hydrophobic"])
    pop.evaluate(fp64=True)
    e64, _ = pop.energies()
    lj64, co64, hy64 = pop.breakdown()
    scale = 1.0 + np.abs(r["lj"]) + np.abs(r["coulomb"])
    assert (np.abs(e64 - r["energy"]) <= 1e-9 * scale).all()
    assert (np.abs(lj64 - r["lj"]) <= 1e-9 * scale).all()
    assert np.array_equal(hy64, r["hydrophobic"])
    f64 = pop.forces_f64()
    for a, b in zip(f64, r["forces"]):
        assert (np.abs(a - b) <= 1e-9 * (1.0 + np.abs(b))).all()


def test_deterministic_and_tile_independent(ctx):
    w = W.config2(n_cand=128)
    ags, pop = W.device_objects(ctx, w)
    pop.evaluate(forces=True)
    e1, _ = pop.energies()
    f1 = pop.forces()
    pop.evaluate(forces=True)
    e2, _ = pop.energies()
    assert np.array_equal(e1.view(np.uint64), e2.view(np.uint64))
    # the same candidates in another order land in other tiles
    perm = np.random.default_rng(0).permutation(w.pop.n_cand)
    prots = []
    off = w.pop.cand_atom_off()
    for c in perm:
        r0, r1 = w.pop.cand_res_off[c], w.pop.cand_res_off[c + 1]
        cl = w.pop.candidate_cloud(int(c))
        res_start = (w.pop.res_atom_off[r0:r1 + 1] - off[c]).astype(np.uint32)
        from dnab.structure import AtomProtein
        prots.append(AtomProtein(cl, res_start, w.pop.res_amino_acid[r0:r1].copy(),
                                 (w.pop.res_ca_atom[r0:r1] - off[c]).astype(np.uint32)))
    pop2 = DevicePopulation(ctx, ags, population_from_proteins(prots))
    pop2.evaluate(forces=True)
    e3, _ = pop2.energies()
    assert np.array_equal(e3.view(np.uint64), e1[perm].view(np.uint64))
    # a one-candidate population of the same atoms gives the same bits (ragged == single)
    from dnab.structure import PopulationHost
    for c in (0, 7, 31):
        r0, r1 = w.pop.cand_res_off[c], w.pop.cand_res_off[c + 1]
        cl = w.pop.candidate_cloud(c)
        base = w.pop.cand_atom_off()[c]
        host = PopulationHost(np.zeros(1, np.uint32), np.array([0, r1 - r0], np.uint32),
                              (w.pop.res_atom_off[r0:r1 + 1] - base).astype(np.uint32),
                              (w.pop.res_ca_atom[r0:r1] - base).astype(np.uint32),
                              w.pop.res_amino_acid[r0:r1].copy(), cl)
        single = DevicePopulation(ctx, ags, host)
        single.evaluate(forces=True)
        es, _ = single.energies()
        assert es[0].view(np.uint64) == e1[c].view(np.uint64)
        fs = single.forces()
        for a, b in zip(fs, f1):
            assert np.array_equal(a.view(np.uint32), b[base:base + len(cl)].view(np.uint32))


def test_score_batch_mirrors_reference(ctx):
    w = W.config2(n_cand=48)
    ag = w.antigens[0].atoms
    cand = w.pop.atoms.gpu_atoms()
    n_ab = int(w.pop.cand_atom_off()[1])     # uniform-length view, as src/gpu.rs:166
    n_cand = len(cand) // n_ab
    e = ctx.score_batch(ag, cand[: n_cand * n_ab], n_ab)
    assert e.dtype == np.float32 and len(e) == n_cand
    off = np.arange(n_cand + 1, dtype=np.uint32) * n_ab
    sub = w.pop.atoms.subset(slice(0, n_cand * n_ab))
    r = O.population_f64(ag, sub, off, 10.0)
    assert tol(e, r["energy"]).all()
    assert len(ctx.score_batch(ag, cand, 0)) == 0          # src/gpu.rs:172-174
    assert len(ctx.score_batch(ag, cand[:0], n_ab)) == 0


def test_empty_inputs(ctx):
    ab = cloud([[0, 0, 0], [1, 1, 1]], 0.5, 1.9, 0.1, 1)
    empty = cloud(np.zeros((0, 3)), 0, 1.9, 0.1, 0)
    assert ctx.interaction_energy_atoms(empty, ab) == 0.0
    assert ctx.interaction_energy_atoms(ab, empty) == 0.0
    fx, fy, fz, e = ctx.atom_forces(empty, ab)
    assert (fx == 0).all() and e == 0.0
    # a population with a candidate of zero residues
    w = W.config2(n_cand=4)
    p = w.pop
    from dnab.structure import PopulationHost
    host = PopulationHost(np.zeros(5, np.uint32),
                          np.concatenate([p.cand_res_off[:3], p.cand_res_off[2:]]).astype(np.uint32),
                          p.res_atom_off, p.res_ca_atom, p.res_amino_acid, p.atoms)
    ags = [DeviceAntigen(ctx, w.antigens[0], cutoff=10.0)]
    pop = DevicePopulation(ctx, ags, host)
    pop.evaluate()
    e, _ = pop.energies()
    assert e[2] == 0.0


def test_err_estimate_reported(ctx):
    w = W.config2(n_cand=64)
    ags, pop, e, err, f, r = eval_vs_oracle(ctx, w)
    actual = np.abs(e - r["energy"])
    assert (err >= 0).all()
    print(f"max actual {actual.max():.3g}, max estimate {err.max():.3g}, "
          f"frac actual<=estimate {(actual <= err + 1e-9).mean():.3f}")


def test_tuning_changes_speed_not_results(ctx):
    # the pair kernels add per-chunk sums as exact integers: any work split
    # gives the same bits; the autotuner only picks the fastest
    w = W.config2(n_cand=256)
    ags, pop = W.device_objects(ctx, w)
    ref = None
    for sv, ch, wv in ((1, 1, 4), (48, 16, 8), (10000, 10000, 16), (24, 8, 16)):
        pop.set_tuning(sv, ch, wv)
        pop.evaluate(forces=True)
        e, err = pop.energies()
        f = pop.forces()
        bits = (e.view(np.uint64).copy(), err.view(np.uint32).copy(),
                [a.view(np.uint32).copy() for a in f])
        if ref is None:
            ref = bits
        else:
            assert np.array_equal(bits[0], ref[0]) and np.array_equal(bits[1], ref[1])
            for a, b in zip(bits[2], ref[2]):
                assert np.array_equal(a, b)
    best, ms = pop.autotune(reps=2)
    assert best == pop.tuning() and ms > 0
    assert best["rest_waves"] in (4, 8, 16)
    pop.evaluate(forces=True)
    e2, _ = pop.energies()
    assert np.array_equal(e2.view(np.uint64), ref[0])
    with pytest.raises(Exception):
        pop.set_tuning(0, 16, 8)
    with pytest.raises(Exception):
        pop.set_tuning(48, 16, 5)
