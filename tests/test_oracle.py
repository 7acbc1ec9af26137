"""CPU tests that pin the oracle (oracle/dnab_oracle.c) independently of it:
closed forms, the reference's own unit tests restated, brute-force set
definitions in numpy, finite differences, published known-answer vectors."""
import math

import numpy as np
import pytest

from oracle import oracle as O
from dnab.structure import AtomCloud, ResidueCloud, protein_from_ca_trace, aa_from_sequence

A = aa_from_sequence


def cloud(xyz, q, rmh, eps, hyd):
    xyz = np.asarray(xyz, np.float32).reshape(-1, 3)
    n = len(xyz)
    f = lambda v: np.broadcast_to(np.asarray(v, np.float32), (n,)).copy()
    return AtomCloud(xyz[:, 0], xyz[:, 1], xyz[:, 2], f(q), f(rmh), f(eps),
                     np.broadcast_to(np.asarray(hyd, np.uint8), (n,)).copy())


# ---- SpatialHashGrid keys and lists (src/spatial.rs) -------------------------

def np_keys(x, cell):
    inv = np.float32(1.0) / np.float32(cell)
    with np.errstate(invalid="ignore", over="ignore"):
        p = np.floor(np.asarray(x, np.float32) * inv).astype(np.float64)
    out = np.where(np.isnan(p), 0, np.clip(p, -2.0 ** 31, 2.0 ** 31 - 1)).astype(np.int64)
    return out.astype(np.int32)


def test_cell_of_closed_forms():
    assert O.cell_of(0.0, 10.0) == 0
    assert O.cell_of(9.999, 10.0) == 0
    assert O.cell_of(10.0, 10.0) == 1
    assert O.cell_of(-0.001, 10.0) == -1
    assert O.cell_of(-10.0, 10.0) == -1
    assert O.cell_of(float("nan"), 10.0) == 0          # Rust `as i32`: NaN -> 0
    assert O.cell_of(1e30, 10.0) == 2 ** 31 - 1        # saturating
    assert O.cell_of(-1e30, 10.0) == -(2 ** 31)
    assert O.cell_of(float("inf"), 10.0) == 2 ** 31 - 1


def test_cell_of_uses_rounded_reciprocal():
    # fl(1/12) * x differs from x / 12 for some x: the reference multiplies by
    # the f32 reciprocal (src/spatial.rs:26, 63), so must the oracle.
    inv = np.float32(1.0) / np.float32(12.0)
    k = np.arange(1, 20000, dtype=np.float32) * np.float32(12.0)
    xs = np.concatenate([np.nextafter(k, np.float32(0)), np.nextafter(-k, np.float32(0))])
    mul = np.floor(xs * inv)
    div = np.floor(xs / np.float32(12.0))
    diff = np.nonzero(mul != div)[0]
    assert len(diff) > 0, "no discriminating value found"
    for i in diff[:20]:
        assert O.cell_of(float(xs[i]), 12.0) == int(mul[i])
    rng = np.random.default_rng(5)
    xs = rng.uniform(-300, 300, 5000).astype(np.float32)
    assert [O.cell_of(float(v), 12.0) for v in xs] == list(np_keys(xs, 12.0))


def test_query_finds_nearby_atom():
    # src/spatial.rs:75-87
    off, idx = O.grid_query([0, 50, 100], [0, 50, 100], [0, 50, 100], 10.0, [1.0], [1.0], [1.0])
    assert 0 in idx and 1 not in idx


@pytest.mark.parametrize("cell", [8.0, 10.0, 12.0, 3.7])
def test_query_matches_definition(cell):
    rng = np.random.default_rng(int(cell * 10))
    p = rng.uniform(-25, 25, (300, 3)).astype(np.float32)
    q = rng.uniform(-35, 35, (40, 3)).astype(np.float32)
    off, idx = O.grid_query(p[:, 0], p[:, 1], p[:, 2], cell, q[:, 0], q[:, 1], q[:, 2])
    kp = np.stack([np_keys(p[:, a], cell) for a in range(3)], 1).astype(np.int64)
    kq = np.stack([np_keys(q[:, a], cell) for a in range(3)], 1).astype(np.int64)
    for j in range(len(q)):
        expect = []
        for dx in (-1, 0, 1):
            for dy in (-1, 0, 1):
                for dz in (-1, 0, 1):
                    c = kq[j] + (dx, dy, dz)
                    expect += list(np.nonzero((kp == c).all(1))[0])
        assert list(idx[off[j]:off[j + 1]]) == expect
        # as a set: every point within one cell in each axis
        cheb = np.abs(kp - kq[j]).max(1)
        assert sorted(expect) == list(np.nonzero(cheb <= 1)[0])


# ---- all-atom energy closed forms (src/energy.rs:261-314) --------------------

def two_atoms(r, q1=0.0, q2=0.0, h1=0, h2=0, rmh=(1.908, 1.824), eps=(0.1094, 0.17)):
    ag = cloud([0, 0, 0], q1, rmh[0], eps[0], h1)
    ab = cloud([r, 0, 0], q2, rmh[1], eps[1], h2)
    return ag, ab


def test_lj_minimum_closed_form():
    R = np.float32(1.908) + np.float32(1.824)
    ag, ab = two_atoms(float(R))
    e32 = O.energy_f32(ag, ab)
    e64, (fx, fy, fz), n = O.energy_f64(ag, ab)
    eps = math.sqrt(float(np.float32(0.1094)) * float(np.float32(0.17)))
    assert n == 1
    assert e64[0] == pytest.approx(-eps, rel=1e-12)
    assert abs(e32[0] - (-eps)) < 1e-6
    # zero slope at R_ij (the f32 position differs from R_ij by < 1 ulp)
    assert abs(fx[0]) < 1e-6
    ag2, ab2 = two_atoms(float(R) * 1.1)
    assert abs(O.energy_f64(ag2, ab2)[1][0][0]) > 1e-2
    assert e64[1] == 0 and e64[2] == 0


def test_coulomb_and_hydrophobic_closed_forms():
    ag, ab = two_atoms(4.0, q1=0.8, q2=-0.5, h1=1, h2=1)
    e64, (fx, _, _), _ = O.energy_f64(ag, ab)
    q1, q2 = float(np.float32(0.8)), float(np.float32(-0.5))
    assert e64[1] == pytest.approx(332.0 * q1 * q2 / 4.0, rel=1e-12)
    assert e64[2] == -0.5
    ag, ab = two_atoms(6.5, h1=1, h2=1)
    assert O.energy_f64(ag, ab)[0][2] == 0.0      # beyond 6 A: no bonus
    ag, ab = two_atoms(4.0, h1=1, h2=0)
    assert O.energy_f64(ag, ab)[0][2] == 0.0      # both atoms must be flagged
    ag, ab = two_atoms(4.0, q1=0.8, q2=-0.5)
    assert fx[0] < 0                               # opposite charges attract (src/energy.rs:445)


def test_cutoff_predicate_edges():
    # r^2 == 100 is inside (reference skips r^2 > 100), r^2 < 0.25 is outside.
    ag, ab = two_atoms(10.0, q1=1.0, q2=1.0)
    assert O.energy_f64(ag, ab)[2] == 1
    ag, ab = two_atoms(float(np.nextafter(np.float32(10.0), np.float32(11))), q1=1.0, q2=1.0)
    assert O.energy_f64(ag, ab)[2] == 0
    ag, ab = two_atoms(0.5, q1=1.0, q2=1.0)
    assert O.energy_f64(ag, ab)[2] == 1
    ag, ab = two_atoms(0.49, q1=1.0, q2=1.0)
    assert O.energy_f64(ag, ab)[2] == 0
    ag, ab = two_atoms(11.9, q1=1.0, q2=1.0)
    assert O.energy_f64(ag, ab, cutoff=12.0)[2] == 1


def test_atom_breakdown_sums_to_total():
    # src/energy.rs:496-507: Asp at the origin, Lys 3 A away.
    ag = protein_from_ca_trace([0.0], [0.0], [0.0], A("D"))
    ab = protein_from_ca_trace([3.0], [0.0], [0.0], A("K"))
    e32 = O.energy_f32(ag.atoms, ab.atoms)
    e64 = O.energy_f64(ag.atoms, ab.atoms)[0]
    tot = float(e32[0] + e32[1] + e32[2])
    assert abs(tot - float(np.float32(e32[0]) + np.float32(e32[1]) + np.float32(e32[2]))) < 1e-4
    assert abs(tot - e64.sum()) <= 1e-3 + 1e-4 * abs(e64.sum()) * 50  # f32 reference vs FP64


def test_f32_mirror_close_to_f64_on_random_system():
    rng = np.random.default_rng(7)
    n_ag, n_ab = 300, 60
    ag = cloud(rng.uniform(-12, 12, (n_ag, 3)), rng.uniform(-0.9, 0.9, n_ag),
               rng.choice([1.908, 1.824, 1.6612], n_ag), rng.choice([0.086, 0.17, 0.21], n_ag),
               rng.integers(0, 2, n_ag))
    ab = cloud(rng.uniform(-12, 12, (n_ab, 3)) + [14, 0, 0], rng.uniform(-0.9, 0.9, n_ab),
               rng.choice([1.908, 1.824], n_ab), rng.choice([0.086, 0.17], n_ab),
               rng.integers(0, 2, n_ab))
    e32 = O.energy_f32(ag, ab)
    e64 = O.energy_f64(ag, ab)[0]
    for a, b in zip(e32, e64):
        assert abs(a - b) <= 1e-2 + 1e-3 * abs(b)


def _energy_lj_coul(ag, ab_xyz, base):
    ab = cloud(ab_xyz, base.charge, base.r_min_half, base.epsilon, 0)
    e = O.energy_f64(ag, ab)[0]
    return e[0] + e[1]


def test_forces_are_negative_gradient_fd():
    rng = np.random.default_rng(11)
    ag = cloud(rng.uniform(-5, 5, (40, 3)), rng.uniform(-0.8, 0.8, 40), 1.908, 0.1094, 0)
    base = cloud([[7.0, 0.5, -0.3]], 0.6, 1.824, 0.17, 0)
    p = np.array([[7.0, 0.5, -0.3]])
    _, (fx, fy, fz), _ = O.energy_f64(ag, base)
    h = 1e-3
    grad = []
    for a in range(3):
        pp, pm = p.copy(), p.copy()
        pp[0, a] += h
        pm[0, a] -= h
        # coordinates are f32: use an exactly representable step
        ep = _energy_lj_coul(ag, pp.astype(np.float32), base)
        em = _energy_lj_coul(ag, pm.astype(np.float32), base)
        hh = float(np.float32(pp[0, a]) - np.float32(pm[0, a]))
        grad.append((ep - em) / hh)
    for f, g in zip((fx[0], fy[0], fz[0]), grad):
        assert f == pytest.approx(-g, rel=2e-5, abs=2e-6)


def test_lj_repulsion_close_range():
    ag, ab = two_atoms(1.0)
    _, (fx, _, _), _ = O.energy_f64(ag, ab)
    assert fx[0] > 0


def test_population_f64_matches_single_calls():
    rng = np.random.default_rng(3)
    ag = cloud(rng.uniform(-8, 8, (120, 3)), rng.uniform(-0.8, 0.8, 120), 1.908, 0.1094,
               rng.integers(0, 2, 120))
    parts = [cloud(rng.uniform(-8, 8, (k, 3)) + [10, 0, 0], rng.uniform(-0.8, 0.8, k), 1.824,
                   0.17, rng.integers(0, 2, k)) for k in (5, 17, 1, 30)]
    atoms = AtomCloud.concat(parts)
    off = np.cumsum([0] + [len(p) for p in parts]).astype(np.uint32)
    r = O.population_f64(ag, atoms, off, 10.0, nthreads=3)
    for c, p in enumerate(parts):
        e = O.energy_f64(ag, p)[0]
        assert r["energy"][c] == pytest.approx(e.sum(), rel=1e-13, abs=1e-12)


# ---- residue-level forces (src/energy.rs:154-234, 322-398, 400-473) ---------

def residue_cloud(points, aas):
    p = np.asarray(points, np.float32).reshape(-1, 3)
    return ResidueCloud.from_points(p[:, 0], p[:, 1], p[:, 2], A(aas))


def test_cg_grid_forces_match_brute_force():
    ag = residue_cloud([[0, 0, 0], [5, 0, 0]], "KL")
    ab = residue_cloud([[3, 0, 0]], "D")
    bf = O.cg_forces(ag, ab)
    gr = O.cg_forces_with_grid(ag, ab, 10.0)
    for a, b in zip(bf, gr):
        assert abs(a[0] - b[0]) < 1e-4
    assert bf[0][0] < 0                                # opposite_charges_attract
    ag = residue_cloud([[0, 0, 0]], "G")
    ab = residue_cloud([[1, 0, 0]], "G")
    assert O.cg_forces(ag, ab)[0][0] > 0               # lj_repulsion_at_close_range


def test_cg_grid_equals_brute_on_random_points():
    rng = np.random.default_rng(9)
    ag = residue_cloud(rng.uniform(-20, 20, (200, 3)), "".join(rng.choice(list("ARNDCQEGHILKMFPSTWYV"), 200)))
    ab = residue_cloud(rng.uniform(-25, 25, (50, 3)), "".join(rng.choice(list("ARNDCQEGHILKMFPSTWYV"), 50)))
    bf = O.cg_forces(ag, ab)
    gr = O.cg_forces_with_grid(ag, ab, 10.0)
    for a, b in zip(bf, gr):
        np.testing.assert_allclose(a, b, rtol=1e-4, atol=1e-4)


# ---- the step's helpers (src/diffusion.rs:102-109, 300-339) ------------------

def test_philox_known_answers():
    # Random123 philox4x32-10 known-answer vectors (Salmon et al. 2011).
    assert list(O.philox([0, 0, 0, 0], [0, 0])) == [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]
    assert list(O.philox([0xffffffff] * 4, [0xffffffff] * 2)) == \
        [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]
    assert list(O.philox([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344],
                         [0xa4093822, 0x299f31d0])) == [0xd16cfe09, 0x94fdcceb, 0x5001e420,
                                                        0x24126ea1]


def test_temperature_schedule_endpoints():
    assert O.sigma(0, 800) == pytest.approx(0.25 * math.sqrt(3.0), rel=1e-6)
    assert O.sigma(800, 800) == pytest.approx(0.25 * math.sqrt(0.02), rel=1e-5)


def _one_residue(ca=(30.0, 0.0, 0.0)):
    x = np.array([ca[0], ca[0] + 1.0], np.float32)
    y = np.array([ca[1], ca[1]], np.float32)
    z = np.array([ca[2], ca[2]], np.float32)
    return x, y, z, np.array([0, 2], np.uint32), np.array([0], np.uint32)


def base_params(**kw):
    p = dict(step=0.08, max_disp=2.0, restraint_k=0.02, restraint_radius=20.0, sigma=0.0,
             with_noise=0, force_model=0, iter=0, seed_lo=0, seed_hi=0, center=(0, 0, 0))
    p.update(kw)
    return p


def test_step_restraint_closed_form():
    x, y, z, off, ca = _one_residue((30.0, 0.0, 0.0))
    zero = np.zeros(2, np.float32)
    O.langevin_step(x, y, z, off, ca, base_params(), forces=(zero, zero, zero))
    # F = -0.02 (30 - 20)/30 * 30 = -0.2; dx = clamp(0.08 * -0.2) = -0.016
    assert x[0] == np.float32(30.0) + np.float32(np.float32(0.08) * np.float32(-0.2))
    assert x[1] - x[0] == pytest.approx(1.0, abs=1e-6)      # rigid translation


def test_step_clamp_and_noise_bounds():
    x, y, z, off, ca = _one_residue((20.0, 0.0, 0.0))
    big = np.array([1e6, 0], np.float32)
    zero = np.zeros(2, np.float32)
    O.langevin_step(x, y, z, off, ca, base_params(), forces=(big, zero, zero))
    assert x[0] == np.float32(22.0)                          # clamped at MAX_DISP
    xs = []
    for it in range(200):
        x, y, z, off, ca = _one_residue((20.0, 0.0, 0.0))
        O.langevin_step(x, y, z, off, ca, base_params(with_noise=1, sigma=0.5, iter=it, seed_lo=7),
                        forces=(zero, zero, zero))
        xs.append(float(x[0]) - 20.0)
    xs = np.array(xs)
    assert (np.abs(xs) <= 0.5).all() and xs.std() > 0.2      # U(-s, s)


# ---- H-bond, EEF1 solvation and disulfide terms -------------------------------

# the reference's f32 constants (src/energy.rs:24-25), as the FP64 oracle widens them
HB_R0, HB_SIGMA = float(np.float32(2.9)), float(np.float32(0.35))


def test_hbond_energy_closed_forms():
    # src/energy.rs:477-491: deepest at the ideal distance, decays, zero beyond 5 A
    assert abs(O.hbond_energy_f32(np.float32(HB_R0) * np.float32(HB_R0)) - (-1.5)) < 1e-5
    assert O.hbond_energy_f32(2.9 ** 2) < O.hbond_energy_f32(3.9 ** 2)
    assert O.hbond_energy_f32(25.0 + 1.0) == 0.0
    assert O.hbond_energy_f32(25.0) != 0.0          # the cutoff is inclusive
    # the Gaussian well: symmetric about r0, half depth at r0 +- sigma sqrt(2 ln 2)
    for d in (0.1, 0.3, 0.7):
        assert abs(O.hbond_energy_f64((HB_R0 - d) ** 2) - O.hbond_energy_f64((HB_R0 + d) ** 2)) < 1e-12
    w = HB_SIGMA * math.sqrt(2 * math.log(2))
    assert abs(O.hbond_energy_f64((HB_R0 + w) ** 2) - (-0.75)) < 1e-6
    assert abs(O.hbond_energy_f64(HB_R0 ** 2) - (-1.5)) < 1e-6
    # the f32 restatement follows the f64 one
    for r in np.linspace(2.0, 5.0, 31):
        assert abs(O.hbond_energy_f32(np.float32(r) ** 2) - O.hbond_energy_f64(float(np.float32(r) ** 2))) < 1e-5


def test_pair_solvation_signs_and_shape():
    # src/solvation.rs:92-117: hydrophobic burial lowers, polar burial raises the energy
    hyd = (1.0, 10.0, 3.5)
    pol = (-5.0, 14.0, 3.15)
    assert O.pair_solvation_f32(hyd, hyd, 9.0) < 0.0
    assert O.pair_solvation_f32(pol, pol, 9.0) > 0.0
    # symmetric in the two atoms; linear in dg_ref; exp(-(r/lambda)^2) decay
    a, b = (-0.187, 11.7, 3.7), (-5.33, 14.2, 3.15)
    assert abs(O.pair_solvation_f64(a, b, 16.0) - O.pair_solvation_f64(b, a, 16.0)) < 1e-15
    two = (2 * hyd[0], hyd[1], hyd[2])
    assert abs(O.pair_solvation_f64(two, two, 9.0) - 2 * O.pair_solvation_f64(hyd, hyd, 9.0)) < 1e-15
    lam = hyd[2]
    ratio = O.pair_solvation_f64(hyd, hyd, lam * lam) / O.pair_solvation_f64(hyd, hyd, 1e-12)
    assert abs(ratio - math.exp(-1.0)) < 1e-9
    # f(0) = 1 / (pi^(3/2) lambda^3) per burial direction
    f0 = -O.pair_solvation_f64(hyd, hyd, 1e-18) / (2 * hyd[0] * hyd[1])
    assert abs(f0 - 1.0 / (float(np.float32(5.568328)) * lam ** 3)) < 1e-9


def _cys_protein(xs):
    return protein_from_ca_trace(list(xs), [0.0] * len(xs), [0.0] * len(xs), A("C" * len(xs)))


def _sg(p):
    n = len(p.amino_acid)
    return O.sg_positions(p.atoms.x, p.atoms.y, p.atoms.z, p.res_start, p.amino_acid, 0, n)


def test_disulfide_energy_reference_cases():
    # src/allatom.rs:590-639
    assert O.disulfide_f32(*_sg(_cys_protein([0.0]))) == 0.0           # no partner
    assert O.disulfide_f32(*_sg(_cys_protein([0.0, 50.0]))) == 0.0     # beyond 4 A
    p = _cys_protein([0.0, 3.8, 7.6])
    x, y, z = _sg(p)
    assert len(x) == 3
    exp = np.float32(0.0)
    for i in range(3):
        for j in range(i + 1, 3):
            d = np.float32(np.sqrt(np.float32((x[i] - x[j]) ** 2 + (y[i] - y[j]) ** 2 + (z[i] - z[j]) ** 2)))
            if d > 4.0:
                continue
            t = (d - np.float32(2.05)) / np.float32(0.4)
            exp -= np.float32(4.0) * np.float32(np.exp(-0.5 * t * t))
    assert abs(O.disulfide_f32(x, y, z) - float(exp)) < 1e-5


def test_disulfide_well_closed_forms():
    # deepest (-4) at the ideal S-S distance, Gaussian in d, zero past 4 A
    x = np.array([0.0, 2.05], np.float32)
    z0 = np.zeros(2, np.float32)
    assert abs(O.disulfide_f64(x, z0, z0) - (-4.0)) < 1e-6
    w = 0.4 * math.sqrt(2 * math.log(2))
    assert abs(O.disulfide_f64(np.array([0.0, 2.05 + w], np.float32), z0, z0) - (-2.0)) < 1e-5
    assert O.disulfide_f64(np.array([0.0, 4.01], np.float32), z0, z0) == 0.0
    # three mutual partners add pairwise
    t = np.array([0.0, 2.05, 1.025], np.float32)
    u = np.array([0.0, 0.0, float(np.float32(2.05 * math.sqrt(3) / 2))], np.float32)
    e3 = O.disulfide_f64(t, u, np.zeros(3, np.float32))
    assert abs(e3 - 3 * (-4.0)) < 1e-4


def test_full_energy_f32_follows_f64_and_sums():
    # five terms of interaction_energy_atoms_breakdown (src/energy.rs:261-314)
    ag = protein_from_ca_trace([0.0, 3.8, 7.6, 11.4], [0.0] * 4, [0.0] * 4, A("DKSN"))
    ab = protein_from_ca_trace([1.0, 4.8, 8.6], [5.0] * 3, [0.5] * 3, A("ERT"))
    e32 = O.energy_full_f32(ag.atoms, ab.atoms)
    e64, per = O.energy_full_f64(ag.atoms, ab.atoms)
    assert e64[3] < 0.0 and e64[4] != 0.0           # H-bonds and solvation present
    for a, b in zip(e32, e64):
        assert abs(float(a) - b) <= 1e-3 + 1e-4 * abs(b)
    assert np.allclose(per.sum(0), e64, rtol=1e-12, atol=1e-12)
    # the first three terms are those of the three-term oracle
    three = O.energy_f64(ag.atoms, ab.atoms)[0]
    assert np.allclose(e64[:3], three, rtol=1e-12, atol=1e-12)
    # H-bond and solvation: brute force over all pairs with the pair functions
    # pinned above, the reference's pair sets (src/energy.rs:284-305: inside
    # the 0.25 <= r^2 <= 100 predicate, donor-acceptor either way;
    # src/solvation.rs:75-80: 0.25 <= r^2 < 81) and the type tables
    tab = O.reference_tables()
    eef = np.asarray(tab["eef1"], np.float32).reshape(-1, 3)    # the f32 table values
    don, acc = tab["hbond_donor"], tab["hbond_acceptor"]
    a, b = ag.atoms, ab.atoms
    hb = sv = 0.0
    for j in range(len(b)):
        for i in range(len(a)):
            dx, dy, dz = (np.float32(getattr(b, k)[j] - getattr(a, k)[i]) for k in "xyz")
            r2f = float(np.float32(np.float32(dx * dx + dy * dy) + np.float32(dz * dz)))
            # the sets by the f32 predicate, the values at the exact r^2
            r2 = sum((float(getattr(b, k)[j]) - float(getattr(a, k)[i])) ** 2 for k in "xyz")
            ta, tb = int(a.atom_type[i]), int(b.atom_type[j])
            if 0.25 <= r2f <= 25.0 and ((don[tb] and acc[ta]) or (acc[tb] and don[ta])):
                hb += O.hbond_energy_f64(r2)
            if 0.25 <= r2f < 81.0:
                sv += O.pair_solvation_f64(eef[ta], eef[tb], r2)
    assert abs(e64[3] - hb) <= 1e-9 * (1 + abs(hb)), (e64[3], hb)
    assert abs(e64[4] - sv) <= 1e-9 * (1 + abs(sv)), (e64[4], sv)


# ---- affinity maturation (src/diffusion.rs:924-961) ----

def _residues(p):
    rs = p.res_start
    return [p.atoms.subset(slice(int(rs[r]), int(rs[r + 1]))) for r in range(len(p.amino_acid))]


def _build(aa, rot, ca):
    return protein_from_ca_trace([ca[0]], [ca[1]], [ca[2]], np.array([aa], np.uint8),
                                 np.array([rot], np.uint8)).atoms


def _state(p):
    ca = [np.array([p.atoms.x[i], p.atoms.y[i], p.atoms.z[i]], np.float32) for i in p.ca_atom_idx]
    return _residues(p), [int(a) for a in p.amino_acid], [0] * len(p.amino_acid), ca


def _straight(start, n):
    # straight_chain_at (src/diffusion.rs:1116-1122): Ala every 3.8 A along x
    xs = [start[0] + i * np.float32(3.8) for i in range(n)]
    return protein_from_ca_trace(xs, [start[1]] * n, [start[2]] * n, A("A" * n))


def test_affinity_maturation_never_worsens_the_score():
    # affinity_maturation_never_worsens_the_score (src/diffusion.rs:1157-1172)
    ag = _straight([0.0, 0.0, 0.0], 6)
    ab = _straight([8.0, 0.0, 0.0], 6)
    res, aa, rot, ca = _state(ab)
    before = O.allatom_energy_f64(ag.atoms, res, aa)
    after = O.affinity_maturation_f64(ag.atoms, res, aa, rot, ca, range(6), _build)
    assert after <= before + 1e-3
    assert abs(after - O.allatom_energy_f64(ag.atoms, res, aa)) < 1e-3
    assert aa != [0] * 6                       # some mutation was accepted


def test_affinity_maturation_single_position_by_enumeration():
    # One position: the trials all see the other residues unchanged, so the
    # kept type is the first in ALL_AA order whose full score is strictly
    # below the start; with none, the undo leaves the old type rebuilt on an
    # ideal backbone (mutate_residue, src/allatom.rs:299-384), not the input
    # atoms.
    ag = protein_from_ca_trace([0.0, 3.8, 7.6, 11.4], [0.0] * 4, [0.0] * 4, A("DKSE"))
    ab = protein_from_ca_trace([1.0, 4.8, 8.6, 12.4], [6.0] * 4, [0.0] * 4, A("RLNT"))
    rng = np.random.default_rng(5)
    for k in "xyz":                           # off-ideal input geometry
        getattr(ab.atoms, k)[:] += rng.uniform(-0.3, 0.3, len(ab.atoms)).astype(np.float32)
    for r, far in ((1, False), (2, False)):
        res, aa, rot, ca = _state(ab)
        e0 = O.allatom_energy_f64(ag.atoms, res, aa)
        trials = []
        for cand in range(20):
            if cand == aa[r]:
                continue
            t = list(res)
            t[r] = _build(cand, 0, ca[r])
            ta = list(aa)
            ta[r] = cand
            trials.append((cand, O.allatom_energy_f64(ag.atoms, t, ta)))
        better = [(c, e) for c, e in trials if e < e0]
        got = O.affinity_maturation_f64(ag.atoms, res, aa, rot, ca, [r], _build)
        if better:
            assert aa[r] == better[0][0] and got == better[0][1]
        else:
            assert got == e0 and aa[r] == int(ab.amino_acid[r])
            ideal = _build(aa[r], 0, ca[r])
            assert np.array_equal(res[r].x, ideal.x)
    # a residue out of every cutoff: all trials tie with the start, none is
    # kept, and the residue comes back rebuilt (its jittered atoms are gone)
    far = protein_from_ca_trace([40.0], [0.0], [0.0], A("L"))
    far.atoms.x[4:] += np.float32(0.2)        # side chain off its ideal place, Calpha kept
    res, aa, rot, ca = _state(far)
    e0 = O.allatom_energy_f64(ag.atoms, res, aa)
    assert e0 == 0.0
    assert O.affinity_maturation_f64(ag.atoms, res, aa, rot, ca, [0], _build) == 0.0
    assert aa == [int(far.amino_acid[0])]
    assert np.array_equal(res[0].x, _build(aa[0], 0, ca[0]).x)
    assert not np.array_equal(res[0].x, far.atoms.x)
