"""CPU tests pinning the prescreen oracle (oracle/mlp_oracle.py): finite
differences of its float64 loss, bf16 rounding against float32 bit patterns,
closed forms of the antigen maps, and the map energy of atoms placed on grid
points against the FP64 all-atom energy oracle."""
import numpy as np
import pytest

from oracle import mlp_oracle as M


def test_bf16_rounding():
    v = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -2.5, 3.0e38, 1e-40], np.float32)
    r = M.bf16_round(v)
    assert r[0] == 1.0
    assert r[1] == 1.0                     # tie -> even
    assert r[2] == 1.0 + 2 ** -6           # tie -> even (up)
    assert r[3] == -2.5
    bits = M.bf16_bits(np.float32(1.0))
    assert int(np.asarray(bits).reshape(-1)[0]) == 0x3F80
    x = np.random.default_rng(0).normal(size=1000).astype(np.float32)
    assert np.all(np.abs(M.bf16_round(x) - x) <= np.abs(x) * 2 ** -8)
    assert np.any(np.abs(M.bf16_round(x) - x) > np.abs(x) * 2 ** -9)


AMBER_NB = [[1.908, 0.086], [1.908, 0.086], [1.908, 0.086], [1.908, 0.086], [1.908, 0.1094],
            [1.824, 0.17], [1.824, 0.17], [1.824, 0.17], [1.824, 0.17], [1.824, 0.17],
            [1.6612, 0.21], [1.6612, 0.21], [1.721, 0.2104], [1.6837, 0.17], [2.0, 0.25],
            [2.0, 0.25]]


def _one_atom(x, q=0.5, t=4, hyd=1):
    rmh, eps = AMBER_NB[t]
    return {k: np.array([v], np.float32) for k, v in
            dict(x=x, y=0.0, z=0.0, charge=q, r_min_half=rmh, epsilon=eps, hydrophobic=hyd).items()}


def test_lj_classes_from_the_table():
    from oracle.oracle import reference_tables
    cls = M.lj_classes(reference_tables()["amber_nb"])
    assert len(cls) == M.N_LJ
    assert cls[0] == (np.float32(1.908), np.float32(0.086))      # C, CA, CB, CC
    assert cls[1] == (np.float32(1.908), np.float32(0.1094))     # CT
    assert cls[6] == (np.float32(2.0), np.float32(0.25))         # S, SH


def test_map_closed_forms():
    cls = M.lj_classes(AMBER_NB)
    # one CT antigen atom (q = 0.5, hydrophobic) at -r on the x axis; the grid
    # point (0, 0, 0) sits at distance r from it
    origin = np.zeros(3, np.float32)
    for c, (rc_, ec) in enumerate(cls):
        R = float(rc_) + 1.908                        # probe class c with the CT atom
        v = M.map_values(_one_atom(-R), [[0, 0, 0]], origin, 0.5, 10.0, cls, [1], round16=False)[0]
        eps = np.sqrt(float(ec) * 0.1094)
        assert v[c] == pytest.approx(-eps, rel=1e-6)  # the LJ minimum
        assert v[M.N_LJ] == pytest.approx(332.0 * 0.5 / R, rel=1e-6)
        assert v[M.N_LJ + 1] == -0.5                  # hydrophobic, r^2 < 36
    # radial densities: a peak of 1 at r = mu_k, Gaussian tails on the others
    r = M.RBF_MU0 + 3 * M.RBF_DMU
    v = M.map_values(_one_atom(-r), [[0, 0, 0]], origin, 0.5, 10.0, cls, [0], round16=False)[0]
    rb = v[M.N_LJ + 2:]
    assert rb[3] == pytest.approx(1.0)
    for m in (1, 2):
        assert rb[3 + m] == pytest.approx(np.exp(-(m * M.RBF_DMU) ** 2 / (2 * M.RBF_SIGMA ** 2)))
    assert v[M.N_LJ + 1] == 0.0                       # not hydrophobic
    # beyond 6 A: no hydrophobic term; beyond the cutoff: nothing at all
    v = M.map_values(_one_atom(-7.0), [[0, 0, 0]], origin, 0.5, 10.0, cls, [1])[0]
    assert v[M.N_LJ + 1] == 0.0 and v[M.N_LJ] != 0.0
    v = M.map_values(_one_atom(-10.5), [[0, 0, 0]], origin, 0.5, 10.0, cls, [1])[0]
    assert not v.any()
    # a wall value is saturated at the fp16 clamp
    v = M.map_values(_one_atom(-0.6), [[0, 0, 0]], origin, 0.5, 10.0, cls, [1])[0]
    assert v[0] == M.MAP_CLAMP


def test_map_energy_on_grid_points_equals_exact_energy():
    """Candidate atoms placed exactly on grid points: the map energy (class LJ
    plane + q * potential + hydrophobic plane, no interpolation) equals the
    FP64 oracle's LJ + Coulomb + hydrophobic energy."""
    from oracle import oracle as O
    from dnab.structure import AtomCloud
    rng = np.random.default_rng(5)
    cls = M.lj_classes(AMBER_NB)
    n_ag, n_ab = 40, 30
    t_ag = rng.integers(0, 16, n_ag)
    ag = dict(x=rng.uniform(0, 12, n_ag), y=rng.uniform(0, 12, n_ag), z=rng.uniform(0, 12, n_ag),
              charge=rng.uniform(-0.8, 0.8, n_ag), r_min_half=[AMBER_NB[t][0] for t in t_ag],
              epsilon=[AMBER_NB[t][1] for t in t_ag], hydrophobic=rng.integers(0, 2, n_ag))
    ag = {k: np.asarray(v, np.float32) for k, v in ag.items()}
    origin = np.array([-6.0, -6.0, -6.0], np.float32)
    h = 0.5
    ijk = rng.integers(0, 48, (n_ab, 3))
    pts = (origin[None, :] + np.float32(h) * ijk.astype(np.float32)).astype(np.float32)
    t_ab = rng.integers(0, 16, n_ab)
    ab = dict(x=pts[:, 0], y=pts[:, 1], z=pts[:, 2], charge=rng.uniform(-0.8, 0.8, n_ab),
              r_min_half=[AMBER_NB[t][0] for t in t_ab], epsilon=[AMBER_NB[t][1] for t in t_ab],
              hydrophobic=rng.integers(0, 2, n_ab))
    ab = {k: np.asarray(v, np.float32) for k, v in ab.items()}
    vals = M.map_values(ag, ijk, origin, h, 10.0, cls, ag["hydrophobic"], round16=False)
    e_map = 0.0
    for a in range(n_ab):
        c = M.atom_class(ab["r_min_half"][a], ab["epsilon"][a], cls)
        e_map += vals[a, c] + float(ab["charge"][a]) * vals[a, M.N_LJ]
        e_map += vals[a, M.N_LJ + 1] if ab["hydrophobic"][a] else 0.0
    cl = lambda d: AtomCloud(d["x"], d["y"], d["z"], d["charge"], d["r_min_half"], d["epsilon"],
                             d["hydrophobic"])
    terms, _, n_pairs = O.energy_f64(cl(ag), cl(ab), 10.0)
    assert n_pairs > 0
    assert e_map == pytest.approx(terms.sum(), rel=1e-5, abs=1e-6)


def test_trilinear_and_lookup():
    corners = np.arange(8, dtype=np.float64)   # v = x + 2 y + 4 z at the corners
    for f in ([0.0, 0.0, 0.0], [0.25, 0.5, 0.75], [1.0, 1.0, 1.0]):
        assert M.trilinear(corners, f) == pytest.approx(f[0] + 2 * f[1] + 4 * f[2])
    origin = np.zeros(3, np.float32)
    i, frac, near = M.atom_map_lookup(1.3, 0.2, 0.9, origin, 0.5, [10, 10, 10])
    assert list(i) == [2, 0, 1] and list(near) == [2, 0, 2]  # nearest point of the 1 A grid
    assert M.atom_map_lookup(4.6, 0.2, 0.2, origin, 0.5, [10, 10, 10]) is None  # last plane
    assert M.atom_map_lookup(-0.1, 0.2, 0.2, origin, 0.5, [10, 10, 10]) is None


def test_gradient_finite_differences():
    rng = np.random.default_rng(1)
    d, h, B = 12, 5, 7
    n = d * h + h + h * h + h + h * h + h + h + 1
    p = rng.normal(scale=0.4, size=n)
    x = rng.normal(size=(B, d))
    t = rng.normal(size=B)
    loss, g, _ = M.loss_grad(p, x, t, d, h, bf16=False)
    eps = 1e-6
    for i in rng.choice(n, 40, replace=False):
        pp, pm = p.copy(), p.copy()
        pp[i] += eps
        pm[i] -= eps
        fd = (M.loss_grad(pp, x, t, d, h, bf16=False)[0] -
              M.loss_grad(pm, x, t, d, h, bf16=False)[0]) / (2 * eps)
        assert g[i] == pytest.approx(fd, rel=1e-5, abs=1e-8)
