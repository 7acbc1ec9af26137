"""GPU prescreen vs the numpy oracle (oracle/mlp_oracle.py).

Tolerances: map values within 2^-10 relative + 0.02 (fp16 storage of FP32
sums against FP64 sums rounded once to fp16); features within 2^-7 relative
+ 0.01 (bf16 of FP32 wave sums against FP64 sums); the map energy within
1e-4 relative + 1e-3; the MLP loss within 1e-3 relative, predictions within
1e-2 + 1e-2 |y|, each parameter tensor's gradient within 3e-2 relative L2 of
the oracle's (bf16 rounding points shared, f32 vs f64 accumulation and the
device exp differ). The fused inference kernel equals the layer-by-layer
GEMM forward pass bit for bit (same MFMA sequence and epilogues)."""
import numpy as np
import pytest

from oracle import mlp_oracle as M
from dnab import Dataset, Mlp, exact_topk, map_energy, population_features
from dnab import workloads as W

pytestmark = pytest.mark.gpu

SEGMENTS = [("W1", M.D * M.H), ("b1", M.H), ("W2", M.H * M.H), ("b2", M.H), ("W3", M.H * M.H),
            ("b3", M.H), ("w4", M.H), ("b4", 1)]


def _cls():
    from oracle.oracle import reference_tables
    return M.lj_classes(reference_tables()["amber_nb"])


def _ag_dict(c):
    return dict(x=c.x, y=c.y, z=c.z, charge=c.charge, r_min_half=c.r_min_half,
                epsilon=c.epsilon, hydrophobic=c.hydrophobic)


def _grid(w, dag):
    a = w.antigens[0].atoms
    lo = np.array([a.x.min(), a.y.min(), a.z.min()], np.float32)
    hi = np.array([a.x.max(), a.y.max(), a.z.max()], np.float32)
    origin, dims = M.map_grid(lo, hi, w.cutoff)
    info = dag.map_info()
    return origin, dims, info


def test_map_points_vs_oracle(ctx):
    w = W.config2(n_cand=8)
    ags, pop = W.device_objects(ctx, w)
    dag = ags[0]
    dag.build_maps()
    origin, dims, info = _grid(w, dag)
    assert info["spacing"] == M.MAP_H
    assert np.array_equal(info["origin"], origin) and np.array_equal(info["dims"], dims)
    a = w.antigens[0].atoms
    rng = np.random.default_rng(11)
    # random grid points, and grid points within 3 A of antigen atoms
    pick = rng.integers(0, len(a.x), 150)
    near = np.stack([a.x[pick], a.y[pick], a.z[pick]], 1) + rng.uniform(-3, 3, (150, 3))
    ijk_near = np.clip(np.round((near - origin) / M.MAP_H), 0, dims - 1).astype(np.int32)
    ijk = np.concatenate([rng.integers(0, dims, (150, 3)).astype(np.int32), ijk_near])
    ijk[::2] -= ijk[::2] % 2            # half the points on the radial-density grid
    got = dag.map_points(ijk).astype(np.float64)
    ref = M.map_values(_ag_dict(a), ijk, origin, M.MAP_H, w.cutoff, _cls(), a.hydrophobic)
    nv = M.N_LJ + 2 + M.N_RBF
    coarse = (ijk % 2 == 0).all(1)
    assert np.isnan(got[~coarse, M.N_LJ + 2:nv]).all() and not np.isnan(got[coarse, :nv]).any()
    got[~coarse, M.N_LJ + 2:nv] = ref[~coarse, M.N_LJ + 2:nv]
    bad = np.abs(got[:, :nv] - ref[:, :nv]) > 2 ** -10 * np.abs(ref[:, :nv]) + 0.02
    assert not bad.any(), (np.argwhere(bad)[:5], got[:, :nv][bad][:5], ref[:, :nv][bad][:5])
    assert (np.abs(ref[:, :M.N_LJ]) == M.MAP_CLAMP).any()   # some points sit in a wall
    # a cell's occupancy bit: some corner has an antigen atom within the cutoff
    lower = (ijk < dims - 1).all(1)
    assert not np.isnan(got[lower, -1]).any() and np.isnan(got[~lower, -1]).all()
    cells = ijk[lower][:60]
    corners = (cells[:, None, :] + M.CORNERS[None]).reshape(-1, 3)
    flags = M.map_values(_ag_dict(a), corners, origin, M.MAP_H, w.cutoff, _cls(),
                         a.hydrophobic)[:, -1].reshape(-1, 8)
    assert np.array_equal(got[lower][:60, -1], flags.max(1))
    assert 0 < flags.max(1).sum() < len(cells)              # both kinds sampled


def test_features_vs_oracle(ctx):
    w = W.config2(n_cand=24)
    ags, pop = W.device_objects(ctx, w)
    X = M.bf16_to_f64(population_features(pop))
    S = map_energy(pop)
    dag = ags[0]
    origin, dims, _ = _grid(w, dag)
    cls = _cls()
    at = w.pop.atoms
    off = w.pop.cand_atom_off()
    # the device's own map values at every grid point the candidates read
    need = []
    for a in range(len(at.x)):
        lk = M.atom_map_lookup(at.x[a], at.y[a], at.z[a], origin, M.MAP_H, dims)
        if lk is not None:
            need.append(lk[0][None, :] + M.CORNERS)
            need.append(lk[2][None, :])
    need = np.unique(np.concatenate(need), axis=0).astype(np.int32)
    vals = dag.map_points(need).astype(np.float64)
    table = {tuple(p): v for p, v in zip(need.tolist(), vals)}
    lookup = lambda ijk: np.stack([table[tuple(p)] for p in np.asarray(ijk).tolist()])
    n_in = 0
    for c in range(w.pop.n_cand):
        sl = slice(int(off[c]), int(off[c + 1]))
        atoms = dict(x=at.x[sl], y=at.y[sl], z=at.z[sl], charge=at.charge[sl],
                     r_min_half=at.r_min_half[sl], epsilon=at.epsilon[sl],
                     hydrophobic=at.hydrophobic[sl])
        row, total = M.features(atoms, origin, M.MAP_H, dims, lookup, cls)
        ref = M.bf16_round(row.astype(np.float32))
        assert (np.abs(X[c] - ref) <= 2 ** -7 * np.abs(ref) + 0.01).all(), (c, np.abs(X[c] - ref).max())
        assert S[c] == pytest.approx(total, rel=1e-4, abs=1e-3)
        n_in += row[M.F_NEAR] > 0
    assert n_in >= w.pop.n_cand // 2


def test_map_energy_tracks_exact_energy(ctx):
    """The screen's score without the network: the map energy follows the
    exact energy of candidates that do not clash."""
    w = W.config2(n_cand=4096, seed=31)
    ags, pop = W.device_objects(ctx, w)
    S = map_energy(pop).astype(np.float64)
    pop.evaluate(forces=False)
    E, _ = pop.energies()
    neg = E < -1.0
    assert neg.sum() > 100
    # the exact best 32 are among the map's best eighth (27 on this pool at
    # the 0.625 A default, 30 at 0.5 A: interpolation errors of a few kcal/mol
    # reorder the weak binders, rank correlation 0.56 over E < -1, less so
    # the strong ones)
    top = set(np.argsort(E, kind="stable")[:32].tolist())
    kept = set(np.argsort(S, kind="stable")[:len(S) // 8].tolist())
    assert len(top & kept) >= 26


def test_mlp_loss_grad_vs_oracle(ctx):
    mlp = Mlp(ctx, seed=3)
    p = mlp.get_params()
    assert len(p) == M.D * M.H + M.H + 2 * (M.H * M.H + M.H) + M.H + 1
    rng = np.random.default_rng(4)
    B = 64
    x = M.bf16_bits(np.log1p(rng.exponential(0.3, (B, M.D)) * (rng.random((B, M.D)) < 0.5))
                    .astype(np.float32))
    t = rng.normal(size=B).astype(np.float32)
    loss, g, pred = mlp.loss_grad(x, t)
    lo, go, yo = M.loss_grad(p, M.bf16_to_f64(x), t)
    assert loss == pytest.approx(lo, rel=1e-3)
    assert (np.abs(pred - yo) <= 1e-2 + 1e-2 * np.abs(yo)).all()
    off = 0
    for name, n in SEGMENTS:
        a, b = g[off:off + n], go[off:off + n]
        rel = np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-12)
        assert rel < 3e-2, (name, rel)
        off += n


def test_fused_inference_equals_gemm_forward(ctx):
    mlp = Mlp(ctx, seed=5)
    w = W.config2(n_cand=200, seed=9)
    ags, pop = W.device_objects(ctx, w)
    X = population_features(pop)
    fused = mlp.predict(pop)
    _, _, gemm = mlp.loss_grad(X, np.zeros(len(X), np.float32))
    assert np.array_equal(fused.view(np.uint32), gemm.view(np.uint32))
    yo = M.predict(mlp.get_params(), M.bf16_to_f64(X))
    assert (np.abs(fused - yo) <= 1e-2 + 1e-2 * np.abs(yo)).all()


def test_train_and_screen(ctx):
    """Training reduces the loss; on a fresh pool the network's kept set finds
    the exact top-64 better than an equally large random subset (the method
    without any screen) and about as well as the map energy alone (the method
    without the network); the re-scored energies are the exact ones."""
    ds = Dataset(ctx, capacity=4 * 4096)
    keep_alive = []
    for k, n_res in enumerate((100, 200, 300, 400)):
        r = W.globule_radius(n_res)
        ag = W.globule_antigen(n_res, r, 9000 + k)
        pop_h = W.shell_population(4096, 20, ag.ca_center_of_mass(), r + W.SHELL_GAP, r,
                                   (9 << 32) | (k << 20))
        w = W.Workload("train", [ag], pop_h, 10.0, 10.0, [r + W.SHELL_GAP], 0, 0, 0, 0, "train")
        ags, pop = W.device_objects(ctx, w)
        ds.add(pop)
        keep_alive.append((ags, pop))
    assert len(ds) == 4 * 4096
    mlp = Mlp(ctx, seed=1)
    hist = mlp.train(ds, n_steps=800, batch=256, lr=1e-3)
    assert np.isfinite(hist).all()
    assert hist[-50:].mean() < 0.5 * hist[:50].mean()
    w = W.config2(n_cand=8192, seed=77)
    ags, pop = W.device_objects(ctx, w)
    idx_e, e_exact = exact_topk(pop, 64)
    keep = 512
    idx_p, e_p = mlp.prescreen(pop, keep, 64)
    recall = len(set(idx_p.tolist()) & set(idx_e.tolist())) / 64
    S = map_energy(pop)
    recall_map = len(set(np.argsort(S, kind="stable")[:keep].tolist()) & set(idx_e.tolist())) / 64
    random_recall = keep / w.pop.n_cand
    print(f"prescreen recall@64 keeping {keep} of {w.pop.n_cand}: network {recall:.3f}, "
          f"map energy {recall_map:.3f}, random subset {random_recall:.3f}")
    assert recall >= random_recall + 0.6
    assert recall >= recall_map - 0.1
    # the exact re-score of the kept candidates gives their exact energies
    assert (np.diff(e_p) >= 0).all()
    pop.evaluate(forces=False)
    e_all, _ = pop.energies()
    assert np.array_equal(e_all[idx_p].view(np.uint64), e_p.view(np.uint64))


@pytest.mark.parametrize("n,m", [(100, 64), (4096, 512), (5000, 1), (65536, 2048), (200000, 512),
                                 (300, 1000)])
def test_select_smallest(ctx, n, m):
    from dnab import select_smallest
    rng = np.random.default_rng(n + m)
    s = np.round(rng.normal(size=n), 2).astype(np.float32)   # many ties
    s[rng.integers(0, n, 5)] = np.nan
    s[rng.integers(0, n, 3)] = -np.inf
    got = select_smallest(ctx, s, m)
    key = np.where(np.isnan(s), np.inf, s).astype(np.float64)
    nan_last = np.isnan(s).astype(np.int8)
    order = np.lexsort((np.arange(n), key, nan_last))
    want = np.sort(order[:min(m, n)])
    assert np.array_equal(got, want)
