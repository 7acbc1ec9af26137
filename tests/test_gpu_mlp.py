"""GPU prescreen vs the numpy oracle (oracle/mlp_oracle.py).

Tolerances: features within 2^-7 relative + 1e-3 (bf16 rounding of log1p
sums whose f32 atomics add in another order); the MLP loss within 1e-3
relative, predictions within 1e-2 + 1e-2 |y|, each parameter tensor's
gradient within 3e-2 relative L2 of the oracle's (bf16 rounding points
shared, f32 vs f64 accumulation and the device exp differ)."""
import numpy as np
import pytest

from oracle import mlp_oracle as M
from oracle import oracle as O
from dnab import Dataset, Mlp, exact_topk, population_features
from dnab import workloads as W

pytestmark = pytest.mark.gpu

SEGMENTS = [("W1", M.D * M.H), ("b1", M.H), ("W2", M.H * M.H), ("b2", M.H), ("W3", M.H * M.H),
            ("b3", M.H), ("w4", M.H), ("b4", 1)]


def test_features_vs_oracle(ctx):
    w = W.config2(n_cand=24)
    ags, pop = W.device_objects(ctx, w)
    X = population_features(pop)
    ag = w.antigens[0]
    ag_ca = ag.ca_pos()
    pos = np.stack([w.pop.atoms.x, w.pop.atoms.y, w.pop.atoms.z], 1)
    n_contact = 0
    for c in range(w.pop.n_cand):
        r0, r1 = w.pop.cand_res_off[c], w.pop.cand_res_off[c + 1]
        ca = pos[w.pop.res_ca_atom[r0:r1]]
        h = M.features(ca, w.pop.res_amino_acid[r0:r1], ag_ca, ag.amino_acid)
        ref = M.bf16_round(np.log1p(h).astype(np.float32))
        got = M.bf16_to_f64(X[c])
        assert (np.abs(got - ref) <= 2 ** -7 * np.abs(ref) + 1e-3).all(), c
        # a candidate with no Calpha within 14 A of the antigen has an all-zero row
        assert (got.sum() > 0) == (h.sum() > 0), c
        n_contact += h.sum() > 0
    assert n_contact >= w.pop.n_cand // 2


def test_mlp_loss_grad_vs_oracle(ctx):
    mlp = Mlp(ctx, seed=3)
    p = mlp.get_params()
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


def test_train_reduces_loss_and_screens_better_than_random(ctx):
    # training data: binders around antigens of 100..400 residues
    ds = Dataset(ctx, capacity=4 * 4096)
    keep = []
    for k, n_res in enumerate((100, 200, 300, 400)):
        r = W.globule_radius(n_res)
        ag = W.globule_antigen(n_res, r, 9000 + k)
        binders = [W.shell_binder(20, ag.ca_center_of_mass(), r + W.SHELL_GAP, r,
                                  (9 << 32) | (k << 20) | c) for c in range(4096)]
        from dnab.structure import population_from_proteins
        w = W.Workload("train", [ag], population_from_proteins(binders), 10.0, 10.0,
                       [r + W.SHELL_GAP], 0, 0, 0, 0, "train")
        ags, pop = W.device_objects(ctx, w)
        ds.add(pop)
        keep.append((ags, pop))
    assert len(ds) == 4 * 4096
    mlp = Mlp(ctx, seed=1)
    hist = mlp.train(ds, n_steps=600, batch=256, lr=1e-3)
    assert np.isfinite(hist).all()
    assert hist[-50:].mean() < 0.75 * hist[:50].mean()
    # screening pool against a fresh 300-residue antigen: the network's kept
    # set against an equally large random subset (the method without it)
    w = W.config2(n_cand=8192, seed=77)
    ags, pop = W.device_objects(ctx, w)
    idx_e, e_exact = exact_topk(pop, 64)
    keep = 2048
    idx_p, e_p = mlp.prescreen(pop, keep, 64)
    recall = len(set(idx_p.tolist()) & set(idx_e.tolist())) / 64
    random_recall = keep / w.pop.n_cand
    print(f"prescreen recall@64 keeping {keep} of {w.pop.n_cand}: {recall:.3f} "
          f"(random subset {random_recall:.3f})")
    assert recall >= random_recall + 0.15
    # the exact re-score of the kept candidates gives their exact energies
    assert (np.diff(e_p) >= 0).all()
    pop.evaluate(forces=False)
    e_all, _ = pop.energies()
    assert np.array_equal(e_all[idx_p].view(np.uint64), e_p.view(np.uint64))
