"""Batched affinity maturation (src/diffusion.rs:924-961) on the GPU vs the
oracle's literal scan, which fully re-scores every trial in FP64.

Bars, as written below: the final sequences, rotamer indices, rebuilt flags
and final atom coordinates are exact; the returned scores satisfy
|x - x64| <= 1e-3 + 1e-4 |x64| (kcal/mol). The workloads avoid clashing
binders, whose FP32 pair terms carry absolute errors large enough to flip a
greedy decision against the FP64 oracle; the smallest decision margin seen is
printed.
"""
import numpy as np
import pytest

from oracle import oracle as O
from dnab import workloads as W
from dnab.structure import AtomProtein, aa_from_sequence, protein_from_ca_trace

pytestmark = pytest.mark.gpu


def tol(x, x64):
    return np.abs(np.asarray(x, np.float64) - x64) <= 1e-3 + 1e-4 * np.abs(x64)


def interface(ab, ag):
    # is_interface (src/filters.rs:96-105): Calpha within 8 A of an antigen Calpha
    p, q = ab.ca_pos(), ag.ca_pos()
    d = p[:, None, :] - q[None, :, :]
    r2 = (d[..., 0] * d[..., 0] + d[..., 1] * d[..., 1]) + d[..., 2] * d[..., 2]
    return [int(r) for r in np.nonzero((r2 < np.float32(64.0)).any(axis=1))[0]]


def build(aa, rot, ca):
    return protein_from_ca_trace([ca[0]], [ca[1]], [ca[2]], np.array([aa], np.uint8),
                                 np.array([rot], np.uint8)).atoms


def oracle_scan(ag, ab, mutable, rot, trace=None):
    rs = ab.res_start
    res = [ab.atoms.subset(slice(int(rs[r]), int(rs[r + 1]))) for r in range(len(ab.amino_acid))]
    aa = [int(a) for a in ab.amino_acid]
    rot = [int(k) for k in rot]
    ca = [np.array([ab.atoms.x[i], ab.atoms.y[i], ab.atoms.z[i]], np.float32)
          for i in ab.ca_atom_idx]
    best = O.affinity_maturation_f64(ag.atoms, res, aa, rot, ca, mutable, build, trace=trace)
    return best, res, aa, rot


def straight(start, n, seq=None):
    # straight_chain_at (src/diffusion.rs:1116-1122)
    xs = [start[0] + i * np.float32(3.8) for i in range(n)]
    return protein_from_ca_trace(xs, [start[1]] * n, [start[2]] * n,
                                 aa_from_sequence(seq or "A" * n))


def test_never_worsens_the_score(ctx):
    # affinity_maturation_never_worsens_the_score (src/diffusion.rs:1157-1172)
    ag = straight([0.0, 0.0, 0.0], 6)
    ab = straight([8.0, 0.0, 0.0], 6)
    before = ctx.allatom_energy(ag, ab)
    prots, e, rots, flags = ctx.affinity_maturation(ag, [ab], [list(range(6))])
    assert e[0] <= before + 1e-3
    assert abs(e[0] - ctx.allatom_energy(ag, prots[0])) < 1e-3 * (1 + abs(e[0]))
    assert flags[0].all()                    # every position tried, so every residue rebuilt


def _binders(n, seed):
    w = W.config2(n_cand=64)
    ag = w.antigens[0]
    ctr = ag.ca_center_of_mass()
    rng = np.random.default_rng(seed)
    out = []
    for c in range(64):
        if len(out) == n:
            break
        r0, r1 = int(w.pop.cand_res_off[c]), int(w.pop.cand_res_off[c + 1])
        aa = w.pop.res_amino_acid[r0:r1].copy()
        ca = w.pop.atoms.subset(w.pop.res_ca_atom[r0:r1])
        rot = np.where((aa == 0) | (aa == 7), 0, rng.integers(0, 5, len(aa))).astype(np.uint8)
        # move the binder radially out from the antigen centre to the first
        # shift that leaves contact without clashes
        v0 = np.stack([ca.x, ca.y, ca.z], 1) - ctr
        for shift in np.arange(0.0, 6.01, 0.5):
            p = ctr + v0 * (1.0 + shift / np.linalg.norm(v0, axis=1, keepdims=True))
            ab = protein_from_ca_trace(p[:, 0].astype(np.float32), p[:, 1].astype(np.float32),
                                       p[:, 2].astype(np.float32), aa, rot)
            e = O.allatom_energy_f64(ag.atoms, [ab.atoms], [255])
            mut = interface(ab, ag)
            if abs(e) < 1e4 or len(mut) < 3:
                break
        if abs(e) < 1e4 and len(mut) >= 3:
            # jitter the side chains of the even residues off their ideal places
            for r in range(0, len(aa), 2):
                s0, s1 = int(ab.res_start[r]) + 5, int(ab.res_start[r + 1])
                for k in "xyz":
                    getattr(ab.atoms, k)[s0:s1] += rng.uniform(-0.15, 0.15, max(0, s1 - s0)).astype(np.float32)
            out.append((ab, rot, mut))
    return ag, out


def test_maturation_matches_the_oracle_scan(ctx):
    ag, items = _binders(4, seed=11)
    assert len(items) == 4
    abs_, rots, muts = zip(*items)
    prots, e, rot_out, flags = ctx.affinity_maturation(ag, list(abs_), list(muts), list(rots))
    n_mut_kept = 0
    margin = np.inf
    for a, (ab, rot, mut) in enumerate(items):
        trace = []
        best, res, aa, rot_o = oracle_scan(ag, ab, mut, rot, trace)
        margin = min(margin, min(abs(t[2] - t[3]) / (1 + abs(t[3])) for t in trace))
        p = prots[a]
        assert list(p.amino_acid) == aa, (a, list(p.amino_acid), aa)
        assert list(rot_out[a]) == rot_o
        assert tol(e[a], best), (a, e[a], best)
        final = O.concat_residues(res)
        for k in "xyz":
            assert np.array_equal(getattr(p.atoms, k), getattr(final, k)), (a, k)
        assert np.array_equal(p.atoms.atom_type, final.atom_type)
        n_mut_kept += sum(int(x != y) for x, y in zip(aa, ab.amino_acid))
        # rebuilt exactly where a position was tried
        assert list(np.nonzero(flags[a])[0]) == sorted(mut)
        print(f"antibody {a}: E {e[a]:.4f} (oracle {best:.4f}), mutations kept "
              f"{sum(int(x != y) for x, y in zip(aa, ab.amino_acid))}/{len(mut)}")
    assert n_mut_kept > 0
    print(f"smallest relative decision margin of the oracle's trials: {margin:.3g}")


def test_maturation_edge_cases(ctx):
    ag = straight([0.0, 0.0, 0.0], 4, "DKSE")
    ab = protein_from_ca_trace([1.0, 4.8, 8.6], [6.0] * 3, [0.0] * 3, aa_from_sequence("RCN"))
    # no mutable positions: the score is allatom_energy and nothing changes
    prots, e, rots, flags = ctx.affinity_maturation(ag, [ab, ab], [[], [1]])
    assert abs(e[0] - ctx.allatom_energy(ag, ab)) < 1e-6 * (1 + abs(e[0]))
    for k in "xyz":
        assert np.array_equal(getattr(prots[0].atoms, k), getattr(ab.atoms, k))
    assert not flags[0].any() and list(np.nonzero(flags[1])[0]) == [1]
    # an empty batch
    prots, e, rots, flags = ctx.affinity_maturation(ag, [], [])
    assert prots == [] and len(e) == 0
    # positions must be ascending and inside the antibody
    with pytest.raises(Exception):
        ctx.affinity_maturation(ag, [ab], [[2, 1]])
    with pytest.raises(Exception):
        ctx.affinity_maturation(ag, [ab], [[3]])
    with pytest.raises(Exception):
        ctx.affinity_maturation(ag, [ab], [[0]], [np.array([9, 0, 0], np.uint8)])
