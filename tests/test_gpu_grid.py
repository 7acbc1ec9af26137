"""GPU SpatialHashGrid and residue-level forces vs the oracle: bit-exact keys,
identical neighbour lists (same order), bit-identical f32 forces."""
import numpy as np
import pytest

from oracle import oracle as O
from dnab import SpatialHashGrid
from dnab.structure import ResidueCloud, aa_from_sequence

pytestmark = pytest.mark.gpu
AAS = "ARNDCQEGHILKMFPSTWYV"


def np_keys(x, cell):
    inv = np.float32(1.0) / np.float32(cell)
    with np.errstate(invalid="ignore", over="ignore"):
        p = np.floor(np.asarray(x, np.float32) * inv).astype(np.float64)
    return np.where(np.isnan(p), 0, np.clip(p, -2.0 ** 31, 2.0 ** 31 - 1)).astype(np.int64)


@pytest.mark.parametrize("cell", [8.0, 10.0, 12.0, 3.7])
def test_cell_of_bit_exact(ctx, cell):
    rng = np.random.default_rng(1)
    k = np.arange(1, 3000, dtype=np.float32) * np.float32(cell)
    xs = np.concatenate([rng.uniform(-300, 300, 4000).astype(np.float32),
                         np.nextafter(k, np.float32(0)), np.nextafter(-k, np.float32(0)), k,
                         np.array([0.0, -0.0, np.nan, np.inf, -np.inf, 1e30, -1e30], np.float32)])
    g = SpatialHashGrid(ctx, cell)
    g.build([0.0], [0.0], [0.0])
    keys = g.cell_of(xs, xs, xs)
    expect = np_keys(xs, cell)
    assert (keys[:, 0] == expect).all()
    sample = xs[:: 97]
    assert [O.cell_of(float(v), cell) for v in sample] == list(np_keys(sample, cell))


def test_query_finds_nearby_atom(ctx):
    # src/spatial.rs:75-87
    g = SpatialHashGrid(ctx, 10.0)
    g.build([0, 50, 100], [0, 50, 100], [0, 50, 100])
    off, idx = g.query_neighbors([1.0], [1.0], [1.0])
    assert 0 in idx and 1 not in idx


@pytest.mark.parametrize("cell,n,seed", [(10.0, 500, 1), (8.0, 2545, 2), (12.0, 1000, 3),
                                         (3.7, 300, 4), (10.0, 1, 5)])
def test_query_lists_identical(ctx, cell, n, seed):
    rng = np.random.default_rng(seed)
    p = rng.uniform(-30, 30, (n, 3)).astype(np.float32)
    q = np.concatenate([rng.uniform(-45, 45, (150, 3)), p[:50] + 0.01]).astype(np.float32)
    g = SpatialHashGrid(ctx, cell)
    g.build(p[:, 0], p[:, 1], p[:, 2])
    off, idx = g.query_neighbors(q[:, 0], q[:, 1], q[:, 2])
    eoff, eidx = O.grid_query(p[:, 0], p[:, 1], p[:, 2], cell, q[:, 0], q[:, 1], q[:, 2])
    assert np.array_equal(off, eoff)
    assert np.array_equal(idx, eidx)
    info = g.info()
    keys = np.stack([np_keys(p[:, a], cell) for a in range(3)], 1)
    assert info["n_cells_nonempty"] == len(np.unique(keys, axis=0))
    assert info["key_min"] == list(keys.min(0)) and info["key_max"] == list(keys.max(0))


def test_empty_grid(ctx):
    g = SpatialHashGrid(ctx, 10.0)
    g.build(np.zeros(0), np.zeros(0), np.zeros(0))
    off, idx = g.query_neighbors([0.0, 5.0], [0.0, 5.0], [0.0, 5.0])
    assert list(off) == [0, 0, 0] and len(idx) == 0


def residue_cloud(points, aas):
    p = np.asarray(points, np.float32).reshape(-1, 3)
    return ResidueCloud.from_points(p[:, 0], p[:, 1], p[:, 2], aa_from_sequence(aas))


def bits(a):
    return np.asarray(a, np.float32).view(np.uint32)


def test_reference_force_tests(ctx):
    # src/energy.rs:417-473
    ag = residue_cloud([[0, 0, 0], [5, 0, 0]], "KL")
    ab = residue_cloud([[3, 0, 0]], "D")
    g = SpatialHashGrid(ctx, 10.0)
    g.build(ag.x, ag.y, ag.z)
    bf = ctx.compute_forces(ag, ab)
    gr = ctx.compute_forces_with_grid(ag, g, ab)
    for a, b in zip(bf, gr):
        assert abs(a[0] - b[0]) < 1e-4
    assert bf[0][0] < 0
    ag1 = residue_cloud([[0, 0, 0]], "G")
    ab1 = residue_cloud([[1, 0, 0]], "G")
    assert ctx.compute_forces(ag1, ab1)[0][0] > 0


@pytest.mark.parametrize("seed", [0, 1])
def test_cg_forces_bit_identical_to_oracle(ctx, seed):
    rng = np.random.default_rng(seed)
    ag = residue_cloud(rng.uniform(-20, 20, (300, 3)), "".join(rng.choice(list(AAS), 300)))
    ab = residue_cloud(rng.uniform(-28, 28, (200, 3)), "".join(rng.choice(list(AAS), 200)))
    g = SpatialHashGrid(ctx, 10.0)
    g.build(ag.x, ag.y, ag.z)
    gr = ctx.compute_forces_with_grid(ag, g, ab)
    ogr = O.cg_forces_with_grid(ag, ab, 10.0)
    for a, b in zip(gr, ogr):
        assert np.array_equal(bits(a), bits(b))
    bf = ctx.compute_forces(ag, ab)
    obf = O.cg_forces(ag, ab)
    for a, b in zip(bf, obf):
        assert np.array_equal(bits(a), bits(b))
