"""The GPU pair path against the committed vectors of
tests/golden/pair_fixtures.npz (no live oracle involved):

- closed forms (one antigen atom, one-atom binders at FP32-exact distances;
  reference formulas src/energy.rs:13-21, 269-301, 382-386): population
  energies and forces, and score_batch (src/gpu.rs:166-267), within
  1e-3 + 1e-4 |E| of the analytic FP64 values;
- the seeded c1/c2 sets: per-candidate energies and per-atom forces within
  1e-3 + 1e-4 |ref| of the committed FP64 oracle vectors, in-cutoff pair
  counts bit-exact;
- the 27-cell neighbour lists and cell keys at 8 A and 10 A cells
  (src/spatial.rs:44-67) bit-exact, order included.
"""
from pathlib import Path

import numpy as np
import pytest

import dnab
from dnab.structure import AtomCloud, PopulationHost

pytestmark = pytest.mark.gpu

FIX = Path(__file__).resolve().parent / "golden" / "pair_fixtures.npz"


@pytest.fixture(scope="module")
def fx():
    return np.load(FIX)


def within(got, ref):
    got, ref = np.asarray(got, np.float64), np.asarray(ref, np.float64)
    return np.abs(got - ref) <= 1e-3 + 1e-4 * np.abs(ref)


def cloud(fx, p):
    return AtomCloud(fx[p + "_x"], fx[p + "_y"], fx[p + "_z"], fx[p + "_q"], fx[p + "_rmh"],
                     fx[p + "_eps"], fx[p + "_hyd"])


def test_closed_forms(ctx, fx):
    a = fx["closed_ag"]
    ag = AtomCloud([a[0]], [a[1]], [a[2]], [a[3]], [a[4]], [a[5]], [int(a[6])])
    cs = fx["closed_cases"]
    n = len(cs)
    ab = AtomCloud(cs[:, 0], np.zeros(n), np.zeros(n), cs[:, 1], cs[:, 2], cs[:, 3],
                   cs[:, 4].astype(np.uint8))
    host = PopulationHost(np.zeros(n, np.uint32), np.arange(n + 1, dtype=np.uint32),
                          np.arange(n + 1, dtype=np.uint32), np.arange(n, dtype=np.uint32),
                          np.zeros(n, np.uint8), ab)
    dag = dnab.DeviceAntigen(ctx, ag, cutoff=10.0)
    pop = dnab.DevicePopulation(ctx, [dag], host)
    pop.evaluate(forces=True)
    e, _ = pop.energies()
    f_x, f_y, f_z = pop.forces()
    assert within(e, fx["closed_energy"]).all(), (e, fx["closed_energy"])
    assert within(f_x, fx["closed_fx"]).all(), (f_x, fx["closed_fx"])
    assert np.all(f_y == 0.0) and np.all(f_z == 0.0)
    n_in, _ = pop.pair_count()
    assert n_in == int(np.count_nonzero(fx["closed_energy"]))
    # the reference-shaped entry: score_batch over GpuAtom rows, one atom each
    sb = ctx.score_batch(ag, ab.gpu_atoms(), 1)
    assert within(sb, fx["closed_energy"]).all(), (sb, fx["closed_energy"])
    pop.close()
    dag.close()


@pytest.mark.parametrize("name", ["c1", "c2"])
def test_seeded_vectors(ctx, fx, name):
    ag, ab = cloud(fx, name + "_ag"), cloud(fx, name + "_ab")
    cro = fx[name + "_cand_res_off"]
    host = PopulationHost(np.zeros(len(cro) - 1, np.uint32), cro, fx[name + "_res_atom_off"],
                          fx[name + "_res_ca_atom"], fx[name + "_res_aa"], ab)
    dag = dnab.DeviceAntigen(ctx, ag, cutoff=float(fx[name + "_cutoff"]),
                             cell_size=float(fx[name + "_cell"]))
    pop = dnab.DevicePopulation(ctx, [dag], host)
    pop.evaluate(forces=True)
    e, err = pop.energies()
    assert within(e, fx[name + "_energy"]).all(), (e, fx[name + "_energy"])
    # the per-result error estimate covers the actual error
    assert np.all(np.abs(e - fx[name + "_energy"]) <= err + 1e-9)
    got = pop.forces()
    for k, c in enumerate("xyz"):
        ok = within(got[k], fx[f"{name}_f{c}"])
        assert ok.all(), (c, np.nonzero(~ok)[0][:5])
    n_in, _ = pop.pair_count()
    assert n_in == int(fx[name + "_pairs"].sum())
    pop.close()
    dag.close()


@pytest.mark.parametrize("cell", [8, 10])
def test_grid_lists(ctx, fx, cell):
    ag = cloud(fx, "c1_ag")
    g = dnab.SpatialHashGrid(ctx, float(cell))
    g.build(ag.x, ag.y, ag.z)
    px, py, pz = fx["grid_px"], fx["grid_py"], fx["grid_pz"]
    off, idx = g.query_neighbors(px, py, pz)
    assert np.array_equal(off, fx[f"grid{cell}_offsets"])
    assert np.array_equal(idx, fx[f"grid{cell}_indices"])
    assert np.array_equal(g.cell_of(px, py, pz), fx[f"grid{cell}_keys"])
    g.close()
