"""The oracle pinned to committed vectors (tests/golden/pair_fixtures.npz,
written by tests/golden/gen_pair_fixtures.py):

- closed forms: one antigen atom against one-atom binders at distances exact
  in FP32, whose energies and forces follow from the reference's formulas
  (src/energy.rs:13-21, 269-301, 382-386): LJ = -eps at r = R_ij, Coulomb
  332 q q / r, -0.5 per hydrophobic pair inside 6 A, and the pair predicate's
  edges (r^2 = 100 and 0.25 kept, beyond them nothing);
- the FP64 oracle reproduces its own committed per-candidate energies, terms,
  pair counts and per-atom forces on the seeded c1/c2 inputs, and the
  committed 27-cell neighbour lists and cell keys (src/spatial.rs:44-67), so
  a change to the oracle cannot silently move the kernels' reference.
"""
from pathlib import Path

import numpy as np
import pytest

from oracle import oracle as O
from dnab.structure import AtomCloud

FIX = Path(__file__).resolve().parent / "golden" / "pair_fixtures.npz"


@pytest.fixture(scope="module")
def fx():
    return np.load(FIX)


def closed_clouds(fx):
    a = fx["closed_ag"]
    ag = AtomCloud([a[0]], [a[1]], [a[2]], [a[3]], [a[4]], [a[5]], [int(a[6])])
    cs = fx["closed_cases"]
    n = len(cs)
    ab = AtomCloud(cs[:, 0], np.zeros(n), np.zeros(n), cs[:, 1], cs[:, 2], cs[:, 3],
                   cs[:, 4].astype(np.uint8))
    return ag, ab


def cloud(fx, p):
    return AtomCloud(fx[p + "_x"], fx[p + "_y"], fx[p + "_z"], fx[p + "_q"], fx[p + "_rmh"],
                     fx[p + "_eps"], fx[p + "_hyd"])


def test_closed_forms_shape(fx):
    e, t = fx["closed_energy"], fx["closed_terms"]
    cs = fx["closed_cases"]
    # r = R_ij: LJ is -eps = -sqrt(0.25 * 0.16)
    assert t[0, 0] == pytest.approx(-0.2, abs=1e-7)
    # r^2 = 100 is kept (Coulomb 332 * 0.5 * -0.5 / 10), 10.5 A is not
    assert t[2, 1] == pytest.approx(-8.3, rel=1e-7) and e[3] == 0.0
    # r^2 = 0.25 kept (no charge: no Coulomb), r^2 < 0.25 dropped
    assert t[4, 1] == 0.0 and t[4, 0] > 0.0 and e[5] == 0.0
    # hydrophobic bonus only inside r^2 < 36
    assert t[6, 2] == 0.0 and t[7, 2] == -0.5 and t[1, 2] == -0.5
    assert np.all(e == t.sum(axis=1))
    assert cs.shape == (10, 5)


def test_oracle_matches_closed_forms(fx):
    ag, ab = closed_clouds(fx)
    ref_e, ref_fx, ref_t = fx["closed_energy"], fx["closed_fx"], fx["closed_terms"]
    for k in range(len(ref_e)):
        one = ab.subset(slice(k, k + 1))
        terms, (f_x, f_y, f_z), n = O.energy_f64(ag, one, 10.0)
        np.testing.assert_allclose(terms, ref_t[k], rtol=1e-9, atol=1e-12)
        np.testing.assert_allclose(f_x[0], ref_fx[k], rtol=1e-9, atol=1e-12)
        assert f_y[0] == 0.0 and f_z[0] == 0.0
        assert n == (1 if ref_e[k] != 0.0 or k == 4 else 0)


@pytest.mark.parametrize("name", ["c1", "c2"])
def test_oracle_reproduces_committed_vectors(fx, name):
    ag, ab = cloud(fx, name + "_ag"), cloud(fx, name + "_ab")
    off = fx[name + "_res_atom_off"][fx[name + "_cand_res_off"]].astype(np.uint32)
    r = O.population_f64(ag, ab, off, float(fx[name + "_cutoff"]), forces=True)
    for key in ("energy", "lj", "coulomb", "hydrophobic"):
        np.testing.assert_allclose(r[key], fx[f"{name}_{key}"], rtol=1e-12, atol=1e-12)
    assert np.array_equal(r["pairs"], fx[name + "_pairs"])
    for k, c in enumerate("xyz"):
        np.testing.assert_allclose(r["forces"][k], fx[f"{name}_f{c}"], rtol=1e-12, atol=1e-9)
    assert fx[name + "_pairs"].sum() > 0


@pytest.mark.parametrize("cell", [8, 10])
def test_oracle_grid_reproduces_committed_lists(fx, cell):
    ag = cloud(fx, "c1_ag")
    px, py, pz = fx["grid_px"], fx["grid_py"], fx["grid_pz"]
    off, idx = O.grid_query(ag.x, ag.y, ag.z, float(cell), px, py, pz)
    assert np.array_equal(off, fx[f"grid{cell}_offsets"])
    assert np.array_equal(idx, fx[f"grid{cell}_indices"])
    keys = np.array([[O.cell_of(float(v), float(cell)) for v in p] for p in zip(px, py, pz)])
    assert np.array_equal(keys, fx[f"grid{cell}_keys"])
