"""Non-finite inputs and saturated sums, against the oracle.

A candidate atom with a NaN coordinate: the reference's skip test
`r_sq > CUTOFF_SQ || r_sq < MIN_R_SQ` (src/energy.rs:284) keeps a NaN
distance, so every pair of that atom is evaluated and the candidate's energy
is NaN; an infinite coordinate gives r_sq = +inf, which the test skips. The
oracle restates that predicate. The GPU result must be NaN for exactly the
candidates and force components where the oracle's is, equal to the oracle
elsewhere (tolerance 1e-3 + 1e-4 |x64|), and the status must say NaN.
Stepping a population with a non-finite coordinate is refused.

A saturated sum: 40 antigen S atoms stacked on one point and a binder S atom
0.5 A away (r^2 = 0.25, the closest pair the predicate keeps) give the
oracle a finite energy of ~6.9e11 kcal/mol, beyond the exact accumulator's
+-2^39; the GPU must report +inf with the status SATURATED, never a finite
number. A shallower clash next to it stays exact.
"""
import numpy as np
import pytest

import dnab
from oracle import oracle as O
from dnab import DeviceAntigen, DevicePopulation
from dnab import workloads as W
from dnab.structure import AtomCloud, PopulationHost

pytestmark = pytest.mark.gpu


def tol(x, x64):
    return np.abs(np.asarray(x, np.float64) - x64) <= 1e-3 + 1e-4 * np.abs(x64)


def with_coordinate(host: PopulationHost, atom: int, axis: int, value: float) -> PopulationHost:
    a = host.atoms
    xyz = [a.x.copy(), a.y.copy(), a.z.copy()]
    xyz[axis][atom] = value
    atoms = AtomCloud(xyz[0], xyz[1], xyz[2], a.charge, a.r_min_half, a.epsilon, a.hydrophobic,
                      a.atom_type, a.residue_idx)
    return PopulationHost(host.cand_antigen, host.cand_res_off, host.res_atom_off,
                          host.res_ca_atom, host.res_amino_acid, atoms)


def test_nan_coordinate_scores_nan(ctx):
    w = W.config2(n_cand=8)
    ags = [DeviceAntigen(ctx, w.antigens[0], cutoff=w.cutoff)]
    off = w.pop.cand_atom_off()
    bad_c = 3
    host = with_coordinate(w.pop, int(off[bad_c]) + 2, 0, np.nan)
    pop = DevicePopulation(ctx, ags, host)
    ref = O.population_f64(w.antigens[0].atoms, host.atoms, off, w.cutoff)
    assert np.isnan(ref["energy"][bad_c]) and np.isfinite(np.delete(ref["energy"], bad_c)).all()

    pop.evaluate(forces=True)
    e, _ = pop.energies()
    assert (np.isnan(e) == np.isnan(ref["energy"])).all()
    fin = ~np.isnan(ref["energy"])
    assert tol(e[fin], ref["energy"][fin]).all()
    st = pop.status()
    assert st[bad_c] == dnab.STATUS_NAN and (st[fin] == 0).all()
    for a, b in zip(pop.forces(), ref["forces"]):
        assert (np.isnan(a) == np.isnan(b)).all()
        m = ~np.isnan(b)
        assert tol(a[m], b[m]).all()

    # the energy-only, breakdown and FP64 evaluations carry it too
    for kw in (dict(forces=False), dict(forces=False, breakdown=True), dict(fp64=True)):
        pop.evaluate(**kw)
        e2, _ = pop.energies()
        assert np.isnan(e2[bad_c]) and np.isfinite(e2[fin]).all(), kw
        assert pop.status()[bad_c] == dnab.STATUS_NAN

    # stepping needs finite coordinates
    with pytest.raises(Exception):
        pop.run(dnab.langevin_params(n_steps=2, score_every=1, top_k=2))
    # finite again: the population steps
    pop.set_coords(w.pop.atoms.x, w.pop.atoms.y, w.pop.atoms.z)
    pop.evaluate(forces=True)
    assert (pop.status() == 0).all()
    idx, et = pop.run(dnab.langevin_params(n_steps=2, score_every=1, top_k=2))
    assert len(idx) == 2 and np.isfinite(et).all()


def test_infinite_coordinate_has_no_pair(ctx):
    # r^2 = +inf: the reference's skip test drops every pair of the atom
    w = W.config2(n_cand=8)
    ags = [DeviceAntigen(ctx, w.antigens[0], cutoff=w.cutoff)]
    off = w.pop.cand_atom_off()
    bad = int(off[3]) + 2
    host = with_coordinate(w.pop, bad, 1, -np.inf)
    pop = DevicePopulation(ctx, ags, host)
    ref = O.population_f64(w.antigens[0].atoms, host.atoms, off, w.cutoff)
    assert np.isfinite(ref["energy"]).all() and all(f[bad] == 0.0 for f in ref["forces"])
    pop.evaluate(forces=True)
    e, _ = pop.energies()
    assert tol(e, ref["energy"]).all() and (pop.status() == 0).all()
    for a, b in zip(pop.forces(), ref["forces"]):
        assert tol(a, b).all() and a[bad] == 0.0
    with pytest.raises(Exception):
        pop.run(dnab.langevin_params(n_steps=2, score_every=1, top_k=2))


def stack(n, r=0.0):
    z = np.zeros(n, np.float32)
    return AtomCloud(z + np.float32(r), z.copy(), z.copy(), np.zeros(n, np.float32),
                     np.full(n, 2.0, np.float32), np.full(n, 0.25, np.float32),
                     np.zeros(n, np.uint8))


def test_saturated_sum_is_reported(ctx):
    deep, shallow = stack(40), stack(5)
    ags = [DeviceAntigen(ctx, deep, cutoff=10.0), DeviceAntigen(ctx, shallow, cutoff=10.0)]
    binder = AtomCloud(np.array([0.5, 1.0], np.float32), np.zeros(2, np.float32),
                       np.zeros(2, np.float32), np.zeros(2, np.float32),
                       np.full(2, 2.0, np.float32), np.full(2, 0.25, np.float32),
                       np.zeros(2, np.uint8))
    one = np.array([0, 1], np.uint32)
    host = PopulationHost(np.array([0, 1], np.uint32), np.array([0, 1, 2], np.uint32),
                          np.array([0, 1, 2], np.uint32), np.array([0, 1], np.uint32),
                          np.zeros(2, np.uint8), binder)
    pop = DevicePopulation(ctx, ags, host)
    r0 = O.population_f64(deep, binder.subset(slice(0, 1)), one, 10.0)
    r1 = O.population_f64(shallow, binder.subset(slice(1, 2)), one, 10.0)
    lim = 2.0 ** 39
    assert lim < r0["energy"][0] < 1e13 and r0["forces"][0][0] > lim
    assert r1["energy"][0] < lim and abs(r1["forces"][0][0]) < lim

    pop.evaluate(forces=True)
    e, _ = pop.energies()
    fx, fy, fz = pop.forces()
    assert e[0] == np.inf and fx[0] == np.inf and fy[0] == 0.0 and fz[0] == 0.0
    assert tol(e[1], r1["energy"][0]) and tol(fx[1], r1["forces"][0][0])
    st = pop.status()
    assert st[0] == dnab.STATUS_SATURATED and st[1] == 0
    pop.evaluate(forces=False)
    e, _ = pop.energies()
    assert e[0] == np.inf and tol(e[1], r1["energy"][0])
    assert pop.status()[0] == dnab.STATUS_SATURATED

    # the FP64 kernel holds the value
    pop.evaluate(fp64=True)
    e64, _ = pop.energies()
    assert abs(e64[0] - r0["energy"][0]) <= 1e-9 * abs(r0["energy"][0])
    assert (pop.status() == 0).all()
