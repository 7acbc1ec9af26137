"""run_allatom (src/diffusion.rs:983-1107) as one call, dnab_run_allatom.

The reference's contract: the result holds max(top_n, 1) candidates, each
an antibody of ab_length residues with its sequence, sorted by ascending
allatom_energy (src/diffusion.rs:1097-1106). Each reported energy must be
the candidate's allatom_energy: within 1e-3 + 1e-4 |E64| of the FP64 oracle
(the five pair terms plus the disulfide wells) on the returned coordinates.
Trajectories use the library's own seeded streams (the reference's SmallRng
is not reproduced), so a run must repeat exactly and differ with the seed.
Config 1's shape: a 16-residue peptide antigen, 64 binders of 12 residues,
600 refinement iterations, no broad phase (the reference's path without a
GPU).
"""
import numpy as np
import pytest

import dnab
from oracle import oracle as O
from dnab import DeviceAntigen
from dnab import workloads as W

pytestmark = pytest.mark.gpu


def allatom_energy_oracle(ag, prot, cutoff):
    t, _ = O.energy_full_f64(ag.atoms, prot.atoms, cutoff)
    sg = O.sg_positions(prot.atoms.x, prot.atoms.y, prot.atoms.z, prot.res_start,
                        prot.amino_acid, 0, len(prot.amino_acid))
    return ((((t[0] + t[1]) + t[2]) + t[3]) + t[4]) + O.disulfide_f64(*sg)


def check_contract(ag, res, n_expect, ab_length, cutoff=10.0):
    assert len(res) == n_expect
    e = np.array([r[1] for r in res])
    assert np.isfinite(e).all() and (np.diff(e) >= 0).all()
    for prot, energy, seq in res:
        assert prot.n_residues() == ab_length and len(seq) == ab_length
        assert seq == prot.sequence()
        ref = allatom_energy_oracle(ag, prot, cutoff)
        assert abs(energy - ref) <= 1e-3 + 1e-4 * abs(ref), (energy, ref)


def test_run_allatom_config1(ctx):
    seq = W.random_sequence(16, 1001)
    ag = W.peptide_antigen(seq)
    dag = DeviceAntigen(ctx, ag, cutoff=10.0, cell_size=8.0)
    res = dnab.run_allatom(ctx, dag, ab_length=12, population=64, cpu_iterations=600, top_n=10,
                           n_broad=0, seed=11)
    check_contract(ag, res, 10, 12)
    # the whole population ranked; its first ten are the same run's, repeated exactly
    full = dnab.run_allatom(ctx, dag, ab_length=12, population=64, cpu_iterations=600,
                            top_n=64, n_broad=0, seed=11)
    assert len(full) == 64
    assert [r[1] for r in full[:10]] == [r[1] for r in res]
    # top_n = 0 keeps one (max(top_n, 1)); a different seed gives another run
    one = dnab.run_allatom(ctx, dag, ab_length=12, population=64, cpu_iterations=600, top_n=0,
                           n_broad=0, seed=12)
    check_contract(ag, one, 1, 12)
    assert one[0][1] != res[0][1]


@pytest.mark.parametrize("moves", [1, 0])
def test_run_allatom_broad_phase(ctx, moves):
    # a broad phase of 256 binders (with the moves, as the reference steps it,
    # or the gradient-only loop), the 16 lowest refined; the antigen's length
    # is the default binder length
    ag = W.peptide_antigen(W.random_sequence(10, 77))
    dag = DeviceAntigen(ctx, ag, cutoff=10.0, cell_size=10.0)
    kw = dict(population=16, cpu_iterations=50, top_n=5, n_broad=256, broad_steps=20,
              broad_moves=moves, seed=5)
    res = dnab.run_allatom(ctx, dag, **kw)
    check_contract(ag, res, 5, 10)
    again = dnab.run_allatom(ctx, dag, **kw)
    assert [r[1] for r in again] == [r[1] for r in res]
    for (p, _, _), (q, _, _) in zip(res, again):
        assert np.array_equal(p.atoms.x, q.atoms.x) and np.array_equal(p.atoms.z, q.atoms.z)
