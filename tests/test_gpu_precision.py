"""Precision study at the config-4 shape (a 5,000-residue globule of ~41.7k
heavy atoms, 30-residue binders, 12 A cutoff): the FP32 path (per-chunk
TwoSum, exact fixed-point sums across chunks, FP64 totals) against the
all-FP64 GPU kernel on the same pair set, and the FP64 kernel against the CPU
FP64 oracle on a few candidates.

Bars (written here): every candidate energy within 1e-3 + 1e-4 |E64| of the
FP64 kernel; at least 99.99 % of atom-force components within 1e-3 + 1e-4
|F64| and the force RMS error at most 1e-5 of the force RMS; the FP64 kernel
within 1e-9 relative of the CPU oracle. The force bar is a fraction because
this geometry has deep clashes (force RMS ~1e8 kcal/mol/A): a component that
is the near-cancellation of pair forces of that size keeps the FP32 pair
terms' own rounding (~1e-7 of each term), whatever the accumulation."""
import numpy as np
import pytest

from oracle import oracle as O
from dnab import precision_study
from dnab import workloads as W

pytestmark = pytest.mark.gpu


def test_precision_config4(ctx):
    w = W.config4(n_cand=96)
    ags, pop = W.device_objects(ctx, w)
    r = precision_study(pop)
    print({k: v for k, v in r.items()})
    assert r["energy_within_tol"] == 1.0
    assert r["force_within_tol"] >= 0.9999
    assert r["force_rms_err"] <= 1e-5 * r["force_rms"]
    assert r["force_rms"] > 0 and r["energy_range"][0] < 0
    # the FP64 kernel (left in the population by precision_study) vs the CPU oracle
    e64, _ = pop.energies()
    off = w.pop.cand_atom_off()
    for c in (0, 17, 95):
        cl = w.pop.candidate_cloud(c)
        ref = O.population_f64(w.antigens[0].atoms, cl, np.array([0, len(cl)], np.uint32),
                               w.cutoff, forces=False)["energy"][0]
        assert abs(e64[c] - ref) <= 1e-9 * (1.0 + abs(ref)), (c, e64[c], ref)
