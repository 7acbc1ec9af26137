"""Precision study at the config-4 shape (a 5,000-residue globule of ~42k
heavy atoms, 1,024 binders of 30 residues, 12 A cutoff): the FP32 path
(per-chunk TwoSum, clash pairs in FP64, exact fixed-point sums across
chunks, FP64 totals) against the all-FP64 GPU kernel on the same pair set,
and the FP64 kernel against the CPU FP64 oracle on a few candidates.

Bars (the tolerance 1e-3 + 1e-4 |x64| in kcal/mol and kcal/mol/A): every
candidate energy and every atom-force component within it of the FP64
kernel; every candidate's a-priori error estimate at least its measured
error; the force RMS error at most 1e-5 of the force RMS; the FP64 kernel
within 1e-9 relative of the CPU oracle. The geometry has deep clashes (force
RMS ~1e8 kcal/mol/A): those pairs are the ones evaluated in FP64."""
import numpy as np
import pytest

from oracle import oracle as O
from dnab import precision_study
from dnab import workloads as W

pytestmark = pytest.mark.gpu


def test_precision_config4(ctx):
    w = W.config4(n_cand=1024)
    ags, pop = W.device_objects(ctx, w)
    r = precision_study(pop)
    print({k: v for k, v in r.items()})
    assert r["candidates"] == 1024
    assert r["energy_within_tol"] == 1.0
    assert r["force_within_tol"] == 1.0
    assert r["estimate_covers"] == 1.0
    assert r["force_rms_err"] <= 1e-5 * r["force_rms"]
    assert r["force_rms"] > 1e6 and r["energy_range"][0] < 0  # clashes and contacts both present
    assert (pop.status() == 0).all()
    # the FP64 kernel (left in the population by precision_study) vs the CPU oracle
    e64, _ = pop.energies()
    off = w.pop.cand_atom_off()
    for c in (0, 17, 95, 1023):
        cl = w.pop.candidate_cloud(c)
        ref = O.population_f64(w.antigens[0].atoms, cl, np.array([0, len(cl)], np.uint32),
                               w.cutoff, forces=False)["energy"][0]
        assert abs(e64[c] - ref) <= 1e-9 * (1.0 + abs(ref)), (c, e64[c], ref)
