"""GPU gradient step and population loop vs the oracle's restatement of
src/diffusion.rs:300-339: coordinates bit-identical to the oracle given the
same forces (all-atom model) or the same residue-level grid forces (the
reference's own step), with the noise on and off; the loop's top-k is
sorted and consistent with a fresh evaluation; graph replay == eager."""
import numpy as np
import pytest

from oracle import oracle as O
from dnab import langevin_params
from dnab import _native as N
from dnab import workloads as W

pytestmark = pytest.mark.gpu


def oracle_params(p, iteration, center, radius):
    return dict(step=p.step, max_disp=p.max_disp, restraint_k=p.restraint_k,
                restraint_radius=radius, with_noise=p.with_noise, force_model=p.force_model,
                iter=iteration, seed_lo=p.seed & 0xFFFFFFFF, seed_hi=p.seed >> 32,
                sigma=O.sigma(iteration + p.schedule_offset, p.schedule_total, p.t_hot,
                              p.t_cold, p.noise_base),
                center=center)


@pytest.mark.parametrize("noise", [0, 1])
@pytest.mark.parametrize("model", [N.DNAB_FORCE_ALLATOM, N.DNAB_FORCE_CG])
def test_step_bit_identical(ctx, noise, model):
    w = W.config1(n_cand=16)
    ags, pop = W.device_objects(ctx, w)
    p = langevin_params(with_noise=noise, force_model=model, seed=0x1234_5678_9abc, restraint_radius=20.0)
    x0, y0, z0 = pop.coords()
    it = 7
    pop.step(p, it)
    x1, y1, z1 = pop.coords()
    x, y, z = x0.copy(), y0.copy(), z0.copy()
    host = w.pop
    kw = {}
    if model == N.DNAB_FORCE_ALLATOM:
        kw["forces"] = pop.forces()
    else:
        from dnab.structure import ResidueCloud
        rc = ResidueCloud.from_points(np.zeros(host.n_res), np.zeros(host.n_res),
                                      np.zeros(host.n_res), host.res_amino_acid)
        kw["ag_ca"] = w.antigens[0].ca_residue_cloud()
        kw["res_cg"] = (rc.charge, rc.epsilon, rc.sigma, rc.hydrophobic)
    O.langevin_step(x, y, z, host.res_atom_off, host.res_ca_atom,
                    oracle_params(p, it, ags[0].center, 20.0), **kw)
    for a, b in zip((x1, y1, z1), (x, y, z)):
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    assert not np.array_equal(x0, x1)


def test_run_topk_consistent(ctx):
    w = W.config2(n_cand=200)
    ags, pop = W.device_objects(ctx, w)
    p = langevin_params(n_steps=20, score_every=10, top_k=16, restraint_radius=w.restraint_radius[0])
    idx, e = pop.run(p)
    assert len(idx) == 16 and len(set(idx.tolist())) == 16
    assert (np.diff(e) >= 0).all()
    pop.evaluate()
    ef, _ = pop.energies()
    assert np.array_equal(ef[idx], e)
    rest = np.setdiff1d(np.arange(w.pop.n_cand), idx)
    assert ef[rest].min() >= e[-1]
    # energies decrease on average under the gradient step
    ags2, pop2 = W.device_objects(ctx, w)
    pop2.evaluate()
    e0, _ = pop2.energies()
    assert np.median(ef) < np.median(e0)
    # the score history: scorings at steps 0 and 10, then the final state;
    # row 0 is the top 16 of the initial state (ascending energy, ties by index)
    hi, he = pop.scores()
    assert hi.shape == (3, 16)
    order = np.lexsort((np.arange(len(e0)), e0))[:16]
    assert np.array_equal(hi[0], order)
    assert np.array_equal(he[0].view(np.uint64), e0[order].view(np.uint64))
    assert np.array_equal(hi[-1], idx)


def _run_state(ctx, w, tuning=None, **kw):
    ags, pop = W.device_objects(ctx, w)
    if tuning:
        pop.set_trajectory_tuning(*tuning)
    p = langevin_params(restraint_radius=w.restraint_radius[0], **kw)
    idx, e = pop.run(p)
    ef, err = pop.energies()
    hi, he = pop.scores()
    return dict(idx=idx, e=e, coords=pop.coords(), forces=pop.forces(), energy=ef, err=err,
                hist_idx=hi, hist_e=he,
                stats=pop.pair_stats() if kw.get("count_pairs") else None)


def _assert_same(a, b):
    assert np.array_equal(a["idx"], b["idx"])
    assert np.array_equal(a["hist_idx"], b["hist_idx"])
    assert np.array_equal(a["hist_e"].view(np.uint64), b["hist_e"].view(np.uint64))
    assert np.array_equal(a["e"].view(np.uint64), b["e"].view(np.uint64))
    assert np.array_equal(a["energy"].view(np.uint64), b["energy"].view(np.uint64))
    assert np.array_equal(a["err"].view(np.uint32), b["err"].view(np.uint32))
    for u, v in zip(a["coords"], b["coords"]):
        assert np.array_equal(u.view(np.uint32), v.view(np.uint32))
    for u, v in zip(a["forces"], b["forces"]):
        assert np.array_equal(np.asarray(u).view(np.uint32), np.asarray(v).view(np.uint32))


def test_force_only_steps_keep_the_trajectory(ctx):
    # the fused loop's steps between scorings form forces only; scoring every
    # step (every step forms energies) gives the same trajectory bit for bit,
    # and the scorings both runs make give the same rows
    w = W.config2(n_cand=96)
    kw = dict(n_steps=20, top_k=8, with_noise=1, seed=5)
    a = _run_state(ctx, w, score_every=10, **kw)
    b = _run_state(ctx, w, score_every=1, **kw)
    for u, v in zip(a["coords"], b["coords"]):
        assert np.array_equal(u.view(np.uint32), v.view(np.uint32))
    for u, v in zip(a["forces"], b["forces"]):
        assert np.array_equal(np.asarray(u).view(np.uint32), np.asarray(v).view(np.uint32))
    assert np.array_equal(a["energy"].view(np.uint64), b["energy"].view(np.uint64))
    assert a["hist_e"].shape[0] == 3 and b["hist_e"].shape[0] == 21
    for ra, rb in ((0, 0), (1, 10), (2, 20)):
        assert np.array_equal(a["hist_idx"][ra], b["hist_idx"][rb])
        assert np.array_equal(a["hist_e"][ra].view(np.uint64), b["hist_e"][rb].view(np.uint64))


def test_graph_replay_equals_eager(ctx):
    # per-step kernels eager, per-step kernels replayed from a graph, and the
    # fused trajectory kernel: the same trajectory bit for bit
    w = W.config2(n_cand=128)
    kw = dict(n_steps=30, score_every=10, top_k=8, with_noise=1, seed=5)
    eager = _run_state(ctx, w, use_graph=0, step_kernels=1, **kw)
    graph = _run_state(ctx, w, use_graph=1, step_kernels=1, **kw)
    fused = _run_state(ctx, w, step_kernels=0, **kw)
    _assert_same(eager, graph)
    _assert_same(eager, fused)
    # scorings at steps 0, 10, 20 and the final state; the last row is the result
    assert fused["hist_idx"].shape == (4, 8)
    assert np.array_equal(fused["hist_idx"][-1], fused["idx"])
    assert (np.diff(fused["hist_e"], axis=1) >= 0).all()


def test_trajectory_tuning_changes_speed_not_results(ctx):
    # ragged multi-antigen population: groups mix antigens and binder lengths
    w = W.config3(n_cand=80)
    kw = dict(n_steps=12, score_every=4, top_k=8, with_noise=1, seed=9)
    ref = None
    # (rebalance 5: the 12 steps run as launches of 5, 5 and 2 steps, regrouped between them)
    # (balance 2: groups by position, atoms bisected into compact tiles; with
    # rebalance 5 regrouped by the new positions between launches)
    for waves, cpw, bal, cells, chunks, reb in ((4, 1, 0, 16, 4, 0), (8, 3, 0, 5, 1, 0),
                                                (16, 4, 0, 64, 3, 0), (16, 16, 0, 1, 2, 0),
                                                (16, 6, 1, 16, 2, 0), (4, 2, 1, 3, 1, 5),
                                                (16, 4, 1, 2, 2, 5), (16, 16, 2, 2, 2, 0),
                                                (8, 5, 2, 1, 4, 0), (4, 3, 2, 8, 1, 5)):
        ags, pop = W.device_objects(ctx, w)
        pop.set_trajectory_tuning(waves, cpw, bal, cells, chunks, reb)
        assert pop.trajectory_tuning() == {"waves": waves, "cand_per_group": cpw, "balance": bal,
                                           "cells_per_item": cells, "chunks_per_item": chunks,
                                           "rebalance_steps": reb, "engine": 0}
        p = langevin_params(restraint_radius=w.restraint_radius[0], **kw)
        idx, e = pop.run(p)
        hi, he = pop.scores()
        st = dict(idx=idx, e=e, coords=pop.coords(), forces=pop.forces(), energy=pop.energies()[0],
                  err=pop.energies()[1], hist_idx=hi, hist_e=he)
        if ref is None:
            ref = st
        else:
            _assert_same(ref, st)
    with pytest.raises(RuntimeError):
        pop.set_trajectory_tuning(12, 4)
    with pytest.raises(RuntimeError):
        pop.set_trajectory_tuning(16, 4, 3)
    with pytest.raises(RuntimeError):
        pop.set_trajectory_tuning(16, 4, 1, 2, 2, 0, 2)


def test_trajectory_autotune_keeps_state_and_results(ctx):
    # the tuner runs trial loops from the population's state and puts the
    # coordinates back; the chosen shape is a valid one and the run after it
    # equals a run with the default shape
    w = W.config3(n_cand=80)
    kw = dict(n_steps=12, score_every=4, top_k=8, with_noise=1, seed=9)
    ags, pop = W.device_objects(ctx, w)
    p = langevin_params(restraint_radius=w.restraint_radius[0], **kw)

    def state(pop):
        idx, e = pop.run(p)
        hi, he = pop.scores()
        return dict(idx=idx, e=e, coords=pop.coords(), forces=pop.forces(),
                    energy=pop.energies()[0], err=pop.energies()[1], hist_idx=hi, hist_e=he)

    ref = state(pop)
    ags, pop = W.device_objects(ctx, w)
    x0 = pop.coords()
    t = pop.autotune_trajectory(p, trial_steps=6)
    assert t["waves"] in (4, 8, 16) and t["cand_per_group"] in (1, 2, 4, 8)
    assert t["chunks_per_item"] in (1, 2, 4) and t["cells_per_item"] in (2, 8)
    assert t["balance"] in (0, 1, 2) and t["rebalance_steps"] == 0
    assert t["engine"] in (0, 1)
    for a, b in zip(x0, pop.coords()):
        assert np.array_equal(a, b)
    with pytest.raises(RuntimeError):
        pop.forces()  # no evaluation is current after tuning
    _assert_same(ref, state(pop))
    with pytest.raises(RuntimeError):
        pop.autotune_trajectory(p, trial_steps=0)


@pytest.mark.parametrize("config", ["config2", "config3", "config2_fine_cells", "config2_packed",
                                    "config3_packed"])
def test_fused_trajectory_equals_step_kernels(ctx, config):
    # no noise (the reference's GPU phase), pair counting on: the in-cutoff
    # pair count is a property of the trajectory, the tested count is not.
    # With 4 A antigen cells a tile's box spans hundreds of cells, more chunks
    # than a tile's chunk list holds: those tiles go through cell items.
    if config == "config2_fine_cells":
        import dataclasses
        w = dataclasses.replace(W.config2(n_cand=96), cell_size=4.0)
    else:
        w = getattr(W, config.replace("_packed", ""))(n_cand=96)
    # packed: groups by position, each group's atoms in compact tiles
    tuning = (16, 16, 2, 2, 2, 0) if config.endswith("_packed") else None
    kw = dict(n_steps=40, score_every=0, top_k=8, count_pairs=1)
    step = _run_state(ctx, w, step_kernels=1, **kw)
    fused = _run_state(ctx, w, step_kernels=0, tuning=tuning, **kw)
    _assert_same(step, fused)
    assert step["stats"][0] == fused["stats"][0] > 0
    print(f"{config}: {fused['stats'][0]} pairs in cutoff over 41 evaluations; tested "
          f"{step['stats'][1]} (step kernels) vs {fused['stats'][1]} (fused)")


def test_multi_antigen_ragged_run(ctx):
    w = W.config3(n_cand=64)
    ags, pop = W.device_objects(ctx, w)
    pop.evaluate()
    e, _ = pop.energies()
    off = w.pop.cand_atom_off()
    for a in range(len(w.antigens)):
        sel = np.nonzero(w.pop.cand_antigen == a)[0]
        for c in sel[:2]:
            cl = w.pop.candidate_cloud(int(c))
            r = O.population_f64(w.antigens[a].atoms, cl, np.array([0, len(cl)], np.uint32), 12.0,
                                 forces=False)
            assert abs(e[c] - r["energy"][0]) <= 1e-3 + 1e-4 * abs(r["energy"][0])
    p = langevin_params(n_steps=10, score_every=5, top_k=8)
    idx, ek = pop.run(p)
    assert len(idx) == 8 and (np.diff(ek) >= 0).all()


def test_fused_loop_relists_as_tiles_move(ctx):
    # large clamped steps move the tile boxes out of their listing boxes
    # (the box widened by 1 A) most steps, so the chunk lists are made again
    # and again; the loop still ends in the per-step path's bits
    w = W.config2(n_cand=64)
    kw = dict(n_steps=12, score_every=4, top_k=8, with_noise=1, seed=21, step=0.5, max_disp=3.0)
    fused = _run_state(ctx, w, step_kernels=0, **kw)
    steps = _run_state(ctx, w, step_kernels=1, use_graph=0, **kw)
    _assert_same(fused, steps)
    start = (w.pop.atoms.x, w.pop.atoms.y, w.pop.atoms.z)
    moved = [np.abs(a - b).max() for a, b in zip(fused["coords"], start)]
    assert max(moved) > 2.0


def test_fused_loop_with_an_empty_antigen(ctx):
    # an antigen without atoms: no cell lists a chunk and every pair sum is
    # zero; the binders still move under the restraint, both engines agree
    from dnab.structure import AtomCloud
    from dnab import DeviceAntigen, DevicePopulation
    w = W.config2(n_cand=8)
    z = np.zeros(0, np.float32)
    empty = AtomCloud(z, z, z, z, z, z, np.zeros(0, np.uint8))
    ctr = w.antigens[0].ca_center_of_mass()
    out = []
    for step_kernels in (0, 1):
        ag = DeviceAntigen(ctx, empty, cutoff=10.0, center=ctr)
        pop = DevicePopulation(ctx, [ag], w.pop)
        p = langevin_params(n_steps=6, score_every=3, top_k=4, restraint_radius=5.0,
                            step_kernels=step_kernels, use_graph=0)
        idx, e = pop.run(p)
        assert (e == 0).all()
        assert (pop.energies()[0] == 0).all()
        out.append(pop.coords())
    for a, b in zip(*out):
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    start = (w.pop.atoms.x, w.pop.atoms.y, w.pop.atoms.z)
    assert any(not np.array_equal(a, b) for a, b in zip(out[0], start))


def test_engine_choice_changes_speed_not_results(ctx):
    # step_kernels = 2 runs the engine the trajectory tuning names; either
    # engine ends in the same bits
    w = W.config2(n_cand=64)
    kw = dict(n_steps=10, score_every=5, top_k=8, with_noise=1, seed=13)
    ref = _run_state(ctx, w, step_kernels=0, **kw)
    for engine in (0, 1):
        ags, pop = W.device_objects(ctx, w)
        pop.set_trajectory_tuning(16, 4, 1, 2, 2, 0, engine)
        assert pop.trajectory_tuning()["engine"] == engine
        p = langevin_params(restraint_radius=w.restraint_radius[0], step_kernels=2, use_graph=1, **kw)
        idx, e = pop.run(p)
        hi, he = pop.scores()
        st = dict(idx=idx, e=e, coords=pop.coords(), forces=pop.forces(), energy=pop.energies()[0],
                  err=pop.energies()[1], hist_idx=hi, hist_e=he)
        _assert_same(ref, st)
    with pytest.raises(RuntimeError):
        pop.run(langevin_params(restraint_radius=w.restraint_radius[0], step_kernels=3, **kw))
