"""ctypes binding of libdnab.so (the C ABI declared in include/dnab.h).

The library is loaded from ``de-novo-antigen-binder_amd/lib/libdnab.so``
(built in-tree by ``make -C de-novo-antigen-binder_amd``). There is no
fallback: if the library is missing, importing this module raises.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

PKG_ROOT = Path(__file__).resolve().parent.parent
LIB_PATH = PKG_ROOT / "lib" / "libdnab.so"

DNAB_OK = 0
DNAB_ERR_INVALID = 1
DNAB_ERR_NO_DEVICE = 2
DNAB_ERR_HIP = 3
DNAB_ERR_ALLOC = 4
DNAB_ERR_CAPACITY = 5
DNAB_ERR_UNSUPPORTED = 6

DNAB_EVAL_FORCES = 1
DNAB_EVAL_BREAKDOWN = 2
DNAB_EVAL_FP64 = 4
DNAB_STATUS_NAN = 1
DNAB_STATUS_SATURATED = 2
DNAB_FORCE_ALLATOM = 0
DNAB_FORCE_CG = 1


class DnabError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"dnab status {status}: {message}")
        self.status = status


f32p = C.POINTER(C.c_float)
f64p = C.POINTER(C.c_double)
u8p = C.POINTER(C.c_uint8)
u32p = C.POINTER(C.c_uint32)
i32p = C.POINTER(C.c_int32)
u64p = C.POINTER(C.c_uint64)
szp = C.POINTER(C.c_size_t)


class GpuAtom(C.Structure):
    _fields_ = [(n, C.c_float) for n in
                ("x", "y", "z", "q", "r_min_half", "epsilon", "hydrophobic", "_pad")]


class AtomCloudC(C.Structure):
    _fields_ = [("n", C.c_size_t), ("x", f32p), ("y", f32p), ("z", f32p), ("charge", f32p),
                ("r_min_half", f32p), ("epsilon", f32p), ("hydrophobic", u8p),
                ("atom_type", u8p), ("residue_idx", u32p)]


class ResidueCloudC(C.Structure):
    _fields_ = [("n", C.c_size_t), ("x", f32p), ("y", f32p), ("z", f32p), ("charge", f32p),
                ("epsilon", f32p), ("sigma", f32p), ("hydrophobic", u8p)]


class EnergyBreakdownC(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("lj", "coulomb", "hydrophobic", "hbond", "solvation")]


class PairTuningC(C.Structure):
    _fields_ = [("direct_survivors", C.c_uint32), ("direct_chunks", C.c_uint32),
                ("rest_waves", C.c_uint32)]


class TrajTuningC(C.Structure):
    _fields_ = [("waves", C.c_uint32), ("cand_per_group", C.c_uint32), ("balance", C.c_uint32),
                ("cells_per_item", C.c_uint32), ("chunks_per_item", C.c_uint32),
                ("rebalance_steps", C.c_uint32), ("engine", C.c_uint32)]


class ProteinOutC(C.Structure):
    _fields_ = [("atom_capacity", C.c_size_t), ("x", f32p), ("y", f32p), ("z", f32p),
                ("charge", f32p), ("r_min_half", f32p), ("epsilon", f32p),
                ("hydrophobic", u8p), ("atom_type", u8p), ("residue_idx", u32p),
                ("res_start", u32p), ("ca_atom_idx", u32p), ("amino_acid", u8p)]


class AntigenDescC(C.Structure):
    _fields_ = [("atoms", AtomCloudC), ("ca", ResidueCloudC), ("center", C.c_float * 3),
                ("cell_size", C.c_float), ("cutoff", C.c_float),
                ("restraint_radius", C.c_float), ("ca_amino_acid", u8p)]


class PopulationDescC(C.Structure):
    _fields_ = [("n_cand", C.c_size_t), ("n_res", C.c_size_t), ("cand_antigen", u32p),
                ("cand_res_off", u32p), ("res_atom_off", u32p), ("res_ca_atom", u32p),
                ("res_amino_acid", u8p), ("atoms", AtomCloudC)]


DNAB_MC_SIDECHAIN, DNAB_MC_BACKBONE = 1, 2


class PdbChainC(C.Structure):
    _fields_ = [("x", f32p), ("y", f32p), ("z", f32p), ("res_start", u32p), ("amino_acid", u8p),
                ("n_res", C.c_size_t), ("chain_id", C.c_char)]


class CandidateReportC(C.Structure):
    _fields_ = [("rank", C.c_size_t), ("sequence", C.c_char_p), ("light_sequence", C.c_char_p),
                ("energy", C.c_float), ("entropy_penalty", C.c_float),
                ("dg_corrected", C.c_float), ("net_charge", C.c_int32),
                ("aggregation_risk", C.c_int), ("n_interface", C.c_size_t),
                ("clashes", C.c_int64), ("rama_outliers", C.c_int64)]


class McParamsC(C.Structure):
    _fields_ = [("mutation_p", C.c_float), ("rotamer_move_p", C.c_float),
                ("backbone_move_p", C.c_float), ("max_torsion_deg", C.c_float),
                ("iface_cutoff_sq", C.c_float), ("moves", C.c_uint32), ("waves", C.c_uint32)]


class RunAllatomParamsC(C.Structure):
    _fields_ = [("ab_length", C.c_uint32), ("population", C.c_uint32),
                ("cpu_iterations", C.c_uint32), ("top_n", C.c_uint32), ("n_broad", C.c_uint32),
                ("broad_steps", C.c_uint32), ("broad_moves", C.c_uint32),
                ("_reserved", C.c_uint32), ("seed", C.c_uint64), ("ring_radius", C.c_float),
                ("_pad", C.c_float)]


class LangevinParamsC(C.Structure):
    _fields_ = [("n_steps", C.c_uint32), ("score_every", C.c_uint32), ("top_k", C.c_uint32),
                ("seed", C.c_uint64), ("with_noise", C.c_int32),
                ("force_model", C.c_int32), ("use_graph", C.c_int32), ("step", C.c_float),
                ("max_disp", C.c_float), ("noise_base", C.c_float), ("t_hot", C.c_float),
                ("t_cold", C.c_float), ("schedule_offset", C.c_uint32),
                ("schedule_total", C.c_uint32), ("restraint_k", C.c_float),
                ("restraint_radius", C.c_float), ("count_pairs", C.c_int32),
                ("step_kernels", C.c_int32)]


class TrainParamsC(C.Structure):
    _fields_ = [("n_steps", C.c_uint32), ("batch", C.c_uint32), ("lr", C.c_float),
                ("beta1", C.c_float), ("beta2", C.c_float), ("eps", C.c_float),
                ("seed", C.c_uint64)]


FEATURE_DIM = 32
MAP_VALUES = 18  # per grid point: 7 LJ classes, potential, hydrophobic, 8 radial densities, cell occupancy
MLP_HIDDEN = 256

assert C.sizeof(GpuAtom) == 32


def _load() -> C.CDLL:
    if not LIB_PATH.exists():
        raise ImportError(
            f"{LIB_PATH} is not built; run `make -C {PKG_ROOT}` (or __graft_entry__.build())")
    return C.CDLL(str(LIB_PATH))


lib = _load()
vp = C.c_void_p
_PROTOS = {
    "dnab_ctx_create": (C.c_int, [C.c_int, C.POINTER(vp)]),
    "dnab_ctx_destroy": (None, [vp]),
    "dnab_ctx_last_error": (C.c_char_p, [vp]),
    "dnab_status_string": (C.c_char_p, [C.c_int]),
    "dnab_ctx_synchronize": (C.c_int, [vp]),
    "dnab_device_count": (C.c_int, [C.POINTER(C.c_int)]),
    "dnab_ctx_device_info": (C.c_int, [vp, C.POINTER(C.c_int), C.POINTER(C.c_int), C.c_char_p,
                                       C.c_size_t]),
    "dnab_ctx_set_timing": (C.c_int, [vp, C.c_int]),
    "dnab_ctx_kernel_time": (C.c_int, [vp, C.c_char_p, f64p, u64p]),
    "dnab_build_protein": (C.c_int, [f32p, f32p, f32p, u8p, u8p, C.c_size_t,
                                     C.POINTER(ProteinOutC), szp]),
    "dnab_place_atom": (C.c_int, [f32p, f32p, f32p, C.c_float, C.c_float, C.c_float, f32p]),
    "dnab_residue_atom_count": (C.c_int, [C.c_uint8]),
    "dnab_residue_cg_params": (C.c_int, [C.c_uint8, f32p, f32p, f32p, u8p]),
    "dnab_ff_table": (C.c_int, [C.c_uint8, C.c_int, u8p, f32p, u8p, f32p, f32p]),
    "dnab_ff_type_params": (C.c_int, [C.c_uint8, f32p, f32p, f32p, u8p, u8p]),
    "dnab_synth_globule_trace": (C.c_int, [C.c_size_t, C.c_float, C.c_uint64, f32p, f32p, f32p,
                                           u8p, u8p]),
    "dnab_synth_binder_ring": (C.c_int, [C.c_size_t, f32p, C.c_float, C.c_uint64, f32p, f32p,
                                         f32p, u8p]),
    "dnab_synth_binder_shell": (C.c_int, [C.c_size_t, f32p, C.c_float, C.c_float, C.c_uint64,
                                          f32p, f32p, f32p, u8p]),
    "dnab_population_get_tuning": (C.c_int, [vp, C.POINTER(PairTuningC)]),
    "dnab_population_set_tuning": (C.c_int, [vp, vp, C.POINTER(PairTuningC)]),
    "dnab_population_get_trajectory_tuning": (C.c_int, [vp, C.POINTER(TrajTuningC)]),
    "dnab_population_set_trajectory_tuning": (C.c_int, [vp, vp, C.POINTER(TrajTuningC)]),
    "dnab_population_autotune_trajectory": (C.c_int, [vp, vp, C.POINTER(LangevinParamsC),
                                                      C.c_uint32, C.POINTER(TrajTuningC)]),
    "dnab_population_autotune": (C.c_int, [vp, vp, C.c_uint32, C.POINTER(PairTuningC),
                                           C.POINTER(C.c_double)]),
    "dnab_synth_shell_population": (C.c_int, [C.c_size_t, C.c_size_t, f32p, C.c_float, C.c_float,
                                               C.c_uint64, C.c_int, C.POINTER(ProteinOutC), u32p,
                                               szp]),
    "dnab_grid_create": (C.c_int, [vp, C.c_float, C.POINTER(vp)]),
    "dnab_grid_build": (C.c_int, [vp, vp, f32p, f32p, f32p, C.c_size_t]),
    "dnab_grid_query": (C.c_int, [vp, vp, f32p, f32p, f32p, C.c_size_t, u32p, u32p, C.c_size_t,
                                  szp]),
    "dnab_grid_cell_of": (C.c_int, [vp, vp, f32p, f32p, f32p, C.c_size_t, i32p]),
    "dnab_grid_info": (C.c_int, [vp, szp, szp, i32p, i32p]),
    "dnab_grid_destroy": (None, [vp]),
    "dnab_cg_forces_with_grid": (C.c_int, [vp, C.POINTER(ResidueCloudC), vp,
                                           C.POINTER(ResidueCloudC), f32p, f32p, f32p]),
    "dnab_cg_forces": (C.c_int, [vp, C.POINTER(ResidueCloudC), C.POINTER(ResidueCloudC), f32p,
                                 f32p, f32p]),
    "dnab_interaction_energy_atoms_breakdown": (C.c_int, [vp, C.POINTER(AtomCloudC),
                                                          C.POINTER(AtomCloudC), C.c_float,
                                                          C.POINTER(EnergyBreakdownC)]),
    "dnab_interaction_energy_atoms": (C.c_int, [vp, C.POINTER(AtomCloudC), C.POINTER(AtomCloudC),
                                                C.c_float, f64p]),
    "dnab_atom_forces": (C.c_int, [vp, C.POINTER(AtomCloudC), C.POINTER(AtomCloudC), C.c_float,
                                   f32p, f32p, f32p, f64p]),
    "dnab_score_batch": (C.c_int, [vp, C.POINTER(AtomCloudC), C.POINTER(GpuAtom), C.c_size_t,
                                   C.c_size_t, f32p, C.c_size_t, szp]),
    "dnab_score_batch_noalloc": (C.c_int, [vp, vp, vp, C.c_size_t, C.c_size_t, vp, C.c_size_t,
                                           szp]),
    "dnab_antigen_create": (C.c_int, [vp, C.POINTER(AntigenDescC), C.POINTER(vp)]),
    "dnab_antigen_info": (C.c_int, [vp, szp, szp, f32p, f32p]),
    "dnab_antigen_destroy": (None, [vp]),
    "dnab_population_create": (C.c_int, [vp, C.POINTER(vp), C.c_size_t,
                                         C.POINTER(PopulationDescC), C.POINTER(vp)]),
    "dnab_population_destroy": (None, [vp]),
    "dnab_population_evaluate": (C.c_int, [vp, vp, C.c_uint]),
    "dnab_population_read_energies": (C.c_int, [vp, vp, f64p, f32p]),
    "dnab_population_read_status": (C.c_int, [vp, vp, u8p]),
    "dnab_population_read_breakdown": (C.c_int, [vp, vp, f64p, f64p, f64p]),
    "dnab_population_read_forces": (C.c_int, [vp, vp, f32p, f32p, f32p]),
    "dnab_population_read_forces_f64": (C.c_int, [vp, vp, f64p, f64p, f64p]),
    "dnab_population_read_coords": (C.c_int, [vp, vp, f32p, f32p, f32p]),
    "dnab_population_write_coords": (C.c_int, [vp, vp, f32p, f32p, f32p]),
    "dnab_population_pair_count": (C.c_int, [vp, vp, u64p, u64p]),
    "dnab_population_info": (C.c_int, [vp, szp, szp, szp]),
    "dnab_population_read_pair_stats": (C.c_int, [vp, vp, u64p, u64p]),
    "dnab_population_reset_pair_stats": (C.c_int, [vp, vp]),
    "dnab_population_features": (C.c_int, [vp, vp, C.POINTER(C.c_uint16)]),
    "dnab_prescreen_feature_count": (C.c_int, [szp]),
    "dnab_antigen_build_maps": (C.c_int, [vp, vp, C.c_float]),
    "dnab_antigen_map_info": (C.c_int, [vp, f32p, f32p, C.POINTER(C.c_int32), szp]),
    "dnab_antigen_map_points": (C.c_int, [vp, vp, C.POINTER(C.c_int32), C.c_size_t, f32p]),
    "dnab_population_map_energy": (C.c_int, [vp, vp, f32p]),
    "dnab_select_smallest": (C.c_int, [vp, f32p, C.c_uint32, C.c_uint32, u32p, u32p]),
    "dnab_mlp_create": (C.c_int, [vp, C.c_uint64, C.POINTER(vp)]),
    "dnab_mlp_destroy": (None, [vp]),
    "dnab_mlp_param_count": (C.c_int, [szp]),
    "dnab_mlp_get_params": (C.c_int, [vp, vp, f32p, C.c_size_t]),
    "dnab_mlp_set_params": (C.c_int, [vp, vp, f32p, C.c_size_t]),
    "dnab_mlp_loss_grad": (C.c_int, [vp, vp, C.POINTER(C.c_uint16), f32p, C.c_uint32, f32p,
                                     f32p, f32p]),
    "dnab_dataset_create": (C.c_int, [vp, C.c_size_t, C.c_float, C.c_float, C.POINTER(vp)]),
    "dnab_dataset_destroy": (None, [vp]),
    "dnab_dataset_add_population": (C.c_int, [vp, vp, vp]),
    "dnab_dataset_size": (C.c_int, [vp, szp]),
    "dnab_train_default_params": (C.c_int, [C.POINTER(TrainParamsC)]),
    "dnab_mlp_train": (C.c_int, [vp, vp, vp, C.POINTER(TrainParamsC), f32p]),
    "dnab_mlp_predict": (C.c_int, [vp, vp, vp, f32p]),
    "dnab_prescreen": (C.c_int, [vp, vp, vp, C.c_uint32, C.c_uint32, u32p, f64p, u32p]),
    "dnab_exact_topk": (C.c_int, [vp, vp, C.c_uint32, u32p, f64p, u32p]),
    "dnab_langevin_default_params": (C.c_int, [C.POINTER(LangevinParamsC)]),
    "dnab_langevin_step": (C.c_int, [vp, vp, C.POINTER(LangevinParamsC), C.c_uint32]),
    "dnab_langevin_run": (C.c_int, [vp, vp, C.POINTER(LangevinParamsC), u32p, f64p, u32p]),
    "dnab_population_read_scores": (C.c_int, [vp, vp, u32p, u32p, u32p, f64p]),
    "dnab_population_read_terms": (C.c_int, [vp, vp, f64p]),
    "dnab_allatom_energy": (C.c_int, [vp, C.POINTER(AtomCloudC), C.POINTER(AtomCloudC), u32p,
                                      u32p, u8p, C.c_size_t, C.c_float, f64p, f64p]),
    "dnab_mc_default_params": (C.c_int, [C.POINTER(McParamsC)]),
    "dnab_parse_pdb": (C.c_int, [C.c_char_p, C.c_size_t, f32p, f32p, f32p, u8p, C.c_size_t,
                                 C.POINTER(C.c_size_t)]),
    "dnab_write_pdb_ca": (C.c_int, [f32p, f32p, f32p, u8p, C.c_size_t, C.c_char, C.c_char_p,
                                    C.c_size_t, C.POINTER(C.c_size_t)]),
    "dnab_write_pdb_chains": (C.c_int, [C.POINTER(PdbChainC), C.c_size_t, C.c_int, C.c_char_p,
                                        C.c_size_t, C.POINTER(C.c_size_t)]),
    "dnab_write_report": (C.c_int, [C.POINTER(CandidateReportC), C.c_size_t, C.c_int, C.c_char_p,
                                    C.c_size_t, C.POINTER(C.c_size_t)]),
    "dnab_refine": (C.c_int, [vp, vp, C.POINTER(PopulationDescC), u8p, C.POINTER(LangevinParamsC),
                              C.POINTER(McParamsC), C.POINTER(ProteinOutC), u8p, f64p,
                              C.POINTER(C.c_size_t)]),
    "dnab_run_allatom_default_params": (C.c_int, [C.POINTER(RunAllatomParamsC)]),
    "dnab_run_allatom": (C.c_int, [vp, vp, C.POINTER(RunAllatomParamsC), C.POINTER(ProteinOutC),
                                   f64p, C.POINTER(C.c_uint32), C.POINTER(C.c_size_t)]),
    "dnab_rebuild_side_chain": (C.c_int, [C.c_uint8, C.c_uint8, f32p, C.c_size_t]),
    "dnab_affinity_maturation": (C.c_int, [vp, vp, C.POINTER(PopulationDescC), u8p, u32p, u32p,
                                           C.POINTER(ProteinOutC), u8p, u8p, f64p,
                                           C.POINTER(C.c_size_t)]),
}
for _name, (_res, _args) in _PROTOS.items():
    _fn = getattr(lib, _name)
    _fn.restype = _res
    _fn.argtypes = _args

EXPORTED = tuple(_PROTOS)


def ptr(a: np.ndarray | None, ctype):
    """Pointer to a contiguous numpy array (or NULL)."""
    if a is None:
        return C.cast(None, C.POINTER(ctype))
    assert a.flags["C_CONTIGUOUS"], "array must be C-contiguous"
    return a.ctypes.data_as(C.POINTER(ctype))


def check(status: int, ctx=None) -> None:
    if status != DNAB_OK:
        msg = lib.dnab_status_string(status).decode()
        if ctx is not None:
            detail = lib.dnab_ctx_last_error(ctx)
            if detail:
                msg += f" ({detail.decode()})"
        raise DnabError(status, msg)


def device_count() -> int:
    n = C.c_int(0)
    check(lib.dnab_device_count(C.byref(n)))
    return n.value


def visible_gpu() -> bool:
    """True when HIP reports a device (does not create a context)."""
    if os.environ.get("HIP_VISIBLE_DEVICES") == "-1":
        return False
    return device_count() > 0
