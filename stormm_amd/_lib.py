"""ctypes binding of libstormm_amd.so, the C-ABI declared in include/stormm_amd.h.

Loading fails loudly when the library has not been built: there is no fallback path.
"""
from __future__ import annotations

import ctypes
import os
import re

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libstormm_amd.so")
HEADER_PATH = os.path.join(os.path.dirname(_HERE), "include", "stormm_amd.h")

c_int32 = ctypes.c_int32
c_int64 = ctypes.c_int64
c_double = ctypes.c_double
vp = ctypes.c_void_p
P = ctypes.POINTER

STORMM_OK = 0
STORMM_ERR_INVALID_ARGUMENT = -1
STORMM_ERR_HIP = -2
STORMM_ERR_CAPACITY = -3
STORMM_ERR_STATE = -4
PREC_DOUBLE = 0
PREC_SINGLE = 1


class StormmError(RuntimeError):
    """A C-ABI call returned an error status (the reference raises std::runtime_error)."""

    def __init__(self, code: int, message: str):
        super().__init__(f"stormm_amd error {code}: {message}")
        self.code = code


class TopologyBatch(ctypes.Structure):
    _fields_ = [
        ("system_count", c_int32), ("atom_counts", vp), ("charge", vp), ("lj_type", vp),
        ("lj_type_counts", vp), ("lja", vp), ("ljb", vp), ("lja14", vp), ("ljb14", vp),
        ("pb_radii", vp), ("gb_screen", vp), ("bond_counts", vp), ("bond_atoms", vp),
        ("bond_k", vp), ("bond_r0", vp), ("angle_counts", vp), ("angle_atoms", vp),
        ("angle_k", vp), ("angle_theta", vp), ("dihedral_counts", vp), ("dihedral_atoms", vp),
        ("dihe_amp", vp), ("dihe_phase", vp), ("dihe_freq", vp), ("dihe_improper", vp),
        ("dihe_scee", vp), ("dihe_scnb", vp), ("exclusion_counts", vp), ("exclusion_pairs", vp),
        ("igb", c_int32), ("dielectric", c_double), ("salt_concentration", c_double),
        ("coulomb_constant", c_double), ("lj_combining", vp), ("lj_sigma", vp),
        ("lj_epsilon", vp), ("restraint_counts", vp), ("restraints", vp),
    ]


class LaunchShape(ctypes.Structure):
    _fields_ = [("wu_atom_capacity", c_int32), ("target_wu_per_cu", c_int32),
                ("block_threads", c_int32)]


class SynthesisInfo(ctypes.Structure):
    _fields_ = [("system_count", c_int32), ("padded_atom_count", c_int32),
                ("work_unit_count", c_int32), ("task_count", c_int32),
                ("max_wu_atoms", c_int32), ("lj_table_systems", c_int32),
                ("pair_count", c_int64), ("valence_term_count", c_int64)]


class PsWriter(ctypes.Structure):
    _fields_ = [("system_count", c_int32), ("gpos_bits", c_int32), ("frc_bits", c_int32),
                ("gpos_scale", c_double), ("inv_gpos_scale", c_double),
                ("frc_scale", c_double), ("inv_frc_scale", c_double),
                ("atom_starts", vp), ("atom_counts", vp),
                ("xcrd", vp), ("ycrd", vp), ("zcrd", vp), ("xfrc", vp), ("yfrc", vp),
                ("zfrc", vp), ("xprv", vp), ("yprv", vp), ("zprv", vp),
                ("x_cg_temp", vp), ("y_cg_temp", vp), ("z_cg_temp", vp),
                ("xcrd_ovrf", vp), ("ycrd_ovrf", vp), ("zcrd_ovrf", vp),
                ("xfrc_ovrf", vp), ("yfrc_ovrf", vp), ("zfrc_ovrf", vp),
                ("xprv_ovrf", vp), ("yprv_ovrf", vp), ("zprv_ovrf", vp),
                ("x_cg_temp_ovrf", vp), ("y_cg_temp_ovrf", vp), ("z_cg_temp_ovrf", vp)]


class ScoreCardWriter(ctypes.Structure):
    _fields_ = [("system_count", c_int32), ("data_stride", c_int32), ("nrg_bits", c_int32),
                ("instantaneous_accumulators", vp), ("sampled_step_count", c_int32),
                ("sample_capacity", c_int32), ("running_accumulators", vp),
                ("squared_accumulators", vp), ("time_series", vp)]


class LinMinWriter(ctypes.Structure):
    _fields_ = [("system_count", c_int32)] + [
        (n, vp) for n in ("l_move", "s_move", "mfac_a", "mfac_b", "mfac_c", "nrg_a", "nrg_b",
                          "nrg_c", "nrg_d")]


class MMKit(ctypes.Structure):
    _fields_ = [("step", c_int32), ("sd_cycles", c_int32)]


class MinimizeControls(ctypes.Structure):
    _fields_ = [("total_cycles", c_int32), ("steepest_descent_cycles", c_int32),
                ("initial_step", c_double), ("convergence_rms_force", c_double),
                ("clash_damping_cycles", c_int32), ("clash_minimum_distance", c_double),
                ("clash_vdw_ratio", c_double), ("diagnostic_print_frequency", c_int32)]


class PoolConfig(ctypes.Structure):
    _fields_ = [("precision", c_int32), ("gpos_bits", c_int32), ("frc_bits", c_int32),
                ("nrg_bits", c_int32), ("hbm_budget_bytes", c_int64), ("max_systems", c_int32),
                ("igb", c_int32), ("dielectric", c_double), ("salt_concentration", c_double),
                ("coulomb_constant", c_double), ("minimize", MinimizeControls),
                ("shape", LaunchShape)]


# name -> (restype, argtypes)
_SIGS = {
    "stormm_last_error": (ctypes.c_char_p, []),
    "stormm_pool_create": (c_int32, [P(PoolConfig), P(vp)]),
    "stormm_pool_destroy": (c_int32, [vp]),
    "stormm_pool_capacity": (c_int32, [vp, P(c_int64), P(c_int32), P(c_int64), P(c_int64)]),
    "stormm_pool_admit": (c_int32, [vp, P(TopologyBatch), vp, vp]),
    "stormm_pool_admit_conformers": (c_int32, [vp, P(TopologyBatch), c_int32, vp, vp]),
    "stormm_pool_release": (c_int32, [vp, vp, c_int32]),
    "stormm_pool_minimize": (c_int32, [vp, c_int32, P(c_int32), vp]),
    "stormm_pool_status": (c_int32, [vp, vp, c_int32, vp, vp, vp, vp]),
    "stormm_pool_coordinates": (c_int32, [vp, c_int32, vp, vp]),
    "stormm_pool_history": (c_int32, [vp, c_int32, c_int32, P(c_int32), vp, vp]),
    "stormm_device_info": (c_int32, [P(c_int32), P(c_int32), ctypes.c_char_p, c_int32]),
    "stormm_device_synchronize": (c_int32, []),
    "stormm_synthesis_create": (c_int32, [P(TopologyBatch), P(LaunchShape), P(vp)]),
    "stormm_synthesis_destroy": (c_int32, [vp]),
    "stormm_synthesis_get_info": (c_int32, [vp, P(SynthesisInfo)]),
    "stormm_synthesis_atom_layout": (c_int32, [vp, vp, vp]),
    "stormm_synthesis_plan": (c_int32, [P(TopologyBatch), P(LaunchShape), P(SynthesisInfo)]),
    "stormm_synthesis_status": (c_int32, [vp, P(c_int32), c_int32]),
    "stormm_autotune_launch_shape": (c_int32, [P(TopologyBatch), c_int32, vp, P(LaunchShape),
                                               P(c_double)]),
    "stormm_phasespace_create": (c_int32, [vp, c_int32, c_int32, P(vp)]),
    "stormm_phasespace_destroy": (c_int32, [vp]),
    "stormm_phasespace_upload_coordinates": (c_int32, [vp, vp]),
    "stormm_phasespace_download_coordinates": (c_int32, [vp, vp]),
    "stormm_phasespace_download_forces": (c_int32, [vp, vp]),
    "stormm_phasespace_upload_fixed": (c_int32, [vp, vp]),
    "stormm_phasespace_download_fixed": (c_int32, [vp, vp, vp]),
    "stormm_phasespace_upload_systems": (c_int32, [vp, c_int32, c_int32, vp]),
    "stormm_phasespace_download_systems": (c_int32, [vp, c_int32, c_int32, vp, vp]),
    "stormm_phasespace_import_system": (c_int32, [vp, c_int32, vp]),
    "stormm_phasespace_export_system": (c_int32, [vp, c_int32, vp, vp]),
    "stormm_phasespace_writer": (c_int32, [vp, P(PsWriter)]),
    "stormm_phasespace_upload_fixed_split": (c_int32, [vp, vp, vp]),
    "stormm_phasespace_download_fixed_split": (c_int32, [vp, vp, vp, vp, vp]),
    "stormm_scorecard_data": (c_int32, [vp, P(ScoreCardWriter)]),
    "stormm_mm_controls_data": (c_int32, [vp, P(MMKit)]),
    "stormm_line_min_data": (c_int32, [vp, P(LinMinWriter)]),
    "stormm_check_ps_writer": (c_int32, [vp, P(PsWriter)]),
    "stormm_launch_nonbonded_kit": (c_int32, [c_int32, vp, P(MMKit), P(PsWriter),
                                              P(ScoreCardWriter), c_int32, c_int32, c_double,
                                              c_double, vp]),
    "stormm_launch_valence_kit": (c_int32, [c_int32, vp, P(MMKit), P(PsWriter),
                                            P(ScoreCardWriter), c_int32, c_int32, c_double,
                                            c_double, vp]),
    "stormm_launch_mm_evaluation_kit": (c_int32, [c_int32, vp, P(MMKit), P(PsWriter),
                                                  P(ScoreCardWriter), c_int32, c_int32,
                                                  c_double, c_double, vp]),
    "stormm_launch_conjugate_gradient_kit": (c_int32, [c_int32, vp, P(PsWriter), P(MMKit), vp]),
    "stormm_launch_line_advance_kit": (c_int32, [c_int32, vp, P(PsWriter), P(ScoreCardWriter),
                                                 P(LinMinWriter), c_int32, vp]),
    "stormm_launch_minimization_kit": (c_int32, [c_int32, vp, P(PsWriter),
                                                 P(MinimizeControls), P(ScoreCardWriter),
                                                 P(LinMinWriter), c_int32, vp]),
    "stormm_launch_minimization_kit_noalloc": (c_int32, [c_int32, vp, P(PsWriter),
                                                         P(MinimizeControls),
                                                         P(ScoreCardWriter), P(LinMinWriter),
                                                         vp]),
    "stormm_ps_writer_initialize_forces": (c_int32, [P(PsWriter), c_int32, vp]),
    "stormm_ps_writer_prime_conjugate_gradient": (c_int32, [P(PsWriter), c_int32, vp]),
    "stormm_scorecard_writer_initialize": (c_int32, [P(ScoreCardWriter), vp]),
    "stormm_phasespace_initialize_forces": (c_int32, [vp, vp]),
    "stormm_phasespace_prime_conjugate_gradient": (c_int32, [vp, vp]),
    "stormm_scorecard_create": (c_int32, [c_int32, c_int32, P(vp)]),
    "stormm_scorecard_destroy": (c_int32, [vp]),
    "stormm_scorecard_initialize": (c_int32, [vp, vp]),
    "stormm_scorecard_commit": (c_int32, [vp, vp]),
    "stormm_scorecard_download": (c_int32, [vp, vp]),
    "stormm_scorecard_download_history": (c_int32, [vp, vp, vp, P(c_int32)]),
    "stormm_scorecard_reserve": (c_int32, [vp, c_int32]),
    "stormm_scorecard_reset_sample_count": (c_int32, [vp, c_int32]),
    "stormm_scorecard_set_last_time_step": (c_int32, [vp, c_int32]),
    "stormm_scorecard_sample_count": (c_int32, [vp, P(c_int32), P(c_int32)]),
    "stormm_scorecard_download_series": (c_int32, [vp, c_int32, c_int32, vp, vp]),
    "stormm_scorecard_device_accumulators": (vp, [vp]),
    "stormm_mm_controls_create": (c_int32, [P(MinimizeControls), P(vp)]),
    "stormm_mm_controls_destroy": (c_int32, [vp]),
    "stormm_mm_controls_get_step": (c_int32, [vp, P(c_int32)]),
    "stormm_mm_controls_set_step": (c_int32, [vp, c_int32]),
    "stormm_line_min_create": (c_int32, [c_int32, c_double, P(vp)]),
    "stormm_line_min_destroy": (c_int32, [vp]),
    "stormm_line_min_download": (c_int32, [vp, vp]),
    "stormm_launch_nonbonded": (c_int32, [c_int32, vp, vp, vp, vp, c_int32, c_int32, c_double,
                                          c_double, vp]),
    "stormm_launch_valence": (c_int32, [c_int32, vp, vp, vp, vp, c_int32, c_int32, c_double,
                                        c_double, vp]),
    "stormm_launch_mm_evaluation": (c_int32, [c_int32, vp, vp, vp, vp, c_int32, c_int32,
                                              c_double, c_double, vp]),
    "stormm_launch_conjugate_gradient": (c_int32, [c_int32, vp, vp, vp, vp]),
    "stormm_launch_line_advance": (c_int32, [c_int32, vp, vp, vp, vp, c_int32, vp]),
    "stormm_launch_minimization": (c_int32, [c_int32, vp, vp, P(MinimizeControls), vp, vp,
                                             c_int32, vp]),
    "stormm_minimization_status": (c_int32, [vp, vp, vp]),
    "stormm_copy_bandwidth": (c_int32, [c_int64, c_int32, P(c_double)]),
    "stormm_fma_throughput": (c_int32, [c_int32, P(c_double)]),
    "stormm_time_launch": (c_int32, [c_int32, vp, vp, vp, c_int32, c_int32, c_int32, vp,
                                     P(c_double)]),
    "stormm_time_mm_evaluation": (c_int32, [c_int32, vp, vp, vp, c_int32, c_int32, vp,
                                            P(c_double)]),
}

_lib = None


def declared_symbols():
    """Names of every function the public header declares."""
    with open(HEADER_PATH, "r") as fh:
        text = fh.read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(stormm_[a-z0-9_]+)\s*\(", text)))


def load():
    """Load the HIP library (raises if it is missing: there is no CPU fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing; build it with `python -m stormm_amd.build` "
                           "(the HIP path has no fallback)")
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in _SIGS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(rc: int) -> None:
    if rc != STORMM_OK:
        msg = load().stormm_last_error()
        raise StormmError(rc, msg.decode() if msg else "")
