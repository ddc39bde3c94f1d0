"""ctypes binding of libgpgmp_hip.so (the C ABI declared in include/gpgmp_hip.h).

The library must be built (``python -m inchman_amd.build`` or
``__graft_entry__.build()``); loading fails loudly otherwise -- there is no
fallback implementation.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libgpgmp_hip.so")

MAX_SPECIES = 8
MAX_REACTIONS = 16

OK = 0
E_INVALID = -1
E_HIP = -2
E_CAPACITY = -3
E_STATE = -4
E_UNSUPPORTED = -5

STOCHASTIC_HOMOGENEOUS = 0
STOCHASTIC_INHOMOGENEOUS = 1
DETERMINISTIC_HOMOGENEOUS_RK4 = 2
DETERMINISTIC_INHOMOGENEOUS_RK4 = 3
DETERMINISTIC_HOMOGENEOUS_AQSS = 4
DETERMINISTIC_INHOMOGENEOUS_AQSS = 5
STOCHASTIC_INHOMOGENEOUS_FPE = 6
AQSS_PADE_FIX = 1


class NewIndividualRule(C.Structure):
    """gpgmp_new_individual_rule: per property (DiffusivityX, DiffusivityY, DriftX, DriftY)
    mode 0 keep, 1 constant a, 2 a + b * RANDOM."""
    _fields_ = [("mode", C.c_int32 * 4), ("a", C.c_double * 4), ("b", C.c_double * 4)]


class GpgmpError(RuntimeError):
    """A negative status from the C ABI (the reference raised gpgmp::Error)."""

    def __init__(self, code, msg):
        super().__init__("gpgmp error %d: %s" % (code, msg))
        self.code = code


class ReactionDesc(C.Structure):
    _fields_ = [
        ("rate", C.c_float),
        ("reactant_stoich", C.c_int32 * MAX_SPECIES),
        ("product_stoich", C.c_int32 * MAX_SPECIES),
    ]


class ModelDesc(C.Structure):
    _fields_ = [
        ("width", C.c_int32), ("height", C.c_int32),
        ("physical_length", C.c_float), ("p0", C.c_float),
        ("num_species", C.c_int32),
        ("diffusion", C.c_float * MAX_SPECIES),
        ("num_reactions", C.c_int32),
        ("reactions", C.POINTER(ReactionDesc)),
        ("boundary_mask", C.c_int32 * 4),
        ("source_mask", C.c_int32 * 4),
        ("max_time", C.c_float),
        ("solver_type", C.c_int32),
        ("diffusivity_x", C.POINTER(C.c_float)), ("diffusivity_y", C.POINTER(C.c_float)),
        ("drift_x", C.POINTER(C.c_float)), ("drift_y", C.POINTER(C.c_float)),
    ]


class RddParams(C.Structure):
    _fields_ = [
        ("width", C.c_int32), ("height", C.c_int32),
        ("h", C.c_float), ("dt", C.c_float),
        ("D0u", C.c_float), ("D0v", C.c_float),
        ("k", C.c_float * 4),
    ]


class IndividualsDesc(C.Structure):
    _fields_ = [
        ("width", C.c_int32), ("height", C.c_int32), ("physical_length", C.c_float),
        ("boundary_mask", C.c_int32 * 4), ("species", C.c_int32),
    ]


class ParticleParams(C.Structure):
    _fields_ = [
        ("side", C.c_int32),
        ("D", C.c_float), ("dt", C.c_float), ("vortex_amp", C.c_float), ("p_annihilate", C.c_float),
        ("seed", C.c_uint32),
    ]


_P = C.c_void_p
_I = C.c_int32
_U = C.c_uint32
_F = C.c_float
_pI = C.POINTER(C.c_int32)
_pU = C.POINTER(C.c_uint32)
_pF = C.POINTER(C.c_float)
_pD = C.POINTER(C.c_double)

# name -> argtypes (all return int status except gpgmp_last_error)
SIGNATURES = {
    "gpgmp_version": [_pI, _pI],
    "gpgmp_model_create": [C.POINTER(ModelDesc), C.POINTER(_P)],
    "gpgmp_model_destroy": [_P],
    "gpgmp_model_rate_literals": [_P, _pI, _pD],
    "gpgmp_model_add_reaction_compartment": [_P, _I, _I, _I, _I, _I],
    "gpgmp_model_localized": [_P, _pU],
    "gpgmp_model_set_species_ids": [_P, C.POINTER(C.c_char_p)],
    "gpgmp_model_set_parameter": [_P, C.c_char_p, _F],
    "gpgmp_model_set_kinetic_law": [_P, _I, C.c_char_p],
    "gpgmp_model_law_hazard": [_P, _I, _pI, _I, _I, _F, _U, _I, _pF],
    "gpgmp_model_law_rhs": [_P, _I, _I, _pF, _I, _I, _U, _pF],
    "gpgmp_model_eval_drift_diffusivity": [_P, C.c_char_p, C.POINTER(C.c_char_p), _I, _pF, _pI, _pF, _F, _pF, _P],
    "gpgmp_model_method_host": [_P, C.c_char_p, C.POINTER(C.c_char_p), _I, _pF, _pI, _pF, _F, _pF],
    "gpgmp_model_method_dependencies": [_P, C.c_char_p, C.POINTER(C.c_char_p), _I, C.POINTER(C.c_uint32)],
    "gpgmp_solver_create": [_P, _P, _U, C.POINTER(_P)],
    "gpgmp_solver_destroy": [_P],
    "gpgmp_solver_set_state": [_P, _pI],
    "gpgmp_solver_get_state": [_P, _pI],
    "gpgmp_solver_set_state_f32": [_P, _pF],
    "gpgmp_solver_get_state_f32": [_P, _pF],
    "gpgmp_solver_state_device": [_P, C.POINTER(_P)],
    "gpgmp_solver_do_reactions": [_P, _F, _pI],
    "gpgmp_solver_diffuse": [_P, _I, _F, _pI],
    "gpgmp_solver_next_diffusion_time": [_P, _pF],
    "gpgmp_solver_set_nonlinear_method": [_P, C.c_char_p, C.POINTER(C.c_char_p), _I, _pF],
    "gpgmp_loop_seeds": [C.c_uint32, _I, _pI],
    "gpgmp_solver_time_steps": [_P, _pF, _pF],
    "gpgmp_solver_run": [_P, _I, _U, _pF],
    "gpgmp_solver_set_output": [_P, _F, _F, _I],
    "gpgmp_solver_set_sweep_spread": [_P, _I],
    "gpgmp_solver_set_sweep_wave_cells": [_P, _I],
    "gpgmp_solver_set_predict_busy": [_P, _I],
    "gpgmp_solver_get_output": [_P, _P, _P, _pI, _pI, C.POINTER(C.c_int64)],
    "gpgmp_solver_run_continue": [_P, _I, _I, _pF],
    "gpgmp_solver_prepare_run": [_P, _I, _I],
    "gpgmp_solver_skip": [_P, _I, _pF],
    "gpgmp_solver_individuals": [_P, _I, C.POINTER(C.c_void_p)],
    "gpgmp_solver_set_launch": [_P, _I, _I],
    "gpgmp_solver_kernel_times": [_P, _I, _pD, C.POINTER(C.c_int64), _pD],
    "gpgmp_solver_set_event_budget": [_P, _U, _I],
    "gpgmp_solver_set_hot_workers": [_P, _F, _I, _I],
    "gpgmp_solver_set_lanes_per_cell": [_P, _I, _I],
    "gpgmp_solver_get_launch": [_P, _pI, _pI, _pI, _pI],
    "gpgmp_batch_set_lanes_per_cell": [_P, _I, _I],
    "gpgmp_batch_get_launch": [_P, _pI, _pI, _pI, _pI],
    "gpgmp_batch_set_event_budget": [_P, _U, _I],
    "gpgmp_solver_synchronize": [_P],
    "gpgmp_batch_create": [_P, _I, _pI, _pI, _pU, _P, C.POINTER(_P)],
    "gpgmp_batch_destroy": [_P],
    "gpgmp_batch_set_state": [_P, _I, _pI],
    "gpgmp_batch_get_state": [_P, _I, _pI],
    "gpgmp_batch_set_launch": [_P, _I, _I],
    "gpgmp_batch_autotune": [_P, _I, _pI, _pI],
    "gpgmp_batch_run": [_P, _I, _I],
    "gpgmp_batch_prepare_run": [_P, _I, _I],
    "gpgmp_batch_reset": [_P],
    "gpgmp_batch_kernel_times": [_P, _I, _pD, C.POINTER(C.c_int64), _pD],
    "gpgmp_batch_total_cells": [_P, C.POINTER(C.c_int64)],
    "gpgmp_batch_synchronize": [_P],
    "gpgmp_min_timestep_device": [_P, _P, _I, _F, _pF, _P],
    "gpgmp_rdd_create": [C.POINTER(RddParams), _I, _I, _P, C.POINTER(_P)],
    "gpgmp_rdd_destroy": [_P],
    "gpgmp_rdd_set_rates": [_P, _pF],
    "gpgmp_rdd_set_steps_per_launch": [_P, _I],
    "gpgmp_rdd_forward": [_P, _P, _P, _P, _I],
    "gpgmp_rdd_backward": [_P, C.c_uint64, _P, _P, _P, _P, _P],
    "gpgmp_rdd_set_rates_device": [_P, _P],
    "gpgmp_rdd_generation": [_P, C.POINTER(C.c_uint64)],
    "gpgmp_det_hom_diffuse_device": [_P, _P, _I, _I, _I, _pF, _F, _F, _P],
    "gpgmp_det_inhom_sweep_device": [_P, _P, _I, _I, _I, _I, _F, _P, _P, _F, _P],
    "gpgmp_det_rk4_device": [_P, _P, _F, _P],
    "gpgmp_det_aqss_device": [_P, _P, _F, _I, _P],
    "gpgmp_moments_x_device": [_P, _I, _I, _I, _P, _P],
    "gpgmp_sort_pairs_device": [_P, _P, _P, _P, _I, _I, _P],
    "gpgmp_sort_pairs_device_noalloc": [_P, _P, _P, _P, _I, _I, _P],
    "gpgmp_cell_offsets_device": [_P, _I, _I, _P, _P],
    "gpgmp_cell_left_right_device": [_P, _I, _I, _P, _P, _P],
    "gpgmp_particles_create": [C.POINTER(ParticleParams), _I, _P, C.POINTER(_P)],
    "gpgmp_particles_destroy": [_P],
    "gpgmp_particles_set": [_P, _pF, _pF, _pU, _I],
    "gpgmp_particles_get": [_P, _pF, _pF, _pU, _pU, _pU, _pI],
    "gpgmp_particles_count": [_P, _pI],
    "gpgmp_particles_step": [_P, _I, _I],
    "gpgmp_particles_processed": [_P, C.POINTER(C.c_uint64)],
    "gpgmp_particles_set_sorted_step": [_P, _I],
    "gpgmp_individuals_create": [C.POINTER(IndividualsDesc), _I, _P, C.POINTER(_P)],
    "gpgmp_individuals_destroy": [_P],
    "gpgmp_individuals_set": [_P, _P, _P, _I],
    "gpgmp_individuals_time_steps": [_P, _pF, _pF],
    "gpgmp_individuals_diffuse": [_P, _I, _F, _pI],
    "gpgmp_individuals_update": [_P, _I, _pF, _pF],
    "gpgmp_individuals_run": [_P, _I, _U, _pF],
    "gpgmp_individuals_get": [_P, _P, _P, _P, _pI],
    "gpgmp_individuals_constants": [_P, _P, _P, _P, _P],
    "gpgmp_individuals_keys_device": [_P, _P, _P, _P],
    "gpgmp_individuals_sweeps": [_P, C.POINTER(C.c_int64)],
    "gpgmp_individuals_maintain": [_P, _P, _I, _I, _P],
    "gpgmp_individuals_get_props": [_P, _P],
    "gpgmp_individuals_get_pool": [_P, _P, _P, _pI],
    "gpgmp_individuals_bind_counts": [_P, _P],
    "gpgmp_individuals_synchronize": [_P],
    "gpgmp_solver_attach_individuals": [_P, _I, _P, _P],
}

_lib = None


def lib():
    """Loads the library once; raises if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            "libgpgmp_hip.so is not built (%s); run `python -m inchman_amd.build`" % LIB_PATH)
    L = C.CDLL(LIB_PATH)
    for name, args in SIGNATURES.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = C.c_int
    L.gpgmp_last_error.argtypes = []
    L.gpgmp_last_error.restype = C.c_char_p
    _lib = L
    return L


def check(code):
    if code != OK:
        msg = lib().gpgmp_last_error()
        raise GpgmpError(code, msg.decode() if msg else "")
    return code


def call(name, *args):
    return check(getattr(lib(), name)(*args))
