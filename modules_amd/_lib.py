"""Loader of the native engine (modules_amd/libccgpu.so) through its C ABI (include/ccgpu.h).

The library is the product: there is no Python or CPU fallback.  Loading fails loudly when the shared
object is missing (run ``python -c "import __graft_entry__ as g; g.build()"`` or ``make -C modules_amd/csrc``).
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libccgpu.so")

dp = C.POINTER(C.c_double)
ip32 = C.POINTER(C.c_int32)
i64p = C.POINTER(C.c_int64)
u64p = C.POINTER(C.c_uint64)
u32p = C.POINTER(C.c_uint32)
vp = C.c_void_p

_lib = None


class CCGPUError(RuntimeError):
    """Raised for a nonzero status of the C ABI; .code holds the CCGPU_ERR_* value."""

    def __init__(self, code: int, msg: str):
        super().__init__(f"ccgpu error {code}: {msg}")
        self.code = code


# status codes (include/ccgpu.h)
OK, ERR_ARG, ERR_STATE, ERR_RANGE, ERR_DEVICE, ERR_MEMORY, ERR_NAN = 0, -1, -2, -3, -4, -5, -6

# function name -> (restype, argtypes); every function of include/ccgpu.h is declared here
SIGNATURES = {
    "ccgpu_last_error": (C.c_char_p, []),
    "ccgpu_get_k": (C.c_double, [C.c_double, C.c_double, C.c_int]),
    "ccgpu_get_ks": (C.c_int, [C.c_int, dp, C.c_double, C.c_int64, dp, ip32, i64p]),
    "ccgpu_reciprocal_basis": (C.c_int, [C.c_int, dp, dp]),
    "ccgpu_cell_volume": (C.c_double, [C.c_int, dp]),
    "ccgpu_random_configuration": (C.c_int, [C.c_uint32, C.c_int64, C.c_int, dp]),
    "ccgpu_potential_create": (C.c_int, [C.c_int, C.POINTER(vp)]),
    "ccgpu_potential_destroy": (C.c_int, [vp]),
    "ccgpu_potential_add_constraint": (C.c_int, [vp, dp, C.c_double]),
    "ccgpu_potential_set_constraints": (C.c_int, [vp, C.c_int64, dp, dp]),
    "ccgpu_potential_num_constraints": (C.c_int, [vp, i64p]),
    "ccgpu_potential_set_shift": (C.c_int, [vp, C.c_double]),
    "ccgpu_potential_set_softcore": (C.c_int, [vp, C.c_double, C.c_double]),
    "ccgpu_potential_set_configuration": (C.c_int, [vp, dp, C.c_int64, dp]),
    "ccgpu_potential_get_rho": (C.c_int, [vp, dp, dp]),
    "ccgpu_potential_energy": (C.c_int, [vp, dp]),
    "ccgpu_potential_force": (C.c_int, [vp, C.c_int64, dp]),
    "ccgpu_potential_all_force": (C.c_int, [vp, dp]),
    "ccgpu_potential_energy_and_all_force": (C.c_int, [vp, dp, dp]),
    "ccgpu_energy_and_all_force_noalloc": (C.c_int, [C.c_int, dp, C.c_int64, dp, C.c_int64, dp, dp, dp, dp]),
    "ccgpu_potential_set_tiles": (C.c_int, [vp, C.c_int, C.c_int, C.c_int]),
    "ccgpu_potential_uses_lattice_path": (C.c_int, [vp, C.POINTER(C.c_int)]),
    "ccgpu_potential_set_rho_kernel": (C.c_int, [vp, C.c_int]),
    "ccgpu_potential_set_precision": (C.c_int, [vp, C.c_int]),
    "ccgpu_potential_set_rho_backend": (C.c_int, [vp, C.c_int, C.c_double, C.c_int64]),
    "ccgpu_potential_rho_backend_used": (C.c_int, [vp, C.POINTER(C.c_int)]),
    "ccgpu_potential_force_backend_used": (C.c_int, [vp, C.POINTER(C.c_int)]),
    "ccgpu_potential_nufft_status": (C.c_int, [vp, C.POINTER(C.c_int), C.c_char_p, C.c_int64, dp, dp]),
    "ccgpu_autotune_backend": (C.c_int, [C.c_int, C.c_int64, dp, C.c_int64, dp, dp, C.c_double, C.c_int,
                                         C.POINTER(C.c_int), dp, dp]),
    "ccgpu_solver_set_rho_backend": (C.c_int, [vp, C.c_int, C.c_double]),
    "ccgpu_copy_probe": (C.c_int, [C.c_int64, C.c_int, dp]),
    "ccgpu_lbfgs_minimize": (C.c_int, [C.c_int, vp, vp, dp, dp, C.c_double, C.c_int, C.c_int,
                                       C.POINTER(C.c_int)]),
    "ccgpu_potential_relax": (C.c_int, [vp, dp, C.c_int, C.c_double, C.c_int, dp, C.POINTER(C.c_int),
                                        C.POINTER(C.c_int)]),
    "ccgpu_potential_set_weights": (C.c_int, [vp, C.c_int64, dp]),
    "ccgpu_potential_set_log_form": (C.c_int, [vp, C.c_int]),
    "ccgpu_potential_try_move": (C.c_int, [vp, C.c_int64, dp, dp, dp]),
    "ccgpu_potential_accept_move": (C.c_int, [vp]),
    "ccgpu_md_create": (C.c_int, [C.c_int, dp, C.c_int64, dp, C.c_double, dp, C.POINTER(vp)]),
    "ccgpu_md_destroy": (C.c_int, [vp]),
    "ccgpu_md_set_velocities": (C.c_int, [vp, dp]),
    "ccgpu_md_get_state": (C.c_int, [vp, dp, dp, dp, dp, dp, dp]),
    "ccgpu_md_restore": (C.c_int, [vp, dp, dp, C.c_double, C.c_double, C.c_double]),
    "ccgpu_md_get_q": (C.c_int, [vp, dp]),
    "ccgpu_md_set_timestep": (C.c_int, [vp, C.c_double]),
    "ccgpu_md_kinetic_energy": (C.c_int, [vp, dp]),
    "ccgpu_md_evolve": (C.c_int, [vp, vp, C.c_int64]),
    "ccgpu_md_init_nose_hoover": (C.c_int, [vp, C.c_double]),
    "ccgpu_md_nose_hoover_evolve": (C.c_int, [vp, vp, C.c_int64, C.c_double]),
    "ccgpu_md_anderson_evolve": (C.c_int, [vp, vp, C.c_int64, C.c_double, C.c_double, dp, C.c_int64, i64p]),
    "ccgpu_md_set_random_speed": (C.c_int, [vp, C.c_double, dp, C.c_int64, i64p]),
    "ccgpu_md_remove_overall_momentum": (C.c_int, [vp]),
    "ccgpu_set_softcore_kernel": (C.c_int, [C.c_int]),
    "ccgpu_potential_set_indexed_configuration": (C.c_int, [vp, dp, C.c_int64, dp, C.c_int64, i64p]),
    "ccgpu_potential_create_s2": (C.c_int, [C.c_int, C.POINTER(vp)]),
    "ccgpu_potential_create_s2_multiconfig": (C.c_int, [C.c_int, C.c_int64, C.POINTER(vp)]),
    "ccgpu_potential_add_constraint_s2": (C.c_int, [vp, dp, C.c_double, C.c_double]),
    "ccgpu_potential_set_constraints_s2": (C.c_int, [vp, C.c_int64, dp, dp, dp]),
    "ccgpu_potential_create_varbox": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double,
                                                C.POINTER(vp)]),
    "ccgpu_potential_get_constraints": (C.c_int, [vp, dp, dp, dp, ip32]),
    "ccgpu_varbox_kpoints": (C.c_int, [C.c_int, dp, C.c_double, C.c_int64, dp, ip32, i64p]),
    "ccgpu_potential_energy_derivative_to_basis": (C.c_int, [vp, C.c_double, dp]),
    "ccgpu_potential_relax_cell": (C.c_int, [vp, dp, dp, C.c_double, C.c_int, C.c_double, C.c_int, C.c_double, C.c_int,
                                             dp, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "ccgpu_solver_create": (C.c_int, [C.c_int, C.c_int64, C.POINTER(vp)]),
    "ccgpu_solver_destroy": (C.c_int, [vp]),
    "ccgpu_solver_submit": (C.c_int, [vp, C.c_int, dp, C.c_int64, dp, C.c_int64, dp, dp, C.c_double, C.c_double,
                                      C.c_double, C.c_int, C.c_double, C.c_int, i64p]),
    "ccgpu_solver_submit_s2": (C.c_int, [vp, C.c_int, dp, C.c_int64, dp, C.c_int64, dp, dp, dp, C.c_double,
                                         C.c_double, C.c_double, C.c_int, C.c_double, C.c_int, i64p]),
    "ccgpu_solver_submit_multi": (C.c_int, [vp, C.c_int, dp, C.c_int64, dp, C.c_int64, i64p, i64p, i64p, dp, dp, dp,
                                            C.c_double, C.c_double, C.c_double, C.c_int, C.c_double, C.c_int, i64p]),
    "ccgpu_solver_step": (C.c_int, [vp, C.c_int, C.POINTER(C.c_int)]),
    "ccgpu_solver_set_check_interval": (C.c_int, [vp, C.c_int]),
    "ccgpu_solver_set_graph": (C.c_int, [vp, C.c_int]),
    "ccgpu_solver_set_lanes": (C.c_int, [vp, C.c_int]),
    "ccgpu_solver_set_precision": (C.c_int, [vp, C.c_int, C.c_double]),
    "ccgpu_solver_switch_eval": (C.c_int, [vp, C.c_int64, C.POINTER(C.c_int)]),
    "ccgpu_solver_result": (C.c_int, [vp, C.c_int64, C.POINTER(C.c_int), dp, C.POINTER(C.c_int),
                                      C.POINTER(C.c_int), C.POINTER(C.c_int), dp, dp]),
    "ccgpu_solver_release": (C.c_int, [vp, C.c_int64]),
    "ccgpu_solver_stats": (C.c_int, [vp, C.POINTER(C.c_int), C.POINTER(C.c_int), i64p]),
    "ccgpu_solver_set_timing": (C.c_int, [vp, C.c_int]),
    "ccgpu_solver_launch_stats": (C.c_int, [vp, i64p, i64p]),
    "ccgpu_solver_set_time_limit": (C.c_int, [vp, C.c_double]),
    "ccgpu_solver_job_bytes": (C.c_int, [vp, C.c_int64, i64p, i64p]),
    "ccgpu_autotune": (C.c_int, [C.c_int, C.c_int64, dp, C.c_int64, dp, dp, C.c_int, C.c_int, C.POINTER(C.c_int),
                                 C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int), dp]),
    "ccgpu_autotune_clear": (C.c_int, []),
    "ccgpu_autotune_record": (C.c_int, [C.c_int, C.c_int64, C.c_int64, C.c_int, C.c_int, C.c_int, C.c_int,
                                        C.c_int]),
    "ccgpu_solver_timing": (C.c_int, [vp, dp, dp, i64p, i64p, C.c_int]),
    "ccgpu_multi_create": (C.c_int, [C.c_int, C.c_double, C.POINTER(vp)]),
    "ccgpu_multi_destroy": (C.c_int, [vp]),
    "ccgpu_multi_set_num_species": (C.c_int, [vp, C.c_int64]),
    "ccgpu_multi_num_species": (C.c_int, [vp, i64p]),
    "ccgpu_multi_set_species_of_particle": (C.c_int, [vp, C.c_int64, C.c_int64]),
    "ccgpu_multi_species_of": (C.c_int, [vp, ip32]),
    "ccgpu_multi_species_indices": (C.c_int, [vp, C.c_int64, C.c_int64, i64p, i64p]),
    "ccgpu_multi_add_constraint": (C.c_int, [vp, C.c_int64, dp, C.c_double]),
    "ccgpu_multi_set_species_constraints": (C.c_int, [vp, C.c_int64, C.c_int64, dp, dp]),
    "ccgpu_multi_num_constraints": (C.c_int, [vp, C.c_int64, i64p]),
    "ccgpu_multi_set_scales": (C.c_int, [vp, C.c_double, C.c_double]),
    "ccgpu_multi_set_configuration": (C.c_int, [vp, dp, C.c_int64, dp]),
    "ccgpu_multi_set_radii": (C.c_int, [vp, C.c_int64, dp]),
    "ccgpu_multi_set_packing": (C.c_int, [vp, dp, C.c_int64, dp, dp]),
    "ccgpu_multi_get_radii": (C.c_int, [vp, dp]),
    "ccgpu_multi_get_rc": (C.c_int, [vp, dp]),
    "ccgpu_multi_energy": (C.c_int, [vp, dp]),
    "ccgpu_multi_force": (C.c_int, [vp, C.c_int64, dp]),
    "ccgpu_multi_all_force": (C.c_int, [vp, dp]),
    "ccgpu_multi_energy_and_all_force": (C.c_int, [vp, dp, dp]),
    "ccgpu_multi_get_rho": (C.c_int, [vp, C.c_int64, dp, dp]),
    "ccgpu_multi_species_energy": (C.c_int, [vp, C.c_int64, dp]),
    "ccgpu_multi_relax": (C.c_int, [vp, dp, C.c_int, C.c_double, C.c_int, dp, C.POINTER(C.c_int),
                                    C.POINTER(C.c_int)]),
    "ccgpu_multi_export": (C.c_int, [vp, C.c_int, vp, C.c_int64, i64p]),
    "ccgpu_pack_count": (C.c_int, [C.c_char_p, C.c_int, i64p]),
    "ccgpu_pack_add": (C.c_int, [C.c_char_p, C.c_int, C.c_int, C.c_double, dp, C.c_int64, dp, vp]),
    "ccgpu_pack_get": (C.c_int, [C.c_char_p, C.c_int, C.c_int64, C.POINTER(C.c_int), dp, dp, i64p, C.c_int64, dp,
                                 vp]),
    "ccgpu_pack_clear": (C.c_int, [C.c_char_p, C.c_int]),
    "ccgpu_mddump_write": (C.c_int, [C.c_char_p, C.c_int32, C.c_int64, C.c_int, C.c_double, dp, C.c_int64, dp, vp,
                                     C.c_double, C.c_double, C.c_double, C.c_double, dp, dp, dp]),
    "ccgpu_mddump_read": (C.c_int, [C.c_char_p, C.POINTER(C.c_int32), i64p, C.POINTER(C.c_int), dp, dp, i64p,
                                    C.c_int64, dp, vp, dp, dp, dp, dp, dp, dp, dp]),
    "ccgpu_mspb_encode": (C.c_int, [C.c_int, dp, C.c_int64, u64p, u32p, dp, dp, C.c_uint64, vp, C.c_int64, i64p]),
    "ccgpu_mspb_decode": (C.c_int, [vp, C.c_int64, C.POINTER(C.c_int), dp, i64p, u64p, C.c_int64, u64p, u32p, dp,
                                    dp]),
    "ccgpu_export_text": (C.c_int, [C.c_int, dp, C.c_int64, ip32, dp, dp, vp, C.c_int64, i64p]),
    "ccgpu_mspbpack_count": (C.c_int, [C.c_char_p, i64p]),
    "ccgpu_mspbpack_add": (C.c_int, [C.c_char_p, vp, C.c_int64]),
    "ccgpu_mspbpack_get": (C.c_int, [C.c_char_p, C.c_int64, vp, C.c_int64, i64p]),
}

OBJECTIVE = C.CFUNCTYPE(C.c_double, C.c_uint, dp, dp, C.c_void_p)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"native engine not built: {LIB_PATH} is missing")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(status: int):
    if status != OK:
        msg = lib().ccgpu_last_error()
        raise CCGPUError(status, msg.decode() if msg else "")
    return status
