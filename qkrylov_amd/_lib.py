"""ctypes binding of the engine's C ABI (include/qkrylov/c_api.h and
include/qkrylov_amd.h). Loading fails loudly when the HIP library is not
built: there is no CPU fallback."""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libqkrylov_amd.so")

c_int, c_double, c_float = ctypes.c_int, ctypes.c_double, ctypes.c_float
c_u64, c_i64, c_u32, c_size = ctypes.c_uint64, ctypes.c_int64, ctypes.c_uint32, ctypes.c_size_t
c_vp, c_cp = ctypes.c_void_p, ctypes.c_char_p
P = ctypes.POINTER
p_dbl, p_flt, p_int, p_i64, p_u64 = P(c_double), P(c_float), P(c_int), P(c_i64), P(c_u64)


class LanczosResultC(ctypes.Structure):
    _fields_ = [("energy", c_float), ("iterations", c_int), ("converged", c_int)]


class LanczosResultF64(ctypes.Structure):
    _fields_ = [("energy", c_double), ("iterations", c_int), ("converged", c_int)]


class DavidsonResultC(ctypes.Structure):
    _fields_ = [("iterations", c_int), ("converged", c_int)]


DavidsonResultF64 = DavidsonResultC


class FTLMResultC(ctypes.Structure):
    _fields_ = [("beta", c_float), ("partition_function", c_float), ("internal_energy", c_float),
                ("specific_heat", c_float)]


class FTLMResultF64(ctypes.Structure):
    _fields_ = [("beta", c_double), ("partition_function", c_double), ("internal_energy", c_double),
                ("specific_heat", c_double)]


class HamInfo(ctypes.Structure):
    _fields_ = [("dim", c_u64), ("nbits", c_int), ("NL", c_int), ("is_real", c_int),
                ("n_offdiag_branches", c_int), ("n_diag_terms", c_int), ("ntiles", c_u32),
                ("table_bytes", c_u64)]


class LanczosOptions(ctypes.Structure):
    _fields_ = [("maxiter", c_int), ("tol", c_double), ("reorthogonalize", c_int), ("fixed_steps", c_int),
                ("want_vector", c_int), ("seed", c_u64), ("check_every", c_int), ("graph_mode", c_int),
                ("split_mode", c_int), ("start_mode", c_int)]


class LanczosReport(ctypes.Structure):
    _fields_ = [("energy", c_double), ("iterations", c_int), ("converged", c_int), ("steps_run", c_int),
                ("n_alpha", c_int), ("n_beta", c_int), ("loop_seconds", c_double), ("ritz_seconds", c_double),
                ("split", c_int), ("vector_complex", c_int)]


class ServeRequest(ctypes.Structure):
    _fields_ = [("L", c_int), ("sz2", c_int), ("delta", c_double), ("kind", c_int), ("arrival", c_double),
                ("tol", c_double), ("cf_steps", c_int), ("maxiter", c_int)]


class ServeResult(ctypes.Structure):
    _fields_ = [("energy", c_double), ("iterations", c_int), ("converged", c_int), ("n_cf", c_int),
                ("cf_norm", c_double), ("start", c_double), ("finish", c_double), ("latency", c_double),
                ("batch", c_int), ("ritz_stored", c_int)]


class ServeOptions(ctypes.Structure):
    _fields_ = [("hbm_budget_bytes", c_size), ("max_active", c_int), ("realtime", c_int),
                ("gs_vectors", P(c_vp)), ("max_batch", c_int)]


# name -> (restype, argtypes)
class TuneReport(ctypes.Structure):
    _fields_ = [("split_bits", c_int), ("apply_waves", c_int), ("schedule", c_int), ("block_waves", c_int),
                ("block_schedule", c_int), ("candidates", c_int), ("ms_before", c_double), ("ms_after", c_double),
                ("block_rows", c_int)]


SIGNATURES = {
    # sector
    "qkrylov_sector_create": (c_vp, []),
    "qkrylov_sector_destroy": (None, [c_vp]),
    "qkrylov_sector_set_sz": (c_int, [c_vp, c_int]),
    "qkrylov_sector_set_hubbard_particles": (c_int, [c_vp, c_int, c_int]),
    "qkrylov_sector_set_n": (c_int, [c_vp, c_int]),
    "qkrylov_amd_sector_set_species": (c_int, [c_vp, c_int, c_int]),
    "qkrylov_sector_set_nb": (c_int, [c_vp, c_int]),
    "qkrylov_sector_get_sz": (c_int, [c_vp, p_int, p_int]),
    "qkrylov_sector_get_hubbard_particles": (c_int, [c_vp, p_int, p_int, p_int]),
    "qkrylov_sector_get_n": (c_int, [c_vp, p_int, p_int]),
    "qkrylov_sector_get_nb": (c_int, [c_vp, p_int, p_int]),
    # basis
    "qkrylov_spinhalf_basis_create": (c_vp, [c_int, c_vp]),
    "qkrylov_fermion_basis_create": (c_vp, [c_int, c_vp]),
    "qkrylov_hubbard_basis_create": (c_vp, [c_int, c_vp]),
    "qkrylov_tj_basis_create": (c_vp, [c_int, c_vp]),
    "qkrylov_basis_destroy": (None, [c_vp]),
    "qkrylov_basis_dimension": (c_u64, [c_vp]),
    "qkrylov_basis_nsites": (c_int, [c_vp]),
    "qkrylov_basis_state": (c_u64, [c_vp, c_u64]),
    "qkrylov_basis_index": (c_i64, [c_vp, c_u64]),
    "qkrylov_basis_contains": (c_int, [c_vp, c_u64]),
    # site
    "qkrylov_spinhalf_site_create": (c_vp, []),
    "qkrylov_fermion_site_create": (c_vp, []),
    "qkrylov_hubbard_site_create": (c_vp, []),
    "qkrylov_tj_site_create": (c_vp, []),
    "qkrylov_site_destroy": (None, [c_vp]),
    # opsum
    "qkrylov_opsum_create": (c_vp, []),
    "qkrylov_opsum_destroy": (None, [c_vp]),
    "qkrylov_opsum_clear": (c_int, [c_vp]),
    "qkrylov_opsum_add_term_1body": (c_int, [c_vp, c_float, c_float, c_cp, c_int]),
    "qkrylov_opsum_add_term_2body": (c_int, [c_vp, c_float, c_float, c_cp, c_int, c_cp, c_int]),
    "qkrylov_opsum_add_term_nbody": (c_int, [c_vp, c_float, c_float, c_int, P(c_cp), p_int]),
    "qkrylov_opsum_add_term_nbody_f64": (c_int, [c_vp, c_double, c_double, c_int, P(c_cp), p_int]),
    # hamiltonian
    "qkrylov_hamiltonian_create": (c_vp, [c_vp, c_vp, c_vp]),
    "qkrylov_hamiltonian_destroy": (None, [c_vp]),
    "qkrylov_hamiltonian_dimension": (c_u64, [c_vp]),
    "qkrylov_hamiltonian_apply": (c_int, [c_vp, p_flt, p_flt, p_flt, p_flt]),
    "qkrylov_hamiltonian_apply_complex": (c_int, [c_vp, p_flt, p_flt]),
    "qkrylov_hamiltonian_diagonal": (c_int, [c_vp, p_flt]),
    "qkrylov_hamiltonian_apply_f64": (c_int, [c_vp, p_dbl, p_dbl, p_dbl, p_dbl]),
    "qkrylov_hamiltonian_apply_complex_f64": (c_int, [c_vp, p_dbl, p_dbl]),
    "qkrylov_hamiltonian_diagonal_f64": (c_int, [c_vp, p_dbl]),
    # solvers FP32
    "qkrylov_lanczos_ground_state": (c_int, [c_vp, c_int, c_float, P(LanczosResultC)]),
    "qkrylov_lanczos_ground_state_complex": (c_int, [c_vp, c_int, c_float, P(LanczosResultC), p_flt]),
    "qkrylov_davidson_lowest_complex": (c_int, [c_vp, c_int, c_int, c_float, p_flt, p_flt, P(DavidsonResultC)]),
    "qkrylov_continued_fraction_coeffs_complex": (c_int, [c_vp, p_flt, c_int, p_flt, p_flt, p_flt, p_int]),
    "qkrylov_evaluate_spectral_function": (c_float, [p_flt, p_flt, c_size, c_float, c_float, c_float, c_float]),
    "qkrylov_ftlm": (c_int, [c_vp, c_float, c_int, c_int, P(FTLMResultC)]),
    # solvers FP64
    "qkrylov_lanczos_ground_state_f64": (c_int, [c_vp, c_int, c_double, P(LanczosResultF64), p_dbl]),
    "qkrylov_davidson_lowest_complex_f64": (c_int, [c_vp, c_int, c_int, c_double, p_dbl, p_dbl,
                                                    P(DavidsonResultF64)]),
    "qkrylov_continued_fraction_coeffs_complex_f64": (c_int, [c_vp, p_dbl, c_int, p_dbl, p_dbl, p_dbl, p_int]),
    "qkrylov_evaluate_spectral_function_f64": (c_double, [p_dbl, p_dbl, c_size, c_double, c_double, c_double,
                                                          c_double]),
    "qkrylov_ftlm_f64": (c_int, [c_vp, c_double, c_int, c_int, P(FTLMResultF64)]),
    # extensions
    "qkrylov_amd_last_error": (c_cp, []),
    "qkrylov_amd_device_count": (c_int, []),
    "qkrylov_amd_set_device": (c_int, [c_int]),
    "qkrylov_amd_malloc": (c_vp, [c_size]),
    "qkrylov_amd_free": (c_int, [c_vp]),
    "qkrylov_amd_memcpy_h2d": (c_int, [c_vp, c_vp, c_size]),
    "qkrylov_amd_memcpy_d2h": (c_int, [c_vp, c_vp, c_size]),
    "qkrylov_amd_synchronize": (c_int, []),
    "qkrylov_amd_fill_random": (c_int, [c_vp, c_size, c_u64, c_int]),
    "qkrylov_amd_copy": (c_int, [c_vp, c_vp, c_size]),
    "qkrylov_amd_gather_probe": (c_int, [c_size, c_int, c_int, p_dbl, p_dbl]),
    "qkrylov_amd_set_start_mode": (c_int, [c_int]),
    "qkrylov_amd_get_start_mode": (c_int, []),
    "qkrylov_amd_reference_start_vector": (c_int, [c_size, c_u64, c_int, p_dbl]),
    "qkrylov_amd_reference_normal_vector": (c_int, [c_size, c_u64, c_int, p_dbl]),
    "qkrylov_hamiltonian_create_ex": (c_vp, [c_vp, c_vp, c_vp, c_int]),
    "qkrylov_operator_create": (c_vp, [c_vp, c_vp, c_vp]),
    "qkrylov_hamiltonian_info": (c_int, [c_vp, P(HamInfo)]),
    "qkrylov_hamiltonian_set_tuning": (c_int, [c_vp, c_int, c_int]),
    "qkrylov_hamiltonian_get_tuning": (c_int, [c_vp, c_int, P(c_int)]),
    "qkrylov_hamiltonian_autotune": (c_int, [c_vp, c_int, c_int, P(TuneReport)]),
    "qkrylov_hamiltonian_apply_device_f64": (c_int, [c_vp, c_vp, c_vp, c_int, c_vp]),
    "qkrylov_hamiltonian_apply_columns_device_f64": (c_int, [c_vp, c_vp, c_vp, c_int, p_dbl, c_vp]),
    "qkrylov_hamiltonian_apply_device_complex_f64": (c_int, [c_vp, c_vp, c_vp, c_vp]),
    "qkrylov_hamiltonian_diagonal_device_f64": (c_int, [c_vp, c_vp, c_int, c_vp]),
    "qkrylov_hamiltonian_row_entries": (c_int, [c_vp, c_u64, c_u64, c_int, p_i64, p_dbl, p_int]),
    "qkrylov_hamiltonian_states": (c_int, [c_vp, c_u64, c_u64, p_u64]),
    "qkrylov_hamiltonian_time_apply": (c_int, [c_vp, c_int, c_int, c_int, p_dbl]),
    "qkrylov_lanczos_run_f64": (c_int, [c_vp, P(LanczosOptions), p_dbl, p_dbl, p_dbl, p_dbl, P(LanczosReport)]),
    "qkrylov_lanczos_run_noalloc_f64": (c_int, [c_vp, P(LanczosOptions), p_dbl, p_dbl, p_dbl, p_dbl,
                                                P(LanczosReport)]),
    "qkrylov_ftlm_batched_f64": (c_int, [c_vp, c_int, c_int, c_u64, p_dbl, c_int, c_int, p_dbl, p_dbl, p_dbl,
                                         p_dbl, p_dbl, p_int, p_dbl]),
    "qkrylov_cf_batched_f64": (c_int, [c_vp, c_int, p_dbl, c_int, p_dbl, p_dbl, p_dbl, p_int, p_dbl]),
    "qkrylov_serve_run": (c_int, [P(ServeRequest), c_int, P(ServeOptions), P(ServeResult), p_dbl, p_dbl, c_int,
                                  p_dbl]),
    "qkrylov_structure_factor_f64": (c_int, [c_vp, P(c_vp), c_int, p_dbl, c_int, c_double, c_int, p_dbl, c_int,
                                             c_double, p_dbl, p_dbl, p_dbl, p_dbl, p_int, p_dbl]),
    "qkrylov_hamiltonian_split_info": (c_int, [c_vp, c_int, p_int, p_int, p_int]),
    "qkrylov_hamiltonian_apply_split_device_f64": (c_int, [c_vp, c_vp, c_vp, c_vp]),
    "qkrylov_hamiltonian_time_split_apply": (c_int, [c_vp, c_int, c_int, p_dbl]),
    # host eigensolvers
    "qkrylov_amd_tridiag_eigen": (c_int, [c_int, p_dbl, p_dbl, p_dbl, p_dbl, c_int]),
    "qkrylov_amd_tridiag_lowest": (c_int, [c_int, p_dbl, p_dbl, p_dbl]),
    "qkrylov_amd_hermitian_eigen": (c_int, [c_int, p_dbl, p_dbl, p_dbl]),
}

_lib = None


def load():
    """Load libqkrylov_amd.so and declare every exported signature."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is not built: run `python qkrylov_amd/build.py` (no CPU fallback exists)")
    L = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def last_error():
    return load().qkrylov_amd_last_error().decode(errors="replace")


def check(rc, what=""):
    if rc != 0:
        raise RuntimeError(f"{what} failed ({rc}): {last_error()}")
    return rc
