"""ctypes binding of libself_amd.so (the C ABI declared in include/self_amd.h).

The library is built in-tree (``self_amd/libself_amd.so``) by ``__graft_entry__.build()``
or ``make -C self_amd/csrc``.  Loading fails loudly when it is missing: there is no
fallback implementation.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libself_amd.so")
IO_LIB_PATH = os.path.join(_HERE, "libself_amd_io.so")

c_int, c_double, c_void_p = ctypes.c_int, ctypes.c_double, ctypes.c_void_p
c_i32p = ctypes.POINTER(ctypes.c_int32)
c_dp = ctypes.POINTER(ctypes.c_double)

# (name, argtypes) for every exported entry point; the header is the source of truth and
# tests/test_abi.py checks that each declared symbol is listed and exported.
c_size_t = ctypes.c_size_t

SIGNATURES = {
    "selfamd_last_error": [],
    "selfamd_device_count": [c_void_p],
    "selfamd_device_name": [c_int, ctypes.c_char_p, c_int],
    "selfamd_malloc": [c_void_p, c_size_t],
    "selfamd_free": [c_void_p],
    "selfamd_memcpy": [c_void_p, c_void_p, c_size_t, c_int],
    "selfamd_memset": [c_void_p, c_int, c_size_t],
    "selfamd_device_synchronize": [],
    "selfamd_copy_f64": [c_void_p, c_void_p, c_size_t, c_void_p],
    "selfamd_time_copy_f64": [c_void_p, c_void_p, c_size_t, c_int, c_void_p],
    "selfamd_version": [],
    "selfamd_lagrange_init": [c_int, c_int, c_int, c_int] + [c_void_p] * 8,
    "selfamd_mesh2d_structured": [c_int, c_int, c_double, c_double, c_double, c_double, c_int, c_int, c_int,
                                  c_void_p, c_void_p, c_void_p],
    "selfamd_mesh2d_recalculate_flip": [c_int, c_int, c_void_p, c_void_p],
    "selfamd_mesh2d_reset_boundary_condition_type": [c_int, c_void_p, c_int],
    "selfamd_mesh2d_from_hopr": [c_int, c_int] + [c_void_p] * 6,
    "selfamd_scalar_boundary_interp_2d": [c_void_p] * 3 + [c_int] * 3 + [c_void_p],
    "selfamd_vector_boundary_interp_2d": [c_void_p] * 3 + [c_int] * 3 + [c_void_p],
    "selfamd_tensor_boundary_interp_2d": [c_void_p] * 3 + [c_int] * 3 + [c_void_p],
    "selfamd_scalar_grid_interp_2d": [c_void_p] * 3 + [c_int] * 4 + [c_void_p],
    "selfamd_vector_grid_interp_2d": [c_void_p] * 3 + [c_int] * 4 + [c_void_p],
    "selfamd_vector_gradient_2d": [c_void_p] * 3 + [c_int] * 3 + [c_void_p],
    "selfamd_vector_divergence_2d": [c_void_p] * 3 + [c_int] * 3 + [c_void_p],
    "selfamd_vector_dg_divergence_2d": [c_void_p] * 6 + [c_int] * 3 + [c_void_p],
    "selfamd_determinant_tensor_2d": [c_void_p] * 2 + [c_int] * 3 + [c_void_p],
    "selfamd_side_exchange_scalar_2d": [c_void_p] * 3 + [c_int] * 3 + [c_void_p],
    "selfamd_side_exchange_vector_2d": [c_void_p] * 3 + [c_int] * 3 + [c_void_p],
    "selfamd_gradient_sf_2d": [c_void_p] * 5 + [c_int] * 3 + [c_void_p],
    "selfamd_jacobian_weight_2d": [c_void_p] * 2 + [c_int] * 3 + [c_void_p],
    "selfamd_contravariant_projection_2d": [c_void_p] * 2 + [c_int] * 3 + [c_void_p],
    "selfamd_update_grk": [c_void_p] * 3 + [c_double] * 3 + [c_int] * 3 + [c_void_p],
    "selfamd_calculate_dsdt": [c_void_p] * 3 + [c_int] * 3 + [c_void_p],
    "selfamd_semquad_generate": [c_int] * 4 + [c_void_p] * 11 + [c_void_p],
    "selfamd_sw2d_velocity": [c_void_p] * 4 + [c_int] * 2 + [c_void_p],
    "selfamd_sw2d_set_boundary_condition": [c_void_p] * 5 + [c_int] * 2 + [c_void_p],
    "selfamd_sw2d_source": [c_void_p] * 3 + [c_double] * 2 + [c_int] * 2 + [c_void_p],
    "selfamd_sw2d_flux": [c_void_p] * 3 + [c_double] + [c_int] * 2 + [c_void_p],
    "selfamd_sw2d_riemann": [c_void_p] * 7 + [c_double] + [c_int] * 2 + [c_void_p],
    "selfamd_lsw2d_set_boundary_condition": [c_void_p] * 4 + [c_int] * 2 + [c_void_p],
    "selfamd_lsw2d_source": [c_void_p] * 3 + [c_int] * 2 + [c_void_p],
    "selfamd_lsw2d_flux": [c_void_p] * 3 + [c_double] + [c_int] * 2 + [c_void_p],
    "selfamd_lsw2d_riemann": [c_void_p] * 6 + [c_double] + [c_int] * 2 + [c_void_p],
    "selfamd_euler2d_diagnostics": [c_void_p] * 2 + [c_double] * 2 + [c_int] * 2 + [c_void_p],
    "selfamd_euler2d_primitive": [c_void_p] * 2 + [c_double] + [c_int] * 2 + [c_void_p],
    "selfamd_euler2d_entropy_vars": [c_void_p] * 2 + [c_double] + [c_int] * 2 + [c_void_p],
    "selfamd_euler2d_set_boundary_condition": [c_void_p] * 11 + [c_int] * 2 + [c_void_p],
    "selfamd_euler2d_source": [c_void_p] * 4 + [c_int] * 2 + [c_void_p],
    "selfamd_euler2d_flux": [c_void_p] * 4 + [c_int] * 2 + [c_void_p],
    "selfamd_euler2d_riemann": [c_void_p] * 9 + [c_int] * 2 + [c_void_p],
    "selfamd_model2d_nvar": [c_int],
    "selfamd_model2d_create": [c_void_p, c_int, c_int, c_int] + [c_void_p] * 5 + [c_void_p],
    "selfamd_model2d_destroy": [c_void_p],
    "selfamd_model2d_set_param": [c_void_p, c_int, c_double],
    "selfamd_model2d_set_field": [c_void_p, c_int, c_void_p],
    "selfamd_model2d_set_solution": [c_void_p, c_void_p],
    "selfamd_model2d_get_solution": [c_void_p, c_void_p],
    "selfamd_model2d_tendency": [c_void_p, c_void_p],
    "selfamd_model2d_flux_registers": [c_void_p, c_void_p, c_void_p],
    "selfamd_model2d_entropy": [c_void_p, c_void_p],
    "selfamd_model2d_integral": [c_void_p, c_int, c_void_p],
    "selfamd_model2d_forward_step": [c_void_p, c_int, c_double],
    "selfamd_model2d_synchronize": [c_void_p],
    "selfamd_model2d_time_steps": [c_void_p, c_int, c_double, c_void_p],
    "selfamd_model2d_stream": [c_void_p, c_void_p],
    "selfamd_model2d_stage": [c_void_p, c_int, c_double, c_void_p, c_void_p],
    "selfamd_model2d_tendency_vjp": [c_void_p] * 7,
    "selfamd_model2d_closure_vjp": [c_void_p] * 4,
    "selfamd_model2d_integral_noalloc": [c_void_p, c_int, c_void_p],
    "selfamd_model2d_get_time": [c_void_p, c_void_p],
    "selfamd_model2d_set_time": [c_void_p, c_double],
    "selfamd_batch_create": [c_void_p, c_int, c_void_p, c_void_p],
    "selfamd_batch_forward_step": [c_void_p, c_int],
    "selfamd_batch_time_steps": [c_void_p, c_int, c_void_p],
    "selfamd_batch_scatter": [c_void_p],
    "selfamd_batch_info": [c_void_p, c_void_p, c_void_p],
    "selfamd_batch_destroy": [c_void_p],
    "selfamd_batch_set_mode": [c_void_p, c_int],
    "selfamd_mlp_create": [c_void_p, c_int, c_int, c_int, c_int],
    "selfamd_mlp_destroy": [c_void_p],
    "selfamd_mlp_num_params": [c_void_p, c_void_p],
    "selfamd_mlp_set_params": [c_void_p, c_void_p],
    "selfamd_mlp_get_params": [c_void_p, c_void_p],
    "selfamd_mlp_get_grads": [c_void_p, c_void_p],
    "selfamd_mlp_forward": [c_void_p, c_void_p, c_void_p, c_int, c_void_p],
    "selfamd_mlp_loss_grad": [c_void_p, c_void_p, c_void_p, c_int, c_void_p, c_void_p],
    "selfamd_mlp_adam_step": [c_void_p, ctypes.c_float, c_int, c_void_p],
    "selfamd_mlp_backward": [c_void_p, c_void_p, c_void_p, c_int, c_int, c_void_p, c_void_p],
    "selfamd_mlp_zero_grads": [c_void_p, c_void_p],
    "selfamd_project_l2_structured_2d": [c_void_p, c_void_p, c_int, c_int, c_int, c_int, c_int, c_int, c_void_p],
    "selfamd_model2d_set_closure": [c_void_p] * 6,
    "selfamd_model2d_closure_features": [c_void_p, c_void_p],
    "selfamd_model2d_set_bands": [c_void_p, c_int],
    "selfamd_tile_count": [c_int, c_int, c_void_p],
    "selfamd_tile_info": [c_int, c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p],
    "selfamd_tile_set": [c_int, c_int, c_int],
    "selfamd_tile_get": [c_int, c_int, c_void_p],
}


class SelfAmdError(RuntimeError):
    """A C entry point returned a non-zero status."""


_lib = None


def load():
    """Load libself_amd.so once; raise ImportError (no fallback) when it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"self_amd native library not found at {LIB_PATH}; build it with "
            "`python -c 'import __graft_entry__ as g; g.build()'` or `make -C self_amd/csrc`")
    lib = ctypes.CDLL(LIB_PATH)
    for name, argtypes in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = argtypes
        fn.restype = ctypes.c_char_p if name == "selfamd_last_error" else c_int
    _lib = lib
    return lib


def call(name, *args):
    """Call an entry point and raise SelfAmdError with the library's message on failure."""
    lib = load()
    rc = getattr(lib, name)(*args)
    if rc != 0:
        msg = lib.selfamd_last_error().decode(errors="replace")
        raise SelfAmdError(f"{name} failed (status {rc}): {msg}")
    return rc


# libself_amd_io.so (include/self_amd_io.h): HOPr mesh files through the HDF5 C library
IO_SIGNATURES = {
    "selfamd_io_last_error": [],
    "selfamd_hopr_attribute_int": [ctypes.c_char_p, ctypes.c_char_p, c_void_p],
    "selfamd_hopr_dataset_dims": [ctypes.c_char_p, ctypes.c_char_p, c_void_p, c_void_p],
    "selfamd_hopr_read_int": [ctypes.c_char_p, ctypes.c_char_p, c_void_p, ctypes.c_int64],
    "selfamd_hopr_read_double": [ctypes.c_char_p, ctypes.c_char_p, c_void_p, ctypes.c_int64],
    "selfamd_pickup_write": [ctypes.c_char_p, c_int, c_int, c_int] + [c_void_p] * 9,
    "selfamd_pickup_write_registers": [ctypes.c_char_p, c_int, c_int, c_int] + [c_void_p] * 5,
    "selfamd_pickup_read": [ctypes.c_char_p, c_int, c_int, c_int, c_void_p],
}
_io = None


def load_io():
    """Load libself_amd_io.so once; ImportError when it was not built (no HDF5 at build time)."""
    global _io
    if _io is not None:
        return _io
    if not os.path.exists(IO_LIB_PATH):
        raise ImportError(f"self_amd HOPr reader not found at {IO_LIB_PATH}; it is built by "
                          "`make -C self_amd/csrc` where the HDF5 C library is installed")
    lib = ctypes.CDLL(IO_LIB_PATH)
    for name, argtypes in IO_SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = argtypes
        fn.restype = ctypes.c_char_p if name == "selfamd_io_last_error" else c_int
    _io = lib
    return lib


def io_available():
    """True when libself_amd_io.so is built and loads (its HDF5 library is present)."""
    try:
        load_io()
        return True
    except (ImportError, OSError):
        return False


def call_io(name, *args):
    lib = load_io()
    rc = getattr(lib, name)(*args)
    if rc != 0:
        raise SelfAmdError(f"{name} failed (status {rc}): {lib.selfamd_io_last_error().decode(errors='replace')}")
    return rc
