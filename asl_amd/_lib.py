"""ctypes binding of ``libaslhip.so`` (the C ABI declared in include/aslhip.h).

The library is the product: there is no fallback.  If it is missing or does
not load, every entry point raises.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libaslhip.so")

# constants of include/aslhip.h
F32, F64, BF16, F16 = 0, 1, 2, 3
X0, XE, Y0, YE, Z0, ZE = 1, 2, 4, 8, 16, 32
BC_NOSLIP, BC_CONSTANT_PRESSURE, BC_CONSTANT_VELOCITY, BC_CONSTANT_PRESSURE_VELOCITY = 1, 2, 3, 4
BC_NOSLIP_MAP, BC_PRESSURE_VELOCITY_MAP, BC_VELOCITY_MAP = 5, 6, 7
STEP_WRITE_MACRO = 1
STEP_DEFER = 2
STEP_GRAPH = 4
GRAPH_CHUNK = 100
CODE_FLUID, CODE_FLUID_BC, CODE_SKIP, CODE_GHOST = 0, 1, 2, 3
TUNE_XCD_REMAP = 1
TUNE_BLOCK = 2
TUNE_GROUP_MB = 3
TUNE_WIDE = 4
TUNE_CPL = 5
TUNE_LDS = 6
DF_SPHERE, DF_CYLINDER, DF_CONE, DF_PLANE, DF_AND, DF_OR, DF_NEG, DF_NORMALIZE = 1, 2, 3, 4, 5, 6, 7, 8
CLOSURE_NPARAM = 1345
CLOSURE_FP32, CLOSURE_BF16_MFMA = 0, 1
VIS_NOSLIP_VEL, VIS_NOSLIP_RHO, VIS_VELOCITY = 1, 2, 3
NOALLOC_SCRATCH_BYTES = 16 << 20

# Every entry point of include/aslhip.h with its ctypes signature.
_vp = ctypes.c_void_p
_dp = ctypes.POINTER(ctypes.c_double)
_fp = ctypes.POINTER(ctypes.c_float)
_ip = ctypes.POINTER(ctypes.c_int)
SIGNATURES = {
    "aslhip_last_error": [ctypes.c_char_p, ctypes.c_size_t],
    "aslhip_device_info": [ctypes.c_char_p, ctypes.c_size_t, _ip, _dp],
    "aslhip_lattice_create": [ctypes.POINTER(_vp), ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                              ctypes.c_int, ctypes.c_uint],
    "aslhip_lattice_destroy": [_vp],
    "aslhip_set_viscosity": [_vp, ctypes.c_double],
    "aslhip_set_compressible": [_vp, ctypes.c_int],
    "aslhip_set_omega": [_vp, _dp],
    "aslhip_set_geometry": [_vp, ctypes.c_double, _dp],
    "aslhip_set_vmap_skip_outer": [_vp, ctypes.c_int],
    "aslhip_macro_device": [_vp, _vp, ctypes.POINTER(_vp), ctypes.POINTER(_vp)],
    "aslhip_set_turbulence": [_vp, ctypes.c_int],
    "aslhip_set_compute_macro": [_vp, ctypes.c_int, ctypes.c_int],
    "aslhip_init_equilibrium": [_vp, ctypes.c_double, _dp],
    "aslhip_init_equilibrium_field": [_vp, _vp, _vp],
    "aslhip_bc_add_slice": [_vp, ctypes.c_int, ctypes.c_uint, ctypes.c_double, _dp, _ip],
    "aslhip_bc_add_map": [_vp, ctypes.c_int, _fp, _fp, ctypes.c_double, _dp, _dp, _ip],
    "aslhip_bc_execute": [_vp, ctypes.c_int],
    "aslhip_finalize": [_vp],
    "aslhip_step": [_vp, ctypes.c_int, ctypes.c_uint],
    "aslhip_synchronize": [_vp],
    "aslhip_read_rho": [_vp, _vp],
    "aslhip_read_velocity": [_vp, _vp],
    "aslhip_read_f": [_vp, _vp],
    "aslhip_read_f_noalloc": [_vp, _vp],
    "aslhip_write_f": [_vp, _vp],
    "aslhip_surface_force": [_vp, _fp, _vp],
    "aslhip_surface_flux": [_vp, _fp, _vp],
    "aslhip_vis_add": [_vp, ctypes.c_int, _fp, _dp, _dp, _ip],
    "aslhip_vis_execute": [_vp, ctypes.c_int],
    "aslhip_template": [ctypes.c_int, _ip, _dp, _ip],
    "aslhip_batch_autotune_groups": [_vp, ctypes.c_int, _ip],
    "aslhip_get_tuning": [_vp, _ip, _ip],
    "aslhip_tuning_value": [_vp, ctypes.c_int, _ip],
    "aslhip_df_map": [_dp, ctypes.c_int, _ip, _dp, ctypes.c_double, ctypes.c_int, _vp],
    "aslhip_stl_count": [ctypes.c_char_p, ctypes.POINTER(ctypes.c_longlong)],
    "aslhip_stl_read": [ctypes.c_char_p, _dp, ctypes.c_longlong],
    "aslhip_vtp_count": [ctypes.c_char_p, ctypes.POINTER(ctypes.c_longlong)],
    "aslhip_vtp_read": [ctypes.c_char_p, _dp, ctypes.c_longlong],
    "aslhip_surface_distance_map": [_dp, ctypes.c_longlong, _ip, _dp, _dp, ctypes.c_double, ctypes.c_double,
                                    ctypes.c_double, ctypes.c_double, _fp],
    "aslhip_read_cell_types": [_vp, _vp],
    "aslhip_reduce_conservation": [_vp, _dp],
    "aslhip_stats": [_vp, ctypes.POINTER(ctypes.c_int64)],
    "aslhip_set_tuning": [_vp, ctypes.c_int, ctypes.c_int],
    "aslhip_autotune": [_vp, ctypes.c_int, _ip, _ip],
    "aslhip_time_steps": [_vp, ctypes.c_int, ctypes.c_uint, _dp, _dp, _dp],
    "aslhip_batch_create": [ctypes.POINTER(_vp), ctypes.POINTER(_vp), ctypes.c_int],
    "aslhip_batch_destroy": [_vp],
    "aslhip_batch_step": [_vp, ctypes.c_int, ctypes.c_uint],
    "aslhip_batch_time_steps": [_vp, ctypes.c_int, ctypes.c_uint, _dp],
    "aslhip_batch_set_tuning": [_vp, ctypes.c_int, ctypes.c_int],
    "aslhip_batch_autotune": [_vp, ctypes.c_int, _ip],
    "aslhip_closure_set": [_vp, _fp, _fp, _fp, ctypes.c_float],
    "aslhip_closure_features": [_vp, _fp],
    "aslhip_closure_set_mode": [_vp, ctypes.c_int],
    "aslhip_mlp_create": [ctypes.POINTER(_vp), _fp],
    "aslhip_mlp_destroy": [_vp],
    "aslhip_mlp_set_data": [_vp, _fp, _fp, ctypes.c_longlong],
    "aslhip_mlp_gradient": [_vp, ctypes.c_longlong, ctypes.c_int, _fp, _dp],
    "aslhip_mlp_train": [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_ulonglong, _dp],
    "aslhip_mlp_get_params": [_vp, _fp],
}

_LIB = None


class ASLError(RuntimeError):
    """Raised for a failed library call (ASL throws std::logic_error("ASL ERROR: ...") there)."""


def load_library(path=LIB_PATH):
    """Load libaslhip.so (raises if it was not built)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(path):
        raise ASLError("libaslhip.so is not built (run `make -C asl_amd` or __graft_entry__.build())")
    lib = ctypes.CDLL(path)
    for name, args in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = ctypes.c_int
    _LIB = lib
    return lib


def last_error():
    lib = load_library()
    buf = ctypes.create_string_buffer(1024)
    lib.aslhip_last_error(buf, 1024)
    return buf.value.decode(errors="replace")


def call(name, *args):
    """Call an entry point; raise ASLError with the library's message on failure."""
    lib = load_library()
    r = getattr(lib, name)(*args)
    if r != 0:
        raise ASLError("%s failed (%d): %s" % (name, r, last_error()))
    return r


def device_info():
    buf = ctypes.create_string_buffer(256)
    cu = ctypes.c_int(0)
    mem = ctypes.c_double(0)
    call("aslhip_device_info", buf, 256, ctypes.byref(cu), ctypes.byref(mem))
    return {"name": buf.value.decode(), "cu": cu.value, "hbm_bytes": mem.value}
