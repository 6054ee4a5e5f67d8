"""ctypes binding of libmesh_hip.so (the C ABI declared in include/mesh_hip.h).

The shared library is built in-tree by ``mesh_amd.build``.  Loading fails
loudly if it is missing: there is no CPU fallback.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
# MESH_HIP_LIB selects another build of the same library (e.g. one built with
# a different site-block length); the default is the in-tree build.
LIB_PATH = os.environ.get("MESH_HIP_LIB") or os.path.join(HERE, "libmesh_hip.so")

MESH_OK, MESH_E_ARG, MESH_E_SIZE, MESH_E_ALLOC, MESH_E_HIP, MESH_E_STATE = 0, -1, -2, -3, -4, -5
F64, F32 = 0, 1


class MeshError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"mesh error {code}: {msg}")
        self.code = code


class mesh_grid(C.Structure):
    _fields_ = [("nd", C.c_int), ("dims", C.c_int * 4)]


_lib = None

vp = C.c_void_p
dp = C.POINTER(C.c_double)
gp = C.POINTER(mesh_grid)
u64, u32, i32, f64 = C.c_uint64, C.c_uint32, C.c_int, C.c_double

_SIGS = {
    "mesh_ctx_create": (i32, [C.POINTER(vp), i32]),
    "mesh_ctx_destroy": (i32, [vp]),
    "mesh_last_error": (C.c_char_p, [vp]),
    "mesh_synchronize": (i32, [vp]),
    "mesh_stream": (vp, [vp]),
    "mesh_set_option": (i32, [vp, C.c_char_p, C.c_long]),
    "mesh_profile_enable": (i32, [vp, i32]),
    "mesh_profile_reset": (i32, [vp]),
    "mesh_profile_read": (i32, [vp, C.c_char_p, dp, C.POINTER(C.c_longlong)]),
    "mesh_malloc": (i32, [vp, C.POINTER(vp), C.c_size_t]),
    "mesh_free": (i32, [vp, vp]),
    "mesh_memcpy": (i32, [vp, vp, vp, C.c_size_t]),
    "mesh_memcpy_2d": (i32, [vp, vp, C.c_size_t, vp, C.c_size_t, C.c_size_t, C.c_size_t]),
    "mesh_memclear": (i32, [vp, vp, C.c_size_t]),
    "mesh_memcpy_to_host": (i32, [vp, vp, vp, C.c_size_t]),
    "mesh_memcpy_from_host": (i32, [vp, vp, vp, C.c_size_t]),
    "mesh_stream_copy": (i32, [vp, vp, vp, C.c_size_t]),
    "mesh_volume": (C.c_longlong, [gp]),
    "mesh_site_stride": (C.c_longlong, [gp]),
    "mesh_site_block": (i32, []),
    "mesh_su3_field_bytes": (C.c_size_t, [gp, i32, i32]),
    "mesh_su3_upload": (i32, [vp, gp, i32, i32, vp, vp]),
    "mesh_su3_download": (i32, [vp, gp, i32, i32, vp, vp]),
    "mesh_su3_convert": (i32, [vp, gp, i32, i32, vp, i32, vp]),
    "mesh_su3_unit": (i32, [vp, gp, i32, i32, vp]),
    "mesh_su3_random_gauge": (i32, [vp, gp, i32, i32, vp, u64, u32]),
    "mesh_su3_momenta": (i32, [vp, gp, i32, vp, vp, u64, u32]),
    "mesh_su3_plaquette": (i32, [vp, gp, i32, vp, dp]),
    "mesh_su3_plaquette_noalloc": (i32, [vp, gp, i32, vp, vp]),
    "mesh_su2_field_bytes": (C.c_size_t, [gp, i32, i32]),
    "mesh_su2_transfer": (i32, [vp, gp, i32, i32, vp, vp, i32]),
    "mesh_su2_random": (i32, [vp, gp, i32, i32, vp, u64]),
    "mesh_su2_unit": (i32, [vp, gp, i32, i32, vp]),
    "mesh_su2_momenta": (i32, [vp, gp, i32, vp, vp, u64, u32]),
    "mesh_su2_force_update": (i32, [vp, gp, i32, vp, vp, f64, f64]),
    "mesh_su2_link_update": (i32, [vp, gp, i32, i32, vp, vp, vp, f64, i32]),
    "mesh_su2_reunit": (i32, [vp, gp, i32, i32, vp]),
    "mesh_su2_plaquette": (i32, [vp, gp, i32, vp, dp]),
    "mesh_su2_delta_h": (i32, [vp, gp, i32, vp, vp, vp, vp, f64, dp]),
    "mesh_su2_hmc_trajectory": (i32, [vp, gp, i32, C.POINTER(vp), C.POINTER(vp), vp, vp, f64, dp, i32, u64, u32,
                                      i32, dp]),
    "mesh_su3_float_energy": (i32, [vp, gp, vp, vp, i32, f64, dp]),
    "mesh_su2_landau_functional": (i32, [vp, gp, i32, vp, vp, dp]),
    "mesh_su2_landau_divsum": (i32, [vp, gp, i32, vp, vp, dp]),
    "mesh_su2_landau_deriv": (i32, [vp, gp, i32, vp, vp, vp]),
    "mesh_su2_gauge_transform": (i32, [vp, gp, i32, vp, vp, vp]),
    "mesh_su2_landau_or_halfsweep": (i32, [vp, gp, i32, vp, vp, i32, f64]),
    "mesh_su2_dot": (i32, [vp, gp, i32, i32, vp, vp, dp]),
    "mesh_su2_axpy": (i32, [vp, gp, i32, i32, vp, f64, vp]),
    "mesh_u1_landau_functional": (i32, [vp, gp, vp, vp, dp]),
    "mesh_u1_landau_divsum": (i32, [vp, gp, vp, vp, dp]),
    "mesh_u1_landau_deriv": (i32, [vp, gp, vp, vp, vp]),
    "mesh_u1_gauge_transform": (i32, [vp, gp, vp, vp, vp]),
    "mesh_u1_landau_or_halfsweep": (i32, [vp, gp, vp, vp, i32, f64]),
    "mesh_u1_gauge_step": (i32, [vp, gp, vp, vp, vp, f64]),
    "mesh_u1_dot": (i32, [vp, C.c_long, vp, vp, dp]),
    "mesh_u1_axpy": (i32, [vp, C.c_long, vp, f64, vp]),
    "mesh_su3_wilson_action": (i32, [vp, gp, i32, vp, f64, dp]),
    "mesh_su3_norm2": (i32, [vp, gp, i32, i32, vp, dp]),
    "mesh_su3_sum_real_trace_prod": (i32, [vp, gp, i32, i32, vp, vp, dp]),
    "mesh_su3_staple": (i32, [vp, gp, i32, vp, i32, vp]),
    "mesh_su3_force": (i32, [vp, gp, i32, vp, i32, f64, vp]),
    "mesh_su3_force_update": (i32, [vp, gp, i32, vp, vp, f64, f64]),
    "mesh_su3_link_update": (i32, [vp, gp, i32, vp, vp, vp, f64, i32]),
    "mesh_su3_reunit": (i32, [vp, gp, i32, i32, vp]),
    "mesh_su3_exp_mul": (i32, [vp, gp, i32, i32, vp, vp, f64]),
    "mesh_su3_axpy": (i32, [vp, gp, i32, i32, vp, f64, vp]),
    "mesh_su3_cshift": (i32, [vp, gp, i32, i32, vp, vp, i32, i32]),
    "mesh_su3_sum": (i32, [vp, gp, i32, i32, vp, dp]),
    "mesh_su3_delta_h": (i32, [vp, gp, i32, vp, vp, vp, vp, f64, dp]),
    "mesh_make_deltas": (i32, [C.c_char_p, f64, i32, dp, i32]),
    "mesh_su3_hmc_trajectory": (i32, [vp, gp, i32, C.POINTER(vp), C.POINTER(vp), vp, vp, f64, dp, i32, u64,
                                      u32, i32, dp]),
    "mesh_landau_functional": (i32, [vp, gp, i32, vp, vp, dp]),
    "mesh_landau_divsum": (i32, [vp, gp, i32, vp, vp, dp]),
    "mesh_landau_deriv": (i32, [vp, gp, i32, vp, vp, vp]),
    "mesh_landau_or_halfsweep": (i32, [vp, gp, i32, vp, vp, i32, f64]),
    "mesh_su3_gauge_transform": (i32, [vp, gp, i32, vp, vp, vp]),
    "mesh_u1_random": (i32, [vp, gp, i32, vp, u64]),
    "mesh_u1_observables": (i32, [vp, gp, i32, vp, vp]),
    "mesh_u1_hmc": (i32, [vp, gp, i32, vp, dp, i32, f64, u64, u32, i32, vp]),
    "mesh_flow_num_params": (i32, [i32]),
    "mesh_flow_init": (i32, [vp, i32, vp, u64]),
    "mesh_flow_sample": (i32, [vp, gp, i32, i32, vp, u64, u32, vp, vp, vp]),
    "mesh_flow_logq": (i32, [vp, gp, i32, i32, vp, vp, vp, vp]),
    "mesh_flow_workspace_bytes": (C.c_size_t, [gp, i32, i32]),
    "mesh_flow_gradient": (i32, [vp, gp, i32, i32, vp, u64, u32, f64, vp, vp, vp]),
    "mesh_flow_adam": (i32, [vp, i32, vp, vp, vp, vp, i32, f64]),
    "mesh_flow_mh": (i32, [vp, gp, i32, vp, vp, vp, vp, f64, u64, u32, vp]),
    "mesh_pflow_num_params": (i32, []),
    "mesh_pflow_init": (i32, [vp, vp]),
    "mesh_pflow_sample": (i32, [vp, gp, i32, vp, u64, u32, vp]),
    "mesh_pflow_logq": (i32, [vp, gp, i32, vp, vp, vp]),
    "mesh_pflow_propose": (i32, [vp, gp, i32, vp, vp, vp, u64, u32, vp, vp]),
    "mesh_pflow_workspace_bytes": (C.c_size_t, [gp, i32]),
    "mesh_pflow_gradient": (i32, [vp, gp, i32, vp, u64, u32, f64, vp, vp, vp]),
    "mesh_pflow_adam": (i32, [vp, vp, vp, vp, vp, i32, f64]),
    "mesh_pcflow_num_params": (i32, [i32]),
    "mesh_pcflow_init": (i32, [vp, i32, vp, u64]),
    "mesh_pcflow_sample": (i32, [vp, gp, i32, i32, vp, u64, u32, vp, vp]),
    "mesh_pcflow_logq": (i32, [vp, gp, i32, i32, vp, vp, vp, vp]),
    "mesh_pcflow_propose": (i32, [vp, gp, i32, i32, vp, vp, vp, u64, u32, vp, vp]),
    "mesh_pcflow_workspace_bytes": (C.c_size_t, [gp, i32, i32]),
    "mesh_pcflow_gradient": (i32, [vp, gp, i32, i32, vp, u64, u32, f64, vp, vp, vp]),
    "mesh_pcflow_adam": (i32, [vp, i32, vp, vp, vp, vp, i32, f64]),
}


def lib():
    """Load libmesh_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with `python -m mesh_amd.build`")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def exported_symbols():
    return list(_SIGS)
