"""ctypes bindings of libgq.so (the C-ABI declared in include/gq_api.h).

The library is built in-tree by `__graft_entry__.build()` (or
`make -C gliomauq_amd/csrc`).  Loading fails loudly if it is missing: there is
no CPU fallback for any entry point.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libgq.so")

GQ_OK = 0
STATUS = {0: "ok", 1: "invalid argument", 2: "HIP runtime error", 3: "out of memory", 4: "invalid state",
          5: "grid topology error", 6: "unsupported configuration", 7: "non-finite parameter"}

FIELD_PHI, FIELD_PWM, FIELD_PGM = 0, 1, 2

EXPORTS = [
    "gq_ctx_create", "gq_ctx_destroy", "gq_status_string", "gq_last_error", "gq_grid_create",
    "gq_grid_block_count", "gq_grid_blocks_info", "gq_grid_write_field", "gq_grid_read_field", "gq_fk_steps",
    "gq_fwt_details", "gq_refine_blocks", "gq_compress_blocks", "gq_auto_refine", "gq_auto_compress",
    "gq_dump_uniform", "gq_obs_ref_upload", "gq_loglik_ref", "gq_loglik_ref_volume", "gq_obs_flip_upload", "gq_loglik_flip",
    "gq_last_step_timing", "gq_set_step_timing", "gq_step_timing_drain", "gq_memory_info", "gq_set_autotune", "gq_autotune_info",
    "gq_set_step_coeffs", "gq_set_step_fusion", "gq_set_step_persistent", "gq_persistent_info",
    "gq_host_stats", "gq_graph_info", "gq_set_step_graphs", "gq_topo_create", "gq_topo_destroy",
    "gq_topo_leaves", "gq_topo_refine", "gq_topo_compress", "gq_topo_point_terms",
    "gq_grid_write_volume", "gq_ensemble_create", "gq_ensemble_reset", "gq_ensemble_size", "gq_sample_select",
    "gq_ensemble_block_counts", "gq_ensemble_run", "gq_ensemble_stats", "gq_ensemble_loglik_flip",
    "gq_ensemble_ic_gaussian", "gq_adjoint_flip", "gq_loglik_flip_volume_noalloc",
    "gq_sur_create", "gq_sur_destroy", "gq_sur_last_error", "gq_sur_set_anatomy", "gq_sur_param_count",
    "gq_sur_get_params", "gq_sur_set_params", "gq_sur_predict", "gq_sur_loss_grad", "gq_sur_train_step",
    "gq_sur_timing", "gq_sur_obs_flip_upload", "gq_sur_loglik_flip",
]

RUN_GRAPH = 1


class BlockInfo(C.Structure):
    """gq_block_info = MRAG::BlockInfo of a 16^3 block."""
    _fields_ = [("block_id", C.c_int32), ("index", C.c_int16 * 3), ("level", C.c_int16),
                ("origin", C.c_float * 3), ("h", C.c_float)]


class FKParams(C.Structure):
    _fields_ = [("Dw", C.c_double), ("Dg", C.c_double), ("rho", C.c_double), ("dt", C.c_double)]


class RunOpts(C.Structure):
    """gq_run_opts"""
    _fields_ = [("adapt_every", C.c_int32), ("max_level", C.c_int32), ("tol_refine", C.c_double),
                ("tol_compress", C.c_double), ("flags", C.c_int32)]


class GQError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__("%s (%d): %s" % (STATUS.get(status, "?"), status, msg))
        self.status = status


_LIB = None


def load():
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("libgq.so is not built (%s): run __graft_entry__.build()" % LIB_PATH)
    L = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    i32p = C.POINTER(C.c_int32)
    f32p = C.POINTER(C.c_float)
    f64p = C.POINTER(C.c_double)
    u8p = C.POINTER(C.c_uint8)
    ip = C.POINTER(C.c_int)
    sig = {
        "gq_ctx_create": [C.POINTER(vp), C.c_int],
        "gq_ctx_destroy": [vp],
        "gq_grid_create": [vp, C.c_int, C.c_int, C.c_int],
        "gq_grid_block_count": [vp, ip],
        "gq_grid_blocks_info": [vp, C.POINTER(BlockInfo), C.c_int, ip],
        "gq_grid_write_field": [vp, C.c_int, f32p, C.c_int],
        "gq_grid_read_field": [vp, C.c_int, f32p, C.c_int],
        "gq_fk_steps": [vp, C.POINTER(FKParams), C.c_int],
        "gq_fwt_details": [vp, f64p, C.c_int],
        "gq_refine_blocks": [vp, i32p, C.c_int, ip],
        "gq_compress_blocks": [vp, i32p, C.c_int, ip],
        "gq_auto_refine": [vp, C.c_double, C.c_int, C.c_int, ip],
        "gq_auto_compress": [vp, C.c_double, C.c_int, ip],
        "gq_dump_uniform": [vp, f64p, C.c_size_t],
        "gq_obs_ref_upload": [vp, C.c_int, f64p, f64p, f64p, i32p, C.c_int],
        "gq_loglik_ref": [vp, f64p, C.c_int, C.c_int, C.c_int, f64p],
        "gq_loglik_ref_volume": [vp, f64p, C.c_int, f64p, C.c_int, C.c_int, C.c_int, f64p],
        "gq_obs_flip_upload": [vp, C.c_int, u8p, u8p, f32p],
        "gq_loglik_flip": [vp, f64p, f64p],
        "gq_last_step_timing": [vp, f64p, ip, C.POINTER(C.c_int64)],
        "gq_set_step_timing": [vp, C.c_int],
        "gq_step_timing_drain": [vp, f32p, C.POINTER(C.c_int64), C.c_int, ip],
        "gq_memory_info": [vp, C.POINTER(C.c_int64), ip],
        "gq_set_autotune": [vp, C.c_int, C.c_int],
        "gq_autotune_info": [vp, ip, ip, f64p],
        "gq_set_step_coeffs": [vp, C.c_int],
        "gq_set_step_fusion": [vp, C.c_int],
        "gq_set_step_persistent": [vp, C.c_int],
        "gq_persistent_info": [vp, ip, f64p],
        "gq_host_stats": [vp, f64p],
        "gq_graph_info": [vp, C.POINTER(C.c_int64), C.POINTER(C.c_int64), ip],
        "gq_set_step_graphs": [vp, C.c_int],
        "gq_topo_create": [C.POINTER(vp), C.c_int, C.c_int, C.c_int],
        "gq_topo_destroy": [vp],
        "gq_topo_leaves": [vp, i32p, i32p, i32p, C.c_int, ip],
        "gq_topo_refine": [vp, i32p, C.c_int, ip],
        "gq_topo_compress": [vp, i32p, C.c_int, ip],
        "gq_topo_point_terms": [vp, C.c_int, i32p, C.c_int, i32p, i32p, f64p, ip],
        "gq_grid_write_volume": [vp, C.c_int, f32p, C.c_int, C.c_int],
        "gq_ensemble_create": [vp, C.c_int, C.c_int, C.c_int, C.c_int],
        "gq_ensemble_size": [vp, ip],
        "gq_ensemble_reset": [vp],
        "gq_sample_select": [vp, C.c_int],
        "gq_ensemble_block_counts": [vp, i32p],
        "gq_ensemble_run": [vp, C.POINTER(FKParams), i32p, C.POINTER(RunOpts)],
        "gq_ensemble_stats": [vp, f64p],
        "gq_ensemble_loglik_flip": [vp, f64p, f64p],
        "gq_ensemble_ic_gaussian": [vp, f64p],
        "gq_adjoint_flip": [vp, C.POINTER(FKParams), C.c_int, C.c_int, f64p, f64p, f64p, f64p],
        "gq_loglik_flip_volume_noalloc": [vp, C.c_int, vp, vp, vp, C.c_int, f64p, vp, vp],
    }
    for name, args in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = C.c_int
    L.gq_status_string.argtypes = [C.c_int]
    L.gq_status_string.restype = C.c_char_p
    L.gq_last_error.argtypes = [vp]
    L.gq_last_error.restype = C.c_char_p
    _LIB = L
    return L


def ptr(a, ctype):
    return a.ctypes.data_as(C.POINTER(ctype))
