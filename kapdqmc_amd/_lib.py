"""ctypes binding of libkapdqmc_hip.so (the C ABI declared in include/kapdqmc.h).

The library is built in-tree by ``kapdqmc_amd.build.build()``; importing the product API
without it raises immediately — there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libkapdqmc_hip.so")

KDQ_OK, KDQ_EARG, KDQ_EHIP, KDQ_ENUM, KDQ_ESTATE = 0, -1, -2, -3, -4
MODEL_TU, MODEL_TV, MODEL_TUV = 0, 1, 2
LATTICE_CODE = {"HoneyComb120": 0, "HoneyComb60": 1, "SQUARE90": 2, "SQUARE45": 3}
PREC_FP64, PREC_MIXED = 0, 1
OPTIONS = {"graph": 0, "delay": 1, "gemm_tile": 2, "defer": 3, "pipe": 4, "flush": 5, "aside": 6, "circ": 7, "qsplit": 8, "maside": 9, "ginv": 10, "qinv": 11, "prewrap": 12, "mchain": 13}   # kdq_set_option keys

# Every symbol include/kapdqmc.h declares (checked by tests/test_capi.py).
EXPORTS = ("kdq_create", "kdq_destroy", "kdq_last_error", "kdq_num_chains", "kdq_num_obs", "kdq_set_fields",
           "kdq_get_fields", "kdq_initial_s", "kdq_phy_update", "kdq_get_G", "kdq_set_G", "kdq_wrap",
           "kdq_update_layer", "kdq_init_chain", "kdq_measure", "kdq_measure_noalloc", "kdq_set_replay", "kdq_get_health", "kdq_check_G", "kdq_get_counters", "kdq_get_near_ties", "kdq_profile",
           "kdq_profile_read", "kdq_get_logdet", "kdq_set_option", "kdq_get_precision_error", "kdq_log_weight", "kdq_scee_create", "kdq_scee_set_fields",
           "kdq_scee_get_fields", "kdq_scee_init", "kdq_scee_sweeps", "kdq_scee_get_G4", "kdq_scee_detg",
           "kdq_dop_create", "kdq_G4", "kdq_model_build", "kdq_model_free", "kdq_net_create", "kdq_net_destroy", "kdq_net_last_error", "kdq_net_num_params", "kdq_net_get_params",
           "kdq_net_set_params", "kdq_net_energy", "kdq_net_loss_grad", "kdq_net_train_step", "kdq_net_sample",
           "kdq_slmc_move")
PROF_FAMILIES = ("wrap", "chain", "get_G", "decide", "flush", "qr", "inv", "meas_wrap", "measure",
                 "gemm_other")


class KdqError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code


class ModelDesc(C.Structure):
    _fields_ = [("kind", C.c_int32), ("ns", C.c_int32), ("nt", C.c_int32), ("n_nodes", C.c_int32),
                ("nodes", C.POINTER(C.c_int64)), ("lattice", C.c_int32), ("site", C.c_int32 * 2),
                ("ht", C.c_double), ("alpha", C.POINTER(C.c_double)), ("eta", C.POINTER(C.c_double)),
                ("gamma", C.POINTER(C.c_double)), ("K", C.POINTER(C.c_double)),
                ("eK", C.POINTER(C.c_double)), ("eKinv", C.POINTER(C.c_double)),
                ("halfeK", C.POINTER(C.c_double)), ("halfeKinv", C.POINTER(C.c_double)),
                ("Pt", C.POINTER(C.c_double)), ("z", C.c_int32), ("bonds", C.POINTER(C.c_int32)),
                ("a", C.c_double), ("is_complex", C.c_int32), ("U", C.c_double), ("V", C.c_double),
                ("K_im", C.POINTER(C.c_double)), ("eK_im", C.POINTER(C.c_double)),
                ("eKinv_im", C.POINTER(C.c_double)), ("halfeK_im", C.POINTER(C.c_double)),
                ("halfeKinv_im", C.POINTER(C.c_double)), ("Pt_im", C.POINTER(C.c_double))]


class ModelParams(C.Structure):
    _fields_ = [("kind", C.c_int32), ("lattice", C.c_int32), ("site", C.c_int32 * 2), ("batch_size", C.c_int32),
                ("initial", C.c_int32), ("ht", C.c_double), ("c1", C.c_double), ("c2", C.c_double),
                ("dt", C.c_double), ("theta_relax", C.c_double), ("theta_quench", C.c_double)]


INITIAL_CODE = {"H0": 0, "V": 1}


class RunDesc(C.Structure):
    _fields_ = [("seed", C.c_uint64), ("chain", C.c_uint32), ("sweep0", C.c_uint32),
                ("precision", C.c_int32), ("delay", C.c_int32)]


class SceeDesc(C.Structure):
    _fields_ = [("nA", C.c_int32), ("indexA", C.POINTER(C.c_int64)), ("nB", C.c_int32),
                ("indexB", C.POINTER(C.c_int64)), ("lam", C.c_double), ("nlam", C.c_int32),
                ("pad", C.c_int32)]


_lib = None


def load(path: str = LIB_PATH):
    """Load the HIP library; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise ImportError(f"{path} is missing: build it with kapdqmc_amd.build.build() "
                          "(the product path has no CPU fallback)")
    L = C.CDLL(path)
    vp, i32, u32, dp = C.c_void_p, C.c_int32, C.c_uint32, C.POINTER(C.c_double)
    L.kdq_create.argtypes = [C.POINTER(ModelDesc), C.POINTER(RunDesc), i32, C.POINTER(vp)]
    L.kdq_destroy.argtypes = [vp]
    L.kdq_last_error.argtypes = [vp]
    L.kdq_last_error.restype = C.c_char_p
    L.kdq_num_chains.argtypes = [vp]
    L.kdq_num_obs.argtypes = [vp]
    L.kdq_set_fields.argtypes = [vp, i32, C.POINTER(C.c_uint8)]
    L.kdq_set_replay.argtypes = [vp, i32, C.POINTER(C.c_uint8), dp, C.c_int64]
    L.kdq_get_fields.argtypes = [vp, i32, C.POINTER(C.c_uint8)]
    L.kdq_initial_s.argtypes = [vp]
    L.kdq_phy_update.argtypes = [vp, i32, i32, dp]
    L.kdq_get_G.argtypes = [vp, i32, dp]
    L.kdq_set_G.argtypes = [vp, i32, dp]
    L.kdq_wrap.argtypes = [vp, i32, i32]
    L.kdq_update_layer.argtypes = [vp, i32, i32, u32]
    L.kdq_init_chain.argtypes = [vp]
    L.kdq_measure.argtypes = [vp, i32, dp]
    L.kdq_measure_noalloc.argtypes = [vp, i32, dp]
    L.kdq_get_health.argtypes = [vp, dp]
    L.kdq_check_G.argtypes = [vp, dp]
    L.kdq_get_counters.argtypes = [vp, dp]
    L.kdq_get_near_ties.argtypes = [vp, dp]
    L.kdq_profile.argtypes = [vp, u32]
    L.kdq_profile_read.argtypes = [vp, dp, C.POINTER(C.c_int64)]
    L.kdq_get_logdet.argtypes = [vp, dp]
    L.kdq_set_option.argtypes = [vp, i32, i32]
    L.kdq_get_precision_error.argtypes = [vp, dp]
    L.kdq_log_weight.argtypes = [vp, dp]
    L.kdq_scee_create.argtypes = [C.POINTER(ModelDesc), C.POINTER(RunDesc), C.POINTER(SceeDesc), i32, C.POINTER(vp)]
    L.kdq_scee_set_fields.argtypes = [vp, i32, i32, C.POINTER(C.c_uint8)]
    L.kdq_scee_get_fields.argtypes = [vp, i32, i32, C.POINTER(C.c_uint8)]
    L.kdq_scee_init.argtypes = [vp]
    L.kdq_scee_sweeps.argtypes = [vp, i32, dp]
    L.kdq_scee_get_G4.argtypes = [vp, i32, i32, dp, dp, dp, dp]
    L.kdq_scee_detg.argtypes = [vp, dp]
    L.kdq_G4.argtypes = [vp, i32, i32, i32, dp, dp, dp, dp]
    L.kdq_dop_create.argtypes = [C.POINTER(ModelDesc), C.POINTER(RunDesc), C.POINTER(SceeDesc), C.c_double, i32,
                                 C.POINTER(vp)]
    L.kdq_model_build.argtypes = [C.POINTER(ModelParams), C.POINTER(vp), C.POINTER(ModelDesc)]
    L.kdq_model_free.argtypes = [vp]
    fp = C.POINTER(C.c_float)
    u8p = C.POINTER(C.c_uint8)
    L.kdq_net_create.argtypes = [i32, i32, i32, i32, C.c_uint64, i32, C.POINTER(vp)]
    L.kdq_net_destroy.argtypes = [vp]
    L.kdq_net_last_error.argtypes = [vp]
    L.kdq_net_last_error.restype = C.c_char_p
    L.kdq_net_num_params.argtypes = [vp]
    L.kdq_net_get_params.argtypes = [vp, fp]
    L.kdq_net_set_params.argtypes = [vp, fp]
    L.kdq_net_energy.argtypes = [vp, u8p, i32, dp]
    L.kdq_net_loss_grad.argtypes = [vp, u8p, dp, i32, dp, fp]
    L.kdq_net_train_step.argtypes = [vp, u8p, dp, i32, C.c_float, dp]
    L.kdq_net_sample.argtypes = [vp, u8p, i32, i32, C.c_uint64, dp, C.POINTER(C.c_int32)]
    L.kdq_slmc_move.argtypes = [vp, vp, i32, dp]
    for name in EXPORTS:
        if name not in ("kdq_last_error", "kdq_net_last_error"):
            getattr(L, name).restype = C.c_int
    _lib = L
    return L


def check_net(rc: int, net=None):
    if rc != KDQ_OK:
        msg = load().kdq_net_last_error(net).decode(errors="replace")
        raise KdqError(rc, msg)


def check(rc: int, handle=None):
    if rc != KDQ_OK:
        msg = load().kdq_last_error(handle).decode(errors="replace")
        raise KdqError(rc, msg)
