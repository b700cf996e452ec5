"""ctypes binding of libgrtm.so (the HIP implementation).

The shared library is built in-tree by ``__graft_entry__.build()`` (or
``make -C grtm_cuda_amd/csrc``).  Nothing here falls back to a CPU path: if the
library is missing, loading raises.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libgrtm.so")

GRTM_OK = 0
GRTM_E_IO = -1
GRTM_E_PARSE = -2
GRTM_E_INPUT = -3
GRTM_E_UNSUPPORTED = -4
GRTM_E_DEVICE = -5
GRTM_E_NOMEM = -6
GRTM_E_TOO_LARGE = -7

PREC_FP64 = 0
PREC_FP32 = 1
PREC_BF16X3 = 2
PREC_BF16X6 = 3
PREC_FP16X3 = 4

P = ctypes.c_void_p
D = ctypes.c_double
I = ctypes.c_int


class GrtmError(RuntimeError):
    def __init__(self, status, what=""):
        self.status = status
        msg = _lib().grtm_status_str(status).decode() if _LIB is not None else str(status)
        super().__init__(f"{what}: {msg} (status {status})" if what else f"{msg} (status {status})")


class Grid(ctypes.Structure):
    _fields_ = [("nw", I), ("dt", D), ("nk", I), ("kmax", D), ("omega_i", D), ("t0", D), ("taper", D)]


class Job(ctypes.Structure):
    _fields_ = [("n_layers", I), ("layers", P), ("src_depth", D), ("rcv_depth", D), ("nr", I), ("dists", P)]


class Model(ctypes.Structure):
    _fields_ = [("n_layers", I), ("layers", P), ("src_depth", D), ("rcv_depth", D)]


_LIB = None


def _lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"libgrtm.so not built at {LIB_PATH}; run __graft_entry__.build()")
        lib = ctypes.CDLL(LIB_PATH)
        lib.grtm_status_str.restype = ctypes.c_char_p
        lib.grtm_version.restype = ctypes.c_char_p
        lib.green.restype = None
        lib.green.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p]
        lib.grtm_green.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p, I]
        lib.grtm_green_mem.argtypes = [ctypes.c_char_p, I, P, P, P]
        lib.grtm_green_mem_noalloc.argtypes = [ctypes.c_char_p, I, P, P, P, P]
        lib.grtm_create.argtypes = [P, I]
        lib.grtm_destroy.argtypes = [P]
        lib.grtm_destroy.restype = None
        lib.grtm_plan_traces.argtypes = [P, P, P, I, I, P]
        lib.grtm_plan_columns.argtypes = [P, P, P, I, I, I, P]
        lib.grtm_plan_execute.argtypes = [P, P]
        lib.grtm_plan_sync.argtypes = [P]
        lib.grtm_plan_traces_out.argtypes = [P, P]
        lib.grtm_plan_kernels_out.argtypes = [P, I, I, I, P]
        lib.grtm_plan_timing.argtypes = [P, I, P]
        lib.grtm_plan_work.argtypes = [P, P]
        lib.grtm_plan_destroy.argtypes = [P]
        lib.grtm_plan_destroy.restype = None
        lib.grtm_selftest_cell.argtypes = [I, I, D, D, D, I, P, I, D, I, D, P, P]
        lib.grtm_selftest_cell_dd.argtypes = [I, I, D, D, D, I, P, I, D, I, D, P, P]
        lib.grtm_opcount_cell.argtypes = [I, I, I, I, I, P]
        lib.grtm_opcount_columns_cell.argtypes = [I, I, I, I, I, I, P]
        lib.grtm_selftest_gemm.argtypes = [I, I, I, I, P, P, P]
        lib.grtm_selftest_gemm_cut.argtypes = [I, I, I, I, P, P, P, P]
        lib.grtm_selftest_mfma_probe.argtypes = [I, P]
        lib.grtm_selftest_copy.argtypes = [I]
        lib.grtm_plan_set_adaptive.argtypes = [P, P, D, I, D]
        lib.grtm_plan_set_learned.argtypes = [P, P, D, D, I, D]
        lib.grtm_plan_set_rows64.argtypes = [P, I]
        lib.grtm_plan_plant_nonfinite.argtypes = [P, P, I]
        lib.grtm_plan_set_ptam.argtypes = [P, I, I]
        lib.grtm_plan_set_flag64.argtypes = [P, I, D]
        lib.grtm_plan_flag_list.argtypes = [P, P, I]
        lib.grtm_plan_ptam_stats.argtypes = [P, P]
        lib.grtm_batch_plant_nonfinite.argtypes = [P, P, I]
        lib.grtm_plan_adaptive_stats.argtypes = [P, P]
        lib.grtm_plan_adaptive_rows.argtypes = [P, P, P]
        lib.grtm_plan_row_labels.argtypes = [P, D, P]
        lib.grtm_mlp_create.argtypes = [I, I, I, I, I, P]
        lib.grtm_mlp_destroy.argtypes = [P]
        lib.grtm_mlp_destroy.restype = None
        lib.grtm_mlp_nparams.argtypes = [P]
        lib.grtm_mlp_nparams.restype = ctypes.c_long
        lib.grtm_mlp_set_params.argtypes = [P, P]
        lib.grtm_mlp_get_params.argtypes = [P, P]
        lib.grtm_mlp_set_data.argtypes = [P, P, P, P, I]
        lib.grtm_mlp_train.argtypes = [P, I, I, ctypes.c_float, ctypes.c_ulonglong, P]
        lib.grtm_mlp_grads.argtypes = [P, P, P, P, I, P, P]
        lib.grtm_mlp_predict.argtypes = [P, P, I, P]
        lib.grtm_mlp_predict_train_forward.argtypes = [P, P, I, P]
        lib.grtm_mlp_set_loss.argtypes = [P, I, ctypes.c_float]
        lib.grtm_batch_create.argtypes = [P, P, I, P, I, P, I, I, I, P]
        lib.grtm_batch_execute.argtypes = [P, P]
        lib.grtm_batch_execute_eager.argtypes = [P]
        lib.grtm_batch_sync.argtypes = [P]
        lib.grtm_batch_traces_out.argtypes = [P, I, I, P]
        lib.grtm_batch_timing.argtypes = [P, I, P]
        lib.grtm_batch_config.argtypes = [P, P]
        lib.grtm_batch_work.argtypes = [P, P]
        lib.grtm_batch_set_adaptive.argtypes = [P, P, P, I, D]
        lib.grtm_batch_set_learned.argtypes = [P, P, D, P, I, D]
        lib.grtm_batch_adaptive_stats.argtypes = [P, P]
        lib.grtm_batch_destroy.argtypes = [P]
        lib.grtm_batch_destroy.restype = None
        _LIB = lib
    return _LIB


# every symbol include/grtm.h declares
EXPORTS = [
    "green", "grtm_green", "grtm_green_mem", "grtm_green_mem_noalloc", "grtm_last_status", "grtm_status_str", "grtm_version",
    "grtm_create", "grtm_destroy", "grtm_plan_traces", "grtm_plan_execute", "grtm_plan_sync",
    "grtm_plan_traces_out", "grtm_plan_kernels_out", "grtm_plan_timing", "grtm_plan_work",
    "grtm_plan_destroy", "grtm_selftest_cell", "grtm_selftest_cell_dd", "grtm_opcount_cell", "grtm_selftest_gemm", "grtm_selftest_gemm_cut",
    "grtm_plan_columns", "grtm_opcount_columns_cell",
    "grtm_selftest_mfma_probe", "grtm_selftest_copy", "grtm_plan_set_adaptive", "grtm_plan_set_learned", "grtm_plan_set_rows64",
    "grtm_plan_adaptive_stats", "grtm_plan_adaptive_rows",
    "grtm_plan_row_labels", "grtm_mlp_create", "grtm_mlp_destroy", "grtm_mlp_nparams", "grtm_mlp_set_params",
    "grtm_mlp_get_params", "grtm_mlp_set_data", "grtm_mlp_train", "grtm_mlp_grads", "grtm_mlp_predict", "grtm_mlp_predict_train_forward",
    "grtm_mlp_set_loss", "grtm_batch_create", "grtm_batch_execute", "grtm_batch_execute_eager", "grtm_batch_sync",
    "grtm_batch_traces_out", "grtm_batch_timing", "grtm_batch_config", "grtm_batch_work", "grtm_batch_destroy",
    "grtm_batch_set_adaptive", "grtm_batch_set_learned", "grtm_batch_adaptive_stats",
    "grtm_plan_plant_nonfinite", "grtm_batch_plant_nonfinite", "grtm_plan_set_ptam", "grtm_plan_ptam_stats",
    "grtm_plan_set_flag64", "grtm_plan_flag_list",
]


def check(status, what=""):
    if status != GRTM_OK:
        raise GrtmError(status, what)
