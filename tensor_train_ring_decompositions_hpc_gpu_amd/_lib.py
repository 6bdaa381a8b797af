"""ctypes binding of the ttx C ABI (``include/ttx.h``).

The library is the in-tree ``libttx.so`` built by ``build.py``.  Loading fails loudly when it is
missing: there is no CPU fallback for any entry point.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libttx.so")

MAX_ORDER = 16
MAX_ELL = 64
NOALLOC_SCRATCH_BYTES = 256 << 20

OK, ERR_BAD_ARG, ERR_OOM, ERR_NO_CONVERGE, RESIDUAL_FALLBACK_USED, ERR_HIP, ERR_UNSUPPORTED = range(7)
PREC_FP32, PREC_BF16, PREC_FP16 = 0, 1, 2
RANK_FIXED, RANK_DELTA, RANK_LEARNED = 0, 1, 2
METHOD_AUTO, METHOD_EXACT, METHOD_RSVD = 0, 1, 2
PATH_EXACT_M, PATH_EXACT_N, PATH_RSVD = 0, 1, 2
FLAG_FAST_RANGE, FLAG_TIMING, FLAG_ADAPTIVE, FLAG_STREAM_TWICE, FLAG_TAIL_X3, FLAG_BATCH_PLAN = 1, 2, 4, 8, 16, 32
NUM_KCLASS = 8
KCLASS = ["sketch", "proj", "gram", "permute", "tall", "small", "unused6", "unused7"]

PRECISIONS = {"fp32": PREC_FP32, "bf16": PREC_BF16, "fp16": PREC_FP16}
RANK_MODES = {"fixed": RANK_FIXED, "delta": RANK_DELTA, "learned": RANK_LEARNED}
METHODS = {"auto": METHOD_AUTO, "exact": METHOD_EXACT, "rsvd": METHOD_RSVD}


class Params(ctypes.Structure):
    _fields_ = [
        ("order", ctypes.c_int32),
        ("dims", ctypes.c_int64 * MAX_ORDER),
        ("ranks", ctypes.c_int32 * (MAX_ORDER + 1)),
        ("oversampling", ctypes.c_int32),
        ("power_iters", ctypes.c_int32),
        ("omega_seed", ctypes.c_uint64),
        ("precision", ctypes.c_int32),
        ("method", ctypes.c_int32),
        ("rank_mode", ctypes.c_int32),
        ("eps", ctypes.c_double),
        ("flags", ctypes.c_int32),
    ]


class Report(ctypes.Structure):
    _fields_ = [
        ("ranks", ctypes.c_int32 * (MAX_ORDER + 1)),
        ("ell", ctypes.c_int32 * MAX_ORDER),
        ("path", ctypes.c_int32 * MAX_ORDER),
        ("power_iters", ctypes.c_int32 * MAX_ORDER),
        ("zero_pivots", ctypes.c_int32 * MAX_ORDER),
        ("jacobi_sweeps", ctypes.c_int32 * MAX_ORDER),
        ("norm2_in", ctypes.c_double * MAX_ORDER),
        ("step_ms", ctypes.c_double * MAX_ORDER),
        ("total_ms", ctypes.c_double),
        ("bytes_alg", ctypes.c_int64),
        ("kernel_ms", ctypes.c_double * NUM_KCLASS),
        ("kernel_launches", ctypes.c_int32 * NUM_KCLASS),
        ("kernel_bytes", ctypes.c_int64 * NUM_KCLASS),
        ("range_resid", ctypes.c_double * MAX_ORDER),
        ("range_checks", ctypes.c_int32 * MAX_ORDER),
        ("passes", ctypes.c_int32 * MAX_ORDER),
        ("fallback", ctypes.c_int32 * MAX_ORDER),
        ("predicted_rank", ctypes.c_int32 * MAX_ORDER),
        ("kernel_top_bytes", ctypes.c_int64 * NUM_KCLASS),
        ("kernel_top_ms", ctypes.c_double * NUM_KCLASS),
    ]


class BatchReport(ctypes.Structure):
    _fields_ = [
        ("total_ms", ctypes.c_double),
        ("tensors_per_s", ctypes.c_double),
        ("p50_ms", ctypes.c_double),
        ("p99_ms", ctypes.c_double),
        ("streams", ctypes.c_int32),
        ("graph", ctypes.c_int32),
        ("graph_reused", ctypes.c_int32),
        ("tuned_shapes", ctypes.c_int32),
        ("engine", ctypes.c_int32),
        ("launches", ctypes.c_int32),
        ("sketch_chunks", ctypes.c_int32),
    ]


BATCH_GRAPH, BATCH_PER_TENSOR, BATCH_AUTOTUNE, BATCH_STREAMS = 1, 2, 4, 8

_P = ctypes.c_void_p
_I32 = ctypes.c_int32
_I64 = ctypes.c_int64
_U32 = ctypes.c_uint32
_U64 = ctypes.c_uint64
_F = ctypes.c_float
_D = ctypes.c_double
_FPP = ctypes.POINTER(ctypes.c_void_p)

# name -> (restype, argtypes)
SIGNATURES = {
    "ttx_create": (_I32, [ctypes.POINTER(_P), _P]),
    "ttx_destroy": (_I32, [_P]),
    "ttx_set_stream": (_I32, [_P, _P]),
    "ttx_status_string": (ctypes.c_char_p, [_I32]),
    "ttx_last_error": (ctypes.c_char_p, [_P]),
    "ttx_version": (_I32, []),
    "ttx_params_init": (None, [ctypes.POINTER(Params), _I32, ctypes.POINTER(_I64), ctypes.POINTER(_I32)]),
    "ttx_core_size": (_I64, [ctypes.POINTER(Params), _I32, _I32]),
    "ttx_tt_svd": (_I32, [_P, _P, ctypes.POINTER(Params), _FPP, ctypes.POINTER(_F), ctypes.POINTER(Report)]),
    "ttx_tt_svd_noalloc": (_I32, [_P, _P, ctypes.POINTER(Params), _FPP, ctypes.POINTER(_F), ctypes.POINTER(Report)]),
    "ttx_tr_svd": (_I32, [_P, _P, ctypes.POINTER(Params), _FPP, ctypes.POINTER(_F), ctypes.POINTER(Report)]),
    "ttx_tt_svd_host": (_I32, [_P, _P, ctypes.POINTER(Params), _I64, _FPP, ctypes.POINTER(_F), ctypes.POINTER(Report)]),
    "ttx_tr_svd_host": (_I32, [_P, _P, ctypes.POINTER(Params), _I64, _FPP, ctypes.POINTER(_F), ctypes.POINTER(Report)]),
    "ttx_rsvd": (_I32, [_P, _P, _I64, _I64, _I32, _I32, _I32, _U64, _I32, _P, _P, _P]),
    "ttx_step": (_I32, [_P, _P, _I64, _I64, _I32, _I32, _I32, _U64, _U32, _I32, _I32, _P, _P,
                        ctypes.POINTER(_F), ctypes.POINTER(_I32)]),
    "ttx_permute_231": (_I32, [_P, _P, _P, _I64, _I64, _I64]),
    "ttx_fill_normal": (_I32, [_P, _P, _I64, _U64, _U32, _F]),
    "ttx_omega": (_I32, [_P, _P, _I64, _I32, _U64, _U32]),
    "ttx_tt_generate": (_I32, [_P, _I32, ctypes.POINTER(_I64), ctypes.POINTER(_I32), _F, _F, _U64, _P, _FPP]),
    "ttx_tr_generate": (_I32, [_P, _I32, ctypes.POINTER(_I64), ctypes.POINTER(_I32), _F, _F, _U64, _P, _FPP]),
    "ttx_tt_generate_ref": (_I32, [_P, _I32, ctypes.POINTER(_I64), ctypes.POINTER(_I32), ctypes.c_uint32, _P, _FPP]),
    "ttx_tr_generate_ref": (_I32, [_P, _I32, ctypes.POINTER(_I64), ctypes.POINTER(_I32), ctypes.c_uint32, _P, _FPP]),
    "ttx_tt_rel_error": (_I32, [_P, _I32, ctypes.POINTER(_I64), ctypes.POINTER(_I32), _FPP, _P,
                                ctypes.POINTER(_D)]),
    "ttx_tr_rel_error": (_I32, [_P, _I32, ctypes.POINTER(_I64), ctypes.POINTER(_I32), _FPP, _P,
                                ctypes.POINTER(_D)]),
    "ttx_tt_full": (_I32, [_P, _I32, ctypes.POINTER(_I64), ctypes.POINTER(_I32), _FPP, _P]),
    "ttx_tt_round": (_I32, [_P, _I32, ctypes.POINTER(_I64), ctypes.POINTER(_I32), _FPP, ctypes.c_double,
                            ctypes.POINTER(_I32), _FPP, ctypes.POINTER(_I32), _P]),
    "ttx_tr_full": (_I32, [_P, _I32, ctypes.POINTER(_I64), ctypes.POINTER(_I32), _FPP, _P]),
    "ttx_batch_tt_svd": (_I32, [_P, _I32, _I32, ctypes.POINTER(_I64), ctypes.POINTER(_I32), _FPP, _FPP,
                                ctypes.POINTER(Params), _I32, _I32, ctypes.POINTER(BatchReport)]),
    "ttx_mlp_init": (_I32, [_P, _U64]),
    "ttx_mlp_set_params": (_I32, [_P, ctypes.POINTER(_F)]),
    "ttx_mlp_get_params": (_I32, [_P, ctypes.POINTER(_F)]),
    "ttx_mlp_load": (_I32, [_P, ctypes.c_char_p]),
    "ttx_mlp_forward": (_I32, [_P, _P, _I32, _P]),
    "ttx_mlp_gradient": (_I32, [_P, _P, _P, _I32, _P, ctypes.POINTER(_D)]),
    "ttx_mlp_train_step": (_I32, [_P, _P, _P, _I32, _F, ctypes.POINTER(_D)]),
    "ttx_mlp_features": (_I32, [_P, _P, _I32, _D, _I64, _I64, _P]),
    "ttx_mlp_predict": (_I32, [_P, _P, _I32, _D, _I64, _I64, _P]),
    "ttx_probe": (_I32, [_P, _P, _I64, _I64, _U64, _U32, ctypes.POINTER(_D)]),
}

MLP_IN, MLP_HIDDEN, MLP_OUT = 40, 128, 2
MLP_PARAMS = MLP_IN * MLP_HIDDEN + MLP_HIDDEN + MLP_HIDDEN * MLP_HIDDEN + MLP_HIDDEN + MLP_HIDDEN * MLP_OUT + MLP_OUT
MODEL_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "rank_mlp.f32")

_lib = None


def load(path=LIB_PATH):
    """Load libttx.so and declare every signature; raises if the library was not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError("ttx: %s is missing; run __graft_entry__.build() (there is no CPU fallback)" % path)
    lib = ctypes.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


class TtxError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__("ttx status %d (%s): %s" % (status, _status_name(status), msg))
        self.status = status


def _status_name(s):
    names = {OK: "ok", ERR_BAD_ARG: "bad argument", ERR_OOM: "out of memory", ERR_NO_CONVERGE: "no convergence",
             RESIDUAL_FALLBACK_USED: "fallback used", ERR_HIP: "HIP error", ERR_UNSUPPORTED: "unsupported"}
    return names.get(s, "unknown")


def i64_array(vals):
    arr = (ctypes.c_int64 * len(vals))(*[int(v) for v in vals])
    return arr


def i32_array(vals):
    return (ctypes.c_int32 * len(vals))(*[int(v) for v in vals])


def ptr_array(ptrs):
    return (ctypes.c_void_p * len(ptrs))(*[ctypes.c_void_p(int(p)) for p in ptrs])
