"""ctypes binding of libttamen.so (the C ABI declared in include/ttamen.h).

The library is built in-tree next to this file (torchtt_amd/libttamen.so). There is no
fallback: if it is missing, loading raises ImportError with the build command.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libttamen.so")

c_i64 = ctypes.c_int64
c_i64p = ctypes.POINTER(ctypes.c_int64)
c_dp = ctypes.c_void_p
c_size = ctypes.c_size_t


class TTAmenStats(ctypes.Structure):
    _fields_ = [
        ("sweeps", ctypes.c_int64),
        ("gmres_iters", ctypes.c_int64),
        ("local_solves", ctypes.c_int64),
        ("direct_solves", ctypes.c_int64),
        ("matvecs", ctypes.c_int64),
        ("matvec_flops", ctypes.c_double),
        ("matvec_bytes", ctypes.c_double),
        ("max_res_last", ctypes.c_double),
        ("elapsed_s", ctypes.c_double),
        ("max_rank", ctypes.c_int64),
        ("refinements", ctypes.c_int64),
        ("matvecs_f32", ctypes.c_int64),
        ("matvec_flops_f32", ctypes.c_double),
    ]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class TTAmenSystem(ctypes.Structure):
    _fields_ = [
        ("d", ctypes.c_int),
        ("N", c_i64p), ("rA", c_i64p), ("rb", c_i64p),
        ("A_cores", ctypes.POINTER(c_dp)), ("b_cores", ctypes.POINTER(c_dp)),
        ("x0_cores", ctypes.POINTER(c_dp)), ("rx0", c_i64p),
        ("z0_cores", ctypes.POINTER(c_dp)),
        ("seed", ctypes.c_uint64),
        ("rmax", ctypes.c_int64),
        ("x_out", c_dp), ("x_out_elems", c_size),
        ("rx_out", c_i64p), ("offsets_out", c_i64p),
        ("stats", ctypes.POINTER(TTAmenStats)),
        ("status", ctypes.c_int),
    ]


# name -> (restype, argtypes)
_SIGNATURES = {
    "ttamen_version": (ctypes.c_char_p, []),
    "ttamen_last_error": (ctypes.c_char_p, []),
    "ttamen_cohort_staging_bytes": (c_size, []),
    "ttamen_amen_solve_prec_workspace_size": (c_size, [ctypes.c_int, c_i64p, c_i64p, c_i64p, c_i64p, c_i64,
                                                       c_i64, c_i64, ctypes.c_int]),
    "ttamen_local_prec": (ctypes.c_int, [ctypes.c_int, c_dp, c_dp, c_dp, c_dp, c_dp, c_i64, c_i64, c_i64,
                                         c_i64, c_i64, c_dp, c_size, c_dp]),
    "ttamen_tt_round_workspace_size": (c_size, [ctypes.c_int, c_i64p, c_i64p]),
    "ttamen_tt_full": (ctypes.c_int, [ctypes.c_int, c_i64p, c_i64p, ctypes.POINTER(c_dp), c_dp, c_size, c_dp,
                                      c_size, c_dp]),
    "ttamen_tt_full_workspace_size": (c_size, [ctypes.c_int, c_i64p, c_i64p]),
    "ttamen_tt_round": (ctypes.c_int, [ctypes.c_int, c_i64p, c_i64p, ctypes.POINTER(c_dp), ctypes.c_double,
                                       c_i64, c_dp, c_size, c_dp, c_size, c_i64p, c_i64p, c_dp]),
    "ttamen_cohort_stats": (ctypes.c_int, [ctypes.POINTER(ctypes.c_double)]),
    "ttamen_cohort_set_params": (ctypes.c_int, [ctypes.c_int] * 5),
    "ttamen_amen_solve": (ctypes.c_int, [
        ctypes.c_int, c_i64p, c_i64p, c_i64p, ctypes.POINTER(c_dp), ctypes.POINTER(c_dp),
        ctypes.POINTER(c_dp), c_i64p, ctypes.POINTER(c_dp), ctypes.c_uint64, c_i64, ctypes.c_double,
        c_i64, c_i64, c_i64, c_i64, c_i64, c_i64, ctypes.c_int, ctypes.c_int, ctypes.c_int,
        c_dp, c_size, c_dp, c_size, c_i64p, c_i64p, ctypes.POINTER(TTAmenStats), c_dp]),
    "ttamen_amen_solve_batch": (ctypes.c_int, [ctypes.POINTER(TTAmenSystem), ctypes.c_int, c_i64, ctypes.c_double,
                                               c_i64, c_i64, c_i64, c_i64, c_i64, ctypes.c_int, ctypes.c_int,
                                               c_dp, c_size, c_dp]),
    "ttamen_amen_solve_batch_workspace_size": (c_size, [ctypes.POINTER(TTAmenSystem), ctypes.c_int, c_i64, c_i64,
                                                        c_i64, c_i64, ctypes.c_int, ctypes.c_int]),
    "ttamen_serve": (ctypes.c_int, [ctypes.POINTER(TTAmenSystem), ctypes.POINTER(ctypes.c_double), ctypes.c_int,
                                    c_i64, ctypes.c_double, c_i64, c_i64, c_i64, c_i64, c_i64, ctypes.c_int,
                                    ctypes.c_int, c_dp, c_size, ctypes.POINTER(ctypes.c_double),
                                    ctypes.POINTER(ctypes.c_double), c_dp]),
    "ttamen_amen_solve_workspace_size": (c_size, [
        ctypes.c_int, c_i64p, c_i64p, c_i64p, c_i64p, c_i64, c_i64, c_i64, c_i64, c_i64, ctypes.c_int]),
    "ttamen_amen_solve_output_elems": (c_size, [ctypes.c_int, c_i64p, c_i64p, c_i64, c_i64, c_i64]),
    "ttamen_local_matvec": (ctypes.c_int, [c_dp, c_dp, c_dp, c_dp, c_dp] + [c_i64] * 8 + [c_dp, c_size, c_dp]),
    "ttamen_local_matvec_noalloc": (ctypes.c_int, [c_dp, c_dp, c_dp, c_dp, c_dp] + [c_i64] * 8 + [c_dp]),
    "ttamen_local_matvec_noalloc_limit": (c_size, []),
    "ttamen_matvec_mode": (ctypes.c_int, [ctypes.c_int]),
    "ttamen_rank_capacity": (ctypes.c_int, [ctypes.c_int]),
    "ttamen_tt_residual_norm": (ctypes.c_int, [ctypes.c_int, c_i64p, c_i64p, c_i64p, c_i64p, ctypes.POINTER(c_dp),
                                               ctypes.POINTER(c_dp), ctypes.POINTER(c_dp), c_dp, c_size,
                                               ctypes.POINTER(ctypes.c_double), c_dp]),
    "ttamen_tt_residual_norm_workspace_size": (c_size, [ctypes.c_int, c_i64p, c_i64p, c_i64p, c_i64p]),
    "ttamen_server_create": (ctypes.c_int, [ctypes.POINTER(ctypes.c_void_p), c_i64, ctypes.c_double, c_i64, c_i64,
                                            c_i64, c_i64, c_i64, ctypes.c_int, ctypes.c_int, c_dp, c_size]),
    "ttamen_server_submit": (ctypes.c_int, [c_dp, ctypes.POINTER(TTAmenSystem), c_dp, ctypes.POINTER(c_i64)]),
    "ttamen_server_poll": (ctypes.c_int, [c_dp, c_i64, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_double)]),
    "ttamen_server_wait": (ctypes.c_int, [c_dp, c_i64]),
    "ttamen_server_destroy": (ctypes.c_int, [c_dp]),
    "ttamen_amen_divide": (ctypes.c_int, [ctypes.c_int, c_i64p, c_i64p, c_i64p, ctypes.POINTER(c_dp),
                                          ctypes.POINTER(c_dp), ctypes.POINTER(c_dp), c_i64p,
                                          ctypes.POINTER(c_dp), c_i64, ctypes.c_double, c_i64p, c_i64, c_i64,
                                          c_i64, c_i64, c_i64, ctypes.c_int, ctypes.c_int, c_dp, c_size, c_dp,
                                          c_size, c_i64p, c_i64p, ctypes.POINTER(TTAmenStats), c_dp]),
    "ttamen_amen_divide_workspace_size": (c_size, [ctypes.c_int, c_i64p, c_i64p, c_i64p, c_i64p, c_i64p, c_i64,
                                                   c_i64, c_i64, c_i64, ctypes.c_int]),
    "ttamen_amen_divide_output_elems": (c_size, [ctypes.c_int, c_i64p, c_i64p, c_i64p, c_i64, c_i64]),
    "ttamen_amen_mv": (ctypes.c_int, [ctypes.c_int, c_i64p, c_i64p, c_i64p, c_i64p, ctypes.POINTER(c_dp),
                                      ctypes.POINTER(c_dp), ctypes.POINTER(c_dp), c_i64p, ctypes.POINTER(c_dp),
                                      c_i64, ctypes.c_double, c_i64p, c_i64, c_i64, ctypes.c_uint64, ctypes.c_int,
                                      c_dp, c_size, c_dp, c_size, c_i64p, c_i64p, ctypes.POINTER(TTAmenStats),
                                      c_dp]),
    "ttamen_amen_mv_workspace_size": (c_size, [ctypes.c_int, c_i64p, c_i64p, c_i64p, c_i64p, c_i64p, c_i64p,
                                               c_i64, c_i64]),
    "ttamen_amen_mv_output_elems": (c_size, [ctypes.c_int, c_i64p, c_i64p, c_i64p, c_i64p, c_i64, c_i64]),
    "ttamen_amen_mm": (ctypes.c_int, [ctypes.c_int, c_i64p, c_i64p, c_i64p, c_i64p, c_i64p, ctypes.POINTER(c_dp),
                                      ctypes.POINTER(c_dp), ctypes.POINTER(c_dp), c_i64p, ctypes.POINTER(c_dp),
                                      c_i64, ctypes.c_double, c_i64p, c_i64, c_i64, ctypes.c_uint64, ctypes.c_int,
                                      c_dp, c_size, c_dp, c_size, c_i64p, c_i64p, ctypes.POINTER(TTAmenStats),
                                      c_dp]),
    "ttamen_amen_mm_workspace_size": (c_size, [ctypes.c_int, c_i64p, c_i64p, c_i64p, c_i64p, c_i64p, c_i64p,
                                               c_i64p, c_i64, c_i64]),
    "ttamen_dmrg_mv": (ctypes.c_int, [ctypes.c_int, c_i64p, c_i64p, c_i64p, c_i64p, ctypes.POINTER(c_dp),
                                      ctypes.POINTER(c_dp), ctypes.POINTER(c_dp), c_i64p, c_i64, ctypes.c_double,
                                      c_i64p, c_i64, ctypes.c_uint64, ctypes.c_int, c_dp, c_size, c_dp, c_size,
                                      c_i64p, c_i64p, ctypes.POINTER(TTAmenStats), c_dp]),
    "ttamen_dmrg_mv_workspace_size": (c_size, [ctypes.c_int, c_i64p, c_i64p, c_i64p, c_i64p, c_i64p, c_i64p,
                                               c_i64]),
    "ttamen_dmrg_mv_output_elems": (c_size, [ctypes.c_int, c_i64p, c_i64p, c_i64p, c_i64]),
    "ttamen_rhs_assemble": (ctypes.c_int, [c_dp] * 4 + [c_i64] * 5 + [ctypes.c_double, c_dp, c_size, c_dp]),
    "ttamen_gmres_solve": (ctypes.c_int, [c_dp] * 5 + [c_i64] * 5 + [ctypes.c_double, c_i64, c_i64, c_dp, c_size,
                                          c_i64p, ctypes.POINTER(ctypes.c_int), c_dp]),
    "ttamen_gmres_solve_workspace_size": (c_size, [c_i64] * 6),
    "ttamen_device_info": (ctypes.c_int, [ctypes.c_char_p, c_size]),
    "ttamen_cohort_gemm_profile": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(ctypes.c_double)]),
    "ttamen_profile_read_ortho": (ctypes.c_int, [ctypes.POINTER(ctypes.c_double)]),
    "ttamen_gemm_autotune": (ctypes.c_int, [ctypes.c_int]),
    "ttamen_gemm_autotune_stats": (ctypes.c_int, [ctypes.POINTER(ctypes.c_double)]),
    "ttamen_gemm_autotune_clear": (None, []),
    "ttamen_local_matvec_f32": (ctypes.c_int, [c_dp, c_dp, c_dp, c_dp, c_dp] + [c_i64] * 8 + [c_dp, c_size, c_dp]),
    "ttamen_local_matvec_bf16": (ctypes.c_int, [c_dp, c_dp, c_dp, c_dp, c_dp] + [c_i64] * 8 + [c_dp, c_size, c_dp]),
    "ttamen_local_matvec_workspace_size": (c_size, [c_i64] * 8),
    "ttamen_phi_fwd_A": (ctypes.c_int, [c_dp] * 5 + [c_i64] * 7 + [c_dp, c_size, c_dp]),
    "ttamen_phi_bck_A": (ctypes.c_int, [c_dp] * 5 + [c_i64] * 7 + [c_dp, c_size, c_dp]),
    "ttamen_phi_fwd_rhs": (ctypes.c_int, [c_dp] * 4 + [c_i64] * 5 + [c_dp, c_size, c_dp]),
    "ttamen_phi_bck_rhs": (ctypes.c_int, [c_dp] * 4 + [c_i64] * 5 + [c_dp, c_size, c_dp]),
    "ttamen_phi_workspace_size": (c_size, [c_i64] * 7),
    "ttamen_qr": (ctypes.c_int, [c_dp, c_i64, c_i64, c_dp, c_dp, c_dp, c_size, c_dp]),
    "ttamen_qr_extend": (ctypes.c_int, [c_dp, c_i64, c_i64, c_i64, c_dp, c_dp, c_dp, c_size, c_dp]),
    "ttamen_svd": (ctypes.c_int, [c_dp, c_i64, c_i64, c_dp, c_dp, c_dp, c_dp, c_size, c_dp]),
    "ttamen_dense_workspace_size": (c_size, [c_i64, c_i64]),
    "ttamen_dense_solve": (ctypes.c_int, [c_dp, c_dp, c_dp, c_i64, c_dp, c_size, c_dp]),
    "ttamen_dense_solve_workspace_size": (c_size, [c_i64]),
    "ttamen_profile_enable": (ctypes.c_int, [ctypes.c_int]),
    "ttamen_profile_read": (ctypes.c_int, [ctypes.POINTER(ctypes.c_double)]),
    "ttamen_gemm_f64": (ctypes.c_int, [c_i64, c_i64, c_i64, c_i64, ctypes.c_double,
                                       c_dp, c_i64, c_i64, c_i64, c_dp, c_i64, c_i64, c_i64,
                                       ctypes.c_double, c_dp, c_i64, c_i64, c_i64, c_dp, c_size, c_dp]),
}

_lib = None


def load():
    """Load libttamen.so once and declare every exported signature."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            "libttamen.so is not built (expected at %s). Build it with "
            "`make -C torchtt_amd/csrc -j8` or `python -c 'import __graft_entry__ as g; g.build()'`." % LIB_PATH)
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in _SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def exported_names():
    return list(_SIGNATURES.keys())


class TTAmenError(RuntimeError):
    """A libttamen call returned a non-zero status."""

    def __init__(self, fn, code, msg):
        super().__init__("%s failed with status %d: %s" % (fn, code, msg))
        self.code = code


class BatchError(TTAmenError):
    """A batched or served call in which at least one system failed.

    `statuses[i]` is system i's own status (0 for a solved system), `results[i]` its solution
    (a TT) or None where it failed; `code` is the first failing system's status."""

    def __init__(self, fn, code, msg, statuses, results):
        super().__init__(fn, code, msg)
        self.statuses = list(statuses)
        self.results = list(results)


def check(fn, code):
    if code != 0:
        msg = load().ttamen_last_error()
        raise TTAmenError(fn, code, msg.decode() if msg else "")


def i64_array(values):
    arr = (ctypes.c_int64 * len(values))(*[int(v) for v in values])
    return arr


def ptr_array(tensors):
    arr = (c_dp * len(tensors))(*[t.data_ptr() for t in tensors])
    return arr


def stream_handle(device):
    import torch
    return torch.cuda.current_stream(device).cuda_stream
