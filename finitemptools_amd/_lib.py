"""ctypes binding of ``libmpsheff.so`` (declarations mirror include/mpsheff.h).

Loading fails loudly when the library is missing: there is no fallback path.
"""
import ctypes
import os

from ._build import lib_path

c_int = ctypes.c_int
c_double = ctypes.c_double
c_size_t = ctypes.c_size_t
c_int64 = ctypes.c_int64
c_int32 = ctypes.c_int32
c_void_p = ctypes.c_void_p
P = ctypes.POINTER

MH_OK = 0
STATUS_NAMES = {
    0: "MH_OK", 1: "MH_ERR_ARG", 2: "MH_ERR_SHAPE", 3: "MH_ERR_DTYPE",
    4: "MH_ERR_HIP", 5: "MH_ERR_OOM", 6: "MH_ERR_NOCONV", 7: "MH_ERR_UNSUPPORTED",
}
MH_F64, MH_C64 = 0, 1
PREC = {"f64": 0, "f32": 1, "bf16x3": 2}
MH_HEFF0, MH_HEFF1, MH_HEFF2 = 0, 1, 2
MH_K_GEMM, MH_K_WSTAGE, MH_K_VEC = 0, 1, 2
MH_BS_X, MH_BS_Y, MH_BS_P1, MH_BS_P2, MH_BS_T1, MH_BS_T2 = range(6)


class HeffOp(ctypes.Structure):
    """mirror of ``struct mh_heff_op``"""
    _fields_ = [
        ("kind", c_int), ("dtype", c_int), ("prec", c_int),
        ("Dl", c_int), ("d1", c_int), ("d2", c_int), ("Dr", c_int),
        ("wl", c_int), ("wm", c_int), ("wr", c_int),
        ("W1", c_void_p), ("W2", c_void_p), ("L", c_void_p), ("R", c_void_p),
    ]


class BsSectors(ctypes.Structure):
    """mirror of ``mh_bs_sectors``"""
    _fields_ = [("n", c_int), ("nq", c_int), ("charges", P(c_int32)), ("dims", P(c_int32))]


# name -> (restype, argtypes)
SIGNATURES = {
    "mh_version": (c_int, []),
    "mh_create": (c_int, [c_int, P(c_void_p)]),
    "mh_destroy": (c_int, [c_void_p]),
    "mh_set_stream": (c_int, [c_void_p, c_void_p]),
    "mh_synchronize": (c_int, [c_void_p]),
    "mh_last_error": (ctypes.c_char_p, [c_void_p]),
    "mh_malloc": (c_int, [c_void_p, c_size_t, P(c_void_p)]),
    "mh_free": (c_int, [c_void_p, c_void_p]),
    "mh_upload": (c_int, [c_void_p, c_void_p, c_void_p, c_size_t]),
    "mh_download": (c_int, [c_void_p, c_void_p, c_void_p, c_size_t]),
    "mh_ac_prime": (c_int, [c_void_p, c_int, c_int, c_int, c_int, c_int, c_int, c_int,
                            c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "mh_ac2_prime": (c_int, [c_void_p, c_int, c_int, c_int, c_int, c_int, c_int, c_int,
                             c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p,
                             c_void_p, c_void_p]),
    "mh_ac2_prime_noalloc": (c_int, [c_void_p, c_int, c_int, c_int, c_int, c_int, c_int, c_int,
                                     c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p,
                                     c_void_p, c_void_p]),
    "mh_c_prime": (c_int, [c_void_p, c_int, c_int, c_int, c_int, c_int, c_void_p,
                           c_void_p, c_void_p, c_void_p]),
    "mh_updateleft": (c_int, [c_void_p, c_int, c_int, c_int, c_int, c_int, c_int, c_int,
                              c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "mh_updateright": (c_int, [c_void_p, c_int, c_int, c_int, c_int, c_int, c_int, c_int,
                               c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "mh_mode_product": (c_int, [c_void_p, c_int, c_int, c_int, c_int, c_int, c_void_p, c_void_p,
                                c_void_p]),
    "mh_ac_prime_open": (c_int, [c_void_p, c_int, c_int, c_int, c_int, c_int, c_int, c_void_p,
                                 c_void_p, c_void_p, c_void_p]),
    "mh_schur_fold": (c_int, [c_int, c_int, c_int, c_int, c_int, P(c_int), P(c_int), P(c_int),
                              P(c_double), c_void_p, c_void_p]),
    "mh_env_gather": (c_int, [c_void_p, c_int, c_int, c_int, c_int, P(c_void_p), c_void_p]),
    "mh_env_scatter": (c_int, [c_void_p, c_int, c_int, c_int, c_int, c_void_p, P(c_void_p)]),
    "mh_heff_size": (c_int, [P(HeffOp), P(c_size_t)]),
    "mh_heff_apply": (c_int, [c_void_p, P(HeffOp), c_void_p, c_void_p]),
    "mh_lanczos_eig": (c_int, [c_void_p, P(HeffOp), c_void_p, c_int, c_double, c_int,
                               P(c_double), c_void_p, P(c_int), P(c_int)]),
    "mh_lanczos_eig_proj": (c_int, [c_void_p, P(HeffOp), c_int, P(c_void_p), c_void_p, c_int,
                                    c_double, c_int, P(c_double), c_void_p, P(c_int), P(c_int)]),
    "mh_lanczos_expm": (c_int, [c_void_p, P(HeffOp), c_void_p, c_double, c_double, c_int,
                                c_double, c_void_p, P(c_int)]),
    "mh_arnoldi_expm": (c_int, [c_void_p, P(HeffOp), c_void_p, c_double, c_double, c_int,
                                c_double, c_void_p, P(c_int)]),
    "mh_set_graphs": (c_int, [c_void_p, c_int]),
    "mh_set_krylov_small": (c_int, [c_void_p, c_int]),
    "mh_set_bs_plain": (c_int, [c_void_p, c_int]),
    "mh_set_fusion": (c_int, [c_void_p, c_int]),
    "mh_gemm_candidates": (c_int, [c_int, P(c_int)]),
    "mh_gemm_force_candidate": (c_int, [c_void_p, c_int]),
    "mh_dot": (c_int, [c_void_p, c_int, c_size_t, c_void_p, c_void_p, P(c_double)]),
    "mh_nrm2": (c_int, [c_void_p, c_int, c_size_t, c_void_p, P(c_double)]),
    "mh_axpby": (c_int, [c_void_p, c_int, c_size_t, c_double, c_double, c_void_p,
                         c_double, c_double, c_void_p]),
    "mh_gemm": (c_int, [c_void_p, c_int, c_int, c_int, c_int, c_int, c_void_p, c_int64,
                        c_int64, c_int, c_void_p, c_int64, c_int64, c_void_p, c_int64,
                        c_int64, c_double, c_double, c_double, c_double]),
    "mh_mpo_grad": (c_int, [c_void_p, c_int, c_int, c_int, c_int, c_int, c_int, c_int,
                            c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "mh_bs_plan_create": (c_int, [c_void_p, c_int, P(c_int64), P(c_void_p)]),
    "mh_bs_plan_add_gemm": (c_int, [c_void_p, c_void_p, c_int, c_int, c_int, c_int, c_int,
                                    c_int, c_void_p, c_void_p, c_int, c_void_p, c_void_p]),
    "mh_bs_plan_add_mode": (c_int, [c_void_p, c_void_p, c_int, c_int, c_int, c_void_p,
                                    c_void_p, c_int, c_void_p, c_void_p, c_void_p]),
    "mh_bs_plan_flops": (c_int, [c_void_p, P(c_double)]),
    "mh_bs_plan_variants": (c_int, [c_void_p, P(c_int)]),
    "mh_bs_plan_force_variant": (c_int, [c_void_p, c_int]),
    "mh_bs_plan_destroy": (c_int, [c_void_p, c_void_p]),
    "mh_bs_plan_sizes": (c_int, [c_void_p, P(c_int64)]),
    "mh_bs_plan_stage_count": (c_int, [c_void_p, P(c_int)]),
    "mh_bs_plan_stage_info": (c_int, [c_void_p, c_int, P(c_int32)]),
    "mh_bs_plan_stage_gemm": (c_int, [c_void_p, c_int, c_void_p, c_void_p, c_void_p, c_void_p]),
    "mh_bs_plan_stage_mode": (c_int, [c_void_p, c_int, c_void_p, c_void_p, c_void_p, c_void_p,
                                      c_void_p]),
    "mh_bs_plan_heff2": (c_int, [c_void_p, c_int, P(BsSectors), P(BsSectors), c_int, c_void_p,
                                 c_int, c_void_p, c_int, c_int, c_void_p, c_void_p, c_void_p,
                                 P(c_void_p)]),
    "mh_bs_plan_updateleft": (c_int, [c_void_p, c_int, P(BsSectors), P(BsSectors), c_int,
                                      c_void_p, c_int, c_void_p, c_int, c_void_p, c_void_p,
                                      P(c_void_p)]),
    "mh_bs_plan_updateright": (c_int, [c_void_p, c_int, P(BsSectors), P(BsSectors), c_int,
                                       c_void_p, c_int, c_void_p, c_int, c_void_p, c_void_p,
                                       P(c_void_p)]),
    "mh_bs_plan_two_site": (c_int, [c_void_p, P(BsSectors), P(BsSectors), P(BsSectors), c_int,
                                    c_void_p, P(c_void_p)]),
    "mh_bs_coupled_info": (c_int, [P(BsSectors), P(BsSectors), c_int, c_void_p, c_int, c_int,
                                   P(c_int), c_void_p, c_void_p, c_void_p, c_void_p]),
    "mh_bs_to_coupled": (c_int, [c_void_p, P(BsSectors), P(BsSectors), c_int, c_void_p, c_int,
                                 c_int, c_void_p, c_void_p]),
    "mh_bs_from_coupled": (c_int, [c_void_p, P(BsSectors), P(BsSectors), c_int, c_void_p, c_int,
                                   c_int, c_void_p, c_void_p]),
    "mh_bs_apply": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "mh_bs_lanczos_eig": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_int,
                                  c_double, c_int, P(c_double), c_void_p, P(c_int), P(c_int)]),
    "mh_sytrd_batched": (c_int, [c_void_p, c_int, P(c_int), P(c_int64), P(c_int64), c_void_p,
                                 c_void_p, c_void_p, c_void_p]),
    "mh_stedc_batched": (c_int, [c_void_p, c_int, P(c_int), P(c_int64), P(c_int64), c_void_p,
                                 c_void_p, c_void_p]),
    "mh_stedc_dc_batched": (c_int, [c_void_p, c_int, P(c_int), P(c_int64), P(c_int64), c_void_p,
                                    c_void_p, c_void_p]),
    "mh_gemm_batched": (c_int, [c_void_p, c_int, P(c_int64), c_double, c_double]),
    "mh_ormtr_batched": (c_int, [c_void_p, c_int, P(c_int), P(c_int64), P(c_int64), c_void_p,
                                 c_void_p, P(c_int64), P(c_int), P(c_int), c_void_p]),
    "mh_eig_check": (c_int, [c_void_p, c_int]),
    "mh_heev_small": (c_int, [c_void_p, c_int, c_int, c_int, c_void_p, c_void_p, c_void_p, c_int,
                              c_void_p]),
    "mh_heev_small_max": (c_int, [c_void_p, c_int, P(c_int)]),
    "mh_svd_small": (c_int, [c_void_p, c_int, c_int, c_int, c_void_p, c_int, c_void_p, c_void_p,
                             c_void_p, c_void_p]),
    "mh_svd_small_fits": (c_int, [c_void_p, c_int, c_int, c_int, P(c_int)]),
    "mh_profile_enable": (c_int, [c_void_p, c_int]),
    "mh_profile_read": (c_int, [c_void_p, c_int, P(c_double), P(c_int64), P(c_double)]),
    "mh_profile_reset": (c_int, [c_void_p]),
    "mh_probe_mfma_f64": (c_int, [c_void_p, P(c_double)]),
}

_LIB = None


def load():
    """Load the engine library (once).  Raises if it has not been built."""
    global _LIB
    if _LIB is not None:
        return _LIB
    path = lib_path()
    if not os.path.exists(path):
        raise RuntimeError(
            f"{path} is missing: build the HIP engine first "
            "(python -c 'import __graft_entry__ as g; g.build()')")
    lib = ctypes.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _LIB = lib
    return lib


class MHError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {msg}")
        self.status = status


class DimensionMismatch(MHError, ValueError):
    pass


def raise_for(status, handle=None, what=""):
    if status == MH_OK:
        return
    lib = load()
    msg = lib.mh_last_error(handle).decode() if handle else ""
    if what:
        msg = f"{what}: {msg}"
    if status == 2:
        raise DimensionMismatch(status, msg)
    raise MHError(status, msg)
