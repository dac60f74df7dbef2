"""ctypes binding of libbsr_gpu.so (the C ABI in include/bsr_gpu.h).

Loading fails loudly when the library has not been built; there is no CPU
fallback for any entry point.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libbsr_gpu.so")

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="F_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="F_CONTIGUOUS")


class BsrGpuError(RuntimeError):
    """A C-ABI call returned a non-zero status."""

    def __init__(self, code, msg):
        super().__init__(f"[{code}] {msg}")
        self.code = code


def _opt_dp(base=_dp):
    """ndpointer that also accepts None (NULL)."""

    class _Opt(base):
        @classmethod
        def from_param(cls, obj):
            if obj is None:
                return None
            return base.from_param(obj)

    return _Opt


_odp = _opt_dp()
_oip = _opt_dp(_ip)

# name -> argtypes (all return int status unless listed in _RESTYPE)
SIGNATURES = {
    "bsrg_last_error": [],
    "bsrg_abi_version": [],
    "bsrg_mkgrid": [C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double, C.c_int, _dp,
                    C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_double)],
    "bsrg_grid_tables": [C.c_int, C.c_int, _dp, _odp, _odp, _odp, _odp],
    "bsrg_create": [C.POINTER(C.c_void_p), C.c_int],
    "bsrg_destroy": [C.c_void_p],
    "bsrg_set_grid": [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, _dp],
    "bsrg_mrk_cell": [C.c_void_p, C.c_int, _odp],
    "bsrg_mrk_cell_noalloc": [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p],
    "bsrg_rk_moments": [C.c_void_p, C.c_int, _odp, _odp, _odp],
    "bsrg_convol": [C.c_void_p, C.c_int, C.c_int, _dp, _dp],
    "bsrg_set_orbitals": [C.c_void_p, C.c_int, _dp, _ip, _ip, _ip, C.c_double],
    "bsrg_alloc_matrix": [C.c_void_p, C.c_int, C.c_int, C.c_int],
    "bsrg_update_hl": [C.c_void_p, C.c_int, C.c_int, C.c_double],
    "bsrg_idata": [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, _ip, _ip, _ip, _ip, _ip, _dp],
    "bsrg_odata": [C.c_void_p, C.c_int, _ip, _ip, _ip, _dp, C.c_double],
    "bsrg_gen_lval": [C.c_void_p, C.c_double, C.c_int],
    "bsrg_cdata_build": [C.c_int, _ip, _ip, _ip, _ip, _dp, _ip, _ip, _ip, _ip, _dp, C.POINTER(C.c_int)],
    "bsrg_cdata_merge": [C.c_int, _ip, _ip, C.c_int, _ip, _ip, _ip, _ip, _dp, C.c_double, _ip, C.POINTER(C.c_int)],
    "bsrg_bs_orth": [C.c_void_p, C.c_int],
    "bsrg_lval_get": [C.c_void_p, C.POINTER(C.c_int), _opt_dp(), _opt_dp(), _opt_dp()],
    "bsrg_ldata": [C.c_void_p, C.c_int, C.c_int, C.c_int, _ip, _ip, _ip, _ip, _ip, _dp, C.c_int, C.c_double],
    "bsrg_symmetrize_diag": [C.c_void_p],
    "bsrg_scale": [C.c_void_p, C.c_double],
    "bsrg_acf_finish": [C.c_void_p, C.c_int],
    "bsrg_get_matrix": [C.c_void_p, _odp, _odp, _odp, _odp, _odp, _odp],
    "bsrg_wave_assemble": [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, _ip, _ip, _ip, _ip, _ip, _ip, _dp,
                           C.c_int, _ip, _ip, _dp, C.c_double, C.c_double, C.c_int, _dp, _ip],
    "bsrg_wave_get": [C.c_void_p, C.c_int, _odp, _odp, _odp, _odp, _odp, _odp],
    "bsrg_wave_write": [C.c_void_p, C.c_char_p, C.c_double, C.c_int, _dp, C.c_double],
    "bsrg_wave_timings": [C.c_void_p, _dp],
    "bsrg_set_precision": [C.c_void_p, C.c_int, C.c_int],
    "bsrg_drop_caches": [C.c_void_p],
    "bsrg_wave_diff": [C.c_void_p, C.c_void_p, C.c_int, C.c_double, C.POINTER(C.c_double)],
    "bsrg_wave_eigen": [C.c_void_p, C.c_int, _dp],
    "bsrg_matfile_eigen": [C.c_void_p, C.c_char_p, C.c_int, _dp],
    "bsrg_set_eigen_method": [C.c_void_p, C.c_int],
    "bsrg_eigen_info": [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_double)],
    "bsrg_dbs_set_grid": [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, _dp],
    "bsrg_dbs_mrk": [C.c_void_p, C.c_int, C.c_int, _dp],
    "bsrg_dbs_moments": [C.c_void_p, C.c_int, C.c_int, _dp, _dp],
    "bsrg_dbs_convol": [C.c_void_p, C.c_int, C.c_int, C.c_int, _dp, _dp],
    "bsrg_dirac_set_orbitals": [C.c_void_p, C.c_int, _dp, _ip, C.c_int, _ip, _ip, _dp],
    "bsrg_dirac_assemble": [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, _ip, C.c_int, _ip, _ip, _ip, _ip,
                            _ip, _ip, _dp],
    "bsrg_dirac_get": [C.c_void_p, _odp, _odp, _odp, _odp, _odp],
    "bsrg_dirac_timings": [C.c_void_p, _dp],
    "bsrg_dirac_engine": [C.c_void_p],
    "bsrg_dirac_diff": [C.c_void_p, C.c_void_p, C.c_double, C.POINTER(C.c_double)],
    "bsrg_dirac_overlap": [C.c_void_p, C.c_int, C.c_int, C.c_int, _ip, C.c_int, _ip, _ip, _ip, _dp, C.c_double,
                           C.c_int],
    "bsrg_dirac_overlap_get": [C.c_void_p, _odp, _odp, _odp, _odp],
    "bsrg_dirac_update_hx": [C.c_void_p, C.c_int, C.c_int, _dp, C.c_int, _ip, _ip, _ip, _dp],
    "bsrg_dirac_energy_shift": [C.c_void_p, C.c_double, _dp],
    "bsrg_dirac_diag_channels": [C.c_void_p, _ip, _ip, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double,
                                 C.c_double, _oip, C.POINTER(C.c_int)],
    "bsrg_dirac_solutions": [C.c_void_p, _odp, _odp, _oip],
    "bsrg_dirac_transform": [C.c_void_p, C.c_int],
    "bsrg_dirac_overlaps": [C.c_void_p, _dp],
    "bsrg_dirac_diag_timings": [C.c_void_p, _dp],
    "bsrg_dirac_record": [C.c_void_p, C.c_char_p, C.c_int, C.c_int],
    "bsrg_matrix_record": [C.c_void_p, C.c_char_p, C.c_int, C.c_int, C.c_int],
    "bsrg_matrix_resume": [C.c_void_p, C.c_char_p, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)],
    "bsrg_lb_create": [],
    "bsrg_lb_destroy": [C.c_void_p],
    "bsrg_lb_read_bank": [C.c_void_p, C.c_char_p, _ip],
    "bsrg_lb_bank_get": [C.c_void_p, _dp, _ip, _ip, _ip, _ip, _ip, _ip, _ip, _ip, _ip],
    "bsrg_lb_set_records": [C.c_void_p, C.c_int, _dp, _ip, _ip, _ip],
    "bsrg_lb_set_dets": [C.c_void_p, C.c_int, C.c_int, _ip, _ip, _ip, C.c_int, C.c_int, _ip, _ip, _ip],
    "bsrg_lb_set_orbitals": [C.c_void_p, C.c_int, _ip, _ip, _ip, _dp, C.c_int],
    "bsrg_lb_set_channels": [C.c_void_p, C.c_int, _ip, C.c_int, _ip],
    "bsrg_lb_set_states": [C.c_void_p, C.c_int, _ip, _ip, _dp, C.c_int, _ip, C.c_int, _ip, _ip, C.c_int, _ip],
    "bsrg_lb_set_params": [C.c_void_p, C.c_double, C.c_double, C.c_double, C.c_int, C.c_int, C.c_int],
    "bsrg_lb_run": [C.c_void_p, C.c_int, C.c_int, _ip],
    "bsrg_lb_expansion": [C.c_void_p, _ip, _ip, _ip, _ip, _ip, _ip, _dp],
    "bsrg_lb_lists": [C.c_void_p, _ip, _ip, _ip, _ip, _ip, _ip, _ip, _dp],
    "bsrg_wave_prepare": [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, _ip, _ip, _ip, _ip, _ip, _ip, _dp,
                          C.c_int, _ip, _ip, _dp, C.c_double, C.c_double, C.c_int, _dp, _ip],
    "bsrg_wave_launch": [C.c_void_p],
    "bsrg_lists_upload": [C.c_void_p, C.c_int, _ip, _ip, _ip, _ip, _ip, _ip, _dp],
    "bsrg_wave_assemble_resident": [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, _ip, _ip, _dp, C.c_double,
                                    C.c_double, C.c_int, _dp, _ip],
    "bsrg_wave_finish": [C.c_void_p],
    "bsrg_wave_engine": [C.c_void_p],
    "bsrg_set_capture_safe": [C.c_void_p, C.c_int],
    "bsrg_set_gemm_engine": [C.c_void_p, C.c_int],
    "bsrg_wave_host_timings": [C.c_void_p, _dp],
    "bsrg_wave_term_counts": [C.c_void_p, C.c_int, np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")],
    "bsrg_graph_capture": [C.c_int, C.POINTER(C.c_void_p), C.POINTER(C.c_void_p)],
    "bsrg_graph_launch": [C.c_void_p],
    "bsrg_graph_sync": [C.c_void_p, _odp, C.POINTER(C.c_double)],
    "bsrg_graph_destroy": [C.c_void_p],
    "bsrg_graph_capture_batched": [C.c_int, C.POINTER(C.c_void_p), C.POINTER(C.c_void_p)],
    "bsrg_graph_batch_info": [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_int)],
}
_RESTYPE = {"bsrg_last_error": C.c_char_p, "bsrg_lb_create": C.c_void_p, "bsrg_lb_destroy": None}

_LIB = None


def lib():
    """The loaded library; raises if it has not been built."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `python -m bsr_amd.build` "
                "(there is no fallback path)")
        L = C.CDLL(LIB_PATH)
        for name, args in SIGNATURES.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = _RESTYPE.get(name, C.c_int)
        _LIB = L
    return _LIB


def check(rc):
    if rc != 0:
        raise BsrGpuError(rc, lib().bsrg_last_error().decode())
    return rc


def f64(a):
    return np.asfortranarray(np.asarray(a, dtype=np.float64))


def i32(a):
    return np.asfortranarray(np.asarray(a, dtype=np.int32))
