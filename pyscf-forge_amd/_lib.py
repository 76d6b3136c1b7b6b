"""ctypes binding of libfcisigma.so (include/fcisigma.h).

The library is built in place by ``__graft_entry__.build()`` (or
``make -C pyscf-forge_amd/csrc``).  There is no fallback: every call fails
loudly if the HIP library is missing or a call returns an error status, the
way the reference turns libfci_gas status codes into RuntimeError
(pyscf/mcscf/fci_gas.py:127-129).
"""

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(_HERE, "libfcisigma.so")

SUCCESS, ERR_INVALID, ERR_MEMORY, ERR_DEVICE = 0, -1, -2, -3
PREC_FP64, PREC_FP32, PREC_BF16X3 = 0, 1, 2
PRECISIONS = {"fp64": PREC_FP64, "fp32": PREC_FP32, "bf16x3": PREC_BF16X3}

_vp = ctypes.c_void_p
_dp = ctypes.POINTER(ctypes.c_double)
_i64 = ctypes.c_int64
_i64p = ctypes.POINTER(ctypes.c_int64)
_ip = ctypes.POINTER(ctypes.c_int)
_u64p = ctypes.POINTER(ctypes.c_uint64)

# name -> (restype, argtypes); every symbol declared in include/fcisigma.h
SIGNATURES = {
    "fcis_last_error": (ctypes.c_char_p, []),
    "fcis_device_info": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int]),
    "fcis_kept_memory": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(ctypes.c_uint64)]),
    "fcis_release_kept_memory": (ctypes.c_int, [ctypes.c_int]),
    "fcis_space_create": (ctypes.c_int, [ctypes.POINTER(_vp), ctypes.c_int,
                                         ctypes.c_int, ctypes.c_int, _vp]),
    "fcis_space_free": (None, [_vp]),
    "fcis_space_create_ordered": (ctypes.c_int, [ctypes.POINTER(_vp), ctypes.c_int,
                                                 ctypes.c_int, ctypes.c_int, _u64p, _u64p, _vp]),
    "fcis_gas_string_order": (ctypes.c_int, [ctypes.c_int, _ip, ctypes.c_int, _u64p]),
    "fcis_rotate_fibers": (ctypes.c_int, [_vp, _i64, _vp, _vp, ctypes.c_int, ctypes.c_int, _vp,
                                          _vp, _vp, _vp, _vp, _vp]),
    "fcis_space_info": (ctypes.c_int, [_vp, _i64p, _i64p, _ip, _ip]),
    "fcis_gen_linkstr_index": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int,
                                              _vp, _vp]),
    "fcis_str2addr": (ctypes.c_int64, [ctypes.c_int, ctypes.c_int, ctypes.c_uint64]),
    "fcis_addr2str": (ctypes.c_uint64, [ctypes.c_int, ctypes.c_int, ctypes.c_int64]),
    "fcis_absorb_h1e": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                       _dp, _dp, _dp]),
    "fcis_sigma_plan_create": (ctypes.c_int, [ctypes.POINTER(_vp), _vp, _dp,
                                              ctypes.c_int, ctypes.c_uint64]),
    "fcis_sigma_plan_execute": (ctypes.c_int, [_vp, _vp, _vp, _vp]),
    "fcis_sigma_plan_free": (None, [_vp]),
    "fcis_sigma_plan_info": (ctypes.c_int, [_vp, _ip, _i64p, _u64p]),
    "fcis_sigma_plan_set_h2e": (ctypes.c_int, [_vp, _dp]),
    "fcis_sigma_plan_path": (ctypes.c_int, [_vp]),
    "fcis_sigma_plan_set_timing": (ctypes.c_int, [_vp, ctypes.c_int]),
    "fcis_sigma_plan_get_timing": (ctypes.c_int, [_vp, _dp, _i64p]),
    "fcis_make_hdiag": (ctypes.c_int, [_vp, _dp, _dp, _vp, _vp]),
    "fcis_trans_rdm12s": (ctypes.c_int, [_vp, _vp, _vp, _dp, _dp, _dp, _dp, _dp, _vp]),
    "fcis_trans_rdm12s_noalloc": (ctypes.c_int, [_vp, _vp, _vp, _dp, _dp, _dp, _dp, _dp, _vp]),
    "fcis_pspace": (ctypes.c_int, [_vp, _dp, _dp, ctypes.c_int, _vp, _vp, _vp]),
    "fcis_pspace_uhf": (ctypes.c_int, [_vp, _dp, _dp, _dp, ctypes.c_int, _vp, _vp, _vp]),
    "fcis_contract_1e_ab": (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp, ctypes.c_int, _vp]),
    "fcis_csf_create": (ctypes.c_int, [ctypes.POINTER(_vp), _vp, ctypes.c_int, _vp]),
    "fcis_csf_free": (None, [_vp]),
    "fcis_csf_info": (ctypes.c_int, [_vp, _i64p, _i64p]),
    "fcis_csf_det2csf": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_int, _vp]),
    "fcis_csf_csf2det": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_int, _vp]),
    "fcis_csf_hdiag": (ctypes.c_int, [_vp, _dp, _vp, _vp, _vp]),
    "fcis_csf_symmetry": (ctypes.c_int, [_vp, _ip, ctypes.c_int, _i64p, _vp]),
    "fcis_csf_pack": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_int, _vp]),
    "fcis_csf_unpack": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_int, _vp]),
    "fcis_csf_irrep": (ctypes.c_int, [_vp, _ip, _i64, _ip]),
    "fcis_csf_describe": (ctypes.c_int, [_vp, _i64, _u64p, _u64p, _u64p]),
    "fcis_csf_max_support": (ctypes.c_int, [_vp, _ip]),
    "fcis_csf_support": (ctypes.c_int, [_vp, ctypes.c_int, _vp, ctypes.c_int, _vp, _vp, _vp, _vp]),
    "fcis_batch_create": (ctypes.c_int, [ctypes.POINTER(_vp), ctypes.c_int, _vp, _vp, _vp,
                                         ctypes.c_int, _vp]),
    "fcis_batch_free": (None, [_vp]),
    "fcis_batch_info": (ctypes.c_int, [_vp, _i64p, _vp, _ip, _ip]),
    "fcis_batch_set_active": (ctypes.c_int, [_vp, _vp, _vp]),
    "fcis_batch_set_graph": (ctypes.c_int, [_vp, ctypes.c_int]),
    "fcis_batch_sigma": (ctypes.c_int, [_vp, _vp, _vp, _vp]),
    "fcis_batch_tune": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_int, _dp, _vp]),
    "fcis_gram_list": (ctypes.c_int, [ctypes.c_int, _vp, ctypes.c_int, _vp, ctypes.c_int, _vp,
                                      ctypes.c_int, _vp, _vp]),
    "fcis_gemm_list": (ctypes.c_int, [ctypes.c_int, _vp, ctypes.c_int, _vp, ctypes.c_int, _vp,
                                      ctypes.c_int, _vp, _vp]),
    "fcis_precond_list": (ctypes.c_int, [ctypes.c_int, _vp, ctypes.c_double, _vp]),
    "fcis_dots": (ctypes.c_int, [ctypes.c_int, _i64, _vp, _i64, _vp, _dp, _vp]),
    "fcis_lincomb": (ctypes.c_int, [ctypes.c_int, _i64, _vp, _i64, _dp,
                                    ctypes.c_double, _vp, _vp]),
    "fcis_precond": (ctypes.c_int, [_i64, _vp, ctypes.c_double, ctypes.c_double,
                                    _vp, _vp, _vp]),
    "fcis_init_guess": (ctypes.c_int, [_i64, _i64, ctypes.c_int, ctypes.c_double,
                                       _vp, _vp]),
    "fcis_gas_space_create": (ctypes.c_int, [ctypes.POINTER(_vp), _vp, ctypes.c_int, _ip,
                                             ctypes.c_int, _ip, _vp]),
    "fcis_gas_space_free": (None, [_vp]),
    "fcis_gas_space_info": (ctypes.c_int, [_vp, _i64p, _ip, _ip]),
    "fcis_gas_block_info": (ctypes.c_int, [_vp, _i64p, _ip, _ip, _i64p, _i64p]),
    "fcis_gas_sector_info": (ctypes.c_int, [_vp, _ip, _i64p]),
    "fcis_gas_addr2str": (ctypes.c_int, [_vp, ctypes.c_int, _i64, _u64p]),
    "fcis_gas_str2addr": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_uint64, _i64p]),
    "fcis_gas_embed": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_int, _vp]),
    "fcis_gas_project": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_int, _vp]),
    "fcis_gas_make_hdiag": (ctypes.c_int, [_vp, _dp, _dp, _vp, _vp]),
    "fcis_gas_trans_rdm12s": (ctypes.c_int, [_vp, _vp, _vp, _dp, _dp, _dp, _dp, _dp, _vp]),
    "fcis_gas_sigma_plan_create": (ctypes.c_int, [ctypes.POINTER(_vp), _vp, _dp, ctypes.c_int,
                                                  ctypes.c_uint64]),
    "fcis_gas_sigma_plan_execute": (ctypes.c_int, [_vp, _vp, _vp, _vp]),
    "fcis_gas_sigma_plan_free": (None, [_vp]),
    "fcis_gas_sigma_execute": (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp, _vp]),
    "fcis_gas_rows_info": (ctypes.c_int, [_vp, _i64p, _i64p, _i64p]),
    "fcis_gas_det_index": (ctypes.c_int, [_vp, _i64, _vp, _vp, _vp]),
    "fcis_gas_contract_ss": (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp, _vp]),
    "fcis_gas_space_create_native": (ctypes.c_int, [ctypes.POINTER(_vp), ctypes.c_int, _ip,
                                                    ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                                    _ip, _vp]),
    "fcis_gas_native_info": (ctypes.c_int, [_vp, _ip, _i64p, _i64p, _i64p, _i64p]),
    "fcis_gas_det_strings": (ctypes.c_int, [_vp, _i64, _i64p, _u64p, _u64p]),
    "fcis_gas_pspace": (ctypes.c_int, [_vp, _dp, _dp, ctypes.c_int, _i64p, _vp, _vp]),
    "fcis_gas_ss_diag": (ctypes.c_int, [_vp, _vp, _vp]),
    "fcis_gas_sigma_plan_path": (ctypes.c_int, [_vp]),
    "fcis_spin_plan_create": (ctypes.c_int, [ctypes.POINTER(_vp), _vp, _vp]),
    "fcis_spin_plan_contract": (ctypes.c_int, [_vp, _vp, _vp, _vp]),
    "fcis_spin_plan_diag": (ctypes.c_int, [_vp, _vp, _vp]),
    "fcis_spin_plan_free": (None, [_vp]),
    "fcis_spin_ladder": (ctypes.c_int, [_vp, _vp, ctypes.c_int, _vp, _vp, _vp]),
}

_LIB = None


def load(path=None):
    """Load and configure the HIP library; raise if it is absent."""
    global _LIB
    if _LIB is not None and path is None:
        return _LIB
    p = path or os.environ.get("FCISIGMA_LIB") or LIBPATH
    if not os.path.exists(p):
        raise RuntimeError(
            "libfcisigma.so not found at %s: build it with "
            "`python -c 'import __graft_entry__ as g; g.build()'`" % p)
    # torch ships its own libamdhip64.so.7; loading torch first makes the
    # library bind to that runtime (same soname), so that device memory,
    # streams and kernels share one HIP runtime in the process.
    import torch  # noqa: F401
    lib = ctypes.CDLL(p)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _LIB = lib
    return lib


def check(status, name):
    if status != SUCCESS:
        msg = load().fcis_last_error()
        raise RuntimeError("%s failed with status %d: %s" % (
            name, status, msg.decode() if msg else ""))


def dptr(a):
    """double* of a C-contiguous float64 numpy array."""
    return a.ctypes.data_as(_dp)


def vptr(t):
    """void* of a device tensor (or None)."""
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def stream_ptr(stream=None):
    import torch
    s = stream if stream is not None else torch.cuda.current_stream()
    return ctypes.c_void_p(s.cuda_stream)
