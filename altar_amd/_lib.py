"""ctypes binding of libaltar_amd.so (declarations: include/altar_amd.h).

The library is required: importing a function from here on a machine without
the built library raises ImportError, and every nonzero status raises
AltarError with the library's message.  There is no fallback path.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libaltar_amd.so")

# status codes, models, whitening kinds, precisions (include/altar_amd.h)
OK, EINVAL, EHIP, ENOMEM, ESTATE = 0, 1, 2, 3, 4
MOGI, CDM, LINEAR, JOINT = 0, 1, 2, 3
WHITEN_DIAG, WHITEN_LOWER, WHITEN_BLOCKS = 1, 2, 3
FP64, FP32, BF16, BF16_REFINED = 0, 1, 2, 3
PRECISIONS = {"fp64": FP64, "fp32": FP32, "bf16": BF16, "bf16r": BF16_REFINED}

# Philox stream tags (altar_amd/csrc/common.hpp)
TAG_PROPOSAL, TAG_ACCEPT, TAG_RESAMPLE, TAG_INIT, TAG_SCREEN = 0, 1, 2, 3, 4
TAG_TRAJ_PROPOSAL, TAG_TRAJ_ACCEPT = 5, 6
TRAJ_MAX_SUBSTEPS = 1024   # sub-step streams use step' = step * 1024 + sub


class AltarError(RuntimeError):
    def __init__(self, code, message):
        super().__init__(f"altar_amd error {code}: {message}")
        self.code = code


_lib = None

_vp = ctypes.c_void_p
_i = ctypes.c_int
_d = ctypes.c_double
_u32 = ctypes.c_uint32
_u64 = ctypes.c_uint64
_sz = ctypes.c_size_t
_ull = ctypes.c_ulonglong

SIGNATURES = {
    "altar_last_error": (ctypes.c_char_p, []),
    "altar_version": (_i, []),
    "altar_source_create": (_i, [ctypes.POINTER(_vp), _i, _i, _i, _d]),
    "altar_source_destroy": (_i, [_vp]),
    "altar_source_layout": (_i, [_vp, _vp, _i]),
    "altar_source_observations": (_i, [_vp, _vp, _vp, _vp, _vp, _vp]),
    "altar_source_green": (_i, [_vp, _vp, _vp, _i]),
    "altar_source_whitening": (_i, [_vp, _i, _vp, _d]),
    "altar_source_whitening_blocks": (_i, [_vp, _i, _vp, _vp, _vp, _d]),
    "altar_source_forward": (_i, [_vp, _vp, _i, _vp, _sz, _i, _vp]),
    "altar_source_residuals": (_i, [_vp, _vp, _i, _vp, _sz, _i, _vp]),
    "altar_source_data_llk": (_i, [_vp, _vp, _i, _vp, _i, _vp]),
    "altar_source_verify": (_i, [_vp, _vp, _i, _vp, _vp]),
    "altar_source_data_llk_fields": (_i, [_vp, _vp, _i, _vp, _i, _vp, _sz, _vp]),
    "altar_source_fp64_sides": (_i, [_vp, _d]),
    "altar_source_side_census": (_i, [_vp, _i, _vp, _vp]),
    "altar_source_timing": (_i, [_vp, _i]),
    "altar_source_timers": (_i, [_vp, _vp, _vp, _i]),
    "altar_prior_create": (_i, [ctypes.POINTER(_vp), _i, _vp, _vp, _vp]),
    "altar_prior_destroy": (_i, [_vp]),
    "altar_prior_sample": (_i, [_vp, _vp, _i, _u64, _u32, _u32, _vp, _vp]),
    "altar_prior_llk": (_i, [_vp, _vp, _i, _vp, _vp]),
    "altar_prior_verify": (_i, [_vp, _vp, _i, _vp, _vp]),
    "altar_mh_propose": (_i, [_vp, _i, _i, _vp, _u64, _u32, _u32, _vp, _vp, _vp]),
    "altar_mh_restore": (_i, [_vp, _vp, _vp, _i, _i, _vp]),
    "altar_mh_accept": (_i, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _d, _i, _i, _u64, _u32, _u32,
                             _vp, _vp, _vp, _vp]),
    "altar_mh_scaling": (_d, [_d, _d, _ull, _ull, _ull]),
    "altar_cov_dbeta_grid": (_i, [_vp, _i, _d, _d, _d, _d, _vp, ctypes.POINTER(_d), ctypes.POINTER(_d)]),
    "altar_cov_dbeta_brent": (_i, [_vp, _i, _d, _d, _d, _d, _i, _vp, ctypes.POINTER(_d), ctypes.POINTER(_d)]),
    "altar_cov_covariance": (_i, [_vp, _vp, _i, _i, _vp, _vp, _vp]),
    "altar_cov_multiplicities": (_i, [_vp, _i, _u64, _u32, _vp]),
    "altar_cov_update_dev": (_i, [_vp, _i, _d, _d, _d, _vp, _vp, _vp]),
    "altar_cov_resample_dev": (_i, [_vp, _i, _u64, _u32, _vp, _vp, _vp, _vp, _vp]),
    "altar_cov_resample_dev_noalloc": (_i, [_vp, _i, _u64, _u32, _vp, _vp, _vp]),
    "altar_resample_gather": (_i, [_vp, _vp, _vp, _vp, _vp, _i, _i, _vp, _vp, _vp, _vp, _vp]),
    "altar_posterior": (_i, [_vp, _vp, _d, _i, _vp, _vp]),
    "altar_random_normals": (_i, [_vp, _i, _i, _u64, _u32, _u32, _u32, _vp]),
    "altar_gemm_f32": (_i, [_vp, _sz, _vp, _sz, _i, _i, _i, _vp, _vp, _sz, _vp]),
    "altar_gather_rows_f32": (_i, [_vp, _sz, _vp, _i, _i, _vp, _sz, _vp]),
    "altar_transpose_f32": (_i, [_vp, _sz, _i, _i, _vp, _sz, _vp]),
    "altar_mlp_create": (_i, [ctypes.POINTER(_vp), _i, _i, _i, _i, _i]),
    "altar_mlp_destroy": (_i, [_vp]),
    "altar_mlp_input_scaling": (_i, [_vp, _vp, _vp]),
    "altar_mlp_set_layer": (_i, [_vp, _i, _vp, _vp]),
    "altar_mlp_get_layer": (_i, [_vp, _i, _vp, _vp, _i]),
    "altar_mlp_forward": (_i, [_vp, _vp, _i, _vp, _i, _vp]),
    "altar_mlp_backward": (_i, [_vp, _vp, _vp, _i, _vp, _i, _vp, _vp]),
    "altar_mlp_adam": (_i, [_vp, ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_float, _vp]),
    "altar_mlp_train_epoch": (_i, [_vp, _vp, _vp, _i, _vp, _vp, _i, _i, _i, _i, ctypes.c_float, ctypes.c_float,
                                   ctypes.c_float, ctypes.c_float, ctypes.c_float, _vp, _vp]),
    "altar_surrogate_llk": (_i, [_vp, _i, _i, _i, _vp, _vp, _d, _d, _vp, _vp]),
    "altar_da_screen": (_i, [_vp, _vp, _vp, _vp, _d, _i, _u64, _u32, _u32, _vp, _vp, _vp, _vp, _vp]),
    "altar_da_pack": (_i, [_vp, _vp, _i, _i, _vp, _vp]),
    "altar_da_accept": (_i, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _i, _d, _i, _u64, _u32, _u32,
                             _vp, _vp, _vp]),
    "altar_da_trajectory_propose": (_i, [_vp, _i, _i, _vp, _u64, _u32, _u32, _u32, _vp, _vp, _vp]),
    "altar_da_trajectory_step": (_i, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _d, _i, _i, _u64, _u32, _u32, _u32,
                                      _vp, _vp, _vp]),
    "altar_da_trajectory_list": (_i, [_vp, _i, _vp, _vp]),
    "altar_da_trajectory_fused": (_i, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _i, _d, _d,
                                       _d, _i, _i, _u64, _u32, _u32, _i, _vp, _vp, _vp]),
}


def lib():
    """The loaded library (raises ImportError when it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"altar_amd: {LIB_PATH} is missing; build it with "
            "`python -c 'import __graft_entry__ as g; g.build()'` (hipcc, gfx950)")
    handle = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(handle, name)
        fn.restype = res
        fn.argtypes = args
    _lib = handle
    return _lib


def check(rc):
    if rc != OK:
        msg = lib().altar_last_error().decode(errors="replace")
        raise AltarError(rc, msg)


def call(name, *args):
    check(getattr(lib(), name)(*args))


def ptr(t):
    """Device or host address of a contiguous torch tensor / numpy array (None -> NULL)."""
    if t is None:
        return None
    if hasattr(t, "data_ptr"):
        assert t.is_contiguous(), "altar_amd needs contiguous tensors"
        return ctypes.c_void_p(t.data_ptr())
    assert t.flags["C_CONTIGUOUS"], "altar_amd needs contiguous arrays"
    return ctypes.c_void_p(t.ctypes.data)


def stream():
    """The current torch stream as a hipStream_t."""
    import torch
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
