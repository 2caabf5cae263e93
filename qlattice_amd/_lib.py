"""Access to the HIP library through its Cython binding (_cdwf.pyx).

The shared objects are the in-tree builds ``qlattice_amd/libqlat_dwf_hip.so``
and ``qlattice_amd/_cdwf.<abi>.so`` (``python -m qlattice_amd.build``).
There is no fallback: if either is missing or cannot be loaded every entry
point raises, so nothing in this package silently computes on the CPU.
"""

import os

import numpy as np

from ._errors import QlatDwfError

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libqlat_dwf_hip.so")

QLAT_DWF_SOLVER_FP64 = 0
QLAT_DWF_SOLVER_MIXED_FP32 = 1
QLAT_DWF_SOLVER_MIXED_FP16 = 2
SRC_POINT, SRC_WALL, SRC_MOM, SRC_U1 = 0, 1, 2, 3

# Every symbol declared in include/qlat_dwf.h (checked against the exports of
# the built library by the tests).
EXPORTED = [
    "qlat_dwf_last_error",
    "qlat_dwf_abi_version",
    "qlat_dwf_inverter_create",
    "qlat_dwf_inverter_destroy",
    "qlat_dwf_inverter_set_params",
    "qlat_dwf_inverter_get_params",
    "qlat_dwf_inverter_set_solver",
    "qlat_dwf_inverter_get_stats",
    "qlat_dwf_set_checking_invert",
    "qlat_dwf_is_checking_invert",
    "qlat_dwf_multiply_m_e_e",
    "qlat_dwf_multiply_mdag_e_e",
    "qlat_dwf_multiply_m_e_e_inv",
    "qlat_dwf_multiply_mdag_e_e_inv",
    "qlat_dwf_multiply_wilson_d_e_o",
    "qlat_dwf_multiply_wilson_ddag_e_o",
    "qlat_dwf_multiply_m_e_o",
    "qlat_dwf_multiply_mdag_e_o",
    "qlat_dwf_multiply_mpc_sym2",
    "qlat_dwf_multiply_mpcdag_sym2",
    "qlat_dwf_multiply_hermop_sym2",
    "qlat_dwf_multiply_hermop_sym2_fp32",
    "qlat_dwf_multiply_m",
    "qlat_dwf_multiply_mdag",
    "qlat_dwf_multiply_d_minus",
    "qlat_dwf_dot_product",
    "qlat_dwf_qnorm",
    "qlat_dwf_cg_with_herm_sym_2",
    "qlat_dwf_invert_5d",
    "qlat_dwf_invert_5d_with_guess",
    "qlat_dwf_invert_4d",
    "qlat_dwf_invert_prop",
    "qlat_dwf_free_invert",
    "qlat_dwf_rng_fill_uniform",
    "qlat_dwf_rng_fill_gaussian",
    "qlat_dwf_multiply_m_dwf_qed",
    "qlat_dwf_cg_with_m_dwf_qed",
    "qlat_dwf_invert_dwf_qed",
    "qlat_dwf_inverter_set_low_modes",
    "qlat_dwf_deflate",
    "qlat_dwf_make_source",
    "qlat_dwf_make_source_device",
    "qlat_dwf_source_propagator",
    "qlat_dwf_source_propagator_device",
    "qlat_dwf_find_max_eigen_value_hermop_sym2",
    "qlat_dwf_find_max_eigen_value_hermop_sym2_rng",
    "qlat_dwf_multiply_hermop_sym2_device",
    "qlat_dwf_multiply_hermop_sym2_device_noalloc",
    "qlat_dwf_invert_5d_device",
    "qlat_dwf_invert_5d_with_guess_device",
    "qlat_dwf_invert_prop_device",
    "qlat_dwf_bench_kernel",
    "qlat_dwf_fermion_force",
    "qlat_dwf_fermion_force_device",
    "qlat_dwf_invert_5d_batch",
    "qlat_dwf_invert_5d_batch_device",
    "qlat_dwf_inverter_set_gauge",
]


_mod = None


def load():
    """The Cython binding module (_cdwf), or raise if it is not built."""
    global _mod
    if _mod is not None:
        return _mod
    if not os.path.exists(LIB_PATH):
        raise QlatDwfError(f"{LIB_PATH} is missing: build it with `python -m qlattice_amd.build`")
    try:
        from . import _cdwf
    except ImportError as e:
        raise QlatDwfError(f"the Cython binding qlattice_amd/_cdwf is not built or cannot load ({e}): "
                           "build it with `python -m qlattice_amd.build`") from e
    _mod = _cdwf
    return _mod


def ptr(a):
    """Address of a C-contiguous complex128 numpy array."""
    if not isinstance(a, np.ndarray) or a.dtype != np.complex128 or not a.flags.c_contiguous:
        raise TypeError("expected a C-contiguous complex128 numpy array")
    return a.ctypes.data


def addr(a, dtype=None):
    """Address of a C-contiguous numpy array (of `dtype` when given); 0 for None."""
    if a is None:
        return 0
    if not isinstance(a, np.ndarray) or not a.flags.c_contiguous:
        raise TypeError("expected a C-contiguous numpy array")
    if dtype is not None and a.dtype != np.dtype(dtype):
        raise TypeError(f"expected a {np.dtype(dtype)} array, got {a.dtype}")
    return a.ctypes.data


def dev_ptr(t):
    """Address of a contiguous complex128 torch tensor on the GPU (device
    memory is plumbing only; the computation is in the HIP library)."""
    import torch

    if not (isinstance(t, torch.Tensor) and t.is_cuda and t.dtype == torch.complex128
            and t.is_contiguous()):
        raise TypeError("expected a contiguous complex128 CUDA/HIP tensor")
    return t.data_ptr()
