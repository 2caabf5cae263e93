"""qlattice_amd -- MI355X (gfx950) domain-wall fermion operator and inverter.

A drop-in for Qlattice's even-odd Moebius domain-wall path
(qlat/qlat/lib/dslash.cpp, qlat/qlat/include/qlat/dslash.h) with the same
Python surface as ``qlat`` for it: ``FermionAction``, ``InverterDomainWall``,
``GaugeField``, ``Prop``, the point / wall / momentum / U(1) sources and
their ``*_src_propagator`` solves, and the ``multiply_*`` / ``invert`` /
``cg_with_herm_sym_2`` functions.  All arithmetic runs in the
hand-written HIP library ``libqlat_dwf_hip.so`` (C ABI: include/qlat_dwf.h),
reached through the Cython binding ``_cdwf``.
"""

from .fermion_action import FermionAction
from .fields import (
    Coordinate,
    CoordinateD,
    FermionField4d,
    FieldComplexD,
    FermionField5d,
    GaugeField,
    Geometry,
    Prop,
    fermion_field_4d_from_5d,
    fermion_field_5d_from_4d,
    get_half_fermion,
    set_half_fermion,
    set_point_src_fermion_field,
)
from .inverter import (
    Inverter,
    InverterDomainWall,
    InverterDwfFreeField,
    LowModes,
    bench_kernel,
    cg_with_herm_sym_2,
    deflate,
    dot_product,
    find_max_eigen_value_hermop_sym2,
    free_invert,
    invert,
    invert_batch,
    invert_with_cg_with_guess,
    multiply_d_minus,
    multiply_hermop_sym2,
    multiply_hermop_sym2_fp32,
    multiply_m,
    multiply_m_dwf,
    multiply_m_e_e,
    multiply_m_e_e_inv,
    multiply_m_e_o,
    multiply_m_eo_eo,
    multiply_m_full,
    multiply_m_with_prec_sym2,
    multiply_mdag,
    multiply_mdag_e_e,
    multiply_mdag_e_e_inv,
    multiply_mdag_e_o,
    multiply_mdag_eo_eo,
    multiply_mpc_sym2,
    multiply_mpcdag_sym2,
    multiply_wilson_d_e_o_no_comm,
    multiply_wilson_ddag_e_o_no_comm,
    qnorm,
)
from ._lib import QlatDwfError, load as load_library
from .force import fermion_force, pseudofermion_action, set_gauge
from .eigen_io import (
    CompressedEigenSystemInfo,
    load_compressed_eigen_vectors,
    load_low_modes,
    read_compressed_eigen_system_info,
    read_eigen_values,
    save_compressed_eigen_vectors,
)
from .gauge_io import (
    GaugeFieldInfo,
    gf_avg_link_trace,
    gf_avg_plaq,
    load_gauge_field,
    read_gauge_field_header,
    save_gauge_field,
    unitarize,
)
from .qed import cg_with_m_dwf_qed, invert_dwf_qed, invert_qed, multiply_m_dwf_qed
from .rng import RngState, get_data_sig, get_data_sig_arr, set_g_rand, set_g_rand_device, set_u_rand, set_u_rand_device
from .sources import (
    mk_point_src,
    mk_rand_vol_u1_src,
    mk_wall_src,
    set_mom_src,
    set_mom_src_propagator,
    set_point_src,
    set_point_src_propagator,
    set_rand_vol_u1,
    set_rand_vol_u1_src,
    set_tslice_mom_src,
    set_tslice_mom_src_propagator,
    set_volume_src,
    set_wall_src,
    set_wall_src_propagator,
)

__all__ = [n for n in dir() if not n.startswith("_")]
