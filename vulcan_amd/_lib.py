"""ctypes binding of libvulcan_amd.so (declarations: include/vulcan_amd.h)."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, 'libvulcan_amd.so')
# the same library with the chemistry kernels' table indexes bounds-checked
# (vk_debug_checks); a debugging aid, not loaded by the product path
CHECKED_LIB_PATH = os.path.join(_HERE, 'libvulcan_amd_checked.so')

c_int_p = C.POINTER(C.c_int)
c_double_p = C.POINTER(C.c_double)


class NetworkDesc(C.Structure):
    _fields_ = [
        ('ni', C.c_int), ('npairs', C.c_int), ('natom', C.c_int),
        ('reac', c_int_p), ('prod', c_int_p), ('kind', c_int_p), ('par', c_double_p), ('rev', c_int_p),
        ('nasa9', c_double_p), ('atom_mat', c_double_p), ('ms', c_double_p), ('alpha', c_double_p),
        ('chem_ptr', c_int_p), ('chem_pair', c_int_p), ('chem_sign', c_int_p), ('nchem', C.c_int),
        ('slot_pair', c_int_p), ('slot_sp', c_int_p), ('nslot', C.c_int),
        ('jac_map', c_int_p), ('jac_ptr', c_int_p), ('jac_slot', c_int_p), ('jac_coef', c_double_p),
        ('njnz', C.c_int), ('njac', C.c_int),
    ]


class Atm(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in ('Tco', 'M', 'Kzz', 'Dzz', 'dzi', 'Ti', 'Hpi', 'g', 'Hp', 'tau',
                                          'vz', 'vm', 'top_flux', 'bot_flux', 'bot_vdep', 'fix_bot')] + \
        [('use_topflux', C.c_int), ('use_botflux', C.c_int), ('esc', C.c_uint * 5), ('max_flux', C.c_double)]


class Geom(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in ('lnp', 'zco', 'dz', 'mu', 'gs', 'Rp', 'pref')]


class Params(C.Structure):
    _fields_ = [(n, C.c_double) for n in (
        'dt_min', 'dt_max', 'dt_var_min', 'dt_var_max', 'rtol', 'atol', 'mtol', 'mtol_conv', 'pos_cut',
        'nega_cut', 'loss_eps', 'trun_min', 'runtime', 'st_factor', 'yconv_cri', 'slope_cri', 'yconv_min',
        'flux_cri')] + [(n, C.c_int) for n in (
            'count_min', 'count_max', 'conv_step', 'update_frq', 'use_moldiff', 'use_conv', 'use_update_mu',
            'refine')] + [('conv_ignore', C.c_uint * 5)]


class PhotoDesc(C.Structure):
    _fields_ = [(n, C.c_int) for n in ('nbin', 'nabs', 'nscat', 'nbranch', 'nT', 'nTJ', 'din12')] + \
        [(n, c_int_p) for n in ('abs_sp', 'abs_T', 'scat_sp', 'br_col', 'br_T')] + \
        [(n, c_double_p) for n in ('bins', 'sflux_top', 'cross', 'cross_T', 'cross_scat', 'cross_J', 'cross_J_T')] + \
        [(n, C.c_double) for n in ('dbin1', 'dbin2', 'cos_sl', 'edd', 'f_diurnal', 'flux_atol', 'hc')]


COLSTATE_DTYPE = np.dtype([
    ('t', 'f8'), ('dt', 'f8'), ('delta', 'f8'), ('longdy', 'f8'), ('longdydt', 'f8'), ('small_y', 'f8'),
    ('nega_y', 'f8'), ('aflux_change', 'f8'), ('count', 'i4'), ('attempts', 'i4'), ('nega_count', 'i4'),
    ('loss_count', 'i4'), ('delta_count', 'i4'), ('end_case', 'i4'), ('retry', 'i4'), ('done', 'i4')])
assert COLSTATE_DTYPE.itemsize == 96

VP = C.c_void_p
_SIGS = {
    'vk_plan_create': [C.POINTER(NetworkDesc), C.POINTER(VP)],
    'vk_plan_destroy': [VP],
    'vk_plan_info': [VP, c_int_p, c_int_p, c_int_p],
    'vk_columns_create': [C.c_int, c_int_p, C.POINTER(VP)],
    'vk_columns_destroy': [VP],
    'vk_columns_layers': [VP, c_int_p],
    'vk_rates': [VP, VP, C.POINTER(Atm), VP, VP],
    'vk_chemdf': [VP, VP, VP, VP, VP, VP, VP],
    'vk_diffdf': [VP, VP, C.POINTER(Atm), VP, C.c_int, VP, VP],
    'vk_rhs': [VP, VP, C.POINTER(Atm), VP, VP, C.c_int, VP, VP, VP, VP],
    'vk_jac': [VP, VP, VP, VP, VP, VP, VP],
    'vk_lhs': [VP, VP, C.POINTER(Atm), VP, VP, C.c_int, VP, VP, VP, VP, VP],
    'vk_block_factor': [VP, VP, VP, VP, VP, VP, VP, VP],
    'vk_plan_set_factor_variant': [VP, C.c_int],
    'vk_plan_factor_variant': [VP],
    'vk_tune_factor': [VP, VP, C.c_int, C.POINTER(C.c_float), VP],
    'vk_block_solve_system': [VP, VP, VP, VP, VP, VP, VP, C.c_int, VP, VP],
    'vk_block_solve_system_noalloc': [VP, VP, VP, VP, VP, VP, VP, C.c_int, VP, VP],
    'vk_selftest_mfma_f64': [VP, VP],
    'vk_selftest_wave_max_u32': [VP, VP, VP],
    'vk_selftest_recip_f64': [C.c_int, VP, VP, VP],
    'vk_block_solve': [VP, VP, VP, VP, VP, VP, VP],
    'vk_block_residual': [VP, VP, VP, VP, VP, VP, VP, VP, VP],
    'vk_integrator_create': [VP, VP, C.c_int, C.POINTER(VP)],
    'vk_integrator_destroy': [VP],
    'vk_ros2_solver': [VP, C.POINTER(Atm), C.POINTER(Params), VP, VP, VP, VP, VP],
    'vk_integrate': [VP, C.POINTER(Atm), C.POINTER(Geom), C.POINTER(Params), VP, VP, VP, VP, VP, VP, VP, VP,
                     C.c_int, VP],
    'vk_conv': [VP, VP, C.POINTER(Atm), C.POINTER(Params), VP, VP, VP, C.c_int, VP, C.c_int, VP, VP],
    'vk_integrator_history': [VP, C.c_int, C.c_int, C.c_int, VP, VP],
    'vk_photo_create': [VP, VP, C.POINTER(PhotoDesc), C.POINTER(VP)],
    'vk_photo_destroy': [VP],
    'vk_photo_update': [VP, VP, VP, VP, VP, VP, VP, VP, C.c_int, C.c_int, C.c_int, C.c_double, VP],
    'vk_photo_read': [VP, C.c_int, VP],
    'vk_integrator_set_photo': [VP, VP, C.c_int, C.c_int],
    'vk_update_mu_dz': [VP, VP, C.POINTER(Atm), C.POINTER(Geom), VP, VP, C.c_int, VP],
    'vk_update_phi_esc': [VP, VP, C.POINTER(Atm), VP, VP],
    'vk_debug_checks': [C.POINTER(C.c_int), C.c_int],
    'vk_atom_sums': [VP, VP, VP, VP, VP],
    'vk_integrator_set_timing': [VP, C.c_int],
    'vk_integrator_set_graph': [VP, C.c_int],
    'vk_integrator_stats': [VP, c_double_p, C.POINTER(C.c_longlong)],
    'vk_mlp_gemm': [C.c_int, C.c_int, C.c_int, VP, C.c_long, C.c_long, VP, C.c_long, C.c_long, VP, C.c_long, VP,
                    C.c_int, VP, C.c_long, VP],
    'vk_mlp_silu_bwd': [C.c_long, VP, VP, VP, VP, VP],
    'vk_mlp_mse': [C.c_long, VP, VP, VP, C.c_float, VP, VP, VP, VP],
    'vk_mlp_colsum': [C.c_int, C.c_int, VP, VP, VP],
    'vk_mlp_to_bf16': [C.c_long, VP, VP, VP],
    'vk_mlp_adam': [C.c_long, VP, VP, VP, VP, C.c_float, C.c_float, C.c_float, C.c_float, C.c_int, C.c_float, VP,
                    VP],
}
EXPORTS = tuple(_SIGS)

_lib = None


class VulcanAmdError(RuntimeError):
    pass


def load() -> C.CDLL:
    """Load the HIP library; raise if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise VulcanAmdError(
            f'{LIB_PATH} is missing: build it with `python vulcan_amd/csrc/build.py` '
            '(or __graft_entry__.build()); there is no CPU fallback')
    lib = C.CDLL(LIB_PATH)
    for name, args in _SIGS.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = C.c_int
    _lib = lib
    return lib


def check(rc: int, what: str) -> None:
    if rc != 0:
        raise VulcanAmdError(f'{what} failed with code {rc}')
