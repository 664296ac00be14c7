"""Numerical knobs, with VULCAN's names and defaults (vulcan_cfg.py:124-159).

``Config`` is a plain attribute bag so that host code reads the same names the
reference reads from ``vulcan_cfg``.  Condensation, settling and ions are
outside the accelerated path; their switches exist only so that a
configuration asking for them is rejected loudly instead of being silently
ignored.
"""
from __future__ import annotations

import dataclasses
from typing import Dict, List

# phy_const.py:3-4 (cgs); the kernels use the same literals
kb = 1.38064852E-16
Navo = 6.02214086E+23
r_jup = 7.1492E9


@dataclasses.dataclass
class Config:
    # vulcan_cfg.py:128; the reference dispatches on the name
    # (getattr(op, ode_solver)), so a registered device solver may carry any
    # name (for example 'Ros2_amd'); the step is Ros2 whatever it says
    ode_solver: str = 'Ros2'
    use_moldiff: bool = True          # vulcan_cfg.py:76
    use_vm_mol: bool = False          # vulcan_cfg.py:77
    use_vz: bool = False              # vulcan_cfg.py:78
    use_settling: bool = False
    use_condense: bool = False
    use_ion: bool = False
    use_topflux: bool = False
    use_botflux: bool = False
    use_fix_sp_bot: Dict[str, float] = dataclasses.field(default_factory=dict)   # vulcan_cfg.py:100
    diff_esc: List[str] = dataclasses.field(default_factory=lambda: ['H'])      # vulcan_cfg.py:101
    max_flux: float = 1e13                                                      # vulcan_cfg.py:102
    atm_base: str = 'H2'
    rocky: bool = False
    gs: float = 2140.0                # vulcan_cfg.py:84
    Rp: float = 1.138 * r_jup         # vulcan_cfg.py:46
    update_frq: int = 100             # vulcan_cfg.py:94
    st_factor: float = 0.5            # vulcan_cfg.py:124
    conv_step: int = 500              # vulcan_cfg.py:125
    dttry: float = 1.E-10             # vulcan_cfg.py:132
    trun_min: float = 1e2
    runtime: float = 1.E22
    dt_min: float = 1.E-14
    dt_max: float = 1.E22 * 1e-5      # runtime*1e-5, vulcan_cfg.py:136
    dt_var_max: float = 2.
    dt_var_min: float = 0.5
    count_min: int = 120
    count_max: int = int(3E4)
    atol: float = 1.E-1
    mtol: float = 1.E-22
    mtol_conv: float = 1.E-20
    pos_cut: float = 0
    nega_cut: float = -1.
    loss_eps: float = 1e-1
    yconv_cri: float = 0.01
    slope_cri: float = 1.e-4
    yconv_min: float = 0.1
    flux_cri: float = 0.1
    conver_ignore: List[str] = dataclasses.field(default_factory=list)
    rtol: float = 0.25                # vulcan_cfg.py:156
    # photochemistry (vulcan_cfg.py:43-60, 151); the radiative transfer runs
    # on the device when the caller provides its tables (vulcan_amd.photo)
    use_photo: bool = False
    ini_update_photo_frq: int = 100
    final_update_photo_frq: int = 5
    f_diurnal: float = 1.
    sl_angle: float = 48 / 180. * 3.14159
    edd: float = 0.5
    flux_atol: float = 1.
    scat_sp: List[str] = dataclasses.field(default_factory=lambda: ['H2', 'He'])
    T_cross_sp: List[str] = dataclasses.field(default_factory=list)
    remove_list: List[int] = dataclasses.field(default_factory=list)
    # iterative-refinement steps per block solve (double-double residual);
    # no counterpart in the reference, whose banded LU needs none
    refine: int = 1

    def check_supported(self) -> None:
        """Refuse switches whose physics is not on the accelerated path."""
        bad = [name for name in ('use_settling', 'use_condense', 'use_ion')
               if getattr(self, name)]
        if bad:
            raise NotImplementedError(
                'configuration switches not supported by the device Ros2 path: '
                + ', '.join(bad))
