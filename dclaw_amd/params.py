"""D-Claw step parameters: the module globals the reference reads on the step path.

The reference passes no physics through arguments: ``geoclaw_module`` (grav,
dry_tolerance, speed_limit, manning/friction), ``digclaw_module``
(digclaw_module.f90:15-52, read by ``set_dig`` :129-166) and ``amr_module``
(method, mthlim, mwaves) are module globals.  Here they are one explicit
record, passed to every call and, for batched ensembles, one record per
member.

Defaults follow the reference's config writer ``DClawInputData``
(src/python/dclaw/data.py:315-351) and the standard D-Claw numerics (order 2,
transverse 2, MC limiter on all 5 waves, f-waves, Godunov source split;
examples/radial_slide/setrun.py:174-223).
"""
from __future__ import annotations

import ctypes
import dataclasses
import math
from typing import List

# Clawpack boundary-condition codes (bc2amr): 1 = extrap, 3 = wall.
BC_EXTRAP = 1
BC_WALL = 3


@dataclasses.dataclass
class DigParams:
    # geoclaw_module
    grav: float = 9.81
    dry_tolerance: float = 1.0e-3
    speed_limit: float = 20.0
    manning_coefficient: float = 0.0
    friction_depth: float = 1.0e6
    friction_forcing: int = 0
    # digclaw_module (data.py:315-351 defaults)
    rho_s: float = 2700.0
    rho_f: float = 1000.0
    m_crit: float = 0.62
    mref: float = 0.60
    kref: float = 1.0e-4
    phi_deg: float = 40.0
    delta: float = 0.01
    mu: float = 0.001
    alpha_c: float = 0.01
    c1: float = 1.0
    sigma_0: float = 1.0e3
    kappa: float = 1.0          # digclaw_module.f90:113
    beta_seg: float = 0.0
    delta_kr_order: float = 2.0
    manning_max: float = 0.06
    dd_manning: int = 0
    src2method: int = 0
    alphamethod: int = 0
    rarecorrectortest: int = 0
    segregation: int = 0
    bed_normal: int = 0
    entrainment: int = 0
    entrainment_method: int = 1      # data.py:339 default (1: no update yet, src2.f90:258)
    entrainment_rate: float = 0.2    # data.py:340
    me: float = 0.65                 # solid fraction of the erodible layer
    chie: float = 0.5                # data.py:346
    curvature: int = 0
    # amr_module / clawdata numerics
    order: int = 2
    transverse: int = 2
    mthlim: List[int] = dataclasses.field(default_factory=lambda: [4, 4, 4, 4, 4])
    mwaves: int = 5

    @property
    def phi(self) -> float:
        """Basal friction angle in radians, as set_dig stores it
        (digclaw_module.f90:135-136: phi = deg2rad*phi)."""
        return (math.pi / 180.0) * self.phi_deg

    def validate(self) -> None:
        """Reject configurations the step path does not implement, with the
        same conflicts the reference's writer rejects (data.py:357-365)."""
        if self.bed_normal != 0:
            raise ValueError("bed_normal=1 is not supported (the reference rejects it too: "
                             "data.py:357-361)")
        if self.src2method == 2 and self.alphamethod < 1:
            raise ValueError("src2method == 2 requires alphamethod = 1 (data.py:362-365)")
        if self.src2method not in (-1, 0, 1, 2):
            raise ValueError("src2method must be -1, 0, 1 or 2")
        if self.curvature not in (0, 1):
            raise ValueError("curvature must be 0 or 1")
        if self.curvature and self.entrainment:
            raise ValueError("curvature with entrainment is not supported: src2 then reads "
                             "the bed of cells it has already eroded in the same sweep")
        if self.entrainment not in (0, 1) or self.entrainment_method not in (0, 1):
            raise ValueError("entrainment and entrainment_method must be 0 or 1")
        if self.mwaves != 5:
            raise ValueError("D-Claw's rpn2 produces exactly 5 waves")
        if self.order not in (1, 2) or self.transverse not in (0, 1, 2):
            raise ValueError("order must be 1 or 2 and transverse 0, 1 or 2")
        if any(m not in (0, 1, 2, 3, 4, 5) for m in self.mthlim):
            raise ValueError("mthlim entries must be 0..5")


class CParams(ctypes.Structure):
    """Mirror of ``dclaw_params`` in include/dclaw_hip.h (same field order)."""

    _fields_ = [
        ("grav", ctypes.c_double),
        ("dry_tolerance", ctypes.c_double),
        ("speed_limit", ctypes.c_double),
        ("rho_s", ctypes.c_double),
        ("rho_f", ctypes.c_double),
        ("m_crit", ctypes.c_double),
        ("mref", ctypes.c_double),
        ("kref", ctypes.c_double),
        ("phi", ctypes.c_double),
        ("delta", ctypes.c_double),
        ("mu", ctypes.c_double),
        ("alpha_c", ctypes.c_double),
        ("c1", ctypes.c_double),
        ("sigma_0", ctypes.c_double),
        ("kappa", ctypes.c_double),
        ("beta_seg", ctypes.c_double),
        ("delta_kr_order", ctypes.c_double),
        ("manning_coefficient", ctypes.c_double),
        ("friction_depth", ctypes.c_double),
        ("manning_max", ctypes.c_double),
        ("src2method", ctypes.c_int),
        ("alphamethod", ctypes.c_int),
        ("rarecorrectortest", ctypes.c_int),
        ("segregation", ctypes.c_int),
        ("friction_forcing", ctypes.c_int),
        ("dd_manning", ctypes.c_int),
        ("order", ctypes.c_int),
        ("transverse", ctypes.c_int),
        ("mthlim", ctypes.c_int * 5),
        ("mwaves", ctypes.c_int),
        ("entrainment_rate", ctypes.c_double),
        ("me", ctypes.c_double),
        ("chie", ctypes.c_double),
        ("entrainment", ctypes.c_int),
        ("entrainment_method", ctypes.c_int),
        ("curvature", ctypes.c_int),
    ]


def to_c(p: DigParams) -> CParams:
    p.validate()
    c = CParams()
    for name, ctype in CParams._fields_:
        if name == "mthlim":
            c.mthlim = (ctypes.c_int * 5)(*[int(x) for x in p.mthlim])
        elif ctype is ctypes.c_int:
            setattr(c, name, int(getattr(p, name)))
        else:
            setattr(c, name, float(getattr(p, name)))
    return c


def config0_params() -> DigParams:
    """Material parameters of the 200x200 inclined-plane workload: rho_s 2700,
    rho_f 1000, m_crit 0.64, phi 40 deg, k0 1e-10 m^2, alpha 0.01, mu 0.005 Pa s,
    dry_tol 1e-3.  The workload's "delta = 2 deg" has no angle counterpart in
    D-Claw (delta is a grain length, data.py:101-107), so delta keeps the
    reference default 0.01 m."""
    return DigParams(rho_s=2700.0, rho_f=1000.0, m_crit=0.64, phi_deg=40.0, kref=1.0e-10,
                     alpha_c=0.01, mu=0.005, dry_tolerance=1.0e-3, delta=0.01)
