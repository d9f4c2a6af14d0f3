"""The reference's physical constants (param.f90) as FP64 values, folded once on the host the
way its compiler folds them (-autodouble -fpconstant, Makefile:7: every literal is FP64).

    pi        = acos(-1.)                                   param.f90:48
    thetainc  = (thetaincd pi) / 180, thetaincd = 70        param.f90:52
    Rep = .0016, ta = 1160000, Ca = .1, lambda = 1          param.f90:57
    visc0     = 1 / Rep                                     param.f90:59
    col       = 5.68533 (10.**-2.), betacol = 15.36,
    rad0      = 2.5 dz                                      param.f90:64
    gscell    = 1 / We, We = Rep Ca; cbend = .3             param.f90:57,88,91

The derived factors are formed as the reference writes them at their use:
    collision factor  2.*(col/(Rep**2.))                    collisions.f90:24
    centrifugal       sin(thetainc), -cos(thetainc), (ta/(Rep**2.))**.25   indicator.f90:164-165
"""
from __future__ import annotations

import math

PI = math.acos(-1.0)
THETAINC = (70.0 * PI) / 180.0
REP = 0.0016
TA = 1160000.0
CA = 0.1
LAMBDA = 1.0
VISC0 = 1.0 / REP
COL = 5.68533 * (10.0 ** -2.0)
BETACOL = 15.36
GSCELL = 1.0 / (REP * CA)
EBEND = 0.3 * GSCELL


def collision_factor(rep: float = REP, col: float = COL) -> float:
    """2.*(col/(Rep**2.)) (collisions.f90:24)."""
    return 2.0 * (col / (rep * rep))


def rad0(dz: float) -> float:
    """The wall force's cutoff and zero-force distance, 2.5 dz (param.f90:64)."""
    return 2.5 * dz


def centrifugal(thetainc: float = THETAINC, ta: float = TA, rep: float = REP):
    """(sin(thetainc), cos(thetainc), (ta/(Rep**2.))**.25) of indicator.f90:164-165."""
    return math.sin(thetainc), math.cos(thetainc), (ta / (rep * rep)) ** 0.25
