"""Component objects named after the reference's `cmpt::` kinds.

Each mirrors the constructor parameters of the reference component it names
(modules/core/c++/collision.hpp, boundary.hpp, psm.hpp, component/psm.hpp,
momentum_gather.hpp, particle/particle_opa.hpp).  ``combine`` turns a
(boundary, collision) pair into the lbm_op_desc the C ABI takes for one flag.
"""
from dataclasses import dataclass, field
from typing import Optional, Sequence, Tuple

from . import abi


# ---- collisions ------------------------------------------------------------
@dataclass
class BGK:
    """cmpt::BGK (collision.hpp:19-97); component{relaxation}."""
    tau: float


@dataclass
class BGKGuo:
    """cmpt::BGKGuo (collision.hpp:99-205); component{relaxation, global_force}.
    ``cell_force`` also reads the per-cell "force" macro field."""
    tau: float
    global_force: Sequence[float] = (0.0, 0.0, 0.0)
    cell_force: bool = False


@dataclass
class Psm:
    """cmpt::Psm (psm.hpp:37-90).  Optional fused companions of the static
    PSM driver: ``stamp_disc`` = cmpt::OneParticleAt(pos, rad, eps) before the
    collision, ``force_gather`` = cmpt::ForceGather after it; ``cell_force``
    selects the cmpt::PSM variant (component/psm.hpp:16-80)."""
    tau: float
    stamp_disc: Optional[Tuple[Sequence[float], float, float]] = None
    force_gather: bool = False
    cell_force: bool = False


@dataclass
class Hlbm:
    """cmpt::Hlbm (hlbm.hpp:62-105): BGK towards eq(rho, u*B + N(1-B)/D).  The
    particle side (HlbmReset, HlbmOneParticleGather/Stamp) is selected with
    ``abi.particle_params(coupling=abi.COUPLING_HLBM)``."""
    tau: float


@dataclass
class FpLbm:
    """cmpt::FpLbm (fplbm.hpp:60-133): Guo forcing with the cell force field.
    The particle side (FpLbmReset, FpLbmOneParticle) is selected with
    ``abi.particle_params(coupling=abi.COUPLING_FPLBM)``."""
    tau: float


# ---- boundaries -------------------------------------------------------------
@dataclass
class BounceBack:
    """cmpt::BounceBack (boundary.hpp:32-76)."""


@dataclass
class AntiBounceBack:
    """cmpt::AntiBounceBack<0> (boundary.hpp:81-133), D2Q9."""


@dataclass
class ZouHePressureX:
    """cmpt::ZouHePressureX<east>{rho} (boundary.hpp:175-756).  ``ramp`` =
    (rho_start, rho_slope, rho_final, ramp_steps) reproduces the drivers'
    time-ramped inlet density (poiseulle_2d_gpu.cpp:182-190)."""
    east: bool
    rho: float = 1.0
    ramp: Optional[Tuple[float, float, float, int]] = None


@dataclass
class ZouHeVelocityX:
    """cmpt::ZouHeVelocityX<east> (boundary.hpp:759-805), D2Q9; the per-cell
    velocity comes from Lattice.set_boundary_velocity."""
    east: bool


@dataclass
class Equilibrium:
    """cmpt::Equilibrium{density, momentum} (boundary.hpp:135-161)."""
    rho: float
    u: Sequence[float] = (0.0, 0.0, 0.0)


def combine(boundary=None, collision=None):
    kw = {}
    flags = 0
    if isinstance(boundary, BounceBack):
        kw["boundary"] = abi.BC_BOUNCE_BACK
    elif isinstance(boundary, AntiBounceBack):
        kw["boundary"] = abi.BC_ANTI_BOUNCE_BACK
    elif isinstance(boundary, ZouHePressureX):
        kw.update(boundary=abi.BC_ZOUHE_PRESSURE_X, east=int(boundary.east),
                  rho=boundary.rho, ramp=boundary.ramp)
    elif isinstance(boundary, ZouHeVelocityX):
        kw.update(boundary=abi.BC_ZOUHE_VELOCITY_X, east=int(boundary.east))
    elif isinstance(boundary, Equilibrium):
        kw.update(boundary=abi.BC_EQUILIBRIUM, rho_eq=boundary.rho, vec=tuple(boundary.u))
    elif boundary is not None:
        raise TypeError(f"unknown boundary component {boundary!r}")
    if isinstance(collision, BGK):
        kw.update(collision=abi.COLL_BGK, tau=collision.tau)
    elif isinstance(collision, BGKGuo):
        kw.update(collision=abi.COLL_BGK_GUO, tau=collision.tau, vec=tuple(collision.global_force))
        if collision.cell_force:
            flags |= abi.OPF_GUO_CELL_FORCE
    elif isinstance(collision, Hlbm):
        kw.update(collision=abi.COLL_HLBM, tau=collision.tau)
    elif isinstance(collision, FpLbm):
        kw.update(collision=abi.COLL_FPLBM, tau=collision.tau)
    elif isinstance(collision, Psm):
        kw.update(collision=abi.COLL_PSM, tau=collision.tau)
        if collision.stamp_disc is not None:
            pos, rad, eps = collision.stamp_disc
            p = list(pos) + [0.0] * (3 - len(pos))
            kw["disc"] = (p[0], p[1], p[2], rad, eps)
            flags |= abi.OPF_PSM_STAMP_DISC
        if collision.force_gather:
            flags |= abi.OPF_PSM_FORCE_GATHER
        if collision.cell_force:
            flags |= abi.OPF_PSM_CELL_FORCE
    elif collision is not None:
        raise TypeError(f"unknown collision component {collision!r}")
    return abi.op_desc(flags=flags, **kw)
