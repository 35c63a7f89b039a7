"""ctypes mirror of include/lbm_mirror_amd/lbm.h (the C ABI).

Plain data definitions only: the structures and enums a caller needs to talk
to liblbm_mirror_amd.so.  Loading the library lives in ``lbm_mirror_amd``.
"""
import ctypes as C

# lbm_status
OK = 0
ERR_CRITICAL = 1
ERR_NOT_FOUND = 2
ERR_OUT_OF_MEMORY = 3
ERR_INVALID_STATE = 4
ERR_NOT_SUPPORTED = 5
ERR_INVALID_ARG = 6
ERR_OVERFLOW = 7

STATUS_NAMES = {
    OK: "ok", ERR_CRITICAL: "critical", ERR_NOT_FOUND: "not_found",
    ERR_OUT_OF_MEMORY: "out_of_memory", ERR_INVALID_STATE: "invalid_state",
    ERR_NOT_SUPPORTED: "not_supported", ERR_INVALID_ARG: "invalid_argument",
    ERR_OVERFLOW: "overflow",
}

# lbm_lattice
D2Q9, D3Q19, D3Q27 = 0, 1, 2
LATTICE_DQ = {D2Q9: (2, 9), D3Q19: (3, 19), D3Q27: (3, 27)}

# lbm_storage
F64, F32, F16S, BF16S = 0, 1, 2, 3

# lbm_boundary_kind
BC_NONE, BC_BOUNCE_BACK, BC_ANTI_BOUNCE_BACK = 0, 1, 2
BC_ZOUHE_PRESSURE_X, BC_ZOUHE_VELOCITY_X, BC_EQUILIBRIUM = 3, 4, 5

# lbm_collision_kind
COLL_NONE, COLL_BGK, COLL_BGK_GUO, COLL_PSM, COLL_HLBM, COLL_FPLBM = 0, 1, 2, 3, 4, 5
COUPLING_PSM, COUPLING_HLBM, COUPLING_FPLBM = 0, 1, 2

OPF_GUO_CELL_FORCE = 0x1
OPF_PSM_STAMP_DISC = 0x2
OPF_PSM_FORCE_GATHER = 0x4
OPF_PSM_CELL_FORCE = 0x8
OPF_ZOUHE_NORMAL_ONLY = 0x10

POROSITY_SCATTER, POROSITY_PRODUCT = 0, 1

class StepTuning(C.Structure):
    _fields_ = [("block_x", C.c_int32), ("form", C.c_int32 * 2), ("us", C.c_double * 2)]


# lbm_layout
LAYOUT_SOA, LAYOUT_REF_AOS_GHOST = 0, 1

# lbm_ceiling_kind
CEILING_COPY, CEILING_PLANES = 0, 1


class OpDesc(C.Structure):
    _fields_ = [
        ("boundary", C.c_int32),
        ("collision", C.c_int32),
        ("east", C.c_int32),
        ("flags", C.c_uint32),
        ("tau", C.c_double),
        ("rho_start", C.c_double),
        ("rho_slope", C.c_double),
        ("rho_final", C.c_double),
        ("ramp_steps", C.c_uint64),
        ("vec", C.c_double * 3),
        ("rho_eq", C.c_double),
        ("disc", C.c_double * 5),
    ]


class Particle(C.Structure):
    _fields_ = [
        ("position", C.c_double * 3),
        ("position_old", C.c_double * 3),
        ("rotation", C.c_double * 3),
        ("rotation_old", C.c_double * 3),
        ("acceleration", C.c_double * 3),
        ("angular_acceleration", C.c_double * 3),
        ("force", C.c_double * 3),
        ("torque", C.c_double * 3),
        ("mass", C.c_double),
        ("inertia", C.c_double),
        ("density", C.c_double),
        ("radius", C.c_double),
    ]


class ParticleParams(C.Structure):
    _fields_ = [
        ("k_n", C.c_double),
        ("zeta_n", C.c_double),
        ("gamma_t", C.c_double),
        ("mu", C.c_double),
        ("walls", C.c_int32),
        ("porosity_mode", C.c_int32),
        ("psm_torque", C.c_int32),
        ("cell_capacity", C.c_int32),
        ("sub_steps", C.c_uint64),
        ("body_acceleration", C.c_double * 3),
        ("eps", C.c_double),
        ("max_radius", C.c_double),
        ("state_fp64", C.c_int32),
        ("coupling", C.c_int32),
        ("fixed_axis", C.c_int32),
        ("spin_axis", C.c_double * 3),
    ]


def op_desc(boundary=BC_NONE, collision=COLL_NONE, east=0, flags=0, tau=1.0,
            rho=None, ramp=None, vec=(0.0, 0.0, 0.0), rho_eq=1.0, disc=None):
    """Build an OpDesc.  ``rho`` is a constant Zou/He set-point; ``ramp`` is
    (rho_start, rho_slope, rho_final, ramp_steps) as the drivers compute it."""
    d = OpDesc()
    d.boundary, d.collision, d.east, d.flags, d.tau = boundary, collision, int(east), flags, tau
    if ramp is not None:
        d.rho_start, d.rho_slope, d.rho_final, d.ramp_steps = ramp
    else:
        r = 1.0 if rho is None else rho
        d.rho_start, d.rho_slope, d.rho_final, d.ramp_steps = r, 0.0, r, 0
    for i in range(3):
        d.vec[i] = vec[i] if i < len(vec) else 0.0
    d.rho_eq = rho_eq
    if disc is not None:
        for i in range(5):
            d.disc[i] = disc[i]
    return d


def particle_params(k_n=1.0, zeta_n=0.3, gamma_t=0.0, mu=0.0, walls=False,
                    porosity_mode=POROSITY_SCATTER, psm_torque=False,
                    cell_capacity=4, sub_steps=1, body_acceleration=(0.0, 0.0, 0.0),
                    eps=1.5, max_radius=1.0, state_fp64=False, coupling=COUPLING_PSM,
                    spin_axis=None):
    """spin_axis: None for ParticleIntegrate (free body), else the axis of
    ParticleAxisIntegrate (pinned body spinning about it; 2D uses its sign)."""
    p = ParticleParams()
    p.k_n, p.zeta_n, p.gamma_t, p.mu = k_n, zeta_n, gamma_t, mu
    p.walls, p.porosity_mode, p.psm_torque = int(walls), porosity_mode, int(psm_torque)
    p.cell_capacity, p.sub_steps = cell_capacity, sub_steps
    for i in range(3):
        p.body_acceleration[i] = body_acceleration[i] if i < len(body_acceleration) else 0.0
    p.eps, p.max_radius = eps, max_radius
    p.state_fp64 = int(state_fp64)
    p.coupling = coupling
    if spin_axis is not None:
        p.fixed_axis = 1
        for i in range(3):
            p.spin_axis[i] = spin_axis[i] if i < len(spin_axis) else 0.0
    return p


class FieldSet(C.Structure):
    """lbm_field_set: host copies of a lattice's fields (SoA doubles)."""
    _fields_ = [
        ("lattice", C.c_int32),
        ("storage", C.c_int32),
        ("extent", C.c_uint64 * 3),
        ("velocity", C.POINTER(C.c_double)),
        ("density", C.POINTER(C.c_double)),
        ("force", C.POINTER(C.c_double)),
        ("porosity", C.POINTER(C.c_double)),
        ("info", C.POINTER(C.c_uint8)),
        ("dfs", C.POINTER(C.c_double)),
    ]


class BatchTuning(C.Structure):
    _fields_ = [
        ("block", C.c_int32),
        ("rows", C.c_int32),
        ("grid", C.c_int32),
        ("ms_per_step", C.c_double),
    ]


class NetConfig(C.Structure):
    _fields_ = [
        ("height", C.c_int32),
        ("width", C.c_int32),
        ("widths", C.c_int32 * 4),
        ("batch", C.c_int32),
        ("capacity", C.c_int32),
        ("seed", C.c_uint64),
    ]


def net_config(height, width, widths=(32, 64, 96, 160), batch=16, capacity=0, seed=1):
    c = NetConfig()
    c.height, c.width = height, width
    for i in range(4):
        c.widths[i] = widths[i]
    c.batch, c.capacity, c.seed = batch, capacity, seed
    return c
