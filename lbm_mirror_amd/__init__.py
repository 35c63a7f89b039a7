"""lbm_mirror_amd — MI355X-native lattice Boltzmann engine (gfx950 HIP).

Host-side mirror of the reference's component API (keldu/lbm_mirror,
modules/core/c++/{collision,boundary,psm,stream}.hpp) over the C ABI declared
in include/lbm_mirror_amd/lbm.h.  The compute path is the HIP library
``_build/liblbm_mirror_amd.so``; there is no CPU fallback: if the library is
missing or no GPU is visible, constructing a lattice raises.
"""
import ctypes as C
import os

from . import abi

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "liblbm_mirror_amd.so")

__all__ = ["abi", "load_library", "LIB_PATH", "LbmError"]


class LbmError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"[{abi.STATUS_NAMES.get(status, status)}] {msg}")
        self.status = status


_lib = None

# every symbol include/lbm_mirror_amd/lbm.h declares
EXPORTS = [
    "lbm_last_error", "lbm_version", "lbm_device_count", "lbm_ctx_create", "lbm_ctx_destroy",
    "lbm_set_flags", "lbm_set_periodic", "lbm_set_operator", "lbm_set_boundary_velocity",
    "lbm_set_force_field", "lbm_get_force_field", "lbm_set_porosity", "lbm_get_porosity",
    "lbm_get_hlbm_sums",
    "lbm_set_populations", "lbm_get_populations", "lbm_init_equilibrium", "lbm_get_macros",
    "lbm_set_macros", "lbm_step", "lbm_sync", "lbm_last_step_ms", "lbm_particles_set",
    "lbm_particles_get", "lbm_spheroid_mass", "lbm_psm_couple", "lbm_particle_reduce",
    "lbm_dem_contact", "lbm_particle_integrate", "lbm_cell_list_overflow",
    "lbm_batch_create", "lbm_batch_destroy", "lbm_batch_member", "lbm_batch_step",
    "lbm_batch_run", "lbm_batch_autotune", "lbm_batch_set_tuning", "lbm_batch_stats",
    "lbm_init_equilibrium_field", "lbm_net_create", "lbm_net_destroy", "lbm_net_param_count",
    "lbm_net_get_params", "lbm_net_set_params", "lbm_net_features", "lbm_net_add_sample",
    "lbm_net_clear_samples", "lbm_net_sample_count", "lbm_net_train", "lbm_net_predict", "lbm_net_correct_flux",
    "lbm_net_forward_raw", "lbm_net_loss_grad_raw", "lbm_net_last_ms",
    "lbm_net_get_output_scale", "lbm_net_set_output_scale",
    "lbm_get_flags", "lbm_ctx_info", "lbm_write_vtk", "lbm_write_csv", "lbm_write_forst",
    "lbm_read_forst", "lbm_write_vtk_fields", "lbm_write_csv_fields", "lbm_write_forst_fields",
    "lbm_read_forst_fields", "lbm_set_step_graph", "lbm_hbm_ceiling",
    "lbm_autotune", "lbm_set_step_tuning", "lbm_get_step_tuning",
    "lbm_population_count", "lbm_set_populations_layout", "lbm_get_populations_layout",
    "lbm_set_residual", "lbm_residual", "lbm_residual_noalloc",
]


def load_library(path=None):
    """Load liblbm_mirror_amd.so and declare its prototypes.  Raises if the
    library has not been built (no silent fallback)."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise RuntimeError(f"HIP library not built: {p} (run __graft_entry__.build())")
    L = C.CDLL(p)
    dp = C.POINTER(C.c_double)
    vp = C.c_void_p
    st = C.c_int
    L.lbm_last_error.restype = C.c_char_p
    L.lbm_version.restype = C.c_char_p
    L.lbm_device_count.restype = C.c_int
    L.lbm_ctx_create.argtypes = [C.POINTER(vp), C.c_int32, C.c_int32, C.POINTER(C.c_uint64), C.c_int32]
    L.lbm_ctx_destroy.argtypes = [vp]
    L.lbm_set_flags.argtypes = [vp, C.POINTER(C.c_uint8)]
    L.lbm_set_periodic.argtypes = [vp, C.c_uint32]
    L.lbm_set_operator.argtypes = [vp, C.c_uint8, C.POINTER(abi.OpDesc)]
    for name in ("lbm_set_boundary_velocity", "lbm_set_force_field", "lbm_set_porosity",
                 "lbm_set_populations"):
        getattr(L, name).argtypes = [vp, dp]
    for name in ("lbm_get_force_field", "lbm_get_porosity", "lbm_get_populations"):
        getattr(L, name).argtypes = [vp, dp]
    L.lbm_get_hlbm_sums.argtypes = [vp, dp, dp]
    L.lbm_init_equilibrium.argtypes = [vp, C.c_double, dp]
    L.lbm_get_macros.argtypes = [vp, dp, dp]
    L.lbm_set_macros.argtypes = [vp, dp, dp]
    L.lbm_step.argtypes = [vp, C.c_uint64, C.c_uint64]
    L.lbm_sync.argtypes = [vp]
    L.lbm_last_step_ms.argtypes = [vp, dp]
    L.lbm_particles_set.argtypes = [vp, C.POINTER(abi.Particle), C.c_uint32,
                                    C.POINTER(abi.ParticleParams)]
    L.lbm_particles_get.argtypes = [vp, C.POINTER(abi.Particle), C.c_uint32]
    L.lbm_spheroid_mass.argtypes = [C.c_int32, C.c_int32, C.c_double, C.c_double, dp, dp]
    L.lbm_psm_couple.argtypes = [vp, C.c_uint64]
    L.lbm_particle_reduce.argtypes = [vp, dp, dp]
    L.lbm_dem_contact.argtypes = [vp]
    L.lbm_particle_integrate.argtypes = [vp]
    L.lbm_cell_list_overflow.argtypes = [vp, C.POINTER(C.c_uint64)]
    u64p = C.POINTER(C.c_uint64)
    L.lbm_batch_create.argtypes = [C.POINTER(vp), C.c_int32, C.c_int32, C.c_uint32, u64p, C.c_int32]
    L.lbm_batch_destroy.argtypes = [vp]
    L.lbm_batch_member.argtypes = [vp, C.c_uint32]
    L.lbm_batch_member.restype = vp
    L.lbm_batch_step.argtypes = [vp, C.c_uint64, C.c_uint64]
    L.lbm_batch_run.argtypes = [vp, C.c_uint64, C.c_double, C.c_uint64, C.c_uint64, u64p, dp]
    L.lbm_batch_autotune.argtypes = [vp, C.POINTER(abi.BatchTuning)]
    L.lbm_batch_set_tuning.argtypes = [vp, C.POINTER(abi.BatchTuning)]
    L.lbm_batch_stats.argtypes = [vp, dp, u64p]
    L.lbm_init_equilibrium_field.argtypes = [vp, dp, dp]
    fp = C.POINTER(C.c_float)
    u8p = C.POINTER(C.c_uint8)
    L.lbm_net_create.argtypes = [C.POINTER(vp), C.POINTER(abi.NetConfig), C.c_int32]
    L.lbm_net_destroy.argtypes = [vp]
    L.lbm_net_param_count.argtypes = [vp, u64p]
    L.lbm_net_get_params.argtypes = [vp, fp]
    L.lbm_net_set_params.argtypes = [vp, fp]
    L.lbm_net_features.argtypes = [vp, u8p, C.c_double, C.c_double, fp]
    L.lbm_net_add_sample.argtypes = [vp, u8p, C.c_double, C.c_double, dp, dp]
    L.lbm_net_clear_samples.argtypes = [vp]
    L.lbm_net_sample_count.argtypes = [vp, C.POINTER(C.c_uint32)]
    L.lbm_net_train.argtypes = [vp, C.c_uint32, C.c_float, C.c_uint64, fp]
    L.lbm_net_predict.argtypes = [vp, C.c_uint32, u8p, dp, dp, dp, dp]
    L.lbm_net_correct_flux.argtypes = [vp, C.c_uint32, u8p, dp, dp]
    L.lbm_net_forward_raw.argtypes = [vp, C.c_uint32, fp, fp]
    L.lbm_net_loss_grad_raw.argtypes = [vp, C.c_uint32, fp, fp, fp, dp, fp]
    L.lbm_net_last_ms.argtypes = [vp, dp]
    L.lbm_net_get_output_scale.argtypes = [vp, fp, fp]
    L.lbm_net_set_output_scale.argtypes = [vp, fp, fp]
    L.lbm_get_flags.argtypes = [vp, u8p]
    L.lbm_set_step_graph.argtypes = [vp, C.c_uint32]
    L.lbm_hbm_ceiling.argtypes = [C.c_int32, C.c_int32, C.c_uint64, C.c_uint32, dp]
    L.lbm_autotune.argtypes = [vp, C.c_uint32, C.POINTER(abi.StepTuning)]
    L.lbm_set_step_tuning.argtypes = [vp, C.POINTER(abi.StepTuning)]
    L.lbm_get_step_tuning.argtypes = [vp, C.POINTER(abi.StepTuning)]
    L.lbm_population_count.argtypes = [vp, C.c_int32, C.POINTER(C.c_uint64)]
    L.lbm_set_populations_layout.argtypes = [vp, dp, C.c_int32]
    L.lbm_get_populations_layout.argtypes = [vp, dp, C.c_int32]
    L.lbm_set_residual.argtypes = [vp, C.c_int32]
    L.lbm_residual.argtypes = [vp, dp]
    L.lbm_residual_noalloc.argtypes = [vp, vp]
    L.lbm_ctx_info.argtypes = [vp, C.POINTER(C.c_int32), C.POINTER(C.c_int32), u64p]
    cs = C.c_char_p
    L.lbm_write_vtk.argtypes = [vp, cs, cs, C.c_uint64, cs]
    L.lbm_write_csv.argtypes = [vp, cs, cs, C.c_uint64, cs]
    L.lbm_write_forst.argtypes = [vp, cs, cs]
    L.lbm_read_forst.argtypes = [vp, cs, cs]
    fsp = C.POINTER(abi.FieldSet)
    L.lbm_write_vtk_fields.argtypes = [fsp, cs, cs, C.c_uint64, cs]
    L.lbm_write_csv_fields.argtypes = [fsp, cs, cs, C.c_uint64, cs]
    L.lbm_write_forst_fields.argtypes = [fsp, cs, cs]
    L.lbm_read_forst_fields.argtypes = [fsp, cs, cs]
    for name in EXPORTS:
        f = getattr(L, name)
        if name not in ("lbm_last_error", "lbm_version", "lbm_device_count", "lbm_batch_member"):
            f.restype = st
    if path is None:
        _lib = L
    return L


def check(status):
    if status != abi.OK:
        msg = load_library().lbm_last_error()
        raise LbmError(status, msg.decode() if msg else "")


def __getattr__(name):
    # Lattice and friends import the library lazily so that ``abi`` stays
    # importable on machines without the built library.
    if name in ("Lattice", "components"):
        from . import lattice as _lat
        if name == "Lattice":
            return _lat.Lattice
        from . import components
        return components
    raise AttributeError(name)
