"""The rest of the C-ABI boundary: the device fault words that lbm_sync and
lbm_step report, the FP64 residual of one lattice (lbm_residual), and the
populations in the reference's own chunk layout with its ghost ring
(lbm_get/set_populations_layout), checked through the C ABI against numpy
recomputations, the oracle and the Forst writer's chunk bytes."""
import numpy as np
import pytest

from lbm_mirror_amd import abi, LbmError, output
from helpers import (channel_flags_2d, duct_flags_3d, parabolic_inlet, c1_ops, c2_ops, make_particle,
                     seeded_discs)

pytestmark = pytest.mark.gpu


def _lattice(*a):
    from lbm_mirror_amd.lattice import Lattice
    return Lattice(*a)


def _channel(make, st, ext=(64, 18)):
    s = make(abi.D2Q9, st, ext)
    s.set_flags(channel_flags_2d(*ext))
    for fl, op in c1_ops(0.8, 0.05).items():
        s.set_operator(fl, op)
    s.set_boundary_velocity(parabolic_inlet(*ext, 0.05))
    s.init_equilibrium(1.0, (0.0, 0.0))
    return s


# --------------------------------------------------------------- fault words
@pytest.mark.parametrize("st", [abi.F64, abi.F32, abi.BF16S])
def test_nonfinite_state_is_reported_by_sync_and_next_step(hip, st):
    g = _channel(_lattice, st)
    g.step(5)
    g.sync()  # a finite run reports nothing
    f = g.get_populations()
    f[2, 300] = np.nan
    g.set_populations(f)
    g.step(3)
    with pytest.raises(LbmError) as e:
        g.sync()
    assert e.value.status == abi.ERR_INVALID_STATE
    g.sync()  # reported once, then cleared
    # lbm_step reports a fault of an earlier call that has run, before queuing
    g.step(2)
    g.last_step_ms()  # waits for the call's events, reports nothing
    with pytest.raises(LbmError) as e:
        g.step(1)
    assert e.value.status == abi.ERR_INVALID_STATE


def test_finite_runs_report_ok_on_every_lattice(hip):
    for lat, ext in ((abi.D3Q19, (40, 8, 6)), (abi.D3Q27, (24, 6, 6))):
        g = _lattice(lat, abi.F32, ext)
        g.set_flags(duct_flags_3d(*ext))
        g.set_periodic(1)
        for fl, op in c2_ops(0.6, 1e-5).items():
            g.set_operator(fl, op)
        g.init_equilibrium()
        for _ in range(3):
            g.step(7)
            g.sync()


def test_cell_list_overflow_is_reported(hip, orc):
    ext = (64, 32)
    g = _channel(_lattice, abi.F64, ext)
    g.set_operator(2, abi.op_desc(collision=abi.COLL_PSM, tau=0.7))
    # three discs in one bin of capacity 1: two are not listed
    parts = [make_particle([x, 15.0], 3.0, 1.5, 2, abi.F64, orc=orc) for x in (31.0, 33.5, 35.0)]
    prm = abi.particle_params(k_n=1e3, zeta_n=0.3, porosity_mode=abi.POROSITY_PRODUCT, cell_capacity=1,
                              sub_steps=4, max_radius=3.0)
    g.particles_set(parts, prm)
    g.psm_couple()
    assert g.overflow() == 2
    with pytest.raises(LbmError) as e:
        g.sync()
    assert e.value.status == abi.ERR_OVERFLOW
    g.sync()


# ------------------------------------------------------------------ residual
def _residual_np(u, up):
    num = float(np.sum((u - up) ** 2))
    den = float(np.sum(u ** 2))
    return (np.sqrt(num / den) if den > 0 else 0.0), num, den


@pytest.mark.parametrize("st", [abi.F64, abi.F32])
@pytest.mark.parametrize("n", [1, 2, 3, 70])
def test_residual_of_one_lattice(hip, orc, st, n):
    """One call of n steps (70: step graphs replay inside it) against the
    oracle run split before its last step; FP64 sums to 1e-12 relative."""
    g = _channel(_lattice, st)
    o = _channel(lambda *a: orc.Sim(*a), st)
    g.set_residual(True)
    g.step(40)
    o.step(40)
    g.step(n)
    o.step(n - 1)
    _, ua = o.get_macros()
    o.step(1)
    _, ub = o.get_macros()
    _, ug = g.get_macros()
    assert np.array_equal(ug, ub)
    r, num, den = g.residual()
    rn, numn, denn = _residual_np(ub, ua)
    assert abs(num - numn) <= 1e-12 * numn and abs(den - denn) <= 1e-12 * denn
    assert abs(r - rn) <= 1e-12 * rn
    assert r > 0
    assert g.residual() == (r, num, den)  # fixed order: equal bits on every call


def test_residual_with_macros_every_step(hip, orc):
    """FPLBM particles keep the macros every step: u_{t-1} is then copied,
    and the run itself is unchanged by tracking."""
    ext = (96, 48)

    def run(track):
        g = _channel(_lattice, abi.F64, ext)
        g.set_operator(2, abi.op_desc(collision=abi.COLL_FPLBM, tau=0.7))
        discs = seeded_discs(4, ext, 4.0, 6.0, 5)
        parts = [make_particle([x, y], r, 1.5, 2, abi.F64, orc=orc) for (x, y, r) in discs]
        g.particles_set(parts, abi.particle_params(k_n=1e3, zeta_n=0.3, walls=True, sub_steps=8,
                                                   max_radius=6.0, coupling=abi.COUPLING_FPLBM))
        if track:
            g.set_residual(True)
        return g

    a = run(False)
    a.step(12)
    _, ua = a.get_macros()
    a.step(1)
    _, ub = a.get_macros()
    b = run(True)
    b.step(13)
    assert np.array_equal(b.get_macros()[1], ub)
    r, num, den = b.residual()
    rn, numn, denn = _residual_np(ub, ua)
    assert abs(num - numn) <= 1e-12 * numn and abs(den - denn) <= 1e-12 * denn
    assert abs(r - rn) <= 1e-12 * rn


def test_residual_off_and_nonfinite(hip):
    g = _channel(_lattice, abi.F64)
    with pytest.raises(LbmError) as e:
        g.residual()
    assert e.value.status == abi.ERR_INVALID_STATE
    g.set_residual(True)
    f = g.get_populations()
    f[1, 100] = np.inf
    g.set_populations(f)
    g.step(4)
    assert np.isnan(g.residual()[0])
    with pytest.raises(LbmError):
        g.sync()
    g.set_residual(False)
    with pytest.raises(LbmError):
        g.residual()


# ------------------------------------------------- the reference's chunk layout
def _aos_from_soa(f, ext, Q):
    D = len(ext)
    gshape = tuple(e + 2 for e in reversed(ext)) + (Q,)  # (z, y, x, Q) ghost-widened
    out = np.zeros(gshape)
    inner = tuple(slice(1, -1) for _ in ext)
    out[inner] = f.reshape((Q,) + tuple(reversed(ext))).transpose(tuple(range(1, D + 1)) + (0,))
    return out.reshape(-1)


@pytest.mark.parametrize("lat,ext", [(abi.D2Q9, (13, 7)), (abi.D3Q19, (9, 6, 5)), (abi.D3Q27, (7, 5, 4))])
@pytest.mark.parametrize("st", [abi.F64, abi.F32])
def test_reference_chunk_layout(hip, orc, tmp_path, lat, ext, st):
    D, Q = abi.LATTICE_DQ[lat]
    N = int(np.prod(ext))
    rng = np.random.default_rng(31 + lat)
    _, _, _, w, _ = orc.lattice(lat)
    g = _lattice(lat, st, ext)
    g.set_flags(np.full(N, 2, np.uint8))
    g.set_operator(2, abi.op_desc(collision=abi.COLL_BGK, tau=0.8))
    f0 = np.outer(w, np.ones(N)) * (1 + 0.05 * rng.uniform(-1, 1, size=(Q, N)))
    g.set_populations(f0)
    assert g.population_count(abi.LAYOUT_REF_AOS_GHOST) == Q * int(np.prod([e + 2 for e in ext]))
    for steps in (0, 1, 2):  # both AA phases
        if steps:
            g.step(1)
        soa = g.get_populations()
        aos = g.get_populations(abi.LAYOUT_REF_AOS_GHOST)
        assert np.array_equal(aos, _aos_from_soa(soa, ext, Q))
        # the same bytes as the reference-format dfs chunk of the Forst writer
        p = str(tmp_path / f"dfs_{steps}.forst")
        output.write_forst(g, p, "dfs")
        T = np.float64 if st == abi.F64 else np.float32
        assert np.array_equal(np.fromfile(p, dtype="<f8" if T is np.float64 else "<f4"), aos.astype(T))
    # set in the chunk layout (ghost ring ignored) == set in SoA
    h = _lattice(lat, st, ext)
    h.set_flags(np.full(N, 2, np.uint8))
    h.set_operator(2, abi.op_desc(collision=abi.COLL_BGK, tau=0.8))
    a = _aos_from_soa(f0, ext, Q).reshape(tuple(e + 2 for e in reversed(ext)) + (Q,))
    ghost = np.ones(a.shape[:-1], bool)
    ghost[tuple(slice(1, -1) for _ in ext)] = False
    a[ghost] = 7.0
    h.set_populations(a.reshape(-1), abi.LAYOUT_REF_AOS_GHOST)
    k = _lattice(lat, st, ext)
    k.set_flags(np.full(N, 2, np.uint8))
    k.set_operator(2, abi.op_desc(collision=abi.COLL_BGK, tau=0.8))
    k.set_populations(f0)
    assert np.array_equal(h.get_populations(), k.get_populations())
    h.step(3)
    k.step(3)
    assert np.array_equal(h.get_populations(abi.LAYOUT_REF_AOS_GHOST),
                          k.get_populations(abi.LAYOUT_REF_AOS_GHOST))
    with pytest.raises(LbmError):
        g.get_populations(5)
