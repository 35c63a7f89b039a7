"""Launch shape of the fused step (lbm_autotune / lbm_set_step_tuning): every
block width and addressing form gives bitwise the same populations, macros
and cell fields, and the autotuner leaves the lattice state as it found it.
The oracle run is the reference for all of them."""
import numpy as np
import pytest

from lbm_mirror_amd import abi
from helpers import duct_flags_3d, c2_ops

pytestmark = pytest.mark.gpu


def _ops_mixture(lat):
    ops = {1: abi.op_desc(boundary=abi.BC_BOUNCE_BACK),
           2: abi.op_desc(collision=abi.COLL_BGK_GUO, tau=0.65, vec=(2e-4, -1e-4, 5e-5),
                          flags=abi.OPF_GUO_CELL_FORCE),
           3: abi.op_desc(collision=abi.COLL_PSM, tau=0.9, flags=abi.OPF_PSM_FORCE_GATHER),
           4: abi.op_desc(collision=abi.COLL_PSM, tau=0.7,
                          flags=abi.OPF_PSM_STAMP_DISC | abi.OPF_PSM_CELL_FORCE,
                          disc=(5.0, 3.0, 2.0, 2.5, 1.5)),
           5: abi.op_desc(collision=abi.COLL_BGK, tau=0.8)}
    if lat == abi.D2Q9:
        ops[6] = abi.op_desc(boundary=abi.BC_ZOUHE_PRESSURE_X, collision=abi.COLL_BGK, east=1,
                             tau=0.8, rho=1.002)
    return ops


def _build(make, lat, st, ext, seed):
    rng = np.random.default_rng(seed)
    N = int(np.prod(ext))
    D, Q = abi.LATTICE_DQ[lat]
    ops = _ops_mixture(lat)
    flags = rng.choice(sorted(ops), size=N).astype(np.uint8)
    from oracle import oracle as orc
    _, _, _, w, _ = orc.lattice(lat)
    f0 = np.outer(w, np.ones(N)) * (1 + 0.05 * rng.uniform(-1, 1, size=(Q, N)))
    sims = []
    for mk in make:
        s = mk(lat, st, ext)
        s.set_flags(flags)
        s.set_periodic(1)
        for fl, op in ops.items():
            s.set_operator(fl, op)
        s.set_populations(f0)
        s.set_force_field(1e-4 * rng.uniform(-1, 1, size=(D, N)) if not sims else sims[0].get_force_field())
        s.set_porosity(rng.uniform(0, 1, size=N) if not sims else sims[0].get_porosity())
        sims.append(s)
    return sims


def _state(s):
    return (s.get_populations(), *s.get_macros(), s.get_force_field(), s.get_porosity())


@pytest.mark.parametrize("lat,st,ext", [(abi.D3Q19, abi.F32, (260, 5, 4)), (abi.D3Q27, abi.F32, (130, 4, 3)),
                                        (abi.D2Q9, abi.F64, (300, 7)), (abi.D2Q9, abi.F32, (70, 9))])
def test_every_shape_and_form_is_bitwise_the_oracle(hip, orc, lat, st, ext):
    from lbm_mirror_amd.lattice import Lattice
    widths = [w for w in (64, 128, 256) if w == 64 or w // 2 < ext[0]]
    cases = [(w, (f0, f1)) for w in widths for f0 in (1, 2) for f1 in (1, 2)]
    o = _build([lambda l, s, e: orc.Sim(l, s, e)], lat, st, ext, 7)[0]
    o.step(5)
    ref = _state(o)
    for w, form in cases:
        g = _build([lambda l, s, e: Lattice(l, s, e)], lat, st, ext, 7)[0]
        g.set_step_tuning(w, form)
        assert g.get_step_tuning() == {"block_x": w, "form": list(form)}
        g.step(5)
        for a, b in zip(_state(g), ref):
            assert np.array_equal(a, b), f"width {w} form {form} differs from the oracle"
        g.close()


@pytest.mark.parametrize("lat,st,ext", [(abi.D3Q19, abi.F32, (260, 5, 4)), (abi.D2Q9, abi.F64, (300, 7))])
def test_autotune_restores_state_and_keeps_results(hip, orc, lat, st, ext):
    from lbm_mirror_amd.lattice import Lattice
    g, o = _build([lambda l, s, e: Lattice(l, s, e), lambda l, s, e: orc.Sim(l, s, e)], lat, st, ext, 11)
    before = _state(g)
    chosen = g.autotune(rounds=2)
    assert chosen["block_x"] in (64, 128, 256) and all(f in (1, 2) for f in chosen["form"])
    assert all(u > 0 for u in chosen["us"])
    for a, b in zip(_state(g), before):
        assert np.array_equal(a, b), "autotune changed the lattice state"
    g.step(7)
    o.step(7)
    for a, b in zip(_state(g), _state(o)):
        assert np.array_equal(a, b)


def test_autotune_needs_even_phase_and_set_up(hip):
    from lbm_mirror_amd import LbmError
    from lbm_mirror_amd.lattice import Lattice
    g = Lattice(abi.D3Q19, abi.F32, (64, 6, 6))
    with pytest.raises(LbmError):
        g.autotune()  # no flags yet
    g.set_flags(duct_flags_3d(64, 6, 6))
    for fl, op in c2_ops(0.6, 1e-6).items():
        g.set_operator(fl, op)
    g.init_equilibrium()
    g.step(1)
    with pytest.raises(LbmError):
        g.autotune()  # odd layout
    g.step(1)
    assert g.autotune(rounds=1)["block_x"] == 64
    with pytest.raises(LbmError):
        g.set_step_tuning(96)


@pytest.mark.parametrize("lat,ext", [(abi.D3Q19, (260, 5, 4)), (abi.D2Q9, (300, 7)), (abi.D3Q27, (130, 4, 3)),
                                     (abi.D3Q19, (64, 5, 4))])
@pytest.mark.parametrize("st", [abi.F16S, abi.BF16S])
def test_16bit_forms_are_bitwise_equal(hip, lat, st, ext):
    """16-bit shifted storage: the scalar kernel in both addressing forms and
    the two-cells-per-thread kernel (form 3) give the same bits, periodic and
    not, over both AA phases and random operator mixtures."""
    from lbm_mirror_amd.lattice import Lattice
    for periodic in (0, 1):
        results = []
        for w, form in [(256, (1, 1)), (64, (2, 2)), (64, (3, 3)), (128, (3, 3)), (256, (3, 1)), (256, (1, 3))]:
            g = _build([lambda l, s, e: Lattice(l, s, e)], lat, st, ext, 13)[0]
            if not periodic:
                g.set_periodic(0)
            g.set_step_tuning(w, form)
            g.step(5)
            results.append(_state(g))
            g.close()
        for k, r in enumerate(results[1:], 1):
            for a, b in zip(r, results[0]):
                assert np.array_equal(a, b), f"case {k} periodic {periodic}"


def test_pair_form_needs_16bit_and_even_rows(hip):
    from lbm_mirror_amd import LbmError
    from lbm_mirror_amd.lattice import Lattice
    with pytest.raises(LbmError):
        Lattice(abi.D2Q9, abi.F32, (64, 8)).set_step_tuning(64, (3, 3))
    with pytest.raises(LbmError):
        Lattice(abi.D2Q9, abi.F16S, (63, 8)).set_step_tuning(64, (3, 3))
    Lattice(abi.D2Q9, abi.F16S, (64, 8)).set_step_tuning(64, (3, 3))
