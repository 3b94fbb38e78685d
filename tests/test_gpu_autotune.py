"""Launch-shape autotuner: it leaves the simulation state untouched, picks valid shapes, and
the run that continues with them matches one that never tuned.

The coupled case is the carbonate tableau of the reference's example 04 with dual-Monod
kinetics and biomass, in both speciation modes and both population precisions. The tuner's
trials all start from the state it found, and it restores that state (populations, biomass,
mask, deferred increments, statistics) when it ends, so the downloads before and after must
be bit-identical. Block shapes do not change any voxel's arithmetic; the fused evaluation
budget changes only how many Newton steps run inside the transport kernel before the
persistent kernel finishes a voxel, so later steps agree to the solver tolerance."""
import numpy as np
import pytest

from conftest import d3q19_kind_omega, rel_l2
from test_oracle_speciation import CARB_COMP, CARB_LOGK, CARB_S, carbonate_exact


def build(clb, mode, prec, layout=0):
    nx, ny, nz = 48, 24, 16
    pack = clb.gen_sphere_pack(nx, ny, nz, 3.0, 5.0, 0.5, seed=0)
    kind, om = d3q19_kind_omega(0.8)
    b = clb.gen_biomass_seed(pack, pore=2, fraction=0.05, bmin=10, bmax=100, seed_sites=1, seed_density=2)
    c_in, c_0 = carbonate_exact(2e-3, 4e-4), carbonate_exact(1e-3, 1e-4)
    lat = clb.Lattice(nx, ny, nz, precision=prec)
    lat.set_layout(layout)
    lat.set_speciation_mode(mode)
    lat.set_dynamics(kind, om)
    lat.upload_mask(pack)
    lat.flow_init(1.0, 0.999)
    lat.flow_steps(200)
    lat.set_solutes(np.full((4, 256), 1.0 / 0.8), [0, 1] * 4, [c_in[0], 0, c_in[1], 0, c_in[2], 0, c_in[3], 0])
    lat.couple_velocity()
    lat.set_biomass(b)
    lat.set_kinetics(clb.KINETICS_PLUGIN, 1.0, 0, 1)
    lat.set_mask_update(True, 10.0, 2, 3)
    lat.set_tableau(CARB_S, CARB_LOGK, CARB_COMP, tol=1e-10, max_iter=200)
    lat.solutes_init(c_0)
    lat.step(31)  # odd: the tuner starts from an odd step with deferred increments pending
    return lat


def snapshot(lat):
    u, rho = lat.velocity()
    return lat.conc(), lat.biomass(), lat.download_mask(), u, rho, lat.speciation_stats()


@pytest.mark.gpu
@pytest.mark.parametrize("mode,prec,layout", [(0, 1, 0), (1, 1, 0), (0, 0, 0), (0, 1, 1)])
def test_autotune_restores_state_and_matches(clb, mode, prec, layout):
    print(f"case: autotune, speciation mode {mode}, precision {prec}, layout {layout}", flush=True)
    lat = build(clb, mode, prec, layout)
    ref = build(clb, mode, prec, layout)
    before = snapshot(lat)
    rep = lat.autotune(4)
    after = snapshot(lat)
    for x, y in zip(before, after):
        assert np.array_equal(x, y)
    assert rep["flow_bx"] * rep["flow_by"] in (64, 128, 256) and rep["flow_ms"] > 0
    if mode == 0:
        assert rep["fused_bx"] * rep["fused_by"] in (256, 512) and 2 <= rep["fused_budget"] <= 4
        assert rep["ade_bx"] == 0 and rep["fused_ms"] > 0
    else:
        assert rep["ade_bx"] * rep["ade_by"] in (64, 128, 256) and rep["fused_bx"] == 0
    # flow: 3 block sizes on the compact layout, 6 shapes on the dense one; fused on the compact
    # layout: 2 block sizes x 3 budgets, on the dense one 6 shapes x 3 budgets; the separate
    # path's transport (dense): 6 shapes
    assert rep["trials"] == (3 if layout == 0 else 6) + ((6 if layout == 0 else 18) if mode == 0 else 6)
    # continue both runs: tuned shapes (and budget) against the defaults
    for l in (lat, ref):
        l.step(40)
        l.flow_steps(20)
    a, b = snapshot(lat), snapshot(ref)
    assert np.array_equal(a[2], b[2])
    for s in range(4):
        assert rel_l2(a[0][s], b[0][s]) < 2e-6
    assert np.array_equal(a[3], b[3]) and np.array_equal(a[4], b[4])  # flow: same arithmetic per voxel
    assert a[5][0] == b[5][0] and a[5][0] == a[5][1]
    lat.close()
    ref.close()


@pytest.mark.gpu
def test_autotune_arguments(clb):
    print("case: autotune argument checks", flush=True)
    lat = clb.Lattice(32, 16, 8, precision=1)
    with pytest.raises(clb.ClbError):
        lat.autotune(4)  # nothing set up
    kind, om = d3q19_kind_omega(0.8)
    lat.set_dynamics(kind, om)
    lat.upload_mask(clb.gen_sphere_pack(32, 16, 8, 3.0, 5.0, 0.5, seed=0))
    lat.flow_init(1.0, 0.999)
    for bad in (0, 3, -2):
        with pytest.raises(clb.ClbError):
            lat.autotune(bad)
    rep = lat.autotune(2)
    # the flow alone, on the compact layout: 3 block sizes
    assert rep["trials"] == 3 and rep["fused_bx"] == 0 and rep["ade_bx"] == 0
    lat.set_layout(clb.LAYOUT_DENSE)
    rep = lat.autotune(2)
    assert rep["trials"] == 6 and rep["fused_bx"] == 0 and rep["ade_bx"] == 0
    lat.close()
