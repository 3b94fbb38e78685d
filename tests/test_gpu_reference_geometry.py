"""The reference's own example geometry on the GPU: examples/input_shared/grains.dat (copied
under tests/golden/), read at its XML size 24 x 24 x 6 into the 26 x 24 x 6 lattice. Its four
grains are no-dynamics voxels (code 0) right next to pore voxels, with no bounce-back layer
between: Palabos' NoDynamics does not collide, but the lattice's streaming still moves its
populations, so populations stream into the grains and on through them (as the oracle's
push stream does). The flow (FP64, 500 steps), transport of two solutes with dual-Monod-free
inlet chemistry (FP64 1e-10, FP32 1e-5) and the coupled speciation of the carbonate tableau
(run as its own kernel there: 2e-6) against the oracle."""
import os

import numpy as np
import pytest

from conftest import d3q19_kind_omega, rel_l2
from test_oracle_speciation import CARB_COMP, CARB_LOGK, CARB_S, carbonate_exact

NX, NY, NZ = 26, 24, 6
PATH = os.path.join(os.path.dirname(__file__), "golden", "grains_24x24x6.dat")


@pytest.fixture(scope="module")
def geo(clb, oracle):
    mask = clb.read_geometry(PATH, 24, 24, 6, clb.GEOM_X_FASTEST)
    kind, om = d3q19_kind_omega(0.8)
    f, _ = oracle.flow_run(NX, NY, NZ, mask, kind, om, 1.0, 0.999, 500)
    u, r = oracle.flow_velocity(NX, NY, NZ, mask, kind, 1.0, 0.999, f)
    return mask, kind, om, u, r


@pytest.mark.gpu
@pytest.mark.parametrize("prec,tol", [(0, 1e-10), (1, 1e-5)])
def test_flow_through_grains(clb, geo, prec, tol):
    print(f"case: flow on the reference grains geometry, prec={prec}", flush=True)
    mask, kind, om, uo, ro = geo
    assert np.any(mask == 0)
    lat = clb.Lattice(NX, NY, NZ, precision=prec)
    lat.set_dynamics(kind, om)
    lat.upload_mask(mask)
    lat.flow_init(1.0, 0.999)
    lat.flow_steps(500)
    u, r = lat.velocity()
    lat.close()
    assert rel_l2(u, uo) < tol
    if prec == 0:
        assert np.max(np.abs(r - ro)) < 1e-12
    ux = u[:, 0].reshape(NZ, NY, NX)
    assert ux[:, 1:23, 1:-1].mean() > 0 and np.all(ux[:, 0, :] == 0) and np.all(ux[:, 23, :] == 0)


@pytest.mark.gpu
@pytest.mark.parametrize("layout", [0, 1])
@pytest.mark.parametrize("prec,tol", [(0, 1e-10), (1, 1e-5)])
def test_transport_through_grains(clb, oracle, geo, prec, tol, layout):
    print(f"case: transport on the reference grains geometry, prec={prec} layout={layout}", flush=True)
    mask, kind, om, u, _ = geo
    bct, bcv, c0 = [0, 1, 0, 1], [1.0, 0.0, 0.5, 0.0], [0.0, 0.2]
    run = oracle.AdeRun(NX, NY, NZ, mask, kind, np.full((2, 256), 1.0 / 0.8), bct, bcv, u, c0)
    run.run(300)
    ref = run.conc()
    lat = clb.Lattice(NX, NY, NZ, precision=prec)
    lat.set_layout(layout)
    lat.set_dynamics(kind, om)
    lat.upload_mask(mask)
    lat.set_solutes(np.full((2, 256), 1.0 / 0.8), bct, bcv)
    lat.set_velocity(u)
    lat.solutes_init(c0)
    lat.step(300)
    c = lat.conc()
    lat.close()
    for s in range(2):
        assert rel_l2(c[s], ref[s]) < tol, s


@pytest.mark.gpu
def test_speciation_through_grains(clb, oracle, geo):
    print("case: coupled carbonate speciation on the reference grains geometry", flush=True)
    mask, kind, om, u, _ = geo
    c_in, c_0 = carbonate_exact(2e-3, 4e-4), carbonate_exact(1e-3, 1e-4)
    bct = [0, 1] * 4
    bcv = [c_in[0], 0, c_in[1], 0, c_in[2], 0, c_in[3], 0]
    run = oracle.AdeRun(NX, NY, NZ, mask, kind, np.full((4, 256), 1.0 / 0.8), bct, bcv, u, c_0,
                        tableau=(CARB_S, CARB_LOGK, CARB_COMP), eq_tol=1e-10)
    run.run(60)
    ref = run.conc()
    lat = clb.Lattice(NX, NY, NZ, precision=0)
    lat.set_dynamics(kind, om)
    lat.upload_mask(mask)
    lat.set_solutes(np.full((4, 256), 1.0 / 0.8), bct, bcv)
    lat.set_velocity(u)
    lat.set_tableau(CARB_S, CARB_LOGK, CARB_COMP, tol=1e-10, max_iter=200)
    lat.solutes_init(c_0)
    lat.set_graph(True, 20)
    lat.step(60)
    c = lat.conc()
    st = lat.speciation_stats()
    lat.close()
    assert st[0] > 0 and st[0] == st[1]
    for s in range(4):
        assert rel_l2(c[s], ref[s]) < 2e-6, s
