"""Fused speciation (increments deferred to the next load) against the separate path.

The same coupled case -- transport, dual-Monod kinetics with biomass, and speciation on the
carbonate tableau of the reference's example 04 -- runs in both speciation modes: fused,
replayed from captured graphs, and separate, eager. Node conversions between the segments
turn pore voxels into walls and back, so the deferred increments are folded into the
populations (including those in flight towards walls) with the geometry changing under
them. Both modes solve every voxel to the same 1e-10 judge, so the fields agree to the
solver tolerance and the masks exactly."""
import numpy as np
import pytest

from conftest import d3q19_kind_omega, rel_l2
from test_oracle_speciation import CARB_COMP, CARB_LOGK, CARB_S, carbonate_exact


def run(clb, oracle, mode, graph, nseg=4, seg=20):
    nx, ny, nz = 32, 16, 16
    pack = clb.gen_sphere_pack(nx, ny, nz, 3.0, 5.0, 0.5, seed=0)
    kind, om = d3q19_kind_omega(0.8)
    f, _ = oracle.flow_run(nx, ny, nz, pack, kind, om, 1.0, 0.999, 300)
    u, _ = oracle.flow_velocity(nx, ny, nz, pack, kind, 1.0, 0.999, f)
    b = clb.gen_biomass_seed(pack, pore=2, fraction=0.05, bmin=10, bmax=100, seed_sites=1, seed_density=2)
    c_in, c_0 = carbonate_exact(2e-3, 4e-4), carbonate_exact(1e-3, 1e-4)
    lat = clb.Lattice(nx, ny, nz, precision=0)
    lat.set_speciation_mode(mode)
    lat.set_dynamics(kind, om)
    lat.upload_mask(pack)
    lat.set_solutes(np.full((4, 256), 1.0 / 0.8), [0, 1] * 4, [c_in[0], 0, c_in[1], 0, c_in[2], 0, c_in[3], 0])
    lat.set_velocity(u)
    lat.set_biomass(b)
    lat.set_tableau(CARB_S, CARB_LOGK, CARB_COMP, tol=1e-10, max_iter=200)
    lat.solutes_init(c_0)
    lat.set_phase(np.zeros(pack.size, dtype=np.uint8))
    if graph:
        lat.set_graph(True, seg)
    changes = []
    for k in range(nseg):
        lat.step(seg)
        # carbonic acid above its 95 % quantile -> wall; back to pore below 90 % of it
        c = lat.conc()[3]
        thr = float(np.quantile(c[pack == 2], 0.95))
        changes.append(lat.node_conversion(3, 2, 1, 5, thr if k % 2 == 0 else -1.0, 0.9 * thr if k % 2 else -1.0))
    c = lat.conc()
    m = lat.download_mask()
    st = lat.speciation_stats()
    lat.close()
    return c, m, st, changes


@pytest.mark.gpu
def test_fused_graph_matches_separate_with_node_conversion(clb, oracle):
    print("case: fused (graph) vs separate (eager), node conversions between segments", flush=True)
    cf, mf, sf, chf = run(clb, oracle, 0, True)
    cs, ms, ss, chs = run(clb, oracle, 1, False)
    assert sum(chf) > 0 and chf == chs
    assert np.array_equal(mf, ms)
    assert sf[0] == sf[1] and ss[0] == ss[1] and sf[0] == ss[0]
    for s in range(4):
        assert rel_l2(cf[s], cs[s]) < 2e-6
