"""Fused speciation (increments deferred to the next load) against the separate path.

The same coupled case -- transport, dual-Monod kinetics with biomass, and speciation on the
carbonate tableau of the reference's example 04 -- runs in both speciation modes: fused,
replayed from captured graphs, and separate, eager. Node conversions between the segments
turn pore voxels into walls and back, so the deferred increments are folded into the
populations (including those in flight towards walls) with the geometry changing under
them. Both modes solve every voxel to the same 1e-10 judge, so the fields agree to the
solver tolerance and the masks exactly, whatever the fused evaluation budget (voxels beyond
it continue in the persistent kernel from where the fused solve stopped) and block shape."""
import numpy as np
import pytest

from conftest import d3q19_kind_omega, rel_l2
from test_oracle_speciation import CARB_COMP, CARB_LOGK, CARB_S, carbonate_exact


def run(clb, oracle, mode, graph, nseg=4, seg=20, budget=None, block=None, straggler_budget=None):
    nx, ny, nz = 32, 16, 16
    pack = clb.gen_sphere_pack(nx, ny, nz, 3.0, 5.0, 0.5, seed=0)
    kind, om = d3q19_kind_omega(0.8)
    f, _ = oracle.flow_run(nx, ny, nz, pack, kind, om, 1.0, 0.999, 300)
    u, _ = oracle.flow_velocity(nx, ny, nz, pack, kind, 1.0, 0.999, f)
    b = clb.gen_biomass_seed(pack, pore=2, fraction=0.05, bmin=10, bmax=100, seed_sites=1, seed_density=2)
    c_in, c_0 = carbonate_exact(2e-3, 4e-4), carbonate_exact(1e-3, 1e-4)
    lat = clb.Lattice(nx, ny, nz, precision=0)
    lat.set_speciation_mode(mode)
    if budget is not None:
        lat.set_fused_budget(budget)
    if straggler_budget is not None:
        lat.set_straggler_budget(straggler_budget)
    if block is not None:
        lat.set_block_shape(clb.KERNEL_FUSED, *block)
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


_separate = {}


@pytest.mark.gpu
@pytest.mark.parametrize("budget,block", [(3, (64, 8)), (2, (64, 4)), (8, (32, 8)), (4, (128, 4))])
def test_fused_graph_matches_separate_with_node_conversion(clb, oracle, budget, block):
    print(f"case: fused (graph, budget {budget}, block {block}) vs separate (eager), node conversions", flush=True)
    cf, mf, sf, chf = run(clb, oracle, 0, True, budget=budget, block=block)
    if not _separate:
        _separate["r"] = run(clb, oracle, 1, False)
    cs, ms, ss, chs = _separate["r"]
    assert sum(chf) > 0 and chf == chs
    assert np.array_equal(mf, ms)
    assert sf[0] == sf[1] and ss[0] == ss[1] and sf[0] == ss[0]
    for s in range(4):
        assert rel_l2(cf[s], cs[s]) < 2e-6


@pytest.mark.gpu
@pytest.mark.parametrize("straggler_budget", [0, 1, 16])
def test_straggler_passes_match_separate(clb, oracle, straggler_budget):
    """The stragglers of a fused budget of 2 (about a quarter of the voxels) finished by the
    first pass alone (16), mostly by the full iteration after a one-evaluation first pass
    (1: the full iteration skips the entries the first pass solved), or by the full iteration
    alone (0): the same fields as the separate path, to the solver tolerance."""
    print(f"case: fused budget 2, straggler first-pass budget {straggler_budget}", flush=True)
    cf, mf, sf, chf = run(clb, oracle, 0, True, budget=2, block=(64, 4), straggler_budget=straggler_budget)
    if not _separate:
        _separate["r"] = run(clb, oracle, 1, False)
    cs, ms, ss, chs = _separate["r"]
    assert sum(chf) > 0 and chf == chs
    assert np.array_equal(mf, ms)
    assert sf[0] == sf[1] and sf[0] == ss[0]
    for s in range(4):
        assert rel_l2(cf[s], cs[s]) < 2e-6


@pytest.mark.gpu
def test_launch_settings_rejected(clb):
    print("case: budget / block shape argument checks", flush=True)
    lat = clb.Lattice(32, 16, 8, precision=0)
    for bad in (1, 0, 65):
        with pytest.raises(clb.ClbError):
            lat.set_fused_budget(bad)
    for bad in (-1, 65):
        with pytest.raises(clb.ClbError):
            lat.set_straggler_budget(bad)
    for kern, bx, by in ((clb.KERNEL_FUSED, 64, 16), (clb.KERNEL_ADE, 64, 8), (clb.KERNEL_FLOW, 32, 3), (7, 64, 4)):
        with pytest.raises(clb.ClbError):
            lat.set_block_shape(kern, bx, by)
    lat.set_block_shape(clb.KERNEL_FLOW, 128, 2)
    lat.set_fused_budget(5)
    lat.close()


@pytest.mark.gpu
def test_statistics_survive_a_block_shape_change(clb, oracle):
    """Fused solves counted before a block-shape change stay in clb_speciation_stats."""
    print("case: speciation statistics across a fused block-shape change", flush=True)
    nx, ny, nz = 32, 16, 16
    pack = clb.gen_sphere_pack(nx, ny, nz, 3.0, 5.0, 0.5, seed=0)
    kind, om = d3q19_kind_omega(0.8)
    c_in, c_0 = carbonate_exact(2e-3, 4e-4), carbonate_exact(1e-3, 1e-4)
    lat = clb.Lattice(nx, ny, nz, precision=0)
    lat.set_dynamics(kind, om)
    lat.upload_mask(pack)
    lat.set_solutes(np.full((4, 256), 1.0 / 0.8), [0, 1] * 4, [c_in[0], 0, c_in[1], 0, c_in[2], 0, c_in[3], 0])
    lat.set_velocity(np.zeros(3 * pack.size))
    lat.set_tableau(CARB_S, CARB_LOGK, CARB_COMP, tol=1e-10, max_iter=200)
    lat.solutes_init(c_0)
    lat.step(3)
    s1 = lat.speciation_stats()
    lat.set_block_shape(clb.KERNEL_FUSED, 64, 4)   # the default is 64 x 8: new wavefront count
    lat.step(2)
    s2 = lat.speciation_stats()
    lat.close()
    x = np.arange(pack.size) % nx
    react = int(((pack == 2) & (x >= 1) & (x <= nx - 2)).sum())
    assert s1[0] == s1[1] == 3 * react
    assert s2[0] == s2[1] == 5 * react
