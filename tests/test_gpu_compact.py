"""The compact (fluid-only) solute layout against the dense one: bit for bit.

The step kernels on the compact layout read and write the same populations as the dense
kernels, in the same order of arithmetic, only at other addresses (the fluid voxels packed
in natural order, a ghost slot standing for each wall slot a bounce-back link uses). So the
concentrations, biomass and mask after any number of steps must be identical, whatever the
step parity, the precision, the captured graphs, the fused speciation's budget, and node
conversions between the segments (each one rebuilds the compact table from the dense state
and gathers it again)."""
import numpy as np
import pytest

from conftest import d3q19_kind_omega

NX, NY, NZ = 64, 32, 32


def kinetics_run(clb, oracle, layout, prec, steps, graph):
    pack = clb.gen_sphere_pack(NX, NY, NZ, 3.0, 5.0, 0.45, seed=0)
    kind, om = d3q19_kind_omega(0.8)
    f, _ = oracle.flow_run(NX, NY, NZ, pack, kind, om, 1.0, 0.999, 300)
    u, _ = oracle.flow_velocity(NX, NY, NZ, pack, kind, 1.0, 0.999, f)
    b = clb.gen_biomass_seed(pack, pore=2, fraction=0.02, bmin=10, bmax=100, seed_sites=1, seed_density=2)
    lat = clb.Lattice(NX, NY, NZ, precision=prec)
    lat.set_layout(layout)
    lat.set_dynamics(kind, om)
    lat.upload_mask(pack)
    lat.set_solutes(np.full((4, 256), 1.0 / 0.8), [0, 1, 0, 1, 1, 1, 1, 1], [1.0, 0, 0.5, 0, 0, 0, 0, 0])
    lat.set_velocity(u)
    lat.set_biomass(b)
    lat.set_kinetics(clb.KINETICS_PLUGIN, 1.0, 0, 1)
    lat.set_mask_update(True, 30.0, 2, 3)
    lat.solutes_init([0.2, 0.1, 0.05, 0.0])
    if graph:
        lat.set_graph(True, 20)
    lat.step(steps)
    out = (lat.conc(), lat.biomass(), lat.download_mask())
    lat.step(7)  # continues after the download (back to the compact layout)
    out2 = (lat.conc(), lat.biomass(), lat.download_mask())
    lat.close()
    return out, out2


@pytest.mark.gpu
@pytest.mark.parametrize("prec", [0, 1])
@pytest.mark.parametrize("graph", [False, True])
def test_kinetics_compact_equals_dense(clb, oracle, prec, graph):
    print(f"case: kinetics + mask rule, compact vs dense, prec={prec} graph={graph}", flush=True)
    dense = kinetics_run(clb, oracle, clb.LAYOUT_DENSE, prec, 121, graph)
    comp = kinetics_run(clb, oracle, clb.LAYOUT_COMPACT, prec, 121, graph)
    for (cd, bd, md), (cc, bc, mc) in zip(dense, comp):
        assert np.array_equal(cd, cc)
        assert np.array_equal(bd, bc)
        assert np.array_equal(md, mc)
    assert not np.array_equal(dense[0][0], dense[1][0])  # the state moved between the checks


def headline_run(clb, oracle, layout, prec, budget, graph):
    from test_gpu_headline import inputs, gpu_lattice, kind_omega
    nx, ny, nz = 48, 24, 24
    pack, mask, b, S, lk, c0, cin = inputs(clb, oracle, nx, ny, nz)
    kind, om = kind_omega()
    f, _ = oracle.flow_run(nx, ny, nz, mask, kind, om, 1.0, 0.999, 300)
    u, _ = oracle.flow_velocity(nx, ny, nz, mask, kind, 1.0, 0.999, f)
    lat = gpu_lattice(clb, nx, ny, nz, prec, mask, u, b, S, lk, cin, c0, 0, budget)
    lat.set_layout(layout)
    lat.set_phase(np.zeros(mask.size, dtype=np.uint8))
    if graph:
        lat.set_graph(True, 10)
    changes = []
    x = np.arange(mask.size) % nx
    inner = (mask >= 2) & (x >= 1) & (x <= nx - 2)
    for k in range(4):
        lat.step(20)
        c = lat.conc()[5]
        thr = float(np.quantile(c[inner], 0.97))
        # precipitation (pore -> wall) and reopening in turn: the fluid set changes
        changes.append(lat.node_conversion(5, 2, 1, 9, thr if k % 2 == 0 else -1.0, 0.9 * thr if k % 2 else -1.0))
    lat.step(13)
    out = (lat.conc(), lat.biomass(), lat.download_mask(), lat.speciation_stats())
    strag = lat.perf().stragglers
    lat.close()
    return out, changes, strag


@pytest.mark.gpu
@pytest.mark.parametrize("prec,budget,graph", [(0, 3, False), (1, 3, True), (1, 2, True), (0, 2, True)])
def test_fused_compact_equals_dense_with_node_conversion(clb, oracle, prec, budget, graph):
    print(f"case: headline shape 48x24x24, compact vs dense, prec={prec} budget={budget} graph={graph}", flush=True)
    (cd, bd, md, sd), chd, strd = headline_run(clb, oracle, clb.LAYOUT_DENSE, prec, budget, graph)
    (cc, bc, mc, sc), chc, strc = headline_run(clb, oracle, clb.LAYOUT_COMPACT, prec, budget, graph)
    assert chd == chc and sum(chd) > 0
    assert np.array_equal(md, mc)
    assert np.array_equal(cd, cc)
    assert np.array_equal(bd, bc)
    assert np.array_equal(sd, sc) and sd[0] == sd[1]
    assert strd == strc
    if budget == 2:
        assert strc > 0
