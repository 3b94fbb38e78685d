"""The surrogate metabolic processor (run_surrogate3D, src/complab3d_processors_surrogate.hh:
74-295) on the GPU against the oracle's restatement (oracle/surrogate.cpp), through the C ABI.

Case: the 64 x 32 x 32 sphere pack with four solutes (donor D, acceptor A, products P1, P2)
and two biomass fields. Field 0 follows the compiled dual-Monod kinetics; field 1 is the
reference's shipped surrogate organism (Geobacter, the trained 2-10-10-10-10-1 network of
surrogateModel.hh, weights from tests/golden/surrogate_geobacter.json), whose increments
join the kinetics' in the same step. Pinned by:
  * FP64 GPU vs the oracle, 1e-10, and FP32 storage, 1e-5 (both with the shipped network,
    Vmax keeping its inputs inside its training box);
  * two surrogate organisms alone (no kinetics plugin): a small random network with the
    growth-limiting closure and a substrate floor, and one without a network (decay), with a
    held substrate (fix_concentration);
  * component totals with the fused speciation path (deferred increments) and the separate
    one, against the oracle;
  * a captured graph replays the eager steps bit for bit; switching the processor off
    restores the plain kinetics run bit for bit;
  * the argument checks of the two setters."""
import json
import os

import numpy as np
import pytest

import oracle as orc
from conftest import d3q19_kind_omega, rel_l2
from test_oracle_speciation import CARB_COMP, CARB_LOGK, CARB_S, carbonate_exact

HERE = os.path.dirname(os.path.abspath(__file__))
NX, NY, NZ = 64, 32, 32
BCT = [0, 1, 0, 1, 1, 1, 1, 1]
BCV = [1.0, 0, 0.5, 0, 0, 0, 0, 0]
C0 = [0.2, 0.1, 0.05, 0.0]
OM = np.full((4, 256), 1.0 / 0.8)
DT = 1.0


def geobacter():
    with open(os.path.join(HERE, "golden", "surrogate_geobacter.json")) as fh:
        return json.load(fh)["networks"]["0"]


@pytest.fixture(scope="module")
def case(clb):
    pack = clb.gen_sphere_pack(NX, NY, NZ, 3.0, 5.0, 0.45, seed=0)
    kind, om = d3q19_kind_omega(0.8)
    f, _ = orc.flow_run(NX, NY, NZ, pack, kind, om, 1.0, 0.999, 500)
    u, _ = orc.flow_velocity(NX, NY, NZ, pack, kind, 1.0, 0.999, f)
    b0 = clb.gen_biomass_seed(pack, pore=2, fraction=0.02, bmin=10, bmax=100, seed_sites=1, seed_density=2)
    b1 = clb.gen_biomass_seed(pack, pore=2, fraction=0.02, bmin=5, bmax=60, seed_sites=2, seed_density=2)
    return pack, kind, u, np.concatenate([b0, b1])


# the shipped organism on field 1: Vmax keeps its inputs (donor, acceptor uptake) in its box
SHIPPED = dict(ids=[1], Kc=[[0.05, 0.02, 0.0, 0.0]], decay=[6.94e-8], vmax=[[8.0, 0.4, 0.0, 0.0]],
               molar_mass=[24.6])


def random_net(seed, n_in, in_sub, hidden):
    rng = np.random.default_rng(seed)
    sizes = [n_in] + hidden + [1]
    params = np.concatenate([np.concatenate([rng.normal(scale=0.8, size=(sizes[k + 1], sizes[k])).ravel(),
                                             rng.normal(scale=0.3, size=sizes[k + 1])])
                             for k in range(len(sizes) - 1)])
    return dict(n_in=n_in, in_sub=in_sub, widths=hidden + [1], params=params, x_offset=[0.0] * n_in,
                x_gain=[0.5] * n_in, x_ymin=-1.0, y_ymin=-1.0, y_gain=4.0, y_offset=0.0)


def gpu_run(clb, case, prec, steps, kinetics, srg, nets, graph=False, off_after=False):
    pack, kind, u, b = case
    lat = clb.Lattice(NX, NY, NZ, precision=prec)
    lat.set_dynamics(kind, np.full(256, 1.25))
    lat.upload_mask(pack)
    lat.set_solutes(OM, BCT, BCV)
    lat.set_velocity(u)
    lat.set_biomass(b.reshape(2, -1))
    lat.set_kinetics(clb.KINETICS_PLUGIN if kinetics else clb.KINETICS_NONE, DT, 0, 1)
    lat.solutes_init(C0)
    lat.set_surrogate(**srg)
    for mid, (net, closure, yields) in nets.items():
        lat.set_surrogate_net(mid, net, closure, yields)
    if off_after:
        lat.set_surrogate([], [], [])
    if graph:
        lat.set_graph(True, 10)
    lat.step(steps)
    return lat


def oracle_run(case, steps, kinetics, srg, nets):
    pack, kind, u, b = case
    ids = srg["ids"]
    net_list = []
    for mid in ids:
        if mid in nets:
            net, closure, yields = nets[mid]
            net = dict(net, closure=closure)
            net["yield"] = yields
            net_list.append(net)
        else:
            net_list.append(None)
    s = orc.Surrogate(4, ids, srg["Kc"], srg["decay"], net_list, vmax=srg.get("vmax"),
                      max_uptake=srg.get("max_uptake"), molar_mass=srg.get("molar_mass"),
                      fix_conc=srg.get("fix_conc"))
    run = orc.AdeRun(NX, NY, NZ, pack, kind, OM, BCT, BCV, u, C0, kinetics=kinetics, dt=DT, biomass=b, surrogate=s)
    run.run(steps)
    return run


@pytest.mark.gpu
@pytest.mark.parametrize("prec,tol", [(0, 1e-10), (1, 1e-5)])
def test_shipped_surrogate_with_kinetics(clb, case, prec, tol):
    print(f"case: shipped surrogate + dual-Monod kinetics, precision {prec}, 40 steps", flush=True)
    nets = {1: (geobacter(), 0, None)}
    lat = gpu_run(clb, case, prec, 40, True, SHIPPED, nets)
    ref = oracle_run(case, 40, True, SHIPPED, nets)
    cg, cr = lat.conc(), ref.conc()
    for s in range(4):
        assert rel_l2(cg[s], cr[s]) < tol, s
    bg = lat.biomass().ravel()
    b0 = case[3]
    assert rel_l2(bg, ref.bmass) < tol
    # the surrogate organism grew where it sits, the donor and acceptor were drawn down
    grown = (ref.bmass - b0)[b0.size // 2:]
    assert grown.max() > 1e-3 and np.abs(bg - b0).max() > 1e-3
    lat.close()


@pytest.mark.gpu
def test_two_surrogate_organisms_without_kinetics(clb, case):
    print("case: two surrogate organisms, closure 2, floor, fix_concentration, FP64, 30 steps", flush=True)
    srg = dict(ids=[0, 1], Kc=[[0.05, 0.02, 0.1, 0.1], [0.02, 0.05, 0.1, 0.1]], decay=[1e-5, 2e-4],
               vmax=[[3.0, 1.0, 0.0, 0.5], [0.0, 0.0, 0.0, 0.0]],
               max_uptake=[[-1e30, -0.3, -1e30, -1e30], [-1e30, -1e30, -1e30, -1e30]],
               molar_mass=[24.6, 30.0], fix_conc=[0, 0, 1, 0])
    nets = {0: (random_net(5, 2, [0, 3], [6, 5]), 2, [0.04, 0.0, 0.0, 0.02])}
    lat = gpu_run(clb, case, 0, 30, False, srg, nets)
    ref = oracle_run(case, 30, False, srg, nets)
    cg, cr = lat.conc(), ref.conc()
    for s in range(4):
        assert rel_l2(cg[s], cr[s]) < 1e-10, s
    assert rel_l2(lat.biomass().ravel(), ref.bmass) < 1e-10
    lat.close()


@pytest.mark.gpu
def test_graph_replay_and_switch_off(clb, case):
    print("case: surrogate graph replay vs eager; processor switched off vs kinetics alone", flush=True)
    nets = {1: (geobacter(), 1, [0.05, 0.0, 0.0, 0.0])}
    eager = gpu_run(clb, case, 1, 40, True, SHIPPED, nets)
    graph = gpu_run(clb, case, 1, 40, True, SHIPPED, nets, graph=True)
    assert np.array_equal(eager.conc(), graph.conc())
    assert np.array_equal(eager.biomass(), graph.biomass())
    off = gpu_run(clb, case, 1, 20, True, SHIPPED, nets, off_after=True)
    plain = gpu_run(clb, case, 1, 20, True, dict(ids=[], Kc=[], decay=[]), {})
    assert np.array_equal(off.conc(), plain.conc())
    assert np.array_equal(off.biomass(), plain.biomass())
    for lat in (eager, graph, off, plain):
        lat.close()


# ------------------------------------------------------------------ totals + speciation
TNX, TNY, TNZ = 24, 12, 10


def totals_setup(clb, mode):
    N = TNX * TNY * TNZ
    mask = np.full(N, 2, dtype=np.uint8)
    m3 = mask.reshape(TNZ, TNY, TNX)
    m3[3:7, 4:8, 9:14] = 1
    kind = np.full(256, 2, dtype=np.int32)
    kind[0], kind[1] = 0, 1
    om = np.full((4, 256), 1.0 / 0.8)
    ci, c0 = carbonate_exact(2e-3, 4e-4), carbonate_exact(1e-3, 1e-4)
    bct = [0, 1] * 4
    bcv = [ci[0], 0, ci[1], 0, ci[2], 0, ci[3], 0]
    rng = np.random.default_rng(9)
    b = np.where((mask == 2) & (rng.random(N) < 0.3), rng.uniform(1.0, 20.0, N), 0.0)
    # one organism drawing on the two components (species named like them: comp_species)
    srg = dict(ids=[0], Kc=[[1e-4] * 4], decay=[1e-6], vmax=[[0.5, 0.2, 0.0, 0.0]], use_totals=True)
    net = random_net(7, 2, [int(CARB_COMP[0]), int(CARB_COMP[1])], [4])
    lat = clb.Lattice(TNX, TNY, TNZ, precision=0)
    lat.set_speciation_mode(mode)
    lat.set_dynamics(kind, np.full(256, 1.25))
    lat.upload_mask(mask)
    lat.set_solutes(om, bct, bcv)
    lat.set_velocity(np.zeros((N, 3)))
    lat.set_biomass(b[None, :])
    lat.set_kinetics(clb.KINETICS_NONE, 1.0, 0, 1)
    lat.set_tableau(CARB_S, CARB_LOGK, CARB_COMP, tol=1e-10, max_iter=200)
    lat.solutes_init(list(c0))
    lat.set_surrogate(**srg)
    lat.set_surrogate_net(0, net, 0, None)
    s = orc.Surrogate(4, [0], srg["Kc"], srg["decay"], [net], vmax=srg["vmax"], use_totals=True,
                      tableau=(CARB_S, CARB_COMP))
    ref = orc.AdeRun(TNX, TNY, TNZ, mask, kind, om, bct, bcv, np.zeros((N, 3)), list(c0), dt=1.0,
                     biomass=b, tableau=(CARB_S, CARB_LOGK, CARB_COMP), surrogate=s)
    return lat, ref


@pytest.mark.gpu
def test_component_totals_with_speciation(clb):
    print("case: surrogate on component totals, fused (graph) and separate speciation vs oracle", flush=True)
    out = []
    for mode, graph in ((0, True), (1, False)):
        lat, ref = totals_setup(clb, mode)
        if graph:
            lat.set_graph(True, 10)
        lat.step(60)
        out.append((lat.conc(), lat.biomass().ravel()))
        lat.close()
    ref.run(60)
    cr = ref.conc()
    for s in range(4):
        assert rel_l2(out[0][0][s], out[1][0][s]) < 2e-6, s
        assert rel_l2(out[1][0][s], cr[s]) < 2e-6, s
    assert rel_l2(out[1][1], ref.bmass) < 2e-6
    assert rel_l2(out[0][1], out[1][1]) < 2e-6


@pytest.mark.gpu
def test_setter_checks(clb):
    print("case: surrogate setter argument checks", flush=True)
    lat = clb.Lattice(8, 4, 4, precision=0)
    lat.set_solutes(np.full((2, 256), 1.0), [0, 1, 0, 1], [0, 0, 0, 0])
    with pytest.raises(clb.ClbError):
        lat.set_surrogate([0, 0], [[0.1, 0.1]] * 2, [0.0, 0.0])             # listed twice
    with pytest.raises(clb.ClbError):
        lat.set_surrogate([0], [[0.1, 0.1]], [0.0], molar_mass=[0.0])      # molar mass <= 0
    lat.set_surrogate([1], [[0.1, 0.1]], [0.0])
    net = random_net(1, 2, [0, 1], [4])
    with pytest.raises(clb.ClbError):
        lat.set_surrogate_net(0, net)                                       # not a surrogate microbe
    with pytest.raises(clb.ClbError):
        lat.set_surrogate_net(1, dict(net, widths=[4, 2]))                  # output of width 2
    with pytest.raises(clb.ClbError):
        lat.set_surrogate_net(1, random_net(1, 2, [0, 1], [17]))            # wider than 16
    with pytest.raises(clb.ClbError):
        lat.set_surrogate_net(1, dict(net, in_sub=[0, 2]))                  # no substrate 2
    lat.set_surrogate_net(1, net)
    lat.set_surrogate_net(1, None)
    lat.close()


@pytest.mark.gpu
def test_autotune_before_the_first_step(clb, case):
    """The autotuner runs the step on its own; the buffers the surrogate kernel reads are set
    up for it as for clb_step (regression: they were only made by the first clb_step)."""
    print("case: surrogate, autotune before any step, then 20 steps vs no autotune", flush=True)
    nets = {1: (geobacter(), 0, None)}
    pack, kind, u, b = case
    out = []
    for tune in (True, False):
        lat = clb.Lattice(NX, NY, NZ, precision=1)
        lat.set_dynamics(kind, np.full(256, 1.25))
        lat.upload_mask(pack)
        lat.set_solutes(OM, BCT, BCV)
        lat.set_velocity(u)
        lat.set_biomass(b.reshape(2, -1))
        lat.set_kinetics(clb.KINETICS_PLUGIN, DT, 0, 1)
        lat.solutes_init(C0)
        lat.set_surrogate(**SHIPPED)
        lat.set_surrogate_net(1, geobacter())
        if tune:
            rep = lat.autotune(2)
            assert rep["trials"] > 0
            lat.set_block_shape(clb.KERNEL_ADE, 64, 4)  # the same arithmetic order either way
        lat.step(20)
        out.append((lat.conc(), lat.biomass()))
        lat.close()
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])
