"""GPU abiotic kinetics, immobile solutes, node conversion, flow statistics, solute reset
and captured-graph replay, each against the FP64 oracle (or the eager path for the graph).

Case: the 64 x 32 x 32 sphere pack with four solutes (donor D, acceptor A, products P1,
P2), dual-Monod biotic kinetics on biomass seeds, and the second-order abiotic reaction
P1 + A -> P2 (run_abiotic_kinetics / surfaceAbioticKinetics3D,
src/complab3d_processors_part1.hh:158-336, src/precipitationVOP.hh:53-145). P2 is made
immobile (vec_immobile, src/complab.cpp:248-266), so it accumulates where it forms and
serves as the precipitate for node conversion (precipNodeConversion3D,
src/precipitationVOP.hh:176-204)."""
import numpy as np
import pytest

from conftest import d3q19_kind_omega, rel_l2

NX, NY, NZ = 64, 32, 32
BCT = [0, 1, 0, 1, 1, 1, 1, 1]
BCV = [1.0, 0, 0.5, 0, 0, 0, 0, 0]
C0 = [0.2, 0.1, 0.05, 0.0]


@pytest.fixture(scope="module")
def case(clb, oracle):
    pack = clb.gen_sphere_pack(NX, NY, NZ, 3.0, 5.0, 0.45, seed=0)
    kind, om = d3q19_kind_omega(0.8)
    f, _ = oracle.flow_run(NX, NY, NZ, pack, kind, om, 1.0, 0.999, 500)
    u, _ = oracle.flow_velocity(NX, NY, NZ, pack, kind, 1.0, 0.999, f)
    b = clb.gen_biomass_seed(pack, pore=2, fraction=0.02, bmin=10, bmax=100, seed_sites=1, seed_density=2)
    return pack, kind, om, u, b


def gpu_run(clb, case, prec, steps, abiotic, immobile, graph=False):
    pack, kind, om, u, b = case
    lat = clb.Lattice(NX, NY, NZ, precision=prec)
    lat.set_dynamics(kind, om)
    lat.upload_mask(pack)
    lat.set_solutes(np.full((4, 256), 1.0 / 0.8), BCT, BCV)
    lat.set_velocity(u)
    lat.set_biomass(b)
    lat.set_kinetics(clb.KINETICS_PLUGIN, 1.0, 0, 1)
    lat.set_abiotic(abiotic)
    lat.set_immobile(immobile)
    lat.solutes_init(C0)
    if graph:
        lat.set_graph(True, 20)
    lat.step(steps)
    return lat


def oracle_run(oracle, case, steps, abiotic, immobile):
    pack, kind, om, u, b = case
    run = oracle.AdeRun(NX, NY, NZ, pack, kind, np.full((4, 256), 1.0 / 0.8), BCT, BCV, u, C0, kinetics=True,
                        biomass=b, immobile=immobile, abiotic=abiotic)
    run.run(steps)
    return run


@pytest.mark.gpu
@pytest.mark.parametrize("abiotic", [1, 2])
@pytest.mark.parametrize("prec,tol", [(0, 1e-10), (1, 1e-5)])
def test_abiotic_and_immobile(clb, oracle, case, abiotic, prec, tol):
    print(f"case: abiotic mode {abiotic} + immobile P2, prec={prec}", flush=True)
    imm = [0, 0, 0, 1]
    run = oracle_run(oracle, case, 121, abiotic, imm)
    ref = run.conc()
    lat = gpu_run(clb, case, prec, 121, abiotic, imm)
    c = lat.conc()
    lat.close()
    assert ref[3].max() > 1e-4                          # the abiotic product formed
    for s in range(4):
        assert rel_l2(c[s], ref[s]) < tol
    # immobile: P2 never leaves the voxels where the abiotic reaction ran
    pack = case[0]
    assert np.all(ref[3][pack != 2] == 0.0)


@pytest.mark.gpu
def test_node_conversion_bit_exact(clb, oracle, case):
    print("case: node conversion", flush=True)
    imm = [0, 0, 0, 1]
    lat = gpu_run(clb, case, 0, 121, 1, imm)
    c = lat.conc()
    mineral = c[3].copy()
    thr = np.quantile(mineral[mineral > 0], 0.9)
    mask0 = lat.download_mask()
    phase0 = np.zeros(mask0.size, dtype=np.uint8)
    lat.set_phase(phase0)
    # precipitation: non-wall voxels (interior or faces) whose mineral reaches thr -> bb, phase 7
    changed = lat.node_conversion(3, 2, 1, 7, thr, -1.0)
    m_ref, ph_ref, n_ref = oracle.node_conversion(mask0, phase0, mineral, 2, 0, 1, 1, 7, thr, -1.0)
    assert changed == n_ref > 0
    assert np.array_equal(lat.download_mask(), m_ref) and np.array_equal(lat.phase(), ph_ref)
    # partial reopen: the stamped voxels below the median of their mineral dissolve back
    thr_open = float(np.median(mineral[ph_ref == 7]))
    changed_p = lat.node_conversion(3, 2, 1, 7, -1.0, thr_open)
    m_ref, ph_ref, n_ref = oracle.node_conversion(m_ref, ph_ref, mineral, 2, 0, 1, 1, 7, -1.0, thr_open)
    assert changed_p == n_ref and 0 < n_ref < changed
    assert np.array_equal(lat.download_mask(), m_ref) and np.array_equal(lat.phase(), ph_ref)
    # the step keeps running on the converted geometry
    lat.step(4)
    assert np.all(np.isfinite(lat.conc()))
    # full reopen: a threshold above every mineral amount dissolves the rest
    changed2 = lat.node_conversion(3, 2, 1, 7, -1.0, 2.0 * lat.conc()[3].max() + 1.0)
    assert changed_p + changed2 == changed
    assert np.array_equal(lat.download_mask(), mask0)
    assert not lat.phase().any()
    lat.close()


@pytest.mark.gpu
def test_flow_stats(clb, oracle):
    print("case: flow stats", flush=True)
    pack = clb.gen_sphere_pack(NX, NY, NZ, 3.0, 5.0, 0.45, seed=0)
    kind, om = d3q19_kind_omega(0.8)
    lat = clb.Lattice(NX, NY, NZ, precision=0)
    lat.set_dynamics(kind, om)
    lat.upload_mask(pack)
    lat.flow_init(1.0, 0.999)
    lat.flow_steps(300)
    st = lat.flow_stats()
    u, _ = lat.velocity()
    lat.close()
    v = u.reshape(NZ, NY, NX, 3)
    fl = (pack.reshape(NZ, NY, NX) == 2)
    fl[:, :, 0] = False
    fl[:, :, NX - 1] = False
    speed = np.linalg.norm(v, axis=3)
    assert abs(st[0] - speed[fl].mean()) < 1e-12 * max(speed.max(), 1e-30) + 1e-18
    assert abs(st[1] - speed[fl].max()) < 1e-15
    assert abs(st[2] - v[:, :, NX - 2, 0].mean()) < 1e-15
    assert st[3] == fl.sum()
    # Darcy permeability as the driver computes it (src/complab.cpp:648)
    nu, L, dP = (0.8 - 0.5) / 3.0, NX - 2, 1e-3 / 3.0
    assert st[2] * nu * L / dP > 0


@pytest.mark.gpu
def test_solutes_reset(clb, oracle, case):
    print("case: solutes reset", flush=True)
    imm = [0, 0, 0, 1]
    pack = case[0]
    lat = gpu_run(clb, case, 0, 31, 0, imm)
    c_new = [0.3, 0.2, 0.1, 0.0]
    lat.solutes_reset(c_new)
    c = lat.conc()
    interior = np.zeros((NZ, NY, NX), dtype=bool)
    interior[:, :, 1:NX - 1] = True
    fl = (pack == 2) & interior.ravel()
    for s in range(3):
        assert np.max(np.abs(c[s][fl] - c_new[s])) < 1e-15
    # the same reset applied to the oracle's state, then both continue
    run = oracle_run(oracle, case, 31, 0, imm)
    g = run.g.reshape(4, 7, -1)
    w = np.array([0.25] + [0.125] * 6)
    for s in range(3):
        g[s][:, pack == 2] = (w * (c_new[s] - 1.0))[:, None]
    # the Neumann faces copy the (reset) interior density at the next step
    rho = run.bc_rho.reshape(4, -1)
    for s in range(3):
        rho[s] = c_new[s]
    run.run(20)
    lat.step(20)
    ref, c = run.conc(), lat.conc()
    lat.close()
    for s in range(4):
        assert rel_l2(c[s], ref[s]) < 1e-10


@pytest.mark.gpu
def test_graph_replay_matches_eager(clb, case):
    print("case: hip graph replay", flush=True)
    imm = [0, 0, 0, 1]
    a = gpu_run(clb, case, 0, 61, 2, imm, graph=False)
    b = gpu_run(clb, case, 0, 61, 2, imm, graph=True)
    ca, cb = a.conc(), b.conc()
    a.close()
    b.close()
    assert np.array_equal(ca, cb)
