"""GPU D3Q7 transport + kinetics + biomass threshold mask vs the FP64 oracle.

Flow + transport case: 2 solutes on the 64 x 32 x 32 pack, tau 0.8, west Dirichlet 1.0,
east Neumann, c0 = 0, 1000 steps after the flow. Kinetics case reactions on a reduced
domain: 4 solutes, dual-Monod (Ks 1e-2, k_max dt 1e-3, Y 0.4), biomass seeded in
1 % of pore voxels with B ~ U[10, 100], pore <-> biofilm reclassification at a total
biomass of 50 (chosen so growth moves voxels across it within the run). FP64 to
rounding, FP32 store within rel-L2 1e-5, the mask bit-exact."""
import numpy as np
import pytest

from conftest import d3q19_kind_omega, rel_l2

NX, NY, NZ = 64, 32, 32


@pytest.fixture(scope="module")
def setup(clb, oracle):
    pack = clb.gen_sphere_pack(NX, NY, NZ, 3.0, 5.0, 0.45, seed=0)
    kind, om = d3q19_kind_omega(0.8)
    f, _ = oracle.flow_run(NX, NY, NZ, pack, kind, om, 1.0, 0.999, 500)
    u, _ = oracle.flow_velocity(NX, NY, NZ, pack, kind, 1.0, 0.999, f)
    return pack, kind, u


def gpu_transport(clb, pack, kind, u, prec, S, bc_type, bc_value, c0, steps, biomass=None, kinetics=False,
                  mask_update=None, omega=None):
    lat = clb.Lattice(NX, NY, NZ, precision=prec)
    lat.set_dynamics(kind, np.full(256, 1.25))
    lat.upload_mask(pack)
    om = omega if omega is not None else np.full((S, 256), 1.0 / 0.8)
    lat.set_solutes(om, bc_type, bc_value)
    lat.set_velocity(u)
    if biomass is not None:
        lat.set_biomass(biomass)
    if kinetics:
        lat.set_kinetics(clb.KINETICS_PLUGIN, 1.0, 0, 1)
    if mask_update is not None:
        lat.set_mask_update(True, *mask_update)
    lat.solutes_init(c0)
    lat.step(steps)
    out = (lat.conc(), lat.biomass() if biomass is not None else None, lat.download_mask())
    lat.close()
    return out


@pytest.fixture(scope="module")
def oracle_transport(oracle, setup):
    pack, kind, u = setup
    run = oracle.AdeRun(NX, NY, NZ, pack, kind, np.full((2, 256), 1.25), [0, 1, 0, 1], [1.0, 0, 1.0, 0], u,
                        [0.0, 0.0])
    run.run(999)
    return run.conc()


@pytest.mark.gpu
@pytest.mark.parametrize("prec,tol", [(0, 1e-10), (1, 1e-5)])
def test_transport_two_solutes(clb, setup, oracle_transport, prec, tol):
    print(f"case: transport 2 solutes 999 steps prec={prec}", flush=True)
    pack, kind, u = setup
    c, _, _ = gpu_transport(clb, pack, kind, u, prec, 2, [0, 1, 0, 1], [1.0, 0, 1.0, 0], [0.0, 0.0], 999)
    for s in range(2):
        assert rel_l2(c[s], oracle_transport[s]) < tol
    assert c.max() <= 1.0 + 1e-6


def reactive_oracle(clb, oracle, setup, mask_update):
    pack, kind, u = setup
    b = clb.gen_biomass_seed(pack, pore=2, fraction=0.01, bmin=10, bmax=100, seed_sites=1, seed_density=2)
    mask = pack.copy()
    mask[b >= 50.0] = 3                                  # dense seeds start as biofilm
    om = np.full((4, 256), 1.0 / 0.8)
    om[:, 3] = 1.0 / 1.0                                 # solutes diffuse slower in biofilm
    bc_type = [0, 1, 0, 1, 1, 1, 1, 1]
    bc_val = [1.0, 0, 0.5, 0, 0, 0, 0, 0]
    c0 = [0.2, 0.1, 0.0, 0.0]
    mu = (50.0, 2, 3) if mask_update else None
    run = oracle.AdeRun(NX, NY, NZ, mask, kind, om, bc_type, bc_val, u, c0, kinetics=True, dt=1.0, biomass=b,
                        mask_update=mu)
    run.run(150)
    return dict(mask=mask, b=b, om=om, bc_type=bc_type, bc_val=bc_val, c0=c0, conc=run.conc(), mu=mu,
                bmass=run.bmass.copy(), mask_out=run.mask.copy())


@pytest.mark.gpu
@pytest.mark.parametrize("prec,tol,mask_update", [(0, 1e-10, True), (1, 1e-5, False)])
def test_kinetics_biomass_mask(clb, oracle, setup, prec, tol, mask_update):
    """FP64: transport, kinetics, biomass and the threshold mask rule (mask bit-exact).
    FP32 store: the same without the threshold rule, whose discrete pore/biofilm switch
    can trip one step apart when the biomass sits within FP32 rounding of it."""
    print(f"case: kinetics + biomass 150 steps prec={prec} mask_update={mask_update}", flush=True)
    _, kind, u = setup
    r = reactive_oracle(clb, oracle, setup, mask_update)
    c, b, m = gpu_transport(clb, r["mask"], kind, u, prec, 4, r["bc_type"], r["bc_val"], r["c0"], 150,
                            biomass=r["b"], kinetics=True, mask_update=r["mu"], omega=r["om"])
    for s in range(4):
        assert rel_l2(c[s], r["conc"][s]) < tol
    assert rel_l2(b[0], r["bmass"]) < tol
    assert r["conc"][2].max() > 1e-3                     # reactions happened
    if mask_update:
        assert np.array_equal(m, r["mask_out"])
        assert (r["mask_out"] != r["mask"]).any()        # growth moved voxels across the threshold
