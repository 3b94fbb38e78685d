"""GPU D3Q19 flow vs the FP64 oracle on the 64 x 32 x 32 sphere pack (porosity 0.45,
radius 3-5, seed 0), tau 0.8, density drop 1e-3 along x, periodic y / z: FP64 storage to
rounding (rel-L2 <= 1e-10), FP32 deviation storage to rel-L2 <= 1e-5 (the stated bound
for the FP32 store)."""
import numpy as np
import pytest

from conftest import d3q19_kind_omega, rel_l2

NX, NY, NZ = 64, 32, 32
RHO_IN, RHO_OUT = 1.0, 0.999


@pytest.fixture(scope="module")
def pack(clb):
    return clb.gen_sphere_pack(NX, NY, NZ, 3.0, 5.0, 0.45, seed=0)


@pytest.fixture(scope="module")
def oracle_flow(oracle, pack):
    kind, om = d3q19_kind_omega(0.8)
    out = {}
    f, e = oracle.flow_run(NX, NY, NZ, pack, kind, om, RHO_IN, RHO_OUT, 301, energy=True)
    out[301] = (oracle.flow_velocity(NX, NY, NZ, pack, kind, RHO_IN, RHO_OUT, f), e)
    f, e2 = oracle.flow_run(NX, NY, NZ, pack, kind, om, RHO_IN, RHO_OUT, 1699, energy=True, f=f)
    out[2000] = (oracle.flow_velocity(NX, NY, NZ, pack, kind, RHO_IN, RHO_OUT, f), np.concatenate([e, e2]))
    return out


def run_gpu(clb, pack, prec, steps):
    lat = clb.Lattice(NX, NY, NZ, precision=prec)
    kind, om = d3q19_kind_omega(0.8)
    lat.set_dynamics(kind, om)
    lat.upload_mask(pack)
    lat.flow_init(RHO_IN, RHO_OUT)
    energy = []
    done = 0
    for n in steps:
        energy.append(lat.flow_steps(n - done, energy=True))
        done = n
    (u, rho) = lat.velocity()
    lat.close()
    return u, rho, np.concatenate(energy)


@pytest.mark.gpu
def test_flow_fp64_matches_oracle(clb, pack, oracle_flow):
    print("case: flow fp64 64x32x32, 301 and 2000 steps", flush=True)
    lat = clb.Lattice(NX, NY, NZ, precision=clb.PREC_FP64)
    kind, om = d3q19_kind_omega(0.8)
    lat.set_dynamics(kind, om)
    lat.upload_mask(pack)
    lat.flow_init(RHO_IN, RHO_OUT)
    e1 = lat.flow_steps(301, energy=True)            # odd count: AA layout mid-pair
    u301, r301 = lat.velocity()
    (uo, ro), eo = oracle_flow[301]
    assert rel_l2(u301, uo) < 1e-10
    assert np.max(np.abs(r301 - ro)) < 1e-12
    # the first step's energy is zero (the state at rest); the oracle's holds rounding noise
    assert np.all(np.abs(e1 - eo) <= 1e-9 * eo + 1e-30)
    lat.flow_steps(1699)
    u, r = lat.velocity()
    (uo, ro), _ = oracle_flow[2000]
    assert rel_l2(u, uo) < 1e-10
    assert np.max(np.abs(r - ro)) < 1e-12
    lat.close()


@pytest.mark.gpu
def test_flow_fp32_deviation_store(clb, pack, oracle_flow):
    print("case: flow fp32 64x32x32, 2000 steps", flush=True)
    lat = clb.Lattice(NX, NY, NZ, precision=clb.PREC_FP32)
    kind, om = d3q19_kind_omega(0.8)
    lat.set_dynamics(kind, om)
    lat.upload_mask(pack)
    lat.flow_init(RHO_IN, RHO_OUT)
    lat.flow_steps(2000)
    u, r = lat.velocity()
    (uo, ro), _ = oracle_flow[2000]
    assert rel_l2(u, uo) < 1e-5
    lat.close()


@pytest.mark.gpu
def test_bounce_back_nodes_untouched(clb, pack):
    # solid and wall nodes carry no velocity; the pore velocity field is nonzero
    print("case: flow fp64 wall nodes", flush=True)
    lat = clb.Lattice(NX, NY, NZ, precision=clb.PREC_FP64)
    kind, om = d3q19_kind_omega(0.8)
    lat.set_dynamics(kind, om)
    lat.upload_mask(pack)
    lat.flow_init(RHO_IN, RHO_OUT)
    lat.flow_steps(40)
    u, _ = lat.velocity()
    assert np.all(u[pack != 2] == 0)
    assert np.abs(u[pack == 2, 0]).max() > 0
    lat.close()


@pytest.mark.gpu
@pytest.mark.parametrize("block", [(64, 3), (192, 1), (64, 1)])
def test_energy_any_block_shape(clb, pack, oracle_flow, block):
    """The device-reduced energy (the ValueTracer's input) is the same for block shapes whose
    thread count is not a power of two: the per-block sum covers every wavefront."""
    print(f"case: flow fp64 energy with block {block}", flush=True)
    lat = clb.Lattice(NX, NY, NZ, precision=clb.PREC_FP64)
    kind, om = d3q19_kind_omega(0.8)
    lat.set_block_shape(clb.KERNEL_FLOW, *block)
    lat.set_dynamics(kind, om)
    lat.upload_mask(pack)
    lat.flow_init(RHO_IN, RHO_OUT)
    e = lat.flow_steps(301, energy=True)
    lat.close()
    _, eo = oracle_flow[301]
    assert np.all(np.abs(e - eo) <= 1e-9 * eo + 1e-30)

