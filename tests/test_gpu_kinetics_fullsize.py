"""The kinetics + biomass workload at its full size (256 x 128 x 128 sphere pack, porosity
0.45, 4 solutes on D3Q7 at tau 0.8, dual-Monod kinetics with Ks 1e-2, kmax 1e-3 per step,
yield 0.4, biomass seeded in 1 % of the pore voxels at densities in [10, 100], the pore <->
biofilm mask rule at 10): the FP32 deviation store against the FP64 store, within the
stated bound of 1e-5 relative L2 on the velocity, the concentrations and the biomass, with
the masks equal.

The step kernels are those of the reference's loop (src/complab.cpp:1252-1849: collide,
run_kinetics, update_rxnLattices, updateLocalMaskNtotalLattices3D, stream); the FP64 store
is the one the oracle matches at 1e-10 on smaller lattices (test_gpu_transport.py,
test_gpu_reactions.py)."""
import numpy as np
import pytest

from conftest import d3q19_kind_omega, rel_l2

NX, NY, NZ = 256, 128, 128
FLOW_STEPS = 2000
STEPS = 1000


def run(clb, prec):
    pack = clb.gen_sphere_pack(NX, NY, NZ, 3.0, 5.0, 0.45, seed=0)
    b = clb.gen_biomass_seed(pack, pore=2, fraction=0.01, bmin=10.0, bmax=100.0, seed_sites=1, seed_density=2)
    mask = pack.copy()
    mask[b >= 10.0] = 3
    kind, om = d3q19_kind_omega(0.8)
    lat = clb.Lattice(NX, NY, NZ, precision=prec)
    lat.set_dynamics(kind, om)
    lat.upload_mask(mask)
    lat.flow_init(1.0, 0.999)
    lat.flow_steps(FLOW_STEPS)
    u, _ = lat.velocity()
    lat.set_solutes(np.full((4, 256), 1.0 / 0.8), [0, 1, 0, 1, 1, 1, 1, 1], [1.0, 0, 0.5, 0, 0, 0, 0, 0])
    lat.couple_velocity()
    lat.set_biomass(b)
    lat.set_kinetics(clb.KINETICS_PLUGIN, 1.0, 0, 1)
    lat.set_mask_update(True, 10.0, 2, 3)
    lat.solutes_init([0.2, 0.1, 0.0, 0.0])
    lat.set_graph(True, 100)
    lat.step(STEPS)
    out = (u, lat.conc(), lat.biomass(), lat.download_mask(), b)
    lat.close()
    return out


@pytest.mark.gpu
def test_fp32_store_against_fp64_at_full_size(clb):
    print(f"case: kinetics {NX}x{NY}x{NZ}, {FLOW_STEPS} flow + {STEPS} coupled steps, FP64 then FP32", flush=True)
    u64, c64, b64, m64, b0 = run(clb, clb.PREC_FP64)
    u32, c32, b32, m32, _ = run(clb, clb.PREC_FP32)
    assert rel_l2(u32, u64) < 1e-5
    for s in range(4):
        assert rel_l2(c32[s], c64[s]) < 1e-5, s
    assert rel_l2(b32, b64) < 1e-5
    assert np.array_equal(m32, m64)
    # the workload did what it is for: the inlet solutes spread, the products formed and the
    # biomass grew
    assert c64[2].max() > 1e-3 and c64[3].max() > 1e-3
    assert b64.sum() > b0.sum() * (1 + 1e-4)
