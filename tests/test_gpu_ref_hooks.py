"""The reference's own example kinetics hooks on the GPU: the library built with
examples/05_biotic_cellular_automaton/defineKinetics.hh and
examples/03_abiotic_kinetics/defineAbioticKinetics.hh compiled in, unchanged, through the
plugin adapter (oracle/_ref/libclb_ref_hooks.so, `make -C oracle ref-hooks`), against the
oracle built from the same two headers (oracle/_ref/liboracle_ref_hooks.so).

Case: the 32 x 16 x 16 sphere pack (porosity 0.5), three solutes (donor, product of the Monod
growth and second reactant of A + B -> C, and C), biomass seeds on 5 % of the pore voxels,
transport with Dirichlet inlets, 100 coupled steps of run_kinetics + run_abiotic_kinetics
(src/complab3d_processors_part1.hh:72-336). Skipped where the reference tree was not there to
build from."""
import os

import numpy as np
import pytest

import oracle as orc
from conftest import d3q19_kind_omega, rel_l2

REF_LIB = os.path.join(os.path.dirname(orc.REF_HOOKS_PATH), "libclb_ref_hooks.so")
pytestmark = pytest.mark.skipif(not (os.path.exists(orc.REF_HOOKS_PATH) and os.path.exists(REF_LIB)),
                                reason="built only where the reference tree is present (make -C oracle ref-hooks)")

NX, NY, NZ = 32, 16, 16
BCT = [0, 1, 0, 1, 0, 1]
BCV = [0.05, 0, 0.01, 0, 0.0, 0]
C0 = [0.02, 0.01, 0.0]
STEPS = 100


@pytest.fixture(scope="module")
def case(clb):
    pack = clb.gen_sphere_pack(NX, NY, NZ, 3.0, 5.0, 0.5, seed=0)
    kind, om = d3q19_kind_omega(0.8)
    f, _ = orc.flow_run(NX, NY, NZ, pack, kind, om, 1.0, 0.999, 300)
    u, _ = orc.flow_velocity(NX, NY, NZ, pack, kind, 1.0, 0.999, f)
    b = clb.gen_biomass_seed(pack, pore=2, fraction=0.05, bmin=0.5, bmax=2.0, seed_sites=1, seed_density=2)
    run = orc.AdeRun(NX, NY, NZ, pack, kind, np.full((3, 256), 1.0 / 0.8), BCT, BCV, u, C0, kinetics=True,
                     biomass=b, abiotic=1, library=orc.REF_HOOKS_PATH)
    run.run(STEPS)
    return pack, kind, om, u, b, run.conc(), run.bmass.copy()


@pytest.mark.gpu
@pytest.mark.parametrize("layout", [0, 1])
@pytest.mark.parametrize("prec,tol", [(0, 1e-10), (1, 1e-5)])
def test_reference_example_hooks_match_oracle(clb, case, prec, tol, layout):
    print(f"case: reference example hooks 05 + 03, prec={prec} layout={layout}", flush=True)
    pack, kind, om, u, b, ref_c, ref_b = case
    lat = clb.Lattice(NX, NY, NZ, precision=prec, library=REF_LIB)
    lat.set_layout(layout)
    lat.set_dynamics(kind, om)
    lat.upload_mask(pack)
    lat.set_solutes(np.full((3, 256), 1.0 / 0.8), BCT, BCV)
    lat.set_velocity(u)
    lat.set_biomass(b)
    lat.set_kinetics(clb.KINETICS_PLUGIN, 1.0, 0, 1)
    lat.set_abiotic(1)
    lat.solutes_init(C0)
    lat.step(STEPS)
    c = lat.conc()
    bm = lat.biomass()
    lat.close()
    # the reactions ran: the product of the Monod growth and of A + B -> C formed, biomass changed
    assert ref_c[2].max() > 1e-5
    assert np.abs(ref_b - b).max() > 1e-4
    for s in range(3):
        assert rel_l2(c[s], ref_c[s]) < tol, s
    assert rel_l2(bm.ravel(), ref_b) < tol
