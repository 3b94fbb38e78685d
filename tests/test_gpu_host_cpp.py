"""The C++ host interface (host/complab_gpu.hh) against the FP64 oracle.

tests/cpp/host_mirror_check drives complab_gpu::EquilibriumChemistry<double> (the surface of
the reference's EquilibriumChemistry<T>, src/complab3d_processors_part4_eqsolver.hh:187-1320)
and complab_gpu::Lattice the way the reference's C++ driver would; this test feeds it
inputs and checks what it wrote: carbonate speciation against the oracle's PCF solve,
flow velocity and coupled kinetics transport against the oracle's restatement."""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT, d3q19_kind_omega, rel_l2
from test_oracle_speciation import CARB_COMP, CARB_LOGK, CARB_S

EXE = os.path.join(ROOT, "tests", "cpp", "host_mirror_check")


@pytest.mark.gpu
def test_cpp_host_interface(clb, oracle, tmp_path):
    print("case: C++ host interface (EquilibriumChemistry + Lattice)", flush=True)
    if not os.path.exists(EXE):
        raise RuntimeError(f"{EXE} missing: run __graft_entry__.build()")
    rng = np.random.default_rng(21)
    conc = np.column_stack([10 ** rng.uniform(-4, -2.5, 300), 10 ** rng.uniform(-7, -4, 300),
                            10 ** rng.uniform(-9, -7, 300), 10 ** rng.uniform(-5, -3, 300)])
    conc.astype(np.float64).tofile(tmp_path / "eq_in.bin")
    r = subprocess.run([EXE, str(tmp_path)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    tag, conv, iters, fluid, energy = r.stdout.split()
    assert tag == "ok" and conv == "1" and int(iters) > 0
    # speciation: the batched and single-voxel calls against the reference's PCF answer
    got = np.fromfile(tmp_path / "eq_out.bin").reshape(-1, 4)
    one = np.fromfile(tmp_path / "eq_one.bin")
    ref, _, _, cv, _ = oracle.speciate_pcf(CARB_S, CARB_LOGK, CARB_COMP, conc, tol=1e-12)
    assert cv.all()
    assert np.max(np.abs(got - ref) / ref) < 1e-8
    assert np.max(np.abs(one - ref[0]) / ref[0]) < 1e-8
    # lattice: the same case through the oracle
    nx, ny, nz = 32, 16, 16
    pack = clb.gen_sphere_pack(nx, ny, nz, 3.0, 5.0, 0.45, seed=0)
    kind, om = d3q19_kind_omega(0.8)
    f, e = oracle.flow_run(nx, ny, nz, pack, kind, om, 1.0, 0.999, 200, energy=True)
    u_ref, _ = oracle.flow_velocity(nx, ny, nz, pack, kind, 1.0, 0.999, f)
    u = np.fromfile(tmp_path / "u.bin").reshape(-1, 3)
    assert rel_l2(u, u_ref) < 1e-10
    assert abs(float(energy) - e[-1]) <= 1e-6 * abs(e[-1])
    fl = (pack.reshape(nz, ny, nx) == 2)
    fl[:, :, [0, nx - 1]] = False
    assert int(fluid) == fl.sum()
    b = clb.gen_biomass_seed(pack, pore=2, fraction=0.02, bmin=10, bmax=100, seed_sites=1, seed_density=2)
    run = oracle.AdeRun(nx, ny, nz, pack, kind, np.full((4, 256), 1.0 / 0.8), [0, 1, 0, 1, 1, 1, 1, 1],
                        [1.0, 0, 0.5, 0, 0, 0, 0, 0], u_ref, [0.2, 0.1, 0.0, 0.0], kinetics=True, biomass=b)
    run.run(30)
    c_ref = run.conc()
    c = np.fromfile(tmp_path / "conc.bin").reshape(4, -1)
    for s in range(4):
        assert rel_l2(c[s], c_ref[s]) < 1e-9
    assert rel_l2(np.fromfile(tmp_path / "bmass.bin"), run.bmass) < 1e-9
