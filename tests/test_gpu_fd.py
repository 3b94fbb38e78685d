"""The FD biofilm step (fdDiffusion3D between two mask / total updates, src/complab.cpp:1716-1731)
on the GPU against the oracle (oracle/ca.cpp), through the C ABI.

Both sides evaluate the same expression with no FMA contraction, so with FP64 biomass the
fields, masks and reports agree bit for bit over several calls; with FP32 biomass the fields
agree to FP32 rounding and the line fixtures of tests/test_oracle_fd.py hold."""
import numpy as np
import pytest

import oracle
from test_gpu_ca import gpu_lattice
from test_oracle_ca import random_case

NX, NY, NZ = 40, 20, 16


@pytest.mark.gpu
def test_fd_bit_exact_vs_oracle(clb):
    print(f"case: FD biomass FP64 vs oracle, {NX}x{NY}x{NZ}, 6 calls", flush=True)
    mask, b = random_case(8, NX, NY, NZ)
    b = b * 0.1  # under the CA's cap: only diffusion and the mask rule act
    lat = gpu_lattice(clb, NX, NY, NZ, mask, b)
    ref = oracle.CaState(NX, NY, NZ, b, mask)
    flips = 0
    for _ in range(6):
        rep = lat.fd_biomass(0.12)
        fl, mx0, mx1 = ref.fd_biomass(0.12)
        assert (rep["sweeps"], rep["mask_changes"], rep["status"]) == (1, fl, 0)
        assert rep["bmax_before"] == mx0 and rep["bmax_after"] == mx1
        assert np.array_equal(lat.biomass(), ref.b)
        assert np.array_equal(lat.download_mask(), ref.mask)
        flips += fl
    assert flips > 0  # the mask rule acted
    lat.close()


@pytest.mark.gpu
def test_fd_fp32_and_line_fixtures(clb):
    print("case: FD biomass FP32 vs oracle; closed-form lines on the GPU", flush=True)
    mask, b = random_case(9, NX, NY, NZ)
    b = b * 0.1
    lat = gpu_lattice(clb, NX, NY, NZ, mask, b.astype(np.float32).astype(np.float64), prec=1)
    ref = oracle.CaState(NX, NY, NZ, b.astype(np.float32).astype(np.float64), mask)
    for _ in range(3):
        lat.fd_biomass(0.12)
        ref.fd_biomass(0.12)
    g = lat.biomass()
    assert np.max(np.abs(g - ref.b)) <= 5e-7 * np.max(np.abs(ref.b))
    lat.close()
    # one wall above mirrors; walls on both sides read the upper wall
    lat = gpu_lattice(clb, 5, 1, 1, np.array([2, 2, 2, 0, 2], np.uint8), np.array([0, 3.0, 5.0, 0, 0]), bthr=1e9)
    lat.fd_biomass(0.2)
    assert lat.biomass()[0, 2] == pytest.approx(5.0 + 0.2 * (3.0 - 5.0), rel=1e-15)
    lat.close()
    lat = gpu_lattice(clb, 5, 1, 1, np.array([2, 0, 2, 1, 2], np.uint8), np.array([0, 0, 4.0, 0, 0]), bthr=1e9)
    lat.fd_biomass(0.2)
    assert lat.biomass()[0, 2] == pytest.approx(4.0 * 0.8, rel=1e-15)
    lat.close()


@pytest.mark.gpu
def test_fd_checks(clb):
    print("case: FD biomass argument and state checks", flush=True)
    lat = clb.Lattice(8, 4, 4, precision=0)
    with pytest.raises(clb.ClbError):
        lat.fd_biomass(0.1)  # no clb_ca_init, no biomass
    lat.close()
    mask, b = random_case(3, 12, 6, 5)
    lat = gpu_lattice(clb, 12, 6, 5, mask, b)
    with pytest.raises(clb.ClbError):
        lat.fd_biomass(-0.1)
    lat.close()
