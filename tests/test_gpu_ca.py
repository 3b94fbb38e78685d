"""CA biomass redistribution on the GPU against the oracle (oracle/ca.cpp), through the C ABI.

Both sides run the same order-independent sweep (decisions from the sweep's starting state,
neighbour orders from the same counter-based draws) with no FMA contraction, so with FP64
biomass the fields, masks, ages and reports agree bit for bit, including a second call that
continues the sweep count. The closed-form fixtures of tests/test_oracle_ca.py run through
the GPU as well. With FP32 biomass, the invariants: mass conserved and the cap met."""
import numpy as np
import pytest

import oracle
from test_oracle_ca import BMAX, bfs_distance, random_case

NX, NY, NZ = 40, 20, 16


def gpu_lattice(clb, nx, ny, nz, mask, b, prec=0, bthr=10.0):
    lat = clb.Lattice(nx, ny, nz, precision=prec)
    kind = np.full(256, 2, dtype=np.int32)
    kind[0], kind[1] = 0, 1
    lat.set_dynamics(kind, np.full(256, 1.25))
    lat.upload_mask(mask)
    lat.set_biomass(b)
    lat.set_kinetics(clb.KINETICS_NONE, 1.0, 0, 1)
    lat.set_mask_update(True, bthr, 2, 3)
    lat.ca_init()
    return lat


@pytest.mark.gpu
@pytest.mark.parametrize("half", [0, 1])
def test_ca_bit_exact_vs_oracle(clb, half):
    print(f"case: CA FP64 vs oracle, {NX}x{NY}x{NZ}, half={half}", flush=True)
    mask, b = random_case(5, NX, NY, NZ)  # with boxed-in biofilm: both stall the same way
    lat = gpu_lattice(clb, NX, NY, NZ, mask, b)
    ref = oracle.CaState(NX, NY, NZ, b, mask)
    assert np.array_equal(lat.ca_age(), ref.age)
    for call, seed in enumerate((17, 17)):
        rep = lat.ca_redistribute(BMAX, half=half, seed=seed)
        sw, passes, status, flips, mx0, mx1 = ref.redistribute(BMAX, half=half, seed=seed)
        assert (rep["sweeps"], rep["age_passes"], rep["status"], rep["mask_changes"]) == (sw, passes, status, flips)
        assert rep["bmax_before"] == mx0 and rep["bmax_after"] == mx1
        assert np.array_equal(lat.biomass(), ref.b)
        assert np.array_equal(lat.download_mask(), ref.mask)
        assert np.array_equal(lat.ca_age(), ref.age)
        if call == 0:
            assert sw > 1 and flips > 0
            # grow the biomass again for the second call, on both sides
            grown = ref.b * 1.3
            lat.set_biomass(grown)
            ref.b[:] = grown
    lat.close()


@pytest.mark.gpu
def test_ca_closed_forms_on_gpu(clb):
    print("case: CA closed-form line fixtures on the GPU", flush=True)
    # single open route
    lat = gpu_lattice(clb, 6, 1, 1, np.array([2, 1, 3, 2, 2, 2], np.uint8), np.array([0, 0, 130, 0, 0, 0.0]))
    rep = lat.ca_redistribute(BMAX, seed=7)
    assert lat.biomass()[0].tolist() == [0, 0, 100, 30, 0, 0]
    assert lat.download_mask().tolist() == [2, 1, 3, 3, 2, 2]
    assert (rep["sweeps"], rep["status"], rep["mask_changes"]) == (1, 0, 1)
    lat.close()
    # rest towards the lower age, then into the empty pore voxel
    lat = gpu_lattice(clb, 7, 1, 1, np.array([2, 2, 3, 3, 3, 3, 2], np.uint8),
                      np.array([0, 0, 100, 140, 100, 100, 0.0]))
    lat.ca_set_age(np.array([0, 0, 1, 2, 3, 4, 0]))
    rep = lat.ca_redistribute(BMAX, seed=3)
    assert lat.biomass()[0].tolist() == [0, 40, 100, 100, 100, 100, 0]
    assert (rep["sweeps"], rep["status"]) == (2, 0)
    lat.close()
    # encased: the age relaxation hits its cap, nothing moves
    lat = gpu_lattice(clb, 5, 1, 1, np.array([2, 1, 3, 1, 2], np.uint8), np.array([0, 0, 150, 0, 0.0]))
    rep = lat.ca_redistribute(BMAX, seed=1)
    assert (rep["status"], rep["sweeps"], rep["age_passes"]) == (clb.CA_AGE_CAP, 1, 1000)
    lat.close()


@pytest.mark.gpu
def test_ca_age_distance_on_gpu(clb):
    print("case: CA age relaxation = distance to the surface", flush=True)
    nx, ny, nz = 14, 11, 9
    mask = np.full((nz, ny, nx), 2, dtype=np.uint8)
    mask[1:8, 1:10, 2:12] = 3
    mask[4, 5, 6] = 2
    b = np.where(mask == 3, BMAX, 0.0).ravel()
    lat = gpu_lattice(clb, nx, ny, nz, mask.ravel(), b)
    lat.ca_redistribute(BMAX)  # fills the total lattice; nothing over the cap
    assert lat.ca_update_age(BMAX) > 0
    want = np.where(mask == 3, bfs_distance(mask == 3, nx, ny, nz), 0).ravel()
    assert np.array_equal(lat.ca_age(), want)
    lat.close()


@pytest.mark.gpu
def test_ca_fp32_invariants(clb):
    print(f"case: CA FP32 invariants, {NX}x{NY}x{NZ}", flush=True)
    mask, b = random_case(9, NX, NY, NZ, solid=0.05)
    lat = gpu_lattice(clb, NX, NY, NZ, mask, b, prec=1)
    b32 = b.astype(np.float32).astype(np.float64)
    rep = lat.ca_redistribute(BMAX, seed=2)
    out = lat.biomass()
    assert rep["status"] == 0 and rep["sweeps"] > 0
    assert np.allclose(out.sum(axis=1), b32.sum(axis=1), rtol=2e-6)
    tot = out.sum(axis=0)
    inner = (np.arange(tot.size) % NX >= 1) & (np.arange(tot.size) % NX <= NX - 2)
    assert tot[inner].max() <= BMAX * (1 + 8 * np.finfo(np.float32).eps) + 1e-12
    m = lat.download_mask()
    assert not out[:, (m == 0) | (m == 1)].any()
    lat.close()
