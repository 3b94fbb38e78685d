"""CPU pins of the oracle's CA biomass redistribution (oracle/ca.cpp).

The reference's push/pull CA (src/complab3d_processors_part1.hh:432-894, sweep loop
src/complab.cpp:1680-1731) has no golden vectors: its sweeps depend on loop order and on
std::rand. These fixtures are built so that the outcome is fixed by the rule itself:
a single open route, capacities that every neighbour order fills the same way, the age rule
as a distance transform, the species split ratio, the stop conditions. A wrong direction,
capacity, split or age condition fails one of them.

Codes: 0 solid, 1 bounce-back, 2 pore, 3 biofilm. Line fixtures are nx x 1 x 1 (the CA does
not wrap in y or z, so only +-x neighbours exist); x = 0 and x = nx-1 are the inlet/outlet
planes, which the CA leaves alone."""
import numpy as np
import pytest

import oracle

BMAX = 100.0


def line(masks, b, bthr=10.0):
    m = np.array(masks, dtype=np.uint8)
    return oracle.CaState(len(m), 1, 1, np.atleast_2d(np.array(b, dtype=np.float64)), m, bthr=bthr)


def test_single_route_takes_the_excess():
    st = line([2, 1, 3, 2, 2, 2], [0, 0, 130, 0, 0, 0])
    sweeps, passes, status, flips, mx0, mx1 = st.redistribute(BMAX, seed=7)
    assert st.b[0].tolist() == [0, 0, 100, 30, 0, 0]
    assert st.mask.tolist() == [2, 1, 3, 3, 2, 2]  # the receiving pore voxel turned biofilm (30 >= 10)
    assert (sweeps, status, flips, mx0, mx1) == (1, 0, 1, 130.0, 100.0)
    assert passes >= 1  # the age relaxation runs after the first sweep


def test_capacity_then_age_rule_then_overflow():
    # x=3 holds 50 over the cap; its neighbours have room 10 and 5. Whatever the order, both
    # fill; the remaining 35 goes to one of them (equal ages: any neighbour), which passes it
    # on to its free pore neighbour in the next sweep.
    for seed in range(1, 9):
        st = line([2, 2, 3, 3, 3, 2, 2], [0, 0, 90, 150, 95, 0, 0])
        sweeps, _, status, _, _, mx1 = st.redistribute(BMAX, seed=seed)
        b = st.b[0]
        assert status == 0 and sweeps == 2 and mx1 == 100.0
        assert b[2:5].tolist() == [100, 100, 100]
        assert sorted([b[1], b[5]]) == [0, 35] and b[0] == 0 and b[6] == 0
        assert b.sum() == 335


def test_age_rule_sends_the_rest_towards_the_surface():
    # x=3 holds 40 over the cap and both neighbours are full: the rest goes to the neighbour
    # of lower age (x=2, towards the pore at x=1), never to x=4, whatever the seed; x=2 then
    # places it in the empty pore voxel x=1 (a pore neighbour counts as empty).
    for seed in range(1, 9):
        st = line([2, 2, 3, 3, 3, 3, 2], [0, 0, 100, 140, 100, 100, 0])
        st.age[:] = [0, 0, 1, 2, 3, 4, 0]
        sweeps, _, status, flips, _, mx1 = st.redistribute(BMAX, seed=seed)
        assert st.b[0].tolist() == [0, 40, 100, 100, 100, 100, 0]
        assert (sweeps, status, flips, mx1) == (2, 0, 1, 100.0)


def test_species_split_in_proportion():
    st = line([2, 1, 3, 2, 2, 2], [[0, 0, 30, 0, 0, 0], [0, 0, 100, 0, 0, 0]])
    st.redistribute(BMAX, seed=3)
    s1, s2 = 30.0 / 130.0 * 30.0, 100.0 / 130.0 * 30.0
    assert st.b[0, 3] == s1 and st.b[1, 3] == s2
    assert st.b[0, 2] == 30.0 - s1 and st.b[1, 2] == 100.0 - s2
    assert st.b[:, 2].sum() == pytest.approx(100.0, abs=1e-12)


def test_half_push_moves_half():
    st = line([2, 1, 3, 2, 2, 2], [0, 0, 120, 0, 0, 0])
    sweeps, _, status, _, _, mx1 = st.redistribute(BMAX, half=1, seed=5)
    assert st.b[0].tolist() == [0, 0, 60, 60, 0, 0]
    assert (sweeps, status, mx1) == (1, 0, 60.0)
    # a pore voxel over the cap does not half-push (halfPushExcessBiomass3D skips pore codes)
    st = line([2, 1, 2, 2, 2, 2], [0, 0, 120, 0, 0, 0], bthr=1e9)
    sweeps, _, status, _, _, _ = st.redistribute(BMAX, half=1, seed=5)
    assert status == 1 and st.b[0, 2] == 120  # stalled: nothing could move it


def test_boxed_in_excess_stops():
    # ping-pong: x=2's rest goes to x=1 (lower age), whose only CA neighbour is x=2 again, so
    # the max never falls: stalled after 100 sweeps (an even number: back where it started)
    st = line([2, 3, 3, 3, 2], [0, 100, 140, 100, 0])
    st.age[:] = [0, 1, 2, 3, 0]
    sweeps, _, status, _, _, mx1 = st.redistribute(BMAX, seed=1)
    assert (status, sweeps, mx1) == (oracle_status("stalled"), 100, 140.0)
    assert st.b[0].tolist() == [0, 100, 140, 100, 0]
    # encased (walls on both sides): nothing moves, and the age of a voxel without
    # neighbours counts up for ever, so the first age relaxation hits its cap, as the
    # reference's does (src/complab.cpp:1705)
    st = line([2, 1, 3, 1, 2], [0, 0, 150, 0, 0])
    sweeps, passes, status, _, _, mx1 = st.redistribute(BMAX, seed=1)
    assert (status, sweeps, passes, mx1) == (oracle_status("age_cap"), 1, 1000, 150.0)
    assert st.b[0].tolist() == [0, 0, 150, 0, 0]


def test_nonfinite_biomass_stops():
    # an infinite total stops the loop before any sweep (src/complab.cpp:1456); a NaN
    # species value fails the > 1e-14 test and stays out of the total, as in the reference
    st = line([2, 2, 3, 2, 2], [0, 0, np.inf, 0, 0])
    sweeps, _, status, _, _, _ = st.redistribute(BMAX, seed=1)
    assert status == oracle_status("nonfinite") and sweeps == 0
    st = line([2, 2, 3, 2, 2], [0, 0, np.nan, 0, 0])
    sweeps, _, status, _, mx0, _ = st.redistribute(BMAX, seed=1)
    assert (status, sweeps, mx0) == (0, 0, 0.0)


def oracle_status(name):
    return {"done": 0, "stalled": 1, "sweep_cap": 2, "age_cap": 3, "nonfinite": 4}[name]


def bfs_distance(region, nx, ny, nz):
    """Face-step distance of every voxel to the nearest voxel outside `region` (no wrap)."""
    INF = 10 ** 9
    d = np.full(region.shape, INF, dtype=np.int64)
    d[~region] = 0
    frontier = list(zip(*np.nonzero(~region)))
    k = 0
    while frontier:
        k += 1
        nxt = []
        for (z, y, x) in frontier:
            for dz, dy, dx in ((0, 0, 1), (0, 0, -1), (0, 1, 0), (0, -1, 0), (1, 0, 0), (-1, 0, 0)):
                zz, yy, xx = z + dz, y + dy, x + dx
                if 0 <= zz < nz and 0 <= yy < ny and 0 <= xx < nx and d[zz, yy, xx] == INF:
                    d[zz, yy, xx] = k
                    nxt.append((zz, yy, xx))
        frontier = nxt
    return d


def test_age_is_the_distance_to_the_surface():
    nx, ny, nz = 14, 11, 9
    mask = np.full((nz, ny, nx), 2, dtype=np.uint8)
    mask[1:8, 1:10, 2:12] = 3
    mask[4, 5, 6] = 2  # a pore pocket inside the block
    b = np.where(mask == 3, BMAX, 0.0)
    st = oracle.CaState(nx, ny, nz, b.ravel()[None, :], mask.ravel())
    st.redistribute(BMAX)  # nothing over the cap: only fills the total lattice
    passes = st.update_age(BMAX)
    assert passes > 0
    want = np.where(mask == 3, bfs_distance(mask == 3, nx, ny, nz), 0).ravel()
    assert np.array_equal(st.age, want)


def random_case(seed, nx=24, ny=12, nz=10, solid=0.2):
    """Scattered walls (0) in pore (2), a quarter of the pore voxels biofilm (3) holding two
    species, up to ~3 Bmax in total; the inlet / outlet planes copy their neighbours. With
    solid = 0.2 some biofilm voxels are boxed in by walls (the CA stalls on them)."""
    rng = np.random.default_rng(seed)
    mask = np.where(rng.random((nz, ny, nx)) < solid, 0, 2).astype(np.uint8)
    mask[:, :, 0] = mask[:, :, 1]
    mask[:, :, -1] = mask[:, :, -2]
    bio = (mask == 2) & (rng.random(mask.shape) < 0.25)
    mask[bio] = 3
    b = np.zeros((2,) + mask.shape)
    b[0][bio] = rng.uniform(20, 260, bio.sum())
    b[1][bio] = rng.uniform(0, 60, bio.sum())
    return mask.ravel(), b.reshape(2, -1)


@pytest.mark.parametrize("half", [0, 1])
def test_random_pack_invariants(half):
    nx, ny, nz = 24, 12, 10
    mask, b = random_case(11, solid=0.05)
    st = oracle.CaState(nx, ny, nz, b, mask)
    total0 = st.b.sum(axis=1)
    sweeps, _, status, flips, mx0, mx1 = st.redistribute(BMAX, half=half, seed=42)
    assert sweeps > 0 and mx0 > BMAX and status == 0
    assert np.allclose(st.b.sum(axis=1), total0, rtol=1e-13, atol=0)  # mass conserved
    walls = (st.mask == 0) | (st.mask == 1)
    assert not st.b[:, walls].any()
    inner = (np.arange(st.btot.size) % nx >= 1) & (np.arange(st.btot.size) % nx <= nx - 2)
    if status == 0:  # the cap holds on the planes the CA acts on
        assert mx1 <= BMAX + 1e-12
        assert st.btot[inner].max() <= BMAX + 1e-12
    # same seed, same answer; another seed, another (equally valid) answer
    st2 = oracle.CaState(nx, ny, nz, b, mask)
    st2.redistribute(BMAX, half=half, seed=42)
    assert np.array_equal(st.b, st2.b) and np.array_equal(st.mask, st2.mask) and np.array_equal(st.age, st2.age)
    st3 = oracle.CaState(nx, ny, nz, b, mask)
    st3.redistribute(BMAX, half=half, seed=43)
    assert not np.array_equal(st.b, st3.b)
    assert np.allclose(st3.b.sum(axis=1), total0, rtol=1e-13, atol=0)
