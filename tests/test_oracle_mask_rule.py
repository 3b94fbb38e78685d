"""Pins the oracle's pore <-> biofilm rule with several material types
(updateLocalMaskNtotalLattices3D, src/complab3d_processors_part2.hh:117-198) on a hand-built
mask, in both places the oracle applies it: the CA's mask / total pass (ca.cpp) and the
step (oracle.cpp, after the voxel loop with several pore types, inside it with one).

Pore types 2 and 5 (pore_dynamics), biomass fields with biofilm codes 3 and 4
(bio_dynamics[i][0]), threshold 10:
  * a pore voxel at or above the threshold takes the sum of the codes of the fields present
    (both: 7; only the second: 4), whatever its pore type;
  * a voxel of no pore type below it takes the type of its first face neighbour of a pore
    type in the order +x -x +y -y +z -z, skipping walls and directions off the domain (no
    wrap), else pore type 2;
  * every voxel reads its neighbours' types from the mask before the pass: a neighbour that
    turns pore in the same pass is not a pore neighbour yet."""
import numpy as np

from conftest import d3q19_kind_omega

NX, NY, NZ = 7, 4, 3
THR = 10.0


def idx(x, y, z):
    return x + NX * (y + NY * z)


def fixture():
    m = np.full((NZ, NY, NX), 2, dtype=np.uint8)
    m[:, :, 0] = 1
    m[:, :, NX - 1] = 1
    m = m.ravel()
    b = np.zeros((2, m.size))
    expect = m.copy()
    # A: pore type 2, both fields present, at the threshold -> 3 + 4
    A = idx(2, 1, 1); b[0, A], b[1, A] = 8.0, 5.0; expect[A] = 7
    # B: pore type 5, only the second field -> 4
    B = idx(4, 1, 1); m[B] = 5; b[1, B] = 12.0; expect[B] = 4
    # C: biofilm 7 below the threshold; +x is a wall, -x is pore type 5 -> 5
    C = idx(2, 2, 1); m[C] = 7; b[0, C] = 1.0; expect[C] = 5
    m[idx(3, 2, 1)] = 1; expect[idx(3, 2, 1)] = 1
    m[idx(1, 2, 1)] = 5; expect[idx(1, 2, 1)] = 5
    # D: biofilm 3, no pore neighbour (+y and +z are off the domain) -> pore type 2
    D = idx(4, 3, 2); m[D] = 3; expect[D] = 2
    for q in (idx(5, 3, 2), idx(3, 3, 2), idx(4, 2, 2), idx(4, 3, 1)):
        m[q] = 1; expect[q] = 1
    # E: another code (9) below the threshold, at y = 0, z = 0; +x is pore type 5 -> 5
    E = idx(3, 0, 0); m[E] = 9; expect[E] = 5
    m[idx(4, 0, 0)] = 5; expect[idx(4, 0, 0)] = 5
    # F, G: two biofilm voxels side by side, both empty. G (+x of F) is biofilm before the
    # pass, so F takes its +y neighbour's type 2; G takes its +x neighbour's 5
    F, G = idx(1, 0, 2), idx(2, 0, 2)
    m[F], m[G] = 4, 3
    m[idx(3, 0, 2)] = 5; expect[idx(3, 0, 2)] = 5
    m[idx(1, 1, 2)] = 2
    expect[F], expect[G] = 2, 5
    return m, b, expect


def test_ca_mask_pass_multi_type(oracle):
    m, b, expect = fixture()
    rule = oracle.mask_rule([2, 5], [3, 4])
    ca = oracle.CaState(NX, NY, NZ, b, m, solid=0, bb=1, pore=2, biofilm=3, bthr=THR, rule=rule)
    ca.redistribute(bmax=1e9)          # nothing above the cap: one mask / total pass
    assert np.array_equal(ca.mask, expect), np.flatnonzero(ca.mask != expect)


def test_step_mask_rule_multi_type(oracle):
    m, b, expect = fixture()
    kind, _ = d3q19_kind_omega(codes_fluid=(2, 3, 4, 5, 7, 9))
    run = oracle.AdeRun(NX, NY, NZ, m, kind, np.full((1, 256), 1.0 / 0.8), [1, 1], [0.0, 0.0],
                        np.zeros((m.size, 3)), [0.0], biomass=b.ravel(), mask_update=(THR, 2, 3),
                        rule=oracle.mask_rule([2, 5], [3, 4]))
    run.run(1)
    assert np.array_equal(run.mask, expect), np.flatnonzero(run.mask != expect)


def test_one_pore_type_sum_labels(oracle):
    # one pore type (2): type 5 is then not a pore type. A -> 7 as before; B (5, above the
    # threshold) is no pore voxel and keeps its code; every voxel of no pore type below the
    # threshold (C, D, E, F, G and the type-5 voxels) returns to 2
    m, b, _ = fixture()
    kind, _ = d3q19_kind_omega(codes_fluid=(2, 3, 4, 5, 7, 9))
    rule = oracle.mask_rule([2], [3, 4])
    exp = m.copy()
    inner = (np.arange(m.size) % NX >= 1) & (np.arange(m.size) % NX <= NX - 2)
    tot = np.where(b > 1e-14, b, 0).sum(axis=0)
    exp[inner & (m != 1) & (m != 2) & (tot < THR)] = 2
    exp[idx(2, 1, 1)] = 7
    ca = oracle.CaState(NX, NY, NZ, b, m, bthr=THR, rule=rule)
    ca.redistribute(bmax=1e9)
    assert np.array_equal(ca.mask, exp)
    run = oracle.AdeRun(NX, NY, NZ, m, kind, np.full((1, 256), 1.0 / 0.8), [1, 1], [0.0, 0.0],
                        np.zeros((m.size, 3)), [0.0], biomass=b.ravel(), mask_update=(THR, 2, 3), rule=rule)
    run.run(1)
    assert np.array_equal(run.mask, exp)
