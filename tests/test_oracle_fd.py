"""CPU pins of the oracle's finite-difference biofilm step (oracle/ca.cpp, fd_step and
orc_fd_biomass): fdDiffusion3D (src/complab3d_processors_part2.hh:257-430) between two mask /
total updates, as the loop runs it (src/complab.cpp:1716-1731).

The reference ships no golden vectors for it. The fixtures pin the rule by what it fixes:
  * a single spike: the centre loses 6 nu b, each face neighbour gains nu b, mass is kept;
  * the discrete cosine modes of the zero-flux line decay by exactly 1 - 2 nu (1 - cos(pi k / L))
    per step, which pins nu, the Laplacian and the mirror treatment at both ends;
  * a wall on one side mirrors; walls on both sides of an axis read the upper wall (the
    reference's branch order), so such a voxel loses nu b there;
  * values below 1e-14 count as zero, and a result at or below 1e-14 is not written;
  * biomass diffused over the threshold turns pore into biofilm (the mask update)."""
import numpy as np
import pytest

import oracle


def line(b, masks=None, nu=0.1, steps=1, bthr=1e9):
    b = np.atleast_2d(np.asarray(b, dtype=np.float64))
    nx = b.shape[1]
    m = np.full(nx, 2, dtype=np.uint8) if masks is None else np.asarray(masks, dtype=np.uint8)
    st = oracle.CaState(nx, 1, 1, b, m, bthr=bthr)
    out = [st.fd_biomass(nu) for _ in range(steps)]
    return st, out


def test_spike_closed_form_and_mass():
    nx, ny, nz = 9, 7, 7
    N = nx * ny * nz
    b = np.zeros((1, N))
    c = 4 + nx * (3 + ny * 3)
    b[0, c] = 10.0
    st = oracle.CaState(nx, ny, nz, b, np.full(N, 2, np.uint8), bthr=1e9)
    flips, mx0, mx1 = st.fd_biomass(0.1)
    assert st.b[0, c] == pytest.approx(10.0 - 6 * 0.1 * 10.0, rel=1e-15)
    for d in (1, -1, nx, -nx, nx * ny, -nx * ny):
        assert st.b[0, c + d] == pytest.approx(0.1 * 10.0, rel=1e-15)
    assert st.b.sum() == pytest.approx(10.0, rel=1e-15)
    assert (flips, mx0, mx1) == (0, 10.0, 4.0)


@pytest.mark.parametrize("k", [1, 2, 5])
def test_cosine_modes_of_the_zero_flux_line(k):
    # x = 1 .. L are updated; x = 0 and nx-1 are the inlet / outlet planes (never read here)
    L, nu, steps = 16, 0.15, 7
    x = np.arange(1, L + 1)
    mode = np.cos(np.pi * k * (x - 1 + 0.5) / L)
    b = np.zeros(L + 2)
    b[1:-1] = 1.0 + 0.5 * mode
    st, _ = line(b, nu=nu, steps=steps)
    lam = 1.0 - 2.0 * nu * (1.0 - np.cos(np.pi * k / L))
    np.testing.assert_allclose(st.b[0, 1:-1], 1.0 + 0.5 * lam ** steps * mode, rtol=1e-13, atol=1e-14)
    assert st.b[0, 1:-1].sum() == pytest.approx(L, rel=1e-14)


def test_walls():
    nu = 0.2
    # one wall above: the +x neighbour is the voxel itself
    st, _ = line([0, 3.0, 5.0, 0, 0], masks=[2, 2, 2, 0, 2], nu=nu)
    assert st.b[0, 2] == pytest.approx(5.0 + nu * (3.0 - 5.0), rel=1e-15)
    assert st.b[0, 1] == pytest.approx(3.0 + nu * (5.0 - 3.0), rel=1e-15)
    # walls on both sides: the upper wall (holding nothing) is read, so nu b is lost
    st, _ = line([0, 0, 4.0, 0, 0], masks=[2, 0, 2, 1, 2], nu=nu)
    assert st.b[0, 2] == pytest.approx(4.0 * (1.0 - nu), rel=1e-15)
    # a wall voxel's own value is never updated
    st, _ = line([0, 7.0, 4.0, 0, 0], masks=[2, 1, 2, 2, 2], nu=nu)
    assert st.b[0, 1] == 7.0


def test_thresholds():
    # a result at or below 1e-14 is not written: the voxel keeps its value
    st, _ = line([0, 0, 1e-13, 0, 0], nu=0.5)
    assert st.b[0, 2] == 1e-13
    # neighbours below 1e-14 count as zero
    st, _ = line([0, 5e-15, 1.0, 5e-15, 0], nu=0.1)
    assert st.b[0, 2] == pytest.approx(1.0 + 0.1 * (0 - 2 + 0), rel=1e-15)


def test_mask_update_around_the_step():
    # threshold 10: a pore voxel at 8 next to biofilm at 30 receives 0.2 * (30 - 16) = 2.8 and
    # turns biofilm; a lone biofilm voxel at 11 loses 0.2 * 11 and turns pore
    st, out = line([0, 30.0, 8.0, 0, 0, 0, 11.0, 0], masks=[2, 3, 2, 2, 2, 2, 3, 2], nu=0.2, bthr=10.0)
    flips, mx0, mx1 = out[0]
    assert st.b[0, 2] == pytest.approx(10.8, rel=1e-15)
    assert st.b[0, 6] == pytest.approx(11.0 - 0.2 * 11.0, rel=1e-15)
    assert st.mask.tolist() == [2, 3, 3, 2, 2, 2, 2, 2]
    assert (flips, mx0) == (2, 30.0)
    assert mx1 == pytest.approx(30.0 + 0.2 * (8.0 - 60.0 + 30.0), rel=1e-15)
