"""Pins the oracle's D3Q19 restatement (Palabos IncBGKdynamics + BounceBack + regularized
pressure faces, src/complab_functions.hh:245-294) without Palabos: closed-form channel
flow, exact conservation in a closed box, rest-state preservation and mirror symmetry."""
import numpy as np

from conftest import d3q19_kind_omega


def channel_mask(nx, ny, nz):
    """Walls (bounce-back) at y = 0 and y = ny-1, periodic z, open x faces."""
    m = np.full((nz, ny, nx), 2, dtype=np.uint8)
    m[:, 0, :] = 1
    m[:, ny - 1, :] = 1
    return m.ravel()


def test_plane_poiseuille_closed_form(oracle):
    # Pressure-driven flow between bounce-back walls. Full-way bounce-back puts the wall
    # half-way between the wall node and the first fluid node, so the fluid width is
    # H = ny - 2 and u(y) = G/(2 nu) (y - 0.5)(ny - 1.5 - y) with G = c_s^2 drho/(nx-1).
    nx, ny, nz, tau = 24, 12, 2, 0.8
    rho_in, rho_out = 1.0, 0.999
    kind, om = d3q19_kind_omega(tau)
    m = channel_mask(nx, ny, nz)
    f, e = oracle.flow_run(nx, ny, nz, m, kind, om, rho_in, rho_out, 6000, energy=True)
    u, rho = oracle.flow_velocity(nx, ny, nz, m, kind, rho_in, rho_out, f)
    u = u.reshape(nz, ny, nx, 3)
    nu = (tau - 0.5) / 3.0
    G = (rho_in - rho_out) / 3.0 / (nx - 1)
    y = np.arange(1, ny - 1)
    exact = G / (2 * nu) * (y - 0.5) * (ny - 1.5 - y)
    prof = u[0, 1:ny - 1, nx // 2, 0]
    assert np.max(np.abs(prof - exact)) / exact.max() < 2e-2
    # converged: energy no longer changing
    assert abs(e[-1] - e[-100]) / e[-1] < 1e-8
    # no flow across z; across y only the small compressibility correction near the faces
    assert np.max(np.abs(u[..., 2])) < 1e-14
    assert np.max(np.abs(u[..., 1])) < 1e-3 * exact.max()
    # the density profile is linear between the imposed face densities
    r = rho.reshape(nz, ny, nx)[0, ny // 2, :]
    assert np.max(np.abs(r - np.linspace(rho_in, rho_out, nx))) < 5e-5


def test_closed_box_conserves_mass_exactly(oracle):
    # All x faces solid: no open boundary; populations only move and reflect, so the sum
    # over every node is invariant (to rounding) however the interior evolves.
    nx, ny, nz = 10, 8, 6
    rng = np.random.default_rng(0)
    m = np.full((nz, ny, nx), 2, dtype=np.uint8)
    m[:, :, 0] = 1
    m[:, :, nx - 1] = 1
    m[(rng.random((nz, ny, nx)) < 0.15)] = 1
    m = m.ravel()
    kind, om = d3q19_kind_omega(0.7)
    N = nx * ny * nz
    f = 0.01 * rng.standard_normal(19 * N)
    f0 = f.sum()
    oracle.flow_run(nx, ny, nz, m, kind, om, 1.0, 1.0, 300, f=f)
    assert abs(f.sum() - f0) < 1e-12 * N


def test_rest_state_is_preserved(oracle):
    nx, ny, nz = 12, 6, 6
    m = channel_mask(nx, ny, nz)
    kind, om = d3q19_kind_omega(0.8)
    f, _ = oracle.flow_run(nx, ny, nz, m, kind, om, 1.0, 1.0, 50)
    u, rho = oracle.flow_velocity(nx, ny, nz, m, kind, 1.0, 1.0, f)
    assert np.max(np.abs(u)) < 1e-15
    assert np.max(np.abs(rho[m == 2] - 1.0)) < 1e-14


def test_mirror_symmetry(oracle):
    # A geometry symmetric under y -> ny-1-y gives u_x symmetric and u_y antisymmetric.
    nx, ny, nz = 20, 11, 4
    rng = np.random.default_rng(1)
    half = (rng.random((nz, ny // 2, nx)) < 0.2)
    solid = np.zeros((nz, ny, nx), dtype=bool)
    solid[:, : ny // 2] = half
    solid[:, ny - ny // 2:] = half[:, ::-1]
    solid[:, :, 0] = solid[:, :, 1]
    solid[:, :, nx - 1] = solid[:, :, nx - 2]
    m = np.where(solid, 1, 2).astype(np.uint8).ravel()
    kind, om = d3q19_kind_omega(0.9)
    f, _ = oracle.flow_run(nx, ny, nz, m, kind, om, 1.0, 0.9995, 400)
    u, _ = oracle.flow_velocity(nx, ny, nz, m, kind, 1.0, 0.9995, f)
    u = u.reshape(nz, ny, nx, 3)
    scale = np.max(np.abs(u[..., 0]))
    assert scale > 0
    assert np.max(np.abs(u[..., 0] - u[:, ::-1, :, 0])) < 1e-12 * scale
    assert np.max(np.abs(u[..., 1] + u[:, ::-1, :, 1])) < 1e-12 * scale
