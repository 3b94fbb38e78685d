// clb_flow_node.h -- the D3Q19 incompressible BGK equilibrium and the regularized pressure
// faces, shared by the dense (clb_flow.hip) and compact (clb_flow_compact.hip) flow kernels.
#pragma once

#include "clb_lattice.h"

namespace clb {

namespace {

// f_eq,i - t_i (1 + r) = t_i [(rho - 1 - r) + 3 c.j + 4.5 (c.j)^2 - 1.5 j^2]   (rho0 = 1), for
// any reference r: called with rhobar = rho - 1 - rref(x)
template <typename R>
__device__ __forceinline__ R feq19(int i, R rhobar, R jx, R jy, R jz, R jsq) {
  const R cj = R(Q19::cx[i]) * jx + R(Q19::cy[i]) * jy + R(Q19::cz[i]) * jz;
  return R(Q19::w(i)) * (rhobar + R(3) * cj + R(4.5) * cj * cj - R(1.5) * jsq);
}

// feq19 with j = (jx, 0, 0), its multiply-adds left unfused (see regularize).
template <typename R>
__device__ __forceinline__ R feq19_x(int i, R rhobar, R jx, R jsq) {
#pragma clang fp contract(off)
  const R cj = R(Q19::cx[i]) * jx;
  return R(Q19::w(i)) * (rhobar + R(3) * cj + R(4.5) * cj * cj - R(1.5) * jsq);
}

// Regularized pressure boundary on an x face (side -1: x = 0, side +1: x = nx-1).
// j_x from the mass relation over the known populations, tangential momentum zero,
// Pi_neq from on-wall (once) and normal (twice) populations, then all 19 populations
// rebuilt as f_eq + t_i * 9/2 * Q_i:Pi_neq. Multiply-adds are not fused here: the
// compiler fused them differently in the dense and compact flow kernels, whose face nodes
// then differed in the last bit.
template <typename R, int SIDE>
__device__ __forceinline__ void regularize(R* p, R rhobar) {
#pragma clang fp contract(off)
  R onwall = 0, normal = 0;
#pragma unroll
  for (int i = 0; i < 19; ++i) {
    if (Q19::cx[i] == 0) onwall += p[i];
    else if (Q19::cx[i] == SIDE) normal += p[i];
  }
  const R jx = (SIDE < 0) ? (rhobar - onwall - R(2) * normal) : -(rhobar - onwall - R(2) * normal);
  const R jsq = jx * jx;
  R pxx = 0, pxy = 0, pxz = 0, pyy = 0, pyz = 0, pzz = 0;
#pragma unroll
  for (int i = 0; i < 19; ++i) {
    if (Q19::cx[i] != 0 && Q19::cx[i] != SIDE) continue;
    const R wgt = (Q19::cx[i] == 0) ? R(1) : R(2);
    const R fneq = wgt * (p[i] - feq19_x<R>(i, rhobar, jx, jsq));
    const R cx = Q19::cx[i], cy = Q19::cy[i], cz = Q19::cz[i];
    pxx += cx * cx * fneq; pxy += cx * cy * fneq; pxz += cx * cz * fneq;
    pyy += cy * cy * fneq; pyz += cy * cz * fneq; pzz += cz * cz * fneq;
  }
#pragma unroll
  for (int i = 0; i < 19; ++i) {
    const R cx = Q19::cx[i], cy = Q19::cy[i], cz = Q19::cz[i];
    const R third = R(1) / R(3);
    const R qpi = (cx * cx - third) * pxx + (cy * cy - third) * pyy + (cz * cz - third) * pzz +
                  R(2) * (cx * cy * pxy + cx * cz * pxz + cy * cz * pyz);
    p[i] = feq19_x<R>(i, rhobar, jx, jsq) + R(Q19::w(i) * 4.5) * qpi;
  }
}

// BGK collision of one node's populations (relative to its reference density) toward feq19 at
// relaxation om: p_i + om (feq_i - p_i), written with an explicit fused multiply-add so that
// every kernel that collides a node (dense and compact layouts) rounds it the same way.
// Returns |j|^2.
template <typename R>
__device__ __forceinline__ R collide19(R* p, R om) {
  R rhobar = 0, jx = 0, jy = 0, jz = 0;
#pragma unroll
  for (int i = 0; i < 19; ++i) {
    rhobar += p[i];
    jx += R(Q19::cx[i]) * p[i];
    jy += R(Q19::cy[i]) * p[i];
    jz += R(Q19::cz[i]) * p[i];
  }
  const R jsq = jx * jx + jy * jy + jz * jz;
#pragma unroll
  for (int i = 0; i < 19; ++i) p[i] = fma(om, feq19<R>(i, rhobar, jx, jy, jz, jsq) - p[i], p[i]);
  return jsq;
}

}  // namespace
}  // namespace clb
