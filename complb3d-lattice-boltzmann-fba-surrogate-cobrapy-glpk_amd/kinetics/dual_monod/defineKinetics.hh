/* defineKinetics.hh -- dual-Monod donor / acceptor kinetics with one biomass pool.
 *
 * Written against the reference's compiled-in kinetics interface
 * (defineKinetics.hh:223-228 of CompLB3D): the same free function, the same by-value
 * std::vector inputs and by-reference outputs, the same plb::plint mask argument. The
 * GPU build compiles this file unchanged into device code; the CPU oracle compiles it
 * as ordinary host C++.
 *
 *     r     = k_max * B * D/(Ks_D + D) * A/(Ks_A + A)      (per second)
 *     dD/dt = dA/dt = -r,  dP1/dt = dP2/dt = +r,  dB/dt = Y * r
 *
 * Substrate order: C[0] = donor D, C[1] = acceptor A, C[2], C[3] = products.
 * With the time step dt = 1 s, k_max * dt = 1e-3 per step.
 *
 * Per-step stability limit, the same safeguard the reference's own kinetics file applies
 * (defineKinetics.hh, "PER-STEP STABILITY LIMIT"): with k_max B / Ks up to 10 per step the
 * explicit update would draw a reactant below zero, so r is capped so that neither D nor
 * A loses more than half of what it holds in one step. Scaling the single rate keeps the
 * stoichiometry (and so every mass balance) exact.
 */
#ifndef DEFINE_KINETICS_HH
#define DEFINE_KINETICS_HH

#include <vector>
#include <cstddef>
#include <cmath>
#include <algorithm>

namespace DualMonod {
    const double k_max = 1.0e-3;   // 1/s, maximum specific rate
    const double Ks_D  = 1.0e-2;   // half saturation on the donor
    const double Ks_A  = 1.0e-2;   // half saturation on the acceptor
    const double Y     = 0.4;      // biomass yield per unit donor
    const double dt    = 1.0;      // s, the transport time step the rates are applied over
    const double MAX_RATE_FRACTION = 0.5;
}

void defineRxnKinetics(std::vector<double> B, std::vector<double> C,
                       std::vector<double>& subsR, std::vector<double>& bioR,
                       plb::plint mask)
{
    for (std::size_t i = 0; i < subsR.size(); ++i) subsR[i] = 0.0;
    for (std::size_t i = 0; i < bioR.size();  ++i) bioR[i]  = 0.0;
    if (C.size() < 4 || subsR.size() < 4 || B.size() < 1 || bioR.size() < 1) return;
    if (mask < 2) return;                        // no reactions in solid or wall voxels

    const double D  = std::max(C[0], 0.0);
    const double A  = std::max(C[1], 0.0);
    const double Bm = std::max(B[0], 0.0);
    double r = DualMonod::k_max * Bm * (D / (DualMonod::Ks_D + D)) * (A / (DualMonod::Ks_A + A));

    const double cap = DualMonod::MAX_RATE_FRACTION * std::min(D, A) / DualMonod::dt;
    if (r > cap) r = cap;

    subsR[0] = -r;
    subsR[1] = -r;
    subsR[2] = +r;
    subsR[3] = +r;
    bioR[0]  = DualMonod::Y * r;

    for (std::size_t i = 0; i < subsR.size(); ++i)
        if (std::isnan(subsR[i]) || std::isinf(subsR[i])) subsR[i] = 0.0;
    for (std::size_t i = 0; i < bioR.size(); ++i)
        if (std::isnan(bioR[i]) || std::isinf(bioR[i])) bioR[i] = 0.0;
}

#endif
