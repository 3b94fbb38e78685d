/* defineAbioticKinetics.hh -- one second-order abiotic reaction between two solutes.
 *
 * Written against the reference's abiotic kinetics interface
 * (defineAbioticKinetics.hh:130-134 of CompLB3D): the same free function, by-value
 * std::vector of concentrations, by-reference rates, plb::plint mask. The GPU build
 * compiles this file unchanged into device code; the CPU oracle compiles it as host C++.
 *
 *     R = k * C[a] * C[b]         C[a] and C[b] consumed, C[p] produced (mol/L/s)
 *
 * With the four-solute dual-Monod case (D, A, P1, P2) the product P1 is re-oxidised by
 * the acceptor A into P2: a = 2, b = 1, p = 3. The per-step stability cap keeps either
 * reactant from losing more than half of what it holds in one step.
 */
#ifndef DEFINE_ABIOTIC_KINETICS_HH
#define DEFINE_ABIOTIC_KINETICS_HH

#include <vector>
#include <cstddef>
#include <cmath>
#include <algorithm>

namespace SecondOrderAbiotic {
    const double k   = 5.0e-2;    // L/mol/s
    const int    ia  = 2;         // first reactant
    const int    ib  = 1;         // second reactant
    const int    ip  = 3;         // product
    const double dt  = 1.0;       // s, the transport time step
    const double MAX_RATE_FRACTION = 0.5;
    // dissolution of solid phase 1: first order in the attacking solute and in the mineral
    // left, releasing one mole of `irel` per mole dissolved (the worked example's form)
    const double k_diss = 2.0e-2;  // L/mol/s
    const int    iatt   = 1;       // the solute that attacks the mineral (consumed)
    const int    irel   = 0;       // the solute the mineral releases
}

void defineAbioticRxnKinetics(std::vector<double> C, std::vector<double>& subsR, plb::plint mask)
{
    using namespace SecondOrderAbiotic;
    for (std::size_t i = 0; i < subsR.size(); ++i) subsR[i] = 0.0;
    if (C.size() < 4 || subsR.size() < 4) return;
    if (mask < 2) return;                       // nothing happens in solid or wall voxels

    const double a = std::max(C[ia], 0.0);
    const double b = std::max(C[ib], 0.0);
    double R = k * a * b;
    const double cap = MAX_RATE_FRACTION * std::min(a, b) / dt;
    if (R > cap) R = cap;

    subsR[ia] -= R;
    subsR[ib] -= R;
    subsR[ip] += R;

    for (std::size_t i = 0; i < subsR.size(); ++i)
        if (std::isnan(subsR[i]) || std::isinf(subsR[i])) subsR[i] = 0.0;
}

/* Mineral dissolution hook (the reference's defineAbioticKinetics.hh:292 interface): given
 * the water in contact with the mineral, the phase id and the inventory left, return the
 * mineral's rate (negative: dissolving) and the released / consumed substrates (mol/L/s). */
inline void defineDissolutionRate(plb::plint phaseId, const std::vector<double>& C, double mineral,
                                  std::vector<double>& subsR, double& mineralR, plb::plint mask)
{
    using namespace SecondOrderAbiotic;
    (void) mask;
    mineralR = 0.0;
    for (std::size_t i = 0; i < subsR.size(); ++i) subsR[i] = 0.0;
    if (phaseId != 1 || C.size() < 2 || subsR.size() < 2) return;
    const double rate = k_diss * std::max(C[iatt], 0.0) * std::max(mineral, 0.0);
    mineralR = -rate;
    subsR[irel] += rate;
    subsR[iatt] -= rate;
    if (std::isnan(mineralR) || std::isinf(mineralR)) mineralR = 0.0;
    for (std::size_t i = 0; i < subsR.size(); ++i)
        if (std::isnan(subsR[i]) || std::isinf(subsR[i])) subsR[i] = 0.0;
}

#endif
