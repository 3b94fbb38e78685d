/* defineAbioticKinetics.hh -- one second-order abiotic reaction between two solutes.
 *
 * Written against the reference's abiotic kinetics interface
 * (defineAbioticKinetics.hh:130-134 of CompLB3D): the same free function, by-value
 * std::vector of concentrations, by-reference rates, plb::plint mask. The GPU build
 * compiles this file unchanged into device code; the CPU checker compiles it as host C++.
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

#endif
