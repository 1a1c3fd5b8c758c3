// Direct element-based construction of the first Galerkin coarse operator.
#pragma once
#include <vector>

#include "common.hpp"

namespace et {

// Constant child tables M_t = W_t^T K0 W_t, layout [t][a][b][r][s] (host).
std::vector<double> galerkin_child_tables(int dim, const double* K0_host);
// True when every element count is even (each fine element has one coarse parent).
bool galerkin_elem_supported(const Grid& g);
// S_coarse = R A_hat R^T from E, the constraint bytes and the device tables.
// eflag: n_elem bytes of scratch.
int galerkin_from_elements(const Grid& gf, const double* K0, const double* Mt, const double* E,
                           const uint8_t* cons, uint8_t* eflag, double* Sc, hipStream_t s);

}  // namespace et
