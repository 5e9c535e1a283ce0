// Host construction of the translation operators uploaded to the device once
// per plan (FP64 on the host; the device stores FP32 or FP64 copies).
#pragma once

#include <array>
#include <vector>

#include "basis.hpp"

namespace cdfmm_hip {

// Raw derivatives D_g G(r), G = 1/(4 pi |r|), for every g of basis(order) in
// storage order.  Same quantity as the reference's laplace_derivatives_raw
// (src/laplace_derivatives.cpp:21-42), computed here with the Cartesian
// Taylor-coefficient recurrence of the Laplace kernel (Duan & Krasny 2001;
// Lindsay & Krasny 2001): for a_k = D_k(1/|r|)/k!,
//   |k| |r|^2 a_k = -(2|k|-1) sum_i r_i a_{k-e_i} - (|k|-1) sum_i a_{k-2e_i}.
std::vector<double> laplace_derivatives(int order, const std::array<double, 3>& r);

// Column-major class matrix T[beta + n*alpha] = D_{alpha+beta} G(R) for an
// integer box offset R (src/static_operators.cpp:159-184).
std::vector<double> m2l_class_matrix(const Basis& basis, const std::array<int, 3>& offset);

// Every list-2 offset of a uniform octree: (dx,dy,dz) in [-3,3]^3 with
// max(|d|) >= 2, lexicographic order (316 offsets).
std::vector<std::array<int, 3>> list2_offsets();

// Dense row-major shift operators (out x in) for one displacement d:
//   M2M  M_parent[a] += sum_{g<=a} d^g/g! M_child[a-g],  d = parent - child
//   L2L  L_child[b]  += sum_{|b+g|<=p} d^g/g! L_parent[b+g], d = child - parent
// (src/static_operators.cpp:257-298).
std::vector<double> m2m_dense(const Basis& basis, const std::array<double, 3>& d);
std::vector<double> l2l_dense(const Basis& basis, const std::array<double, 3>& d);

// r^g / g! (include/cdfmm/multi_index.hpp:112-115).
double monomial_over_factorial(const std::array<double, 3>& r, const MultiIndex& g);

}  // namespace cdfmm_hip
