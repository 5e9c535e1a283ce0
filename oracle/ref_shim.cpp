// Test infrastructure only: a C-ABI shim over the reference's own C++ API.
//
// This file is compiled together with the reference sources where they lie
// (see oracle/Makefile) into oracle/_ref/libcdfmm_ref.so.  It lets the Python
// tests drive the real reference CpuStatic path and read its plan (Morton
// permutations, node ranges, list1/list2, M2L matrices) so that the oracle
// restatement (oracle/cpustatic.cpp) and the HipFull host plan can be pinned
// against it.  Nothing in the product links or loads this library.
//
// Reference interfaces used:
//   cdfmm::UniformFmm            include/cdfmm/uniform_fmm.hpp:153-442
//   cdfmm::UniformTree           include/cdfmm/uniform_tree.hpp:82-157
//   cdfmm::build_static_m2l_matrix  src/static_operators.cpp:159-184
//   cdfmm::laplace_derivatives_raw  src/laplace_derivatives.cpp:21-42
//   cdfmm::direct_p2p_reference     src/validation.cpp:71-111

#include <cstring>
#include <exception>
#include <memory>
#include <string>
#include <vector>

#include "cdfmm/laplace_derivatives.hpp"
#include "cdfmm/static_operators.hpp"
#include "cdfmm/uniform_fmm.hpp"
#include "cdfmm/validation.hpp"

namespace {

thread_local std::string g_last_error;

struct RefHandle {
  std::unique_ptr<cdfmm::UniformFmm> fmm;
  bool float32{false};
};

std::vector<cdfmm::Vec3> to_vec3(const double* xyz, int n) {
  std::vector<cdfmm::Vec3> out(static_cast<std::size_t>(n));
  for (int i = 0; i < n; ++i) {
    out[static_cast<std::size_t>(i)] = {xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]};
  }
  return out;
}

template <class F>
int guarded(F&& f) {
  try {
    f();
    return 0;
  } catch (const std::invalid_argument& e) {
    g_last_error = e.what();
    return 2;
  } catch (const std::exception& e) {
    g_last_error = e.what();
    return 1;
  }
}

}  // namespace

extern "C" {

const char* ref_last_error() { return g_last_error.c_str(); }

// precision: 32 or 64.  root_half_width <= 0 selects the inferred root.
void* ref_fmm_create(const double* src, int ns, const double* tgt, int nt, int order,
                     int max_level, int precision, int has_root, double cx, double cy,
                     double cz, double root_half_width, const int* fixed_ids, int n_fixed) {
  RefHandle* handle = nullptr;
  const int rc = guarded([&] {
    cdfmm::UniformFmmOptions options;
    options.expansion_order = order;
    options.tree.max_level = max_level;
    options.backend = cdfmm::ExecutionBackend::CpuStatic;
    options.precision = precision == 32 ? cdfmm::StaticPrecision::Float32
                                        : cdfmm::StaticPrecision::Float64;
    if (has_root) {
      options.tree.root_centre = cdfmm::Vec3{cx, cy, cz};
    }
    if (root_half_width > 0.0) {
      options.tree.root_half_width = root_half_width;
    }
    if (fixed_ids != nullptr) {
      options.fixed_target_source_indices =
          std::vector<int>(fixed_ids, fixed_ids + n_fixed);
    }
    auto h = std::make_unique<RefHandle>();
    h->float32 = precision == 32;
    h->fmm = std::make_unique<cdfmm::UniformFmm>(to_vec3(src, ns), to_vec3(tgt, nt), options);
    handle = h.release();
  });
  return rc == 0 ? handle : nullptr;
}

void ref_fmm_destroy(void* h) { delete static_cast<RefHandle*>(h); }

// output: 1 potential, 2 field, 3 both.  H_out is nt*3 doubles, phi_out nt.
int ref_fmm_evaluate(void* hv, const double* moments, int ns, const int* ids, int n_ids,
                     int output, double* H_out, double* phi_out) {
  auto* h = static_cast<RefHandle*>(hv);
  return guarded([&] {
    const std::vector<cdfmm::Vec3> m = to_vec3(moments, ns);
    const std::span<const int> id_span =
        ids == nullptr ? std::span<const int>{} : std::span<const int>(ids, n_ids);
    const auto flags = static_cast<cdfmm::OutputFlags>(output);
    if (h->float32) {
      const auto r = h->fmm->evaluate_float32(m, flags, id_span);
      for (std::size_t t = 0; t < r.size(); ++t) {
        H_out[3 * t] = r[t].H.x;
        H_out[3 * t + 1] = r[t].H.y;
        H_out[3 * t + 2] = r[t].H.z;
        if (phi_out) phi_out[t] = r[t].phi;
      }
    } else {
      const auto r = h->fmm->evaluate(m, flags, id_span);
      for (std::size_t t = 0; t < r.size(); ++t) {
        H_out[3 * t] = r[t].H.x;
        H_out[3 * t + 1] = r[t].H.y;
        H_out[3 * t + 2] = r[t].H.z;
        if (phi_out) phi_out[t] = r[t].phi;
      }
    }
  });
}

// Coefficients of one node after the last evaluate (which = 0 multipole, 1 local).
int ref_fmm_coefficients(void* hv, int which, int node, double* out) {
  auto* h = static_cast<RefHandle*>(hv);
  return guarded([&] {
    if (h->float32) {
      const auto c = which == 0 ? h->fmm->multipole_float32(node) : h->fmm->local_float32(node);
      for (std::size_t i = 0; i < c.size(); ++i) out[i] = c[i];
    } else {
      const auto c = which == 0 ? h->fmm->multipole(node) : h->fmm->local(node);
      for (std::size_t i = 0; i < c.size(); ++i) out[i] = c[i];
    }
  });
}

// sizes: [node_count, n_sources, n_targets, list1_total, list2_total, basis_size]
void ref_fmm_sizes(void* hv, long long* out) {
  auto* h = static_cast<RefHandle*>(hv);
  const auto& tree = h->fmm->tree();
  long long l1 = 0, l2 = 0;
  for (const auto& node : tree.nodes()) {
    l1 += static_cast<long long>(node.list1.size());
    l2 += static_cast<long long>(node.list2.size());
  }
  out[0] = static_cast<long long>(tree.nodes().size());
  out[1] = static_cast<long long>(tree.sorted_source_positions().size());
  out[2] = static_cast<long long>(tree.sorted_target_positions().size());
  out[3] = l1;
  out[4] = l2;
  out[5] = h->fmm->basis().size();
}

// Root geometry: [cx, cy, cz, half_width]
void ref_fmm_root(void* hv, double* out) {
  auto* h = static_cast<RefHandle*>(hv);
  const auto& tree = h->fmm->tree();
  out[0] = tree.root_centre().x;
  out[1] = tree.root_centre().y;
  out[2] = tree.root_centre().z;
  out[3] = tree.root_half_width();
}

// Named integer arrays of the tree plan.
int ref_fmm_tree_ints(void* hv, const char* name, long long* out) {
  auto* h = static_cast<RefHandle*>(hv);
  return guarded([&] {
    const auto& tree = h->fmm->tree();
    const std::string n(name);
    auto copy_span = [&](std::span<const int> s) {
      for (std::size_t i = 0; i < s.size(); ++i) out[i] = s[i];
    };
    if (n == "source_permutation") {
      copy_span(tree.source_permutation());
    } else if (n == "source_inverse_permutation") {
      copy_span(tree.source_inverse_permutation());
    } else if (n == "target_permutation") {
      copy_span(tree.target_permutation());
    } else if (n == "target_inverse_permutation") {
      copy_span(tree.target_inverse_permutation());
    } else if (n == "occupied_source_leaves") {
      copy_span(tree.occupied_source_leaves());
    } else if (n == "occupied_target_leaves") {
      copy_span(tree.occupied_target_leaves());
    } else if (n == "node_ranges") {
      // 4 per node: source_begin, source_end, target_begin, target_end
      std::size_t i = 0;
      for (const auto& node : tree.nodes()) {
        out[i++] = static_cast<long long>(node.source_begin);
        out[i++] = static_cast<long long>(node.source_end);
        out[i++] = static_cast<long long>(node.target_begin);
        out[i++] = static_cast<long long>(node.target_end);
      }
    } else if (n == "list1" || n == "list2") {
      // CSR: node_count+1 offsets followed by the concatenated entries.
      const bool first = n == "list1";
      const auto nodes = tree.nodes();
      std::size_t off = 0;
      out[0] = 0;
      for (std::size_t k = 0; k < nodes.size(); ++k) {
        off += (first ? nodes[k].list1 : nodes[k].list2).size();
        out[k + 1] = static_cast<long long>(off);
      }
      std::size_t i = nodes.size() + 1;
      for (const auto& node : nodes) {
        for (const int v : (first ? node.list1 : node.list2)) out[i++] = v;
      }
    } else {
      throw std::invalid_argument("unknown tree array " + n);
    }
  });
}

int ref_m2l_matrix(int order, double dx, double dy, double dz, double* out) {
  return guarded([&] {
    const cdfmm::MultiIndexSet basis(order);
    const auto m = cdfmm::build_static_m2l_matrix(basis, cdfmm::Vec3{dx, dy, dz});
    std::memcpy(out, m.data(), m.size() * sizeof(double));
  });
}

int ref_laplace_derivatives(int order, double x, double y, double z, double* out) {
  return guarded([&] {
    const cdfmm::MultiIndexSet basis(order);
    const auto d = cdfmm::laplace_derivatives_raw(basis, cdfmm::Vec3{x, y, z});
    std::memcpy(out, d.data(), d.size() * sizeof(double));
  });
}

int ref_direct(const double* tgt, int nt, const double* src, int ns, const double* moments,
               const int* ids, int output, double* H_out, double* phi_out) {
  return guarded([&] {
    const auto t = to_vec3(tgt, nt);
    const auto s = to_vec3(src, ns);
    const auto m = to_vec3(moments, ns);
    const std::span<const int> id_span =
        ids == nullptr ? std::span<const int>{} : std::span<const int>(ids, nt);
    const auto r = cdfmm::direct_p2p_reference(t, s, m, static_cast<cdfmm::OutputFlags>(output),
                                               id_span);
    for (std::size_t i = 0; i < r.size(); ++i) {
      H_out[3 * i] = r[i].H.x;
      H_out[3 * i + 1] = r[i].H.y;
      H_out[3 * i + 2] = r[i].H.z;
      if (phi_out) phi_out[i] = r[i].phi;
    }
  });
}

}  // extern "C"
