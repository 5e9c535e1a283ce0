// Total-degree Cartesian multi-index basis (host side of HipFull).
//
// Storage order is the reference's coefficient layout: by total degree, then
// alpha_x, then alpha_y (include/cdfmm/multi_index.hpp:59-129).  The device
// kernels and every uploaded operator use this order.
#pragma once

#include <array>
#include <stdexcept>
#include <vector>

namespace cdfmm_hip {

using MultiIndex = std::array<int, 3>;

class Basis {
 public:
  explicit Basis(int order) : p_(order) {
    if (order < 0) throw std::invalid_argument("expansion order must be >= 0");
    for (int d = 0; d <= p_; ++d)
      for (int ax = 0; ax <= d; ++ax)
        for (int ay = 0; ay <= d - ax; ++ay) entries_.push_back({ax, ay, d - ax - ay});
    lookup_.assign(static_cast<std::size_t>((p_ + 1) * (p_ + 1) * (p_ + 1)), -1);
    for (int i = 0; i < size(); ++i) lookup_[key(entries_[i])] = i;
  }
  int order() const { return p_; }
  int size() const { return static_cast<int>(entries_.size()); }
  const MultiIndex& operator[](int i) const { return entries_[static_cast<std::size_t>(i)]; }
  // Index of a multi-index, or -1 when outside the basis.
  int find(int ax, int ay, int az) const {
    if (ax < 0 || ay < 0 || az < 0 || ax + ay + az > p_) return -1;
    return lookup_[key({ax, ay, az})];
  }
  static int count(int order) { return (order + 1) * (order + 2) * (order + 3) / 6; }

 private:
  std::size_t key(const MultiIndex& m) const {
    return static_cast<std::size_t>((m[0] * (p_ + 1) + m[1]) * (p_ + 1) + m[2]);
  }
  int p_;
  std::vector<MultiIndex> entries_;
  std::vector<int> lookup_;
};

inline int degree(const MultiIndex& m) { return m[0] + m[1] + m[2]; }

inline double factorial(int n) {
  double v = 1.0;
  for (int i = 2; i <= n; ++i) v *= i;
  return v;
}

}  // namespace cdfmm_hip
