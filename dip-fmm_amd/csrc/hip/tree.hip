// Device-side leaf order of the uniform tree: the stable sort of points by
// the Morton index of their leaf, and each leaf's [begin, end) range, equal
// to the host build (the leaf-key and counting-sort steps of
// src/uniform_tree.cpp, restated in host/uniform_fmm.cpp).
//
// Keys are computed in FP64 exactly as on the host: per axis
// floor((x - (c - h)) / cell), clamped to [0, 2^L - 1] (0 for a root of zero
// width, where all points coincide), with a point outside
// the root box (beyond a 1e-12 tolerance) reported as an error.  The sort is
// hipCUB's LSD radix sort of (key, index) pairs over the 3L key bits, which
// is stable, so equal keys keep the input order as the host's counting sort
// does.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>

#include "common.hpp"

namespace hipfmm {

namespace {

__global__ void leaf_key_kernel(const double* __restrict__ pos, int count, double cx, double cy, double cz, double h,
                                int L, uint32_t* __restrict__ keys, int* __restrict__ index, int* __restrict__ outside) {
  const int i = int(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= count) return;
  const double c[3] = {cx, cy, cz};
  const int nb = 1 << L;
  const double cell = (2.0 * h) / double(nb);
  int idx[3] = {0, 0, 0};
  bool out = false;
  for (int a = 0; a < 3; ++a) {
    const double v = pos[3 * int64_t(i) + a], lo = c[a] - h, hi = c[a] + h;
    if (!(v >= lo - 1.0e-12 && v <= hi + 1.0e-12)) out = true;
    // A root of zero width (all points coincide) has every point in leaf 0.
    if (L > 0 && cell > 0.0) idx[a] = min(max(int(floor((v - lo) / cell)), 0), nb - 1);
  }
  if (out) *outside = 1;  // every writer stores the same value
  keys[i] = morton3(uint32_t(idx[0]), uint32_t(idx[1]), uint32_t(idx[2]), L);
  index[i] = i;
}

__global__ void fill_ranges_kernel(int64_t* __restrict__ ranges, int64_t leaves, int64_t empty) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < 2 * leaves) ranges[i] = empty;
}

// Leaf l occupies the sorted positions [begin, end): a position starts its
// leaf's range where the key changes from the previous one, and ends it
// where it changes to the next.
__global__ void leaf_ranges_kernel(const uint32_t* __restrict__ keys, int count, int64_t* __restrict__ ranges) {
  const int i = int(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= count) return;
  const uint32_t k = keys[i];
  if (i == 0 || keys[i - 1] != k) ranges[2 * int64_t(k)] = i;
  if (i == count - 1 || keys[i + 1] != k) ranges[2 * int64_t(k) + 1] = i + 1;
}

}  // namespace

void device_leaf_order(const double* positions, int count, const double* centre, double half_width, int L,
                       int* permutation, int64_t* leaf_ranges, hipStream_t s) {
  if (L < 0 || L > 9) throw invalid_argument("max_level must be in [0, 9]");
  if (count < 0) throw invalid_argument("point count is negative");
  if (!(half_width >= 0.0) || !std::isfinite(half_width)) throw invalid_argument("root half width must be finite and >= 0");
  const int64_t leaves = int64_t(1) << (3 * L);
  fill_ranges_kernel<<<unsigned((2 * leaves + 255) / 256), 256, 0, s>>>(leaf_ranges, leaves, count);
  HIPFMM_CHECK(hipGetLastError());
  if (count == 0) return;
  // Scratch: keys in/out, indices in, the outside flag and the sort's
  // temporary storage, freed before returning.
  uint32_t* keys = nullptr;
  int* index = nullptr;
  int* outside = nullptr;
  void* temp = nullptr;
  size_t temp_bytes = 0;
  HIPFMM_CHECK(hipcub::DeviceRadixSort::SortPairs(nullptr, temp_bytes, static_cast<const uint32_t*>(nullptr),
                                                  static_cast<uint32_t*>(nullptr), static_cast<const int*>(nullptr),
                                                  static_cast<int*>(nullptr), count, 0, std::max(3 * L, 1), s));
  HIPFMM_CHECK(hipMalloc(&keys, 2 * size_t(count) * sizeof(uint32_t)));
  HIPFMM_CHECK(hipMalloc(&index, size_t(count) * sizeof(int)));
  HIPFMM_CHECK(hipMalloc(&outside, sizeof(int)));
  HIPFMM_CHECK(hipMalloc(&temp, std::max<size_t>(temp_bytes, 16)));
  int flag = 0;
  try {
    HIPFMM_CHECK(hipMemsetAsync(outside, 0, sizeof(int), s));
    leaf_key_kernel<<<unsigned((count + 255) / 256), 256, 0, s>>>(positions, count, centre[0], centre[1], centre[2],
                                                                  half_width, L, keys, index, outside);
    HIPFMM_CHECK(hipGetLastError());
    HIPFMM_CHECK(hipcub::DeviceRadixSort::SortPairs(temp, temp_bytes, keys, keys + count, index, permutation, count,
                                                    0, std::max(3 * L, 1), s));
    leaf_ranges_kernel<<<unsigned((count + 255) / 256), 256, 0, s>>>(keys + count, count, leaf_ranges);
    HIPFMM_CHECK(hipGetLastError());
    HIPFMM_CHECK(hipMemcpyAsync(&flag, outside, sizeof(int), hipMemcpyDeviceToHost, s));
    HIPFMM_CHECK(hipStreamSynchronize(s));
  } catch (...) {
    (void)hipStreamSynchronize(s);
    (void)hipFree(temp);
    (void)hipFree(outside);
    (void)hipFree(index);
    (void)hipFree(keys);
    throw;
  }
  HIPFMM_CHECK(hipFree(temp));
  HIPFMM_CHECK(hipFree(outside));
  HIPFMM_CHECK(hipFree(index));
  HIPFMM_CHECK(hipFree(keys));
  if (flag != 0) throw invalid_argument("Point lies outside the requested root box");
}

}  // namespace hipfmm
