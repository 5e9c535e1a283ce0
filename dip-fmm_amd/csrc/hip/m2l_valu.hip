// Plain M2L on the vector ALUs (FP32 or FP64 FMA): the baseline path.
//
// Reference: apply_static_m2l_plan / apply_scaled_m2l_rows_kernel
// (src/static_operators.cpp:16-102, src/cuda_fmm.cu:142-189):
//   L[t][b] += ls[l][b] * sum_{s in list2(t)} sum_a T_{t-s}[b][a] ms[l][a] M[s][a]
// with class matrices keyed by the integer offset t - s
// (src/uniform_fmm.cpp:502-625).
//
// Regrouping: list2(t) is every box s = t - o inside the level's grid for
// the offsets o in [q-3, q+2]^3 minus [-1,1]^3 per axis, q = t's coordinate
// parity (189 offsets per parity class, 316 overall).  The boxes of one
// parity class therefore share one offset list, and a workgroup that owns a
// tile of same-parity targets walks the 189 offsets once, streaming each
// class matrix T_o from L2 and the tile's 16 gathered source columns from
// LDS.  Each thread owns one output coefficient b for all 16 columns
// (column = target box x moment state); accumulation is in registers and
// every local is written once, without atomics.  Empty source boxes carry
// exact zero multipoles, so they need no test.  One launch covers every
// level 2..L, so the small upper levels overlap the leaf level.
#include "common.hpp"

namespace hipfmm {

namespace {

constexpr int kCols = 16;

struct ValuLevels {
  int count;
  int block_begin[10];
  int tiles[10];  // tiles per parity
  int boxes_per_tile;
  int k_per_tile;
  int kchunks;
};

template <class S>
__global__ void m2l_valu_kernel(const DevicePlan d, const ValuLevels lv, int K) {
  extern __shared__ unsigned char smem_raw[];
  S* ms_tile = reinterpret_cast<S*>(smem_raw);  // n x kCols
  __shared__ int4 tbox[kCols];                   // target box coords + valid flag
  const int n = d.n;
  int li = lv.count - 1;
  while (li > 0 && int(blockIdx.x) < lv.block_begin[li]) --li;
  const int level = 2 + li;
  const int lb = int(blockIdx.x) - lv.block_begin[li];
  const int parity = lb & 7, rest = lb >> 3;
  const int tile = rest % lv.tiles[li], kc = rest / lv.tiles[li];
  const int px = parity & 1, py = (parity >> 1) & 1, pz = (parity >> 2) & 1;
  const int nbox = 1 << level;
  const int64_t parents = int64_t(1) << (3 * (level - 1));
  const int bpt = lv.boxes_per_tile, kpt = lv.k_per_tile;
  const int64_t q0 = int64_t(tile) * bpt;
  const int k0 = kc * kpt;
  const int64_t lvl0 = level_offset(level);
  if (threadIdx.x < kCols) {
    const int b = threadIdx.x / kpt;
    int4 tb{0, 0, 0, 0};
    if (b < bpt && q0 + b < parents && k0 + threadIdx.x % kpt < K) {
      int x, y, z;
      demorton3(uint32_t(q0 + b), level - 1, x, y, z);
      tb = int4{2 * x + px, 2 * y + py, 2 * z + pz, 1};
    }
    tbox[threadIdx.x] = tb;
  }
  const S* M = static_cast<const S*>(d.M);
  const S* msl = static_cast<const S*>(d.ms) + int64_t(level) * n;
  const S* T = static_cast<const S*>(d.m2l_T);
  const int4* offs = d.parity_offsets + parity * 189;
  const int beta = threadIdx.x;
  S acc[kCols];
#pragma unroll
  for (int c = 0; c < kCols; ++c) acc[c] = S(0);
  for (int oi = 0; oi < 189; ++oi) {
    const int4 o = offs[oi];
    __syncthreads();
    for (int w = threadIdx.x; w < n * kCols; w += blockDim.x) {
      const int a = w / kCols, col = w % kCols;
      const int4 tb = tbox[col];
      const int sx = tb.x - o.x, sy = tb.y - o.y, sz = tb.z - o.z;
      S v = S(0);
      if (tb.w && sx >= 0 && sy >= 0 && sz >= 0 && sx < nbox && sy < nbox && sz < nbox) {
        const int64_t snode = lvl0 + morton3(uint32_t(sx), uint32_t(sy), uint32_t(sz), level);
        v = msl[a] * M[(snode * n + a) * K + k0 + col % kpt];
      }
      ms_tile[w] = v;
    }
    __syncthreads();
    if (beta < n) {
      const S* Tb = T + int64_t(o.w) * n * n + beta;
#pragma unroll 4
      for (int a = 0; a < n; ++a) {
        const S t = Tb[int64_t(a) * n];
        const S* mrow = ms_tile + a * kCols;
#pragma unroll
        for (int c = 0; c < kCols; ++c) acc[c] += t * mrow[c];
      }
    }
  }
  if (beta < n) {
    const S lsb = static_cast<const S*>(d.ls)[int64_t(level) * n + beta];
    S* Lc = static_cast<S*>(d.Lc);
#pragma unroll
    for (int c = 0; c < kCols; ++c) {
      const int4 tb = tbox[c];
      if (!tb.w) continue;
      const int64_t tnode = lvl0 + morton3(uint32_t(tb.x), uint32_t(tb.y), uint32_t(tb.z), level);
      Lc[(tnode * n + beta) * K + k0 + c % kpt] = lsb * acc[c];
    }
  }
}

}  // namespace

template <class S>
void launch_m2l_valu(const DevicePlan& d, int K, hipStream_t s) {
  if (d.L < 2) return;
  ValuLevels lv{};
  lv.count = d.L - 1;
  lv.k_per_tile = K >= kCols ? kCols : K;
  lv.boxes_per_tile = kCols / lv.k_per_tile;
  lv.kchunks = (K + lv.k_per_tile - 1) / lv.k_per_tile;
  int begin = 0;
  for (int i = 0; i < lv.count; ++i) {
    const int64_t parents = int64_t(1) << (3 * (i + 1));
    lv.tiles[i] = int((parents + lv.boxes_per_tile - 1) / lv.boxes_per_tile);
    lv.block_begin[i] = begin;
    begin += 8 * lv.tiles[i] * lv.kchunks;
  }
  lv.block_begin[lv.count] = begin;
  const int threads = ((d.n + 63) / 64) * 64;
  const size_t lds = size_t(d.n) * kCols * sizeof(S);
  m2l_valu_kernel<S><<<unsigned(begin), threads, lds, s>>>(d, lv, K);
  HIPFMM_CHECK(hipGetLastError());
}
template void launch_m2l_valu<float>(const DevicePlan&, int, hipStream_t);
template void launch_m2l_valu<double>(const DevicePlan&, int, hipStream_t);

}  // namespace hipfmm
