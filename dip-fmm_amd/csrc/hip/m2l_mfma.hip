// M2L on the CDNA4 matrix cores (v_mfma_f32_32x32x16_bf16).
//
// Same regrouped contraction as the VALU path (m2l_valu.hip):
//   L[t][b] = ls[l][b] * sum_{o in list2 offsets} sum_a T_o[b][a] * Mt[t-o][a],
//   Mt[s][a] = ms[l][a] * M[s][a]
// written as one GEMM per offset class: A = T_o (rows b, depth a), B = the
// gathered source columns (depth a, columns = target box x moment state).
// A workgroup holds 4 waves; each wave owns 32 output columns and every row
// tile of L for them (n rounded up to 32 rows, 16 accumulators per tile per
// lane) and walks the 189 offsets of its parity class, so each local is
// written once and no atomics are needed.
//
// Precision: FP32 operands are split x = hi + lo with hi = bf16(x) and
// lo = bf16(x - hi) (16 significant bits), and each product is formed as
// hi*hi + hi*lo + lo*hi with FP32 accumulation on the matrix cores.
//
// Data layout:
//   T_hi/T_lo  [class][row tile][depth step][lane][8] bf16 -- the exact MFMA
//              A-fragment order, so a lane fetches a fragment with one 16-byte
//              load (lane i: row i%32, depth 8*(i/32)..+7 of the 16-deep step)
//   mt_hi/lo   [node][state][apad] bf16 -- a B fragment (column = lane%32,
//              depth 8*(i/32)..+7) is one 16-byte load as well
#include <cstring>

#include "common.hpp"

namespace hipfmm {

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

__host__ __device__ inline uint16_t bf16_bits(float x) {
  uint32_t u = __builtin_bit_cast(uint32_t, x);
  if ((u & 0x7f800000u) == 0x7f800000u) return uint16_t(u >> 16);  // inf / nan
  u += 0x7fffu + ((u >> 16) & 1u);                                   // round to nearest even
  return uint16_t(u >> 16);
}
__host__ __device__ inline float bf16_value(uint16_t h) {
  return __builtin_bit_cast(float, uint32_t(h) << 16);
}

// One workgroup per (box, chunk of kc moment states): transpose
// M[node][a][k] to [k][a], scale by ms and split into hi/lo bf16 halves;
// padding depth entries are zero.
__global__ void __launch_bounds__(256) split_kernel(const DevicePlan d, int level, int K, int kc) {
  extern __shared__ float tile[];  // n x kc
  const int n = d.n, AP = d.apad;
  const int k0 = blockIdx.y * kc;
  const int64_t node = level_offset(level) + blockIdx.x;
  const float* M = static_cast<const float*>(d.M) + node * int64_t(n) * K;
  for (int w = threadIdx.x; w < n * kc; w += blockDim.x) {
    const int a = w / kc, k = k0 + w % kc;
    tile[w] = k < K ? M[a * K + k] : 0.0f;
  }
  __syncthreads();
  const float* ms = static_cast<const float*>(d.ms) + int64_t(level) * n;
  uint16_t* hi = static_cast<uint16_t*>(d.mt_hi) + (node * K + k0) * int64_t(AP);
  uint16_t* lo = static_cast<uint16_t*>(d.mt_lo) + (node * K + k0) * int64_t(AP);
  const int kn = min(kc, K - k0);
  for (int w = threadIdx.x; w < kn * AP; w += blockDim.x) {
    const int k = w / AP, a = w % AP;
    const float x = a < n ? ms[a] * tile[a * kc + k] : 0.0f;
    const uint16_t h = bf16_bits(x);
    hi[w] = h;
    lo[w] = bf16_bits(x - bf16_value(h));
  }
}

__device__ inline bf16x8 as_bf16x8(uint4 v) {
  bf16x8 r;
  __builtin_memcpy(&r, &v, 16);
  return r;
}

template <int MT>
__global__ void __launch_bounds__(256) m2l_mfma_kernel(const DevicePlan d, int level, int K, int cols_per_box,
                                                       int kchunks, int tiles) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int parity = blockIdx.y;
  const int tile = blockIdx.x * 4 + wave;
  if (tile >= tiles) return;  // no workgroup barriers below
  const int px = parity & 1, py = (parity >> 1) & 1, pz = (parity >> 2) & 1;
  const int n = d.n, AP = d.apad, BP = d.bpad, KS = AP / 16;
  const int nbox = 1 << level;
  const int64_t parents = int64_t(1) << (3 * (level - 1));
  const int boxes_per_tile = 32 / cols_per_box;
  const int bg = tile / kchunks, kc = tile % kchunks;
  const int col = lane & 31, g = lane >> 5;
  const int64_t q = int64_t(bg) * boxes_per_tile + col / cols_per_box;
  const int kk = kc * cols_per_box + col % cols_per_box;
  const bool valid = q < parents && kk < K;
  int tx = 0, ty = 0, tz = 0;
  if (valid) {
    demorton3(uint32_t(q), level - 1, tx, ty, tz);
    tx = 2 * tx + px;
    ty = 2 * ty + py;
    tz = 2 * tz + pz;
  }
  const int64_t lvl0 = level_offset(level);
  const uint4* Thi = static_cast<const uint4*>(d.T_hi);
  const uint4* Tlo = static_cast<const uint4*>(d.T_lo);
  const uint16_t* mhi = static_cast<const uint16_t*>(d.mt_hi);
  const uint16_t* mlo = static_cast<const uint16_t*>(d.mt_lo);
  f32x16 acc[MT];
#pragma unroll
  for (int m = 0; m < MT; ++m)
    for (int r = 0; r < 16; ++r) acc[m][r] = 0.0f;
  const uint4 zero4 = {0u, 0u, 0u, 0u};
  for (int ox = px - 3; ox <= px + 2; ++ox)
    for (int oy = py - 3; oy <= py + 2; ++oy)
      for (int oz = pz - 3; oz <= pz + 2; ++oz) {
        if (ox >= -1 && ox <= 1 && oy >= -1 && oy <= 1 && oz >= -1 && oz <= 1) continue;
        const int cls = d.m2l_slot[offset_slot(ox, oy, oz)];
        const int sx = tx - ox, sy = ty - oy, sz = tz - oz;
        const bool sv = valid && sx >= 0 && sy >= 0 && sz >= 0 && sx < nbox && sy < nbox && sz < nbox;
        const int64_t snode = sv ? lvl0 + morton3(uint32_t(sx), uint32_t(sy), uint32_t(sz), level) : 0;
        const uint4* bh = reinterpret_cast<const uint4*>(mhi + (snode * K + kk) * AP) + g;
        const uint4* bl = reinterpret_cast<const uint4*>(mlo + (snode * K + kk) * AP) + g;
        const uint4* ah = Thi + int64_t(cls) * MT * KS * 64 + lane;
        const uint4* al = Tlo + int64_t(cls) * MT * KS * 64 + lane;
        for (int ks = 0; ks < KS; ++ks) {
          const bf16x8 bhv = as_bf16x8(sv ? bh[2 * ks] : zero4);
          const bf16x8 blv = as_bf16x8(sv ? bl[2 * ks] : zero4);
#pragma unroll
          for (int m = 0; m < MT; ++m) {
            const bf16x8 ahv = as_bf16x8(ah[(m * KS + ks) * 64]);
            const bf16x8 alv = as_bf16x8(al[(m * KS + ks) * 64]);
            acc[m] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ahv, bhv, acc[m], 0, 0, 0);
            acc[m] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ahv, blv, acc[m], 0, 0, 0);
            acc[m] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(alv, bhv, acc[m], 0, 0, 0);
          }
        }
      }
  if (!valid) return;
  const int64_t tnode = lvl0 + morton3(uint32_t(tx), uint32_t(ty), uint32_t(tz), level);
  const float* lsp = static_cast<const float*>(d.ls_pad) + int64_t(level) * BP;
  float* Lc = static_cast<float*>(d.Lc) + tnode * int64_t(n) * K + kk;
#pragma unroll
  for (int m = 0; m < MT; ++m)
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int row = m * 32 + 8 * (r >> 2) + 4 * g + (r & 3);
      if (row < n) Lc[int64_t(row) * K] = lsp[row] * acc[m][r];
    }
}

template <int MT>
void launch_mfma_t(const DevicePlan& d, int level, int K, hipStream_t s) {
  int cols_per_box = 1;
  while (cols_per_box < K && cols_per_box < 32) cols_per_box *= 2;
  const int kchunks = (K + cols_per_box - 1) / cols_per_box;
  const int boxes_per_tile = 32 / cols_per_box;
  const int64_t parents = int64_t(1) << (3 * (level - 1));
  const int tiles = int((parents + boxes_per_tile - 1) / boxes_per_tile) * kchunks;
  const dim3 grid(unsigned((tiles + 3) / 4), 8);
  m2l_mfma_kernel<MT><<<grid, 256, 0, s>>>(d, level, K, cols_per_box, kchunks, tiles);
  HIPFMM_CHECK(hipGetLastError());
}

}  // namespace

void launch_m2l_split(const DevicePlan& d, int level, int K, int mode, hipStream_t s) {
  (void)mode;
  const unsigned boxes = unsigned(int64_t(1) << (3 * level));
  int kc = K;
  while (kc > 1 && size_t(d.n) * kc * sizeof(float) > 32768) kc = (kc + 1) / 2;
  const dim3 grid(boxes, unsigned((K + kc - 1) / kc));
  split_kernel<<<grid, 256, size_t(d.n) * kc * sizeof(float), s>>>(d, level, K, kc);
  HIPFMM_CHECK(hipGetLastError());
}

void launch_m2l_mfma(const DevicePlan& d, int level, int K, int mode, hipStream_t s) {
  (void)mode;
  switch (d.bpad / 32) {
    case 1: launch_mfma_t<1>(d, level, K, s); break;
    case 2: launch_mfma_t<2>(d, level, K, s); break;
    case 3: launch_mfma_t<3>(d, level, K, s); break;
    case 4: launch_mfma_t<4>(d, level, K, s); break;
    case 5: launch_mfma_t<5>(d, level, K, s); break;
    case 6: launch_mfma_t<6>(d, level, K, s); break;
    case 7: launch_mfma_t<7>(d, level, K, s); break;
    case 8: launch_mfma_t<8>(d, level, K, s); break;
    case 9: launch_mfma_t<9>(d, level, K, s); break;
    default: throw invalid_argument("expansion order too large for the MFMA M2L path");
  }
}

void prepare_mfma_operators(DevicePlan& d, const std::vector<double>& matrices, int classes,
                            const std::vector<int>& slot_of_class, const std::vector<double>& local_scaling,
                            std::vector<void*>& allocations, size_t& bytes) {
  (void)slot_of_class;
  const int n = d.n;
  d.apad = ((n + 15) / 16) * 16;
  d.bpad = ((n + 31) / 32) * 32;
  const int MT = d.bpad / 32, KS = d.apad / 16;
  const size_t per_class = size_t(MT) * KS * 64 * 8;
  std::vector<uint16_t> hi(per_class * classes), lo(per_class * classes);
  for (int c = 0; c < classes; ++c) {
    const double* T = matrices.data() + size_t(c) * n * n;  // T[beta + n*alpha]
    for (int m = 0; m < MT; ++m)
      for (int ks = 0; ks < KS; ++ks)
        for (int lane = 0; lane < 64; ++lane)
          for (int e = 0; e < 8; ++e) {
            const int row = m * 32 + lane % 32;
            const int dep = ks * 16 + 8 * (lane / 32) + e;
            const float v = (row < n && dep < n) ? float(T[row + size_t(n) * dep]) : 0.0f;
            const size_t i = size_t(c) * per_class + ((size_t(m) * KS + ks) * 64 + lane) * 8 + e;
            hi[i] = bf16_bits(v);
            lo[i] = bf16_bits(v - bf16_value(hi[i]));
          }
  }
  const int levels = d.L + 1;
  std::vector<float> lsp(size_t(levels) * d.bpad, 0.0f);
  for (int l = 0; l < levels; ++l)
    for (int b = 0; b < n; ++b) lsp[size_t(l) * d.bpad + b] = float(local_scaling[size_t(l) * n + b]);
  auto up = [&](const void* host, size_t count) {
    void* ptr = nullptr;
    HIPFMM_CHECK(hipMalloc(&ptr, std::max<size_t>(count, 16)));
    allocations.push_back(ptr);
    HIPFMM_CHECK(hipMemcpy(ptr, host, count, hipMemcpyHostToDevice));
    bytes += count;
    return ptr;
  };
  d.T_hi = up(hi.data(), hi.size() * 2);
  d.T_lo = up(lo.data(), lo.size() * 2);
  d.ls_pad = up(lsp.data(), lsp.size() * 4);
}

}  // namespace hipfmm
