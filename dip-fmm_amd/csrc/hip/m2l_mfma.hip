// M2L on the CDNA4 matrix cores (v_mfma_f32_32x32x16_bf16).
//
// Same regrouped contraction as the VALU path (m2l_valu.hip):
//   L[t][b] = ls[l][b] * sum_{o in list2 offsets} sum_a T_o[b][a] * Mt[t-o][a],
//   Mt[s][a] = ms[l][a] * M[s][a]
// written as one GEMM per offset class: A = T_o (rows b, depth a), B = the
// gathered source columns (depth a, columns = target box x moment state).
//
// Work decomposition (one launch covers every level 2..L):
//   * a workgroup holds 4 waves and owns 4*NB tiles of 32 output columns of
//     one parity class at one level; each wave owns NB tiles and every row
//     tile of L for them (n rounded up to 32 rows), so each local is written
//     once and no atomics are needed;
//   * the workgroup walks the 189 list-2 offsets of its parity; the class
//     matrix of each offset is streamed through LDS in chunks of CH depth
//     steps, double buffered, and read by all 4 waves (A is the operand with
//     the most reuse: every box of the parity shares it);
//   * B fragments (source multipoles) are prefetched one step ahead from
//     HBM/L2; a wave skips the products of an offset whose source boxes all
//     lie outside the level's grid;
//   * block b -> parity b % 8, so the 8 XCDs each serve one parity class and
//     keep its 189 class matrices in their own L2; consecutive workgroups of
//     one XCD hold spatially consecutive tiles.
//
// Precision: FP32 operands are split x = hi + lo with hi = bf16(x) and
// lo = bf16(x - hi) (16 significant bits), and each product is formed as
// hi*hi + hi*lo + lo*hi with FP32 accumulation on the matrix cores.
//
// Data layout:
//   T_hi/T_lo  [class][depth step][row tile][lane][8] bf16 -- MFMA A-fragment
//              order (lane i: row i%32, depth 8*(i/32)..+7 of the 16-deep
//              step), so a chunk of depth steps is one contiguous block
//   mt_hi/lo   [node][depth step][state][16] bf16 -- a B fragment (column =
//              lane%32 -> state, depth 8*(i/32)..+7) is one 16-byte load and
//              the 32 columns of a step form one contiguous 1 KB block
#include "common.hpp"

namespace hipfmm {

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int kNB = 2;  // 32-column tiles per wave
constexpr int kCH = 1;  // depth steps per LDS chunk

__host__ __device__ inline uint16_t bf16_bits(float x) {
  uint32_t u = __builtin_bit_cast(uint32_t, x);
  if ((u & 0x7f800000u) == 0x7f800000u) return uint16_t(u >> 16);  // inf / nan
  u += 0x7fffu + ((u >> 16) & 1u);                                   // round to nearest even
  return uint16_t(u >> 16);
}
__host__ __device__ inline float bf16_value(uint16_t h) { return __builtin_bit_cast(float, uint32_t(h) << 16); }

__device__ inline int level_of_node(int64_t node) {
  int level = 0;
  while (level_offset(level + 1) <= node) ++level;
  return level;
}

// One workgroup per (box at level >= 2, chunk of kc moment states):
// transpose M[node][a][k] into the fragment layout, scale by ms and split
// into hi/lo bf16 halves; padding depth entries are zero.
__global__ void __launch_bounds__(256) split_kernel(const DevicePlan d, int K, int kc) {
  extern __shared__ float tile[];  // n x kc
  const int n = d.n, AP = d.apad, KS = AP / 16;
  const int k0 = blockIdx.y * kc;
  const int64_t node = level_offset(2) + blockIdx.x;
  const int level = level_of_node(node);
  const float* M = static_cast<const float*>(d.M) + node * int64_t(n) * K;
  for (int w = threadIdx.x; w < n * kc; w += blockDim.x) {
    const int a = w / kc, k = k0 + w % kc;
    tile[w] = k < K ? M[a * K + k] : 0.0f;
  }
  __syncthreads();
  const float* ms = static_cast<const float*>(d.ms) + int64_t(level) * n;
  uint16_t* hi = static_cast<uint16_t*>(d.mt_hi);
  uint16_t* lo = static_cast<uint16_t*>(d.mt_lo);
  const int kn = min(kc, K - k0);
  for (int w = threadIdx.x; w < kn * AP; w += blockDim.x) {
    const int k = w / AP, a = w % AP;  // consecutive threads: consecutive depth
    const float x = a < n ? ms[a] * tile[a * kc + k] : 0.0f;
    const uint16_t h = bf16_bits(x);
    const int64_t at = ((node * KS + a / 16) * K + k0 + k) * 16 + a % 16;
    hi[at] = h;
    lo[at] = bf16_bits(x - bf16_value(h));
  }
}

__device__ inline bf16x8 as_bf16x8(uint4 v) { return __builtin_bit_cast(bf16x8, v); }

struct LevelTable {
  int count;
  int first_level;
  int block_begin[10];
  int blocks_per_parity[10];
  int tiles[10];  // 32-column tiles per parity
  int cols_per_box[10];
  int kchunks[10];
};

template <int MT>
__global__ void __launch_bounds__(256, 1) m2l_mfma_kernel(const DevicePlan d, const LevelTable lt, int K) {
  constexpr int UNITS_PER_KS = MT * 64;  // uint4 per depth step per half (hi or lo)
  __shared__ uint4 abuf[2][2][kCH][UNITS_PER_KS];  // [buffer][hi/lo][step][row tile x lane]
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  int li = lt.count - 1;
  while (li > 0 && int(blockIdx.x) < lt.block_begin[li]) --li;
  const int level = lt.first_level + li;
  const int lb = int(blockIdx.x) - lt.block_begin[li];
  const int parity = lb & 7, wgp = lb >> 3;
  const int px = parity & 1, py = (parity >> 1) & 1, pz = (parity >> 2) & 1;
  const int n = d.n, AP = d.apad, BP = d.bpad, KS = AP / 16;
  const int NCH = (KS + kCH - 1) / kCH;
  const int nbox = 1 << level;
  const int64_t parents = int64_t(1) << (3 * (level - 1));
  const int cpb = lt.cols_per_box[li], kchunks = lt.kchunks[li], tiles = lt.tiles[li];
  const int boxes_per_tile = 32 / cpb;
  const int col = lane & 31, g = lane >> 5;
  // Per owned tile: column -> (target box, state).
  bool valid[kNB];
  int tx[kNB], ty[kNB], tz[kNB], kk[kNB];
#pragma unroll
  for (int nb = 0; nb < kNB; ++nb) {
    const int tile = (wgp * 4 + wave) * kNB + nb;
    const int bg = tile / kchunks, kc = tile % kchunks;
    const int64_t q = int64_t(bg) * boxes_per_tile + col / cpb;
    kk[nb] = kc * cpb + col % cpb;
    valid[nb] = tile < tiles && q < parents && kk[nb] < K;
    int x = 0, y = 0, z = 0;
    if (valid[nb]) demorton3(uint32_t(q), level - 1, x, y, z);
    tx[nb] = 2 * x + px;
    ty[nb] = 2 * y + py;
    tz[nb] = 2 * z + pz;
  }
  const int64_t lvl0 = level_offset(level);
  const int4* offs = d.parity_offsets + parity * 189;
  const uint4* Thi = static_cast<const uint4*>(d.T_hi);
  const uint4* Tlo = static_cast<const uint4*>(d.T_lo);
  const uint4* mhi = static_cast<const uint4*>(d.mt_hi);
  const uint4* mlo = static_cast<const uint4*>(d.mt_lo);
  const uint4 zero4 = {0u, 0u, 0u, 0u};

  // A staging: step s = (offset oi, chunk c); this thread's share of the
  // chunk's A fragments (hi then lo halves) is loaded into registers one step
  // ahead and stored into the free LDS buffer after the current step's
  // products.
  constexpr int STAGE_UNITS = 2 * kCH * UNITS_PER_KS;
  constexpr int PER_THREAD = (STAGE_UNITS + 255) / 256;
  uint4 stage[PER_THREAD];
#define HIPFMM_STAGE_LOAD(step)                                                                     \
  do {                                                                                               \
    const int oi_ = (step) / NCH, c_ = (step) % NCH;                                                 \
    const int cls_ = offs[oi_].w;                                                                    \
    const int ks0_ = c_ * kCH, units_ = min(kCH, KS - ks0_) * UNITS_PER_KS;                         \
    const int64_t base_ = (int64_t(cls_) * KS + ks0_) * UNITS_PER_KS;                                \
    _Pragma("unroll") for (int i = 0; i < PER_THREAD; ++i) {                                        \
      const int u = int(threadIdx.x) + 256 * i;                                                      \
      const int half = u / (kCH * UNITS_PER_KS), r = u % (kCH * UNITS_PER_KS);                       \
      stage[i] = zero4;                                                                              \
      if (u < STAGE_UNITS && r < units_) stage[i] = (half ? Tlo : Thi)[base_ + r];                   \
    }                                                                                                \
  } while (0)
#define HIPFMM_STAGE_STORE(buf)                                                                     \
  do {                                                                                               \
    _Pragma("unroll") for (int i = 0; i < PER_THREAD; ++i) {                                        \
      const int u = int(threadIdx.x) + 256 * i;                                                      \
      if (u < STAGE_UNITS) (&abuf[(buf)][0][0][0])[u] = stage[i];                                    \
    }                                                                                                \
  } while (0)
  // B fragments of one step for this lane's NB columns.
  uint4 bcur[kNB][kCH][2], bnext[kNB][kCH][2];
  auto load_b_next = [&](int s) -> bool {
    const int oi = s / NCH, c = s % NCH;
    const int4 o = offs[oi];
    const int ks0 = c * kCH, nks = min(kCH, KS - ks0);
    bool any = false;
#pragma unroll
    for (int nb = 0; nb < kNB; ++nb) {
      const int sx = tx[nb] - o.x, sy = ty[nb] - o.y, sz = tz[nb] - o.z;
      const bool sv = valid[nb] && sx >= 0 && sy >= 0 && sz >= 0 && sx < nbox && sy < nbox && sz < nbox;
      any |= sv;
      const int64_t snode = sv ? lvl0 + morton3(uint32_t(sx), uint32_t(sy), uint32_t(sz), level) : 0;
#pragma unroll
      for (int j = 0; j < kCH; ++j) {
        const int64_t at = (((snode * KS + ks0 + j) * K + kk[nb]) * 2 + g);
        const bool ok = sv && j < nks;
        bnext[nb][j][0] = ok ? mhi[at] : zero4;
        bnext[nb][j][1] = ok ? mlo[at] : zero4;
      }
    }
    return __any(any) != 0;
  };

  f32x16 acc[kNB][MT];
#pragma unroll
  for (int nb = 0; nb < kNB; ++nb)
#pragma unroll
    for (int m = 0; m < MT; ++m)
#pragma unroll
      for (int r = 0; r < 16; ++r) acc[nb][m][r] = 0.0f;

  const int nsteps = 189 * NCH;
  HIPFMM_STAGE_LOAD(0);
  HIPFMM_STAGE_STORE(0);
  bool active = load_b_next(0);
#pragma unroll
  for (int nb = 0; nb < kNB; ++nb)
#pragma unroll
    for (int j = 0; j < kCH; ++j) {
      bcur[nb][j][0] = bnext[nb][j][0];
      bcur[nb][j][1] = bnext[nb][j][1];
    }
  __syncthreads();
  for (int s = 0; s < nsteps; ++s) {
    const int buf = s & 1;
    const bool more = s + 1 < nsteps;
    if (more) HIPFMM_STAGE_LOAD(s + 1);
    const bool active_next = more ? load_b_next(s + 1) : false;
    if (active) {
      const int nks = min(kCH, KS - (s % NCH) * kCH);
#pragma unroll
      for (int j = 0; j < kCH; ++j) {
        if (j >= nks) break;
#pragma unroll
        for (int m = 0; m < MT; ++m) {
          const bf16x8 ah = as_bf16x8(abuf[buf][0][j][m * 64 + lane]);
          const bf16x8 al = as_bf16x8(abuf[buf][1][j][m * 64 + lane]);
#pragma unroll
          for (int nb = 0; nb < kNB; ++nb) {
            const bf16x8 bh = as_bf16x8(bcur[nb][j][0]);
            const bf16x8 bl = as_bf16x8(bcur[nb][j][1]);
            acc[nb][m] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, bh, acc[nb][m], 0, 0, 0);
            acc[nb][m] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, bl, acc[nb][m], 0, 0, 0);
            acc[nb][m] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al, bh, acc[nb][m], 0, 0, 0);
          }
        }
      }
    }
    if (more) HIPFMM_STAGE_STORE(buf ^ 1);
    __syncthreads();
#pragma unroll
    for (int nb = 0; nb < kNB; ++nb)
#pragma unroll
      for (int j = 0; j < kCH; ++j) {
        bcur[nb][j][0] = bnext[nb][j][0];
        bcur[nb][j][1] = bnext[nb][j][1];
      }
    active = active_next;
  }
#undef HIPFMM_STAGE_LOAD
#undef HIPFMM_STAGE_STORE
  const float* lsp = static_cast<const float*>(d.ls_pad) + int64_t(level) * BP;
  float* Lc = static_cast<float*>(d.Lc);
#pragma unroll
  for (int nb = 0; nb < kNB; ++nb) {
    if (!valid[nb]) continue;
    const int64_t tnode = lvl0 + morton3(uint32_t(tx[nb]), uint32_t(ty[nb]), uint32_t(tz[nb]), level);
    float* out = Lc + tnode * int64_t(n) * K + kk[nb];
#pragma unroll
    for (int m = 0; m < MT; ++m)
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int row = m * 32 + 8 * (r >> 2) + 4 * g + (r & 3);
        if (row < n) out[int64_t(row) * K] = lsp[row] * acc[nb][m][r];
      }
  }
}

LevelTable level_table(const DevicePlan& d, int K) {
  LevelTable lt{};
  lt.first_level = 2;
  lt.count = d.L - 1;
  int begin = 0;
  for (int i = 0; i < lt.count; ++i) {
    const int level = 2 + i;
    int cpb = 1;
    while (cpb < K && cpb < 32) cpb *= 2;
    const int kchunks = (K + cpb - 1) / cpb;
    const int64_t parents = int64_t(1) << (3 * (level - 1));
    const int boxes_per_tile = 32 / cpb;
    const int tiles = int((parents + boxes_per_tile - 1) / boxes_per_tile) * kchunks;
    const int bpp = (tiles + 4 * kNB - 1) / (4 * kNB);
    lt.block_begin[i] = begin;
    lt.blocks_per_parity[i] = bpp;
    lt.tiles[i] = tiles;
    lt.cols_per_box[i] = cpb;
    lt.kchunks[i] = kchunks;
    begin += 8 * bpp;
  }
  lt.block_begin[lt.count] = begin;
  return lt;
}

template <int MT>
void launch_mfma_t(const DevicePlan& d, int K, hipStream_t s) {
  const LevelTable lt = level_table(d, K);
  m2l_mfma_kernel<MT><<<unsigned(lt.block_begin[lt.count]), 256, 0, s>>>(d, lt, K);
  HIPFMM_CHECK(hipGetLastError());
}

}  // namespace

void launch_m2l_split(const DevicePlan& d, int K, int mode, hipStream_t s) {
  (void)mode;
  if (d.L < 2) return;
  const unsigned boxes = unsigned(d.nodes - level_offset(2));
  int kc = K;
  while (kc > 1 && size_t(d.n) * kc * sizeof(float) > 32768) kc = (kc + 1) / 2;
  const dim3 grid(boxes, unsigned((K + kc - 1) / kc));
  split_kernel<<<grid, 256, size_t(d.n) * kc * sizeof(float), s>>>(d, K, kc);
  HIPFMM_CHECK(hipGetLastError());
}

void launch_m2l_mfma(const DevicePlan& d, int K, int mode, hipStream_t s) {
  (void)mode;
  if (d.L < 2) return;
  switch (d.bpad / 32) {
    case 1: launch_mfma_t<1>(d, K, s); break;
    case 2: launch_mfma_t<2>(d, K, s); break;
    case 3: launch_mfma_t<3>(d, K, s); break;
    case 4: launch_mfma_t<4>(d, K, s); break;
    case 5: launch_mfma_t<5>(d, K, s); break;
    case 6: launch_mfma_t<6>(d, K, s); break;
    default: throw invalid_argument("expansion order too large for the MFMA M2L path (p <= 8)");
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
    for (int ks = 0; ks < KS; ++ks)
      for (int m = 0; m < MT; ++m)
        for (int lane = 0; lane < 64; ++lane)
          for (int e = 0; e < 8; ++e) {
            const int row = m * 32 + lane % 32;
            const int dep = ks * 16 + 8 * (lane / 32) + e;
            const float v = (row < n && dep < n) ? float(T[row + size_t(n) * dep]) : 0.0f;
            const size_t i = size_t(c) * per_class + ((size_t(ks) * MT + m) * 64 + lane) * 8 + e;
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
