// Far-field kernels of the HipFull path: moment gather, leaf P2M, level-wise
// M2M and L2L, and leaf L2P fused with the near-field combine and the unsort
// to user target order.  All are HBM-bandwidth kernels: one workgroup owns
// one box, reads its inputs once and writes its outputs once, with no atomics.
//
// Reference operations restated here (values computed on the fly instead of
// streamed from stored entry lists):
//   permute_moments / prepare_moments_float   src/cuda_fmm.cu:56-64, src/far_field.cpp:334-365
//   P2M   M_a = (-1)^|a| sum_j sum_{c: a_c>0} m_jc d_j^(a-e_c)/(a-e_c)!
//         src/static_operators.cpp:186-220, src/operators.cpp:16-52
//   M2M / L2L  shared per-(level, class) shift operators
//         src/static_operators.cpp:257-298, src/far_field.cpp:395-446, :555-574
//   L2P   H_c = -sum_{b: b_c>0} L_b dx^(b-e_c)/(b-e_c)!
//         src/static_operators.cpp:300-335
//   combine + unsort  src/cuda_fmm.cu:503-516, src/uniform_fmm.cpp:1439-1452
#include "common.hpp"

namespace hipfmm {

namespace {

constexpr int kThreads = 256;
constexpr int kOutputsPerThread = 8;  // outputs per thread in the coefficient kernels

// d^g/g! for one axis, g <= p, by the recurrence x^i/i! = x^(i-1)/(i-1)! * x / i.
template <class S>
__device__ inline S axis_power(S x, int g) {
  S v = S(1);
  for (int i = 1; i <= g; ++i) v = v * x / S(i);
  return v;
}

template <class S>
__global__ void __launch_bounds__(kThreads) gather_kernel(const S* __restrict__ mu, const int* __restrict__ perm,
                                                          S scale, int ns, int K, S* __restrict__ sm) {
  const int64_t idx = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (idx >= int64_t(ns) * K) return;
  const int j = int(idx / K), k = int(idx % K);
  const S* src = mu + (int64_t(k) * ns + perm[j]) * 3;
  S* dst = sm + int64_t(j) * 3 * K + k;
  dst[0] = src[0] * scale;
  dst[K] = src[1] * scale;
  dst[2 * K] = src[2] * scale;
}

// One workgroup per (leaf, group of kThreads*kOutputsPerThread outputs).
// Outputs o = alpha*K + k.  Sources are processed in chunks whose monomials
// d^g/g! (|g| <= p-1) and moments are staged in LDS.
template <class S>
__global__ void __launch_bounds__(kThreads) p2m_kernel(const DevicePlan d, int K, int chunk) {
  extern __shared__ unsigned char smem_raw[];
  S* mono = reinterpret_cast<S*>(smem_raw);   // chunk x nm
  S* mom = mono + chunk * d.nm;                // chunk x 3 x K
  const int leaf = blockIdx.x;
  const int2 range = d.leaf_src[leaf];
  const int n = d.n, nm = d.nm;
  const int total = n * K;
  const int base = blockIdx.y * kThreads * kOutputsPerThread;
  const int64_t node = level_offset(d.L) + leaf;
  S acc[kOutputsPerThread];
#pragma unroll
  for (int r = 0; r < kOutputsPerThread; ++r) acc[r] = S(0);
  const auto* rel = static_cast<const typename Vec4<S>::type*>(d.src_rel);
  const S* sm = static_cast<const S*>(d.sm);
  for (int j0 = range.x; j0 < range.y; j0 += chunk) {
    const int cnt = min(chunk, range.y - j0);
    for (int w = threadIdx.x; w < cnt * nm; w += kThreads) {
      const int j = w / nm, g = w % nm;
      const int4 e = d.p2m_tab[n + g];  // exponents of basis entry g (second table block)
      const auto r = rel[j0 + j];
      mono[j * nm + g] = axis_power<S>(r.x, e.x) * axis_power<S>(r.y, e.y) * axis_power<S>(r.z, e.z);
    }
    for (int w = threadIdx.x; w < cnt * 3 * K; w += kThreads) mom[w] = sm[int64_t(j0) * 3 * K + w];
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kOutputsPerThread; ++r) {
      const int o = base + r * kThreads + threadIdx.x;
      if (o >= total) break;
      const int alpha = o / K, k = o % K;
      const int4 t = d.p2m_tab[alpha];
      S a = acc[r];
      for (int j = 0; j < cnt; ++j) {
        const S* mj = mom + j * 3 * K + k;
        const S* pj = mono + j * nm;
        if (t.x >= 0) a += pj[t.x] * mj[0];
        if (t.y >= 0) a += pj[t.y] * mj[K];
        if (t.z >= 0) a += pj[t.z] * mj[2 * K];
      }
      acc[r] = a;
    }
    __syncthreads();
  }
  S* M = static_cast<S*>(d.M) + node * total;
#pragma unroll
  for (int r = 0; r < kOutputsPerThread; ++r) {
    const int o = base + r * kThreads + threadIdx.x;
    if (o >= total) break;
    const int sign = d.p2m_tab[o / K].w;
    M[o] = sign > 0 ? acc[r] : -acc[r];
  }
}

// Parents of one level; output (alpha, k) gathers all 8 children through the
// shared CSR pattern of the M2M operator.
template <class S>
__global__ void __launch_bounds__(kThreads) m2m_kernel(const DevicePlan d, int level, int K) {
  const int parent = blockIdx.x;
  const int n = d.n, total = n * K;
  const int64_t pnode = level_offset(level) + parent;
  const int64_t cnode0 = level_offset(level + 1) + int64_t(parent) * 8;
  const S* M = static_cast<const S*>(d.M);
  const S* vals = static_cast<const S*>(d.m2m.values) + int64_t(level + 1) * 8 * d.m2m.nnz;
  S* out = static_cast<S*>(d.M) + pnode * total;
  const int base = blockIdx.y * kThreads * kOutputsPerThread;
#pragma unroll
  for (int r = 0; r < kOutputsPerThread; ++r) {
    const int o = base + r * kThreads + threadIdx.x;
    if (o >= total) break;
    const int alpha = o / K, k = o % K;
    const int e0 = d.m2m.row_offsets[alpha], e1 = d.m2m.row_offsets[alpha + 1];
    S a = S(0);
    for (int c = 0; c < 8; ++c) {
      const S* child = M + (cnode0 + c) * total + k;
      const S* v = vals + c * d.m2m.nnz;
      for (int e = e0; e < e1; ++e) a += v[e] * child[d.m2m.columns[e] * K];
    }
    out[o] = a;
  }
}

// Children of one level: L_child += L2L(L_parent).
template <class S>
__global__ void __launch_bounds__(kThreads) l2l_kernel(const DevicePlan d, int level, int K) {
  const int child = blockIdx.x;
  const int n = d.n, total = n * K;
  const int cls = child & 7;
  const int64_t cnode = level_offset(level) + child;
  const int64_t pnode = level_offset(level - 1) + (child >> 3);
  const S* P = static_cast<const S*>(d.Lc) + pnode * total;
  S* out = static_cast<S*>(d.Lc) + cnode * total;
  const S* v = static_cast<const S*>(d.l2l.values) + (int64_t(level) * 8 + cls) * d.l2l.nnz;
  const int base = blockIdx.y * kThreads * kOutputsPerThread;
#pragma unroll
  for (int r = 0; r < kOutputsPerThread; ++r) {
    const int o = base + r * kThreads + threadIdx.x;
    if (o >= total) break;
    const int beta = o / K, k = o % K;
    S a = S(0);
    for (int e = d.l2l.row_offsets[beta]; e < d.l2l.row_offsets[beta + 1]; ++e)
      a += v[e] * P[d.l2l.columns[e] * K + k];
    out[o] += a;
  }
}

// One workgroup per (target leaf, chunk of kc moment states): far field from
// the leaf local, plus the near field, written to out[k][user target][3].
template <class S>
__global__ void __launch_bounds__(kThreads) l2p_kernel(const DevicePlan d, S* __restrict__ out, int K, int kc,
                                                       int tpass) {
  extern __shared__ unsigned char smem_raw[];
  const int n = d.n, nm = d.nm;
  S* loc = reinterpret_cast<S*>(smem_raw);  // n x kc
  S* mono = loc + n * kc;                    // tpass x nm
  const int leaf = blockIdx.x;
  const int k0 = blockIdx.y * kc;
  const int2 range = d.leaf_tgt[leaf];
  if (range.x == range.y) return;
  const int64_t node = level_offset(d.L) + leaf;
  const S* Lg = static_cast<const S*>(d.Lc) + node * int64_t(n) * K;
  for (int w = threadIdx.x; w < n * kc; w += kThreads) {
    const int b = w / kc, k = k0 + w % kc;
    loc[w] = k < K ? Lg[b * K + k] : S(0);
  }
  const auto* rel = static_cast<const typename Vec4<S>::type*>(d.tgt_rel);
  const S* Hn = static_cast<const S*>(d.Hn);
  for (int t0 = range.x; t0 < range.y; t0 += tpass) {
    const int cnt = min(tpass, range.y - t0);
    __syncthreads();
    for (int w = threadIdx.x; w < cnt * nm; w += kThreads) {
      const int t = w / nm, g = w % nm;
      const int4 e = d.p2m_tab[n + g];
      const auto r = rel[t0 + t];
      mono[t * nm + g] = axis_power<S>(r.x, e.x) * axis_power<S>(r.y, e.y) * axis_power<S>(r.z, e.z);
    }
    __syncthreads();
    const int t = threadIdx.x / kc, kk = threadIdx.x % kc, k = k0 + kk;
    if (t < cnt && k < K) {
      const S* pm = mono + t * nm;
      S hx = S(0), hy = S(0), hz = S(0);
      for (int b = 0; b < n; ++b) {
        const int4 tb = d.p2m_tab[b];
        const S lv = loc[b * kc + kk];
        if (tb.x >= 0) hx -= lv * pm[tb.x];
        if (tb.y >= 0) hy -= lv * pm[tb.y];
        if (tb.z >= 0) hz -= lv * pm[tb.z];
      }
      const int ts = t0 + t;
      const S* hn = Hn + int64_t(ts) * 3 * K + k;
      S* o = out + (int64_t(k) * d.nt + d.tgt_perm[ts]) * 3;
      o[0] = hx + hn[0];
      o[1] = hy + hn[K];
      o[2] = hz + hn[2 * K];
    }
  }
}

template <class S>
__global__ void zero_kernel(S* p, size_t count) {
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < count; i += size_t(gridDim.x) * blockDim.x)
    p[i] = S(0);
}

inline unsigned groups_for(int total) {
  return unsigned((total + kThreads * kOutputsPerThread - 1) / (kThreads * kOutputsPerThread));
}

}  // namespace

template <class S>
void launch_gather(const DevicePlan& d, const S* user_moments, S scale, int K, hipStream_t s) {
  const int64_t work = int64_t(d.ns) * K;
  if (work == 0) return;
  gather_kernel<S><<<unsigned((work + kThreads - 1) / kThreads), kThreads, 0, s>>>(
      user_moments, d.src_perm, scale, d.ns, K, static_cast<S*>(d.sm));
  HIPFMM_CHECK(hipGetLastError());
}

template <class S>
void launch_p2m(const DevicePlan& d, int K, hipStream_t s) {
  const int chunk = sizeof(S) == 4 ? 32 : 16;
  const size_t lds = size_t(chunk) * (d.nm + 3 * K) * sizeof(S);
  const dim3 grid(unsigned(d.leaves), groups_for(d.n * K));
  p2m_kernel<S><<<grid, kThreads, lds, s>>>(d, K, chunk);
  HIPFMM_CHECK(hipGetLastError());
}

template <class S>
void launch_m2m(const DevicePlan& d, int parent_level, int K, hipStream_t s) {
  const dim3 grid(unsigned(int64_t(1) << (3 * parent_level)), groups_for(d.n * K));
  m2m_kernel<S><<<grid, kThreads, 0, s>>>(d, parent_level, K);
  HIPFMM_CHECK(hipGetLastError());
}

template <class S>
void launch_l2l(const DevicePlan& d, int child_level, int K, hipStream_t s) {
  const dim3 grid(unsigned(int64_t(1) << (3 * child_level)), groups_for(d.n * K));
  l2l_kernel<S><<<grid, kThreads, 0, s>>>(d, child_level, K);
  HIPFMM_CHECK(hipGetLastError());
}

template <class S>
void launch_l2p_unsort(const DevicePlan& d, S* fields, int K, hipStream_t s) {
  // Stage at most 32 KB of the leaf local per workgroup.
  int kc = K;
  while (kc > 1 && size_t(d.n) * kc * sizeof(S) > 32768) kc = (kc + 1) / 2;
  // Targets per pass: every thread busy, monomial table within 24 KB.
  const int tpass = std::max(1, std::min(kThreads / kc, int(24576 / (std::max(d.nm, 1) * sizeof(S)))));
  const size_t lds = (size_t(d.n) * kc + size_t(tpass) * d.nm) * sizeof(S);
  const dim3 grid(unsigned(d.leaves), unsigned((K + kc - 1) / kc));
  l2p_kernel<S><<<grid, kThreads, lds, s>>>(d, fields, K, kc, tpass);
  HIPFMM_CHECK(hipGetLastError());
}

template <class S>
void launch_zero(S* ptr, size_t count, hipStream_t s) {
  if (count == 0) return;
  const unsigned blocks = unsigned(std::min<size_t>((count + kThreads - 1) / kThreads, 65536));
  zero_kernel<S><<<blocks, kThreads, 0, s>>>(ptr, count);
  HIPFMM_CHECK(hipGetLastError());
}

#define HIPFMM_INSTANTIATE(S)                                                       \
  template void launch_gather<S>(const DevicePlan&, const S*, S, int, hipStream_t); \
  template void launch_p2m<S>(const DevicePlan&, int, hipStream_t);                \
  template void launch_m2m<S>(const DevicePlan&, int, int, hipStream_t);           \
  template void launch_l2l<S>(const DevicePlan&, int, int, hipStream_t);           \
  template void launch_l2p_unsort<S>(const DevicePlan&, S*, int, hipStream_t);     \
  template void launch_zero<S>(S*, size_t, hipStream_t);
HIPFMM_INSTANTIATE(float)
HIPFMM_INSTANTIATE(double)
#undef HIPFMM_INSTANTIATE

}  // namespace hipfmm
