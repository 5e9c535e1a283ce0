#include "operator_tables.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <stdexcept>

#include "operators.hpp"

namespace cdfmm_hip {

namespace {

constexpr char kMagic[8] = {'H', 'I', 'P', 'F', 'M', 'M', 'O', 'T'};

uint64_t fnv1a(const unsigned char* p, size_t n) {
  uint64_t h = 1469598103934665603ull;
  for (size_t i = 0; i < n; ++i) {
    h ^= p[i];
    h *= 1099511628211ull;
  }
  return h;
}

struct Writer {
  std::vector<unsigned char> out;
  template <class T>
  void put(const T& v) {
    const auto* b = reinterpret_cast<const unsigned char*>(&v);
    out.insert(out.end(), b, b + sizeof(T));
  }
  template <class T>
  void put_all(const std::vector<T>& v) {
    const auto* b = reinterpret_cast<const unsigned char*>(v.data());
    out.insert(out.end(), b, b + v.size() * sizeof(T));
  }
};

struct Reader {
  const std::vector<unsigned char>& in;
  size_t at = 0;
  bool ok = true;
  template <class T>
  T get() {
    T v{};
    if (at + sizeof(T) > in.size()) {
      ok = false;
      return v;
    }
    std::memcpy(&v, in.data() + at, sizeof(T));
    at += sizeof(T);
    return v;
  }
  template <class T>
  std::vector<T> get_all(size_t count) {
    if (!ok || count > (in.size() - at) / sizeof(T)) {
      ok = false;
      return {};
    }
    std::vector<T> v(count);
    std::memcpy(v.data(), in.data() + at, count * sizeof(T));
    at += count * sizeof(T);
    return v;
  }
};

}  // namespace

OperatorTables build_operator_tables(const Basis& basis, int L, double W) {
  OperatorTables t;
  t.order = basis.order();
  t.max_level = L;
  t.frame_root_width = W;
  const int n = basis.size();
  // M2L classes: every list-2 offset (src/uniform_fmm.cpp:502-625).
  std::vector<std::array<int, 3>> offsets;
  if (L >= 2) offsets = list2_offsets();
  t.class_offsets.resize(3 * offsets.size());
  for (size_t k = 0; k < offsets.size(); ++k)
    for (int a = 0; a < 3; ++a) t.class_offsets[3 * k + a] = offsets[k][a];
  t.m2l.resize(offsets.size() * size_t(n) * n);
#pragma omp parallel for schedule(dynamic, 4)
  for (std::ptrdiff_t k = 0; k < static_cast<std::ptrdiff_t>(offsets.size()); ++k) {
    const auto m = m2l_class_matrix(basis, offsets[size_t(k)]);
    std::copy(m.begin(), m.end(), t.m2l.begin() + k * static_cast<std::ptrdiff_t>(m.size()));
  }
  // Level scalings ms = w^-|alpha|, ls = w^-(|beta|+1) (src/uniform_fmm.cpp:529-566).
  t.ms.resize(size_t(L + 1) * n);
  t.ls.resize(t.ms.size());
  std::vector<double> iwp(size_t(basis.order() + 2), 1.0);
  for (int level = 0; level <= L; ++level) {
    const double w = W / double(1 << level);
    for (int d = 1; d <= basis.order() + 1; ++d) iwp[d] = iwp[d - 1] / w;
    for (int i = 0; i < n; ++i) {
      t.ms[size_t(level) * n + i] = iwp[degree(basis[i])];
      t.ls[size_t(level) * n + i] = iwp[degree(basis[i]) + 1];
    }
  }
  // Shift operators per level and child class (src/uniform_fmm.cpp:457-500):
  // child centre = parent centre +- h per axis, h = half the child width.
  t.m2m.resize(size_t(L) * 8 * n * n);
  t.l2l.resize(t.m2m.size());
  for (int level = 1; level <= L; ++level) {
    const double h = 0.5 * W / double(1 << level);
    for (int cls = 0; cls < 8; ++cls) {
      const std::array<double, 3> off{(cls & 1) ? h : -h, (cls & 2) ? h : -h, (cls & 4) ? h : -h};
      const auto up = m2m_dense(basis, {-off[0], -off[1], -off[2]});
      const auto down = l2l_dense(basis, off);
      const size_t at = (size_t(level - 1) * 8 + cls) * n * n;
      std::copy(up.begin(), up.end(), t.m2m.begin() + static_cast<std::ptrdiff_t>(at));
      std::copy(down.begin(), down.end(), t.l2l.begin() + static_cast<std::ptrdiff_t>(at));
    }
  }
  return t;
}

std::vector<unsigned char> serialise_operator_tables(const OperatorTables& t) {
  const int n = (t.order + 1) * (t.order + 2) * (t.order + 3) / 6;
  Writer w;
  w.out.insert(w.out.end(), kMagic, kMagic + 8);
  w.put(kOperatorTablesVersion);
  w.put(uint32_t(t.order));
  w.put(uint32_t(t.max_level));
  w.put(uint32_t(n));
  w.put(uint32_t(t.class_offsets.size() / 3));
  w.put(t.frame_root_width);
  w.put_all(t.class_offsets);
  w.put_all(t.m2l);
  w.put_all(t.ms);
  w.put_all(t.ls);
  w.put_all(t.m2m);
  w.put_all(t.l2l);
  w.put(fnv1a(w.out.data() + 8, w.out.size() - 8));
  return std::move(w.out);
}

void save_operator_tables(const OperatorTables& t, const std::string& path) {
  const auto blob = serialise_operator_tables(t);
  const std::string tmp = path + ".tmp";
  {
    std::ofstream f(tmp, std::ios::binary | std::ios::trunc);
    if (!f) throw std::runtime_error("cannot write operator tables to " + tmp);
    f.write(reinterpret_cast<const char*>(blob.data()), static_cast<std::streamsize>(blob.size()));
    if (!f) throw std::runtime_error("cannot write operator tables to " + tmp);
  }
  if (std::rename(tmp.c_str(), path.c_str()) != 0) throw std::runtime_error("cannot rename " + tmp + " to " + path);
}

std::optional<OperatorTables> parse_operator_tables(const std::vector<unsigned char>& blob, int order, int L,
                                                    double W, std::string* why) {
  auto fail = [&](const char* reason) -> std::optional<OperatorTables> {
    if (why) *why = reason;
    return std::nullopt;
  };
  if (blob.size() < 8 + 4 * 5 + 8 + 8 || std::memcmp(blob.data(), kMagic, 8) != 0) return fail("not an operator-table blob");
  uint64_t stored = 0;
  std::memcpy(&stored, blob.data() + blob.size() - 8, 8);
  if (fnv1a(blob.data() + 8, blob.size() - 16) != stored) return fail("checksum mismatch");
  Reader r{blob, 8};
  const uint32_t version = r.get<uint32_t>();
  if (version != kOperatorTablesVersion) return fail("format version differs");
  OperatorTables t;
  t.order = int(r.get<uint32_t>());
  t.max_level = int(r.get<uint32_t>());
  const uint32_t n = r.get<uint32_t>();
  const uint32_t classes = r.get<uint32_t>();
  t.frame_root_width = r.get<double>();
  if (t.order != order || t.max_level != L || t.frame_root_width != W) return fail("built for another order, depth or frame");
  if (int(n) != (order + 1) * (order + 2) * (order + 3) / 6 || (L >= 2 ? classes != 316u : classes != 0u))
    return fail("inconsistent sizes");
  const size_t nn = size_t(n) * n;
  t.class_offsets = r.get_all<int>(3 * size_t(classes));
  t.m2l = r.get_all<double>(classes * nn);
  t.ms = r.get_all<double>(size_t(L + 1) * n);
  t.ls = r.get_all<double>(size_t(L + 1) * n);
  t.m2m = r.get_all<double>(size_t(L) * 8 * nn);
  t.l2l = r.get_all<double>(size_t(L) * 8 * nn);
  if (!r.ok || r.at + 8 != blob.size()) return fail("truncated or oversized blob");
  return t;
}

std::optional<OperatorTables> load_operator_tables(const std::string& path, int order, int L, double W,
                                                   std::string* why) {
  std::ifstream f(path, std::ios::binary);
  if (!f) {
    if (why) *why = "no such file";
    return std::nullopt;
  }
  std::vector<unsigned char> blob((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
  return parse_operator_tables(blob, order, L, W, why);
}

}  // namespace cdfmm_hip
