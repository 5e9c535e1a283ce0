// Serialised translation tables of a uniform FMM plan: the M2L class
// matrices, level scalings and M2M/L2L shift operators that the host builds
// before the device upload (src/uniform_fmm.cpp:502-625 classes and
// scalings, :457-500 shift operators; src/static_operators.cpp:159-298).
// They depend on the expansion order, the leaf level and the width of the
// root box in the operator frame only -- 1 for FP32 plans, whose frame is
// normalised by the root width, the physical width for FP64 plans -- so one
// blob serves every geometry with the same three keys.
//
// Blob (little-endian):
//   char[8]  "HIPFMMOT"
//   u32      format version (kOperatorTablesVersion)
//   u32      order, max_level, n = C(order+3, 3), class_count
//   f64      frame root width
//   i32      class_offsets[3 * class_count]
//   f64      m2l[class_count * n * n]        column-major T[beta + n*alpha]
//   f64      ms[(max_level+1) * n], ls[(max_level+1) * n]
//   f64      m2m[max_level * 8 * n * n], l2l[max_level * 8 * n * n]  row-major (out x in)
//   u64      FNV-1a of every byte after the magic
#pragma once

#include <array>
#include <cstdint>
#include <optional>
#include <string>
#include <vector>

#include "basis.hpp"

namespace cdfmm_hip {

constexpr uint32_t kOperatorTablesVersion = 1;

struct OperatorTables {
  int order = 0;
  int max_level = 0;
  double frame_root_width = 1.0;
  std::vector<int> class_offsets;  // 3 per class
  std::vector<double> m2l;
  std::vector<double> ms, ls;
  std::vector<double> m2m, l2l;
};

// Build the tables (FP64, OpenMP over classes).
OperatorTables build_operator_tables(const Basis& basis, int max_level, double frame_root_width);

// Serialise / write atomically (temporary file, then rename).
std::vector<unsigned char> serialise_operator_tables(const OperatorTables& t);
void save_operator_tables(const OperatorTables& t, const std::string& path);

// Parse a blob; nullopt with the reason in *why when it is malformed, of
// another format version, fails its checksum, or was built for other keys.
std::optional<OperatorTables> parse_operator_tables(const std::vector<unsigned char>& blob, int order, int max_level,
                                                    double frame_root_width, std::string* why);
std::optional<OperatorTables> load_operator_tables(const std::string& path, int order, int max_level,
                                                   double frame_root_width, std::string* why);

}  // namespace cdfmm_hip
