// MPS reader (reference include/sankhya/io.hpp:35-46; src/io/mps_reader.cpp, mps_bounds.cpp).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "lp.hpp"

namespace sk {

// What a file says beyond the LP itself.
struct MpsInfo {
  std::string name;
  std::vector<std::string> row_names, col_names;
  int32_t integer_columns = 0;  // columns marked integer (an LP solve relaxes them)
  bool fixed_format = false;    // the free-format reading failed and fixed columns succeeded
};

// Reads `path` into `lp`; on failure returns false with "path:line: reason" in `error`.
bool read_mps(const std::string& path, HostLp* lp, MpsInfo* info, std::string* error);

}  // namespace sk
