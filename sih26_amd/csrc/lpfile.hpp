// Writing an LP out: free-format MPS (read back to the same arrays by read_mps) and a binary
// instance cache with a checksum.
#pragma once

#include <string>

#include "lp.hpp"
#include "mps.hpp"

namespace sk {

// Free-format MPS of `lp`; names from `info` when it holds a full set of blank-free, distinct
// names, else R<i> / C<j>. Values carry 17 significant digits, so read_mps gives back the same
// arrays (entries below 1e-11 in magnitude excepted: readers drop them; a two-sided row whose
// bound difference rounds comes back within one rounding of its range).
bool write_mps(const HostLp& lp, const MpsInfo* info, const std::string& path, std::string* error);

// The LP's arrays as stored in memory (host byte order) after an 8-byte magic, followed by an
// FNV-1a checksum; load_lp checks sizes, checksum and the model before accepting a file.
bool save_lp(const HostLp& lp, const std::string& path, std::string* error);
bool load_lp(const std::string& path, HostLp* lp, std::string* error);

}  // namespace sk
