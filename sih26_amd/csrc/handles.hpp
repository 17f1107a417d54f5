// The opaque handle types of include/sk_gpu.h that more than one translation unit opens.
#pragma once

#include "lp.hpp"
#include "mps.hpp"

struct sk_lp {
  sk::HostLp lp;
  sk::MpsInfo file;  // names and markers when the LP was read from a file
};
