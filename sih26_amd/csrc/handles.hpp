// The opaque handle types of include/sk_gpu.h that more than one translation unit opens.
#pragma once

#include "lp.hpp"

struct sk_lp {
  sk::HostLp lp;
};
