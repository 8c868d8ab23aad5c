// Error plumbing shared by every translation unit of the library: a status code is returned
// from each C entry point and a thread-local message describes the last failure. The reference
// prints and calls exit() on any CUDA error (inc/_cuda.hxx:221-235); this library never exits.
#pragma once
#include <cstdint>
#include <cstdio>
#include <string>

#include "rlc.h"

namespace rlc {

void set_error(const std::string &msg);
void clear_error();

inline rlc_status fail(rlc_status s, const std::string &msg) {
  set_error(msg);
  return s;
}

}  // namespace rlc
