// io.hpp -- the parsed edge list shared by the text readers (io.hip) and the
// binary caches (serial.hip).
#pragma once
#include <cstdint>
#include <functional>
#include <vector>

#include "../../include/suminc.h"

struct suminc_edges {
  uint32_t n = 0;
  int wkind = SUMINC_W_NONE;
  uint64_t skipped = 0;     // lines that did not parse (skipped, as the reference's loaders do)
  std::vector<uint64_t> off;
  std::vector<uint32_t> col;
  std::vector<uint32_t> w;  // weight bits (u32 or f32), empty without weights
};

namespace suminc {
// Runs f for a call without a context: returns SUMINC_OK or the error's code, and
// keeps the message for suminc_last_error(NULL).
int host_call(const std::function<void()>& f);
}  // namespace suminc
