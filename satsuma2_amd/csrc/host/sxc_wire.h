// sxc_wire.h -- the master/slave wire records, byte-compatible with the reference on LP64.
#pragma once
#include <stddef.h>
#include <stdint.h>

#include "../../../include/sxc.h"

namespace sxc_host {

// t_pair (analysis/WorkQueue.h:17-22): one block of target chunks x query chunks, both
// ranges inclusive.  28 bytes, host byte order.
struct WirePair {
  int32_t target_from, target_to, query_from, query_to;
  uint8_t fast;
  uint8_t pad_[3];
  int32_t slave_id;
  uint8_t status;
  uint8_t pad2_[3];
};
static_assert(sizeof(WirePair) == 28, "t_pair is 28 bytes");
static_assert(offsetof(WirePair, fast) == 16 && offsetof(WirePair, slave_id) == 20 &&
                  offsetof(WirePair, status) == 24,
              "t_pair layout");

// t_result (analysis/WorkQueue.h:23-33) is sxc_result.
static_assert(sizeof(sxc_result) == 72, "t_result is 72 bytes");

// Default master port (analysis/WorkQueue.h:42).
constexpr int kDefaultPort = 3491;

}  // namespace sxc_host
