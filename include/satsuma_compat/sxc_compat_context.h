/* sxc_compat_context.h -- the process-wide context behind the reference-compatible classes
 * (CCSignal, CrossCorrelation, SeqAnalyzer, GetMatchProbability, PrintMatch).
 *
 * The reference classes are self-contained CPU objects; their replacements forward to the
 * C ABI (sxc.h) and share one context on HIP device $SXC_DEVICE (default 0), created on first
 * use.  A failing call throws std::runtime_error with sxc_last_error(): the reference has no
 * error returns, and results must never be silently missing. */
#ifndef SXC_COMPAT_CONTEXT_H
#define SXC_COMPAT_CONTEXT_H

#include <cstdlib>
#include <mutex>
#include <stdexcept>
#include <string>

#include "../sxc.h"

namespace sxc_compat {

inline sxc_ctx* context() {
  static std::once_flag once;
  static sxc_ctx* ctx = nullptr;
  static std::string err;
  std::call_once(once, [] {
    const char* d = std::getenv("SXC_DEVICE");
    if (sxc_create(&ctx, d ? std::atoi(d) : 0) != SXC_OK) {
      err = sxc_last_error();
      ctx = nullptr;
    }
  });
  if (!ctx) throw std::runtime_error("sxc: no GPU context: " + err);
  return ctx;
}

inline void check(int rc, const char* what) {
  if (rc != SXC_OK) throw std::runtime_error(std::string("sxc: ") + what + ": " + sxc_last_error());
}

}  // namespace sxc_compat

#endif
