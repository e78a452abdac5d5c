// Per-thread error text and the exception-to-status translation used by every C ABI entry point
// (host only: no HIP calls, so the loader can also be built on its own, e.g. with sanitizers).
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <new>
#include <stdexcept>
#include "common.h"

namespace cn {

static thread_local char g_err[kErrorTextSize];

cn_status set_error(cn_status st, const char* fmt, ...) {
  char buf[kErrorTextSize];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  std::memcpy(g_err, buf, sizeof buf);
  return st;
}

void clear_error() { g_err[0] = 0; }

const char* last_error_text() { return g_err; }

cn_status status_of_exception() noexcept {
  try {
    throw;
  } catch (const std::bad_alloc&) {
    return set_error(CN_ERR_OUT_OF_MEMORY, "host allocation failed");
  } catch (const std::length_error&) {
    return set_error(CN_ERR_OUT_OF_MEMORY, "host allocation too large");
  } catch (const std::exception& ex) {
    return set_error(CN_ERR_STATE, "internal error: %s", ex.what());
  } catch (...) {
    return set_error(CN_ERR_STATE, "internal error");
  }
}

}  // namespace cn

extern "C" const char* cn_last_error(void) { return cn::last_error_text(); }
