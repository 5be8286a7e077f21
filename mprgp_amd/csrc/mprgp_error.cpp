// mprgp_error.cpp -- the thread-local last-error string and the mapping of
// C++ exceptions to status codes at the C ABI.
#include "mprgp_error.h"

#include <cstdarg>
#include <cstdio>
#include <exception>
#include <new>
#include <string>

#include "../../include/mprgp_amd.h"

namespace mprgp {

static thread_local std::string g_error;

static void vset(const char *fmt, va_list ap) {
  char buf[1024];
  vsnprintf(buf, sizeof(buf), fmt, ap);
  g_error = buf;
}

void set_error(const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vset(fmt, ap);
  va_end(ap);
}

int fail(int code, const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vset(fmt, ap);
  va_end(ap);
  return code;
}

int on_exception() noexcept {
  try {
    throw;
  } catch (const std::bad_alloc &) {
    return fail(MPRGP_ENOMEM, "host memory allocation failed");
  } catch (const std::exception &e) {
    return fail(MPRGP_EBADINPUT, "%s", e.what());
  } catch (...) {
    return fail(MPRGP_EBADINPUT, "unknown exception");
  }
}

}  // namespace mprgp

extern "C" const char *mprgp_last_error(void) { return mprgp::g_error.c_str(); }
