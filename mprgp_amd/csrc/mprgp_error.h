// mprgp_error.h -- status codes and the last-error string of the C ABI
// (internal header; no HIP dependency, so host-only units such as the QP
// file reader build with any C++ compiler).
#pragma once

namespace mprgp {

// Sets the calling thread's last-error message (mprgp_last_error).
void set_error(const char *fmt, ...);
// set_error, then returns `code`.
int fail(int code, const char *fmt, ...);

// No C++ exception crosses the C ABI: every extern "C" entry point runs its
// body between MPRGP_TRY and MPRGP_CATCH, and on_exception() turns what was
// thrown into a status code (std::bad_alloc: MPRGP_ENOMEM; any other
// exception: MPRGP_EBADINPUT with its message) and the last-error string.
int on_exception() noexcept;

}  // namespace mprgp

#define MPRGP_TRY try {
#define MPRGP_CATCH \
  }                 \
  catch (...) { return ::mprgp::on_exception(); }
#define MPRGP_CATCH_VOID \
  }                      \
  catch (...) { (void)::mprgp::on_exception(); }
