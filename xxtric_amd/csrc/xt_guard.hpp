// xt_guard.hpp -- keeps C++ exceptions inside libxtric: every int-returning
// entry point of the C ABI runs its body through xt_guarded, so a failed host
// allocation or thread start becomes a status code instead of unwinding into
// a C caller.
#pragma once

#include <exception>
#include <new>

#include "../../include/xtric.h"

template <typename F>
inline int xt_guarded(F&& body) noexcept {
  try {
    return body();
  } catch (const std::bad_alloc&) {
    return XT_ERR_NOMEM;
  } catch (...) {
    return XT_ERR_INTERNAL;
  }
}
