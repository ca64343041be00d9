// The device context of the calling thread, shared by the drop-in entry
// points (dropin.cpp) and load_functions (plugin.cpp): one context per host
// thread, on the device rin::set_device() chose (default 0).
#pragma once
#include "rin.h"

namespace rin {
namespace detail {
rin_ctx* thread_context();  // throws std::runtime_error when no device is usable
}  // namespace detail
}  // namespace rin
