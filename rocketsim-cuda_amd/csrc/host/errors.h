// rsim_last_error() for the host-only sources of librsim.so (the message store lives in rsim.cpp).
#pragma once

namespace rs {
namespace host {
// Sets the calling thread's rsim_last_error() message and returns code.
int set_error(int code, const char* what);
}  // namespace host
}  // namespace rs
