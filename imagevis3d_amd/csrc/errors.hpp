// errors.hpp -- the last-error channel of the C ABI (tvk_last_error).
#pragma once
#include <string>

namespace tvk {
// Record the message returned by tvk_last_error() on this thread.
void set_last_error(const std::string& msg);
// The message last recorded on this thread.
const std::string& last_error();
}  // namespace tvk
