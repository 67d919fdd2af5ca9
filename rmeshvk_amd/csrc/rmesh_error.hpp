// Last-error message shared by the library's C entry points (host code only).
#pragma once
#include <string>

namespace rmesh {
// Sets the message rmesh_last_error() returns on the calling thread.
void set_last_error(const std::string& msg);
}  // namespace rmesh
