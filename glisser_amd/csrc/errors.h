// Thread-local last-error text shared by every C entry point of the library.
#pragma once

#include <string>

namespace glx {

// Records `msg` as the calling thread's last error and returns `code`.
int set_error(int code, const std::string& msg);
void clear_error();

}  // namespace glx
