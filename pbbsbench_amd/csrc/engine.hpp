// Internal helpers shared by the library's translation units.
#pragma once

namespace pbbs {
// Records a message for pbbs_last_error() and returns `code`.
int fail(int code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
void clear_error();
const char* last_error();
}  // namespace pbbs
