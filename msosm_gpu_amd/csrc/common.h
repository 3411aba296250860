// common.h -- error reporting shared by the C ABI translation units.
#pragma once

namespace msosm {
// Records the message for msosm_last_error() (per thread) and returns code.
int set_error(int code, const char *fmt, ...) __attribute__((format(printf, 2, 3)));
} // namespace msosm
