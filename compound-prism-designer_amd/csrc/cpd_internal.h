// cpd_internal.h -- declarations shared by the library's translation units.
#pragma once

#include <stdarg.h>

#include "cpd.h"

namespace cpd {

// Record a message for cpd_last_error() and return st.
cpd_status fail(cpd_status st, const char *fmt, ...) __attribute__((format(printf, 2, 3)));
void clear_error();

}  // namespace cpd
