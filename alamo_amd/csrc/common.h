// common.h -- error state shared by the translation units of libalamo_elastic.so.
#pragma once
#include <string>

namespace alamo_detail {
// Message of the last failed call on this thread (alamo_el_last_error).
extern thread_local std::string g_err;
inline int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}
}  // namespace alamo_detail
