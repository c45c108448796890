// Host-side error plumbing shared by the C-ABI translation units.
#pragma once

namespace dfm {
// Records a thread-local message for dfm_last_error() and returns `code`.
int set_error(int code, const char* what, const char* detail = nullptr);
}  // namespace dfm
