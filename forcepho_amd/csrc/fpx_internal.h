// Declarations shared by the translation units of libfpx.so (not part of
// the C ABI).
#pragma once

#include <hip/hip_runtime.h>

namespace fpx {

// Record msg as the text fpx_last_error returns and return code.
int report_error(int code, const char *msg);
// Record a HIP runtime error (where: the failing call) and return FPX_E_HIP.
int report_hip_error(hipError_t e, const char *where);

}  // namespace fpx
