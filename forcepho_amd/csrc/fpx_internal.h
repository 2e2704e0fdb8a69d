// Declarations shared by the translation units of libfpx.so (not part of
// the C ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <string>

#include "forcepho_amd.h"

namespace fpx {

// Kernel configuration used when a caller passes cfg < 0 (8 wavefronts x 4
// pixels per lane).
constexpr int kDefaultConfig = 5;

// Record msg as the text fpx_last_error returns and return code.
int report_error(int code, const char *msg);
// Record a HIP runtime error (where: the failing call) and return FPX_E_HIP.
int report_hip_error(hipError_t e, const char *where);
// Text of a nonzero FpxSlot status word.
std::string status_text(uint32_t status);

// The tile kernel and (finalize_outputs) the output pass for one slot passed
// by value (host struct): no slot table in device memory.  launch_flags:
// FPX_LAUNCH_PERSISTENT, FPX_LAUNCH_SPLIT.  With a flag (a host-visible word;
// needs queue and n_tiles > 0) the launch's last workgroup stores seq into
// it once every result of the launch is visible to the host.
int launch_one(const FpxSlot &slot, int32_t n_tiles, int32_t max_gauss, int32_t cfg,
               int32_t launch_flags, uint32_t *queue, void *stream, bool finalize_outputs = true,
               uint32_t *flag = nullptr, uint32_t seq = 0);
// Pixels per tile of kernel configuration cfg (0 if cfg is invalid).
int config_pixels(int32_t cfg);

}  // namespace fpx
