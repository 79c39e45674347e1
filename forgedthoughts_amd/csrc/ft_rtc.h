// ft_rtc.h — hiprtc compilation of emitted scene sources.
#pragma once

#include <string>
#include <vector>

namespace fth {

// Kernel selectors (FT_RTC_KERNEL in ft_kernels.h): one code object per kernel.
enum RtcKernel { kRtcDepth = 0, kRtcWhitted16 = 1, kRtcWhitted48 = 2, kRtcWhitted128 = 3, kRtcKernelCount = 4 };

// Compiles `source` for gfx950 with only kernel `kernel` instantiated.  Returns the code
// object in `code`; `compile_ms` is 0 when it came from a cache.
// `counted` builds the *_counted variant (FT_COUNTED).
bool rtc_compile(const std::string& source, int kernel, bool counted, std::vector<char>& code, double& compile_ms,
                 std::string& err);

}  // namespace fth
