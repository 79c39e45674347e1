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
// `waves_per_eu`, `unroll` and `shade_inline` set FT_WAVES_PER_EU / FT_MARCH_UNROLL /
// FT_SHADE_INLINE (launch-shape variants).
bool rtc_compile(const std::string& source, int kernel, bool counted, int waves_per_eu, int unroll,
                 int shade_inline, std::vector<char>& code, double& compile_ms, std::string& err);

// FT_STRICT_MATH=1 in the environment: every march kernel is runtime-compiled with
// correctly rounded square roots and no multiply-add contraction (the oracle's arithmetic).
bool strict_math();

// A file `name` in the code-object cache directory ("" when there is none).
std::string cache_file(const std::string& name);

}  // namespace fth
