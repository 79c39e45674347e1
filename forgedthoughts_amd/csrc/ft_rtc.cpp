// ft_rtc.cpp — compiles emitted scene sources with hiprtc for gfx950.
//
// The kernel headers are embedded in the library (ft_rtc_headers.inc, generated from the
// csrc headers by the build), so a scene can be specialised without any file next to the
// library.  Code objects are cached in memory for the process and, when a cache directory
// is writable, on disk keyed by a hash of the complete source and options.
#include "ft_rtc.h"

#include <hip/hiprtc.h>

#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <mutex>
#include <sstream>
#include <string>
#include <sys/stat.h>
#include <time.h>
#include <unistd.h>
#include <unordered_map>
#include <vector>

#include "ft_rtc_headers.inc"  // kRtcHeaderCount, kRtcHeaderNames[], kRtcHeaderSources[]

namespace fth {
namespace {

// FNV-1a, 64-bit
uint64_t fnv1a(const std::string& s, uint64_t h = 1469598103934665603ull) {
  for (unsigned char c : s) {
    h ^= c;
    h *= 1099511628211ull;
  }
  return h;
}

std::string cache_dir() {
  if (const char* d = std::getenv("FT_RTC_CACHE")) return d;
  if (const char* x = std::getenv("XDG_CACHE_HOME")) return std::string(x) + "/forgedthoughts_amd";
  if (const char* h = std::getenv("HOME")) return std::string(h) + "/.cache/forgedthoughts_amd";
  return "";
}

std::mutex g_mu;
std::unordered_map<uint64_t, std::vector<char>> g_mem;

const std::vector<std::string>& options() {
  // the same code-generation flags as the precompiled kernels: packed FP32 instructions off
  static const std::vector<std::string> o = {
      "--offload-arch=gfx950", "-O3", "-std=c++17", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"};
  return o;
}

}  // namespace

bool strict_math() {
  const char* v = std::getenv("FT_STRICT_MATH");
  return v && v[0] == '1';
}

std::string cache_file(const std::string& name) {
  const std::string dir = cache_dir();
  return dir.empty() ? std::string() : dir + "/" + name;
}

bool rtc_compile(const std::string& source, int kernel, bool counted, int waves_per_eu, int unroll,
                 int shade_inline, std::vector<char>& code, double& compile_ms, std::string& err) {
  compile_ms = 0.0;
  std::vector<std::string> opts = options();
  opts.push_back("-DFT_RTC_KERNEL=" + std::to_string(kernel));
  if (counted) opts.push_back("-DFT_COUNTED=1");
  if (strict_math()) {
    opts.push_back("-ffp-contract=off");
    opts.push_back("-DFT_STRICT_MATH=1");
  }
  if (waves_per_eu != 4) opts.push_back("-DFT_WAVES_PER_EU=" + std::to_string(waves_per_eu));
  if (unroll != 1) opts.push_back("-DFT_MARCH_UNROLL=" + std::to_string(unroll));
  if (shade_inline) opts.push_back("-DFT_SHADE_INLINE=1");
  std::string key_text = source;
  for (const auto& o : opts) key_text += "\n" + o;
  for (int i = 0; i < kRtcHeaderCount; ++i) key_text += kRtcHeaderSources[i];
  const uint64_t key = fnv1a(key_text);
  {
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_mem.find(key);
    if (it != g_mem.end()) {
      code = it->second;
      return true;
    }
  }
  char name[32];
  std::snprintf(name, sizeof(name), "%016llx.co", (unsigned long long)key);
  const std::string dir = cache_dir();
  if (!dir.empty() && std::getenv("FT_RTC_NO_DISK_CACHE") == nullptr) {
    std::ifstream f(dir + "/" + name, std::ios::binary);
    if (f) {
      std::vector<char> buf((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
      if (buf.size() > 4 && buf[0] == 0x7f && buf[1] == 'E' && buf[2] == 'L' && buf[3] == 'F') {
        std::lock_guard<std::mutex> lk(g_mu);
        g_mem[key] = buf;
        code = std::move(buf);
        return true;
      }
    }
  }

  hiprtcProgram prog;
  if (hiprtcCreateProgram(&prog, source.c_str(), "ft_scene.hip", kRtcHeaderCount, kRtcHeaderSources,
                          kRtcHeaderNames) != HIPRTC_SUCCESS) {
    err = "hiprtcCreateProgram failed";
    return false;
  }
  std::vector<const char*> argv;
  for (const auto& o : opts) argv.push_back(o.c_str());
  struct timespec t0, t1;
  clock_gettime(CLOCK_MONOTONIC, &t0);
  hiprtcResult r = hiprtcCompileProgram(prog, (int)argv.size(), argv.data());
  clock_gettime(CLOCK_MONOTONIC, &t1);
  compile_ms = (t1.tv_sec - t0.tv_sec) * 1e3 + (t1.tv_nsec - t0.tv_nsec) * 1e-6;
  if (r != HIPRTC_SUCCESS) {
    size_t n = 0;
    hiprtcGetProgramLogSize(prog, &n);
    std::string log(n, '\0');
    if (n) hiprtcGetProgramLog(prog, &log[0]);
    hiprtcDestroyProgram(&prog);
    err = std::string("scene specialisation failed to compile: ") + hiprtcGetErrorString(r) + "\n" + log.substr(0, 4000);
    return false;
  }
  size_t n = 0;
  hiprtcGetCodeSize(prog, &n);
  code.assign(n, 0);
  hiprtcGetCode(prog, code.data());
  hiprtcDestroyProgram(&prog);
  {
    std::lock_guard<std::mutex> lk(g_mu);
    g_mem[key] = code;
  }
  if (!dir.empty() && std::getenv("FT_RTC_NO_DISK_CACHE") == nullptr) {
    // best effort: create the directory chain, write to a temporary name, then rename
    std::string path = dir[0] == '/' ? "" : ".";
    std::stringstream ss(dir);
    std::string part;
    while (std::getline(ss, part, '/')) {
      if (part.empty()) continue;
      path += "/" + part;
      mkdir(path.c_str(), 0755);
    }
    const std::string tmp = dir + "/" + name + ".tmp" + std::to_string((long)getpid());
    {
      std::ofstream f(tmp, std::ios::binary);
      if (f) f.write(code.data(), (std::streamsize)code.size());
    }
    std::rename(tmp.c_str(), (dir + "/" + name).c_str());
  }
  return true;
}

}  // namespace fth
