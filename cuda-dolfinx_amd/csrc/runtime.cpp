// HIP runtime layer: hipRTC compilation with a content-hashed code-object
// cache, module loading.  Replaces CUDA::compile_cuda_cpp_to_ptx and
// CUDA::Module (cpp/cudolfinx/common/CUDA.cpp:79-436); the cache is keyed on
// a hash of the full source and options rather than a factory name
// (CUDAFormIntegral.cpp:2499-2545 keys on the name only).
#include <hip/hiprtc.h>

#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <sstream>
#include <sys/stat.h>
#include <unistd.h>

#include "mf_internal.h"

namespace mf {

uint64_t fnv1a(const std::string& s, uint64_t h) {
  for (unsigned char c : s) {
    h ^= c;
    h *= 1099511628211ULL;
  }
  return h;
}

std::string default_cache_dir() {
  if (const char* e = std::getenv("MIFEM_CACHE_DIR")) return e;
  if (const char* x = std::getenv("XDG_CACHE_HOME")) return std::string(x) + "/mifem";
  if (const char* h = std::getenv("HOME")) return std::string(h) + "/.cache/mifem";
  return "";
}

static void mkdirs(const std::string& path) {
  std::string cur;
  std::stringstream ss(path);
  std::string part;
  if (!path.empty() && path[0] == '/') cur = "/";
  while (std::getline(ss, part, '/')) {
    if (part.empty()) continue;
    cur += part + "/";
    ::mkdir(cur.c_str(), 0755);
  }
}

static bool read_file(const std::string& p, std::vector<char>& out) {
  std::ifstream f(p, std::ios::binary);
  if (!f) return false;
  out.assign(std::istreambuf_iterator<char>(f), std::istreambuf_iterator<char>());
  return !out.empty();
}

std::vector<char> rtc_compile(const std::string& src, const std::string& name,
                              const std::vector<std::string>& opts, const std::string& cache_dir,
                              bool* from_cache) {
  int maj = 0, min = 0;
  hiprtcVersion(&maj, &min);
  std::string key = src;
  for (auto& o : opts) key += "\n" + o;
  key += "\nhiprtc " + std::to_string(maj) + "." + std::to_string(min);
  char hex[32];
  std::snprintf(hex, sizeof hex, "%016llx", static_cast<unsigned long long>(fnv1a(key)));
  std::string path;
  if (!cache_dir.empty()) {
    path = cache_dir + "/" + name + "_" + hex + ".co";
    std::vector<char> code;
    if (read_file(path, code)) {
      if (from_cache) *from_cache = true;
      return code;
    }
  }
  if (from_cache) *from_cache = false;
  hiprtcProgram prog;
  if (hiprtcCreateProgram(&prog, src.c_str(), (name + ".hip").c_str(), 0, nullptr, nullptr) !=
      HIPRTC_SUCCESS)
    MF_THROW(MF_ECOMPILE, "hiprtcCreateProgram failed");
  std::vector<const char*> o;
  for (auto& s : opts) o.push_back(s.c_str());
  hiprtcResult r = hiprtcCompileProgram(prog, static_cast<int>(o.size()), o.data());
  size_t log_size = 0;
  hiprtcGetProgramLogSize(prog, &log_size);
  std::string log(log_size, '\0');
  if (log_size) hiprtcGetProgramLog(prog, &log[0]);
  if (r != HIPRTC_SUCCESS) {
    hiprtcDestroyProgram(&prog);
    // dump the failing source for inspection
    std::string dump = "/tmp/mifem_failed_" + std::string(hex) + ".hip";
    std::ofstream(dump) << src;
    MF_THROW(MF_ECOMPILE, std::string("hipRTC compile of ") + name + " failed (source in " + dump +
                              "):\n" + log);
  }
  size_t size = 0;
  hiprtcGetCodeSize(prog, &size);
  std::vector<char> code(size);
  hiprtcGetCode(prog, code.data());
  hiprtcDestroyProgram(&prog);
  if (!path.empty()) {
    mkdirs(cache_dir);
    std::string tmp = path + ".tmp" + std::to_string(::getpid());
    {
      std::ofstream f(tmp, std::ios::binary);
      f.write(code.data(), static_cast<std::streamsize>(code.size()));
    }
    std::rename(tmp.c_str(), path.c_str());
  }
  return code;
}

void Module::load(const std::vector<char>& code) {
  if (mod) {
    (void)hipModuleUnload(mod);
    mod = nullptr;
  }
  MF_HIP(hipModuleLoadData(&mod, code.data()));
}

hipFunction_t Module::get(const std::string& name) const {
  hipFunction_t f = nullptr;
  hipError_t e = hipModuleGetFunction(&f, mod, name.c_str());
  if (e != hipSuccess) MF_THROW(MF_ECOMPILE, "kernel " + name + " not found in module");
  return f;
}

}  // namespace mf
