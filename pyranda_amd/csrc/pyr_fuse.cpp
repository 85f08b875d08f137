// Runtime-compiled elementwise kernels for the pointwise algebra of EOM lines.
//
// The reference evaluates every pointwise expression of an equation with numpy
// temporaries (pyranda/pyrandaEq.py:39, pyranda/pyranda.py:376-427). Here each
// maximal pointwise subtree of an equation becomes one generated HIP kernel
// that reads its input fields once and writes its result once. The generator
// lives in pyranda_amd/fuse.py; this unit compiles the generated source with
// hiprtc for gfx950, loads it and launches it on a stream.
#include <hip/hip_runtime.h>
#include <hip/hiprtc.h>

#include <cstdint>
#include <cstring>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/pyranda_amd.h"

namespace {

thread_local std::string g_fuse_err;

int ffail(const std::string& m) {
  g_fuse_err = m;
  return 1;
}

struct Loaded {
  hipModule_t module = nullptr;
  hipFunction_t fn = nullptr;
};

std::mutex g_mu;
std::unordered_map<std::string, Loaded> g_cache;  // keyed by source text

}  // namespace

// Kernel argument block shared with the generated code (fuse.py writes the
// same layout): element count, pointers, double scalars, and whether every
// pointer is 16-byte aligned (the elementwise loops then use vector accesses).
struct PyrPwArgs {
  long long n;
  void* p[PYR_PW_MAX_PTRS];
  double s[PYR_PW_MAX_SCALARS];
  int vec;
};

extern "C" {

const char* pyr_pw_last_error(void) { return g_fuse_err.c_str(); }

int pyr_pw_compile(const char* source, const char* kernel_name, void** fn) {
  if (!source || !kernel_name || !fn) return ffail("null argument to pyr_pw_compile");
  std::lock_guard<std::mutex> lk(g_mu);
  const std::string key = std::string(kernel_name) + "\n" + source;
  auto it = g_cache.find(key);
  if (it != g_cache.end()) {
    *fn = (void*)it->second.fn;
    return 0;
  }
  hiprtcProgram prog;
  if (hiprtcCreateProgram(&prog, source, kernel_name, 0, nullptr, nullptr) != HIPRTC_SUCCESS)
    return ffail("hiprtcCreateProgram failed");
  const char* opts[] = {"--offload-arch=gfx950", "-O3", "-ffp-contract=off", "-std=c++17",
                        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"};
  const hiprtcResult cr = hiprtcCompileProgram(prog, (int)(sizeof(opts) / sizeof(opts[0])), opts);
  if (cr != HIPRTC_SUCCESS) {
    size_t ls = 0;
    hiprtcGetProgramLogSize(prog, &ls);
    std::string log(ls, '\0');
    if (ls) hiprtcGetProgramLog(prog, &log[0]);
    hiprtcDestroyProgram(&prog);
    return ffail("hiprtc compile failed: " + log);
  }
  size_t cs = 0;
  hiprtcGetCodeSize(prog, &cs);
  std::vector<char> code(cs);
  hiprtcGetCode(prog, code.data());
  hiprtcDestroyProgram(&prog);
  Loaded ld;
  if (hipModuleLoadData(&ld.module, code.data()) != hipSuccess) return ffail("hipModuleLoadData failed");
  if (hipModuleGetFunction(&ld.fn, ld.module, kernel_name) != hipSuccess)
    return ffail("hipModuleGetFunction failed");
  g_cache[key] = ld;
  *fn = (void*)ld.fn;
  return 0;
}

// Host-only check that a generated source compiles (no device needed).
int pyr_pw_check_source(const char* source, const char* kernel_name) {
  hiprtcProgram prog;
  if (hiprtcCreateProgram(&prog, source, kernel_name, 0, nullptr, nullptr) != HIPRTC_SUCCESS)
    return ffail("hiprtcCreateProgram failed");
  const char* opts[] = {"--offload-arch=gfx950", "-O3", "-ffp-contract=off", "-std=c++17",
                        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"};
  const hiprtcResult cr = hiprtcCompileProgram(prog, (int)(sizeof(opts) / sizeof(opts[0])), opts);
  if (cr != HIPRTC_SUCCESS) {
    size_t ls = 0;
    hiprtcGetProgramLogSize(prog, &ls);
    std::string log(ls, '\0');
    if (ls) hiprtcGetProgramLog(prog, &log[0]);
    hiprtcDestroyProgram(&prog);
    return ffail("hiprtc compile failed: " + log);
  }
  hiprtcDestroyProgram(&prog);
  return 0;
}

static int pw_launch(void* fn, void* stream, long long n, long long blocks, int nptr,
                     void* const* ptrs, int nscal, const double* scal) {
  if (!fn) return ffail("null kernel");
  if (nptr < 0 || nptr > PYR_PW_MAX_PTRS || nscal < 0 || nscal > PYR_PW_MAX_SCALARS)
    return ffail("too many kernel arguments");
  if (n <= 0) return 0;
  if (blocks < 1 || blocks > 65535) return ffail("bad block count");
  PyrPwArgs a;
  std::memset(&a, 0, sizeof(a));
  a.n = n;
  a.vec = 1;
  for (int i = 0; i < nptr; ++i) {
    a.p[i] = ptrs[i];
    if ((uintptr_t)ptrs[i] % 16 != 0) a.vec = 0;
  }
  for (int i = 0; i < nscal; ++i) a.s[i] = scal[i];
  void* params[] = {&a};
  if (hipModuleLaunchKernel((hipFunction_t)fn, (unsigned)blocks, 1, 1, PYR_PW_BLOCK, 1, 1, 0,
                            (hipStream_t)stream, params, nullptr) != hipSuccess)
    return ffail("hipModuleLaunchKernel failed");
  return 0;
}

int pyr_pw_launch(void* fn, void* stream, long long n, int nptr, void* const* ptrs, int nscal,
                  const double* scal) {
  // one group of four elements per thread (the grid-stride loops of the
  // generated kernels cover what a grid of at most 65535 workgroups does not):
  // measured faster than a capped, persistent-style grid (8192 workgroups:
  // +0.36 ms per 256^3 Taylor-Green step; 1024: +0.9 ms)
  long long blocks = (n + PYR_PW_BLOCK * 4 - 1) / (PYR_PW_BLOCK * 4);
  if (blocks > 65535) blocks = 65535;
  if (blocks < 1) blocks = 1;
  return pw_launch(fn, stream, n, blocks, nptr, ptrs, nscal, scal);
}

int pyr_pw_launch_blocks(void* fn, void* stream, long long n, int blocks, int nptr,
                         void* const* ptrs, int nscal, const double* scal) {
  return pw_launch(fn, stream, n, blocks, nptr, ptrs, nscal, scal);
}

}  // extern "C"
