// zh_jit.cpp -- specialised kernels compiled at run time.
//
// A flat pointwise stage whose micro-op signature is not one of the
// registered kernels of zh_spec.h would otherwise run on the general
// interpreter (k_pointwise), which dispatches every instruction per pixel
// group.  Instead, the first program built with such a signature compiles the
// same kernel bodies (zh_spec_kernels.h) for it with hiprtc -- the kernel is
// `spec_stage<Prog<signature>>`, straight-line code with no dispatch -- and
// every later program with the signature reuses the code object.  This is
// the HIP form of what the reference does when wgpu builds a pipeline for a
// shader at run time (program/encoder.rs:1375-1389): the backend's kernels
// cover any chain of the operations the specialised bodies implement, not a
// fixed list.
//
// The compile needs the package's device headers (next to this library, in
// csrc/) and the ROCm headers (ROCM_PATH, default /opt/rocm).  A signature
// whose compile fails is remembered and its stages stay on the interpreter;
// zh_program_describe says why.
#include <dlfcn.h>
#include <hip/hip_runtime.h>
#include <hip/hiprtc.h>

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "zh_internal.h"
#include "zh_spec.h"

using namespace zh;

namespace {

struct JitEntry {
  std::vector<uint16_t> sig;
  bool compiled = false;
  std::string error;             // why the compile failed (compiled and no shaders)
  zh_shader* stage = nullptr;    // kernel zh_jit_stage (held for the process)
  zh_shader* batch = nullptr;    // kernel zh_jit_batch
  double compile_s = 0.0;
};

std::mutex g_mu;
std::vector<JitEntry>& entries() {
  static std::vector<JitEntry> e;
  return e;
}

// The directory of this library (the package), for its device headers.
std::string package_dir() {
  Dl_info info;
  if (!dladdr((const void*)&zh_version, &info) || !info.dli_fname) return ".";
  std::string p = info.dli_fname;
  const size_t k = p.rfind('/');
  return k == std::string::npos ? std::string(".") : p.substr(0, k);
}

std::string source_for(const std::vector<uint16_t>& sig) {
  std::string list;
  for (size_t i = 0; i < sig.size(); ++i) list += (i ? ", " : "") + std::to_string(sig[i]);
  return "#include \"zh_spec_kernels.h\"\n"
         "typedef zh::spec::Prog<" + list + "> P;\n"
         "extern \"C\" __global__ __launch_bounds__(256) void zh_jit_stage(const zh::StageDesc* d) {\n"
         "  zh::spec::spec_stage<P>(d);\n}\n"
         "extern \"C\" __global__ __launch_bounds__(256) void zh_jit_batch(const zh::BatchDesc* b) {\n"
         "  zh::spec::spec_batch<P>(b);\n}\n";
}

// Compile entry e (called with g_mu held, once).
void compile(JitEntry& e) {
  e.compiled = true;
  const std::string src = source_for(e.sig);
  hiprtcProgram prog = nullptr;
  if (hiprtcCreateProgram(&prog, src.c_str(), "zh_jit.hip", 0, nullptr, nullptr) != HIPRTC_SUCCESS) {
    e.error = "hiprtcCreateProgram failed";
    return;
  }
  const char* rocm = std::getenv("ROCM_PATH");
  const std::string inc_pkg = "-I" + package_dir() + "/csrc";
  const std::string inc_rocm = std::string("-I") + (rocm && *rocm ? rocm : "/opt/rocm") + "/include";
  // The same device flags as the library's own kernels: no packed FP32.
  const char* opts[] = {"--offload-arch=gfx950", "-O3", "-std=c++17", inc_pkg.c_str(), inc_rocm.c_str(),
                        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"};
  const auto t0 = std::chrono::steady_clock::now();
  const hiprtcResult r = hiprtcCompileProgram(prog, (int)(sizeof(opts) / sizeof(opts[0])), opts);
  e.compile_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  if (r != HIPRTC_SUCCESS) {
    size_t n = 0;
    hiprtcGetProgramLogSize(prog, &n);
    std::string log(n, '\0');
    if (n) hiprtcGetProgramLog(prog, &log[0]);
    e.error = std::string("hiprtc: ") + hiprtcGetErrorString(r) + ": " + log.substr(0, 400);
    hiprtcDestroyProgram(&prog);
    return;
  }
  size_t n = 0;
  hiprtcGetCodeSize(prog, &n);
  std::vector<char> code(n);
  hiprtcGetCode(prog, code.data());
  hiprtcDestroyProgram(&prog);
  if (zh_shader_create(code.data(), code.size(), "zh_jit_stage", &e.stage) != ZH_OK ||
      zh_shader_create(code.data(), code.size(), "zh_jit_batch", &e.batch) != ZH_OK)
    e.error = "code object rejected";
}

}  // namespace

namespace zh {

bool jit_supported(const uint16_t* sig, int n) {
  if (n <= 0 || n > kMaxInstr) return false;
  for (int i = 0; i < n; ++i) {
    const uint32_t op = sig[i] & 0xffu;
    if (op != OP_LOAD && op != OP_STORE && op != OP_MAT3 && op != OP_CBRT && op != OP_CUBE && op != OP_CLAMP01 &&
        op != OP_PREMUL && op != OP_UNPREMUL && op != OP_OVER)
      return false;
  }
  return true;
}

int jit_signature(const uint16_t* sig, int n) {
  std::lock_guard<std::mutex> lock(g_mu);
  std::vector<JitEntry>& es = entries();
  for (size_t i = 0; i < es.size(); ++i)
    if ((int)es[i].sig.size() == n && std::memcmp(es[i].sig.data(), sig, sizeof(uint16_t) * (size_t)n) == 0)
      return (int)i;
  JitEntry e;
  e.sig.assign(sig, sig + n);
  es.push_back(std::move(e));
  return (int)es.size() - 1;
}

zh_status jit_kernels(int id, int device, hipFunction_t* stage, hipFunction_t* batch, std::string* why) {
  zh_shader *s = nullptr, *b = nullptr;
  {
    std::lock_guard<std::mutex> lock(g_mu);
    std::vector<JitEntry>& es = entries();
    if (id < 0 || id >= (int)es.size()) return fail(ZH_ERR_INVALID_ARGUMENT, "unknown run-time kernel");
    JitEntry& e = es[(size_t)id];
    if (!e.compiled) compile(e);
    if (!e.stage) {
      if (why) *why = e.error;
      return fail(ZH_ERR_UNSUPPORTED, "run-time kernel unavailable: " + e.error);
    }
    s = e.stage;
    b = e.batch;
  }
  zh_status st = shader_function(s, device, stage);
  if (st == ZH_OK) st = shader_function(b, device, batch);
  if (st != ZH_OK && why) *why = zh_last_error();
  return st;
}

double jit_compile_seconds(int id) {
  std::lock_guard<std::mutex> lock(g_mu);
  const std::vector<JitEntry>& es = entries();
  return id >= 0 && id < (int)es.size() ? es[(size_t)id].compile_s : 0.0;
}

}  // namespace zh
