// zh_jit.cpp -- specialised kernels compiled at run time.
//
// A flat pointwise stage whose micro-op signature is not one of the
// registered kernels of zh_spec.h would otherwise run on the general
// interpreter (k_pointwise), which dispatches every instruction per pixel
// group.  Instead, the first program built with such a signature compiles the
// same kernel bodies (zh_spec_kernels.h) for it with hiprtc -- the kernel is
// `spec_stage<Prog<signature>>`, straight-line code with no dispatch -- and
// every later program with the signature reuses the code object.  Resample
// stages whose affine layer-stack form (source form, format, matrix, post
// program) has no registered warp kernel get theirs the same way from the
// warp bodies (zh_warp_kernels.h).  This is
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
#include <initializer_list>
#include <mutex>
#include <string>
#include <vector>

#include "zh_internal.h"
#include "zh_spec.h"

using namespace zh;

namespace {

// A pointwise signature (warp = false) or an affine layer-stack form: source
// form, source format, mat3 flag and the post program's signature.
struct JitEntry {
  bool warp = false;
  bool pair = false;  // two warp forms in one grid: the first in src..sig, the second in src2..sig2
  uint32_t src = 0, fmt = 0, mat = 0;
  std::vector<uint16_t> sig;
  bool sep = false, sep2 = false;
  uint32_t src2 = 0, fmt2 = 0, mat2 = 0;
  std::vector<uint16_t> sig2;
  bool compiled = false;
  std::string error;  // why the compile failed (compiled and no kernels)
  // Kernels held for the process: pointwise {stage, batch}; warp {stage,
  // stage of separable maps, batch, batch of separable maps} and, for
  // one-source forms, the LDS-staged {stage, batch} of separable maps.
  std::vector<zh_shader*> kernels;
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

const char* const kSpecKernels[] = {"zh_jit_stage", "zh_jit_batch"};
const char* const kWarpKernels[] = {"zh_jit_warp",       "zh_jit_warp_sep",    "zh_jit_warp_batch",
                                    "zh_jit_warp_batch_sep", "zh_jit_warp_staged", "zh_jit_warp_batch_staged"};

const char* const kPairKernels[] = {"zh_jit_warp_pair"};

int kernel_count(const JitEntry& e) { return e.pair ? 1 : !e.warp ? 2 : e.src == WSRC_ONE ? 6 : 4; }

std::string prog_list(const std::vector<uint16_t>& sig) {
  std::string list;
  for (size_t i = 0; i < sig.size(); ++i) list += (i ? ", " : "") + std::to_string(sig[i]);
  return list;
}

std::string warp_type(uint32_t src, uint32_t fmt, uint32_t mat, const std::vector<uint16_t>& sig, bool sep) {
  return "zh::spec::Warp<" + std::to_string(src) + "u, " + std::to_string(fmt) + "u, " + (mat ? "true" : "false") +
         ", zh::spec::Prog<" + prog_list(sig) + ">, " + (sep ? "true" : "false") + ">";
}

std::string source_for(const JitEntry& e) {
  if (e.pair)
    return "#include \"zh_warp_kernels.h\"\n"
           "typedef " + warp_type(e.src, e.fmt, e.mat, e.sig, e.sep) + " VA;\n"
           "typedef " + warp_type(e.src2, e.fmt2, e.mat2, e.sig2, e.sep2) + " VB;\n"
           "extern \"C\" __global__ __launch_bounds__(256) void zh_jit_warp_pair(const zh::BatchDesc* a, "
           "const zh::BatchDesc* b) { zh::spec::warp_pair<VA, VB>(a, b); }\n";
  std::string list;
  for (size_t i = 0; i < e.sig.size(); ++i) list += (i ? ", " : "") + std::to_string(e.sig[i]);
  const std::string k = "extern \"C\" __global__ __launch_bounds__(256) void ";
  if (!e.warp)
    return "#include \"zh_spec_kernels.h\"\n"
           "typedef zh::spec::Prog<" + list + "> P;\n" +
           k + "zh_jit_stage(const zh::StageDesc* d) { zh::spec::spec_stage<P>(d); }\n" +
           k + "zh_jit_batch(const zh::BatchDesc* b) { zh::spec::spec_batch<P>(b); }\n";
  const std::string form = std::to_string(e.src) + "u, " + std::to_string(e.fmt) + "u, " + (e.mat ? "true" : "false") +
                           ", zh::spec::Prog<" + list + ">, ";
  std::string src = "#include \"zh_warp_kernels.h\"\n"
                    "typedef zh::spec::Warp<" + form + "false> V;\n"
                    "typedef zh::spec::Warp<" + form + "true> VS;\n" +
                    k + "zh_jit_warp(const zh::StageDesc* d) { zh::spec::warp_stage<V>(d); }\n" +
                    k + "zh_jit_warp_sep(const zh::StageDesc* d) { zh::spec::warp_stage<VS>(d); }\n" +
                    k + "zh_jit_warp_batch(const zh::BatchDesc* b) { zh::spec::warp_batch<V>(b); }\n" +
                    k + "zh_jit_warp_batch_sep(const zh::BatchDesc* b) { zh::spec::warp_batch<VS>(b); }\n";
  if (e.src == WSRC_ONE)
    src += k + "zh_jit_warp_staged(const zh::StageDesc* d) { zh::spec::warp_staged_stage<VS>(d); }\n" +
           k + "zh_jit_warp_batch_staged(const zh::BatchDesc* b) { zh::spec::warp_staged_batch<VS>(b); }\n";
  return src;
}

// Compile entry e (called with g_mu held, once).
void compile(JitEntry& e) {
  e.compiled = true;
  const std::string src = source_for(e);
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
  const char* const* names = e.pair ? kPairKernels : e.warp ? kWarpKernels : kSpecKernels;
  const int count = kernel_count(e);
  for (int i = 0; i < count; ++i) {
    zh_shader* k = nullptr;
    if (zh_shader_create(code.data(), code.size(), names[i], &k) != ZH_OK) {
      e.error = "code object rejected";
      for (zh_shader* x : e.kernels) zh_shader_release(x);
      e.kernels.clear();
      return;
    }
    e.kernels.push_back(k);
  }
}

int find_or_add(const JitEntry& want) {
  std::vector<JitEntry>& es = entries();
  for (size_t i = 0; i < es.size(); ++i) {
    const JitEntry& e = es[i];
    if (e.warp == want.warp && e.pair == want.pair && e.src == want.src && e.fmt == want.fmt && e.mat == want.mat &&
        e.sig == want.sig && e.sep == want.sep && e.src2 == want.src2 && e.fmt2 == want.fmt2 &&
        e.mat2 == want.mat2 && e.sig2 == want.sig2 && e.sep2 == want.sep2)
      return (int)i;
  }
  es.push_back(want);
  return (int)es.size() - 1;
}

// The kernels `which` of entry `id` on `device`, compiling the entry first.
zh_status kernels_of(int id, bool warp, std::initializer_list<int> which, int device, hipFunction_t* out,
                     std::string* why, bool pair = false) {
  std::vector<zh_shader*> ks;
  {
    std::lock_guard<std::mutex> lock(g_mu);
    std::vector<JitEntry>& es = entries();
    if (id < 0 || id >= (int)es.size() || es[(size_t)id].warp != warp || es[(size_t)id].pair != pair)
      return fail(ZH_ERR_INVALID_ARGUMENT, "unknown run-time kernel");
    JitEntry& e = es[(size_t)id];
    if (!e.compiled) compile(e);
    if (e.kernels.empty()) {
      if (why) *why = e.error;
      return fail(ZH_ERR_UNSUPPORTED, "run-time kernel unavailable: " + e.error);
    }
    ks = e.kernels;
  }
  int k = 0;
  for (int w : which) {
    zh_status st = shader_function(ks[(size_t)w], device, &out[k++]);
    if (st != ZH_OK) {
      if (why) *why = zh_last_error();
      return st;
    }
  }
  return ZH_OK;
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
  JitEntry e;
  e.sig.assign(sig, sig + n);
  return find_or_add(e);
}

zh_status jit_kernels(int id, int device, hipFunction_t* stage, hipFunction_t* batch, std::string* why) {
  hipFunction_t f[2] = {nullptr, nullptr};
  zh_status st = kernels_of(id, false, {0, 1}, device, f, why);
  *stage = f[0];
  *batch = f[1];
  return st;
}

bool jit_warp_supported(const uint16_t* post, int n) {
  if (n <= 0 || n > kMaxInstr) return false;
  for (int i = 0; i < n; ++i) {
    const uint32_t op = post[i] & 0xffu;
    if (op != OP_STORE && op != OP_MAT3 && op != OP_CBRT && op != OP_CUBE && op != OP_CLAMP01 && op != OP_PREMUL &&
        op != OP_UNPREMUL)
      return false;
  }
  return (post[n - 1] & 0xffu) == OP_STORE;
}

int jit_warp_form(uint32_t src, uint32_t fmt, uint32_t mat, const uint16_t* post, int n) {
  std::lock_guard<std::mutex> lock(g_mu);
  JitEntry e;
  e.warp = true;
  e.src = src;
  e.fmt = fmt;
  e.mat = mat;
  e.sig.assign(post, post + n);
  return find_or_add(e);
}

zh_status jit_warp_kernels(int id, bool sep, bool staged, int device, hipFunction_t* stage, hipFunction_t* batch,
                           std::string* why) {
  hipFunction_t f[2] = {nullptr, nullptr};
  if (staged && !jit_warp_has_staged(id)) return fail(ZH_ERR_INVALID_ARGUMENT, "run-time warp form without staged tiles");
  zh_status st = staged ? kernels_of(id, true, {4, 5}, device, f, why)
                 : sep  ? kernels_of(id, true, {1, 3}, device, f, why)
                        : kernels_of(id, true, {0, 2}, device, f, why);
  *stage = f[0];
  *batch = f[1];
  return st;
}

bool warp_form_of(int variant, WarpForm* f) {
  if (variant < 0) return false;
  if (variant >= kJitWarpBase) {
    const int id = (variant - kJitWarpBase) / 2;
    std::lock_guard<std::mutex> lock(g_mu);
    const std::vector<JitEntry>& es = entries();
    if (id >= (int)es.size() || !es[(size_t)id].warp || es[(size_t)id].pair) return false;
    const JitEntry& e = es[(size_t)id];
    f->src = e.src;
    f->fmt = e.fmt;
    f->mat = e.mat;
    f->post = e.sig;
    f->sep = (variant & 1) != 0;
    return true;
  }
  const uint16_t* post = nullptr;
  int n = 0;
  if (!zh_warp_form_of(variant, &f->src, &f->fmt, &f->mat, &post, &n)) return false;
  f->post.assign(post, post + n);
  f->sep = (variant & 1) != 0;
  return true;
}

int jit_warp_pair(const WarpForm& a, const WarpForm& b) {
  std::lock_guard<std::mutex> lock(g_mu);
  JitEntry e;
  e.warp = true;
  e.pair = true;
  e.src = a.src;
  e.fmt = a.fmt;
  e.mat = a.mat;
  e.sig = a.post;
  e.sep = a.sep;
  e.src2 = b.src;
  e.fmt2 = b.fmt;
  e.mat2 = b.mat;
  e.sig2 = b.post;
  e.sep2 = b.sep;
  return find_or_add(e);
}

zh_status jit_pair_kernel(int id, int device, hipFunction_t* fn, std::string* why) {
  return kernels_of(id, true, {0}, device, fn, why, true);
}

bool jit_warp_has_staged(int id) {
  std::lock_guard<std::mutex> lock(g_mu);
  const std::vector<JitEntry>& es = entries();
  return id >= 0 && id < (int)es.size() && es[(size_t)id].warp && !es[(size_t)id].pair && es[(size_t)id].src == WSRC_ONE;
}

double jit_compile_seconds(int id) {
  std::lock_guard<std::mutex> lock(g_mu);
  const std::vector<JitEntry>& es = entries();
  return id >= 0 && id < (int)es.size() ? es[(size_t)id].compile_s : 0.0;
}

}  // namespace zh
