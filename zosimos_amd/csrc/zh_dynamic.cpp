// zh_dynamic.cpp -- user stage kernels: the HIP code-object form of the
// reference's ShaderSource (lib/zosimos/src/command/dynamic.rs:7-51).  A
// zh_shader holds a copy of the code object and the kernel's name; each
// device loads the code object once, on the first program built with it
// there, and unloads it when the last reference goes.
#include <hip/hip_runtime.h>

#include <atomic>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "zh_internal.h"

using namespace zh;

struct zh_shader {
  std::atomic<int> refs{1};
  std::vector<uint8_t> code;
  std::string kernel;
  std::mutex mu;
  struct Loaded {
    int device;
    hipModule_t module;
    hipFunction_t fn;
  };
  std::vector<Loaded> loaded;
};

static void shader_drop(zh_shader* s) {
  if (!s || s->refs.fetch_sub(1) != 1) return;
  int cur = 0;
  const bool have_cur = hipGetDevice(&cur) == hipSuccess;
  for (const zh_shader::Loaded& l : s->loaded) {
    hipSetDevice(l.device);
    hipModuleUnload(l.module);
  }
  if (have_cur && !s->loaded.empty()) hipSetDevice(cur);
  delete s;
}

namespace zh {

ShaderRef::ShaderRef(zh_shader* s) : s_(s) {
  if (s_) s_->refs.fetch_add(1);
}
ShaderRef::ShaderRef(const ShaderRef& o) : s_(o.s_) {
  if (s_) s_->refs.fetch_add(1);
}
ShaderRef& ShaderRef::operator=(const ShaderRef& o) {
  if (o.s_) o.s_->refs.fetch_add(1);
  shader_drop(s_);
  s_ = o.s_;
  return *this;
}
ShaderRef::~ShaderRef() { shader_drop(s_); }

zh_status shader_function(zh_shader* s, int device, hipFunction_t* fn) {
  std::lock_guard<std::mutex> lock(s->mu);
  for (const zh_shader::Loaded& l : s->loaded)
    if (l.device == device) {
      *fn = l.fn;
      return ZH_OK;
    }
  hipModule_t m = nullptr;
  hipError_t e = hipModuleLoadData(&m, s->code.data());
  if (e != hipSuccess) {
    (void)hipGetLastError();  // the runtime keeps the error for the next launch's check otherwise
    return fail(ZH_ERR_HIP, std::string("hipModuleLoadData of the shader's code object failed: ") + hipGetErrorString(e));
  }
  hipFunction_t f = nullptr;
  e = hipModuleGetFunction(&f, m, s->kernel.c_str());
  if (e != hipSuccess) {
    (void)hipGetLastError();
    hipModuleUnload(m);
    return fail(ZH_ERR_INVALID_ARGUMENT, "the shader's code object has no kernel named '" + s->kernel + "'");
  }
  s->loaded.push_back({device, m, f});
  *fn = f;
  return ZH_OK;
}

bool shader_loaded(zh_shader* s, int device) {
  std::lock_guard<std::mutex> lock(s->mu);
  for (const zh_shader::Loaded& l : s->loaded)
    if (l.device == device) return true;
  return false;
}

const char* shader_kernel_name(zh_shader* s) { return s->kernel.c_str(); }

}  // namespace zh

extern "C" {

zh_status zh_shader_create(const void* code_object, size_t len, const char* kernel, zh_shader** out) {
  if (!code_object || !kernel || !out) return fail(ZH_ERR_INVALID_ARGUMENT, "null argument");
  if (len < 4 || !*kernel) return fail(ZH_ERR_INVALID_ARGUMENT, "empty code object or kernel name");
  zh_shader* s = new zh_shader();
  const uint8_t* b = static_cast<const uint8_t*>(code_object);
  s->code.assign(b, b + len);
  s->kernel = kernel;
  *out = s;
  return ZH_OK;
}

zh_status zh_shader_retain(zh_shader* shader) {
  if (!shader) return fail(ZH_ERR_INVALID_ARGUMENT, "null shader");
  shader->refs.fetch_add(1);
  return ZH_OK;
}

zh_status zh_shader_release(zh_shader* shader) {
  shader_drop(shader);
  return ZH_OK;
}

}  // extern "C"
