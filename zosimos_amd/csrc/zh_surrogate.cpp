// zh_surrogate.cpp -- learning, judging and installing the colour surrogate
// of a program's pointwise stages (zosimos_hip.h, zh_mlp.h).
//
// For an eligible stage [load RGBA8 sRGB, colour micro-ops, store RGBA8 sRGB]:
//   1. samples: 2^20 points uniform in the encoded input cube plus 2^16 on
//      its faces (seeded), run through the stage's own micro-op program by the
//      general pointwise kernel with RGBA32F sRGB slots (the plain chain on
//      continuous inputs);
//   2. training: Adam on the mean squared error, minibatches from a seeded
//      permutation, cosine learning-rate decay, all on the GPU;
//   3. check: the trained network on a lattice^3 grid of the input cube
//      (matrix-core inference kernel) against the stage's program evaluated
//      on the host in FP64, compared in Oklab; the stage switches to the
//      network only if the max Delta-E stays below the bound.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

#include "zh_internal.h"
#include "zh_mlp.h"

extern "C" {
hipError_t zh_launch_pointwise(const zh::StageDesc* d_dev, uint32_t width, uint32_t height, int flat, int fast,
                               uint32_t cus, hipStream_t stream);
}

using namespace zh;

namespace {

zh_status hip_err(hipError_t e, const char* what) {
  return fail(ZH_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}
#define ZS_HIP(call)                                  \
  do {                                                \
    hipError_t _e = (call);                           \
    if (_e != hipSuccess) return hip_err(_e, #call);  \
  } while (0)

// splitmix64 (Steele, Lea & Flood 2014): seeded, platform independent.
struct Rng {
  uint64_t s;
  uint64_t next() {
    uint64_t z = (s += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
  }
  double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
};

double eotf(double v) { return v <= 0.04045 ? v / 12.92 : std::pow((v + 0.055) / 1.055, 2.4); }
double oetf(double l) { return l <= 0.0031308 ? l * 12.92 : 1.055 * std::pow(l, 1.0 / 2.4) - 0.055; }

// Oklab (oklab.frag:14-24, Ottosson 2020) of linear RGB through `to_xyz`.
void oklab(const Mat3& to_xyz, const double rgb[3], double lab[3]) {
  static const double M1[9] = {0.8189330101, 0.3618667424, -0.1288597137, 0.0329845436, 0.9293118715,
                               0.0361456387, 0.0482003018, 0.2643662691, 0.6338517070};
  static const double M2[9] = {0.2104542553, 0.7936177850, -0.0040720468, 1.9779984951, -2.4285922050,
                               0.4505937099, 0.0259040371, 0.7827717662, -0.8086757660};
  double xyz[3], lms[3];
  for (int i = 0; i < 3; ++i) xyz[i] = to_xyz[3 * i] * rgb[0] + to_xyz[3 * i + 1] * rgb[1] + to_xyz[3 * i + 2] * rgb[2];
  for (int i = 0; i < 3; ++i) lms[i] = std::cbrt(M1[3 * i] * xyz[0] + M1[3 * i + 1] * xyz[1] + M1[3 * i + 2] * xyz[2]);
  for (int i = 0; i < 3; ++i) lab[i] = M2[3 * i] * lms[0] + M2[3 * i + 1] * lms[1] + M2[3 * i + 2] * lms[2];
}

// Delta-E OK between two encoded (sRGB transfer) colours.
double delta_e(const Mat3& to_xyz, const double a[3], const double b[3]) {
  double la[3], lb[3], ra[3], rb[3];
  for (int i = 0; i < 3; ++i) {
    ra[i] = eotf(std::min(std::max(a[i], 0.0), 1.0));
    rb[i] = eotf(std::min(std::max(b[i], 0.0), 1.0));
  }
  oklab(to_xyz, ra, la);
  oklab(to_xyz, rb, lb);
  return std::sqrt((la[0] - lb[0]) * (la[0] - lb[0]) + (la[1] - lb[1]) * (la[1] - lb[1]) +
                   (la[2] - lb[2]) * (la[2] - lb[2]));
}

// [load RGBA8 sRGB, (mat3 | cbrt | cube | clamp01)*, store RGBA8 sRGB], flat.
bool eligible(const StagePlan& sp) {
  const StageDesc& d = sp.desc;
  if (d.kind != STAGE_POINTWISE || !d.flat || d.n_instr < 2) return false;
  const Instr& ld = d.code[0];
  const Instr& st = d.code[d.n_instr - 1];
  if (ld.op != OP_LOAD || ld.b != 0 || st.op != OP_STORE) return false;
  if (d.slot[ld.a].fmt != FMT_U8_SRGB_RGBA || d.slot[st.a].fmt != FMT_U8_SRGB_RGBA) return false;
  for (uint32_t i = 1; i + 1 < d.n_instr; ++i) {
    const uint8_t o = d.code[i].op;
    if (o != OP_MAT3 && o != OP_CBRT && o != OP_CUBE && o != OP_CLAMP01) return false;
  }
  return true;
}

// The stage's program in FP64 on one encoded input colour.
void eval_fp64(const StageDesc& d, const double in[3], double out[3]) {
  double v[3] = {eotf(in[0]), eotf(in[1]), eotf(in[2])};
  for (uint32_t i = 1; i + 1 < d.n_instr; ++i) {
    const Instr& c = d.code[i];
    switch (c.op) {
      case OP_MAT3: {
        const float* m = d.params + c.arg;
        const double r[3] = {m[0] * v[0] + m[1] * v[1] + m[2] * v[2], m[3] * v[0] + m[4] * v[1] + m[5] * v[2],
                             m[6] * v[0] + m[7] * v[1] + m[8] * v[2]};
        v[0] = r[0];
        v[1] = r[1];
        v[2] = r[2];
        break;
      }
      case OP_CBRT: for (double& x : v) x = std::cbrt(x); break;
      case OP_CUBE: for (double& x : v) x = x * x * x; break;
      case OP_CLAMP01: for (double& x : v) x = std::min(std::max(x, 0.0), 1.0); break;
      default: break;
    }
  }
  for (int k = 0; k < 3; ++k) out[k] = oetf(v[k]);
}

void init_params(Rng& rng, std::vector<float>& p) {
  using namespace zh::mlp;
  p.assign(kParams, 0.0f);
  auto u = [&](double s) { return (float)((2.0 * rng.uniform() - 1.0) * s); };
  for (int i = 0; i < kIn * kHid; ++i) p[kOffW1 + i] = u(std::sqrt(6.0 / kIn));
  // First-layer hyperplanes through random points of the input cube (a zero
  // bias would put them all through its corner at the origin).
  for (int f = 0; f < kHid; ++f) {
    double b = 0.0;
    for (int k = 0; k < kIn; ++k) b -= p[kOffW1 + k * kHid + f] * rng.uniform();
    p[kOffB1 + f] = (float)b;
  }
  for (int i = 0; i < kHid * kHid; ++i) p[kOffW2 + i] = u(std::sqrt(6.0 / kHid));
  for (int i = 0; i < kHid * kOut; ++i) p[kOffW3 + i] = u(0.1 * std::sqrt(6.0 / kHid));
  for (int i = 0; i < kOut; ++i) p[kOffB3 + i] = 0.5f;
}

// Device buffers of one training run.
struct Trainer {
  hipStream_t s = nullptr;
  std::vector<void*> bufs;
  ~Trainer() {
    if (s) hipStreamSynchronize(s);
    for (void* b : bufs) hipFree(b);
  }
  template <typename T>
  hipError_t alloc(T** p, size_t count) {
    void* v = nullptr;
    hipError_t e = hipMalloc(&v, std::max<size_t>(count, 1) * sizeof(T));
    if (e == hipSuccess) bufs.push_back(v);
    *p = (T*)v;
    return e;
  }
};

// Surrogate outputs on n float4 inputs (device), downloaded as float4.
zh_status infer_host(hipStream_t s, uint32_t cus, const uint16_t* d_ops, const float* d_in, float* d_out, uint64_t n,
                     std::vector<float>& host) {
  ZS_HIP(zh_mlp_infer(d_ops, d_in, d_out, n, 6, 6, cus, s));
  host.resize(n * 4);
  ZS_HIP(hipMemcpyAsync(host.data(), d_out, n * 16, hipMemcpyDeviceToHost, s));
  ZS_HIP(hipStreamSynchronize(s));
  return ZH_OK;
}

zh_status train_stage(zh_program* p, uint32_t stage, const zh_learn_options& o, zh_learn_report* rep) {
  using namespace zh::mlp;
  zh_context* ctx = p->ctx;
  const StagePlan& sp = p->plan.stages[stage];
  const hipStream_t s = ctx->stream;
  const uint32_t cus = (uint32_t)ctx->cus;
  const auto t0 = std::chrono::steady_clock::now();
  Trainer T;
  T.s = s;

  // 1. samples, through the stage's own program on RGBA32F sRGB slots.
  const uint32_t M = o.uniform_points + o.face_points;
  Rng rng{0x5eedull * 0x100000001b3ull ^ o.seed};
  std::vector<float> hx((size_t)M * 4);
  for (uint32_t i = 0; i < M; ++i) {
    float* q = &hx[(size_t)i * 4];
    for (int k = 0; k < 3; ++k) q[k] = (float)rng.uniform();
    if (i >= o.uniform_points) {
      const uint32_t f = (uint32_t)(rng.next() % 6);
      q[f / 2] = (float)(f % 2);
    }
    q[3] = 1.0f;
  }
  float *d_x = nullptr, *d_t = nullptr;
  ZS_HIP(T.alloc(&d_x, (size_t)M * 4));
  ZS_HIP(T.alloc(&d_t, (size_t)M * 4));
  ZS_HIP(hipMemcpyAsync(d_x, hx.data(), hx.size() * 4, hipMemcpyHostToDevice, s));
  StageDesc td = sp.desc;
  td.width = M;
  td.height = 1;
  td.flat = 1;
  td.srgb_lut = ctx->d_lut;
  td.error_flag = ctx->d_error;
  td.pre_slot[0] = td.pre_slot[1] = -1;
  for (size_t k = 0; k < sp.slot_ref.size(); ++k) {
    Slot& sl = td.slot[k];
    sl.bits = ZH_BITS_FLOAT32X4;
    sl.parts = ZH_PARTS_RGBA;
    sl.transfer = ZH_TRANSFER_SRGB;
    sl.bpp = 16;
    sl.native = 0;
    sl.fmt = FMT_GENERIC;
    sl.width = M;
    sl.height = 1;
    sl.row_stride = (uint64_t)M * 16;
    sl.vec_ok = 1;
    sl.ptr = 0;
  }
  // The targets are the chain before its final clamps (oklab.frag:63):
  // smooth where the clamped chain has kinks at the gamut boundary.  The
  // inference kernel clamps the network's output when it stores 8-bit
  // texels, which applies the clamps exactly.
  for (uint32_t i = 0; i < td.n_instr; ++i)
    if (td.code[i].op == OP_CLAMP01) td.code[i].op = OP_NOP;
  td.slot[td.code[0].a].ptr = (uint64_t)(uintptr_t)d_x;
  td.slot[td.code[td.n_instr - 1].a].ptr = (uint64_t)(uintptr_t)d_t;
  StageDesc* d_td = nullptr;
  ZS_HIP(T.alloc(&d_td, 1));
  ZS_HIP(hipMemcpyAsync(d_td, &td, sizeof(td), hipMemcpyHostToDevice, s));
  ZS_HIP(zh_launch_pointwise(d_td, M, 1, 1, 0, cus, s));

  // 2. training.
  const uint32_t B = std::min(o.batch, M);
  std::vector<uint32_t> perm((size_t)M + B);
  for (uint32_t i = 0; i < M; ++i) perm[i] = i;
  for (uint32_t i = M - 1; i > 0; --i) std::swap(perm[i], perm[rng.next() % (i + 1)]);
  for (uint32_t i = 0; i < B; ++i) perm[M + i] = perm[i];
  std::vector<float> init;
  init_params(rng, init);
  uint32_t* d_perm = nullptr;
  float *d_p, *d_q, *d_g, *d_m, *d_v, *d_loss, *act;
  ZS_HIP(T.alloc(&d_perm, perm.size()));
  ZS_HIP(T.alloc(&d_p, kParams));
  ZS_HIP(T.alloc(&d_q, kParams));
  ZS_HIP(T.alloc(&d_g, kParams + zh_mlp_part_floats(B)));
  ZS_HIP(T.alloc(&d_m, kParams));
  ZS_HIP(T.alloc(&d_v, kParams));
  ZS_HIP(T.alloc(&d_loss, 1));
  ZS_HIP(T.alloc(&act, (size_t)B * (6 * kHid + kOut)));
  ZS_HIP(hipMemcpyAsync(d_perm, perm.data(), perm.size() * 4, hipMemcpyHostToDevice, s));
  ZS_HIP(hipMemcpyAsync(d_p, init.data(), kParams * 4, hipMemcpyHostToDevice, s));
  ZS_HIP(hipMemsetAsync(d_m, 0, kParams * 4, s));
  ZS_HIP(hipMemsetAsync(d_v, 0, kParams * 4, s));
  ZS_HIP(hipMemsetAsync(d_loss, 0, 4, s));
  zh_mlp_batch b{};
  b.x = d_x;
  b.t = d_t;
  b.x_stride = 4;
  b.n = B;
  b.a1 = act;
  b.h1 = act + (size_t)B * kHid;
  b.a2 = act + (size_t)2 * B * kHid;
  b.h2 = act + (size_t)3 * B * kHid;
  b.da1 = act + (size_t)4 * B * kHid;
  b.da2 = act + (size_t)5 * B * kHid;
  b.dy = act + (size_t)6 * B * kHid;
  b.loss = d_loss;
  const uint32_t tail = std::min<uint32_t>(100, o.steps);
  float first = 0.0f, last = 0.0f;
  double b1t = 1.0, b2t = 1.0;
  for (uint32_t t = 0; t < o.steps; ++t) {
    b.idx = d_perm + (size_t)(((uint64_t)t * B) % M);
    if (t == o.steps - tail) ZS_HIP(hipMemsetAsync(d_loss, 0, 4, s));
    ZS_HIP(zh_mlp_quantise(d_p, d_q, s));
    ZS_HIP(zh_mlp_loss_grad(d_q, &b, d_g, s));
    if (t == 0) {
      ZS_HIP(hipMemcpyAsync(&first, d_loss, 4, hipMemcpyDeviceToHost, s));
      ZS_HIP(hipStreamSynchronize(s));
    }
    const double lr_min = o.lr * 0.01;
    const double lr = lr_min + 0.5 * (o.lr - lr_min) * (1.0 + std::cos(M_PI * (double)t / (double)o.steps));
    b1t *= 0.9;
    b2t *= 0.999;
    ZS_HIP(zh_mlp_adam(d_p, d_g, d_m, d_v, (float)lr, (float)(1.0 - b1t), (float)(1.0 - b2t), s));
  }
  ZS_HIP(hipMemcpyAsync(&last, d_loss, 4, hipMemcpyDeviceToHost, s));
  std::vector<float> trained(kParams);
  ZS_HIP(hipMemcpyAsync(trained.data(), d_p, kParams * 4, hipMemcpyDeviceToHost, s));
  ZS_HIP(hipStreamSynchronize(s));
  const auto t1 = std::chrono::steady_clock::now();

  // 3. check on the lattice, for the trained and the initial network.
  Mat3 to_xyz = mat_identity();
  uint32_t prim = ZH_PRIMARIES_BT709, wp = ZH_WHITEPOINT_D65;
  const int64_t ref = sp.slot_ref[sp.desc.code[sp.desc.n_instr - 1].a];
  if (ref < 0 && ref > -1000000) {
    const zh_descriptor& od = p->plan.output_desc[(size_t)(~ref)];
    prim = od.color.primaries;
    wp = od.color.whitepoint;
  }
  if (!primaries_to_xyz(prim, wp, &to_xyz)) primaries_to_xyz(ZH_PRIMARIES_BT709, ZH_WHITEPOINT_D65, &to_xyz);
  const uint32_t L = o.lattice;
  const uint64_t nl = (uint64_t)L * L * L;
  std::vector<float> hl(nl * 4);
  std::vector<double> want(nl * 3);
  for (uint64_t i = 0; i < nl; ++i) {
    const double in[3] = {(double)(i % L) / (L - 1), (double)((i / L) % L) / (L - 1), (double)(i / ((uint64_t)L * L)) / (L - 1)};
    for (int k = 0; k < 3; ++k) hl[i * 4 + k] = (float)in[k];
    hl[i * 4 + 3] = 1.0f;
    eval_fp64(sp.desc, in, &want[i * 3]);
  }
  float *d_l = nullptr, *d_ly = nullptr;
  uint16_t* d_ops = nullptr;
  ZS_HIP(T.alloc(&d_l, nl * 4));
  ZS_HIP(T.alloc(&d_ly, nl * 4));
  ZS_HIP(T.alloc(&d_ops, kOps));
  ZS_HIP(hipMemcpyAsync(d_l, hl.data(), nl * 16, hipMemcpyHostToDevice, s));
  auto lattice_check = [&](const float* params_dev, double* mx, double* mean) -> zh_status {
    ZS_HIP(zh_mlp_pack(params_dev, d_ops, s));
    std::vector<float> got;
    zh_status st = infer_host(s, cus, d_ops, d_l, d_ly, nl, got);
    if (st != ZH_OK) return st;
    double m = 0.0, sum = 0.0;
    for (uint64_t i = 0; i < nl; ++i) {
      const double g[3] = {got[i * 4], got[i * 4 + 1], got[i * 4 + 2]};
      const double e = delta_e(to_xyz, g, &want[i * 3]);
      m = std::max(m, std::isfinite(e) ? e : 1e9);
      sum += e;
    }
    *mx = m;
    *mean = sum / (double)nl;
    return ZH_OK;
  };
  double mx0 = 0, mean0 = 0, mx = 0, mean = 0;
  float* d_init = d_q;  // reuse: the initial parameters
  ZS_HIP(hipMemcpyAsync(d_init, init.data(), kParams * 4, hipMemcpyHostToDevice, s));
  zh_status st = lattice_check(d_init, &mx0, &mean0);
  if (st != ZH_OK) return st;
  st = lattice_check(d_p, &mx, &mean);  // leaves the trained operands in d_ops
  if (st != ZH_OK) return st;

  rep->stage = stage;
  rep->steps = o.steps;
  rep->samples = M;
  rep->initial_loss = first;
  rep->final_loss = last / (float)tail;
  rep->max_delta_e = mx;
  rep->mean_delta_e = mean;
  rep->max_delta_e_initial = mx0;
  rep->train_ms = (float)std::chrono::duration<double, std::milli>(t1 - t0).count();
  rep->accepted = mx < (double)o.max_delta_e ? 1u : 0u;

  // Install: keep the operands (owned by the program) and the weights.
  if (p->d_surrogate[stage]) hipFree(p->d_surrogate[stage]);
  p->d_surrogate[stage] = nullptr;
  ZS_HIP(hipMalloc(&p->d_surrogate[stage], kOps * sizeof(uint16_t)));
  ZS_HIP(hipMemcpyAsync(p->d_surrogate[stage], d_ops, kOps * sizeof(uint16_t), hipMemcpyDeviceToDevice, s));
  ZS_HIP(hipStreamSynchronize(s));
  p->surrogate_ok[stage] = rep->accepted ? 1 : 0;
  p->surrogate_params[stage] = trained;
  return ZH_OK;
}

zh_status mlp_batch_on_host(zh_context* ctx, const float* params, uint32_t n, const float* x, const float* t,
                            float* grad, float* loss) {
  using namespace zh::mlp;
  Trainer T;
  const hipStream_t s = ctx->stream;
  T.s = s;
  std::vector<float> hx((size_t)n * 4, 0.0f), ht((size_t)n * 4, 0.0f);
  for (uint32_t i = 0; i < n; ++i)
    for (int k = 0; k < 3; ++k) {
      hx[(size_t)i * 4 + k] = x[(size_t)i * 3 + k];
      ht[(size_t)i * 4 + k] = t[(size_t)i * 3 + k];
    }
  float *d_x, *d_t, *d_p, *d_q, *d_g, *d_loss, *act;
  ZS_HIP(T.alloc(&d_x, hx.size()));
  ZS_HIP(T.alloc(&d_t, ht.size()));
  ZS_HIP(T.alloc(&d_p, kParams));
  ZS_HIP(T.alloc(&d_q, kParams));
  ZS_HIP(T.alloc(&d_g, kParams + zh_mlp_part_floats(n)));
  ZS_HIP(T.alloc(&d_loss, 1));
  ZS_HIP(T.alloc(&act, (size_t)n * (6 * kHid + kOut)));
  ZS_HIP(hipMemcpyAsync(d_x, hx.data(), hx.size() * 4, hipMemcpyHostToDevice, s));
  ZS_HIP(hipMemcpyAsync(d_t, ht.data(), ht.size() * 4, hipMemcpyHostToDevice, s));
  ZS_HIP(hipMemcpyAsync(d_p, params, kParams * 4, hipMemcpyHostToDevice, s));
  ZS_HIP(hipMemsetAsync(d_loss, 0, 4, s));
  zh_mlp_batch b{};
  b.x = d_x;
  b.t = d_t;
  b.idx = nullptr;
  b.x_stride = 4;
  b.n = n;
  b.a1 = act;
  b.h1 = act + (size_t)n * kHid;
  b.a2 = act + (size_t)2 * n * kHid;
  b.h2 = act + (size_t)3 * n * kHid;
  b.da1 = act + (size_t)4 * n * kHid;
  b.da2 = act + (size_t)5 * n * kHid;
  b.dy = act + (size_t)6 * n * kHid;
  b.loss = d_loss;
  ZS_HIP(zh_mlp_quantise(d_p, d_q, s));
  ZS_HIP(zh_mlp_loss_grad(d_q, &b, d_g, s));
  ZS_HIP(hipMemcpyAsync(grad, d_g, kParams * 4, hipMemcpyDeviceToHost, s));
  ZS_HIP(hipMemcpyAsync(loss, d_loss, 4, hipMemcpyDeviceToHost, s));
  ZS_HIP(hipStreamSynchronize(s));
  return ZH_OK;
}

}  // namespace

extern "C" {

zh_status zh_program_learn(zh_program* prog, const zh_learn_options* options, zh_learn_report* reports,
                           uint32_t max_reports, uint32_t* num_reports) {
  if (!prog) return fail(ZH_ERR_INVALID_ARGUMENT, "null program");
  if (!prog->ctx) return fail(ZH_ERR_NO_DEVICE, "program was built without a context");
  zh_learn_options o{};
  if (options) o = *options;
  if (!options) o.seed = 0;
  if (o.uniform_points == 0) o.uniform_points = 1u << 20;
  if (o.face_points == 0 && !(options && options->uniform_points)) o.face_points = 1u << 16;
  if (o.steps == 0) o.steps = 20000;
  if (o.batch == 0) o.batch = 8192;
  if (!(o.lr > 0.0f)) o.lr = 3e-3f;
  if (o.lattice < 2) o.lattice = 65;
  if (!(o.max_delta_e > 0.0f)) o.max_delta_e = 0.005f;
  if (o.lattice > 257) return fail(ZH_ERR_INVALID_ARGUMENT, "lattice larger than 257 per axis");
  ZS_HIP(hipSetDevice(prog->ctx->device));
  const size_t ns = prog->plan.stages.size();
  prog->d_surrogate.resize(ns, nullptr);
  prog->surrogate_ok.resize(ns, 0);
  prog->surrogate_params.resize(ns);
  // Stages with the same colour program (ops and matrices) share one
  // training run: the surrogate depends on nothing else.
  auto program_of = [&](uint32_t s) {
    const StageDesc& d = prog->plan.stages[s].desc;
    std::vector<float> key;
    for (uint32_t i = 1; i + 1 < d.n_instr; ++i) {
      key.push_back((float)d.code[i].op);
      if (d.code[i].op == OP_MAT3) key.insert(key.end(), d.params + d.code[i].arg, d.params + d.code[i].arg + 9);
    }
    return key;
  };
  std::vector<std::pair<std::vector<float>, zh_learn_report>> done;
  uint32_t count = 0;
  for (uint32_t s = 0; s < ns; ++s) {
    if (!eligible(prog->plan.stages[s])) continue;
    const std::vector<float> key = program_of(s);
    zh_learn_report rep{};
    bool shared = false;
    for (const auto& dn : done) {
      if (dn.first != key) continue;
      const uint32_t from = dn.second.stage;
      rep = dn.second;
      rep.stage = s;
      rep.train_ms = 0.0f;
      if (prog->d_surrogate[s]) hipFree(prog->d_surrogate[s]);
      prog->d_surrogate[s] = nullptr;
      ZS_HIP(hipMalloc(&prog->d_surrogate[s], zh::mlp::kOps * sizeof(uint16_t)));
      ZS_HIP(hipMemcpy(prog->d_surrogate[s], prog->d_surrogate[from], zh::mlp::kOps * sizeof(uint16_t),
                       hipMemcpyDeviceToDevice));
      prog->surrogate_ok[s] = prog->surrogate_ok[from];
      prog->surrogate_params[s] = prog->surrogate_params[from];
      shared = true;
      break;
    }
    if (!shared) {
      zh_status st = train_stage(prog, s, o, &rep);
      if (st != ZH_OK) return st;
      done.emplace_back(key, rep);
    }
    if (reports && count < max_reports) reports[count] = rep;
    ++count;
  }
  if (num_reports) *num_reports = count;
  prog->uploaded = false;  // re-bind: the stage kernels may change
  return ZH_OK;
}

zh_status zh_program_use_surrogates(zh_program* prog, int32_t enable) {
  if (!prog) return fail(ZH_ERR_INVALID_ARGUMENT, "null program");
  if (enable < 0 || enable > 2) return fail(ZH_ERR_INVALID_ARGUMENT, "mode is 0, 1 or 2");
  prog->use_surrogates = enable;
  prog->uploaded = false;
  return ZH_OK;
}

zh_status zh_program_surrogate_params(zh_program* prog, uint32_t stage, float* params, uint32_t len) {
  if (!prog || !params) return fail(ZH_ERR_INVALID_ARGUMENT, "null argument");
  if (stage >= prog->surrogate_params.size() || prog->surrogate_params[stage].empty())
    return fail(ZH_ERR_INVALID_ARGUMENT, "stage has no surrogate");
  if (len < (uint32_t)zh::mlp::kParams) return fail(ZH_ERR_INVALID_ARGUMENT, "buffer shorter than 4611 floats");
  std::memcpy(params, prog->surrogate_params[stage].data(), zh::mlp::kParams * sizeof(float));
  return ZH_OK;
}

zh_status zh_mlp_gradient(zh_context* ctx, const float* params, uint32_t n, const float* x, const float* t,
                          float* grad, float* loss) {
  if (!ctx) return fail(ZH_ERR_NO_DEVICE, "no context");
  if (!params || !x || !t || !grad || !loss || n == 0) return fail(ZH_ERR_INVALID_ARGUMENT, "null argument");
  ZS_HIP(hipSetDevice(ctx->device));
  return mlp_batch_on_host(ctx, params, n, x, t, grad, loss);
}

zh_status zh_mlp_evaluate(zh_context* ctx, const float* params, uint32_t n, const float* x, float* y) {
  using namespace zh::mlp;
  if (!ctx) return fail(ZH_ERR_NO_DEVICE, "no context");
  if (!params || !x || !y || n == 0) return fail(ZH_ERR_INVALID_ARGUMENT, "null argument");
  ZS_HIP(hipSetDevice(ctx->device));
  Trainer T;
  const hipStream_t s = ctx->stream;
  T.s = s;
  std::vector<float> hx((size_t)n * 4, 1.0f);
  for (uint32_t i = 0; i < n; ++i)
    for (int k = 0; k < 3; ++k) hx[(size_t)i * 4 + k] = x[(size_t)i * 3 + k];
  float *d_p, *d_x, *d_y;
  uint16_t* d_ops;
  ZS_HIP(T.alloc(&d_p, kParams));
  ZS_HIP(T.alloc(&d_x, hx.size()));
  ZS_HIP(T.alloc(&d_y, hx.size()));
  ZS_HIP(T.alloc(&d_ops, kOps));
  ZS_HIP(hipMemcpyAsync(d_p, params, kParams * 4, hipMemcpyHostToDevice, s));
  ZS_HIP(hipMemcpyAsync(d_x, hx.data(), hx.size() * 4, hipMemcpyHostToDevice, s));
  ZS_HIP(zh_mlp_pack(d_p, d_ops, s));
  std::vector<float> got;
  zh_status st = infer_host(s, (uint32_t)ctx->cus, d_ops, d_x, d_y, n, got);
  if (st != ZH_OK) return st;
  for (uint32_t i = 0; i < n; ++i)
    for (int k = 0; k < 3; ++k) y[(size_t)i * 3 + k] = got[(size_t)i * 4 + k];
  return ZH_OK;
}

}  // extern "C"
