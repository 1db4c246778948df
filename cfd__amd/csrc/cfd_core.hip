// Solver runtime behind the C ABI of include/cfd_mi355x.h.
//
// Mirrors the resource model of SolverWebGPU (src/backend/webgpu/solver.js):
// two population sets with a front flag, a type field, a macro field and a
// reduction scratch area, all owned by the handle.  Each handle has its own
// HIP stream (default priority); everything it does is ordered on it.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/cfd_mi355x.h"
#include "kernels.h"
#include "lattice.h"

using namespace cfd;

namespace {

thread_local std::string g_err;

int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

#define HIPCK(expr)                                                                  \
  do {                                                                               \
    hipError_t e_ = (expr);                                                          \
    if (e_ != hipSuccess)                                                            \
      return fail(e_ == hipErrorOutOfMemory ? CFD_E_OOM : CFD_E_HIP, "%s: %s (%s)",   \
                  __func__, hipGetErrorString(e_), #expr);                           \
  } while (0)

}  // namespace

struct cfd_solver {
  int nx = 0, ny = 0, nz = 0;
  size_t N = 0;
  Store store = Store::F32;
  size_t pop_bytes = 0;  // bytes of one population set
  Geom geom{};
  Tuning tuning{256, 8, kLayoutBlocked, 0, 0, 0};  // cfd_create adjusts it per population set
  StepParams params{};
  bool configured = false;
  int front = 0;  // 0 = A, 1 = B
  uint32_t sr_step = 0;  // steps taken (seeds the stochastic float16 rounding)
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  void* pop[2] = {nullptr, nullptr};
  uint8_t* tipo8 = nullptr;
  uint2* links = nullptr;     // body links of the current types, cell order
  uint32_t nlinks = 0;
  size_t link_cap = 0;
  float4* macros = nullptr;
  double* partials = nullptr;
  double* result = nullptr;   // device, 4 doubles
  double* host_res = nullptr; // pinned, 4 doubles
  // learned closure (off while mlp == nullptr)
  MlpPacked* mlp = nullptr;
  int mlp_act = CFD_ACT_SILU;  // its hidden activation
  uint8_t* wall8 = nullptr;   // wall distance, 1/8 cell, saturating at 31.875
  bool wall_given = false;    // wall8 holds the caller's distances, not the types' transform
  unsigned long long* fallbacks = nullptr;
  // The macro field is computed when something reads it, not in every
  // batch of steps: a step records the set it read and its parameters, and
  // materialize_macros() pulls that step's moments from that set, bit for
  // bit what the step would have stored (the same pull and moments code).
  // Anything that overwrites that set, the types or the layout first
  // materializes a pending field.
  bool macros_pending = false;
  int macros_src = 0;
  StepParams macros_params{};
};

namespace {

void release(cfd_solver* s) {
  if (!s) return;
  if (s->stream) (void)hipStreamSynchronize(s->stream);
  for (void* p : s->pop)
    if (p) (void)hipFree(p);
  if (s->tipo8) (void)hipFree(s->tipo8);
  if (s->links) (void)hipFree(s->links);
  if (s->macros) (void)hipFree(s->macros);
  if (s->partials) (void)hipFree(s->partials);
  if (s->result) (void)hipFree(s->result);
  if (s->host_res) (void)hipHostFree(s->host_res);
  if (s->mlp) (void)hipFree(s->mlp);
  if (s->wall8) (void)hipFree(s->wall8);
  if (s->fallbacks) (void)hipFree(s->fallbacks);
  if (s->ev0) (void)hipEventDestroy(s->ev0);
  if (s->ev1) (void)hipEventDestroy(s->ev1);
  if (s->stream) (void)hipStreamDestroy(s->stream);
  delete s;
}

int64_t device_bytes(const cfd_solver* s) {
  return (int64_t)(2 * s->pop_bytes + s->N * (1 + 16) +
                   (kReduceGrid * 3 + 4) * sizeof(double));
}

// c_i . v in the emitted order (ir.js:116-126), float32, host side.
float dotc_host(int i, const float v[3]) {
  float acc = 0.0f;
  bool first = true;
  const int c[3] = {cx(i), cy(i), cz(i)};
  for (int a = 0; a < 3; ++a) {
    if (c[a] == 0) continue;
    if (first) {
      acc = c[a] > 0 ? v[a] : -v[a];
      first = false;
    } else {
      acc = c[a] > 0 ? acc + v[a] : acc - v[a];
    }
  }
  return acc;
}

bool is_f32(Store s) { return s == Store::F32 || s == Store::F32X; }
size_t elem_bytes(const cfd_solver* s) { return is_f32(s->store) ? 4 : 2; }

// 36 % of the step's launch grid (workgroups, padded to whole XCD chunk
// spans as launch_step pads it).
int default_rotation(size_t n, int block, int remap) {
  const uint64_t nb = (n + block - 1) / block;
  const uint64_t span = remap > 1 ? (uint64_t)kXcds * remap : 1;
  const uint64_t grid = (nb + span - 1) / span * span;
  return (int)(grid * 36 / 100);
}

}  // namespace

namespace cfd {

// The Params uniform (solver.js:277-293) rounded to float32 as
// Float32Array.set() does, plus the uniform sub-expressions of the emitted
// shader folded once, in float32 and in the shader's operation order
// (volatile keeps each rounding where the shader has it).
StepParams make_step_params(const cfd_params& p) {
  StepParams q{};
  q.omega_plus = (float)p.omega_plus;
  q.magic = (float)p.magic;
  q.les_cs = (float)p.les_cs;
  for (int a = 0; a < 3; ++a) {
    q.belt[a] = (float)p.belt[a];
    q.inlet[a] = (float)p.inlet[a];
  }
  q.sponge_start = (float)p.sponge_start;
  const float len = (float)p.sponge_len;
  q.sponge_len_m = len > 1.0f ? len : 1.0f;  // max(uSpongeLen, 1.0)  (ir.js:461)
  q.bl = (float)p.bl_thickness;
  q.bl_m = q.bl > 1.0f ? q.bl : 1.0f;         // max(P.sponge.z, 1.0) (wgsl.js:217)
  volatile float om = q.omega_plus;
  q.tau0 = 1.0f / om;                         // ir.js:290
  volatile float t0 = q.tau0;
  q.tau0sq = t0 * t0;                         // ir.js:292
  volatile float cs = q.les_cs;
  volatile float c18 = 18.0f * cs;
  q.c18cs2 = c18 * cs;                        // ir.js:293
  // free-stream equilibrium of the sponge (ir.js:464-475)
  const float* ui = q.inlet;
  volatile float uu_a = ui[0] * ui[0];
  volatile float uu_b = uu_a + ui[1] * ui[1];
  volatile float uuinf = uu_b + ui[2] * ui[2];
  {
    volatile float m15 = -1.5f * uuinf;
    q.tinf[0] = wf(0) * m15;
  }
  for (int i = 1; i < Q; ++i) {
    volatile float si = dotc_host(i, ui);
    volatile float a = 3.0f * si;
    volatile float b = 4.5f * si;
    volatile float b2 = b * si;
    volatile float ab = a + b2;
    volatile float c = 1.5f * uuinf;
    volatile float r = ab - c;
    q.tinf[i] = wf(i) * r;
  }
  q.sponge_all = 0;
  for (int i = 0; i < Q; ++i)
    if (!std::isfinite(q.tinf[i])) q.sponge_all = 1;
  return q;
}

}  // namespace cfd

extern "C" {

const char* cfd_last_error(void) { return g_err.c_str(); }

// Lets the voxelizer translation unit report through cfd_last_error().
void cfd__set_error(const char* msg) { g_err = msg ? msg : ""; }

// 0.2: cfd_tuning gained rotate; 0.3: cfd_abi_version, default launch shape per set;
// 0.4: cfd_mlp.activation
const char* cfd_version(void) { return "cfd_mi355x 0.4 gfx950"; }

int cfd_abi_version(void) { return CFD_ABI_VERSION; }

int cfd_create(int nx, int ny, int nz, uint32_t flags, cfd_solver** out) {
  g_err.clear();
  if (!out) return fail(CFD_E_ARG, "cfd_create: out is null");
  *out = nullptr;
  if (nx < 1 || ny < 1 || nz < 1)
    return fail(CFD_E_ARG, "cfd_create: grid %dx%dx%d must be positive", nx, ny, nz);
  const uint64_t N = (uint64_t)nx * (uint64_t)ny * (uint64_t)nz;
  // Byte offsets of one population slice are 32-bit in the kernels.
  if (N >= (1ull << 30))
    return fail(CFD_E_ARG, "cfd_create: %llu cells exceed the 2^30 limit of one handle",
                (unsigned long long)N);
  if (flags != CFD_STORE_FP32 && flags != CFD_STORE_FP16 &&
      flags != (CFD_STORE_FP16 | CFD_ROUND_STOCHASTIC) &&
      flags != (CFD_STORE_FP32 | CFD_ROUND_EXACT))
    return fail(CFD_E_ARG, "cfd_create: unknown storage flags 0x%x", flags);
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
    return fail(CFD_E_HIP, "cfd_create: no HIP device visible");

  cfd_solver* s = new cfd_solver();
  s->nx = nx;
  s->ny = ny;
  s->nz = nz;
  s->N = (size_t)N;
  s->store = flags == CFD_STORE_FP32                        ? Store::F32
             : flags == (CFD_STORE_FP32 | CFD_ROUND_EXACT) ? Store::F32X
             : (flags & CFD_ROUND_STOCHASTIC)              ? Store::F16S
                                                           : Store::F16;
  const size_t eb = is_f32(s->store) ? 4 : 2;
  // A float16 set of at most 128 MB (both sets fit the 256 MB last-level
  // cache): SoA, whose per-load address is one add where the blocked layout's
  // is a shift and a multiply-add; these steps are bound by arithmetic, and
  // at 240x120x96 the closure step measured 75.2 us against 79.5 us blocked,
  // the plain float16 step 39.0 us against 40.3 us.  The 249 MB float16
  // vehicle set stays blocked (97.1 us against 99.0 us SoA).
  if (!is_f32(s->store) && (size_t)Q * (size_t)N * eb <= (128ull << 20))
    s->tuning.layout = kLayoutSoA;
  s->geom = make_geom(nx, ny, nz, s->tuning.layout, (uint32_t)s->tuning.slice_pad);
  if (s->geom.pop_elems * eb >= (1ull << 32)) {
    // the blocked layout addresses a set with 32-bit byte offsets
    s->tuning.layout = kLayoutSoA;
    s->geom = make_geom(nx, ny, nz, kLayoutSoA, (uint32_t)s->tuning.slice_pad);
  }
  s->pop_bytes = s->geom.pop_elems * eb;
  // Alternate the cell order per step where it measured faster: a set larger
  // than half the 256 MB last-level cache (below that both sets stay cached
  // in either order) that fits it whole, so that the next step first reads
  // what is still cached, or any such float16 set, whose step is bound by
  // arithmetic and gained at every size measured.  A float32 set larger than
  // the cache lost 1.5-2 % with it, one of 62 MB about 1 %.
  s->tuning.sweep = s->pop_bytes > (128ull << 20) &&
                    (!is_f32(s->store) || s->pop_bytes <= (256ull << 20));
  // A float32 set larger than the cache (the 6.55 M-cell vehicle's 498 MB):
  // 512-thread workgroups with the start rotated by 36 % of the launch grid
  // per step, the launch-order autotuner's pick on every box measured
  // (174.3-175.1 us per step on the vehicle against 175.5-177.5 us for
  // 256-thread workgroups in ascending order).  Every shape gives the same
  // bits, so this only sets what a run that does not autotune gets.
  if (is_f32(s->store) && s->pop_bytes > (256ull << 20) &&
      s->tuning.layout == kLayoutBlocked) {
    s->tuning.block = 512;
    s->tuning.rotate = default_rotation(s->N, 512, s->tuning.remap);
  }

  auto bail = [&](int code) {
    release(s);
    return code;
  };
  hipError_t e = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking);
  if (e != hipSuccess) return bail(fail(CFD_E_HIP, "cfd_create: %s", hipGetErrorString(e)));
  if (hipEventCreate(&s->ev0) != hipSuccess || hipEventCreate(&s->ev1) != hipSuccess)
    return bail(fail(CFD_E_HIP, "cfd_create: event creation failed"));

  struct {
    void** p;
    size_t bytes;
  } allocs[] = {
      {&s->pop[0], s->pop_bytes},
      {&s->pop[1], s->pop_bytes},
      {(void**)&s->tipo8, s->N},
      {(void**)&s->macros, s->N * sizeof(float4)},
      {(void**)&s->partials, (size_t)kReduceGrid * 3 * sizeof(double)},
      {(void**)&s->result, 4 * sizeof(double)},
  };
  for (auto& a : allocs) {
    e = hipMalloc(a.p, a.bytes);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      return bail(fail(CFD_E_OOM,
                       "device memory insufficient for %.2f GB (%dx%dx%d); choose a "
                       "smaller grid",
                       device_bytes(s) / 1e9, nx, ny, nz));
    }
    e = hipMemsetAsync(*a.p, 0, a.bytes, s->stream);
    if (e != hipSuccess) return bail(fail(CFD_E_HIP, "cfd_create: memset failed"));
  }
  if (hipHostMalloc((void**)&s->host_res, 4 * sizeof(double), hipHostMallocDefault) !=
      hipSuccess)
    return bail(fail(CFD_E_HIP, "cfd_create: pinned allocation failed"));
  e = hipStreamSynchronize(s->stream);
  if (e != hipSuccess) return bail(fail(CFD_E_HIP, "cfd_create: %s", hipGetErrorString(e)));
  *out = s;
  return CFD_OK;
}

int cfd_configure(cfd_solver* s, const cfd_params* p) {
  g_err.clear();
  if (!s || !p) return fail(CFD_E_ARG, "cfd_configure: null argument");
  const StepParams q = make_step_params(*p);
  s->params = q;
  s->configured = true;
  return CFD_OK;
}

namespace {
// The body-link list of the current types (momentum exchange reads it):
// count per chunk, scan, then write the entries into a list grown as needed.
int build_links(cfd_solver* s) {
  const uint32_t nb = link_chunks(s->geom);
  uint32_t* tmp = nullptr;  // counts[nb], offsets[nb + 1]
  HIPCK(hipMalloc(&tmp, (2 * (size_t)nb + 1) * sizeof(uint32_t)));
  uint32_t total = 0;
  hipError_t e = launch_link_count(s->tipo8, s->geom, tmp, tmp + nb, s->stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(&total, tmp + 2 * (size_t)nb, sizeof(uint32_t), hipMemcpyDeviceToHost,
                       s->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
  if (e == hipSuccess && total > s->link_cap) {
    if (s->links) (void)hipFree(s->links);
    s->links = nullptr;
    s->link_cap = 0;
    s->nlinks = 0;
    e = hipMalloc(&s->links, (size_t)total * sizeof(uint2));
    if (e == hipSuccess) s->link_cap = total;
  }
  if (e == hipSuccess && total)
    e = launch_link_write(s->tipo8, s->geom, tmp + nb, s->links, s->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
  (void)hipFree(tmp);
  if (e != hipSuccess) {
    s->nlinks = 0;
    return fail(e == hipErrorOutOfMemory ? CFD_E_OOM : CFD_E_HIP, "cfd_set_types: %s",
                hipGetErrorString(e));
  }
  s->nlinks = total;
  return CFD_OK;
}
}  // namespace

static int compute_wall(cfd_solver* s);
static int materialize_macros(cfd_solver* s);

int cfd_set_types(cfd_solver* s, const uint32_t* tipo, size_t n) {
  g_err.clear();
  if (!s || !tipo) return fail(CFD_E_ARG, "cfd_set_types: null argument");
  if (n != s->N)
    return fail(CFD_E_ARG, "cfd_set_types: type field has %zu cells, expected %zu", n, s->N);
  const int mrc = materialize_macros(s);
  if (mrc) return mrc;
  uint32_t* d32 = nullptr;
  HIPCK(hipMalloc(&d32, n * sizeof(uint32_t)));
  hipError_t e = hipMemcpyAsync(d32, tipo, n * sizeof(uint32_t), hipMemcpyHostToDevice,
                                s->stream);
  if (e == hipSuccess) e = launch_prep_types(d32, s->tipo8, s->geom, s->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
  (void)hipFree(d32);
  if (e != hipSuccess) return fail(CFD_E_HIP, "cfd_set_types: %s", hipGetErrorString(e));
  const int rc = build_links(s);
  if (rc) return rc;
  // a wall distance derived from the types follows them
  return s->wall8 && !s->wall_given ? compute_wall(s) : CFD_OK;
}

int cfd_init_equilibrium(cfd_solver* s) {
  g_err.clear();
  if (!s) return fail(CFD_E_ARG, "cfd_init_equilibrium: null handle");
  if (!s->configured)
    return fail(CFD_E_STATE, "cfd_init_equilibrium: cfd_configure was never called");
  const int rc = materialize_macros(s);
  if (rc) return rc;
  HIPCK(launch_init(s->store, s->pop[1], s->tipo8, s->geom, s->params, s->stream));
  s->front = 1;  // init writes B (solver.js:342-346)
  return CFD_OK;
}

// Dense [19][N] float32 on the device <-> the handle's storage layout.
static int upload_dense(cfd_solver* s, const float* host, void* dst) {
  const size_t n = (size_t)Q * s->N;
  float* tmp = nullptr;
  HIPCK(hipMalloc(&tmp, n * sizeof(float)));
  hipError_t e = hipMemcpyAsync(tmp, host, n * sizeof(float), hipMemcpyHostToDevice, s->stream);
  if (e == hipSuccess) e = launch_f32_to_store(s->store, tmp, dst, s->geom, s->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
  (void)hipFree(tmp);
  if (e != hipSuccess) return fail(CFD_E_HIP, "population upload: %s", hipGetErrorString(e));
  return CFD_OK;
}

static int download_dense(cfd_solver* s, const void* src, float* host) {
  const size_t n = (size_t)Q * s->N;
  float* tmp = nullptr;
  HIPCK(hipMalloc(&tmp, n * sizeof(float)));
  hipError_t e = launch_store_to_f32(s->store, src, tmp, s->geom, s->stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(host, tmp, n * sizeof(float), hipMemcpyDeviceToHost, s->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
  (void)hipFree(tmp);
  if (e != hipSuccess) return fail(CFD_E_HIP, "population read-back: %s", hipGetErrorString(e));
  return CFD_OK;
}

int cfd_set_populations(cfd_solver* s, const float* soa, size_t n) {
  g_err.clear();
  if (!s || !soa) return fail(CFD_E_ARG, "cfd_set_populations: null argument");
  if (n != (size_t)Q * s->N)
    return fail(CFD_E_ARG, "cfd_set_populations: %zu values, expected %zu", n,
                (size_t)Q * s->N);
  const int rc = materialize_macros(s);
  if (rc) return rc;
  return upload_dense(s, soa, s->pop[s->front]);
}

int cfd_get_populations(cfd_solver* s, float* out, size_t n) {
  g_err.clear();
  if (!s || !out) return fail(CFD_E_ARG, "cfd_get_populations: null argument");
  if (n != (size_t)Q * s->N)
    return fail(CFD_E_ARG, "cfd_get_populations: %zu values, expected %zu", n,
                (size_t)Q * s->N);
  return download_dense(s, s->pop[s->front], out);
}

int cfd_set_tuning(cfd_solver* s, const cfd_tuning* t) {
  g_err.clear();
  if (!s || !t) return fail(CFD_E_ARG, "cfd_set_tuning: null argument");
  if (t->block != 256 && t->block != 512)
    return fail(CFD_E_ARG, "cfd_set_tuning: block %d must be 256 or 512", t->block);
  if (t->layout != kLayoutSoA && t->layout != kLayoutBlocked)
    return fail(CFD_E_ARG, "cfd_set_tuning: layout %d must be 0 (SoA) or 1 (blocked)",
                t->layout);
  if (t->remap < 0 || t->remap > 1024 || (t->remap & (t->remap - 1)) != 0)
    return fail(CFD_E_ARG, "cfd_set_tuning: remap %d must be 0 or a power of two <= 1024",
                t->remap);
  if (t->slice_pad < 0 || t->slice_pad > (1 << 20))
    return fail(CFD_E_ARG, "cfd_set_tuning: slice_pad %d outside [0, 2^20]", t->slice_pad);
  if (t->sweep != 0 && t->sweep != 1)
    return fail(CFD_E_ARG, "cfd_set_tuning: sweep %d must be 0 or 1", t->sweep);
  if (t->rotate < 0)
    return fail(CFD_E_ARG, "cfd_set_tuning: rotate %d must not be negative", t->rotate);
  if (t->rotate > 0 && t->sweep != 0)
    return fail(CFD_E_ARG, "cfd_set_tuning: rotate needs sweep 0 (the two orders exclude "
                           "each other)");
  const size_t eb = elem_bytes(s);
  const Geom ng = make_geom(s->nx, s->ny, s->nz, t->layout, (uint32_t)t->slice_pad);
  if (t->layout == kLayoutBlocked && ng.pop_elems * eb >= (1ull << 32))
    return fail(CFD_E_ARG, "cfd_set_tuning: the blocked layout needs a population set below "
                           "4 GiB (this grid: %.2f GB)", ng.pop_elems * eb / 1e9);
  if (ng.lb != s->geom.lb || ng.S != s->geom.S) {
    const int rc = materialize_macros(s);
    if (rc) return rc;
    // Re-lay both population sets (contents preserved, walls included).
    const size_t n = (size_t)Q * s->N;
    float* dense = nullptr;
    void* np[2] = {nullptr, nullptr};
    const size_t nbytes = ng.pop_elems * eb;
    hipError_t e = hipMalloc(&dense, 2 * n * sizeof(float));
    for (int k = 0; k < 2 && e == hipSuccess; ++k)
      e = launch_store_to_f32(s->store, s->pop[k], dense + k * n, s->geom, s->stream);
    for (int k = 0; k < 2 && e == hipSuccess; ++k) {
      e = hipMalloc(&np[k], nbytes);
      if (e == hipSuccess) e = hipMemsetAsync(np[k], 0, nbytes, s->stream);
      if (e == hipSuccess) e = launch_f32_to_store(s->store, dense + k * n, np[k], ng, s->stream);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    if (dense) (void)hipFree(dense);
    if (e != hipSuccess) {
      for (void* p : np)
        if (p) (void)hipFree(p);
      (void)hipGetLastError();
      return fail(e == hipErrorOutOfMemory ? CFD_E_OOM : CFD_E_HIP,
                  "cfd_set_tuning: re-layout failed: %s", hipGetErrorString(e));
    }
    for (int k = 0; k < 2; ++k) {
      (void)hipFree(s->pop[k]);
      s->pop[k] = np[k];
    }
    s->geom = ng;
    s->pop_bytes = nbytes;
  }
  s->tuning.block = t->block;
  s->tuning.remap = t->remap > 1 ? t->remap : 0;
  s->tuning.layout = t->layout;
  s->tuning.slice_pad = t->slice_pad;
  s->tuning.sweep = t->sweep;
  s->tuning.rotate = t->rotate;
  return CFD_OK;
}

int cfd_get_tuning(cfd_solver* s, cfd_tuning* t) {
  g_err.clear();
  if (!s || !t) return fail(CFD_E_ARG, "cfd_get_tuning: null argument");
  t->block = s->tuning.block;
  t->remap = s->tuning.remap;
  t->layout = s->tuning.layout;
  t->slice_pad = s->tuning.slice_pad;
  t->sweep = s->tuning.sweep;
  t->rotate = s->tuning.rotate;
  return CFD_OK;
}

int cfd_set_front(cfd_solver* s, int front) {
  g_err.clear();
  if (!s) return fail(CFD_E_ARG, "cfd_set_front: null handle");
  if (front != 0 && front != 1) return fail(CFD_E_ARG, "cfd_set_front: front must be 0 or 1");
  const int rc = materialize_macros(s);
  if (rc) return rc;
  s->front = front;
  return CFD_OK;
}

int cfd_get_front(cfd_solver* s, int* front) {
  g_err.clear();
  if (!s || !front) return fail(CFD_E_ARG, "cfd_get_front: null argument");
  *front = s->front;
  return CFD_OK;
}

// The pending macro field of the last step into s->macros (see cfd_solver).
static int materialize_macros(cfd_solver* s) {
  if (!s->macros_pending) return CFD_OK;
  HIPCK(launch_pull_macros(s->store, s->pop[s->macros_src], s->tipo8, s->macros, s->geom,
                           s->macros_params, s->stream));
  s->macros_pending = false;
  return CFD_OK;
}

static int enqueue_steps(cfd_solver* s, int n) {
  for (int k = 0; k < n; ++k) {
    const int f = s->front;
    s->params.sr_seed = ++s->sr_step * 0x9E3779B9u;   // a new rounding draw per step
    if (s->mlp)
      HIPCK(launch_step_closure(s->store, s->pop[f], s->pop[f ^ 1], s->tipo8, s->wall8, s->mlp,
                                s->mlp_act, s->macros, false, s->geom, s->params, s->fallbacks,
                                s->stream));
    else
      HIPCK(launch_step(s->store, s->pop[f], s->pop[f ^ 1], s->tipo8, s->macros, false,
                        s->geom, s->params, s->tuning, s->tuning.sweep && (s->sr_step & 1),
                        s->sr_step, s->stream));
    s->front = f ^ 1;
    // this step's macro field supersedes any pending one
    s->macros_pending = true;
    s->macros_src = f;
    s->macros_params = s->params;
  }
  return CFD_OK;
}

int cfd_step(cfd_solver* s, int n) {
  g_err.clear();
  if (!s) return fail(CFD_E_ARG, "cfd_step: null handle");
  if (n < 0) return fail(CFD_E_ARG, "cfd_step: negative step count %d", n);
  if (!s->configured) return fail(CFD_E_STATE, "cfd_step: cfd_configure was never called");
  return enqueue_steps(s, n);
}

int cfd_step_timed(cfd_solver* s, int n, double* ms_out) {
  g_err.clear();
  if (!s || !ms_out) return fail(CFD_E_ARG, "cfd_step_timed: null argument");
  if (n < 1) return fail(CFD_E_ARG, "cfd_step_timed: step count %d must be >= 1", n);
  if (!s->configured)
    return fail(CFD_E_STATE, "cfd_step_timed: cfd_configure was never called");
  HIPCK(hipEventRecord(s->ev0, s->stream));
  int rc = enqueue_steps(s, n);
  if (rc) return rc;
  HIPCK(hipEventRecord(s->ev1, s->stream));
  HIPCK(hipEventSynchronize(s->ev1));
  float ms = 0.f;
  HIPCK(hipEventElapsedTime(&ms, s->ev0, s->ev1));
  *ms_out = ms;
  return CFD_OK;
}

int cfd_forces(cfd_solver* s, double out_f[3]) {
  g_err.clear();
  if (!s || !out_f) return fail(CFD_E_ARG, "cfd_forces: null argument");
  HIPCK(launch_forces(s->store, s->pop[s->front], s->links, s->nlinks, s->geom, s->partials,
                      s->result, s->stream));
  HIPCK(hipMemcpyAsync(s->host_res, s->result, 3 * sizeof(double), hipMemcpyDeviceToHost,
                       s->stream));
  HIPCK(hipStreamSynchronize(s->stream));
  for (int a = 0; a < 3; ++a) out_f[a] = s->host_res[a];
  return CFD_OK;
}

int cfd_update_macros(cfd_solver* s) {
  g_err.clear();
  if (!s) return fail(CFD_E_ARG, "cfd_update_macros: null handle");
  HIPCK(launch_macros(s->store, s->pop[s->front], s->macros, s->geom, s->stream));
  s->macros_pending = false;   // overwritten
  return CFD_OK;
}

int cfd_read_macros(cfd_solver* s, float* out, size_t n) {
  g_err.clear();
  if (!s || !out) return fail(CFD_E_ARG, "cfd_read_macros: null argument");
  if (n != 4 * s->N)
    return fail(CFD_E_ARG, "cfd_read_macros: %zu values, expected %zu", n, 4 * s->N);
  int rc = materialize_macros(s);
  if (rc) return rc;
  HIPCK(hipMemcpyAsync(out, s->macros, n * sizeof(float), hipMemcpyDeviceToHost, s->stream));
  HIPCK(hipStreamSynchronize(s->stream));
  return CFD_OK;
}

int cfd_mass(cfd_solver* s, double* mass_out, int64_t* n_nonfinite_out) {
  g_err.clear();
  if (!s || !mass_out) return fail(CFD_E_ARG, "cfd_mass: null argument");
  HIPCK(launch_mass(s->store, s->pop[s->front], s->tipo8, s->geom, s->partials, s->result,
                    s->stream));
  HIPCK(hipMemcpyAsync(s->host_res, s->result, 2 * sizeof(double), hipMemcpyDeviceToHost,
                       s->stream));
  HIPCK(hipStreamSynchronize(s->stream));
  *mass_out = s->host_res[0];
  if (n_nonfinite_out) *n_nonfinite_out = (int64_t)s->host_res[1];
  return CFD_OK;
}

int cfd_synchronize(cfd_solver* s) {
  g_err.clear();
  if (!s) return fail(CFD_E_ARG, "cfd_synchronize: null handle");
  HIPCK(hipStreamSynchronize(s->stream));
  return CFD_OK;
}

int cfd_bytes(cfd_solver* s, int64_t* bytes_out) {
  g_err.clear();
  if (!s || !bytes_out) return fail(CFD_E_ARG, "cfd_bytes: null argument");
  *bytes_out = device_bytes(s);
  return CFD_OK;
}

void cfd_destroy(cfd_solver* s) { release(s); }

// The closure's wall feature: distance in eighths of a cell, rounded, 0 for
// walls and non-positive values, saturating at 255 (31.875 cells).
static uint8_t quantize_wall(float d) {
  const float v = d * 8.0f;
  return !(v > 0.0f) ? 0 : (v >= 255.0f ? 255 : (uint8_t)(v + 0.5f));
}

// wall8 = the quantized exact distance transform of the current types.
static int compute_wall(cfd_solver* s) {
  // d^2 -> wall8 for d^2 = 0..1023, and 255 at 1024 (every d >= 32)
  constexpr size_t kTable = 1088;  // 1025 entries, padded to keep gx aligned
  uint8_t table[kTable] = {};
  for (uint32_t d2 = 0; d2 < 1024; ++d2) table[d2] = quantize_wall((float)std::sqrt((double)d2));
  table[1024] = 255;
  uint8_t* tmp = nullptr;
  HIPCK(hipMalloc(&tmp, kTable + 3 * s->N + 2));
  uint8_t* gx = tmp + kTable;
  uint16_t* h2 = reinterpret_cast<uint16_t*>(gx + ((s->N + 1) & ~(size_t)1));
  hipError_t e = hipMemcpyAsync(tmp, table, kTable, hipMemcpyHostToDevice, s->stream);
  if (e == hipSuccess) e = launch_wall_distance(s->tipo8, s->wall8, gx, h2, tmp, s->geom, s->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
  (void)hipFree(tmp);
  if (e != hipSuccess) return fail(CFD_E_HIP, "wall distance: %s", hipGetErrorString(e));
  s->wall_given = false;
  return CFD_OK;
}

// The closure's buffers; a new wall-distance field starts as the transform
// of the current types.
static int ensure_closure_buffers(cfd_solver* s) {
  if (!s->wall8) {
    HIPCK(hipMalloc(&s->wall8, s->N));
    int rc = compute_wall(s);
    if (rc) return rc;
  }
  if (!s->fallbacks) {
    HIPCK(hipMalloc(&s->fallbacks, kFallbackSlots * sizeof(unsigned long long)));
    HIPCK(hipMemsetAsync(s->fallbacks, 0, kFallbackSlots * sizeof(unsigned long long),
                         s->stream));
  }
  return CFD_OK;
}

int cfd_set_closure(cfd_solver* s, const cfd_closure* c) {
  g_err.clear();
  if (!s) return fail(CFD_E_ARG, "cfd_set_closure: null handle");
  HIPCK(hipStreamSynchronize(s->stream));
  if (!c) {
    if (s->mlp) (void)hipFree(s->mlp);
    s->mlp = nullptr;
    return CFD_OK;
  }
  for (int k = 0; k < 8; ++k)
    if (!std::isfinite(c->in_shift[k]) || !std::isfinite(c->in_scale[k]) ||
        !(c->in_lo[k] <= c->in_hi[k]))
      return fail(CFD_E_ARG, "cfd_set_closure: feature %d has a bad normalization or range", k);
  if (!(c->out_scale >= 0.0f)) return fail(CFD_E_ARG, "cfd_set_closure: out_scale must be >= 0");
  if (!known_activation(c->mlp.activation))
    return fail(CFD_E_ARG, "cfd_set_closure: unknown activation %d", (int)c->mlp.activation);
  int rc = ensure_closure_buffers(s);
  if (rc) return rc;
  MlpPacked pk;
  pack_closure(*c, &pk);
  if (!s->mlp) HIPCK(hipMalloc(&s->mlp, sizeof(MlpPacked)));
  // on the handle's stream (non-blocking: a null-stream copy would not order
  // before the steps that read it), complete before pk goes out of scope
  HIPCK(hipMemcpyAsync(s->mlp, &pk, sizeof pk, hipMemcpyHostToDevice, s->stream));
  HIPCK(hipStreamSynchronize(s->stream));
  s->mlp_act = c->mlp.activation;
  return CFD_OK;
}

int cfd_set_wall_distance(cfd_solver* s, const float* d, size_t n) {
  g_err.clear();
  if (!s || !d) return fail(CFD_E_ARG, "cfd_set_wall_distance: null argument");
  if (n != s->N)
    return fail(CFD_E_ARG, "cfd_set_wall_distance: %zu values, expected %zu", n, s->N);
  int rc = ensure_closure_buffers(s);
  if (rc) return rc;
  std::vector<uint8_t> q(n);
  for (size_t k = 0; k < n; ++k) q[k] = quantize_wall(d[k]);
  // on the handle's stream: steps already enqueued read the previous field
  HIPCK(hipMemcpyAsync(s->wall8, q.data(), n, hipMemcpyHostToDevice, s->stream));
  HIPCK(hipStreamSynchronize(s->stream));
  s->wall_given = true;
  return CFD_OK;
}

int cfd_compute_wall_distance(cfd_solver* s) {
  g_err.clear();
  if (!s) return fail(CFD_E_ARG, "cfd_compute_wall_distance: null handle");
  if (!s->wall8) return ensure_closure_buffers(s);  // computes the new field
  return compute_wall(s);
}

int cfd_get_wall_distance(cfd_solver* s, uint8_t* q, size_t n) {
  g_err.clear();
  if (!s || !q) return fail(CFD_E_ARG, "cfd_get_wall_distance: null argument");
  if (n != s->N)
    return fail(CFD_E_ARG, "cfd_get_wall_distance: %zu values, expected %zu", n, s->N);
  if (!s->wall8) {
    std::memset(q, 0, n);
    return CFD_OK;
  }
  HIPCK(hipStreamSynchronize(s->stream));
  HIPCK(hipMemcpy(q, s->wall8, n, hipMemcpyDeviceToHost));
  return CFD_OK;
}

int cfd_closure_fallbacks(cfd_solver* s, int64_t* count, int reset) {
  g_err.clear();
  if (!s || !count) return fail(CFD_E_ARG, "cfd_closure_fallbacks: null argument");
  *count = 0;
  if (!s->fallbacks) return CFD_OK;
  unsigned long long v[kFallbackSlots];
  HIPCK(hipStreamSynchronize(s->stream));
  HIPCK(hipMemcpy(v, s->fallbacks, sizeof v, hipMemcpyDeviceToHost));
  // the reset on the handle's stream, so that it lands before the next step
  if (reset) HIPCK(hipMemsetAsync(s->fallbacks, 0, sizeof v, s->stream));
  unsigned long long sum = 0;
  for (int k = 0; k < kFallbackSlots; ++k) sum += v[k];
  *count = (int64_t)sum;
  return CFD_OK;
}

int cfd_copy_bandwidth(size_t bytes, int iters, double* gbps_out, double* ms_out) {
  g_err.clear();
  if (!gbps_out || bytes < 16 || iters < 1)
    return fail(CFD_E_ARG, "cfd_copy_bandwidth: bad arguments");
  const size_t n4 = bytes / 16;
  float4 *a = nullptr, *b = nullptr;
  hipStream_t st = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  HIPCK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
  HIPCK(hipEventCreate(&e0));
  HIPCK(hipEventCreate(&e1));
  hipError_t e = hipMalloc(&a, n4 * 16);
  if (e == hipSuccess) e = hipMalloc(&b, n4 * 16);
  if (e == hipSuccess) e = hipMemsetAsync(a, 0, n4 * 16, st);
  if (e == hipSuccess) e = launch_copy(a, b, n4, st);  // warm-up
  if (e == hipSuccess) e = hipEventRecord(e0, st);
  for (int k = 0; k < iters && e == hipSuccess; ++k) e = launch_copy(a, b, n4, st);
  if (e == hipSuccess) e = hipEventRecord(e1, st);
  if (e == hipSuccess) e = hipEventSynchronize(e1);
  float ms = 0.f;
  if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
  if (a) (void)hipFree(a);
  if (b) (void)hipFree(b);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipStreamDestroy(st);
  if (e != hipSuccess) return fail(CFD_E_HIP, "cfd_copy_bandwidth: %s", hipGetErrorString(e));
  const double per = ms / iters;
  *gbps_out = 2.0 * (double)(n4 * 16) / (per * 1e-3) / 1e9;
  if (ms_out) *ms_out = per;
  return CFD_OK;
}

}  // extern "C"
