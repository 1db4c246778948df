// Learned sub-grid closure: a per-cell MLP 8 -> 32 -> 32 -> 1 (SiLU, SiLU,
// softplus) that replaces the Smagorinsky eddy viscosity inside the step.
//
// Inference runs on the matrix cores: bf16 operands, float32 accumulation,
// v_mfma_f32_16x16x32_bf16.  A wave steps 64 cells; each lane holds its own
// cell's 8 features as the B operand (k = 8 * (lane / 16) + v), and layer 1 is
// evaluated as four 16-cell blocks with the A operand (W1) read from a zero
// row in every lane group but one, so no lane ever gathers another lane's
// features.  A 16x16 result tile leaves hidden unit 4 * (lane / 16) + v of
// cell (lane % 16) in each lane, which is already the B layout of the next
// layer once the next weight matrix has its input index permuted (done on
// the host).  Layer 3 puts w3 in every row of A, so every lane group receives
// each cell's output and the owning lane keeps its own: no shuffle.
//
// The hidden activation is the network's own (cfd_mlp.activation).  A SiLU
// network runs in base 2 (pack_closure): layer 1 and 2 produce
// u = -log2(e) z, whose SiLU is u / (1 + 2^u) = -log2(e) SiLU(z), and w3
// carries the -ln 2 back, so a SiLU is one exponential, one add, one
// reciprocal and one multiply.  A hard-SiLU network, z * clamp(z/6 + 1/2,
// 0, 1), runs on z itself: one fma whose result the hardware clamps to
// [0, 1] and one multiply.
//
// The closure step (k_step_closure) is the step of lbm_kernels.hip with the
// eddy viscosity taken from the MLP: features are the non-equilibrium stress
// Pi_ab / rho (6), the strain rate |S| = 3 Q_s / (2 rho tau_s) implied by the
// Smagorinsky solution tau_s, and the wall distance; a fluid cell whose
// normalized features leave the trained range, or whose output is not finite,
// falls back to the Smagorinsky tau_s (counted per launch).  Every lane of a
// wave reaches the matrix instructions: tail lanes and wall cells run with
// zero features and store nothing.
//
// Training is hand-written too (float32): k_mlp_grad evaluates a minibatch,
// forward and backward, one sample per lane, and reduces the per-sample
// gradients of a workgroup through LDS in a fixed order; k_mlp_adam sums the
// workgroup partials in a fixed order and applies Adam.  The result does not
// depend on scheduling.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cmath>
#include <cstring>
#include <vector>

#include "../../include/cfd_mi355x.h"
#include "kernels.h"
#include "lattice.h"
#include "lbm_device.h"

namespace cfd {

typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
typedef float v4f __attribute__((ext_vector_type(4)));

// ------------------------------------------------------------------ inference

constexpr float kSixth = 1.0f / 6.0f, kThird = 1.0f / 3.0f;

// The hidden activation of the matrix-core path.  SiLU in base 2 on
// u = -log2(e) z: u / (1 + 2^u) = -log2(e) SiLU(z); for u beyond the
// exponential's range 2^u is +inf and the result 0.  Hard-SiLU on z:
// z * clamp(z/6 + 1/2, 0, 1) (a NaN z gives NaN).
template <int ACT>
__device__ __forceinline__ float act_fast(float u) {
  if constexpr (ACT == CFD_ACT_HARD_SILU)
    return u * __builtin_amdgcn_fmed3f(fmaf(u, kSixth, 0.5f), 0.0f, 1.0f);
  else
    return u * __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(u));
}

__device__ __forceinline__ float softplus_fast(float z) {
  if (z > 20.0f) return z;
  if (z < -15.0f) return __expf(z);
  return __logf(1.0f + __expf(z));
}

__device__ __forceinline__ v8bf ld8(const uint16_t* p) {
  return *reinterpret_cast<const v8bf*>(p);
}

__device__ __forceinline__ v4f ld4(const float* p) { return *reinterpret_cast<const v4f*>(p); }

template <int ACT>
__device__ __forceinline__ v8bf pack_act(const v4f& a, const v4f& b) {
  v8bf r;
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    r[v] = (__bf16)act_fast<ACT>(a[v]);
    r[v + 4] = (__bf16)act_fast<ACT>(b[v]);
  }
  return r;
}

// z (before softplus) of this lane's cell; xb = its normalized features in
// bf16 (zero for lanes that do not use the result).  All 64 lanes must call.
template <int ACT>
__device__ __forceinline__ float mlp_wave(const v8bf& xb, const MlpPacked& W) {
  const int l = threadIdx.x & 63, grp = l >> 4, row = l & 15;
  const v4f zero4 = {0.f, 0.f, 0.f, 0.f};
  // operands shared by the four 16-cell blocks
  const v8bf w2a = ld8(&W.w2[0][row][8 * grp]), w2b = ld8(&W.w2[1][row][8 * grp]);
  const v8bf a3 = ld8(&W.w3[8 * grp]);
  const v4f b1a = ld4(&W.b1[4 * grp]), b1b = ld4(&W.b1[16 + 4 * grp]);
  const v4f b2a = ld4(&W.b2[4 * grp]), b2b = ld4(&W.b2[16 + 4 * grp]);
  float mine = 0.0f;
#pragma unroll 1
  for (int gs = 0; gs < 4; ++gs) {
    // layer 1, cells 16 gs .. 16 gs + 15: W1 rows in lane group gs, the zero
    // row 16 in the others
    const int r1 = grp == gs ? row : 16;
    const v4f h0 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ld8(W.w1[0][r1]), xb, b1a, 0, 0, 0);
    const v4f h1 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ld8(W.w1[1][r1]), xb, b1b, 0, 0, 0);
    const v8bf x2 = pack_act<ACT>(h0, h1);
    // layer 2
    const v4f g0 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w2a, x2, b2a, 0, 0, 0);
    const v4f g1 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w2b, x2, b2b, 0, 0, 0);
    const v8bf x3 = pack_act<ACT>(g0, g1);
    // layer 3: every row of A is w3, so every row of the result holds the
    // outputs of the 16 cells; the lanes of group gs own them
    const v4f zz = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a3, x3, zero4, 0, 0, 0);
    if (grp == gs) mine = zz[0];
  }
  return mine + W.b3;
}

// Copy the packed MLP into LDS (all threads of the workgroup).
__device__ __forceinline__ void load_mlp(MlpPacked& dst, const MlpPacked* __restrict__ src) {
  const uint32_t* s = reinterpret_cast<const uint32_t*>(src);
  uint32_t* d = reinterpret_cast<uint32_t*>(&dst);
  for (uint32_t k = threadIdx.x; k < sizeof(MlpPacked) / 4; k += blockDim.x) d[k] = s[k];
  __syncthreads();
}

// ------------------------------------------------------------------ closure step

constexpr int kClosureBlock = 256;

template <class S, bool MACROS, int LB, int ACT>
__device__ __forceinline__ void step_cell_closure(uint32_t cell_in, bool live,
                                                  const GLOBAL typename S::T* __restrict__ src,
                                                  GLOBAL typename S::T* __restrict__ dst,
                                                  const GLOBAL uint8_t* __restrict__ tipo,
                                                  const GLOBAL uint8_t* __restrict__ wall8,
                                                  float4* __restrict__ macros, const Geom& g,
                                                  const StepParams& p, const MlpPacked& W,
                                                  unsigned long long* fallbacks) {
  const uint32_t cell = live ? cell_in : g.N - 1;  // tail lanes read a valid cell
  const Coord c = split(cell, g);
  const Wrap w = wrap_offsets(c, g);

  // pull with halfway bounce-back (ir.js:166-189), as in step_cell
  constexpr int kSz = (int)sizeof(typename S::T);
  const uint32_t tb = tipo[cell];
  const uint32_t cb = cell * (uint32_t)kSz;
  const Wrap wb = scaled(w, kSz);
  const uint32_t own = cell_byte_off<LB, typename S::T>(cb);
  const PopView<S, LB> in(src, g), out(dst, g);
  float gi[Q];
  uint32_t sb[Q];
  gi[0] = in.ld(0, own);
#pragma unroll
  for (int i = 1; i < Q; ++i) {
    sb[i] = cb + (uint32_t)nb_offset(wb, -cx(i), -cy(i), -cz(i));
    gi[i] = in.ld(i, cell_byte_off<LB, typename S::T>(sb[i]));
  }
  const uint32_t ct = tb & TYPE_MASK;
  const bool wall = is_wall(ct);
  if (!wall && (tb & NEAR_WALL)) bounce_back<false>(gi, sb, own, tipo, in, p, ct, nullptr);

  // the arithmetic runs in the stored units (g * unit), as the plain step's
  constexpr bool F = S::kFast;
  constexpr float unit = unit_of<S>();
  const Moments M = moments<F, S::kScaleLog2>(gi);
  float pab[6];
  if constexpr (F) {
    noneq_stress_fast(M, pab);
  } else {
    // the exact form's equilibrium is recomputed after the MLP rather than
    // held across it (fewer registers at the matrix instructions)
    float e[Q];
    equilibrium(M, e);
    noneq_stress(gi, e, pab);
  }
  const float qs = stress_norm<F>(pab);
  const float tau_s = smagorinsky_tau<F>(qs, M.rho, p);
  const bool fluid = live && !wall && ct != ESPELHO_Y && ct != ESPELHO_Z && ct != ENTRADA &&
                     ct != SAIDA;

  // features, normalized; outside the trained box the cell keeps tau_s.
  // Pi / rho and Q_s / rho are ratios of two values in the same units.
  float x[kMlpIn];
  const float inv_rho = F ? __builtin_amdgcn_rcpf(M.rho) : 1.0f / M.rho;
#pragma unroll
  for (int k = 0; k < 6; ++k) x[k] = pab[k] * inv_rho;
  x[6] = F ? 1.5f * qs * inv_rho * __builtin_amdgcn_rcpf(tau_s) : 1.5f * qs * inv_rho / tau_s;
  x[7] = (float)wall8[cell] * 0.125f;
  bool inside = true;
  v8bf xb;
#pragma unroll
  for (int k = 0; k < kMlpIn; ++k) {
    const float xn = (x[k] - W.shift[k]) * W.scale[k];
    inside = inside && xn >= W.lo[k] && xn <= W.hi[k];
    xb[k] = (__bf16)xn;
  }
  const bool use = fluid && inside;
  if (!use) xb = v8bf{};
  // the 19 pulled values wait in LDS while the MLP runs (fewer live registers
  // at the matrix instructions); the compiler barrier keeps them from being
  // forwarded in registers around it
  __shared__ float s_gi[Q][kClosureBlock];
#pragma unroll
  for (int i = 0; i < Q; ++i) s_gi[i][threadIdx.x] = gi[i];
  asm volatile("" ::: "memory");
  const float nu = W.out_scale * softplus_fast(mlp_wave<ACT>(xb, W));
  asm volatile("" ::: "memory");
#pragma unroll
  for (int i = 0; i < Q; ++i) gi[i] = s_gi[i][threadIdx.x];
  const bool ok = use && isfinite(nu);
  // fallback count: per wave, then per workgroup (below), then one atomic
  // per workgroup into one of kFallbackSlots counters (a single counter took
  // one atomic per wave on one address: ~30 us per step at 2.8 M cells)
  __shared__ uint32_t s_fb[kClosureBlock / 64];
  const unsigned long long fb = __ballot(fluid && !ok);
  if ((threadIdx.x & 63) == 0) s_fb[threadIdx.x >> 6] = (uint32_t)__popcll(fb);
  const float tau = ok ? p.tau0 + 3.0f * nu : tau_s;

  float d[Q];
  if (!boundary_cell(ct, gi, M, c.z, p, d, unit)) {
    if constexpr (F) {
      collide_fast<(S::kScaleLog2 > 0)>(gi, M, tau, p, d);
    } else {
      float e[Q];
      equilibrium(M, e);
      trt(gi, e, tau, p, d);
    }
    sponge<F>(c.x, p, d, unit);
  }
  // wall lanes of a wave that also holds other cells store zeros (never
  // read), so that every line of dst is written whole (no partial-line fetch)
  const bool mixed = __ballot(live && !wall) != 0;
  if (live) {
    if (!wall) {
      const uint32_t h = S::kStochastic ? sr_hash(cell, p.sr_seed) : 0u;
      f16x2 r = {};
      out.st_first(own, d[0], h, r);
#pragma unroll
      for (int i = 1; i < Q; i += 2) out.st_pair(i, own, d[i], d[i + 1], h, r);
    } else if (mixed) {
      out.st_zero(own);
    }
  }
  if (MACROS && live)
    macros[cell] = wall ? make_float4(0.f, 0.f, 0.f, 0.f)
                        : make_float4(M.ux, M.uy, M.uz, M.delta * (1.0f / unit));
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t n = 0;
#pragma unroll
    for (int w = 0; w < kClosureBlock / 64; ++w) n += s_fb[w];
    if (n) atomicAdd(&fallbacks[blockIdx.x & (kFallbackSlots - 1)], (unsigned long long)n);
  }
}

template <class S, bool MACROS, int LB, int ACT>
__global__ __launch_bounds__(kClosureBlock) __attribute__((amdgpu_waves_per_eu(5))) void k_step_closure(
    const typename S::T* __restrict__ src, typename S::T* __restrict__ dst,
    const uint8_t* __restrict__ tipo, const uint8_t* __restrict__ wall8,
    const MlpPacked* __restrict__ mlp, float4* __restrict__ macros, Geom g, StepParams p,
    unsigned long long* fallbacks) {
  __shared__ MlpPacked W;
  load_mlp(W, mlp);
  const uint32_t cell = blockIdx.x * kClosureBlock + threadIdx.x;
  step_cell_closure<S, MACROS, LB, ACT>(cell, cell < g.N, gptr(src), gptr(dst), gptr(tipo),
                                        gptr(wall8), macros, g, p, W, fallbacks);
}

template <int ACT>
static void launch_closure_act(Store sto, const void* src, void* dst, const uint8_t* tipo8,
                               const uint8_t* wall8, const MlpPacked* mlp, float4* macros,
                               bool write_macros, const Geom& g, const StepParams& p,
                               unsigned long long* fallbacks, hipStream_t st) {
  const dim3 grid((g.N + kClosureBlock - 1) / kClosureBlock);
  CFD_DISPATCH(sto, g, {
    if (write_macros)
      hipLaunchKernelGGL((k_step_closure<S, true, L, ACT>), grid, dim3(kClosureBlock), 0, st,
                         static_cast<const typename S::T*>(src), static_cast<typename S::T*>(dst),
                         tipo8, wall8, mlp, macros, g, p, fallbacks);
    else
      hipLaunchKernelGGL((k_step_closure<S, false, L, ACT>), grid, dim3(kClosureBlock), 0, st,
                         static_cast<const typename S::T*>(src), static_cast<typename S::T*>(dst),
                         tipo8, wall8, mlp, macros, g, p, fallbacks);
  });
}

hipError_t launch_step_closure(Store sto, const void* src, void* dst, const uint8_t* tipo8,
                               const uint8_t* wall8, const MlpPacked* mlp, int act,
                               float4* macros, bool write_macros, const Geom& g,
                               const StepParams& p, unsigned long long* fallbacks,
                               hipStream_t st) {
  if (act == CFD_ACT_HARD_SILU)
    launch_closure_act<CFD_ACT_HARD_SILU>(sto, src, dst, tipo8, wall8, mlp, macros, write_macros,
                                          g, p, fallbacks, st);
  else
    launch_closure_act<CFD_ACT_SILU>(sto, src, dst, tipo8, wall8, mlp, macros, write_macros, g,
                                     p, fallbacks, st);
  return hipGetLastError();
}

// Inference alone on given raw features (the same device path as the step).
template <int ACT>
__global__ __launch_bounds__(64) void k_mlp_infer(const float* __restrict__ X, uint32_t n,
                                                  const MlpPacked* __restrict__ mlp,
                                                  float* __restrict__ nu,
                                                  uint8_t* __restrict__ inside_out) {
  __shared__ MlpPacked W;
  load_mlp(W, mlp);
  const uint32_t s = blockIdx.x * 64 + threadIdx.x;
  const bool live = s < n;
  const uint32_t r = live ? s : 0;
  bool inside = true;
  v8bf xb;
#pragma unroll
  for (int k = 0; k < kMlpIn; ++k) {
    const float xn = (X[(size_t)r * kMlpIn + k] - W.shift[k]) * W.scale[k];
    inside = inside && xn >= W.lo[k] && xn <= W.hi[k];
    xb[k] = (__bf16)xn;
  }
  const float out = W.out_scale * softplus_fast(mlp_wave<ACT>(xb, W));
  if (live) {
    nu[s] = out;
    inside_out[s] = inside ? 1 : 0;
  }
}

// ------------------------------------------------------------------ training

// Parameter offsets in the cfd_mlp order.
constexpr int kOffW1 = 0, kOffB1 = 256, kOffW2 = 288, kOffB2 = 1312, kOffW3 = 1344,
              kOffB3 = 1376;
static_assert(kOffB3 + 1 == kMlpParams, "cfd_mlp layout");
constexpr int kGradSlots = kMlpParams + 1;  // + the loss

__device__ __forceinline__ float sigmoidf_exact(float a) { return 1.0f / (1.0f + expf(-a)); }

// The hidden activation in float32 and its derivative (training).
__device__ __forceinline__ float act_exact(float a, int act) {
  if (act == CFD_ACT_HARD_SILU) return a * fminf(fmaxf(fmaf(a, kSixth, 0.5f), 0.0f), 1.0f);
  return a * sigmoidf_exact(a);
}

__device__ __forceinline__ float act_grad(float a, int act) {
  if (act == CFD_ACT_HARD_SILU) return a <= -3.0f ? 0.0f : (a >= 3.0f ? 1.0f : fmaf(a, kThird, 0.5f));
  const float sg = sigmoidf_exact(a);
  return sg * (1.0f + a * (1.0f - sg));
}

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// One workgroup = one wave = 64 samples.  random = 1: sample index drawn
// from (seed, step, slot); random = 0: samples blockIdx.x * 64 + lane (full
// batch).  inv_count = 1 / (samples in the mean).
__global__ __launch_bounds__(64) void k_mlp_grad(const float* __restrict__ params,
                                                 const float* __restrict__ X,
                                                 const float* __restrict__ Y, uint64_t n,
                                                 int random, uint64_t seed, uint32_t step,
                                                 float inv_count, int act,
                                                 float* __restrict__ partial) {
  __shared__ float P[kMlpParams];
  __shared__ float xs[64][kMlpIn];
  __shared__ float h1s[64][kMlpHidden], d1s[64][kMlpHidden];
  __shared__ float h2s[64][kMlpHidden], d2s[64][kMlpHidden];
  __shared__ float dzs[64], ls[64];
  const int t = threadIdx.x;
  for (int k = t; k < kMlpParams; k += 64) P[k] = params[k];
  __syncthreads();

  uint64_t s;
  bool live;
  if (random) {
    s = mix64(seed ^ ((uint64_t)step << 32) ^ (uint64_t)(blockIdx.x * 64 + t)) % n;
    live = true;
  } else {
    s = (uint64_t)blockIdx.x * 64 + t;
    live = s < n;
    if (!live) s = 0;
  }
  float x[kMlpIn];
#pragma unroll
  for (int k = 0; k < kMlpIn; ++k) x[k] = X[s * kMlpIn + k];
  const float y = Y[s];

  // forward
  float a1[kMlpHidden], h1[kMlpHidden], a2[kMlpHidden], h2[kMlpHidden];
#pragma unroll
  for (int o = 0; o < kMlpHidden; ++o) {
    float acc = P[kOffB1 + o];
#pragma unroll
    for (int k = 0; k < kMlpIn; ++k) acc += P[kOffW1 + o * kMlpIn + k] * x[k];
    a1[o] = acc;
    h1[o] = act_exact(acc, act);
  }
#pragma unroll
  for (int o = 0; o < kMlpHidden; ++o) {
    float acc = P[kOffB2 + o];
#pragma unroll
    for (int k = 0; k < kMlpHidden; ++k) acc += P[kOffW2 + o * kMlpHidden + k] * h1[k];
    a2[o] = acc;
    h2[o] = act_exact(acc, act);
  }
  float z = P[kOffB3];
#pragma unroll
  for (int k = 0; k < kMlpHidden; ++k) z += P[kOffW3 + k] * h2[k];
  const float yhat = z > 20.0f ? z : log1pf(expf(z));
  const float err = yhat - y;

  // backward: L = mean (yhat - y)^2
  const float dz = live ? 2.0f * err * inv_count * sigmoidf_exact(z) : 0.0f;
#pragma unroll
  for (int o = 0; o < kMlpHidden; ++o) {
    d2s[t][o] = dz * P[kOffW3 + o] * act_grad(a2[o], act);
    h2s[t][o] = h2[o];
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < kMlpHidden; ++k) {
    float acc = 0.0f;
#pragma unroll
    for (int o = 0; o < kMlpHidden; ++o) acc += P[kOffW2 + o * kMlpHidden + k] * d2s[t][o];
    d1s[t][k] = acc * act_grad(a1[k], act);
    h1s[t][k] = h1[k];
  }
#pragma unroll
  for (int k = 0; k < kMlpIn; ++k) xs[t][k] = x[k];
  dzs[t] = dz;
  ls[t] = live ? err * err : 0.0f;
  __syncthreads();

  // workgroup sums over the 64 samples, in sample order
  float* out = partial + (size_t)blockIdx.x * kGradSlots;
  for (int q = t; q < kGradSlots; q += 64) {
    float acc = 0.0f;
    if (q < kOffB1) {
      const int o = q / kMlpIn, k = q % kMlpIn;
      for (int r = 0; r < 64; ++r) acc += d1s[r][o] * xs[r][k];
    } else if (q < kOffW2) {
      const int o = q - kOffB1;
      for (int r = 0; r < 64; ++r) acc += d1s[r][o];
    } else if (q < kOffB2) {
      const int o = (q - kOffW2) / kMlpHidden, k = (q - kOffW2) % kMlpHidden;
      for (int r = 0; r < 64; ++r) acc += d2s[r][o] * h1s[r][k];
    } else if (q < kOffW3) {
      const int o = q - kOffB2;
      for (int r = 0; r < 64; ++r) acc += d2s[r][o];
    } else if (q < kOffB3) {
      const int k = q - kOffW3;
      for (int r = 0; r < 64; ++r) acc += dzs[r] * h2s[r][k];
    } else if (q == kOffB3) {
      for (int r = 0; r < 64; ++r) acc += dzs[r];
    } else {
      for (int r = 0; r < 64; ++r) acc += ls[r];
    }
    out[q] = acc;
  }
}

// Sum of the workgroup partials (fixed order) -> gradient, then Adam when
// lr > 0.  Slot kMlpParams carries the summed squared error.
__global__ __launch_bounds__(256) void k_mlp_adam(float* __restrict__ params,
                                                  float* __restrict__ m, float* __restrict__ v,
                                                  const float* __restrict__ partial, int nwg,
                                                  uint32_t step, float lr, float beta1,
                                                  float beta2, float eps, float loss_scale,
                                                  float* __restrict__ grad_out,
                                                  float* __restrict__ loss_out) {
  const int q = blockIdx.x * 256 + threadIdx.x;
  if (q >= kGradSlots) return;
  float gsum = 0.0f;
  for (int w = 0; w < nwg; ++w) gsum += partial[(size_t)w * kGradSlots + q];
  if (q == kMlpParams) {
    if (loss_out) loss_out[step - 1] = gsum * loss_scale;
    return;
  }
  if (grad_out) grad_out[q] = gsum;
  if (lr > 0.0f) {
    const float mq = beta1 * m[q] + (1.0f - beta1) * gsum;
    const float vq = beta2 * v[q] + (1.0f - beta2) * gsum * gsum;
    m[q] = mq;
    v[q] = vq;
    const float mh = mq / (1.0f - powf(beta1, (float)step));
    const float vh = vq / (1.0f - powf(beta2, (float)step));
    params[q] -= lr * mh / (sqrtf(vh) + eps);
  }
}

// ------------------------------------------------------------------ wall distance
//
// Exact Euclidean distance transform of the wall cells (types 1, 4, 7), kept
// only as far as the closure's wall8 feature can tell distances apart: wall8
// is 8 d rounded and saturates at 255, so every d >= 32 (d^2 >= 1024) maps to
// 255 and the search stops there.  Three separable passes over squared
// distances (rows along x, then y, then z): the nearest wall of a cell within
// 32 cells lies within 31 cells on each axis, so windows of +-31 cells (and a
// row distance capped at 32) give the exact d^2 whenever d^2 < 1024 and a
// value >= 1024 otherwise.  The last pass maps d^2 through a table built on
// the host with the same quantization as cfd_set_wall_distance, so the field
// equals the host quantization of the exact transform bit for bit.
constexpr uint32_t kWdCap = 32, kWdCap2 = kWdCap * kWdCap;

__device__ __forceinline__ bool wd_wall(const uint8_t* t8, size_t cell) {
  return is_wall(t8[cell] & TYPE_MASK);
}

// Distance along x to the nearest wall cell of the row, capped at 32.  One
// wave per 64-cell segment of a row: three ballots hold the wall bits of the
// segment and of its two neighbours, and each lane reads its distance from
// the leading / trailing zeros of the 64-bit windows that end / start at its
// own cell.
__global__ __launch_bounds__(256) void k_wd_rows(const uint8_t* __restrict__ t8,
                                                 uint8_t* __restrict__ gx, uint32_t nx,
                                                 uint32_t segs, uint32_t waves) {
  const uint32_t wave = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (wave >= waves) return;  // uniform across the wave
  const int lane = (int)(threadIdx.x & 63);
  const uint32_t row = wave / segs;
  const int x = (int)((wave - row * segs) * 64) + lane;
  const size_t base = (size_t)row * nx;
  auto wall = [&](int xx) { return xx >= 0 && xx < (int)nx && wd_wall(t8, base + xx); };
  const uint64_t prev = __ballot(wall(x - 64));
  const uint64_t cur = __ballot(wall(x));
  const uint64_t next = __ballot(wall(x + 64));
  // bit 63 - k of left = cell x - k; bit k of right = cell x + k
  const uint64_t left = (cur << (63 - lane)) | (lane == 63 ? 0ull : prev >> (lane + 1));
  const uint64_t right = (cur >> lane) | (lane == 0 ? 0ull : next << (64 - lane));
  uint32_t d = kWdCap;
  if (left) d = min(d, (uint32_t)__clzll((long long)left));
  if (right) d = min(d, (uint32_t)(__ffsll((long long)right) - 1));
  if (x < (int)nx) gx[base + x] = (uint8_t)d;
}

// The windowed searches below take offsets in groups of kWdGroup, all loads
// of a group issued together (one memory round trip per group), and stop
// once the group's first offset alone reaches the best value.  Candidates
// past that point have d^2 >= best and cannot lower it, and out-of-range
// neighbours contribute a sentinel whose square is >= 1024, so the result
// is the plain min over the window.
constexpr uint32_t kWdGroup = 8;

// Squared distance over the x-y plane: min over |dy| <= 31 of dy^2 + gx^2,
// capped at 1024.
__global__ __launch_bounds__(256) void k_wd_cols(const uint8_t* __restrict__ gx,
                                                 uint16_t* __restrict__ h2, Geom g) {
  const uint32_t cell = blockIdx.x * 256 + threadIdx.x;
  if (cell >= g.N) return;
  const uint32_t y = (cell / g.nx) % g.ny;
  const uint32_t gc = gx[cell];
  uint32_t best = min(gc * gc, kWdCap2);
  for (uint32_t d0 = 1; d0 < kWdCap && d0 * d0 < best; d0 += kWdGroup) {
    uint32_t v[kWdGroup];
#pragma unroll
    for (uint32_t j = 0; j < kWdGroup; ++j) {
      const uint32_t d = d0 + j;
      const uint32_t lo = (d < kWdCap && y >= d) ? gx[cell - d * g.nx] : kWdCap;
      const uint32_t hi = (d < kWdCap && y + d < g.ny) ? gx[cell + d * g.nx] : kWdCap;
      v[j] = min(lo, hi);
    }
#pragma unroll
    for (uint32_t j = 0; j < kWdGroup; ++j) {
      const uint32_t d = d0 + j;
      best = min(best, d * d + v[j] * v[j]);
    }
  }
  h2[cell] = (uint16_t)best;
}

// Squared distance over the volume (min over |dz| <= 31 of dz^2 + h2), then
// the quantized wall distance from the host-built table (index 1024 = any
// d^2 >= 1024).
__global__ __launch_bounds__(256) void k_wd_planes(const uint16_t* __restrict__ h2,
                                                   const uint8_t* __restrict__ table,
                                                   uint8_t* __restrict__ wall8, Geom g) {
  const uint32_t cell = blockIdx.x * 256 + threadIdx.x;
  if (cell >= g.N) return;
  const uint32_t z = cell / g.nxny;
  uint32_t best = h2[cell];
  for (uint32_t d0 = 1; d0 < kWdCap && d0 * d0 < best; d0 += kWdGroup) {
    uint32_t v[kWdGroup];
#pragma unroll
    for (uint32_t j = 0; j < kWdGroup; ++j) {
      const uint32_t d = d0 + j;
      const uint32_t lo = (d < kWdCap && z >= d) ? (uint32_t)h2[cell - d * g.nxny] : kWdCap2;
      const uint32_t hi = (d < kWdCap && z + d < g.nz) ? (uint32_t)h2[cell + d * g.nxny] : kWdCap2;
      v[j] = min(lo, hi);
    }
#pragma unroll
    for (uint32_t j = 0; j < kWdGroup; ++j) best = min(best, (d0 + j) * (d0 + j) + v[j]);
  }
  wall8[cell] = table[min(best, kWdCap2)];
}

hipError_t launch_wall_distance(const uint8_t* tipo8, uint8_t* wall8, uint8_t* gx, uint16_t* h2,
                                const uint8_t* table, const Geom& g, hipStream_t st) {
  const uint32_t segs = (g.nx + 63) / 64, waves = g.ny * g.nz * segs;
  hipLaunchKernelGGL(k_wd_rows, dim3((waves + 3) / 4), dim3(256), 0, st, tipo8, gx, g.nx, segs,
                     waves);
  const dim3 cells((g.N + 255) / 256);
  hipLaunchKernelGGL(k_wd_cols, cells, dim3(256), 0, st, gx, h2, g);
  hipLaunchKernelGGL(k_wd_planes, cells, dim3(256), 0, st, h2, table, wall8, g);
  return hipGetLastError();
}

}  // namespace cfd

// ------------------------------------------------------------------ host side

using namespace cfd;

extern "C" void cfd__set_error(const char* msg);

namespace {

uint16_t bf16_rne(float f) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)((u >> 16) | ((u & 0xffff) ? 0x40 : 0));
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

int perm_k(int k) {
  const int q = k >> 3, v = k & 7;
  return v < 4 ? 4 * q + v : 16 + 4 * q + (v - 4);
}

int efail(int code, const char* what, hipError_t e) {
  char buf[256];
  snprintf(buf, sizeof buf, "%s: %s", what, hipGetErrorString(e));
  cfd__set_error(buf);
  return code;
}

}  // namespace

namespace cfd {

// The matrix operands (float32 products, then bf16).  A SiLU network in its
// base-2 form: W1, b1, b2 times -log2(e), so layers 1 and 2 produce
// u = -log2(e) z; W2 unchanged (its input is -log2(e) SiLU and its output is
// scaled like its bias); w3 times -ln 2.  A hard-SiLU network as it is.  Row
// 16 of each W1 tile stays zero.
void pack_closure(const cfd_closure& c, MlpPacked* out) {
  std::memset(out, 0, sizeof *out);
  const cfd_mlp& m = c.mlp;
  const bool base2 = m.activation == CFD_ACT_SILU;
  const float nl2e = base2 ? -kLog2e : 1.0f, nln2 = base2 ? -kLn2 : 1.0f;
  for (int t = 0; t < 2; ++t)
    for (int i = 0; i < 16; ++i) {
      for (int v = 0; v < 8; ++v)
        out->w1[t][i][v] = bf16_rne(nl2e * m.w1[(16 * t + i) * 8 + v]);
      for (int k = 0; k < 32; ++k) out->w2[t][i][k] = bf16_rne(m.w2[(16 * t + i) * 32 + perm_k(k)]);
    }
  for (int k = 0; k < 32; ++k) out->w3[k] = bf16_rne(nln2 * m.w3[perm_k(k)]);
  for (int o = 0; o < 32; ++o) {
    out->b1[o] = nl2e * m.b1[o];
    out->b2[o] = nl2e * m.b2[o];
  }
  out->b3 = m.b3;
  out->out_scale = c.out_scale;
  for (int k = 0; k < 8; ++k) {
    out->shift[k] = c.in_shift[k];
    out->scale[k] = c.in_scale[k];
    out->lo[k] = c.in_lo[k];
    out->hi[k] = c.in_hi[k];
  }
}

bool known_activation(int32_t act) {
  return act == CFD_ACT_SILU || act == CFD_ACT_HARD_SILU;
}

}  // namespace cfd

extern "C" {

int cfd_mlp_infer(const float* X, size_t n, const cfd_closure* c, float* nu, uint8_t* inside) {
  cfd__set_error("");
  if (!X || !c || !nu || !inside || n == 0 || n >= (1ull << 31) ||
      !known_activation(c->mlp.activation)) {
    cfd__set_error("cfd_mlp_infer: bad arguments");
    return CFD_E_ARG;
  }
  MlpPacked pk;
  pack_closure(*c, &pk);
  float *dX = nullptr, *dnu = nullptr;
  uint8_t* din = nullptr;
  MlpPacked* dm = nullptr;
  hipError_t e = hipMalloc(&dX, n * kMlpIn * sizeof(float));
  if (e == hipSuccess) e = hipMalloc(&dnu, n * sizeof(float));
  if (e == hipSuccess) e = hipMalloc(&din, n);
  if (e == hipSuccess) e = hipMalloc(&dm, sizeof pk);
  if (e == hipSuccess) e = hipMemcpy(dX, X, n * kMlpIn * sizeof(float), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(dm, &pk, sizeof pk, hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    const dim3 grid((unsigned)((n + 63) / 64));
    if (c->mlp.activation == CFD_ACT_HARD_SILU)
      hipLaunchKernelGGL(k_mlp_infer<CFD_ACT_HARD_SILU>, grid, dim3(64), 0, 0, dX, (uint32_t)n,
                         dm, dnu, din);
    else
      hipLaunchKernelGGL(k_mlp_infer<CFD_ACT_SILU>, grid, dim3(64), 0, 0, dX, (uint32_t)n, dm,
                         dnu, din);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpy(nu, dnu, n * sizeof(float), hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(inside, din, n, hipMemcpyDeviceToHost);
  (void)hipFree(dX);
  (void)hipFree(dnu);
  (void)hipFree(din);
  (void)hipFree(dm);
  if (e != hipSuccess) return efail(e == hipErrorOutOfMemory ? CFD_E_OOM : CFD_E_HIP,
                                    "cfd_mlp_infer", e);
  return CFD_OK;
}

// Shared driver of cfd_mlp_train / cfd_mlp_grad.
static int mlp_run(const float* X, const float* Y, size_t n, const cfd_train_cfg* cfg,
                   cfd_mlp* params, cfd_mlp* grad, double* loss_full, float* loss_curve) {
  const bool train = cfg != nullptr;
  const int batch = train ? cfg->batch : 0;
  const int nwg = train ? batch / 64 : (int)((n + 63) / 64);
  float *dX = nullptr, *dY = nullptr, *dP = nullptr, *dm = nullptr, *dv = nullptr,
        *dpart = nullptr, *dgrad = nullptr, *dloss = nullptr;
  const int steps = train ? cfg->steps : 1;
  const int act = params->activation;
  hipError_t e = hipMalloc(&dX, n * kMlpIn * sizeof(float));
  if (e == hipSuccess) e = hipMalloc(&dY, n * sizeof(float));
  if (e == hipSuccess) e = hipMalloc(&dP, kMlpParams * sizeof(float));
  if (e == hipSuccess) e = hipMalloc(&dm, kMlpParams * sizeof(float));
  if (e == hipSuccess) e = hipMalloc(&dv, kMlpParams * sizeof(float));
  if (e == hipSuccess) e = hipMalloc(&dpart, (size_t)nwg * kGradSlots * sizeof(float));
  if (e == hipSuccess) e = hipMalloc(&dgrad, kMlpParams * sizeof(float));
  if (e == hipSuccess) e = hipMalloc(&dloss, (size_t)steps * sizeof(float));
  if (e == hipSuccess) e = hipMemcpy(dX, X, n * kMlpIn * sizeof(float), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(dY, Y, n * sizeof(float), hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(dP, params, kMlpParams * sizeof(float), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemset(dm, 0, kMlpParams * sizeof(float));
  if (e == hipSuccess) e = hipMemset(dv, 0, kMlpParams * sizeof(float));
  const unsigned agrid = (kGradSlots + 255) / 256;
  for (int s = 1; e == hipSuccess && s <= steps; ++s) {
    if (train) {
      hipLaunchKernelGGL(k_mlp_grad, dim3(nwg), dim3(64), 0, 0, dP, dX, dY, (uint64_t)n, 1,
                         (uint64_t)cfg->seed, (uint32_t)s, 1.0f / (float)batch, act, dpart);
      hipLaunchKernelGGL(k_mlp_adam, dim3(agrid), dim3(256), 0, 0, dP, dm, dv, dpart, nwg,
                         (uint32_t)s, cfg->lr, cfg->beta1, cfg->beta2, cfg->eps,
                         1.0f / (float)batch, (float*)nullptr, dloss);
    } else {
      hipLaunchKernelGGL(k_mlp_grad, dim3(nwg), dim3(64), 0, 0, dP, dX, dY, (uint64_t)n, 0,
                         (uint64_t)0, 1u, 1.0f / (float)n, act, dpart);
      hipLaunchKernelGGL(k_mlp_adam, dim3(agrid), dim3(256), 0, 0, dP, dm, dv, dpart, nwg, 1u,
                         0.0f, 0.9f, 0.999f, 1e-8f, 1.0f / (float)n, dgrad, dloss);
    }
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipDeviceSynchronize();
  if (e == hipSuccess && train)
    e = hipMemcpy(params, dP, kMlpParams * sizeof(float), hipMemcpyDeviceToHost);
  if (e == hipSuccess && train && loss_curve)
    e = hipMemcpy(loss_curve, dloss, (size_t)steps * sizeof(float), hipMemcpyDeviceToHost);
  if (e == hipSuccess && !train && grad) {
    e = hipMemcpy(grad, dgrad, kMlpParams * sizeof(float), hipMemcpyDeviceToHost);
    grad->activation = act;
  }
  if (e == hipSuccess && !train && loss_full) {
    float l = 0.f;
    e = hipMemcpy(&l, dloss, sizeof(float), hipMemcpyDeviceToHost);
    *loss_full = l;
  }
  for (float* p : {dX, dY, dP, dm, dv, dpart, dgrad, dloss}) (void)hipFree(p);
  if (e != hipSuccess) return efail(e == hipErrorOutOfMemory ? CFD_E_OOM : CFD_E_HIP,
                                    train ? "cfd_mlp_train" : "cfd_mlp_grad", e);
  return CFD_OK;
}

int cfd_mlp_train(const float* X, const float* y, size_t n, const cfd_train_cfg* cfg,
                  cfd_mlp* params, float* loss_curve) {
  cfd__set_error("");
  if (!X || !y || !cfg || !params || n == 0 || cfg->steps < 1 || cfg->batch < 64 ||
      cfg->batch % 64 != 0 || !(cfg->lr > 0.0f) || !known_activation(params->activation)) {
    cfd__set_error("cfd_mlp_train: bad arguments (batch a multiple of 64, steps >= 1, lr > 0, "
                   "a known activation)");
    return CFD_E_ARG;
  }
  return mlp_run(X, y, n, cfg, params, nullptr, nullptr, loss_curve);
}

int cfd_mlp_grad(const float* X, const float* y, size_t n, const cfd_mlp* params, cfd_mlp* grad,
                 double* loss) {
  cfd__set_error("");
  if (!X || !y || !params || !grad || n == 0 || n >= (1ull << 31) ||
      !known_activation(params->activation)) {
    cfd__set_error("cfd_mlp_grad: bad arguments");
    return CFD_E_ARG;
  }
  cfd_mlp p = *params;
  return mlp_run(X, y, n, nullptr, &p, grad, loss, nullptr);
}

}  // extern "C"
