// Device helpers shared by the lattice kernels (lbm_kernels.hip) and the
// learned-closure kernels (closure.hip): storage formats, addressing, the
// (x, y, z) split with periodic neighbour offsets, and the per-cell pieces of
// the emitted step (moments, equilibrium, non-equilibrium stress, TRT pair
// collision), written once so that every kernel rounds identically (each in
// its storage format's kFast form).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "kernels.h"
#include "lattice.h"

namespace cfd {

// ------------------------------------------------------------------ storage

// Global address space.  Pointers that the kernels read from device memory
// (the batch descriptors) carry no address space; accessing through a
// GLOBAL-qualified pointer keeps their loads and stores in the global form
// with a wave-uniform base in SGPRs instead of FLAT with per-lane 64-bit
// addresses.
#define GLOBAL __attribute__((address_space(1)))

// Constant address space: tables that no kernel writes while it runs (the
// batch schedule and tunnel descriptors).  Read through such a pointer at a
// wave-uniform address, they come through scalar loads, whose waits do not
// also wait for the vector stores in flight.
#define CONSTANT __attribute__((address_space(4)))

template <class T>
__device__ __forceinline__ T cload(const T* p, size_t i) {
  T v;
  __builtin_memcpy(&v, (const CONSTANT T*)p + i, sizeof(T));
  return v;
}

template <class T>
__device__ __forceinline__ const GLOBAL T* gptr(const T* p) {
  return (const GLOBAL T*)p;
}
template <class T>
__device__ __forceinline__ GLOBAL T* gptr(T* p) {
  return (GLOBAL T*)p;
}

// kFast: the per-cell arithmetic of the step uses single-instruction
// reciprocals and square roots (1 ulp) and explicit fused multiply-adds: a
// quotient within ~2 ulp, inside the 2.5 ulp WGSL allows for the division of
// the reference's shader, and 8 correctly rounded divisions fewer per cell.
// On for float32 and float16 storage; StoreF32X (CFD_ROUND_EXACT) turns it
// off and rounds like the reference's shader text read with correctly
// rounded division and square root, operation for operation (the oracle's
// arithmetic, bit for bit).
struct StoreF32 {
  static constexpr bool kStochastic = false;
  static constexpr bool kFast = true;
  using T = float;
  using Bits = uint32_t;
  template <class P>
  static __device__ __forceinline__ float ld(P p) { return *p; }
  template <class P>
  static __device__ __forceinline__ void st(P p, float v) { *p = v; }
  // the step's arithmetic runs in the stored units: value * 2^kScaleLog2
  static constexpr int kScaleLog2 = 0;
  template <class P>
  static __device__ __forceinline__ float ld_s(P p) { return *p; }
  static __device__ __forceinline__ float dec_s(Bits b) { return __uint_as_float(b); }
  static __device__ __forceinline__ Bits enc_s(float v) { return __float_as_uint(v); }
};

struct StoreF32X : StoreF32 {
  static constexpr bool kFast = false;
};

// float16 deviations g_i = f_i - w_i scaled by 2^12 (exact power of two, so
// the scaling itself adds no rounding): |g| up to 16 stays finite and values
// down to 1.5e-8 stay normal.  Collision still runs in float32.
struct StoreF16 {
  static constexpr bool kStochastic = false;
  static constexpr bool kFast = true;
  using T = _Float16;
  template <class P>
  static __device__ __forceinline__ float ld(P p) {
    return static_cast<float>(*p) * 0.000244140625f;
  }
  template <class P>
  static __device__ __forceinline__ void st(P p, float v) { *p = static_cast<T>(v * 4096.0f); }
  using Bits = uint16_t;
  // The step keeps the populations in the stored units (g * 2^12) through
  // its arithmetic: scaling by a power of two commutes with every rounding,
  // so the collision rounds exactly as it would on g itself, without a
  // multiply per load and per store.
  static constexpr int kScaleLog2 = 12;
  template <class P>
  static __device__ __forceinline__ float ld_s(P p) { return static_cast<float>(*p); }
  static __device__ __forceinline__ float dec_s(Bits b) {
    return static_cast<float>(__builtin_bit_cast(T, b));
  }
  static __device__ __forceinline__ Bits enc_s(float v) {
    return __builtin_bit_cast(Bits, static_cast<T>(v));
  }
};

// The same float16 storage with stochastic rounding in the step's stores: the
// float32 value is rounded up or down to a neighbouring float16 with
// probability proportional to its distance, so the stored value is unbiased.
// A deterministic rounding loses every update smaller than half a float16 ulp
// (~2e-4 relative): a slowly decaying flow then stops decaying.  The rounding
// is the hardware's stochastic conversion (v_cvt_sr_f16_f32), which compares
// the discarded bits with the upper bits of a 32-bit random word; it is
// unbiased over the whole float16 range, subnormals included (measured on
// this GPU against 65 536 draws per value).  Loads, init and conversions are
// those of StoreF16.
typedef _Float16 f16x2 __attribute__((ext_vector_type(2)));

struct StoreF16S : StoreF16 {
  static constexpr bool kStochastic = true;
};

// A 32-bit random word for population i of a cell: one avalanche hash per
// cell and step, then the Weyl sequence h + i * 0x9E3779B9 (2^32 / golden
// ratio): every word is as uniform as the hash, and the 19 words of a cell
// are spread evenly around the circle, for one add per population.
__device__ __forceinline__ uint32_t sr_hash(uint32_t cell, uint32_t seed) {
  uint32_t h = cell * 0x9E3779B1u ^ seed;
  h ^= h >> 16;
  h *= 0x85EBCA6Bu;
  h ^= h >> 13;
  h *= 0xC2B2AE35u;
  h ^= h >> 16;
  return h;
}
__device__ __forceinline__ uint32_t sr_bits(uint32_t h, int i) {
  return h + (uint32_t)i * 0x9E3779B9u;
}

// 2^kScaleLog2 of a storage format: the step's unit of a population.
template <class S>
__device__ __forceinline__ constexpr float unit_of() {
  return (float)(1u << S::kScaleLog2);
}

// ------------------------------------------------------------------ indexing

// Element `idx` of the slice starting at `base`: a wave-uniform 64-bit base
// plus a 32-bit byte offset, so the load uses the scalar-base addressing form
// (no per-lane 64-bit address arithmetic).  idx * sizeof(T) < 2^32 because a
// handle holds fewer than 2^30 cells.
template <class T>
__device__ __forceinline__ const T* at(const T* base, uint32_t idx) {
  return reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) +
                                    (uint32_t)(idx * (uint32_t)sizeof(T)));
}
template <class T>
__device__ __forceinline__ T* at(T* base, uint32_t idx) {
  return reinterpret_cast<T*>(reinterpret_cast<char*>(base) +
                              (uint32_t)(idx * (uint32_t)sizeof(T)));
}
template <class T>
__device__ __forceinline__ const GLOBAL T* at(const GLOBAL T* base, uint32_t idx) {
  return (const GLOBAL T*)((const GLOBAL char*)base + (uint32_t)(idx * (uint32_t)sizeof(T)));
}
template <class T>
__device__ __forceinline__ GLOBAL T* at(GLOBAL T* base, uint32_t idx) {
  return (GLOBAL T*)((GLOBAL char*)base + (uint32_t)(idx * (uint32_t)sizeof(T)));
}

// Address of population i of cell c.  LB = 0: SoA, slice stride g.S, a
// wave-uniform 64-bit slice base plus a 32-bit offset.  LB > 0: blocks of
// 2^LB cells, pop[(c >> LB)][i][c & (2^LB - 1)], a 32-bit byte offset from
// the set's base (the host only selects it when a population set is below
// 4 GiB).
template <int LB, class T>
__device__ __forceinline__ T* pop_at(T* base, int i, uint32_t c, const Geom& g) {
  if constexpr (LB == 0) {
    return at(base + (size_t)i * g.S, c);
  } else {
    constexpr uint32_t B = 1u << LB;
    return at(base, (c >> LB) * (B * (uint32_t)Q) + (uint32_t)i * B + (c & (B - 1)));
  }
}

// Byte addressing of the step's 19 loads and 19 stores.  A population's
// address is a direction base (wave-uniform: i * S elements for SoA, i * 2^LB
// for blocks) plus a per-lane byte offset of the cell.  With blocks,
// the element index (s >> LB) * Q * 2^LB + i * 2^LB + (s & (2^LB - 1)) equals
// s + (Q - 1) * 2^LB * (s >> LB) + i * 2^LB, so from the cell's byte index
// sb = s * sizeof(T) the offset is one shift and one 24-bit multiply-add
// (s >> LB < 2^21 for a handle's fewer than 2^30 cells; the product stays
// below 2^32 because blocks are only used for sets below 4 GiB).
template <int LB, class T>
__device__ __forceinline__ uint32_t cell_byte_off(uint32_t sb) {
  if constexpr (LB == 0) {
    return sb;
  } else {
    constexpr uint32_t kLog = sizeof(T) == 4 ? 2u : (sizeof(T) == 2 ? 1u : 0u);
    static_assert((1u << kLog) == sizeof(T), "2- or 4-byte populations");
    constexpr uint32_t kStride = (uint32_t)(Q - 1) * (1u << LB) * (uint32_t)sizeof(T);
    return __umul24(sb >> (LB + kLog), kStride) + sb;
  }
}
template <class T>
__device__ __forceinline__ T* at_byte(T* base, uint32_t off) {
  return (T*)((GLOBAL char*)base + off);
}
template <class T>
__device__ __forceinline__ const T* at_byte(const T* base, uint32_t off) {
  return (const T*)((const GLOBAL char*)base + off);
}

// The step's view of a population set.  Blocked sets (LB > 0, below 4 GiB)
// are read and written with raw buffer instructions: one resource for the
// set, the direction's base as the scalar offset and the cell's byte offset
// as the only per-lane address (no 64-bit address arithmetic per access).
// SoA sets (which may exceed 4 GiB) use a wave-uniform 64-bit base per
// direction and the same 32-bit byte offset.  Values are in the stored units
// (g * unit_of<S>()), the units of the step's arithmetic.
constexpr int kBufferWord3 = 0x00020000;  // raw dword access, gfx9-family resource

template <class S, int LB>
struct PopView {
  using T = typename S::T;
  GLOBAL T* base;
  size_t slice;  // SoA: elements between direction slices
  __amdgpu_buffer_rsrc_t rsrc;

  __device__ __forceinline__ PopView(const GLOBAL T* p, const Geom& g)
      : base(const_cast<GLOBAL T*>(p)), slice(g.S) {
    if constexpr (LB > 0)
      rsrc = __builtin_amdgcn_make_buffer_rsrc((void*)base, (short)0, (int)0xffffffff,
                                               kBufferWord3);
  }
  static __device__ __forceinline__ int dir_off(int i) {
    return i * (int)(1u << LB) * (int)sizeof(T);
  }
  __device__ __forceinline__ float ld(int i, uint32_t off) const {
    if constexpr (LB == 0) {
      return S::ld_s(at_byte(base + (size_t)i * slice, off));
    } else if constexpr (sizeof(T) == 4) {
      return S::dec_s((typename S::Bits)__builtin_amdgcn_raw_buffer_load_b32(rsrc, (int)off,
                                                                             dir_off(i), 0));
    } else {
      return S::dec_s((typename S::Bits)__builtin_amdgcn_raw_buffer_load_b16(rsrc, (int)off,
                                                                             dir_off(i), 0));
    }
  }
  // raw bits of population i of the cell at byte offset off
  __device__ __forceinline__ void st_bits(int i, uint32_t off, typename S::Bits b) const {
    if constexpr (LB == 0) {
      *(GLOBAL typename S::Bits*)at_byte(base + (size_t)i * slice, off) = b;
    } else if constexpr (sizeof(T) == 4) {
      __builtin_amdgcn_raw_buffer_store_b32((int)b, rsrc, (int)off, dir_off(i), 0);
    } else {
      __builtin_amdgcn_raw_buffer_store_b16((short)b, rsrc, (int)off, dir_off(i), 0);
    }
  }
  // Stores of the cell at byte offset off, values in step units: rounded to
  // nearest, or stochastically for StoreF16S (h = sr_hash of the cell for
  // this step).  st_first stores population 0, st_pair the opposite pair
  // (i, i + 1): float16 sets convert the pair into the two halves of one
  // register r (to nearest with one packed conversion, stochastically with one
  // conversion into each half, so no register is cleared per value).
  __device__ __forceinline__ void st_first(uint32_t off, float v, uint32_t h, f16x2& r) const {
    if constexpr (sizeof(T) == 4) {
      st_bits(0, off, S::enc_s(v));
    } else if constexpr (S::kStochastic) {
      r = __builtin_amdgcn_cvt_sr_f16_f32(r, v, sr_bits(h, 0), false);
      st_bits(0, off, __builtin_bit_cast(uint32_t, r) & 0xffffu);
    } else {
      st_bits(0, off, S::enc_s(v));
    }
  }
  __device__ __forceinline__ void st_pair(int i, uint32_t off, float a, float b, uint32_t h,
                                          f16x2& r) const {
    if constexpr (sizeof(T) == 4) {
      st_bits(i, off, S::enc_s(a));
      st_bits(i + 1, off, S::enc_s(b));
    } else {
      if constexpr (S::kStochastic) {
        r = __builtin_amdgcn_cvt_sr_f16_f32(r, a, sr_bits(h, i), false);
        r = __builtin_amdgcn_cvt_sr_f16_f32(r, b, sr_bits(h, i + 1), true);
      } else {
        r = f16x2{(_Float16)a, (_Float16)b};
      }
      const uint32_t w = __builtin_bit_cast(uint32_t, r);
      st_bits(i, off, w & 0xffffu);
      st_bits(i + 1, off, w >> 16);
    }
  }
  // zeros in every population of the cell (wall lanes of a mixed wave)
  __device__ __forceinline__ void st_zero(uint32_t off) const {
#pragma unroll
    for (int i = 0; i < Q; ++i) st_bits(i, off, 0);
  }
};

__device__ __forceinline__ uint32_t fast_div(uint32_t n, uint32_t m, uint32_t s) {
  return (__umulhi(n, m) + n) >> s;
}

// Logical workgroup of this launch: physical workgroup b runs on XCD b % 8;
// XCD k owns logical workgroups [k*per_xcd, (k+1)*per_xcd).
__device__ __forceinline__ uint32_t logical_block(const Geom& g) {
  const uint32_t b = blockIdx.x;
  return (b % kXcds) * g.per_xcd + b / kXcds;
}
// Chunked XCD mapping: XCD k steps through chunks k, k + 8, k + 16, ... of
// 2^chunk_log2 consecutive workgroups (chunk_log2 = 0: the launch order).  A
// workgroup's x-neighbour lines at its block boundary then belong to a block
// stepped by the same XCD, whose L2 already holds them.
__device__ __forceinline__ uint32_t logical_block_chunked(uint32_t chunk_log2) {
  const uint32_t b = blockIdx.x;
  const uint32_t xcd = b % kXcds, idx = b / kXcds;
  const uint32_t chunk = idx >> chunk_log2, within = idx & ((1u << chunk_log2) - 1u);
  return ((chunk * kXcds + xcd) << chunk_log2) + within;
}

struct Coord {
  uint32_t x, y, z;
};

__device__ __forceinline__ Coord split(uint32_t cell, const Geom& g) {
  Coord c;
  c.z = fast_div(cell, g.mnxy, g.snxy);
  const uint32_t r = cell - c.z * g.nxny;
  c.y = fast_div(r, g.mnx, g.snx);
  c.x = r - c.y * g.nx;
  return c;
}

// Signed index offsets to the periodic neighbours x-1, x+1 (and y, z), the
// select-based wrap of wgsl.js:159-182 (no modulo).
struct Wrap {
  int xm, xp, ym, yp, zm, zp;
};

__device__ __forceinline__ Wrap wrap_offsets(const Coord& c, const Geom& g) {
  Wrap w;
  const int nx = (int)g.nx, ny = (int)g.ny, nz = (int)g.nz, nxny = (int)g.nxny;
  w.xm = c.x == 0 ? nx - 1 : -1;
  w.xp = c.x == g.nx - 1 ? -(nx - 1) : 1;
  w.ym = c.y == 0 ? (ny - 1) * nx : -nx;
  w.yp = c.y == g.ny - 1 ? -(ny - 1) * nx : nx;
  w.zm = c.z == 0 ? (nz - 1) * nxny : -nxny;
  w.zp = c.z == g.nz - 1 ? -(nz - 1) * nxny : nxny;
  return w;
}

// The same offsets in bytes of a T-sized population.
__device__ __forceinline__ Wrap scaled(const Wrap& w, int sz) {
  return Wrap{w.xm * sz, w.xp * sz, w.ym * sz, w.yp * sz, w.zm * sz, w.zp * sz};
}

// Offset of the cell at x + (dx, dy, dz), each component in {-1, 0, 1}.
__device__ __forceinline__ int nb_offset(const Wrap& w, int dx, int dy, int dz) {
  const int ox = dx < 0 ? w.xm : (dx > 0 ? w.xp : 0);
  const int oy = dy < 0 ? w.ym : (dy > 0 ? w.yp : 0);
  const int oz = dz < 0 ? w.zm : (dz > 0 ? w.zp : 0);
  return ox + oy + oz;
}

// c_i . v written as the emitted literal arithmetic (ir.js:116-126): the
// non-zero components in x, y, z order, a leading minus for a negative first
// component, then + or -.
__device__ __forceinline__ float dotc(int i, float vx, float vy, float vz) {
  float acc = 0.0f;
  bool first = true;
  const int c[3] = {cx(i), cy(i), cz(i)};
  const float v[3] = {vx, vy, vz};
#pragma unroll
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

// Halfway bounce-back of a cell flagged NEAR_WALL (ir.js:166-189): for every
// direction i whose upstream neighbour x - c_i is a wall (SOLIDO, SOLIDO_MOVEL,
// PAREDE), g_i = src[opp i][x] + 2 w_i (c_i . u_w) / c_s^2 with u_w = beltU on
// SOLIDO_MOVEL and 0 elsewhere.  sb[i] is the byte offset of x - c_i in units
// of the population type, own the cell's own byte offset in the set.
//
// The 18 neighbour types are requested together and reduced to masks, then
// every bounced direction is re-loaded under its mask bit before any is used:
// two memory round trips however many links the cell has (a test-and-load per
// direction waits on memory up to 36 times, which made the no-slip floor's
// first fluid plane cost ~3 % of a 320x160x128 step).
//
// FORCE also accumulates, into f, the momentum exchange c_j 2 (g_j + w_j) of
// every SOLIDO link of a FLUIDO cell (wgsl.js:358-379), j = opp(i), from the
// re-loaded g_j before the wall term is added.  gi is in the step's units
// (g * unit_of<S>()).
template <bool FORCE, class S, int LB>
__device__ __forceinline__ void bounce_back(float (&gi)[Q], const uint32_t (&sb)[Q],
                                            uint32_t own, const GLOBAL uint8_t* __restrict__ tipo,
                                            const PopView<S, LB>& in, const StepParams& p,
                                            uint32_t ct, double* f) {
  constexpr uint32_t kSz = (uint32_t)sizeof(typename S::T);
  uint32_t tn[Q];
#pragma unroll
  for (int i = 1; i < Q; ++i) tn[i] = tipo[sb[i] / kSz];
  uint32_t wall_m = 0u, moving_m = 0u, solid_m = 0u;
#pragma unroll
  for (int i = 1; i < Q; ++i) {
    const uint32_t t = tn[i] & TYPE_MASK;
    if (is_wall(t)) wall_m |= 1u << i;
    if (t == SOLIDO_MOVEL) moving_m |= 1u << i;
    if (t == SOLIDO) solid_m |= 1u << i;
  }
#pragma unroll
  for (int i = 1; i < Q; ++i)
    if ((wall_m >> i) & 1u) gi[i] = in.ld(opp(i), own);
#pragma unroll
  for (int i = 1; i < Q; ++i) {
    if (!((wall_m >> i) & 1u)) continue;
    const bool moving = (moving_m >> i) & 1u;
    const float cu = dotc(i, moving ? p.belt[0] : 0.f, moving ? p.belt[1] : 0.f,
                          moving ? p.belt[2] : 0.f);
    if (FORCE && ((solid_m >> i) & 1u) && ct == FLUIDO) {
      const int j = opp(i);
      const double q = 2.0 * ((double)gi[i] * (1.0 / (double)unit_of<S>()) + wd(j));
      if (cx(j)) f[0] += cx(j) > 0 ? q : -q;
      if (cy(j)) f[1] += cy(j) > 0 ? q : -q;
      if (cz(j)) f[2] += cz(j) > 0 ? q : -q;
    }
    gi[i] = gi[i] + (unit_of<S>() * kwall(i)) * cu;
  }
}

// ------------------------------------------------------------------ cell physics

// Moments of the pulled populations (ir.js:192-213).
//
// Exact form: left-to-right sums as emitted, delta = sum g_i, rho = 1 + delta,
// m = sum c_i g_i, u = m / rho.
//
// FAST form, in the step's units (g * 2^SL2, so rho = 2^SL2 + delta): delta
// as emitted; the opposite pairs (2k-1, 2k) give sums a_k and differences
// b_k, then m_a = sum_k c_a(2k-1) b_k and the second moments S_ab =
// sum_i c_ia c_ib g_i = sum_k c_a(2k-1) c_b(2k-1) a_k: 63 adds give delta,
// the momentum and all six second moments (the emitted form takes 45 for
// delta and m alone, and Pi another 54 through the 19 equilibria).
struct Moments {
  float delta, rho, ux, uy, uz, uu;
  float m[3];   // FAST: momentum rho u
  float s2[6];  // FAST: sum c_ia c_ib g_i in the order xx, yy, zz, xy, xz, yz
};

CFD_HD constexpr int c_axis(int i, int a) { return a == 0 ? cx(i) : (a == 1 ? cy(i) : cz(i)); }

// Signed sum of v[k] over the pairs k = 1..9 with weight c(2k-1) in {-1, 0, 1},
// left to right.
template <class C>
__device__ __forceinline__ float pair_sum(const float (&v)[10], C coef) {
  float acc = 0.0f;
  bool first = true;
#pragma unroll
  for (int k = 1; k <= 9; ++k) {
    const int c = coef(2 * k - 1);
    if (c == 0) continue;
    if (first) {
      acc = c > 0 ? v[k] : -v[k];
      first = false;
    } else {
      acc = c > 0 ? acc + v[k] : acc - v[k];
    }
  }
  return acc;
}

template <bool FAST = false, int SL2 = 0>
__device__ __forceinline__ Moments moments(const float (&gi)[Q]) {
  Moments M;
  if constexpr (FAST) {
    float a[10], b[10];
#pragma unroll
    for (int k = 1; k <= 9; ++k) {
      a[k] = gi[2 * k - 1] + gi[2 * k];
      b[k] = gi[2 * k - 1] - gi[2 * k];
    }
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) {
      M.m[ax] = pair_sum(b, [ax](int i) { return c_axis(i, ax); });
      M.s2[ax] = pair_sum(a, [ax](int i) { return c_axis(i, ax) * c_axis(i, ax); });
    }
    M.s2[3] = pair_sum(a, [](int i) { return cx(i) * cy(i); });
    M.s2[4] = pair_sum(a, [](int i) { return cx(i) * cz(i); });
    M.s2[5] = pair_sum(a, [](int i) { return cy(i) * cz(i); });
    // delta in the emitted order: rho = 1 + delta is the divisor of u, and a
    // cell whose rho is small (a far-from-equilibrium state) would magnify
    // any other rounding of it
    float delta = gi[0];
#pragma unroll
    for (int i = 1; i < Q; ++i) delta = delta + gi[i];
    M.delta = delta;
    M.rho = (float)(1u << SL2) + delta;
    const float r = __builtin_amdgcn_rcpf(M.rho);
    M.ux = M.m[0] * r;
    M.uy = M.m[1] * r;
    M.uz = M.m[2] * r;
    M.uu = fmaf(M.ux, M.ux, fmaf(M.uy, M.uy, M.uz * M.uz));
    return M;
  } else {
    static_assert(SL2 == 0, "the exact form runs on float32 storage");
    float delta = gi[0];
#pragma unroll
    for (int i = 1; i < Q; ++i) delta = delta + gi[i];
    M.delta = delta;
    M.rho = 1.0f + delta;
    float m[3];
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) {
      float acc = 0.0f;
      bool first = true;
#pragma unroll
      for (int i = 1; i < Q; ++i) {
        const int ca = c_axis(i, ax);
        if (ca == 0) continue;
        if (first) {
          acc = ca > 0 ? gi[i] : -gi[i];
          first = false;
        } else {
          acc = ca > 0 ? acc + gi[i] : acc - gi[i];
        }
      }
      m[ax] = acc;
    }
    M.ux = m[0] / M.rho;
    M.uy = m[1] / M.rho;
    M.uz = m[2] / M.rho;
    M.uu = M.ux * M.ux + M.uy * M.uy + M.uz * M.uz;
    return M;
  }
}

// Shifted equilibrium (ir.js:216-232), the exact form's (the fast form never
// forms the 19 equilibria: noneq_stress_fast and collide_fast).
__device__ __forceinline__ void equilibrium(const Moments& M, float (&e)[Q]) {
  e[0] = wf(0) * (M.delta - 1.5f * M.rho * M.uu);
#pragma unroll
  for (int i = 1; i < Q; ++i) {
    const float cu = dotc(i, M.ux, M.uy, M.uz);
    e[i] = wf(i) * (M.delta + M.rho * (3.0f * cu + 4.5f * cu * cu - 1.5f * M.uu));
  }
}

// Non-equilibrium momentum flux Pi_ab = sum c_ia c_ib (g_i - e_i) in the
// order xx, yy, zz, xy, xz, yz (ir.js:262-285).
__device__ __forceinline__ void noneq_stress(const float (&gi)[Q], const float (&e)[Q],
                                             float (&pab)[6]) {
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    const int a = k < 3 ? k : (k == 5 ? 1 : 0);
    const int b = k < 3 ? k : (k == 3 ? 1 : 2);
    float acc = 0.0f;
    bool first = true;
#pragma unroll
    for (int i = 0; i < Q; ++i) {
      const int ca = a == 0 ? cx(i) : (a == 1 ? cy(i) : cz(i));
      const int cb = b == 0 ? cx(i) : (b == 1 ? cy(i) : cz(i));
      const int coef = ca * cb;
      if (coef == 0) continue;
      const float n = gi[i] - e[i];
      if (first) {
        acc = coef > 0 ? n : -n;
        first = false;
      } else {
        acc = coef > 0 ? acc + n : acc - n;
      }
    }
    pab[k] = acc;
  }
}

// FAST: Pi from the moments.  The shifted equilibrium's second moment is
// sum c_ia c_ib e_i = delta/3 delta_ab + rho u_a u_b (the lattice's
// fourth-order isotropy, sum w c_a c_b c_c c_d = (delta_ab delta_cd +
// delta_ac delta_bd + delta_ad delta_bc) / 9), so Pi_ab = S_ab - m_a u_b -
// delta/3 delta_ab without forming the 19 equilibria (the same quantity as
// noneq_stress up to float32 rounding).
__device__ __forceinline__ void noneq_stress_fast(const Moments& M, float (&pab)[6]) {
  const float d3 = M.delta * 0.333333333f;
  pab[0] = fmaf(-M.m[0], M.ux, M.s2[0] - d3);
  pab[1] = fmaf(-M.m[1], M.uy, M.s2[1] - d3);
  pab[2] = fmaf(-M.m[2], M.uz, M.s2[2] - d3);
  pab[3] = fmaf(-M.m[0], M.uy, M.s2[3]);
  pab[4] = fmaf(-M.m[0], M.uz, M.s2[4]);
  pab[5] = fmaf(-M.m[1], M.uz, M.s2[5]);
}

// Q_s = sqrt(2 Pi:Pi) (ir.js:286-289).
template <bool FAST = false>
__device__ __forceinline__ float stress_norm(const float (&pab)[6]) {
  if constexpr (FAST) {
    const float dia = fmaf(pab[0], pab[0], fmaf(pab[1], pab[1], pab[2] * pab[2]));
    const float off = fmaf(pab[3], pab[3], fmaf(pab[4], pab[4], pab[5] * pab[5]));
    return __builtin_amdgcn_sqrtf(2.0f * fmaf(2.0f, off, dia));
  } else {
    const float pp = pab[0] * pab[0] + pab[1] * pab[1] + pab[2] * pab[2] +
                     2.0f * (pab[3] * pab[3] + pab[4] * pab[4] + pab[5] * pab[5]);
    return sqrtf(2.0f * pp);
  }
}

// Closed-form Smagorinsky total relaxation time (ir.js:290-297).
template <bool FAST = false>
__device__ __forceinline__ float smagorinsky_tau(float qs, float rho, const StepParams& p) {
  if constexpr (FAST)
    return 0.5f * (p.tau0 + __builtin_amdgcn_sqrtf(
                               fmaf(p.c18cs2 * qs, __builtin_amdgcn_rcpf(rho), p.tau0sq)));
  else
    return 0.5f * (p.tau0 + sqrtf(p.tau0sq + p.c18cs2 * qs / rho));
}

// TRT collision per opposite pair with omega- from the fixed magic
// parameter (ir.js:298-323), the exact form's.
__device__ __forceinline__ void trt(const float (&gi)[Q], const float (&e)[Q], float tau,
                                    const StepParams& p, float (&d)[Q]) {
  const float omP = 1.0f / tau;
  const float omM = 1.0f / (p.magic / (1.0f / omP - 0.5f) + 0.5f);
  d[0] = gi[0] - omP * (gi[0] - e[0]);
#pragma unroll
  for (int i = 1; i < Q; i += 2) {
    const int j = i + 1;
    const float pe = 0.5f * ((gi[i] + gi[j]) - (e[i] + e[j]));
    const float mo = 0.5f * ((gi[i] - gi[j]) - (e[i] - e[j]));
    d[i] = gi[i] - omP * pe - omM * mo;
    d[j] = gi[j] - omP * pe + omM * mo;
  }
}

// FAST TRT (ir.js:298-323) from the moments, per opposite pair (i, j = i + 1)
// with cu = c_i . u (c_j = -c_i): the equilibrium's pair sum
// e_i + e_j = w (2 delta + rho (9 cu^2 - 3 uu)) and difference
// e_i - e_j = 6 w rho cu, so no equilibrium array is formed.  Then, with
// 1-ulp reciprocals for omega+ = 1/tau and omega- = 1 / (Lambda / (tau - 1/2) + 1/2),
// g -/+ (omega+/2) (sum_g - sum_e) -/+ (omega-/2) (dif_g - dif_e).
//
// EXPANDED (the float16 steps): the same update with the g terms gathered,
// d_i = A g_i + B g_j + (P + M), d_j = A g_j + B g_i + (P - M), where
// A = 1 - omega+/2 - omega-/2, B = omega-/2 - omega+/2, P = (omega+/2)(e_i + e_j)
// and M = (omega-/2)(e_i - e_j): two instructions fewer per pair.  The terms
// then cancel to the size of the result (a few float32 ulp of it, against
// the 2^-11 of float16 storage), so the float32 steps keep the form above.
template <bool EXPANDED = false>
__device__ __forceinline__ void collide_fast(const float (&gi)[Q], const Moments& M, float tau,
                                             const StepParams& p, float (&d)[Q]) {
  const float omP = __builtin_amdgcn_rcpf(tau);
  const float omM =
      __builtin_amdgcn_rcpf(fmaf(p.magic, __builtin_amdgcn_rcpf(tau - 0.5f), 0.5f));
  const float hp = 0.5f * omP, hm = 0.5f * omM;
  const float e0 = wf(0) * fmaf(-M.rho, 1.5f * M.uu, M.delta);
  d[0] = fmaf(-omP, gi[0] - e0, gi[0]);
  // t = 2 delta - 3 rho uu; per weight class A = w t, B = 9 w rho, C = 6 w rho
  const float t = fmaf(-3.0f * M.rho, M.uu, 2.0f * M.delta);
  if constexpr (EXPANDED) {
    const float A = (1.0f - hp) - hm, B = hm - hp;
    const float hpt = hp * t, hpr = hp * M.rho, hmr = hm * M.rho;
    const float HAa = wf(1) * hpt, HAd = wf(7) * hpt;
    const float HBa = (9.0f * wf(1)) * hpr, HBd = (9.0f * wf(7)) * hpr;
    const float HCa = (6.0f * wf(1)) * hmr, HCd = (6.0f * wf(7)) * hmr;
#pragma unroll
    for (int i = 1; i < Q; i += 2) {
      const int j = i + 1;
      const bool axis = i <= 6;
      const float cu = dotc(i, M.ux, M.uy, M.uz);
      const float P = fmaf(axis ? HBa : HBd, cu * cu, axis ? HAa : HAd);
      const float Mm = (axis ? HCa : HCd) * cu;
      d[i] = fmaf(A, gi[i], fmaf(B, gi[j], P + Mm));
      d[j] = fmaf(A, gi[j], fmaf(B, gi[i], P - Mm));
    }
    return;
  }
  const float Aa = wf(1) * t, Ad = wf(7) * t;
  const float Ba = (9.0f * wf(1)) * M.rho, Bd = (9.0f * wf(7)) * M.rho;
  const float Ca = (6.0f * wf(1)) * M.rho, Cd = (6.0f * wf(7)) * M.rho;
#pragma unroll
  for (int i = 1; i < Q; i += 2) {
    const int j = i + 1;
    const bool axis = i <= 6;
    const float cu = dotc(i, M.ux, M.uy, M.uz);
    const float ep = fmaf(axis ? Ba : Bd, cu * cu, axis ? Aa : Ad);
    const float em = (axis ? Ca : Cd) * cu;
    // g_i + (g_j - e+): not the moments' pair sum, which would otherwise stay
    // live from the moments to here (18 registers, two waves per SIMD)
    const float sp = gi[i] + (gi[j] - ep);
    const float sm = gi[i] - (gi[j] + em);
    d[i] = fmaf(-hm, sm, fmaf(-hp, sp, gi[i]));
    d[j] = fmaf(hm, sm, fmaf(-hp, sp, gi[j]));
  }
}

// Sponge (ir.js:458-478): mix(dst, teq, sigma^2).  For x <= start sigma^2 = 0
// and mix returns dst unchanged while teq is finite, so the blend is skipped
// there; a non-finite free stream (sponge_all, uniform) blends every fluid
// cell as the reference's mix does.
// d in the step's units (g * unit); the free-stream state is scaled to them.
// FAST: sigma from one reciprocal, clamped by a median, and the blend as
// d + sigma^2 (unit tinf - d) = fma(sigma^2 unit, tinf, fma(-sigma^2, d, d)),
// two instructions per population with tinf read as a scalar operand.
template <bool FAST = false>
__device__ __forceinline__ void sponge(uint32_t x, const StepParams& p, float (&d)[Q],
                                       float unit = 1.0f) {
  const float dxs = (float)x - p.sponge_start;
  if (dxs > 0.0f || p.sponge_all) {
    if constexpr (FAST) {
      const float sigma =
          __builtin_amdgcn_fmed3f(dxs * __builtin_amdgcn_rcpf(p.sponge_len_m), 0.0f, 1.0f);
      const float sig2 = sigma * sigma, sig2u = sig2 * unit;
#pragma unroll
      for (int i = 0; i < Q; ++i) d[i] = fmaf(sig2u, p.tinf[i], fmaf(-sig2, d[i], d[i]));
    } else {
      const float sigma = fminf(fmaxf(dxs / p.sponge_len_m, 0.0f), 1.0f);
      const float sig2 = sigma * sigma;
#pragma unroll
      for (int i = 0; i < Q; ++i) {
        const float ti = p.tinf[i] * unit;
        d[i] = d[i] * (1.0f - sig2) + ti * sig2;
      }
    }
  }
}

// The non-fluid branches of emitirPasso (ir.js:506-529): free-slip mirrors,
// the inlet with its 1/7-power profile (wgsl.js:213-220) and the outlet.
// Returns false for the fluid fall-through.
// gi and d in the step's units (g * unit).
__device__ __forceinline__ bool boundary_cell(uint32_t ct, const float (&gi)[Q], const Moments& M,
                                              uint32_t z, const StepParams& p, float (&d)[Q],
                                              float unit = 1.0f) {
  if (ct == ESPELHO_Y) {
    // free-slip mirror (ir.js:383-390): reflect the y component
#pragma unroll
    for (int i = 0; i < Q; ++i) d[i] = gi[mirror(i, 1)];
    return true;
  }
  if (ct == ESPELHO_Z) {
#pragma unroll
    for (int i = 0; i < Q; ++i) d[i] = gi[mirror(i, 2)];
    return true;
  }
  if (ct == ENTRADA) {
    // inlet (ir.js:405-422): equilibrium at the prescribed velocity,
    // delta = 0, rho = 1
    const float zc = (float)z + 0.5f;
    const float blf = zc < p.bl ? powf(zc / p.bl_m, 0.1428571429f) : 1.0f;
    const float ex = p.inlet[0] * blf, ey = p.inlet[1] * blf, ez = p.inlet[2] * blf;
    const float uuin = ex * ex + ey * ey + ez * ez;
    d[0] = wf(0) * (0.0f - 1.5f * 1.0f * uuin) * unit;
#pragma unroll
    for (int i = 1; i < Q; ++i) {
      const float ci = dotc(i, ex, ey, ez);
      d[i] = wf(i) * (0.0f + 1.0f * (3.0f * ci + 4.5f * ci * ci - 1.5f * uuin)) * unit;
    }
    return true;
  }
  if (ct == SAIDA) {
    // outlet (ir.js:432-444): equilibrium at the local u with delta = 0
    d[0] = wf(0) * (-1.5f * M.uu) * unit;
#pragma unroll
    for (int i = 1; i < Q; ++i) {
      const float cu = dotc(i, M.ux, M.uy, M.uz);
      d[i] = wf(i) * (3.0f * cu + 4.5f * cu * cu - 1.5f * M.uu) * unit;
    }
    return true;
  }
  return false;
}

// Instantiate a block for each storage format (S) and layout (L).
#define CFD_DISPATCH(sto, g, ...)                                          \
  do {                                                                     \
    if ((sto) == Store::F32) {                                             \
      using S = StoreF32;                                                  \
      if ((g).lb) { constexpr int L = kBlockLog2; __VA_ARGS__ }            \
      else { constexpr int L = 0; __VA_ARGS__ }                            \
    } else if ((sto) == Store::F32X) {                                     \
      using S = StoreF32X;                                                 \
      if ((g).lb) { constexpr int L = kBlockLog2; __VA_ARGS__ }            \
      else { constexpr int L = 0; __VA_ARGS__ }                            \
    } else if ((sto) == Store::F16) {                                      \
      using S = StoreF16;                                                  \
      if ((g).lb) { constexpr int L = kBlockLog2; __VA_ARGS__ }            \
      else { constexpr int L = 0; __VA_ARGS__ }                            \
    } else {                                                               \
      using S = StoreF16S;                                                 \
      if ((g).lb) { constexpr int L = kBlockLog2; __VA_ARGS__ }            \
      else { constexpr int L = 0; __VA_ARGS__ }                            \
    }                                                                      \
  } while (0)

}  // namespace cfd
