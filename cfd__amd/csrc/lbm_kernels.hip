// D3Q19 TRT-LES kernels for gfx950 (MI355X).
//
// Physics follows the reference's IR (src/core/emit/ir.js) as emitted to WGSL
// by src/core/emit/wgsl.js; citations are given per block.  The layout and
// launch shape are this library's own:
//
//  * populations are stored per cell block: cells are grouped in blocks of
//    512 consecutive linear indices (cell = (z*ny + y)*nx + x) and a block
//    holds its 19 directions one after the other, pop[blk][i][512].  A wave
//    then reads its 19 streamed values from one ~38 KB region instead of 19
//    regions one population slice apart, which keeps the HBM page traffic of
//    the 38 concurrent streams local (measured: the pure pull data movement
//    drops from 190 us to 173 us per 6.55 M-cell step).  A plain SoA layout
//    (pop[i*S + cell]) remains selectable; A/B ping-pong as solver.js:10-18;
//  * one thread per cell over a flat index; the flat index is split into
//    (x, y, z) with two multiply-high divisions;
//  * the 19 pull loads of a cell are issued together with its type byte: one
//    memory round trip for every cell away from walls; blocked sets are read
//    and written with raw buffer instructions (one resource per set, the
//    direction as scalar offset, one 32-bit byte offset per direction);
//  * the cell type is one byte with two precomputed flags (near-wall, body
//    link), so the 18 neighbour-type reads of the bounce-back test happen only
//    for cells that touch a wall.
#include <hip/hip_cooperative_groups.h>
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "kernels.h"
#include "lattice.h"
#include "lbm_device.h"

namespace cfd {

// ------------------------------------------------------------------ types

// Pack the caller's u32 CELULA field into one byte per cell with the
// near-wall and body-link flags (see lattice.h).
__global__ __launch_bounds__(kBlock) void k_prep_types(const uint32_t* __restrict__ t32,
                                                       uint8_t* __restrict__ t8, Geom g) {
  const uint32_t cell = logical_block(g) * kBlock + threadIdx.x;
  if (cell >= g.N) return;
  const uint32_t t = t32[cell];
  uint8_t code = t < 8u ? (uint8_t)t : (uint8_t)OUTRO;
  const Coord c = split(cell, g);
  const Wrap w = wrap_offsets(c, g);
  bool near = false, body = false;
#pragma unroll
  for (int i = 1; i < Q; ++i) {
    const uint32_t up = t32[(int)cell + nb_offset(w, -cx(i), -cy(i), -cz(i))];
    near |= is_wall(up);
    const uint32_t dn = t32[(int)cell + nb_offset(w, cx(i), cy(i), cz(i))];
    body |= (dn == SOLIDO);
  }
  uint8_t out = code;
  if (near) out |= NEAR_WALL;
  if (body && code == FLUIDO) out |= BODY_LINK;
  t8[cell] = out;
}

// ------------------------------------------------------------------ step

// The fused hot kernel (wgsl.js:230-259, body = emitirPasso, ir.js:493-533):
// pull streaming with halfway bounce-back at the fluid node, moments, then
// the branch chain ESPELHO_Y -> ESPELHO_Z -> ENTRADA -> SAIDA -> fluid
// (equilibrium, closed-form Smagorinsky, TRT per opposite pair, sponge).
//
// FORCE = true also accumulates, into f, the momentum exchange of the state
// being pulled (the previous step's post-collision populations): every pull
// of a FLUIDO cell from a SOLIDO neighbour x + c_j re-reads g_j(x), which is
// exactly the link term c_j * 2 (g_j + w_j) of wgsl.js:358-379.
//
// PULL_ONLY = true stops after the moments: it writes only the macro field
// (what MACROS stores), nothing into dst.  Run on the set a step read, with
// the parameters that step used, it gives that step's macro field bit for
// bit (the same pull, bounce-back and moments code).
template <class S, bool MACROS, int LB, bool FORCE, bool PULL_ONLY = false>
__device__ __forceinline__ void step_cell(uint32_t cell,
                                          const GLOBAL typename S::T* __restrict__ src,
                                          GLOBAL typename S::T* __restrict__ dst,
                                          const GLOBAL uint8_t* __restrict__ tipo,
                                          float4* __restrict__ macros, const Geom& g,
                                          const StepParams& p, double* f) {
  const Coord c = split(cell, g);
  const Wrap w = wrap_offsets(c, g);

  // 1. pull (ir.js:166-189): g_i = src[i][x - c_i], or, when x - c_i is a
  //    wall, the population this cell sent the other way last step plus the
  //    moving-wall term 2 w_i (c_i . u_w) / c_s^2 with u_w = beltU on
  //    SOLIDO_MOVEL and 0 elsewhere.
  //    The pull addresses depend only on the coordinates, so the cell's type
  //    byte and all 19 streamed values are requested together: a cell away
  //    from walls costs one memory round trip.  Only cells flagged NEAR_WALL
  //    look at their neighbours' types and re-load the bounced directions.
  //    Addresses are formed in bytes: the cell's byte index plus the
  //    neighbour's byte offset, one shift and one multiply-add to the
  //    cell's place in its block, and a wave-uniform base per direction.
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

  // Walls are not simulated: nothing ever reads their populations
  // (wgsl.js:201-204); the macro field is zero there.  A wave of 64
  // consecutive cells that mixes walls and other cells still stores zeros in
  // its wall lanes (never read): every line of dst is then written whole, and
  // the L2 does not fetch lines from HBM to merge partial writes.
  const bool wall = is_wall(ct);
  const bool mixed = __ballot(!wall) != 0;
  if (wall) {
    if (MACROS) macros[cell] = make_float4(0.f, 0.f, 0.f, 0.f);
    if (!PULL_ONLY && mixed) out.st_zero(own);
    return;
  }
  if (tb & NEAR_WALL) bounce_back<FORCE>(gi, sb, own, tipo, in, p, ct, f);

  // 2. moments, then the branch chain ESPELHO_Y -> ESPELHO_Z -> ENTRADA ->
  //    SAIDA -> fluid (equilibrium, Smagorinsky, TRT, sponge).
  //    The arithmetic runs in the stored units (g * unit), see PopView.
  constexpr bool F = S::kFast;
  constexpr float unit = unit_of<S>();
  const Moments M = moments<F, S::kScaleLog2>(gi);
  if (PULL_ONLY) {
    macros[cell] = make_float4(M.ux, M.uy, M.uz, M.delta * (1.0f / unit));
    return;
  }
  float d[Q];
  if (!boundary_cell(ct, gi, M, c.z, p, d, unit)) {
    float pab[6];
    if constexpr (F) {
      noneq_stress_fast(M, pab);
      collide_fast<(S::kScaleLog2 > 0)>(gi, M, smagorinsky_tau<F>(stress_norm<F>(pab), M.rho, p), p, d);
    } else {
      float e[Q];
      equilibrium(M, e);
      noneq_stress(gi, e, pab);
      trt(gi, e, smagorinsky_tau<F>(stress_norm<F>(pab), M.rho, p), p, d);
    }
    sponge<F>(c.x, p, d, unit);
  }

  const uint32_t h = S::kStochastic ? sr_hash(cell, p.sr_seed) : 0u;
  f16x2 r = {};
  out.st_first(own, d[0], h, r);
#pragma unroll
  for (int i = 1; i < Q; i += 2) out.st_pair(i, own, d[i], d[i + 1], h, r);
  if (MACROS) macros[cell] = make_float4(M.ux, M.uy, M.uz, M.delta * (1.0f / unit));
}

template <class S, bool MACROS, int BLOCK, int LB>
__global__ __launch_bounds__(BLOCK) void k_step(const typename S::T* __restrict__ src,
                                                typename S::T* __restrict__ dst,
                                                const uint8_t* __restrict__ tipo,
                                                float4* __restrict__ macros, Geom g,
                                                StepParams p, uint32_t chunk_log2,
                                                uint32_t descending, uint32_t start) {
  // descending: the first workgroups dispatched take the last cells (a
  // bijection of the padded grid), so a step that follows an ascending one
  // starts on the lines that step wrote last.  start (< gridDim.x) rotates
  // the order instead: the first workgroups take the blocks from start on, so
  // a step that starts where the previous one's last blocks were reads those
  // lines first.
  uint32_t lb = logical_block_chunked(chunk_log2) + start;
  if (lb >= gridDim.x) lb -= gridDim.x;
  const uint32_t cell = (descending ? gridDim.x - 1u - lb : lb) * BLOCK + threadIdx.x;
  if (cell >= g.N) return;
  step_cell<S, MACROS, LB, false>(cell, gptr(src), gptr(dst), gptr(tipo), macros, g, p,
                                  nullptr);
}

// The macro field of the step that read `src` with parameters `p`, without
// collision or stores into a population set.
template <class S, int LB>
__global__ __launch_bounds__(kBlock) void k_pull_macros(const typename S::T* __restrict__ src,
                                                        const uint8_t* __restrict__ tipo,
                                                        float4* __restrict__ macros, Geom g,
                                                        StepParams p) {
  const uint32_t cell = logical_block(g) * kBlock + threadIdx.x;
  if (cell >= g.N) return;
  step_cell<S, true, LB, false, true>(cell, gptr(src), nullptr, gptr(tipo), macros, g, p,
                                      nullptr);
}

// Index of the calling lane in its wave.
__device__ __forceinline__ uint32_t lane_id() {
  return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
}

// Workgroup sum of 3 doubles for any BLOCK (wave64 butterfly, then the
// waves in fixed order through LDS); the result is valid in thread 0.
// wave_base = the first thread index of the calling wave (wave-uniform).
template <int BLOCK>
__device__ __forceinline__ void block_sum_into(double (&v)[3], double (&out)[3],
                                               uint32_t wave_base) {
#pragma unroll
  for (int k = 0; k < 3; ++k) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v[k] += __shfl_xor(v[k], off, 64);
  }
  __shared__ double sh[BLOCK / 64][3];
  const uint32_t wave = wave_base >> 6;
  const bool first_lane = lane_id() == 0;
  if (first_lane)
    for (int k = 0; k < 3; ++k) sh[wave][k] = v[k];
  __syncthreads();
  if (wave == 0 && first_lane)
    for (int k = 0; k < 3; ++k) {
      double s = sh[0][k];
      for (int q = 1; q < BLOCK / 64; ++q) s += sh[q][k];
      out[k] = s;
    }
  __syncthreads();
}

// ------------------------------------------------------------------ batch

// Many independent tunnels in ONE cooperative launch.  The host builds a
// schedule of rounds: in round r every admitted tunnel advances one step,
// its cells split into tiles of `tile_cells` consecutive cells; the tiles of
// all admitted tunnels form one flat list, then the whole grid synchronizes
// (all workgroups are co-resident: the runtime refuses a cooperative launch
// that does not fit).  Rounds with the same admitted set form a phase.  A
// tunnel's parity (A or B as source) is its own step count, so tunnels
// admitted at different rounds coexist.
//
// Workgroups take tiles (`grab` consecutive ones at a time) from a per-round
// counter, in list order, the way the
// hardware dispatcher hands out workgroups of an ordinary launch: the tiles
// in flight stay a contiguous window of the list, so the z-neighbour planes
// a tile pulls from are being read by its neighbours in the window at the
// same time and come from cache, not again from HBM.  (A static stride lets
// workgroups drift apart over a round and those re-reads miss.)  Each
// workgroup requests its next tile before stepping the current one, so the
// counter's latency is hidden behind the tile.
//
// Sampled steps also accumulate the body force (float64 per workgroup, one
// atomic add per workgroup and component into the tunnel's sample slot).
constexpr uint32_t kNoTile = 0xffffffffu;  // no counter value pending

// ---- grid barrier of the batch launch
//
// cooperative_groups' grid sync makes every workgroup write back its XCD's
// whole L2 and then invalidate it (an agent-scope release and acquire each):
// 96 write-backs per XCD per round at 3 workgroups per CU, ~95 us per round
// measured.  Here the workgroups of an XCD first meet on a per-XCD counter;
// the last to arrive writes back that XCD's L2 once (every workgroup of the
// XCD has waited for its own stores to reach that L2 before arriving), meets
// the other XCDs' last arrivals on a global counter, invalidates, and
// releases its XCD; every other workgroup then invalidates (agent-scope
// acquire) for its own CU.  ~24 us per round at 3 workgroups per CU, and
// bit-identical to the single-tunnel solver.  XCD membership is read from the hardware
// (HW_REG_XCC_ID) and counted once at launch; if the counts do not cover the
// grid, every barrier of the launch falls back to the cooperative grid sync.
// Every wait is bounded: a wait that times out sets the fault word (checked
// by the host) and lets the launch finish.
constexpr uint32_t kSpinLimit = 1u << 22;

__device__ __forceinline__ uint32_t* bar_word(uint32_t* bar, uint32_t slot) {
  return bar + slot * kBarStride;
}

// Waits until *w >= target; false after kSpinLimit polls or as soon as the
// fault word is set (one timed-out wait then ends every later wait at once,
// so a broken launch finishes instead of timing out round after round).
__device__ __forceinline__ bool spin_until(uint32_t* bar, uint32_t* w, uint32_t target) {
  for (uint32_t k = 0; k < kSpinLimit; ++k) {
    if (__hip_atomic_load(w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= target) return true;
    if (__hip_atomic_load(bar_word(bar, kBarFault), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
      return false;
    __builtin_amdgcn_s_sleep(2);
  }
  return false;
}

struct XcdGroup {
  uint32_t x;     // this workgroup's XCD
  uint32_t n_x;   // workgroups of the launch on XCD x
  uint32_t nxcd;  // XCDs holding workgroups of the launch
  bool ok;        // counts cover the grid: the per-XCD barrier is usable
};

// Once per launch, by every workgroup (ends with a cooperative grid sync).
__device__ XcdGroup xcd_group(uint32_t* bar, bool leader, cooperative_groups::grid_group& grid) {
  XcdGroup g{};
  g.x = __builtin_amdgcn_s_getreg(20 | (0 << 6) | (3 << 11)) & 15u;  // HW_REG_XCC_ID[3:0]
  if (leader && g.x < kXcds)
    __hip_atomic_fetch_add(bar_word(bar, kBarCount + g.x), 1u, __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
  grid.sync();
  uint32_t total = 0;
  for (uint32_t k = 0; k < kXcds; ++k) {
    const uint32_t c =
        __hip_atomic_load(bar_word(bar, kBarCount + k), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    total += c;
    g.nxcd += c > 0 ? 1u : 0u;
    if (k == g.x) g.n_x = c;
  }
  g.ok = g.x < kXcds && total == gridDim.x;
  return g;
}

// Barrier number `gen` (1, 2, ...) of the launch; every workgroup calls it
// the same number of times.
__device__ __forceinline__ void batch_barrier(uint32_t* bar, const XcdGroup& g, uint32_t gen,
                                              bool leader,
                                              cooperative_groups::grid_group& grid) {
  if (!g.ok) {
    grid.sync();
    return;
  }
  __builtin_amdgcn_s_waitcnt(0);  // this wave's stores have reached the L2
  __syncthreads();
  if (leader) {
    bool ok;
    const uint32_t a = __hip_atomic_fetch_add(bar_word(bar, kBarArrive + g.x), 1u,
                                              __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (a + 1 == gen * g.n_x) {
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");  // write back this XCD's L2
      __hip_atomic_fetch_add(bar_word(bar, kBarGlobal), 1u, __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_AGENT);
      ok = spin_until(bar, bar_word(bar, kBarGlobal), gen * g.nxcd);
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");  // invalidate this XCD's L2 (and L1)
      __hip_atomic_store(bar_word(bar, kBarRelease + g.x), gen, __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_AGENT);
    } else {
      ok = spin_until(bar, bar_word(bar, kBarRelease + g.x), gen);
      // invalidate this CU's L1 (and the XCD's L2 again): an L1-only
      // invalidate measured stale reads
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    }
    if (!ok)
      __hip_atomic_store(bar_word(bar, kBarFault), 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  __syncthreads();
}

// True in every thread when `pred` holds in any thread of the workgroup (one
// ballot per wave through LDS; ends with a barrier, so every thread has read
// the flags before any later write of them).
template <int BLOCK>
__device__ __forceinline__ bool block_any(bool pred, uint32_t wave_base) {
  __shared__ uint32_t s_any[BLOCK / 64];
  const bool w = __ballot(pred) != 0;
  if (lane_id() == 0) s_any[wave_base >> 6] = w ? 1u : 0u;
  __syncthreads();
  uint32_t a = 0;
#pragma unroll
  for (int q = 0; q < BLOCK / 64; ++q) a |= s_any[q];
  __syncthreads();
  return a != 0;
}

template <int BLOCK>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(BLOCK == 512 ? 6 : 5))) void k_batch_run(const BatchTunnel* __restrict__ tun,
                                                     const BatchLaunch L) {
  namespace cg = cooperative_groups;
  cg::grid_group grid = cg::this_grid();
  __shared__ uint32_t s_tile[2];
  // threadIdx.x as a wave-uniform base plus the lane index: the lane index
  // is recomputed where it is used, so no register holds the thread index
  // across the loops (held, it was spilled to scratch and reloaded per tile,
  // and the reload waited for the previous tile's stores)
  const uint32_t wave_base = __builtin_amdgcn_readfirstlane(threadIdx.x) & ~63u;
  const bool leader = threadIdx.x == 0;
  if (blockIdx.x == 0 && leader) L.clocks[0] = wall_clock64();
  const XcdGroup xg = xcd_group(L.bar, leader, grid);
  for (int ph = 0; ph < L.nphases; ++ph) {
    const BatchPhase P = cload(L.phases, ph);
    for (uint32_t r = P.r0; r < P.r1; ++r) {
      uint32_t* ctr = L.counters + r;
      const uint32_t grab = (uint32_t)L.grab;
      if (leader) s_tile[0] = atomicAdd(ctr, grab);
      __syncthreads();
      // the tile index is the same in every lane: made wave-uniform, the
      // tile's table, entry and tunnel descriptor come through scalar loads,
      // whose waits do not wait for the previous tile's vector stores
      uint32_t t0 = __builtin_amdgcn_readfirstlane(s_tile[0]);
      for (uint32_t it = 1; t0 < P.total_tiles; ++it) {
        uint32_t next = 0;
        if (leader) next = atomicAdd(ctr, grab);  // consumed after the tiles
        const uint32_t tend = min(t0 + grab, P.total_tiles);
        for (uint32_t t = t0; t < tend; ++t) {
        const BatchEntry E = cload(L.entries, P.first + cload(L.tile_entry, P.table + t));
        const BatchTunnel& D = tun[E.tunnel];
        const uint32_t step = r - E.start_round;
        const uint32_t from = (step + L.parity0) & 1u;
        const GLOBAL float* src = gptr(static_cast<const float*>(D.pop[from]));
        GLOBAL float* dst = gptr(static_cast<float*>(D.pop[from ^ 1u]));
        const GLOBAL uint8_t* tipo = gptr(D.tipo8);
        const uint32_t c0 = (t - E.tile_prefix) * (uint32_t)L.tile_cells;
        const uint32_t c1 = min(c0 + (uint32_t)L.tile_cells, D.g.N);
        const bool sample =
            L.sample_every > 0 && ((step + 1) % (uint32_t)L.sample_every) == 0;
        if (sample) {
          // the next tile index goes to LDS now: held in a register across
          // the force accumulation, which needs every register, it would be
          // spilled to scratch (and the reload waits for the tile's stores)
          if (leader && next != kNoTile) s_tile[it & 1] = next;
          next = kNoTile;
          double f[3] = {0.0, 0.0, 0.0};
          for (uint32_t cell = c0 + wave_base + lane_id(); cell < c1; cell += BLOCK)
            step_cell<StoreF32, false, kBlockLog2, true>(cell, src, dst, tipo, nullptr, D.g,
                                                         D.p, f);
          // most tiles hold no body link: their workgroup skips the
          // reduction and the atomics (adding exact zeros changes no sum)
          const bool links =
              block_any<BLOCK>(f[0] != 0.0 || f[1] != 0.0 || f[2] != 0.0, wave_base);
          double part[3] = {0.0, 0.0, 0.0};
          if (links) block_sum_into<BLOCK>(f, part, wave_base);  // uniform: the whole workgroup
          if (links && leader) {
            const int k = (int)((step + 1) / (uint32_t)L.sample_every) - 1;
            if (k < L.max_samples) {
              double* slot = L.samples + ((size_t)E.tunnel * L.max_samples + k) * 3;
              atomicAdd(slot + 0, part[0]);
              atomicAdd(slot + 1, part[1]);
              atomicAdd(slot + 2, part[2]);
            }
          }
        } else {
          for (uint32_t cell = c0 + wave_base + lane_id(); cell < c1; cell += BLOCK)
            step_cell<StoreF32, false, kBlockLog2, false>(cell, src, dst, tipo, nullptr,
                                                          D.g, D.p, nullptr);
        }
        }
        // ping-pong slot: a wave still reading the previous slot is not raced
        if (leader && next != kNoTile) s_tile[it & 1] = next;
        __syncthreads();
        t0 = __builtin_amdgcn_readfirstlane(s_tile[it & 1]);
      }
      batch_barrier(L.bar, xg, r + 1, leader, grid);
      if (blockIdx.x == 0 && leader) L.clocks[1 + r] = wall_clock64();
    }
  }
}

// ------------------------------------------------------------------ init

// Free-stream equilibrium everywhere (wgsl.js:269-302): u = inletU, 0 in
// SOLIDO, beltU in SOLIDO_MOVEL (PAREDE keeps inletU, as in the reference).
template <class S, int LB>
__global__ __launch_bounds__(kBlock) void k_init(typename S::T* __restrict__ dst,
                                                 const uint8_t* __restrict__ tipo, Geom g,
                                                 StepParams p) {
  const uint32_t cell = logical_block(g) * kBlock + threadIdx.x;
  if (cell >= g.N) return;
  const uint32_t ct = tipo[cell] & TYPE_MASK;
  float ux = p.inlet[0], uy = p.inlet[1], uz = p.inlet[2];
  if (ct == SOLIDO) ux = uy = uz = 0.0f;
  if (ct == SOLIDO_MOVEL) {
    ux = p.belt[0];
    uy = p.belt[1];
    uz = p.belt[2];
  }
  const float uu = ux * ux + uy * uy + uz * uz;
  S::st(pop_at<LB>(dst, 0, cell, g), wf(0) * (0.0f - 1.5f * 1.0f * uu));
#pragma unroll
  for (int i = 1; i < Q; ++i) {
    const float cu = dotc(i, ux, uy, uz);
    S::st(pop_at<LB>(dst, i, cell, g),
          wf(i) * (0.0f + 1.0f * (3.0f * cu + 4.5f * cu * cu - 1.5f * uu)));
  }
}

// ------------------------------------------------------------------ macros

// Moments of the stored populations without streaming (wgsl.js:442-476).
template <class S, int LB>
__global__ __launch_bounds__(kBlock) void k_macros(const typename S::T* __restrict__ src,
                                                   float4* __restrict__ macros, Geom g) {
  const uint32_t cell = logical_block(g) * kBlock + threadIdx.x;
  if (cell >= g.N) return;
  float gi[Q];
#pragma unroll
  for (int i = 0; i < Q; ++i) gi[i] = S::ld(pop_at<LB>(src, i, cell, g));
  float delta = gi[0];
#pragma unroll
  for (int i = 1; i < Q; ++i) delta = delta + gi[i];
  const float rho = 1.0f + delta;
  float m[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    float acc = 0.0f;
    bool first = true;
#pragma unroll
    for (int i = 1; i < Q; ++i) {
      const int ca = a == 0 ? cx(i) : (a == 1 ? cy(i) : cz(i));
      if (ca == 0) continue;
      if (first) {
        acc = ca > 0 ? gi[i] : -gi[i];
        first = false;
      } else {
        acc = ca > 0 ? acc + gi[i] : acc - gi[i];
      }
    }
    m[a] = acc;
  }
  macros[cell] = make_float4(m[0] / rho, m[1] / rho, m[2] / rho, delta);
}

// ------------------------------------------------------------------ reductions

// Wave64 butterfly sum, then the four waves of the workgroup through LDS in
// fixed order: the result does not depend on scheduling.
template <int NV>
__device__ __forceinline__ void block_sum(double (&v)[NV], double* out) {
#pragma unroll
  for (int k = 0; k < NV; ++k) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v[k] += __shfl_xor(v[k], off, 64);
  }
  __shared__ double sh[kBlock / 64][NV];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) sh[wave][k] = v[k];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) {
      double s = sh[0][k];
      for (int q = 1; q < kBlock / 64; ++q) s += sh[q][k];
      out[k] = s;
    }
  }
}

// ------------------------------------------------------------------ body links

// Momentum exchange (wgsl.js:332-396): for a FLUIDO cell x and a SOLIDO
// neighbour at x + c_i, the link delivers c_i * 2 (g_i(x) + w_i) to the body
// (the 2 w_i c_i term does not cancel per link: lattice.js:37-45);
// accumulation is float64.  The links depend on the cell types only, so they
// are listed once when the types are set: one entry per cell flagged
// BODY_LINK, in cell order, holding the cell and a bit mask of the directions
// i whose neighbour x + c_i is SOLIDO.  A force call then reads the list and
// one population per link instead of scanning the whole type field.
//
// The list is built in three passes that keep cell order: count the link
// cells of each 4096-cell chunk (16 per lane), scan the chunk counts, write
// each chunk's entries from its offset.

constexpr uint32_t kLinkCellsPerLane = 16;
constexpr uint32_t kLinkChunk = kBlock * kLinkCellsPerLane;
constexpr int kLinkBit = 5;
static_assert(BODY_LINK == 1u << kLinkBit, "flag bit of BODY_LINK");

// 16-bit mask of the BODY_LINK cells among this lane's 16 cells.  A whole
// 16-byte word is read with one aligned load (the type array starts on an
// allocation boundary); the cells of a last partial word one by one.
__device__ __forceinline__ uint32_t link_mask16(const GLOBAL uint8_t* tipo, uint32_t first,
                                                uint32_t N) {
  uint32_t mask = 0;
  if (first + kLinkCellsPerLane <= N) {
    typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
    const u32x4 t4 = *reinterpret_cast<const GLOBAL u32x4*>(tipo + first);
    const uint32_t t[4] = {t4.x, t4.y, t4.z, t4.w};
#pragma unroll
    for (int k = 0; k < 16; ++k) mask |= ((t[k >> 2] >> (8 * (k & 3) + kLinkBit)) & 1u) << k;
  } else {
    for (uint32_t k = 0; first + k < N; ++k) mask |= ((tipo[first + k] >> kLinkBit) & 1u) << k;
  }
  return mask;
}

// Exclusive prefix sum over the workgroup (wave64 shuffles, then the four
// wave totals through LDS); *total receives the workgroup sum.
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t* total) {
  __shared__ uint32_t wsum[kBlock / 64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  uint32_t inc = v;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const uint32_t o = __shfl_up(inc, off, 64);
    if (lane >= off) inc += o;
  }
  if (lane == 63) wsum[wave] = inc;
  __syncthreads();
  uint32_t before = 0, all = 0;
#pragma unroll
  for (int w = 0; w < kBlock / 64; ++w) {
    if (w < wave) before += wsum[w];
    all += wsum[w];
  }
  __syncthreads();
  *total = all;
  return before + inc - v;
}

__global__ __launch_bounds__(kBlock) void k_link_count(const uint8_t* __restrict__ tipo_, Geom g,
                                                       uint32_t* __restrict__ counts) {
  const GLOBAL uint8_t* tipo = gptr(tipo_);
  const uint32_t first = blockIdx.x * kLinkChunk + threadIdx.x * kLinkCellsPerLane;
  const uint32_t c = first < g.N ? __popc(link_mask16(tipo, first, g.N)) : 0u;
  uint32_t total;
  (void)block_exclusive_scan(c, &total);
  if (threadIdx.x == 0) counts[blockIdx.x] = total;
}

// One workgroup: offsets[k] = sum of counts[0..k), offsets[n] = total.
__global__ __launch_bounds__(kBlock) void k_link_scan(const uint32_t* __restrict__ counts,
                                                      uint32_t n, uint32_t* __restrict__ offsets) {
  uint32_t carry = 0;
  for (uint32_t base = 0; base < n; base += kBlock) {
    const uint32_t k = base + threadIdx.x;
    const uint32_t v = k < n ? counts[k] : 0u;
    uint32_t total;
    const uint32_t ex = block_exclusive_scan(v, &total);
    if (k < n) offsets[k] = carry + ex;
    carry += total;
  }
  if (threadIdx.x == 0) offsets[n] = carry;
}

__global__ __launch_bounds__(kBlock) void k_link_write(const uint8_t* __restrict__ tipo_, Geom g,
                                                       const uint32_t* __restrict__ offsets,
                                                       uint2* __restrict__ links) {
  const GLOBAL uint8_t* tipo = gptr(tipo_);
  const uint32_t first = blockIdx.x * kLinkChunk + threadIdx.x * kLinkCellsPerLane;
  uint32_t mask = first < g.N ? link_mask16(tipo, first, g.N) : 0u;
  uint32_t total;
  uint32_t pos = offsets[blockIdx.x] + block_exclusive_scan(__popc(mask), &total);
  while (mask) {
    const uint32_t cell = first + __builtin_ctz(mask);
    mask &= mask - 1u;
    const Coord c = split(cell, g);
    const Wrap w = wrap_offsets(c, g);
    uint32_t dirs = 0;
#pragma unroll
    for (int i = 1; i < Q; ++i) {
      const uint32_t nb = (uint32_t)((int)cell + nb_offset(w, cx(i), cy(i), cz(i)));
      if ((tipo[nb] & TYPE_MASK) == SOLIDO) dirs |= 1u << i;
    }
    links[pos++] = make_uint2(cell, dirs);
  }
}

template <class S, int LB>
__global__ __launch_bounds__(kBlock) void k_forces(const typename S::T* __restrict__ pop_,
                                                   const uint2* __restrict__ links, uint32_t n,
                                                   Geom g, double* __restrict__ partials) {
  const GLOBAL typename S::T* pop = gptr(pop_);
  double f[3] = {0.0, 0.0, 0.0};
  for (uint32_t e = blockIdx.x * kBlock + threadIdx.x; e < n; e += gridDim.x * kBlock) {
    const uint2 L = links[e];
#pragma unroll
    for (int i = 1; i < Q; ++i) {
      if (!((L.y >> i) & 1u)) continue;
      const double q = 2.0 * ((double)S::ld(pop_at<LB>(pop, i, L.x, g)) + wd(i));
      if (cx(i)) f[0] += cx(i) > 0 ? q : -q;
      if (cy(i)) f[1] += cy(i) > 0 ? q : -q;
      if (cz(i)) f[2] += cz(i) > 0 ? q : -q;
    }
  }
  block_sum<3>(f, partials + (size_t)blockIdx.x * 3);
}

// Mass: float64 sum of delta = sum_i g_i over non-wall cells, and a count of
// cells holding a non-finite population.
template <class S, int LB>
__global__ __launch_bounds__(kBlock) void k_mass(const typename S::T* __restrict__ pop,
                                                 const uint8_t* __restrict__ tipo, Geom g,
                                                 double* __restrict__ partials) {
  double v[2] = {0.0, 0.0};
  for (uint32_t cell = blockIdx.x * kBlock + threadIdx.x; cell < g.N;
       cell += gridDim.x * kBlock) {
    if (is_wall(tipo[cell] & TYPE_MASK)) continue;
    double s = 0.0;
    bool bad = false;
#pragma unroll
    for (int i = 0; i < Q; ++i) {
      const float x = S::ld(pop_at<LB>(pop, i, cell, g));
      bad |= !isfinite(x);
      s += (double)x;
    }
    if (bad) {
      v[1] += 1.0;
    } else {
      v[0] += s;
    }
  }
  block_sum<2>(v, partials + (size_t)blockIdx.x * 2);
}

// Second stage: one workgroup adds the kReduceGrid partials in a fixed order.
template <int NV>
__global__ __launch_bounds__(kBlock) void k_reduce(const double* __restrict__ partials,
                                                   int n, double* __restrict__ out) {
  double v[NV];
#pragma unroll
  for (int k = 0; k < NV; ++k) v[k] = 0.0;
  for (int j = threadIdx.x; j < n; j += kBlock) {
#pragma unroll
    for (int k = 0; k < NV; ++k) v[k] += partials[(size_t)j * NV + k];
  }
  block_sum<NV>(v, out);
}

// ------------------------------------------------------------------ layout

template <class S, int LB>
__global__ void k_from_f32(const float* __restrict__ in, typename S::T* __restrict__ out,
                           Geom g) {
  const size_t n = (size_t)Q * g.N;
  for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < n;
       k += (size_t)gridDim.x * blockDim.x)
    S::st(pop_at<LB>(out, (int)(k / g.N), (uint32_t)(k % g.N), g), in[k]);
}
template <class S, int LB>
__global__ void k_to_f32(const typename S::T* __restrict__ in, float* __restrict__ out,
                         Geom g) {
  const size_t n = (size_t)Q * g.N;
  for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < n;
       k += (size_t)gridDim.x * blockDim.x)
    out[k] = S::ld(pop_at<LB>(in, (int)(k / g.N), (uint32_t)(k % g.N), g));
}

// ------------------------------------------------------------------ probe

typedef float vf4 __attribute__((ext_vector_type(4)));

__global__ __launch_bounds__(kBlock) void k_copy(const vf4* __restrict__ in,
                                                 vf4* __restrict__ out, size_t n4) {
  const size_t base = (size_t)blockIdx.x * kBlock * 4 + threadIdx.x;
  vf4 v[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const size_t j = base + (size_t)k * kBlock;
    if (j < n4) v[k] = __builtin_nontemporal_load(in + j);
  }
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const size_t j = base + (size_t)k * kBlock;
    if (j < n4) __builtin_nontemporal_store(v[k], out + j);
  }
}

// ------------------------------------------------------------------ host side

// Round-up multiplier for n / d with n < 2^31 (Granlund & Montgomery 1994,
// section 4): m = floor(2^32 (2^l - d) / d) + 1 with l = ceil(log2 d); then
// n / d = (umulhi(n, m) + n) >> l.
static void magic_u32(uint32_t d, uint32_t* m, uint32_t* s) {
  uint32_t l = 0;
  while ((1ull << l) < d) ++l;
  const uint64_t num = (1ull << 32) * ((1ull << l) - d);
  *m = (uint32_t)(num / d + 1);
  *s = l;
}

Geom make_geom(int nx, int ny, int nz, int layout, uint32_t slice_pad) {
  Geom g;
  g.nx = (uint32_t)nx;
  g.ny = (uint32_t)ny;
  g.nz = (uint32_t)nz;
  g.nxny = g.nx * g.ny;
  g.N = g.nxny * g.nz;
  // SoA slices start on 256-element boundaries, plus the staggering pad
  g.S = ((g.N + 255u) & ~255u) + slice_pad;
  g.lb = layout == kLayoutBlocked ? kBlockLog2 : 0;
  const size_t bcells = (size_t)1 << kBlockLog2;
  g.pop_elems = g.lb ? ((g.N + bcells - 1) / bcells) * bcells * Q : (size_t)Q * g.S;
  magic_u32(g.nx, &g.mnx, &g.snx);
  magic_u32(g.nxny, &g.mnxy, &g.snxy);
  g.nblocks = (g.N + kBlock - 1) / kBlock;
  g.per_xcd = (g.nblocks + kXcds - 1) / kXcds;
  return g;
}

static inline dim3 cell_grid(const Geom& g) { return dim3(g.per_xcd * kXcds); }

// Instantiates the statement block passed as the variadic argument with
// S = storage trait and L = layout exponent.

hipError_t launch_prep_types(const uint32_t* t32, uint8_t* t8, const Geom& g, hipStream_t st) {
  hipLaunchKernelGGL(k_prep_types, cell_grid(g), dim3(kBlock), 0, st, t32, t8, g);
  return hipGetLastError();
}

template <class S, bool M, int L>
static void step_dispatch(const void* src, void* dst, const uint8_t* tipo8, float4* macros,
                          const Geom& g, const StepParams& p, const Tuning& tu,
                          bool descending, uint64_t step, hipStream_t st) {
  auto s = static_cast<const typename S::T*>(src);
  auto d = static_cast<typename S::T*>(dst);
  const uint32_t nb = (g.N + tu.block - 1) / tu.block;
  uint32_t lg = 0;
  while ((1 << lg) < tu.remap) ++lg;        // remap = chunk size, a power of two
  const uint32_t span = kXcds << lg;
  const dim3 grid(lg ? (nb + span - 1) / span * span : nb);
  // step k starts k * rotate workgroups before step 0 (modulo the grid)
  const uint64_t back = (uint64_t)(tu.rotate % grid.x) * (step % grid.x) % grid.x;
  const uint32_t start = back ? grid.x - (uint32_t)back : 0u;
  if (tu.block == 512)
    hipLaunchKernelGGL((k_step<S, M, 512, L>), grid, dim3(512), 0, st, s, d, tipo8, macros, g,
                       p, lg, (uint32_t)descending, start);
  else
    hipLaunchKernelGGL((k_step<S, M, 256, L>), grid, dim3(256), 0, st, s, d, tipo8, macros, g,
                       p, lg, (uint32_t)descending, start);
}

hipError_t launch_step(Store sto, const void* src, void* dst, const uint8_t* tipo8,
                       float4* macros, bool write_macros, const Geom& g,
                       const StepParams& p, const Tuning& tu, bool descending,
                       uint64_t step, hipStream_t st) {
  if (tu.block != 256 && tu.block != 512) return hipErrorInvalidValue;
  CFD_DISPATCH(sto, g, {
    if (write_macros)
      step_dispatch<S, true, L>(src, dst, tipo8, macros, g, p, tu, descending, step, st);
    else
      step_dispatch<S, false, L>(src, dst, tipo8, macros, g, p, tu, descending, step, st);
  });
  return hipGetLastError();
}

hipError_t launch_init(Store sto, void* dst, const uint8_t* tipo8, const Geom& g,
                       const StepParams& p, hipStream_t st) {
  CFD_DISPATCH(sto, g, {
    hipLaunchKernelGGL((k_init<S, L>), cell_grid(g), dim3(kBlock), 0, st,
                       static_cast<typename S::T*>(dst), tipo8, g, p);
  });
  return hipGetLastError();
}

hipError_t launch_pull_macros(Store sto, const void* src, const uint8_t* tipo8, float4* macros,
                              const Geom& g, const StepParams& p, hipStream_t st) {
  CFD_DISPATCH(sto, g, {
    hipLaunchKernelGGL((k_pull_macros<S, L>), cell_grid(g), dim3(kBlock), 0, st,
                       static_cast<const typename S::T*>(src), tipo8, macros, g, p);
  });
  return hipGetLastError();
}

hipError_t launch_macros(Store sto, const void* src, float4* macros, const Geom& g,
                         hipStream_t st) {
  CFD_DISPATCH(sto, g, {
    hipLaunchKernelGGL((k_macros<S, L>), cell_grid(g), dim3(kBlock), 0, st,
                       static_cast<const typename S::T*>(src), macros, g);
  });
  return hipGetLastError();
}

static int reduce_grid(const Geom& g) {
  const uint32_t b = (g.N + kBlock - 1) / kBlock;
  return b < (uint32_t)kReduceGrid ? (int)b : kReduceGrid;
}

uint32_t link_chunks(const Geom& g) { return (g.N + kLinkChunk - 1) / kLinkChunk; }

hipError_t launch_link_count(const uint8_t* tipo8, const Geom& g, uint32_t* counts,
                             uint32_t* offsets, hipStream_t st) {
  const uint32_t nb = link_chunks(g);
  hipLaunchKernelGGL(k_link_count, dim3(nb), dim3(kBlock), 0, st, tipo8, g, counts);
  hipLaunchKernelGGL(k_link_scan, dim3(1), dim3(kBlock), 0, st, counts, nb, offsets);
  return hipGetLastError();
}

hipError_t launch_link_write(const uint8_t* tipo8, const Geom& g, const uint32_t* offsets,
                             uint2* links, hipStream_t st) {
  hipLaunchKernelGGL(k_link_write, dim3(link_chunks(g)), dim3(kBlock), 0, st, tipo8, g,
                     offsets, links);
  return hipGetLastError();
}

hipError_t launch_forces(Store sto, const void* pop, const uint2* links, uint32_t nlinks,
                         const Geom& g, double* partials, double* out, hipStream_t st) {
  uint32_t gb = (nlinks + kBlock - 1) / kBlock;
  if (gb < 1) gb = 1;
  if (gb > (uint32_t)kReduceGrid) gb = kReduceGrid;
  CFD_DISPATCH(sto, g, {
    hipLaunchKernelGGL((k_forces<S, L>), dim3(gb), dim3(kBlock), 0, st,
                       static_cast<const typename S::T*>(pop), links, nlinks, g, partials);
  });
  hipLaunchKernelGGL(k_reduce<3>, dim3(1), dim3(kBlock), 0, st, partials, (int)gb, out);
  return hipGetLastError();
}

hipError_t launch_mass(Store sto, const void* pop, const uint8_t* tipo8, const Geom& g,
                       double* partials, double* out, hipStream_t st) {
  const int gb = reduce_grid(g);
  CFD_DISPATCH(sto, g, {
    hipLaunchKernelGGL((k_mass<S, L>), dim3(gb), dim3(kBlock), 0, st,
                       static_cast<const typename S::T*>(pop), tipo8, g, partials);
  });
  hipLaunchKernelGGL(k_reduce<2>, dim3(1), dim3(kBlock), 0, st, partials, gb, out);
  return hipGetLastError();
}

static dim3 flat_grid(size_t n) {
  size_t b = (n + kBlock - 1) / kBlock;
  if (b > 65536) b = 65536;
  if (b == 0) b = 1;
  return dim3((unsigned)b);
}

hipError_t launch_f32_to_store(Store sto, const float* in, void* out, const Geom& g,
                               hipStream_t st) {
  const size_t n = (size_t)Q * g.N;
  CFD_DISPATCH(sto, g, {
    hipLaunchKernelGGL((k_from_f32<S, L>), flat_grid(n), dim3(kBlock), 0, st, in,
                       static_cast<typename S::T*>(out), g);
  });
  return hipGetLastError();
}

hipError_t launch_store_to_f32(Store sto, const void* in, float* out, const Geom& g,
                               hipStream_t st) {
  const size_t n = (size_t)Q * g.N;
  CFD_DISPATCH(sto, g, {
    hipLaunchKernelGGL((k_to_f32<S, L>), flat_grid(n), dim3(kBlock), 0, st,
                       static_cast<const typename S::T*>(in), out, g);
  });
  return hipGetLastError();
}

hipError_t batch_max_grid(int block, int* max_blocks_per_cu, int* cus) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  e = hipDeviceGetAttribute(cus, hipDeviceAttributeMultiprocessorCount, dev);
  if (e != hipSuccess) return e;
  if (block == 512)
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(max_blocks_per_cu, k_batch_run<512>, 512,
                                                        0);
  return hipOccupancyMaxActiveBlocksPerMultiprocessor(max_blocks_per_cu, k_batch_run<256>, 256, 0);
}

hipError_t launch_batch_run(const BatchTunnel* tun, const BatchLaunch& L, int block, int grid,
                            hipStream_t st) {
  BatchLaunch l = L;
  void* args[] = {(void*)&tun, (void*)&l};
  if (block == 512)
    return hipLaunchCooperativeKernel(reinterpret_cast<const void*>(&k_batch_run<512>),
                                      dim3(grid), dim3(512), args, 0, st);
  return hipLaunchCooperativeKernel(reinterpret_cast<const void*>(&k_batch_run<256>), dim3(grid),
                                    dim3(256), args, 0, st);
}

hipError_t launch_copy(const float4* in, float4* out, size_t n4, hipStream_t st) {
  const size_t b = (n4 + kBlock * 4 - 1) / (kBlock * 4);
  hipLaunchKernelGGL(k_copy, dim3((unsigned)b), dim3(kBlock), 0, st,
                     reinterpret_cast<const vf4*>(in), reinterpret_cast<vf4*>(out), n4);
  return hipGetLastError();
}

}  // namespace cfd
