// Host-side launch interface of the device kernels.  Structs here are passed
// by value as kernel arguments, so they stay PODs.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "../../include/cfd_mi355x.h"

namespace cfd {

constexpr int kBlock = 256;  // threads per workgroup of the cell kernels
constexpr int kXcds = 8;     // MI355X: workgroups go round-robin over 8 XCDs

// Grid geometry plus the magic numbers of the two divisions that turn a
// linear cell index into (x, y, z) without integer division instructions.
constexpr int kLayoutSoA = 0;      // pop[i*S + cell]
constexpr int kLayoutBlocked = 1;  // pop[cell >> LB][i][cell & (2^LB - 1)]
#ifndef CFD_BLOCK_LOG2
#define CFD_BLOCK_LOG2 9
#endif
constexpr int kBlockLog2 = CFD_BLOCK_LOG2;  // 2^9 = 512-cell blocks

struct Geom {
  uint32_t nx, ny, nz, nxny, N;
  uint32_t S;           // SoA: elements between population slices (S >= N)
  int lb;               // layout exponent: 0 = SoA, kBlockLog2 = blocked
  size_t pop_elems;     // elements of one population set
  uint32_t mnx, snx;    // n / nx   == (umulhi(n, mnx) + n) >> snx
  uint32_t mnxy, snxy;  // n / nxny == (umulhi(n, mnxy) + n) >> snxy
  uint32_t nblocks;     // logical workgroups covering N cells
  uint32_t per_xcd;     // logical workgroups owned by each XCD
};

// Uniform values of one step, in float32 as the reference's Params uniform
// holds them (solver.js:277-293), plus per-launch constants folded on the
// host in exactly the float32 operation order the emitted shader uses.
struct StepParams {
  float omega_plus, magic, les_cs;
  float tau0;       // 1.0 / omegaPlus        (ir.js:290)
  float tau0sq;     // tau0 * tau0            (ir.js:292)
  float c18cs2;     // 18.0 * Cs * Cs         (ir.js:293, left to right)
  float belt[3];
  float inlet[3];
  float sponge_start, sponge_len_m;  // len_m = max(len, 1.0) (ir.js:461)
  float bl, bl_m;                    // bl_m = max(bl, 1.0)   (wgsl.js:217)
  float tinf[19];   // free-stream equilibrium of the sponge  (ir.js:464-475)
  uint32_t sr_seed; // per-step seed of the stochastic float16 rounding
  uint32_t sponge_all;  // some tinf is not finite: blend every fluid cell, as
                        // mix(a, b, 0) = a*1 + b*0 is then not a (ir.js:477)
};

Geom make_geom(int nx, int ny, int nz, int layout, uint32_t slice_pad);

struct StepParams;
StepParams make_step_params(const cfd_params& p);

// Launch shape of the step kernel (chosen by the autotuner or the caller).
struct Tuning {
  int block;        // threads per workgroup: 256 or 512
  int remap;        // 1: contiguous slab of workgroups per XCD
  int layout;       // kLayoutSoA or kLayoutBlocked
  int slice_pad;    // SoA only: extra elements between population slices
  int sweep;        // 1: alternate the workgroup order (ascending / descending) per step
  int rotate;       // > 0: step k starts k * rotate workgroups before step 0 (sweep 0 only)
};

// Population storage: float32 (152 B/cell/step) or scaled float16 (76 B),
// the latter rounded to nearest or stochastically by the step; F32X is
// float32 with the step's arithmetic rounded exactly as the oracle's.
enum class Store { F32 = 0, F16 = 1, F16S = 2, F32X = 3 };

hipError_t launch_prep_types(const uint32_t* tipo32, uint8_t* tipo8, const Geom& g,
                             hipStream_t st);
hipError_t launch_step(Store sto, const void* src, void* dst, const uint8_t* tipo8,
                       float4* macros, bool write_macros, const Geom& g,
                       const StepParams& p, const Tuning& tu, bool descending,
                       uint64_t step, hipStream_t st);
hipError_t launch_init(Store sto, void* dst, const uint8_t* tipo8, const Geom& g,
                       const StepParams& p, hipStream_t st);
// The macro field a step that read `src` with `p` would have written.
hipError_t launch_pull_macros(Store sto, const void* src, const uint8_t* tipo8, float4* macros,
                              const Geom& g, const StepParams& p, hipStream_t st);
hipError_t launch_macros(Store sto, const void* src, float4* macros, const Geom& g,
                         hipStream_t st);
// Body links: one entry {cell, mask of directions i with a SOLIDO neighbour
// at x + c_i} per BODY_LINK cell, in cell order.  Built in two launches:
// launch_link_count fills counts[link_chunks] and offsets[link_chunks + 1]
// (offsets[link_chunks] = number of entries), then launch_link_write writes
// the entries into `links`.
uint32_t link_chunks(const Geom& g);
hipError_t launch_link_count(const uint8_t* tipo8, const Geom& g, uint32_t* counts,
                             uint32_t* offsets, hipStream_t st);
hipError_t launch_link_write(const uint8_t* tipo8, const Geom& g, const uint32_t* offsets,
                             uint2* links, hipStream_t st);
// Momentum exchange (3 doubles) over the link list and mass (sum, non-finite
// count) reductions.  `partials` must hold kReduceGrid * 3 doubles; result
// lands in out[0..2].
constexpr int kReduceGrid = 1024;
// The learned closure counts fallback cells into this many counters (one
// atomic per workgroup, spread over the slots); the count is their sum.
constexpr int kFallbackSlots = 1024;
hipError_t launch_forces(Store sto, const void* pop, const uint2* links, uint32_t nlinks,
                         const Geom& g, double* partials, double* out, hipStream_t st);
hipError_t launch_mass(Store sto, const void* pop, const uint8_t* tipo8, const Geom& g,
                       double* partials, double* out, hipStream_t st);
// Dense host layout [19][N] <-> device layout [19][S] (S >= N).
hipError_t launch_f32_to_store(Store sto, const float* in, void* out, const Geom& g,
                               hipStream_t st);
hipError_t launch_store_to_f32(Store sto, const void* in, float* out, const Geom& g,
                               hipStream_t st);
hipError_t launch_copy(const float4* in, float4* out, size_t n4, hipStream_t st);

// ---- batched tunnels (one cooperative launch) ----
struct BatchTunnel {       // one tunnel: geometry, uniforms, buffers
  Geom g;
  StepParams p;
  void* pop[2];            // float32, blocked layout; step s reads pop[(s + parity0) & 1]
  const uint8_t* tipo8;
};
struct BatchPhase {        // rounds [r0, r1) with a fixed admitted set
  uint32_t r0, r1;
  uint32_t first;          // index of its first BatchEntry
  uint32_t n_active;
  uint32_t total_tiles;
  uint32_t table;          // offset of its tile -> entry table
};
struct BatchEntry {        // one admitted tunnel inside a phase
  uint32_t tunnel;
  uint32_t start_round;    // round of its first step
  uint32_t tile_prefix;    // first tile index of this tunnel in the phase
  uint32_t pad;
};
struct BatchLaunch {       // everything k_batch_run reads besides the tunnels
  const BatchPhase* phases;
  const BatchEntry* entries;
  const uint32_t* tile_entry;  // per phase: tile -> index of its entry in the phase
  uint32_t* counters;          // one work counter per round, zero at launch
  uint32_t* bar;               // grid barrier words (kBarWords), zero at launch
  double* samples;             // [tunnel][sample][3]
  unsigned long long* clocks;  // [0] launch start, [1 + r] end of round r
  int nphases, tile_cells, sample_every, max_samples;
  int grab;                    // consecutive tiles taken per counter increment
  uint32_t parity0;            // every tunnel's current state is in pop[parity0]
};
// Grid barrier of the batch launch: one word per slot, each on its own
// 64-byte line (kBarStride uint32 apart).  Slots: per-XCD arrival counters,
// per-XCD release generations, the global XCD count, per-XCD workgroup
// counts, and a fault flag (a barrier wait that timed out).
constexpr uint32_t kBarStride = 16;
constexpr uint32_t kBarArrive = 0, kBarRelease = kXcds, kBarGlobal = 2 * kXcds,
                   kBarCount = 2 * kXcds + 1, kBarFault = 3 * kXcds + 1;
constexpr uint32_t kBarWords = (3 * kXcds + 2) * kBarStride;
hipError_t batch_max_grid(int block, int* max_blocks_per_cu, int* cus);
hipError_t launch_batch_run(const BatchTunnel* tun, const BatchLaunch& L, int block, int grid,
                            hipStream_t st);

// ---- learned closure (closure.hip) ----
constexpr int kMlpIn = 8, kMlpHidden = 32;
constexpr int kMlpParams = kMlpHidden * kMlpIn + kMlpHidden + kMlpHidden * kMlpHidden +
                           kMlpHidden + kMlpHidden + 1;  // 1377, the cfd_mlp order
// The MLP in the operand layouts of v_mfma_f32_16x16x32_bf16, built on the
// host (bf16 = round-to-nearest-even of the float32 weights).  Hidden units
// of a layer's output reach the next layer in the permuted order
// k = 8q + v -> (v < 4 ? 4q + v : 16 + 4q + v - 4), which is how a 16x16
// result tile is spread over the lanes; w2 and w3 are stored with their input
// index permuted accordingly.
// log2(e) and ln(2) in float32: the closure network runs in base 2.
constexpr float kLog2e = 1.44269504f;
constexpr float kLn2 = 0.693147181f;
// The closure MLP as the matrix-core kernels read it (bf16 weights of the
// base-2 form, see pack_closure).
struct MlpPacked {
  uint16_t w1[2][17][8];    // M-tile m, row i = hidden 16m + i, input v; row 16 zero
  uint16_t w2[2][16][32];   // M-tile m, row i = output 16m + i, permuted input k
  uint16_t w3[32];          // permuted input k
  float b1[32], b2[32];
  float b3, out_scale;
  float shift[8], scale[8], lo[8], hi[8];
};
void pack_closure(const ::cfd_closure& c, MlpPacked* out);
bool known_activation(int32_t act);  // CFD_ACT_SILU or CFD_ACT_HARD_SILU
hipError_t launch_step_closure(Store sto, const void* src, void* dst, const uint8_t* tipo8,
                               const uint8_t* wall8, const MlpPacked* mlp, int act,
                               float4* macros, bool write_macros, const Geom& g,
                               const StepParams& p, unsigned long long* fallbacks,
                               hipStream_t st);
// Wall distance of the closure from the type field: gx (N bytes) and h2
// (N uint16) are scratch, table[0..1024] maps d^2 (capped at 1024) to wall8.
hipError_t launch_wall_distance(const uint8_t* tipo8, uint8_t* wall8, uint8_t* gx, uint16_t* h2,
                                const uint8_t* table, const Geom& g, hipStream_t st);

}  // namespace cfd
