// D3Q19 velocity set with opposite pairs adjacent, as fixed by
// src/core/lattice.js:57-89 of the reference: (0), (1,2), (3,4), ... so that
// opp(i) = ((i-1)^1)+1.  Everything here is constexpr so fully unrolled loops
// over i fold to literal offsets and literal weights.
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#define CFD_HD __host__ __device__
#else
#define CFD_HD
#endif

namespace cfd {

constexpr int Q = 19;

CFD_HD constexpr int cx(int i) {
  constexpr int t[Q] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0};
  return t[i];
}
CFD_HD constexpr int cy(int i) {
  constexpr int t[Q] = {0, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0, 1, -1, 1, -1};
  return t[i];
}
CFD_HD constexpr int cz(int i) {
  constexpr int t[Q] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1};
  return t[i];
}
CFD_HD constexpr int opp(int i) { return i == 0 ? 0 : (((i - 1) ^ 1) + 1); }

// Index of the direction with component `axis` negated (ir.js:86-93):
// the specular reflection used by the free-slip ESPELHO_Y / ESPELHO_Z walls.
CFD_HD constexpr int mirror(int i, int axis) {
  int a = cx(i), b = cy(i), c = cz(i);
  if (axis == 0) a = -a;
  if (axis == 1) b = -b;
  if (axis == 2) c = -c;
  for (int j = 0; j < Q; ++j)
    if (cx(j) == a && cy(j) == b && cz(j) == c) return j;
  return -1;
}

// Weights as float32 literals.  The reference bakes them into the generated
// shader with num() = toPrecision(9) (ir.js:99-106): 1/3 -> 0.333333333,
// 1/18 -> 0.0555555556, 1/36 -> 0.0277777778.  Each of those strings rounds
// to the same float32 as the exact fraction.
CFD_HD constexpr float wf(int i) {
  return i == 0 ? 0.333333333f : (i <= 6 ? 0.0555555556f : 0.0277777778f);
}
CFD_HD constexpr double wd(int i) {
  return i == 0 ? 1.0 / 3.0 : (i <= 6 ? 1.0 / 18.0 : 1.0 / 36.0);
}
// Moving-wall bounce-back coefficient 2 w_i / c_s^2 (ir.js:185):
// 1/3 on the axes, 1/6 on the diagonals.
CFD_HD constexpr float kwall(int i) {
  return i == 0 ? 0.0f : (i <= 6 ? 0.333333333f : 0.166666667f);
}

// Cell-type codes (CELULA, ir.js:71-80).  Code 8 stands for any value the
// caller passed outside 0..7: it streams and collides as fluid (the step
// kernel's final `else`) but is never FLUIDO for the force kernel.
enum : uint8_t {
  FLUIDO = 0, SOLIDO = 1, ENTRADA = 2, SAIDA = 3, SOLIDO_MOVEL = 4,
  ESPELHO_Y = 5, ESPELHO_Z = 6, PAREDE = 7, OUTRO = 8,
};
// Packed per-cell byte used by the kernels: low nibble = type code,
// bit 4 = some upstream neighbour x - c_i is a wall (1, 4 or 7) so the pull
// needs the bounce-back test, bit 5 = FLUIDO cell with a SOLIDO neighbour at
// x + c_i (contributes to the momentum-exchange force).
constexpr uint8_t TYPE_MASK = 0x0f;
constexpr uint8_t NEAR_WALL = 0x10;
constexpr uint8_t BODY_LINK = 0x20;

CFD_HD constexpr bool is_wall(uint32_t t) {
  return t == SOLIDO || t == SOLIDO_MOVEL || t == PAREDE;
}

}  // namespace cfd
