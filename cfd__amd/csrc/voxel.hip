// GPU voxelizer, bit-identical to voxelizar() of src/geom/voxel.js.
//
// This file MUST be compiled with -ffp-contract=off: the reference computes
// in IEEE float64 (JavaScript numbers) with every product rounded before the
// following add, and the shell classification compares those results against
// 0.75 and against the separating-axis radii.  Division and sqrt are the
// correctly rounded float64 operations.
//
// Stages (each one a kernel over cells or triangles):
//   raster   one workgroup per triangle walks the triangle's padded box; per
//            cell: float64 point-triangle distance (voxel.js:105-151), the
//            running minimum kept as float32 bits with atomicMin (the minimum
//            of rounded values equals the rounded minimum, so the order of
//            triangles cannot change it), and the exact triangle/box overlap
//            (voxel.js:61-98) for cells within sqrt(0.75) -> CASCA.
//   dilate   26-neighbour dilation of the shell, the flood barrier (244-267).
//   flood    6-connected reachability from the six faces (278-327), computed
//            as a monotone fixpoint: each workgroup loads a 32x8x4 tile plus
//            halo into LDS and floods it to local convergence; tiles repeat
//            until no cell changes anywhere.  Reachability is unique, so the
//            result equals the reference's depth-first search.
//   skin     one 6-neighbour growth of the exterior from a snapshot, not
//            through original shell cells (333-355).
//   final    tipo = CASCA / DENTRO / FLUIDO, counts, and the signed distance.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <mutex>
#include <string>

#include "../../include/cfd_mi355x.h"

namespace {

constexpr uint8_t CASCA = 1;
constexpr uint32_t INF_BITS = 0x7f800000u;  // +Infinity as float32 bits

__device__ __forceinline__ double dmin3(double a, double b, double c) {
  double m = a < b ? a : b;
  return m < c ? m : c;
}
__device__ __forceinline__ double dmax3(double a, double b, double c) {
  double m = a > b ? a : b;
  return m > c ? m : c;
}

// Akenine-Moller triangle / AABB overlap with the box centred at the origin
// and half-size h on every axis (voxel.js:61-98, same operation order).
__device__ bool tri_box(const double v0[3], const double v1[3], const double v2[3], double h) {
  const double e0[3] = {v1[0] - v0[0], v1[1] - v0[1], v1[2] - v0[2]};
  const double e1[3] = {v2[0] - v1[0], v2[1] - v1[1], v2[2] - v1[2]};
  const double e2[3] = {v0[0] - v2[0], v0[1] - v2[1], v0[2] - v2[2]};
  const double* es[3] = {e0, e1, e2};
  for (int k = 0; k < 3; ++k) {
    const double* e = es[k];
    const double fx = fabs(e[0]), fy = fabs(e[1]), fz = fabs(e[2]);
    // axis (a, b) applied to components (i, j), radius fa*h + fb*h
    {
      const double a = e[2], b = -e[1];
      const double p0 = a * v0[1] + b * v0[2];
      const double p1 = a * v1[1] + b * v1[2];
      const double p2 = a * v2[1] + b * v2[2];
      const double r = fz * h + fy * h;
      if (dmin3(p0, p1, p2) > r || dmax3(p0, p1, p2) < -r) return false;
    }
    {
      const double a = -e[2], b = e[0];
      const double p0 = a * v0[0] + b * v0[2];
      const double p1 = a * v1[0] + b * v1[2];
      const double p2 = a * v2[0] + b * v2[2];
      const double r = fz * h + fx * h;
      if (dmin3(p0, p1, p2) > r || dmax3(p0, p1, p2) < -r) return false;
    }
    {
      const double a = e[1], b = -e[0];
      const double p0 = a * v0[0] + b * v0[1];
      const double p1 = a * v1[0] + b * v1[1];
      const double p2 = a * v2[0] + b * v2[1];
      const double r = fy * h + fx * h;
      if (dmin3(p0, p1, p2) > r || dmax3(p0, p1, p2) < -r) return false;
    }
  }
  for (int k = 0; k < 3; ++k) {
    if (dmin3(v0[k], v1[k], v2[k]) > h) return false;
    if (dmax3(v0[k], v1[k], v2[k]) < -h) return false;
  }
  const double n0 = e0[1] * e1[2] - e0[2] * e1[1];
  const double n1 = e0[2] * e1[0] - e0[0] * e1[2];
  const double n2 = e0[0] * e1[1] - e0[1] * e1[0];
  const double d = n0 * v0[0] + n1 * v0[1] + n2 * v0[2];
  const double r = fabs(n0) * h + fabs(n1) * h + fabs(n2) * h;
  return fabs(d) <= r;
}

// Squared point-triangle distance over the seven Voronoi regions (Ericson,
// Real-Time Collision Detection 5.1.5), voxel.js:105-151 operation by operation.
__device__ double dist2_point_tri(const double p[3], const double a[3], const double b[3],
                                  const double c[3]) {
  const double abx = b[0] - a[0], aby = b[1] - a[1], abz = b[2] - a[2];
  const double acx = c[0] - a[0], acy = c[1] - a[1], acz = c[2] - a[2];
  const double apx = p[0] - a[0], apy = p[1] - a[1], apz = p[2] - a[2];
  const double d1 = abx * apx + aby * apy + abz * apz;
  const double d2 = acx * apx + acy * apy + acz * apz;
  if (d1 <= 0 && d2 <= 0) return apx * apx + apy * apy + apz * apz;

  const double bpx = p[0] - b[0], bpy = p[1] - b[1], bpz = p[2] - b[2];
  const double d3 = abx * bpx + aby * bpy + abz * bpz;
  const double d4 = acx * bpx + acy * bpy + acz * bpz;
  if (d3 >= 0 && d4 <= d3) return bpx * bpx + bpy * bpy + bpz * bpz;

  const double vc = d1 * d4 - d3 * d2;
  if (vc <= 0 && d1 >= 0 && d3 <= 0) {
    const double v = d1 / (d1 - d3);
    const double qx = apx - v * abx, qy = apy - v * aby, qz = apz - v * abz;
    return qx * qx + qy * qy + qz * qz;
  }

  const double cpx = p[0] - c[0], cpy = p[1] - c[1], cpz = p[2] - c[2];
  const double d5 = abx * cpx + aby * cpy + abz * cpz;
  const double d6 = acx * cpx + acy * cpy + acz * cpz;
  if (d6 >= 0 && d5 <= d6) return cpx * cpx + cpy * cpy + cpz * cpz;

  const double vb = d5 * d2 - d1 * d6;
  if (vb <= 0 && d2 >= 0 && d6 <= 0) {
    const double w = d2 / (d2 - d6);
    const double qx = apx - w * acx, qy = apy - w * acy, qz = apz - w * acz;
    return qx * qx + qy * qy + qz * qz;
  }

  const double va = d3 * d6 - d5 * d4;
  if (va <= 0 && (d4 - d3) >= 0 && (d5 - d6) >= 0) {
    const double w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
    const double qx = bpx + w * (cpx - bpx), qy = bpy + w * (cpy - bpy),
                 qz = bpz + w * (cpz - bpz);
    return qx * qx + qy * qy + qz * qz;
  }

  const double den = 1 / (va + vb + vc);
  const double v = vb * den, w = vc * den;
  const double qx = apx - (v * abx + w * acx);
  const double qy = apy - (v * aby + w * acy);
  const double qz = apz - (v * abz + w * acz);
  return qx * qx + qy * qy + qz * qz;
}

constexpr int kRasterBlock = 64;

__global__ __launch_bounds__(kRasterBlock) void k_raster(
    const float* __restrict__ pos, size_t nv, const uint32_t* __restrict__ idx, size_t nt,
    int nx, int ny, int nz, double band, uint8_t* __restrict__ flags,
    uint32_t* __restrict__ dist2bits, unsigned long long* __restrict__ outside) {
  for (size_t t = blockIdx.x; t < nt; t += gridDim.x) {
    const size_t i0 = idx[t * 3], i1 = idx[t * 3 + 1], i2 = idx[t * 3 + 2];
    // An index past the vertex array reads `undefined` in the reference,
    // which poisons the finiteness sum below: the triangle is skipped.
    if (i0 >= nv || i1 >= nv || i2 >= nv) continue;
    const double a[3] = {pos[i0 * 3], pos[i0 * 3 + 1], pos[i0 * 3 + 2]};
    const double b[3] = {pos[i1 * 3], pos[i1 * 3 + 1], pos[i1 * 3 + 2]};
    const double c[3] = {pos[i2 * 3], pos[i2 * 3 + 1], pos[i2 * 3 + 2]};
    const double sum = a[0] + a[1] + a[2] + b[0] + b[1] + b[2] + c[0] + c[1] + c[2];
    if (!isfinite(sum)) continue;

    // padded triangle box clipped to the grid (voxel.js:195-201)
    double lo[3], hi[3];
    const int n[3] = {nx, ny, nz};
    bool empty = false;
    for (int k = 0; k < 3; ++k) {
      const double f = floor(dmin3(a[k], b[k], c[k]) - band);
      lo[k] = f > 0.0 ? f : 0.0;
      const double g = ceil(dmax3(a[k], b[k], c[k]) + band);
      hi[k] = g < (double)(n[k] - 1) ? g : (double)(n[k] - 1);
      if (hi[k] < lo[k]) empty = true;
    }
    if (empty) {
      if (threadIdx.x == 0) atomicAdd(outside, 1ull);
      continue;
    }
    // the box lies inside the grid, whose cell count is below 2^31: 32-bit
    // index arithmetic throughout
    const int x0 = (int)lo[0], y0 = (int)lo[1], z0 = (int)lo[2];
    const uint32_t bx = (uint32_t)((int)hi[0] - x0 + 1), by = (uint32_t)((int)hi[1] - y0 + 1),
                   bz = (uint32_t)((int)hi[2] - z0 + 1);
    const uint32_t ncell = bx * by * bz;
    // Without a signed distance the exact distance only decides d2 <= 0.75.
    // A cell whose centre lies farther than sqrt(0.7575) from the triangle's
    // plane is farther than that from the triangle, so its test is false: the
    // 1 % margin covers the rounding of the exact distance, whose error grows
    // like 1 / sin^2 of the triangle's sharpest angle, so only triangles with
    // sin^2 > 1e-8 use the shortcut (error below 1e-7 cell^2 there); slivers
    // and degenerate triangles always take the exact distance.
    const double e1x = b[0] - a[0], e1y = b[1] - a[1], e1z = b[2] - a[2];
    const double e2x = c[0] - a[0], e2y = c[1] - a[1], e2z = c[2] - a[2];
    const double nx_ = e1y * e2z - e1z * e2y;
    const double ny_ = e1z * e2x - e1x * e2z;
    const double nz_ = e1x * e2y - e1y * e2x;
    const double nn = nx_ * nx_ + ny_ * ny_ + nz_ * nz_;
    const double ll = (e1x * e1x + e1y * e1y + e1z * e1z) * (e2x * e2x + e2y * e2y + e2z * e2z);
    const bool shortcut = !dist2bits && nn > 1e-8 * ll;
    const double far2 = 0.7575 * nn;
    for (uint32_t k = threadIdx.x; k < ncell; k += kRasterBlock) {
      const uint32_t r = k / bx;
      const int x = x0 + (int)(k - r * bx);
      const uint32_t zr = r / by;
      const int y = y0 + (int)(r - zr * by);
      const int z = z0 + (int)zr;
      const double p[3] = {x + 0.5, y + 0.5, z + 0.5};
      const uint32_t cell = ((uint32_t)z * (uint32_t)ny + (uint32_t)y) * (uint32_t)nx + (uint32_t)x;
      if (shortcut) {
        const double sp = nx_ * (p[0] - a[0]) + ny_ * (p[1] - a[1]) + nz_ * (p[2] - a[2]);
        if (sp * sp > far2) continue;
      }
      const double d2 = dist2_point_tri(p, a, b, c);
      // the distance field is kept only when a signed distance is asked for
      if (dist2bits && d2 == d2) atomicMin(&dist2bits[cell], __float_as_uint((float)d2));
      if (d2 <= 0.75 && flags[cell] != CASCA) {
        const double la[3] = {a[0] - p[0], a[1] - p[1], a[2] - p[2]};
        const double lb[3] = {b[0] - p[0], b[1] - p[1], b[2] - p[2]};
        const double lc[3] = {c[0] - p[0], c[1] - p[1], c[2] - p[2]};
        if (tri_box(la, lb, lc, 0.5)) flags[cell] = CASCA;
      }
    }
  }
}

constexpr int kCellBlock = 256;
constexpr uint32_t kNoCell = 0xffffffffu;

// This thread's cell of a one-thread-per-cell launch and its coordinates, in
// 32-bit arithmetic (the grid holds fewer than 2^31 cells); kNoCell past the
// end.
__device__ __forceinline__ uint32_t cell_xyz(int nx, int ny, int nz, int* x, int* y, int* z) {
  const uint32_t N = (uint32_t)nx * (uint32_t)ny * (uint32_t)nz;
  const uint32_t cell = blockIdx.x * (uint32_t)kCellBlock + threadIdx.x;
  if (cell >= N) return kNoCell;
  const uint32_t r = cell / (uint32_t)nx;
  *x = (int)(cell - r * (uint32_t)nx);
  const uint32_t zz = r / (uint32_t)ny;
  *y = (int)(r - zz * (uint32_t)ny);
  *z = (int)zz;
  return cell;
}

// 26-neighbour dilation of the shell, clamped at the faces (voxel.js:244-267).
// The 3x3x3 box maximum is separable: one pass per axis, each reading the
// cell and its two neighbours along that axis (clamped), the first pass
// testing for CASCA.
__global__ __launch_bounds__(kCellBlock) void k_dilate_axis(const uint8_t* __restrict__ in,
                                                            uint8_t* __restrict__ out, int nx,
                                                            int ny, int nz, int axis,
                                                            int shell) {
  int x, y, z;
  const uint32_t cell = cell_xyz(nx, ny, nz, &x, &y, &z);
  if (cell == kNoCell) return;
  const int pos = axis == 0 ? x : (axis == 1 ? y : z);
  const int n = axis == 0 ? nx : (axis == 1 ? ny : nz);
  const uint32_t step = axis == 0 ? 1u : (axis == 1 ? (uint32_t)nx : (uint32_t)nx * ny);
  auto hit = [&](uint32_t c) -> uint8_t { return shell ? (in[c] == CASCA) : (in[c] != 0); };
  uint8_t v = hit(cell);
  if (pos > 0) v |= hit(cell - step);
  if (pos < n - 1) v |= hit(cell + step);
  out[cell] = v;
}

// Seed: every face cell that is not a barrier is exterior (voxel.js:291-309).
__global__ __launch_bounds__(kCellBlock) void k_seed(const uint8_t* __restrict__ barrier,
                                                     uint8_t* __restrict__ ext, int nx, int ny,
                                                     int nz) {
  int x, y, z;
  const uint32_t cell = cell_xyz(nx, ny, nz, &x, &y, &z);
  if (cell == kNoCell) return;
  const bool face = x == 0 || x == nx - 1 || y == 0 || y == ny - 1 || z == 0 || z == nz - 1;
  ext[cell] = (face && !barrier[cell]) ? 1 : 0;
}

// Tile flood, bit-parallel.  A tile is 32 x 16 x 16 cells; each of its 256
// x-rows is one 64-bit word in LDS, bits 1..32 the row's cells and bits 0
// and 33 its x-neighbours outside the tile, with a one-row halo in y and z.
// Per row two words: E (exterior) and O (open: in the grid and not a
// barrier).  One thread owns one row; an iteration takes the exterior bits of
// the four y/z neighbour rows where the row is open, then fills the row along
// x in both directions through its open runs (Kogge-Stone occluded fill:
// six shift steps per direction).  The tile iterates until no row changes;
// the exterior only grows, so reading a neighbour row while it is rewritten
// sees an old or a new value, both subsets of the fixpoint, and the loop ends
// only after an iteration in which nothing was written.  The fixpoint is
// unique, so the result does not depend on the tile shape or the order.
constexpr int FX = 32, FY = 16, FZ = 16;        // tile (x bits, y rows, z rows)
constexpr int RY = FY + 2, RZ = FZ + 2;         // with the y/z halo
constexpr uint64_t kInterior = 0x1FFFFFFFEull;  // bits 1..32
static_assert(FY * FZ == kCellBlock, "one thread per interior row");

__device__ __forceinline__ uint64_t fill_x(uint64_t gen, uint64_t pro) {
  uint64_t g = gen, p = pro;
  g |= p & (g << 1);  p &= p << 1;
  g |= p & (g << 2);  p &= p << 2;
  g |= p & (g << 4);  p &= p << 4;
  g |= p & (g << 8);  p &= p << 8;
  g |= p & (g << 16); p &= p << 16;
  g |= p & (g << 32);
  p = pro;
  g |= p & (g >> 1);  p &= p >> 1;
  g |= p & (g >> 2);  p &= p >> 2;
  g |= p & (g >> 4);  p &= p >> 4;
  g |= p & (g >> 8);  p &= p >> 8;
  g |= p & (g >> 16); p &= p >> 16;
  g |= p & (g >> 32);
  return g;
}

// A sweep works only on tiles that may change: those with a face neighbour
// (or themselves) changed in the previous sweep, whose halo may then hold new
// exterior cells (`prev`: one byte per tile, all ones for the first sweep).
// A tile that no neighbour changed saw every exterior cell of its halo
// already, so it is at its fixpoint.  `next` records this sweep's changes.
__global__ __launch_bounds__(kCellBlock) void k_flood_tile(const uint8_t* __restrict__ barrier,
                                                           uint8_t* __restrict__ ext, int nx,
                                                           int ny, int nz,
                                                           const uint8_t* __restrict__ prev,
                                                           uint8_t* __restrict__ next,
                                                           int* __restrict__ changed) {
  __shared__ uint64_t E[RZ][RY], O[RZ][RY];
  __shared__ uint32_t grown[FZ][FY];
  const int gx = (nx + FX - 1) / FX, gy = (ny + FY - 1) / FY, gz = (nz + FZ - 1) / FZ;
  const int bxi = blockIdx.x % gx;
  const int byi = (blockIdx.x / gx) % gy;
  const int bzi = blockIdx.x / (gx * gy);
  const int t = blockIdx.x;
  const bool active = prev[t] || (bxi > 0 && prev[t - 1]) || (bxi < gx - 1 && prev[t + 1]) ||
                      (byi > 0 && prev[t - gx]) || (byi < gy - 1 && prev[t + gx]) ||
                      (bzi > 0 && prev[t - gx * gy]) || (bzi < gz - 1 && prev[t + gx * gy]);
  if (!active) {  // uniform over the workgroup
    if (threadIdx.x == 0) next[t] = 0;
    return;
  }
  const int x0 = bxi * FX, y0 = byi * FY - 1, z0 = bzi * FZ - 1;   // y0, z0: halo row 0
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;

  // load: a wave takes two rows at a time, lanes 0-31 the first row's cells,
  // 32-63 the second's; ballots assemble the bits.  The halo cells x0 - 1
  // and x0 + 32 of the same two rows come from lanes 0-3 of a second load.
  // Eight row pairs are loaded before any is assembled, and the barrier and
  // exterior bytes are loaded together, so the loads overlap (the loop is
  // latency-bound otherwise).
  constexpr int kPairs = RY * RZ / 2, kWaves = kCellBlock / 64, kBatch = 8;
  // every lane loads from a clamped, in-grid address and the validity is
  // applied afterwards: no branch around a load, so the loads of a batch go
  // out before the first wait
  auto clampi = [](int v, int n) { return v < 0 ? 0 : (v >= n ? n - 1 : v); };
  for (int p0 = wave; p0 < kPairs; p0 += kWaves * kBatch) {
    uint8_t bar[kBatch], exv[kBatch], hbar[kBatch], hexv[kBatch];
    bool ok[kBatch], hok[kBatch];
#pragma unroll
    for (int k = 0; k < kBatch; ++k) {
      const int pair = p0 + kWaves * k;
      const int row = 2 * pair + (lane >> 5);
      const int y = y0 + row % RY, z = z0 + row / RY;
      const int x = x0 + (lane & 31);
      ok[k] = pair < kPairs && y >= 0 && y < ny && z >= 0 && z < nz && x < nx;
      const size_t c = ((size_t)clampi(z, nz) * ny + clampi(y, ny)) * nx + clampi(x, nx);
      bar[k] = barrier[c];
      exv[k] = ext[c];
      const int hrow = 2 * pair + ((lane >> 1) & 1);
      const int hy = y0 + hrow % RY, hz = z0 + hrow / RY;
      const int hx = (lane & 1) ? x0 + FX : x0 - 1;
      hok[k] = lane < 4 && pair < kPairs && hy >= 0 && hy < ny && hz >= 0 && hz < nz &&
               hx >= 0 && hx < nx;
      const size_t hc = ((size_t)clampi(hz, nz) * ny + clampi(hy, ny)) * nx + clampi(hx, nx);
      hbar[k] = barrier[hc];
      hexv[k] = ext[hc];
    }
#pragma unroll
    for (int k = 0; k < kBatch; ++k) {
      const int pair = p0 + kWaves * k;
      if (pair >= kPairs) break;   // uniform over the wave
      const bool open = ok[k] && !bar[k], hopen = hok[k] && !hbar[k];
      const uint64_t bo = __ballot(open), be = __ballot(open && exv[k]);
      const uint64_t ho = __ballot(hopen), he = __ballot(hopen && hexv[k]);
      if (lane == 0) {
        for (int j = 0; j < 2; ++j) {
          const int r = 2 * pair + j;
          const uint64_t o = ((bo >> (32 * j)) & 0xffffffffull) << 1 |
                             ((ho >> (2 * j)) & 1ull) | (((ho >> (2 * j + 1)) & 1ull) << 33);
          const uint64_t e = ((be >> (32 * j)) & 0xffffffffull) << 1 |
                             ((he >> (2 * j)) & 1ull) | (((he >> (2 * j + 1)) & 1ull) << 33);
          O[r / RY][r % RY] = o;
          E[r / RY][r % RY] = e;
        }
      }
    }
  }
  __syncthreads();

  const int ly = 1 + threadIdx.x % FY, lz = 1 + threadIdx.x / FY;
  const uint64_t pro = O[lz][ly] & kInterior;
  const uint64_t e0 = E[lz][ly];
  uint64_t e = e0;
  for (;;) {
    const uint64_t nb = E[lz][ly - 1] | E[lz][ly + 1] | E[lz - 1][ly] | E[lz + 1][ly];
    const uint64_t g = fill_x(e | (nb & pro), pro);
    const int mine = g != e;
    if (mine) {
      E[lz][ly] = g;
      e = g;
    }
    // barrier + OR: every thread sees the same decision, and the next
    // iteration starts only after every write of this one
    if (!__syncthreads_or(mine)) break;
  }
  const uint32_t gw = (uint32_t)((e & ~e0) >> 1);
  grown[lz - 1][ly - 1] = gw;
  const int any = __syncthreads_or(gw != 0);
  if (any) {
    // store the new exterior cells, two rows per wave instruction
    for (int pair = wave; pair < FY * FZ / 2; pair += kCellBlock / 64) {
      const int row = 2 * pair + (lane >> 5);
      const int ry = row % FY, rz = row / FY;
      const int x = x0 + (lane & 31);
      if ((grown[rz][ry] >> (lane & 31)) & 1u)
        ext[((size_t)(z0 + 1 + rz) * ny + (y0 + 1 + ry)) * nx + x] = 1;
    }
    if (threadIdx.x == 0) *changed = 1;
  }
  if (threadIdx.x == 0) next[t] = any ? 1 : 0;
}

// Skin recovery from a snapshot (voxel.js:333-355).
__global__ __launch_bounds__(kCellBlock) void k_skin(const uint8_t* __restrict__ before,
                                                     const uint8_t* __restrict__ flags,
                                                     uint8_t* __restrict__ after, int nx,
                                                     int ny, int nz) {
  int x, y, z;
  const uint32_t c = cell_xyz(nx, ny, nz, &x, &y, &z);
  if (c == kNoCell) return;
  const uint32_t nxny = (uint32_t)nx * (uint32_t)ny;
  uint8_t v = before[c];
  if (!v && flags[c] != CASCA) {
    v = (x > 0 && before[c - 1]) || (x < nx - 1 && before[c + 1]) ||
        (y > 0 && before[c - nx]) || (y < ny - 1 && before[c + nx]) ||
        (z > 0 && before[c - nxny]) || (z < nz - 1 && before[c + nxny]);
  }
  after[c] = v;
}

// tipo, counts and signed distance (voxel.js:401-415).
__global__ __launch_bounds__(kCellBlock) void k_final(
    const uint8_t* __restrict__ flags, const uint8_t* __restrict__ ext,
    const uint32_t* __restrict__ dist2bits, size_t N, double band, uint8_t* __restrict__ tipo,
    float* __restrict__ sdf, unsigned long long* __restrict__ counts) {
  unsigned long long casca = 0, dentro = 0;
  for (size_t c = blockIdx.x * (size_t)kCellBlock + threadIdx.x; c < N;
       c += (size_t)gridDim.x * kCellBlock) {
    uint8_t t = 0;
    if (flags[c] == CASCA) {
      t = 1;
      ++casca;
    } else if (!ext[c]) {
      t = 2;
      ++dentro;
    }
    tipo[c] = t;
    if (sdf) {
      const uint32_t bits = dist2bits[c];
      const double d = bits == INF_BITS ? band + 1 : sqrt((double)__uint_as_float(bits));
      sdf[c] = (float)(t == 0 ? d : -d);
    }
  }
  // workgroup-aggregated counters: one atomic per workgroup and counter (the
  // grid is capped, so few atomics meet at the two addresses)
  for (int off = 32; off > 0; off >>= 1) {
    casca += __shfl_xor(casca, off, 64);
    dentro += __shfl_xor(dentro, off, 64);
  }
  __shared__ unsigned long long part[2][kCellBlock / 64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) {
    part[0][wave] = casca;
    part[1][wave] = dentro;
  }
  __syncthreads();
  if (threadIdx.x < 2) {
    unsigned long long v = 0;
    for (int w = 0; w < kCellBlock / 64; ++w) v += part[threadIdx.x][w];
    if (v) atomicAdd(&counts[threadIdx.x], v);
  }
}

__global__ void k_fill_u32(uint32_t* p, size_t n, uint32_t v) {
  for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < n;
       k += (size_t)gridDim.x * blockDim.x)
    p[k] = v;
}

}  // namespace

// The solver translation unit owns cfd_last_error(); the voxelizer reports
// through it by writing the shared message with this helper.
extern "C" void cfd__set_error(const char* msg);

namespace {
int vfail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  cfd__set_error(buf);
  return code;
}
__global__ void k_fill_u8(uint8_t* p, size_t n, uint8_t v) {
  for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < n;
       k += (size_t)gridDim.x * blockDim.x)
    p[k] = v;
}

// Device buffers of one voxelization.  pos/idx: the mesh; tipo (and sdf,
// nullable): the result; the rest is scratch: flags, barrier, ext, ext2 (N
// bytes each), dist2 (N words, only with sdf), counters (3), changed (1),
// tile_chg (2 bytes per flood tile).
struct VoxBuffers {
  const float* pos;
  const uint32_t* idx;
  uint8_t *flags, *barrier, *ext, *ext2, *tipo;
  uint32_t* dist2;
  float* sdf;
  unsigned long long* counters;  // [0] casca [1] dentro [2] triangles outside
  int* changed;
  uint8_t* tile_chg;
};

unsigned flood_tiles(int nx, int ny, int nz) {
  return (unsigned)(((nx + FX - 1) / FX) * ((ny + FY - 1) / FY) * ((nz + FZ - 1) / FZ));
}

hipError_t fill_u8(uint8_t* p, size_t n, uint8_t v, hipStream_t st) {
  const size_t blocks = (n + 255) / 256;
  hipLaunchKernelGGL(k_fill_u8, dim3((unsigned)(blocks < 4096 ? blocks : 4096)), dim3(256), 0,
                     st, p, n, v);
  return hipGetLastError();
}

// The whole pipeline on `st`, shared by both entry points: clear the scratch,
// rasterize, seal, flood to the fixpoint (reading the change flag back after
// each sweep), recover the skin and write the types.  ev (nullable): four
// events recorded around raster / flood / final; iters: flood sweeps.
hipError_t run_voxelize(const VoxBuffers& b, size_t nv, size_t nt, int nx, int ny, int nz,
                        double band, int seal, hipStream_t st, hipEvent_t* ev, int* iters) {
  const size_t N = (size_t)nx * ny * nz;
  const unsigned cgrid = (unsigned)((N + kCellBlock - 1) / kCellBlock);
  const unsigned tgrid = flood_tiles(nx, ny, nz);
  hipError_t e;
#define VRET(expr)                 \
  do {                             \
    if ((e = (expr)) != hipSuccess) \
      return e;                    \
  } while (0)
  VRET(fill_u8(b.flags, N, 0, st));
  VRET(fill_u8(reinterpret_cast<uint8_t*>(b.counters), 3 * sizeof(unsigned long long), 0, st));
  if (b.dist2) {
    hipLaunchKernelGGL(k_fill_u32, dim3(4096), dim3(256), 0, st, b.dist2, N, INF_BITS);
    VRET(hipGetLastError());
  }
  if (ev) VRET(hipEventRecord(ev[0], st));
  if (nt) {
    const size_t g = nt < 262144 ? nt : 262144;
    hipLaunchKernelGGL(k_raster, dim3((unsigned)g), dim3(kRasterBlock), 0, st, b.pos, nv, b.idx,
                       nt, nx, ny, nz, band, b.flags, b.dist2, b.counters + 2);
    VRET(hipGetLastError());
  }
  if (ev) VRET(hipEventRecord(ev[1], st));
  const uint8_t* bar = b.flags;
  if (seal) {
    // x pass into ext, y pass into ext2 (both rewritten before use), z pass
    // into the barrier
    hipLaunchKernelGGL(k_dilate_axis, dim3(cgrid), dim3(kCellBlock), 0, st, b.flags, b.ext, nx,
                       ny, nz, 0, 1);
    hipLaunchKernelGGL(k_dilate_axis, dim3(cgrid), dim3(kCellBlock), 0, st, b.ext, b.ext2, nx,
                       ny, nz, 1, 0);
    hipLaunchKernelGGL(k_dilate_axis, dim3(cgrid), dim3(kCellBlock), 0, st, b.ext2, b.barrier,
                       nx, ny, nz, 2, 0);
    VRET(hipGetLastError());
    bar = b.barrier;
  }
  hipLaunchKernelGGL(k_seed, dim3(cgrid), dim3(kCellBlock), 0, st, bar, b.ext, nx, ny, nz);
  VRET(hipGetLastError());
  VRET(fill_u8(b.tile_chg, tgrid, 1, st));   // first sweep: every tile
  int n_it = 0;
  for (;;) {
    VRET(fill_u8(reinterpret_cast<uint8_t*>(b.changed), sizeof(int), 0, st));
    const uint8_t* prev = b.tile_chg + (size_t)(n_it & 1) * tgrid;
    uint8_t* next = b.tile_chg + (size_t)((n_it + 1) & 1) * tgrid;
    hipLaunchKernelGGL(k_flood_tile, dim3(tgrid), dim3(kCellBlock), 0, st, bar, b.ext, nx, ny,
                       nz, prev, next, b.changed);
    VRET(hipGetLastError());
    int hchg = 0;
    VRET(hipMemcpyAsync(&hchg, b.changed, sizeof(int), hipMemcpyDeviceToHost, st));
    VRET(hipStreamSynchronize(st));
    ++n_it;
    if (!hchg) break;
  }
  uint8_t* ext_final = b.ext;
  if (seal) {
    hipLaunchKernelGGL(k_skin, dim3(cgrid), dim3(kCellBlock), 0, st, b.ext, b.flags, b.ext2, nx,
                       ny, nz);
    VRET(hipGetLastError());
    ext_final = b.ext2;
  }
  if (ev) VRET(hipEventRecord(ev[2], st));
  hipLaunchKernelGGL(k_final, dim3(cgrid < 2048u ? cgrid : 2048u), dim3(kCellBlock), 0, st,
                     b.flags, ext_final, b.dist2, N, band, b.tipo, b.sdf, b.counters);
  VRET(hipGetLastError());
  if (ev) VRET(hipEventRecord(ev[3], st));
#undef VRET
  *iters = n_it;
  return hipSuccess;
}

void fill_stats(const unsigned long long hc[3], size_t N, int iters, float ms_raster,
                float ms_flood, float ms_total, cfd_vox_stats* stats) {
  stats->n_cells = (int64_t)N;
  stats->n_casca = (int64_t)hc[0];
  stats->n_dentro = (int64_t)hc[1];
  stats->n_solido = (int64_t)(hc[0] + hc[1]);
  stats->triangles_outside = (int64_t)hc[2];
  stats->hollow_suspect = (hc[1] < hc[0] * 0.25 && hc[0] > 500) ? 1 : 0;
  stats->flood_iterations = iters;
  stats->ms_raster = ms_raster;
  stats->ms_flood = ms_flood;
  stats->ms_total = ms_total;
}

int check_args(const char* fn, const void* pos, size_t nv, const void* idx, size_t nt, int nx,
               int ny, int nz, double band, const void* tipo_out) {
  if (nx < 1 || ny < 1 || nz < 1)
    return vfail(CFD_E_ARG, "%s: grid %dx%dx%d must be positive", fn, nx, ny, nz);
  if (!tipo_out) return vfail(CFD_E_ARG, "%s: tipo_out is null", fn);
  if ((nv && !pos) || (nt && !idx)) return vfail(CFD_E_ARG, "%s: null mesh arrays", fn);
  if (!std::isfinite(band)) return vfail(CFD_E_ARG, "%s: band must be finite", fn);
  if ((uint64_t)nx * ny * nz >= (1ull << 31)) return vfail(CFD_E_ARG, "%s: grid too large", fn);
  return CFD_OK;
}

}  // namespace

#define VCK(expr)                                                                    \
  do {                                                                               \
    hipError_t e_ = (expr);                                                          \
    if (e_ != hipSuccess) {                                                          \
      rc = vfail(e_ == hipErrorOutOfMemory ? CFD_E_OOM : CFD_E_HIP, "cfd_voxelize: %s (%s)", \
                 hipGetErrorString(e_), #expr);                                      \
      goto done;                                                                     \
    }                                                                                \
  } while (0)

extern "C" int cfd_voxelize(const float* pos, size_t nv, const uint32_t* idx, size_t nt,
                            int nx, int ny, int nz, double band, int seal, uint8_t* tipo_out,
                            float* sdf_out, cfd_vox_stats* stats) {
  cfd__set_error("");
  int rc = check_args("cfd_voxelize", pos, nv, idx, nt, nx, ny, nz, band, tipo_out);
  if (rc) return rc;
  const size_t N = (size_t)nx * ny * nz;
  hipStream_t st = nullptr;
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
  float* d_pos = nullptr;
  uint32_t* d_idx = nullptr;
  VoxBuffers b{};
  unsigned long long hc[3] = {0, 0, 0};
  int iters = 0;
  const unsigned tgrid = flood_tiles(nx, ny, nz);

  VCK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
  for (auto& e : ev) VCK(hipEventCreate(&e));
  if (nv) VCK(hipMalloc(&d_pos, nv * 3 * sizeof(float)));
  if (nt) VCK(hipMalloc(&d_idx, nt * 3 * sizeof(uint32_t)));
  VCK(hipMalloc(&b.flags, N));
  VCK(hipMalloc(&b.barrier, N));
  VCK(hipMalloc(&b.ext, N));
  VCK(hipMalloc(&b.ext2, N));
  VCK(hipMalloc(&b.tipo, N));
  // the float32 distance field (atomic minimum per cell) only feeds the
  // signed distance; without sdf_out the rasterizer skips it
  if (sdf_out) VCK(hipMalloc(&b.dist2, N * sizeof(uint32_t)));
  if (sdf_out) VCK(hipMalloc(&b.sdf, N * sizeof(float)));
  VCK(hipMalloc(&b.counters, 3 * sizeof(unsigned long long)));
  VCK(hipMalloc(&b.changed, sizeof(int)));
  VCK(hipMalloc(&b.tile_chg, 2 * (size_t)tgrid));
  if (nv) VCK(hipMemcpyAsync(d_pos, pos, nv * 3 * sizeof(float), hipMemcpyHostToDevice, st));
  if (nt) VCK(hipMemcpyAsync(d_idx, idx, nt * 3 * sizeof(uint32_t), hipMemcpyHostToDevice, st));
  b.pos = d_pos;
  b.idx = d_idx;
  VCK(run_voxelize(b, nv, nt, nx, ny, nz, band, seal, st, ev, &iters));
  VCK(hipMemcpyAsync(tipo_out, b.tipo, N, hipMemcpyDeviceToHost, st));
  if (sdf_out) VCK(hipMemcpyAsync(sdf_out, b.sdf, N * sizeof(float), hipMemcpyDeviceToHost, st));
  VCK(hipMemcpyAsync(hc, b.counters, sizeof hc, hipMemcpyDeviceToHost, st));
  VCK(hipStreamSynchronize(st));
  if (stats) {
    float t01 = 0, t12 = 0, t03 = 0;
    VCK(hipEventElapsedTime(&t01, ev[0], ev[1]));
    VCK(hipEventElapsedTime(&t12, ev[1], ev[2]));
    VCK(hipEventElapsedTime(&t03, ev[0], ev[3]));
    fill_stats(hc, N, iters, t01, t12, t03, stats);
  }

done:
  if (st) (void)hipStreamSynchronize(st);
  for (void* p : {(void*)d_pos, (void*)d_idx, (void*)b.flags, (void*)b.barrier, (void*)b.ext,
                  (void*)b.ext2, (void*)b.tipo, (void*)b.dist2, (void*)b.sdf,
                  (void*)b.counters, (void*)b.changed, (void*)b.tile_chg})
    if (p) (void)hipFree(p);
  for (auto& e : ev)
    if (e) (void)hipEventDestroy(e);
  if (st) (void)hipStreamDestroy(st);
  return rc;
}
#undef VCK

// ---- the same voxelization without allocation ----
// Scratch in one fixed device array per GPU: flags, barrier, ext, ext2 (one
// byte per cell each), dist2 (a word per cell), counters, the change flag and
// the flood tiles' change bytes, at fixed offsets sized for
// CFD_VOXELIZE_NOALLOC_MAX_CELLS cells and kNoAllocMaxTiles flood tiles.
namespace {
constexpr size_t kNoAllocMaxCells = CFD_VOXELIZE_NOALLOC_MAX_CELLS;
constexpr size_t kNoAllocMaxTiles = 1u << 18;
constexpr size_t kOffCounters = 8 * kNoAllocMaxCells;  // after 4 byte planes + dist2
constexpr size_t kOffChanged = kOffCounters + 256;
constexpr size_t kOffTiles = kOffChanged + 256;
constexpr size_t kNoAllocBytes = kOffTiles + 2 * kNoAllocMaxTiles;
static_assert(kNoAllocMaxCells % 256 == 0, "byte planes keep 256-byte alignment");

static __device__ __align__(256) uint8_t scratch[kNoAllocBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
uint8_t* g_scratch[kMaxDevices] = {};

// The device address of `scratch` on the current device, looked up once per
// device (each GPU holds its own copy of the array at its own address).
int scratch_of_current_device(uint8_t** out) {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess)
    return vfail(CFD_E_HIP, "cfd_voxelize_noalloc: hipGetDevice: %s", hipGetErrorString(e));
  if (dev < 0 || dev >= kMaxDevices)
    return vfail(CFD_E_HIP, "cfd_voxelize_noalloc: device %d outside 0..%d", dev,
                 kMaxDevices - 1);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return vfail(CFD_E_HIP, "cfd_voxelize_noalloc: hipGetSymbolAddress(scratch): %s",
                   hipGetErrorString(e));
    g_scratch[dev] = static_cast<uint8_t*>(p);
  }
  *out = g_scratch[dev];
  return CFD_OK;
}
}  // namespace

extern "C" int cfd_voxelize_noalloc(const float* d_pos, size_t nv, const uint32_t* d_idx,
                                    size_t nt, int nx, int ny, int nz, double band, int seal,
                                    uint8_t* d_tipo_out, float* d_sdf_out,
                                    cfd_vox_stats* stats, void* stream) {
  cfd__set_error("");
  int rc = check_args("cfd_voxelize_noalloc", d_pos, nv, d_idx, nt, nx, ny, nz, band,
                      d_tipo_out);
  if (rc) return rc;
  const size_t N = (size_t)nx * ny * nz;
  if (N > kNoAllocMaxCells)
    return vfail(CFD_E_ARG,
                 "cfd_voxelize_noalloc: %zu cells exceed the %zu cells of its fixed scratch "
                 "(cfd_voxelize has no such limit)", N, kNoAllocMaxCells);
  const unsigned tgrid = flood_tiles(nx, ny, nz);
  if (tgrid > kNoAllocMaxTiles)
    return vfail(CFD_E_ARG,
                 "cfd_voxelize_noalloc: grid %dx%dx%d needs %u flood tiles of 32x16x16 cells, "
                 "more than the %zu of its fixed scratch", nx, ny, nz, tgrid, kNoAllocMaxTiles);
  uint8_t* base = nullptr;
  rc = scratch_of_current_device(&base);
  if (rc) return rc;
  VoxBuffers b{};
  b.pos = d_pos;
  b.idx = d_idx;
  b.flags = base;
  b.barrier = base + kNoAllocMaxCells;
  b.ext = base + 2 * kNoAllocMaxCells;
  b.ext2 = base + 3 * kNoAllocMaxCells;
  b.dist2 = d_sdf_out ? reinterpret_cast<uint32_t*>(base + 4 * kNoAllocMaxCells) : nullptr;
  b.tipo = d_tipo_out;
  b.sdf = d_sdf_out;
  b.counters = reinterpret_cast<unsigned long long*>(base + kOffCounters);
  b.changed = reinterpret_cast<int*>(base + kOffChanged);
  b.tile_chg = base + kOffTiles;
  hipStream_t st = static_cast<hipStream_t>(stream);
  int iters = 0;
  unsigned long long hc[3] = {0, 0, 0};
  hipError_t e = run_voxelize(b, nv, nt, nx, ny, nz, band, seal, st, nullptr, &iters);
  if (e == hipSuccess)
    e = hipMemcpyAsync(hc, b.counters, sizeof hc, hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e != hipSuccess)
    return vfail(CFD_E_HIP, "cfd_voxelize_noalloc: %s", hipGetErrorString(e));
  if (stats) fill_stats(hc, N, iters, 0.0f, 0.0f, 0.0f, stats);
  return CFD_OK;
}
