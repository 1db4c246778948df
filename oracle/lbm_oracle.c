/*
 * lbm_oracle.c - CPU oracle of the reference's D3Q19 solver and voxelizer.
 *
 * TEST INFRASTRUCTURE ONLY.  This file restates, in plain C, the algorithm of
 * ceschmiedel/CFD_ for the hot path; it is the checker the GPU kernels are
 * compared against.  Only the test suite, the smoke check and the CPU-baseline
 * leg of the benchmark load it; the product (cfd__amd/libcfdmi.so) never
 * links, calls or falls back to it.
 *
 * Contents and the reference code each part follows:
 *   orc_step_f32 / orc_step_f64   the generated `passo` kernel:
 *                                 src/core/emit/ir.js:166-323, 383-533 and
 *                                 src/core/emit/wgsl.js:159-259
 *   orc_init_*                    wgsl.js:269-302 (shaderInit)
 *   orc_macros_*                  wgsl.js:442-476 (shaderMacros)
 *   orc_forces_*                  wgsl.js:332-435 (shaderForcas + shaderReduzir)
 *   orc_voxelize                  src/geom/voxel.js:61-433 (voxelizar)
 *   orc_equilibrium / orc_moments src/core/lattice.js:150-174 (float64)
 *
 * The float32 step uses the float32 literals the generator bakes in through
 * num() = toPrecision(9) (ir.js:99-106) and the operation order of the
 * emitted text; the float64 step uses exact fractions and is the error
 * reference of the reduced-precision storage.  Compile with
 * -ffp-contract=off (see Makefile) so products are rounded before adds.
 *
 * How each function is pinned (tests/test_oracle.py):
 *   step: uniform flow preserved to float32 round-off; shear-wave decay at
 *         nu = c_s^2 (1/omega - 1/2); Couette slope and wall position at
 *         z = 0.5 and nz - 1.5 independent of omega (Lambda = 3/16); mass
 *         conservation with an obstacle; rest state (g = 0) is a fixed point;
 *         a brute-force float64 collision of one cell against the moments /
 *         equilibrium identities of lattice.js:184-231.
 *   init / macros: closed form eq(0, u) and its moments.
 *   forces: rest offset = sum over links of 2 w_i c_i, zero for a closed
 *         immersed body; symmetry of a centred sphere.
 *   voxelize: the 12-triangle examples/cube_example.stl (hand-checked shell and
 *         interior), axis-aligned boxes, a single triangle, analytic sphere
 *         volume/area, and hashes of the reference's own voxelizar() output on
 *         seeded meshes (tests/golden).
 */
#include <math.h>
#include <stddef.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define Q 19

/* D3Q19 with opposite pairs adjacent (lattice.js:57-79). */
static const int CV[Q][3] = {
    {0, 0, 0},  {1, 0, 0},  {-1, 0, 0}, {0, 1, 0},  {0, -1, 0}, {0, 0, 1},   {0, 0, -1},
    {1, 1, 0},  {-1, -1, 0}, {1, -1, 0}, {-1, 1, 0}, {1, 0, 1},  {-1, 0, -1}, {1, 0, -1},
    {-1, 0, 1}, {0, 1, 1},  {0, -1, -1}, {0, 1, -1}, {0, -1, 1}};
static int OPPV[Q];
static int MIRV[3][Q];

static void tables(void) {
  static int done = 0;
  if (done) return;
  for (int i = 0; i < Q; ++i) {
    OPPV[i] = i == 0 ? 0 : (((i - 1) ^ 1) + 1); /* lattice.js:89 */
    for (int ax = 0; ax < 3; ++ax) {
      int c[3] = {CV[i][0], CV[i][1], CV[i][2]};
      c[ax] = -c[ax]; /* espelharIdx, ir.js:86-93 */
      for (int j = 0; j < Q; ++j)
        if (CV[j][0] == c[0] && CV[j][1] == c[1] && CV[j][2] == c[2]) MIRV[ax][i] = j;
    }
  }
  done = 1;
}

/* ---- learned closure (no reference counterpart) ----
 * The MLP 8 -> 32 -> 32 -> 1 of the library's closure, in float32 with the
 * bf16 rounding of its matrix-core path made explicit: the normalized
 * features, the weights and each hidden activation are rounded to bf16
 * (round to nearest even); products accumulate in float32 from the bias in
 * increasing input order (the matrix cores sum in another order, so results
 * agree to float32 accumulation error, not bit for bit).  A SiLU network
 * (activation 0) is evaluated in its base-2 form, as the matrix-core path:
 * W1, b1 and b2 are multiplied by -log2(e) in float32 (before the bf16
 * rounding of W1), so layers 1 and 2 give u = -log2(e) z and the SiLU is
 * u / (1 + 2^u) = -log2(e) SiLU(z); W2 is unchanged and w3 is multiplied by
 * -ln 2.  A hard-SiLU network (activation 1) is evaluated as it is, each
 * hidden unit z * clamp(fma(z, 1/6, 1/2), 0, 1).  Returns
 * out_scale * softplus(z), or NaN when a normalized feature leaves
 * [lo, hi] (the caller then keeps the Smagorinsky tau). */
typedef struct {
  float p[1377];                  /* w1[32][8] b1[32] w2[32][32] b2[32] w3[32] b3 */
  float shift[8], scale[8], lo[8], hi[8], out_scale;
  float activation;               /* 0 SiLU, 1 hard-SiLU */
} orc_closure;

static float bf16r(float f) {
  uint32_t u;
  memcpy(&u, &f, 4);
  if ((u & 0x7f800000u) == 0x7f800000u) return f;
  u += 0x7fffu + ((u >> 16) & 1u);
  u &= 0xffff0000u;
  memcpy(&f, &u, 4);
  return f;
}

static const float kNegLog2e = -1.44269504f, kNegLn2 = -0.693147181f;
static float silu2_ref(float u) { return u * (1.0f / (1.0f + exp2f(u))); }
static float hard_silu_ref(float z) {
  return z * fminf(fmaxf(fmaf(z, 1.0f / 6.0f, 0.5f), 0.0f), 1.0f);
}

static float orc_closure_nu(const orc_closure* c, const float x[8]) {
  const float *w1 = c->p, *b1 = c->p + 256, *w2 = c->p + 288, *b2 = c->p + 1312,
              *w3 = c->p + 1344, b3 = c->p[1376];
  const int hard = c->activation == 1.0f;
  const float s1 = hard ? 1.0f : kNegLog2e, s3 = hard ? 1.0f : kNegLn2;
  float xb[8], h1[32], h2[32];
  for (int k = 0; k < 8; ++k) {
    const float xn = (x[k] - c->shift[k]) * c->scale[k];
    if (!(xn >= c->lo[k] && xn <= c->hi[k])) return NAN;
    xb[k] = bf16r(xn);
  }
  for (int o = 0; o < 32; ++o) {
    float acc = s1 * b1[o];
    for (int k = 0; k < 8; ++k) acc += bf16r(s1 * w1[o * 8 + k]) * xb[k];
    h1[o] = bf16r(hard ? hard_silu_ref(acc) : silu2_ref(acc));
  }
  for (int o = 0; o < 32; ++o) {
    float acc = s1 * b2[o];
    for (int k = 0; k < 32; ++k) acc += bf16r(w2[o * 32 + k]) * h1[k];
    h2[o] = bf16r(hard ? hard_silu_ref(acc) : silu2_ref(acc));
  }
  float z = b3;
  for (int k = 0; k < 32; ++k) z += bf16r(s3 * w3[k]) * h2[k];
  const float sp = z > 20.0f ? z : (z < -15.0f ? expf(z) : log1pf(expf(z)));
  return c->out_scale * sp;
}

void orc_closure_eval(const orc_closure* c, const float* X, size_t n, float* nu) {
  for (size_t s = 0; s < n; ++s) nu[s] = orc_closure_nu(c, X + 8 * s);
}

/* ---- float32 instantiation: literals as the generator prints them ---- */
#define R float
#define SFX(n) n##_f32
#define W0 0.333333333f
#define WA 0.0555555556f
#define WD 0.0277777778f
#define KA 0.333333333f
#define KD 0.166666667f
#define C18 18.0f
#define PEXP 0.1428571429f
#define SQRT sqrtf
#define POW powf
#define FMIN fminf
#define FMAX fmaxf
#include "lbm_step.inc"
#undef R
#undef SFX
#undef W0
#undef WA
#undef WD
#undef KA
#undef KD
#undef C18
#undef PEXP
#undef SQRT
#undef POW
#undef FMIN
#undef FMAX

/* ---- float64 instantiation: exact fractions ---- */
#define R double
#define SFX(n) n##_f64
#define W0 (1.0 / 3.0)
#define WA (1.0 / 18.0)
#define WD (1.0 / 36.0)
#define KA (1.0 / 3.0)
#define KD (1.0 / 6.0)
#define C18 18.0
#define PEXP (1.0 / 7.0)
#define SQRT sqrt
#define POW pow
#define FMIN fmin
#define FMAX fmax
#include "lbm_step.inc"

/* Exported wrappers make sure the tables exist before any use. */
void orc_step_f32(int nx, int ny, int nz, const float* src, float* dst, const uint32_t* tipo,
                  const double* params, float* macros);
void orc_step_f64(int nx, int ny, int nz, const double* src, double* dst, const uint32_t* tipo,
                  const double* params, double* macros);
void orc_step_closure_f32(int nx, int ny, int nz, const float* src, float* dst,
                          const uint32_t* tipo, const double* params, float* macros,
                          const orc_closure* cl, const float* wdist) {
  tables();
  step_impl_f32(nx, ny, nz, src, dst, tipo, params, macros, cl, wdist);
}

int orc_setup(void) {
  tables();
  return Q;
}

int orc_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

/* lattice.js:150-160: shifted equilibrium, float64. */
void orc_equilibrium(double delta, const double u[3], double g[Q]) {
  const double W[3] = {1.0 / 3.0, 1.0 / 18.0, 1.0 / 36.0};
  const double rho = 1 + delta;
  const double uu = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
  for (int i = 0; i < Q; ++i) {
    const double cu = CV[i][0] * u[0] + CV[i][1] * u[1] + CV[i][2] * u[2];
    g[i] = W[i == 0 ? 0 : (i <= 6 ? 1 : 2)] * (delta + rho * (3 * cu + 4.5 * cu * cu - 1.5 * uu));
  }
}

/* lattice.js:163-174: moments of shifted populations, float64. */
void orc_moments(const double g[Q], double out[4]) {
  double delta = 0, mx = 0, my = 0, mz = 0;
  for (int i = 0; i < Q; ++i) {
    delta += g[i];
    mx += g[i] * CV[i][0];
    my += g[i] * CV[i][1];
    mz += g[i] * CV[i][2];
  }
  const double inv = 1 / (1 + delta);
  out[0] = delta;
  out[1] = mx * inv;
  out[2] = my * inv;
  out[3] = mz * inv;
}

/* ======================================================== voxelizer ===== */
/* src/geom/voxel.js, float64 (JavaScript numbers) on float32 input. */

static double min3(double a, double b, double c) {
  double m = a < b ? a : b;
  return m < c ? m : c;
}
static double max3(double a, double b, double c) {
  double m = a > b ? a : b;
  return m > c ? m : c;
}

/* trianguloTocaCaixa, voxel.js:61-98 */
static int tri_box(const double v0[3], const double v1[3], const double v2[3], const double h[3]) {
  const double e0[3] = {v1[0] - v0[0], v1[1] - v0[1], v1[2] - v0[2]};
  const double e1[3] = {v2[0] - v1[0], v2[1] - v1[1], v2[2] - v1[2]};
  const double e2[3] = {v0[0] - v2[0], v0[1] - v2[1], v0[2] - v2[2]};
  const double* es[3] = {e0, e1, e2};
  for (int k = 0; k < 3; ++k) {
    const double* e = es[k];
    const double fx = fabs(e[0]), fy = fabs(e[1]), fz = fabs(e[2]);
    /* eixo(a, b, 0, fa, fb, i, j) */
    const double A[3] = {e[2], -e[2], e[1]};
    const double B[3] = {-e[1], e[0], -e[0]};
    const double FA[3] = {fz, fz, fy};
    const double FB[3] = {fy, fx, fx};
    const int I[3] = {1, 0, 0}, J[3] = {2, 2, 1};
    for (int q = 0; q < 3; ++q) {
      const double p0 = A[q] * v0[I[q]] + B[q] * v0[J[q]];
      const double p1 = A[q] * v1[I[q]] + B[q] * v1[J[q]];
      const double p2 = A[q] * v2[I[q]] + B[q] * v2[J[q]];
      const double r = FA[q] * h[I[q]] + FB[q] * h[J[q]];
      if (min3(p0, p1, p2) > r || max3(p0, p1, p2) < -r) return 0;
    }
  }
  for (int k = 0; k < 3; ++k) {
    if (min3(v0[k], v1[k], v2[k]) > h[k]) return 0;
    if (max3(v0[k], v1[k], v2[k]) < -h[k]) return 0;
  }
  const double n[3] = {e0[1] * e1[2] - e0[2] * e1[1], e0[2] * e1[0] - e0[0] * e1[2],
                       e0[0] * e1[1] - e0[1] * e1[0]};
  const double d = n[0] * v0[0] + n[1] * v0[1] + n[2] * v0[2];
  const double r = fabs(n[0]) * h[0] + fabs(n[1]) * h[1] + fabs(n[2]) * h[2];
  return fabs(d) <= r;
}

/* distQuadPontoTriangulo, voxel.js:105-151 */
static double dist2pt(const double p[3], const double a[3], const double b[3], const double c[3]) {
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
    const double qx = bpx + w * (cpx - bpx), qy = bpy + w * (cpy - bpy), qz = bpz + w * (cpz - bpz);
    return qx * qx + qy * qy + qz * qz;
  }
  const double den = 1 / (va + vb + vc);
  const double v = vb * den, w = vc * den;
  const double qx = apx - (v * abx + w * acx);
  const double qy = apy - (v * aby + w * acy);
  const double qz = apz - (v * abz + w * acz);
  return qx * qx + qy * qy + qz * qz;
}

/* voxelizar (voxel.js:376-433).  stats: [nCasca, nDentro, nSolido,
 * trianguloFora, ocoSuspeito].  sdf may be NULL. */
int orc_voxelize(const float* pos, size_t nv, const uint32_t* idx, size_t nt, int nx, int ny,
                 int nz, double banda, int selar, uint8_t* tipo, float* sdf, int64_t stats[5]) {
  const size_t nxny = (size_t)nx * ny, N = nxny * nz;
  uint8_t* flags = calloc(N, 1);
  float* dist2 = malloc(N * sizeof(float));
  uint8_t* barreira = NULL;
  uint8_t* visitado = calloc(N, 1);
  int32_t* pilha = malloc(N * sizeof(int32_t));
  if (!flags || !dist2 || !visitado || !pilha) return -1;
  for (size_t i = 0; i < N; ++i) dist2[i] = INFINITY;

  /* rasterizar (169-228) */
  const double h[3] = {0.5, 0.5, 0.5};
  int64_t fora = 0;
  for (size_t t = 0; t < nt; ++t) {
    const size_t i0 = idx[t * 3], i1 = idx[t * 3 + 1], i2 = idx[t * 3 + 2];
    if (i0 >= nv || i1 >= nv || i2 >= nv) continue; /* undefined -> NaN -> skipped */
    const double a[3] = {pos[i0 * 3], pos[i0 * 3 + 1], pos[i0 * 3 + 2]};
    const double b[3] = {pos[i1 * 3], pos[i1 * 3 + 1], pos[i1 * 3 + 2]};
    const double c[3] = {pos[i2 * 3], pos[i2 * 3 + 1], pos[i2 * 3 + 2]};
    if (!isfinite(a[0] + a[1] + a[2] + b[0] + b[1] + b[2] + c[0] + c[1] + c[2])) continue;
    const int n[3] = {nx, ny, nz};
    double lo[3], hi[3];
    int empty = 0;
    for (int k = 0; k < 3; ++k) {
      const double f = floor(min3(a[k], b[k], c[k]) - banda);
      lo[k] = f > 0 ? f : 0; /* Math.max(0, .) */
      const double g = ceil(max3(a[k], b[k], c[k]) + banda);
      hi[k] = g < n[k] - 1 ? g : n[k] - 1; /* Math.min(n-1, .) */
      if (hi[k] < lo[k]) empty = 1;
    }
    if (empty) {
      fora++;
      continue;
    }
    for (int z = (int)lo[2]; z <= (int)hi[2]; ++z)
      for (int y = (int)lo[1]; y <= (int)hi[1]; ++y)
        for (int x = (int)lo[0]; x <= (int)hi[0]; ++x) {
          const double p[3] = {x + 0.5, y + 0.5, z + 0.5};
          const size_t cell = ((size_t)z * ny + y) * nx + x;
          const double d2 = dist2pt(p, a, b, c);
          if (d2 < dist2[cell]) dist2[cell] = (float)d2; /* Float32Array store */
          if (d2 <= 0.75 && flags[cell] != 1) {
            const double la[3] = {a[0] - p[0], a[1] - p[1], a[2] - p[2]};
            const double lb[3] = {b[0] - p[0], b[1] - p[1], b[2] - p[2]};
            const double lc[3] = {c[0] - p[0], c[1] - p[1], c[2] - p[2]};
            if (tri_box(la, lb, lc, h)) flags[cell] = 1;
          }
        }
  }

  /* dilatar (244-267) */
  if (selar) {
    barreira = calloc(N, 1);
    for (int z = 0; z < nz; ++z)
      for (int y = 0; y < ny; ++y)
        for (int x = 0; x < nx; ++x) {
          if (flags[((size_t)z * ny + y) * nx + x] != 1) continue;
          for (int k = z > 0 ? z - 1 : 0; k <= (z < nz - 1 ? z + 1 : nz - 1); ++k)
            for (int j = y > 0 ? y - 1 : 0; j <= (y < ny - 1 ? y + 1 : ny - 1); ++j)
              for (int i = x > 0 ? x - 1 : 0; i <= (x < nx - 1 ? x + 1 : nx - 1); ++i)
                barreira[((size_t)k * ny + j) * nx + i] = 1;
        }
  }
  const uint8_t* bar = selar ? barreira : flags;

  /* inundarExterior (278-327): explicit-stack DFS from the six faces */
  size_t topo = 0;
#define PUSH(c)                                   \
  do {                                            \
    size_t c_ = (c);                              \
    if (!visitado[c_] && !bar[c_]) {              \
      visitado[c_] = 1;                           \
      pilha[topo++] = (int32_t)c_;                \
    }                                             \
  } while (0)
  for (int z = 0; z < nz; ++z)
    for (int y = 0; y < ny; ++y) {
      PUSH(((size_t)z * ny + y) * nx);
      PUSH(((size_t)z * ny + y) * nx + nx - 1);
    }
  for (int z = 0; z < nz; ++z)
    for (int x = 0; x < nx; ++x) {
      PUSH(((size_t)z * ny) * nx + x);
      PUSH(((size_t)z * ny + ny - 1) * nx + x);
    }
  for (int y = 0; y < ny; ++y)
    for (int x = 0; x < nx; ++x) {
      PUSH((size_t)y * nx + x);
      PUSH(((size_t)(nz - 1) * ny + y) * nx + x);
    }
  while (topo > 0) {
    const size_t cell = (size_t)pilha[--topo];
    const int x = (int)(cell % nx);
    const int y = (int)((cell / nx) % ny);
    const int z = (int)(cell / nxny);
    if (x > 0) PUSH(cell - 1);
    if (x < nx - 1) PUSH(cell + 1);
    if (y > 0) PUSH(cell - nx);
    if (y < ny - 1) PUSH(cell + nx);
    if (z > 0) PUSH(cell - nxny);
    if (z < nz - 1) PUSH(cell + nxny);
  }
#undef PUSH

  /* recuperarPele (333-355) */
  if (selar) {
    uint8_t* antes = malloc(N);
    memcpy(antes, visitado, N);
    for (int z = 0; z < nz; ++z)
      for (int y = 0; y < ny; ++y)
        for (int x = 0; x < nx; ++x) {
          const size_t c = ((size_t)z * ny + y) * nx + x;
          if (antes[c] || flags[c] == 1) continue;
          const int viz = (x > 0 && antes[c - 1]) || (x < nx - 1 && antes[c + 1]) ||
                          (y > 0 && antes[c - nx]) || (y < ny - 1 && antes[c + nx]) ||
                          (z > 0 && antes[c - nxny]) || (z < nz - 1 && antes[c + nxny]);
          if (viz) visitado[c] = 1;
        }
    free(antes);
  }

  /* classification and signed distance (401-415) */
  int64_t nCasca = 0, nDentro = 0;
  for (size_t i = 0; i < N; ++i) {
    uint8_t t = 0;
    if (flags[i] == 1) {
      t = 1;
      nCasca++;
    } else if (!visitado[i]) {
      t = 2;
      nDentro++;
    }
    tipo[i] = t;
    if (sdf) {
      const double d = isinf(dist2[i]) ? banda + 1 : sqrt((double)dist2[i]);
      sdf[i] = (float)(t == 0 ? d : -d);
    }
  }
  stats[0] = nCasca;
  stats[1] = nDentro;
  stats[2] = nCasca + nDentro;
  stats[3] = fora;
  stats[4] = (nDentro < nCasca * 0.25 && nCasca > 500) ? 1 : 0;
  free(flags);
  free(dist2);
  free(barreira);
  free(visitado);
  free(pilha);
  return 0;
}
