/*
 * cfd_mi355x.h — C ABI of the MI355X D3Q19 TRT-LES solver and GPU voxelizer.
 *
 * Every entry point returns 0 on success and a negative code on failure; the
 * failure text is available from cfd_last_error() (thread-local).  No C++ or
 * torch types cross this boundary: plain pointers, sizes and PODs only.
 *
 * Threading: one solver handle is driven by one host thread.  All device work
 * of a handle is issued on that handle's own HIP stream; cfd_step() and
 * cfd_update_macros() only enqueue, every function that returns data to the
 * host synchronizes that stream first.
 *
 * Reference interface each function replaces (ceschmiedel/CFD_, the
 * SolverWebGPU object of src/backend/webgpu/solver.js and the voxelizer of
 * src/geom/voxel.js) is cited beside each declaration.
 */
#ifndef CFD_MI355X_H
#define CFD_MI355X_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Error codes (negative). */
#define CFD_OK 0
#define CFD_E_ARG -1     /* bad argument: size mismatch, null pointer, range   */
#define CFD_E_HIP -2     /* HIP runtime failure (no device, launch failure)    */
#define CFD_E_OOM -3     /* device allocation failed                           */
#define CFD_E_STATE -4   /* call not valid in the handle's current state       */

/* Storage flags for cfd_create. */
#define CFD_STORE_FP32 0u   /* populations g_i = f_i - w_i stored as float32    */
#define CFD_STORE_FP16 1u   /* populations stored as scaled float16 deviations */
#define CFD_ROUND_STOCHASTIC 2u /* with FP16: the step rounds each stored value up or
                                   down at random in proportion (unbiased)      */
#define CFD_ROUND_EXACT 4u   /* with FP32: the step rounds every operation as the
                                reference's shader text with correctly rounded
                                division and square root (default: 1-ulp
                                reciprocal / square root and fused multiply-adds,
                                within WGSL's 2.5 ulp division accuracy)        */

/* Cell types, identical to CELULA in src/core/emit/ir.js:71-80. */
#define CFD_FLUIDO 0u
#define CFD_SOLIDO 1u
#define CFD_ENTRADA 2u
#define CFD_SAIDA 3u
#define CFD_SOLIDO_MOVEL 4u
#define CFD_ESPELHO_Y 5u
#define CFD_ESPELHO_Z 6u
#define CFD_PAREDE 7u

typedef struct cfd_solver cfd_solver;

/* The 80-byte Params uniform of src/backend/webgpu/solver.js:277-293
 * (wgsl.js:125-134).  Values are rounded to float32 on upload exactly as the
 * reference's Float32Array.set() does. */
typedef struct cfd_params {
  double omega_plus;    /* even relaxation rate (molecular viscosity)          */
  double magic;         /* TRT Lambda; 3/16 puts the wall half-way             */
  double les_cs;        /* Smagorinsky constant; 0 disables the LES term       */
  double belt[3];       /* SOLIDO_MOVEL wall velocity (rolling road)           */
  double inlet[3];      /* free stream, lattice units                          */
  double sponge_start;  /* first x of the sponge ramp                          */
  double sponge_len;    /* sponge ramp length in cells                         */
  double bl_thickness;  /* inlet 1/7-power boundary-layer thickness (cells)    */
} cfd_params;

/* Launch shape and layout of the step kernel.  No reference counterpart:
 * the WGSL path has one fixed 64x1x1 workgroup (wgsl.js:186).  cfd_create
 * picks the shape measured fastest for the set: float32 sets above 256 MB
 * get 512 threads with the start rotated by 36 % of the grid, float16 sets
 * above 128 MB and float32 sets of 128-256 MB alternate the order, others
 * run 256 threads in ascending order; all with XCD chunks of 8.  Float16
 * sets up to 128 MB are SoA, every other set below 4 GiB blocked. */
typedef struct cfd_tuning {
  int32_t block;        /* threads per workgroup: 256 or 512                   */
  int32_t remap;        /* 0/1: launch order; K (power of two): each XCD steps
                           chunks of K consecutive workgroups                   */
  int32_t layout;       /* 0: SoA pop[i][cell]; 1: 512-cell blocks
                           pop[cell/512][i][cell%512] (default; needs a
                           population set below 4 GiB, else SoA is used)        */
  int32_t slice_pad;    /* SoA only: extra elements between the 19 slices      */
  int32_t sweep;        /* 0: every step walks the cells in ascending order;
                           1: odd steps walk them in descending order, so a
                           step first reads what the previous one wrote last */
  int32_t rotate;       /* 0: off; R > 0 (sweep 0 only): step k starts k*R
                           workgroups before step 0's start (modulo the grid),
                           so its first R workgroups read what the previous
                           step's last R wrote                                 */
} cfd_tuning;

typedef struct cfd_vox_stats {
  int64_t n_cells;
  int64_t n_casca;          /* shell voxels (VOXEL.CASCA)                      */
  int64_t n_dentro;         /* interior voxels (VOXEL.DENTRO)                  */
  int64_t n_solido;         /* n_casca + n_dentro                              */
  int64_t triangles_outside;/* triangles whose padded box misses the grid     */
  int32_t hollow_suspect;   /* ocoSuspeito: n_dentro < 0.25 n_casca, n_casca>500 */
  int32_t flood_iterations; /* global sweeps of the exterior flood fill        */
  double ms_raster;         /* device time of the rasterization               */
  double ms_flood;          /* device time of seal + flood + skin recovery     */
  double ms_total;          /* device time of the whole voxelization           */
} cfd_vox_stats;

/* SolverWebGPU.criar (solver.js:93-124) + _alocar (132-174).  Allocates two
 * population sets (A and B), the cell-type field and the macro field, all
 * zero-filled: g_i = 0 is fluid at rest with rho = 1 exactly (lattice.js:32-34).
 * The front starts at A.  flags: CFD_STORE_FP32, CFD_STORE_FP16 or
 * CFD_STORE_FP16 | CFD_ROUND_STOCHASTIC or CFD_STORE_FP32 | CFD_ROUND_EXACT. */
int cfd_create(int nx, int ny, int nz, uint32_t flags, cfd_solver **out);

/* SolverWebGPU.configurar (solver.js:277-293). */
int cfd_configure(cfd_solver *s, const cfd_params *p);

/* SolverWebGPU.definirTipos (solver.js:296-301): n must equal nx*ny*nz.
 * Values 0..7 are CELULA codes; any other value behaves like the reference's
 * fallthrough: streamed and collided as fluid, never a force source. */
int cfd_set_types(cfd_solver *s, const uint32_t *tipo, size_t n);

/* SolverWebGPU.inicializar (solver.js:338-349) with the init kernel of
 * wgsl.js:269-302: equilibrium at the inlet velocity (0 in SOLIDO, beltU in
 * SOLIDO_MOVEL) written to B; the front becomes B. */
int cfd_init_equilibrium(cfd_solver *s);

/* SolverWebGPU.definirCampo (solver.js:363-389) upload half: writes a full
 * population field pop[i*N + cell] (i = 0..18, cell = (z*ny+y)*nx+x) into the
 * current front.  n must equal 19*N. */
int cfd_set_populations(cfd_solver *s, const float *soa, size_t n);

/* Diagnostic read-back of the current front in the same layout (no reference
 * counterpart; used by the parity tests).  n must equal 19*N. */
int cfd_get_populations(cfd_solver *s, float *soa_out, size_t n);

/* The writable `frente` field (solver.js:245, tests write it at
 * nucleo.html:149): 0 = A, 1 = B. */
int cfd_set_front(cfd_solver *s, int front);
int cfd_get_front(cfd_solver *s, int *front);

/* SolverWebGPU.passo (solver.js:395-415) device half: enqueue n steps of the
 * fused pull-stream + TRT + Smagorinsky + boundary kernel (wgsl.js:230-259),
 * flipping the front after each.  The macro field of the last step (its
 * pulled, pre-collision moments; zero in solids, wgsl.js:255) is what
 * cfd_read_macros returns; it is computed from the set that step read when
 * it is first read (or before that set, the types or the layout change),
 * not stored by every step.  Non-blocking. */
int cfd_step(cfd_solver *s, int n);

/* Launch tuning and population layout of the step kernel (autotuner hook;
 * no reference counterpart).  Changing layout or slice_pad re-lays both
 * population sets (contents preserved).  Results do not depend on it. */
int cfd_set_tuning(cfd_solver *s, const cfd_tuning *t);
int cfd_get_tuning(cfd_solver *s, cfd_tuning *t);

/* Same as cfd_step, then synchronizes and returns the device time between HIP
 * events recorded on the handle's stream around the n launches (ms). */
int cfd_step_timed(cfd_solver *s, int n, double *ms_out);

/* SolverWebGPU.medirForcas (solver.js:486-508) with shaderForcas/shaderReduzir
 * (wgsl.js:332-435): momentum-exchange force on SOLIDO cells from the current
 * front, accumulated in float64.  Raw value: the rest offset of
 * calibrarRepouso (solver.js:480-484) is subtracted by the host wrapper. */
int cfd_forces(cfd_solver *s, double out_f[3]);

/* SolverWebGPU.atualizarMacros (solver.js:511-520): recompute the macro field
 * from the current front without streaming (wgsl.js:442-476).  Non-blocking. */
int cfd_update_macros(cfd_solver *s);

/* SolverWebGPU.lerMacros (solver.js:427-442): copies the N*4 floats
 * (ux, uy, uz, delta) to the host.  n must equal 4*N. */
int cfd_read_macros(cfd_solver *s, float *out, size_t n);

/* Float64 mass-conservation reduction (the sum of delta that the reference's
 * nucleo.html:213-222 does on the host): sum over all non-solid cells of
 * sum_i g_i of the current front, and the number of non-solid cells holding
 * a non-finite population (left out of the sum): the divergence signal the
 * reference's stability page reads from the macro field (estabilidade.html:
 * 82-88). */
int cfd_mass(cfd_solver *s, double *mass_out, int64_t *n_nonfinite_out);

/* Blocks until all work queued on the handle's stream has finished. */
int cfd_synchronize(cfd_solver *s);

/* Device bytes held by the handle (SolverWebGPU.bytesTotais, solver.js:126-128). */
int cfd_bytes(cfd_solver *s, int64_t *bytes_out);

/* SolverWebGPU.destruir (solver.js:523-527). */
void cfd_destroy(cfd_solver *s);

/* ---- learned sub-grid closure ----
 * No reference counterpart: the reference's only closure is the closed-form
 * Smagorinsky of ir.js:262-301.  A per-cell MLP 8 -> 32 -> 32 -> 1 (two
 * hidden activations, then softplus) gives the eddy viscosity nu_t; the step
 * then relaxes with
 * tau = 1/omega_plus + 3 nu_t instead of the Smagorinsky tau.  Features of a
 * fluid cell: Pi_xx, Pi_yy, Pi_zz, Pi_xy, Pi_xz, Pi_yz of the non-equilibrium
 * stress divided by rho, the strain rate |S| = 3 Q_s / (2 rho tau_s) implied
 * by the Smagorinsky solution tau_s, and the wall distance in cells.  A cell
 * whose normalized features leave [in_lo, in_hi], or whose output is not
 * finite, keeps the Smagorinsky tau (counted by cfd_closure_fallbacks).
 * Inference uses bf16 operands with float32 accumulation on the matrix
 * cores; training (cfd_mlp_train) is float32 Adam on minibatches.  The hidden
 * activation is the network's own (cfd_mlp.activation):
 *   CFD_ACT_SILU       z / (1 + e^-z), evaluated in base 2 on the matrix-core
 *                      path (w1, b1, b2 times -log2(e) and w3 times -ln 2 in
 *                      float32, then bf16: the same function, with the bf16
 *                      roundings of that form);
 *   CFD_ACT_HARD_SILU  z * clamp(z / 6 + 1/2, 0, 1), the piecewise-quadratic
 *                      form of SiLU (hardswish): one clamped fma and one
 *                      multiply per hidden unit instead of an exponential and
 *                      a reciprocal. */
#define CFD_ACT_SILU 0
#define CFD_ACT_HARD_SILU 1
typedef struct cfd_mlp {   /* row-major [out][in]; 1377 floats in this order */
  float w1[32 * 8];
  float b1[32];
  float w2[32 * 32];
  float b2[32];
  float w3[32];
  float b3;
  int32_t activation;      /* CFD_ACT_SILU or CFD_ACT_HARD_SILU            */
} cfd_mlp;

typedef struct cfd_closure {
  cfd_mlp mlp;
  float in_shift[8], in_scale[8]; /* x_n = (x - shift) * scale                  */
  float in_lo[8], in_hi[8];       /* Smagorinsky fallback outside [lo, hi] (x_n) */
  float out_scale;                /* nu_t = out_scale * softplus(mlp(x_n))        */
} cfd_closure;

typedef struct cfd_train_cfg {
  int32_t steps;                  /* Adam steps                                  */
  int32_t batch;                  /* samples per step, a multiple of 64          */
  float lr, beta1, beta2, eps;
  uint64_t seed;                  /* minibatch draw                              */
} cfd_train_cfg;

/* Use the closure in every following step (NULL: back to Smagorinsky).  Its
 * wall-distance feature starts as cfd_compute_wall_distance's field. */
int cfd_set_closure(cfd_solver *s, const cfd_closure *c);
/* Distance of every cell to the nearest wall cell, in cells (n = N), from
 * the caller; it stays until cfd_compute_wall_distance replaces it. */
int cfd_set_wall_distance(cfd_solver *s, const float *d, size_t n);
/* The same field computed on the device from the types: exact Euclidean
 * distance to the nearest wall cell (1, 4, 7) centre, quantized as
 * cfd_set_wall_distance quantizes it, and recomputed by every later
 * cfd_set_types.  No reference counterpart (the learned closure is an
 * addition); it replaces a host distance transform. */
int cfd_compute_wall_distance(cfd_solver *s);
/* The quantized field the closure reads, in eighths of a cell (n = N). */
int cfd_get_wall_distance(cfd_solver *s, uint8_t *q, size_t n);
/* Fluid cell-steps that fell back to Smagorinsky since the last reset. */
int cfd_closure_fallbacks(cfd_solver *s, int64_t *count, int reset);
/* Train params in place on X (n x 8, normalized features) and y (n targets
 * of softplus(mlp(x)), i.e. nu_t / out_scale); loss_curve: steps floats (the
 * minibatch mean squared error), nullable. */
int cfd_mlp_train(const float *X, const float *y, size_t n, const cfd_train_cfg *cfg,
                  cfd_mlp *params, float *loss_curve);
/* Full-batch gradient of the mean squared error of softplus(mlp(X)) - y. */
int cfd_mlp_grad(const float *X, const float *y, size_t n, const cfd_mlp *params, cfd_mlp *grad,
                 double *loss);
/* The closure's inference path on raw features X (n x 8): nu_t and whether the
 * normalized features lie inside the trained range. */
int cfd_mlp_infer(const float *X, size_t n, const cfd_closure *c, float *nu, uint8_t *inside);

/* ---- batched tunnels: many independent float32 tunnels, one launch ----
 * No reference counterpart (the reference runs one tunnel per page).  Each
 * tunnel is a grid, a cell-type field and a parameter set; all tunnels share
 * the kernel of cfd_step.  cfd_batch_run advances every tunnel by `steps`
 * steps inside ONE cooperative launch: tunnels are admitted smallest first
 * while the admitted tiles stay within max_active_tiles, each round steps
 * every admitted tunnel once, then the grid synchronizes; a finished tunnel
 * leaves and the next is admitted.  Force samples (momentum exchange of the
 * state after every `sample_every` steps, float64) and each tunnel's
 * completion time are kept for cfd_batch_results. */
typedef struct cfd_batch cfd_batch;

typedef struct cfd_tunnel_spec {
  int32_t nx, ny, nz;
  const uint32_t *tipo;   /* nx*ny*nz CELULA codes, copied at creation       */
  cfd_params params;
} cfd_tunnel_spec;

typedef struct cfd_batch_tuning {
  int32_t block;            /* threads per workgroup: 256 or 512             */
  int32_t tile_cells;       /* consecutive cells per tile (64 .. 2^20)       */
  int32_t wg_per_cu;        /* resident workgroups per CU; 0 = maximum       */
  int32_t max_active_tiles; /* admission cap; 0 = 8 tiles per workgroup      */
  int32_t tiles_per_grab;   /* consecutive tiles a workgroup takes at once;
                               0 = 1                                          */
} cfd_batch_tuning;

/* Allocates every tunnel (blocked float32 layout), uploads its types and
 * initializes it to the free-stream equilibrium. */
int cfd_batch_create(const cfd_tunnel_spec *specs, int n, cfd_batch **out);
/* Back to the free-stream equilibrium (e.g. after autotuning runs). */
int cfd_batch_reset(cfd_batch *b);
/* One cooperative launch running `steps` steps of every tunnel; tuning may
 * be NULL (last used / default).  ms_out: device time of the launch.
 * CFD_E_HIP also when a wait of the launch's grid barrier timed out (the
 * launch then finishes early and the tunnels' state is not to be used). */
int cfd_batch_run(cfd_batch *b, int steps, int sample_every, const cfd_batch_tuning *t,
                  double *ms_out);
/* forces: n * n_samples * 3 doubles (nullable); finish_ms: n doubles, time
 * from launch start to the end of the tunnel's last step (nullable). */
int cfd_batch_results(cfd_batch *b, double *forces, double *finish_ms, int *n_samples,
                      int *rounds);
/* Current populations of tunnel k, layout pop[i*N + cell]; n = 19*N.  Runs
 * continue from the current state (cfd_batch_reset returns to the
 * free-stream equilibrium). */
int cfd_batch_get_populations(cfd_batch *b, int k, float *out, size_t n);
/* Replace the current state of tunnel k (e.g. a perturbed initial field). */
int cfd_batch_set_populations(cfd_batch *b, int k, const float *in, size_t n);
/* Residency limit of the batch kernel for a workgroup size. */
int cfd_batch_max_grid(int block, int *workgroups_per_cu, int *compute_units);
void cfd_batch_destroy(cfd_batch *b);

/* voxelizar (src/geom/voxel.js:376-433) on the GPU: band rasterization with
 * the exact triangle/box overlap test (61-98) and point-triangle distance
 * (105-151) in float64 without contraction, 26-neighbour seal dilation
 * (244-267), 6-connected exterior flood (278-327) and skin recovery (333-355).
 * pos: nv*3 floats in lattice space; idx: nt*3 vertex indices.  tipo_out: N
 * bytes (0 FLUIDO, 1 CASCA, 2 DENTRO); sdf_out (nullable): N floats.
 * Bit-identical to the reference for the same inputs. */
int cfd_voxelize(const float *pos, size_t nv, const uint32_t *idx, size_t nt,
                 int nx, int ny, int nz, double band, int seal,
                 uint8_t *tipo_out, float *sdf_out, cfd_vox_stats *stats);

/* cfd_voxelize for callers that cannot allocate device memory: code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc.
 * No counterpart in the reference.  The mesh (d_pos, d_idx) and the results
 * (d_tipo_out: N bytes, d_sdf_out: N floats, nullable) are DEVICE pointers
 * owned by the caller; the work runs on `stream` (a hipStream_t; NULL is the
 * null stream), which the call synchronizes before it returns (the flood
 * reads its change flag back after every sweep).  It calls no hipMalloc,
 * hipFree, stream or event creation: its scratch is one fixed device array
 * of 64 MiB + 512 KiB + 512 B per GPU, held from library load to process
 * end, which limits the grid to CFD_VOXELIZE_NOALLOC_MAX_CELLS cells (and
 * 2^18 flood tiles of 32x16x16 cells); larger grids fail with CFD_E_ARG.
 * Every call on one GPU shares that array: calls that overlap on one GPU
 * race.  Results are bit-identical to cfd_voxelize; stats carries the counts
 * and flood_iterations, its ms fields are 0 (no events are created). */
#define CFD_VOXELIZE_NOALLOC_MAX_CELLS 8388608 /* 2^23: 320x160x128 fits */
int cfd_voxelize_noalloc(const float *d_pos, size_t nv, const uint32_t *d_idx, size_t nt,
                         int nx, int ny, int nz, double band, int seal,
                         uint8_t *d_tipo_out, float *d_sdf_out, cfd_vox_stats *stats,
                         void *stream);

/* Achievable-HBM probe: a plain device-to-device copy kernel over `bytes`
 * bytes, `iters` times; returns the mean copy bandwidth (read+write GB/s).
 * Used to calibrate the roofline; no reference counterpart. */
int cfd_copy_bandwidth(size_t bytes, int iters, double *gbps_out, double *ms_out);

/* Thread-local text of the last failure ("" if none). */
const char *cfd_last_error(void);

/* Library build identification ("cfd_mi355x <version> gfx950"). */
const char *cfd_version(void);

/* Layout version of the structs this header declares (cfd_params,
 * cfd_tuning, cfd_closure, cfd_tunnel_spec, ...).  A binding compares the
 * CFD_ABI_VERSION of the header it was built against with cfd_abi_version()
 * of the library it loads and refuses a mismatch, so that no caller passes a
 * struct of another size.  No reference counterpart.
 *   2: cfd_tuning without rotate (library 0.1)
 *   3: cfd_tuning.rotate (library 0.2 and 0.3)
 *   4: cfd_mlp.activation (library 0.4 and later) */
#define CFD_ABI_VERSION 4
int cfd_abi_version(void);

#ifdef __cplusplus
}
#endif
#endif /* CFD_MI355X_H */
