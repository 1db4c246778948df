// Batched tunnels behind the C ABI: many independent float32 tunnels of
// uneven size advanced by ONE cooperative launch (k_batch_run).
//
// The schedule is built on the host.  Tunnels are admitted smallest first
// while the tiles of the admitted set stay within `max_active_tiles`; every
// round steps each admitted tunnel once; a tunnel leaves after `steps`
// rounds and the next one in line is admitted.  Small tunnels therefore
// finish early (their time to solution is the clock of their last round),
// and the admitted set always holds enough tiles to fill the device.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/cfd_mi355x.h"
#include "kernels.h"
#include "lattice.h"

using namespace cfd;

extern "C" void cfd__set_error(const char* msg);

namespace {

int bfail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  cfd__set_error(buf);
  return code;
}

#define BCK(expr)                                                                    \
  do {                                                                               \
    hipError_t e_ = (expr);                                                          \
    if (e_ != hipSuccess)                                                            \
      return bfail(e_ == hipErrorOutOfMemory ? CFD_E_OOM : CFD_E_HIP, "%s: %s (%s)",  \
                   __func__, hipGetErrorString(e_), #expr);                          \
  } while (0)


}  // namespace

struct cfd_batch {
  int n = 0;
  std::vector<BatchTunnel> host;     // descriptors (device pointers inside)
  std::vector<uint32_t> cells;
  BatchTunnel* dev = nullptr;
  hipStream_t stream = nullptr;
  int cur = 0;  // every tunnel's current state is in pop[cur]
  // last run
  int steps = 0, sample_every = 0, max_samples = 0, rounds = 0;
  std::vector<uint32_t> finish_round;
  double* samples = nullptr;
  unsigned long long* clocks = nullptr;
  size_t clocks_len = 0;
  cfd_batch_tuning tuning{512, 512, 0, 0, 1};
};

namespace {

void free_batch(cfd_batch* b) {
  if (!b) return;
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  for (auto& t : b->host) {
    if (t.pop[0]) (void)hipFree(t.pop[0]);
    if (t.pop[1]) (void)hipFree(t.pop[1]);
    if (t.tipo8) (void)hipFree((void*)t.tipo8);
  }
  if (b->dev) (void)hipFree(b->dev);
  if (b->samples) (void)hipFree(b->samples);
  if (b->clocks) (void)hipFree(b->clocks);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
}

// Admission schedule: phases of rounds with a fixed admitted set.
void build_schedule(const cfd_batch* b, int steps, uint32_t tile_cells, uint32_t cap,
                    std::vector<BatchPhase>* phases, std::vector<BatchEntry>* entries,
                    std::vector<uint32_t>* table, std::vector<uint32_t>* finish,
                    uint32_t* rounds) {
  std::vector<int> order(b->n);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(),
                   [&](int a, int c) { return b->cells[a] < b->cells[c]; });
  auto tiles = [&](int k) { return (b->cells[k] + tile_cells - 1) / tile_cells; };
  finish->assign(b->n, 0);
  std::vector<std::pair<int, uint32_t>> active;  // tunnel, start round
  size_t next = 0;
  uint32_t round = 0;
  auto admit = [&]() {
    uint32_t sum = 0;
    for (auto& a : active) sum += tiles(a.first);
    while (next < order.size() && (active.empty() || sum + tiles(order[next]) <= cap)) {
      sum += tiles(order[next]);
      active.push_back({order[next], round});
      ++next;
    }
  };
  admit();
  while (!active.empty()) {
    uint32_t r1 = UINT32_MAX;
    for (auto& a : active) r1 = std::min(r1, a.second + (uint32_t)steps);
    BatchPhase ph{};
    ph.r0 = round;
    ph.r1 = r1;
    ph.first = (uint32_t)entries->size();
    ph.n_active = (uint32_t)active.size();
    uint32_t pre = 0;
    ph.table = (uint32_t)table->size();
    for (size_t e = 0; e < active.size(); ++e) {
      const auto& a = active[e];
      entries->push_back({(uint32_t)a.first, a.second, pre, 0u});
      pre += tiles(a.first);
      table->resize(ph.table + pre, (uint32_t)e);
    }
    ph.total_tiles = pre;
    phases->push_back(ph);
    round = r1;
    std::vector<std::pair<int, uint32_t>> keep;
    for (auto& a : active) {
      if (a.second + (uint32_t)steps == round)
        (*finish)[a.first] = round;
      else
        keep.push_back(a);
    }
    active.swap(keep);
    admit();
  }
  *rounds = round;
}

}  // namespace

extern "C" {

int cfd_batch_create(const cfd_tunnel_spec* specs, int n, cfd_batch** out) {
  cfd__set_error("");
  if (!out || !specs || n < 1) return bfail(CFD_E_ARG, "cfd_batch_create: bad arguments");
  *out = nullptr;
  cfd_batch* b = new cfd_batch();
  b->n = n;
  auto bail = [&](int rc) {
    free_batch(b);
    return rc;
  };
  if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess)
    return bail(bfail(CFD_E_HIP, "cfd_batch_create: no HIP device visible"));
  for (int k = 0; k < n; ++k) {
    const cfd_tunnel_spec& sp = specs[k];
    if (sp.nx < 1 || sp.ny < 1 || sp.nz < 1 || !sp.tipo)
      return bail(bfail(CFD_E_ARG, "cfd_batch_create: tunnel %d has a bad grid or no types", k));
    const uint64_t N = (uint64_t)sp.nx * sp.ny * sp.nz;
    BatchTunnel t{};
    t.g = make_geom(sp.nx, sp.ny, sp.nz, kLayoutBlocked, 0);
    if (N >= (1ull << 30) || t.g.pop_elems * 4 >= (1ull << 32))
      return bail(bfail(CFD_E_ARG, "cfd_batch_create: tunnel %d is too large for a batch", k));
    t.p = make_step_params(sp.params);
    b->host.push_back(t);
    b->cells.push_back((uint32_t)N);
    BatchTunnel& h = b->host.back();
    const size_t pb = h.g.pop_elems * sizeof(float);
    hipError_t e = hipMalloc(&h.pop[0], pb);
    if (e == hipSuccess) e = hipMalloc(&h.pop[1], pb);
    uint8_t* t8 = nullptr;
    if (e == hipSuccess) e = hipMalloc(&t8, N);
    h.tipo8 = t8;
    uint32_t* t32 = nullptr;
    if (e == hipSuccess) e = hipMalloc(&t32, N * 4);
    if (e == hipSuccess)
      e = hipMemcpyAsync(t32, sp.tipo, N * 4, hipMemcpyHostToDevice, b->stream);
    if (e == hipSuccess) e = launch_prep_types(t32, t8, h.g, b->stream);
    if (e == hipSuccess) e = hipMemsetAsync(h.pop[1], 0, pb, b->stream);
    if (e == hipSuccess) e = launch_init(Store::F32, h.pop[0], t8, h.g, h.p, b->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
    if (t32) (void)hipFree(t32);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      return bail(bfail(e == hipErrorOutOfMemory ? CFD_E_OOM : CFD_E_HIP,
                        "cfd_batch_create: tunnel %d: %s", k, hipGetErrorString(e)));
    }
  }
  // the descriptors on the batch's stream (non-blocking: a null-stream copy
  // would not order before the runs that read them)
  if (hipMalloc(&b->dev, sizeof(BatchTunnel) * n) != hipSuccess ||
      hipMemcpyAsync(b->dev, b->host.data(), sizeof(BatchTunnel) * n, hipMemcpyHostToDevice,
                     b->stream) != hipSuccess ||
      hipStreamSynchronize(b->stream) != hipSuccess)
    return bail(bfail(CFD_E_HIP, "cfd_batch_create: descriptor upload failed"));
  *out = b;
  return CFD_OK;
}

int cfd_batch_reset(cfd_batch* b) {
  cfd__set_error("");
  if (!b) return bfail(CFD_E_ARG, "cfd_batch_reset: null handle");
  for (auto& h : b->host) {
    BCK(hipMemsetAsync(h.pop[1], 0, h.g.pop_elems * sizeof(float), b->stream));
    BCK(launch_init(Store::F32, h.pop[0], h.tipo8, h.g, h.p, b->stream));
  }
  BCK(hipStreamSynchronize(b->stream));
  b->cur = 0;
  return CFD_OK;
}

int cfd_batch_run(cfd_batch* b, int steps, int sample_every, const cfd_batch_tuning* tu,
                  double* ms_out) {
  cfd__set_error("");
  if (!b || steps < 1 || sample_every < 0)
    return bfail(CFD_E_ARG, "cfd_batch_run: bad arguments");
  cfd_batch_tuning t = tu ? *tu : b->tuning;
  if (t.block != 256 && t.block != 512)
    return bfail(CFD_E_ARG, "cfd_batch_run: block %d must be 256 or 512", t.block);
  if (t.tiles_per_grab < 0 || t.tiles_per_grab > 64)
    return bfail(CFD_E_ARG, "cfd_batch_run: tiles_per_grab %d outside [0, 64]",
                 t.tiles_per_grab);
  if (t.tile_cells < 64 || t.tile_cells > (1 << 20))
    return bfail(CFD_E_ARG, "cfd_batch_run: tile_cells %d outside [64, 2^20]", t.tile_cells);
  int per_cu = 0, cus = 0;
  BCK(batch_max_grid(t.block, &per_cu, &cus));
  if (per_cu < 1) return bfail(CFD_E_HIP, "cfd_batch_run: the batch kernel cannot be resident");
  const int wg_cu = t.wg_per_cu > 0 ? std::min(t.wg_per_cu, per_cu) : per_cu;
  const int grid = wg_cu * cus;
  const uint32_t cap = t.max_active_tiles > 0 ? (uint32_t)t.max_active_tiles : (uint32_t)grid * 8;

  std::vector<BatchPhase> phases;
  std::vector<BatchEntry> entries;
  std::vector<uint32_t> table;
  uint32_t rounds = 0;
  build_schedule(b, steps, (uint32_t)t.tile_cells, cap, &phases, &entries, &table,
                 &b->finish_round, &rounds);
  b->steps = steps;
  b->sample_every = sample_every;
  b->max_samples = sample_every > 0 ? steps / sample_every : 0;
  b->rounds = (int)rounds;

  BatchPhase* dph = nullptr;
  BatchEntry* den = nullptr;
  uint32_t* dtab = nullptr;
  uint32_t* dctr = nullptr;
  uint32_t* dbar = nullptr;
  BCK(hipMalloc(&dph, sizeof(BatchPhase) * phases.size()));
  hipError_t e = hipMalloc(&den, sizeof(BatchEntry) * entries.size());
  if (e == hipSuccess) e = hipMalloc(&dtab, sizeof(uint32_t) * table.size());
  if (e == hipSuccess) e = hipMalloc(&dctr, sizeof(uint32_t) * rounds);
  if (e == hipSuccess) e = hipMalloc(&dbar, sizeof(uint32_t) * kBarWords);
  if (b->samples) (void)hipFree(b->samples);
  if (b->clocks) (void)hipFree(b->clocks);
  b->samples = nullptr;
  b->clocks = nullptr;
  const size_t ns = (size_t)std::max(1, b->n * b->max_samples * 3);
  if (e == hipSuccess) e = hipMalloc(&b->samples, ns * sizeof(double));
  b->clocks_len = rounds + 1;
  if (e == hipSuccess) e = hipMalloc(&b->clocks, b->clocks_len * sizeof(unsigned long long));
  if (e == hipSuccess)
    e = hipMemcpyAsync(dph, phases.data(), sizeof(BatchPhase) * phases.size(),
                       hipMemcpyHostToDevice, b->stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(den, entries.data(), sizeof(BatchEntry) * entries.size(),
                       hipMemcpyHostToDevice, b->stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(dtab, table.data(), sizeof(uint32_t) * table.size(),
                       hipMemcpyHostToDevice, b->stream);
  if (e == hipSuccess) e = hipMemsetAsync(dctr, 0, sizeof(uint32_t) * rounds, b->stream);
  if (e == hipSuccess) e = hipMemsetAsync(dbar, 0, sizeof(uint32_t) * kBarWords, b->stream);
  if (e == hipSuccess) e = hipMemsetAsync(b->samples, 0, ns * sizeof(double), b->stream);
  BatchLaunch L{};
  L.phases = dph;
  L.entries = den;
  L.tile_entry = dtab;
  L.counters = dctr;
  L.bar = dbar;
  L.samples = b->samples;
  L.clocks = b->clocks;
  L.nphases = (int)phases.size();
  L.tile_cells = t.tile_cells;
  L.sample_every = sample_every;
  L.max_samples = b->max_samples;
  L.parity0 = (uint32_t)b->cur;
  L.grab = t.tiles_per_grab > 0 ? t.tiles_per_grab : 1;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (e == hipSuccess) e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  if (e == hipSuccess) e = hipEventRecord(e0, b->stream);
  if (e == hipSuccess) e = launch_batch_run(b->dev, L, t.block, grid, b->stream);
  if (e == hipSuccess) e = hipEventRecord(e1, b->stream);
  if (e == hipSuccess) e = hipEventSynchronize(e1);
  float ms = 0.f;
  if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
  uint32_t fault = 0;
  if (e == hipSuccess)
    e = hipMemcpy(&fault, dbar + kBarFault * kBarStride, sizeof fault, hipMemcpyDeviceToHost);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (dph) (void)hipFree(dph);
  if (den) (void)hipFree(den);
  if (dtab) (void)hipFree(dtab);
  if (dctr) (void)hipFree(dctr);
  if (dbar) (void)hipFree(dbar);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return bfail(CFD_E_HIP, "cfd_batch_run: %s", hipGetErrorString(e));
  }
  if (fault) return bfail(CFD_E_HIP, "cfd_batch_run: a grid barrier wait timed out");
  b->tuning = t;
  b->cur ^= steps & 1;
  if (ms_out) *ms_out = ms;
  return CFD_OK;
}

int cfd_batch_results(cfd_batch* b, double* forces, double* finish_ms, int* n_samples,
                      int* rounds) {
  cfd__set_error("");
  if (!b || !b->clocks) return bfail(CFD_E_STATE, "cfd_batch_results: no run yet");
  if (n_samples) *n_samples = b->max_samples;
  if (rounds) *rounds = b->rounds;
  if (forces && b->max_samples > 0)
    BCK(hipMemcpy(forces, b->samples, sizeof(double) * b->n * b->max_samples * 3,
                  hipMemcpyDeviceToHost));
  if (finish_ms) {
    std::vector<unsigned long long> clk(b->clocks_len);
    BCK(hipMemcpy(clk.data(), b->clocks, clk.size() * sizeof(unsigned long long),
                  hipMemcpyDeviceToHost));
    int dev = 0, khz = 0;
    BCK(hipGetDevice(&dev));
    BCK(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev));
    for (int k = 0; k < b->n; ++k) {
      const uint32_t r = b->finish_round[k];  // clocks[r] = end of round r-1
      finish_ms[k] = (double)(clk[r] - clk[0]) / (double)khz;
    }
  }
  return CFD_OK;
}

int cfd_batch_get_populations(cfd_batch* b, int k, float* out, size_t n) {
  cfd__set_error("");
  if (!b || !out || k < 0 || k >= b->n)
    return bfail(CFD_E_ARG, "cfd_batch_get_populations: bad arguments");
  const BatchTunnel& h = b->host[k];
  if (n != (size_t)Q * h.g.N)
    return bfail(CFD_E_ARG, "cfd_batch_get_populations: %zu values, expected %zu", n,
                 (size_t)Q * h.g.N);
  const void* src = h.pop[b->cur];
  float* tmp = nullptr;
  BCK(hipMalloc(&tmp, n * sizeof(float)));
  hipError_t e = launch_store_to_f32(Store::F32, src, tmp, h.g, b->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(out, tmp, n * sizeof(float), hipMemcpyDeviceToHost,
                                          b->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
  (void)hipFree(tmp);
  if (e != hipSuccess) return bfail(CFD_E_HIP, "cfd_batch_get_populations: %s",
                                    hipGetErrorString(e));
  return CFD_OK;
}

int cfd_batch_set_populations(cfd_batch* b, int k, const float* in, size_t n) {
  cfd__set_error("");
  if (!b || !in || k < 0 || k >= b->n)
    return bfail(CFD_E_ARG, "cfd_batch_set_populations: bad arguments");
  const BatchTunnel& h = b->host[k];
  if (n != (size_t)Q * h.g.N)
    return bfail(CFD_E_ARG, "cfd_batch_set_populations: %zu values, expected %zu", n,
                 (size_t)Q * h.g.N);
  float* tmp = nullptr;
  BCK(hipMalloc(&tmp, n * sizeof(float)));
  hipError_t e = hipMemcpyAsync(tmp, in, n * sizeof(float), hipMemcpyHostToDevice, b->stream);
  if (e == hipSuccess) e = launch_f32_to_store(Store::F32, tmp, h.pop[b->cur], h.g, b->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
  (void)hipFree(tmp);
  if (e != hipSuccess) return bfail(CFD_E_HIP, "cfd_batch_set_populations: %s",
                                    hipGetErrorString(e));
  return CFD_OK;
}

int cfd_batch_max_grid(int block, int* workgroups_per_cu, int* compute_units) {
  cfd__set_error("");
  if (!workgroups_per_cu || !compute_units)
    return bfail(CFD_E_ARG, "cfd_batch_max_grid: null argument");
  BCK(batch_max_grid(block, workgroups_per_cu, compute_units));
  return CFD_OK;
}

void cfd_batch_destroy(cfd_batch* b) { free_batch(b); }

}  // extern "C"
