// Context and C-ABI of the gfx950 time-stepping backend (include/insem_hip.h).
//
// Device layout (chosen for MI355X, not the reference's CPU layout):
//  * Global points of each domain are renumbered: points with no boundary,
//    coupling, source or receiver work come first ("regular"), the rest
//    ("special") after them.  The bulk tail kernel then runs branch-free;
//    a small kernel handles the special points.  The numbering is internal:
//    every field crossing the ABI is in the reference numbering.
//  * Element-local forces f_local are [block position][component][k] with
//    the blocks of a domain concatenated, so each force launch writes one
//    contiguous range; the transpose CSR stores, per global point and in the
//    reference's ascending slot order, the f_local address of each
//    contribution.
//  * Per-point block data (metric, coefficients, wr, hoop, r) are stored
//    component-major; components that are bitwise constant over every
//    element are stored once per element.
//  * The predictor of step n+1 is fused into the tail of step n (except on
//    the last step of an advance call), so one step is: element forces ->
//    tail (assembly, M^-1, boundary terms, corrector, predictor) per domain,
//    then the receivers.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <functional>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../../include/insem_hip.h"
#include "device.hpp"
#include "launch.hpp"

using namespace insem_hip;

namespace {

thread_local std::string g_error;

// Graph captures, allocations and frees are serialised across the threads
// of the process: HIP may refuse an allocation or free that synchronises
// with a stream another thread is capturing.  Captures are short (one chunk
// of launches), so the lock is rarely contended.
std::mutex g_capture_mu;

int set_error(int code, const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_error = buf;
  return code;
}

struct HipFailure {
  int code;
};

#define HIP_TRY(expr)                                                                                  \
  do {                                                                                                 \
    hipError_t e_ = (expr);                                                                            \
    if (e_ != hipSuccess) {                                                                            \
      set_error(e_ == hipErrorOutOfMemory ? INSEM_HIP_ENOMEM : INSEM_HIP_EDEVICE, "%s failed: %s", #expr, \
                hipGetErrorString(e_));                                                                \
      throw HipFailure{e_ == hipErrorOutOfMemory ? INSEM_HIP_ENOMEM : INSEM_HIP_EDEVICE};              \
    }                                                                                                  \
  } while (0)

// Device blocks of destroyed contexts are kept per device and handed to
// later contexts instead of going back through hipFree, which waits for
// every stream of the device: a context destroyed while other contexts have
// steps queued (a finished batch, a power iteration's context) would stall
// its thread until all of them finish.  A cached block serves a request of
// at least 2/3 of its size; a device's cache holds at most kCacheCap bytes
// (the largest blocks go back to the device beyond that) and is emptied
// when an allocation fails for want of memory.  Guarded by g_capture_mu.
constexpr int kCacheDevices = 64;
constexpr size_t kCacheCap = size_t{32} << 30;

struct BlockCache {
  std::multimap<size_t, void*> blocks;  // size -> block
  size_t bytes = 0;
};
BlockCache g_cache[kCacheDevices];

// The device's cache (devices beyond kCacheDevices allocate and free directly).
BlockCache* cache_of(int device) { return device >= 0 && device < kCacheDevices ? &g_cache[device] : nullptr; }

// Blocks reserved for staged contexts (insem_hip_reserve): address ->
// device, size and the staged context laid out in it (null until
// insem_hip_stage claims it).  Guarded by g_capture_mu.
struct Reservation {
  int device = 0;
  size_t size = 0;
  const void* owner = nullptr;
};
std::map<void*, Reservation> g_reserved;

// Allocation granularity: 4 KiB up to 1 MiB, 2 MiB above.
size_t block_bytes(size_t bytes) {
  const size_t g = bytes <= (size_t{1} << 20) ? size_t{4} << 10 : size_t{2} << 20;
  return (bytes + g - 1) / g * g;
}

// Frees every cached block of the device (waits for the device's streams).
// Caller holds g_capture_mu.
void release_cache_locked(int device) {
  BlockCache* bc = cache_of(device);
  if (!bc) return;
  for (auto& b : bc->blocks) (void)hipFree(b.second);
  bc->blocks.clear();
  bc->bytes = 0;
}

// A block of at least `bytes` on the current device (`device`); its size in
// *size.  *reused tells whether it held another context's data.
void* take_block(int device, size_t bytes, size_t* size, bool* reused) {
  std::lock_guard<std::mutex> lock(g_capture_mu);
  BlockCache* bc = cache_of(device);
  const size_t want = block_bytes(bytes);
  if (bc) {
    auto it = bc->blocks.lower_bound(want);
    if (it != bc->blocks.end() && it->first <= want + want / 2) {
      void* p = it->second;
      *size = it->first;
      bc->bytes -= it->first;
      bc->blocks.erase(it);
      *reused = true;
      return p;
    }
  }
  void* p = nullptr;
  hipError_t e = hipMalloc(&p, want);
  if (e == hipErrorOutOfMemory && bc && !bc->blocks.empty()) {
    (void)hipGetLastError();
    release_cache_locked(device);
    e = hipMalloc(&p, want);
  }
  HIP_TRY(e);
  *size = want;
  *reused = false;
  return p;
}

// Returns a block whose last use on the device has finished.
void give_block(int device, void* p, size_t size) {
  std::lock_guard<std::mutex> lock(g_capture_mu);
  BlockCache* bc = cache_of(device);
  if (!bc || size > kCacheCap) {
    (void)hipFree(p);
    return;
  }
  while (bc->bytes + size > kCacheCap && !bc->blocks.empty()) {
    auto last = std::prev(bc->blocks.end());
    (void)hipFree(last->second);
    bc->bytes -= last->first;
    bc->blocks.erase(last);
  }
  bc->blocks.emplace(size, p);
  bc->bytes += size;
}

struct BlockRT {
  int phys = 0;
  bool rect = false;
  int nb = 0;
  int pos0 = 0;
  int64_t n_owned = 0;  // slots whose point the fused launch updates (owned points, batch-local leaders)
  int64_t n_local = 0;  // other slots of batch-local points (force kept in LDS)
  int64_t n_read = 0;   // distinct points the launch gathers
  ForceArgs args;
  int32_t* d_elem = nullptr;
  std::vector<int32_t> h_emodel;  // batch: model of each block element
  // batch: elements a time-step launch covers at each level (the leading
  // force batches that hold work at the level's steps; see insem_hip_ctx::floors)
  std::vector<int> nb_at;
};

struct DomainRT {
  int medium = 0, ncomp = 1, am = 0, npe = 1;
  int64_t ng = 0, n_int = 0, n_reg = 0, n_sp = 0, ne = 0;
  int64_t contrib_len = 0;  // floats of the group rows
  std::vector<int32_t> perm_new, perm_old;
  std::vector<uint8_t> h_dir_mask;  // per old point: 1 Dirichlet, 2 axis (host copy for set_state)
  std::vector<int32_t> h_pmodel;    // batch: model of each point (new numbering)
  int axis_mask = 0;
  DomDev dd;
  SpecialDev sp;
  std::vector<BlockRT> blocks;
  float* tmp = nullptr;  // [ncomp][ng] work field (seams, power iteration)
  float* fl_tmp = nullptr;
  float* contrib = nullptr;  // group rows of the grouped points' forces
  int32_t* d_dir = nullptr;
  int n_dir = 0;
  int32_t* d_axis = nullptr;
  int n_axis = 0;
  // power-iteration coupling (reference order)
  int32_t *cpl_pts = nullptr, *cpl_other = nullptr;
  float *cpl_cx = nullptr, *cpl_cz = nullptr;
  int n_cpl = 0;
  int cpl_role = 0;
};

// The longest captured chunk (graph keys hold k in 14 bits).
constexpr int kMaxChunk = 1024;

struct TraceCol {
  int receiver, quantity, raw_ncomp, col, m1_elastic;
};

}  // namespace

struct insem_hip_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::vector<std::pair<void*, size_t>> allocs;  // device blocks and their sizes
  int64_t footprint = 0;
  // Staged set-up (insem_hip_stage_size / insem_hip_stage): no device call
  // is made; the device arrays are laid end to end (256-byte aligned) in one
  // reserved block at stage_base (a stand-in address while only measuring),
  // and every upload lands in the host image stage_host at the same offset
  // for insem_hip_realize to copy.  0: a device context; 1: measuring; 2:
  // staging into a reserved block.
  int staging = 0;
  uintptr_t stage_base = 0;
  size_t stage_cap = 0, stage_used = 0;
  std::vector<unsigned char> stage_host;
  int ngll = 0;
  float *D = nullptr, *Dj = nullptr, *w2 = nullptr;
  std::vector<DomainRT> doms;
  int solid = -1, fluid = -1;
  int n_sources = 0;
  // receivers
  ReceiverDev rec;
  std::vector<TraceCol> cols;
  int has_pv = 0;
  int n_traces = 0;
  // time
  double dt = 0.0;
  int n_steps = 0, step = 0, initialized = 0, have_time = 0;
  float* amps = nullptr;
  int record_every = 1, energy_every = 0;
  int max_samples = 0;
  double* traces = nullptr;
  // step clock and captured step chunks
  int32_t* d_clock = nullptr;  // steps done, on the device
  int chunk = 32;
  int chunk_set = 0;  // steps per chunk fixed by the caller (else the tuned table's, or 32)
  int use_graphs = 1;
  int fuse = 1;  // owned points updated inside the force launches
  int pair = 1;  // a step's elastic and acoustic force work in one launch
  int local = 1;  // points whose slots share one batch are updated in the force launch
  int force_variant = -1;  // autotuner variant of the paired force launch (-1: the tuned table's)
  // batch of independent models (0: a single model)
  int n_models = 0;
  ModelCoef* d_models = nullptr;
  // The models' distinct step counts, largest first.  A chunk whose first
  // step is s runs at level j, the place of the smallest count >= s: every
  // force batch past the last one holding a model of count >= floors[j] is
  // idle for the whole chunk, so the chunk's force launches leave them out.
  std::vector<int32_t> floors;
  int level = 0;  // of the chunk being launched or captured
  // power iteration in flight: runs started, host start vectors kept alive
  int lam_runs = 0;
  std::vector<std::vector<float>> lam_host;
  std::vector<hipGraphExec_t> lam_execs;
  std::map<int, hipGraphExec_t> graphs;
  // energy
  double* partial = nullptr;
  int partial_rows = 0;
  double* energy_rows = nullptr;  // [row][2]: sum u.f, kinetic
  int max_energy_rows = 0;
  double* red_out = nullptr;  // small reduction outputs
  // timing
  int timing = 0;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> ev_force, ev_update;
  std::vector<int> ev_force_tag;           // domain * 64 + block of each force event pair
  std::map<int, std::pair<double, int>> block_ms;
  int cur_tag = 0;
  std::vector<hipEvent_t> ev_pool;
  // progress of queued advance calls: (event recorded after the call's
  // launches, the step count it reaches), oldest first; the steps the device
  // has completed as far as finished events show
  std::deque<std::pair<hipEvent_t, int>> progress;
  std::vector<hipEvent_t> progress_pool;
  int device_step = 0;
  bool device_started = false;    // an advance call has finished on the device
  hipStream_t copy_stream = nullptr;  // reads of finished trace rows (insem_hip_read_traces_done)
  double force_ms = 0.0, update_ms = 0.0;
  int force_n = 0, update_n = 0;

  void* stage_alloc(size_t bytes) {
    const size_t off = (stage_used + 255) / 256 * 256;
    stage_used = off + bytes;
    footprint += static_cast<int64_t>(bytes);
    if (staging == 2 && stage_used > stage_cap) {
      set_error(INSEM_HIP_EINVAL, "the staged set-up needs more than the %zu bytes reserved for it", stage_cap);
      throw HipFailure{INSEM_HIP_EINVAL};
    }
    constexpr uintptr_t kMeasureBase = uintptr_t{1} << 20;  // never dereferenced
    return reinterpret_cast<void*>((staging == 2 ? stage_base : kMeasureBase) + off);
  }
  // n values on the context's device; a block that held another
  // context's data is cleared first unless `clear` is false (the caller
  // writes all n values before any kernel reads them).
  template <class T>
  T* alloc(size_t n, bool clear = true) {
    if (n == 0) n = 1;
    if (staging) return static_cast<T*>(stage_alloc(n * sizeof(T)));  // the host image starts zeroed
    size_t size = 0;
    bool reused = false;
    void* p = take_block(device, n * sizeof(T), &size, &reused);
    allocs.emplace_back(p, size);
    footprint += static_cast<int64_t>(n * sizeof(T));
    if (reused && clear) zero(p, size);
    return static_cast<T*>(p);
  }
  // Copies and fills go through the context's own stream and wait for it:
  // the legacy stream would wait for every blocking stream of the process,
  // and HIP refuses it while another thread captures a graph (contexts are
  // set up on several threads at once).
  void h2d(void* d, const void* src, size_t bytes) {
    if (!bytes) return;
    if (staging) {
      if (staging == 2) std::memcpy(stage_host.data() + (reinterpret_cast<uintptr_t>(d) - stage_base), src, bytes);
      return;
    }
    HIP_TRY(hipMemcpyAsync(d, src, bytes, hipMemcpyHostToDevice, stream));
    HIP_TRY(hipStreamSynchronize(stream));
  }
  void d2h(void* d, const void* src, size_t bytes) {
    if (!bytes) return;
    HIP_TRY(hipMemcpyAsync(d, src, bytes, hipMemcpyDeviceToHost, stream));
    HIP_TRY(hipStreamSynchronize(stream));
  }
  void zero(void* d, size_t bytes) {
    if (staging) {
      if (staging == 2) std::memset(stage_host.data() + (reinterpret_cast<uintptr_t>(d) - stage_base), 0, bytes);
      return;
    }
    HIP_TRY(hipMemsetAsync(d, 0, bytes, stream));
  }
  template <class T>
  T* upload(const T* src, size_t n) {
    T* d = alloc<T>(n, n == 0);
    if (n) h2d(d, src, n * sizeof(T));
    return d;
  }
  template <class T>
  T* upload(const std::vector<T>& v) {
    return upload(v.data(), v.size());
  }
  hipEvent_t event() {
    if (!ev_pool.empty()) {
      hipEvent_t e = ev_pool.back();
      ev_pool.pop_back();
      return e;
    }
    hipEvent_t e;
    HIP_TRY(hipEventCreate(&e));
    return e;
  }
};

namespace {

bool bits_equal(float a, float b) {
  uint32_t x, y;
  std::memcpy(&x, &a, 4);
  std::memcpy(&y, &b, 4);
  return x == y;
}

// Uploads one per-point quantity get(b, k) of a block, per element when it is
// bitwise constant over every element.
template <class Get>
PF upload_pf(insem_hip_ctx* c, int nb, int npe, Get get) {
  bool per_elem = true;
  for (int b = 0; b < nb && per_elem; ++b) {
    const float v0 = get(b, 0);
    for (int k = 1; k < npe; ++k)
      if (!bits_equal(get(b, k), v0)) {
        per_elem = false;
        break;
      }
  }
  PF pf;
  if (per_elem) {
    std::vector<float> h(static_cast<size_t>(nb));
    for (int b = 0; b < nb; ++b) h[b] = get(b, 0);
    pf.p = c->upload(h);
    pf.sb = 1;
    pf.sk = 0;
  } else {
    std::vector<float> h(static_cast<size_t>(nb) * npe);
    for (int b = 0; b < nb; ++b)
      for (int k = 0; k < npe; ++k) h[static_cast<size_t>(b) * npe + k] = get(b, k);
    pf.p = c->upload(h);
    pf.sb = npe;
    pf.sk = 1;
  }
  return pf;
}

int phys_of(int medium, int kind, int am) {
  const bool el = medium == INSEM_MEDIUM_ELASTIC;
  if (kind == INSEM_KIND_AXISYM) {
    if (am == 1) return el ? kElasticAxi1 : kAcousticAxi1;
    return el ? kElasticAxi0 : kAcousticAxi0;
  }
  return el ? kElasticCart : kAcousticCart;
}

void free_all(insem_hip_ctx* c) {
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  for (auto& g : c->graphs) (void)hipGraphExecDestroy(g.second);
  c->graphs.clear();
  for (hipGraphExec_t g : c->lam_execs) (void)hipGraphExecDestroy(g);
  c->lam_execs.clear();
  // the context's stream is idle: its blocks can serve other contexts
  for (auto& b : c->allocs) give_block(c->device, b.first, b.second);
  c->allocs.clear();
  for (auto& e : c->ev_force) {
    (void)hipEventDestroy(e.first);
    (void)hipEventDestroy(e.second);
  }
  for (auto& e : c->ev_update) {
    (void)hipEventDestroy(e.first);
    (void)hipEventDestroy(e.second);
  }
  for (hipEvent_t e : c->ev_pool) (void)hipEventDestroy(e);
  for (auto& e : c->progress) (void)hipEventDestroy(e.first);
  c->progress.clear();
  for (hipEvent_t e : c->progress_pool) (void)hipEventDestroy(e);
  c->progress_pool.clear();
  if (c->copy_stream) (void)hipStreamDestroy(c->copy_stream);
  if (c->stream) (void)hipStreamDestroy(c->stream);
}

int validate_setup(const insem_hip_setup* s) {
  if (!s) return set_error(INSEM_HIP_EINVAL, "null setup");
  if (s->ngll < 2 || s->ngll > 9) return set_error(INSEM_HIP_EINVAL, "no kernel for %d GLL points (supported: 2..9)", s->ngll);
  if (s->n_domains < 0 || s->n_domains > 2) return set_error(INSEM_HIP_EINVAL, "a model has one or two domains");
  if (s->record_every < 1) return set_error(INSEM_HIP_EINVAL, "record_every must be >= 1");
  if (s->energy_every < 0) return set_error(INSEM_HIP_EINVAL, "energy_every must be >= 0");
  int seen[2] = {0, 0};
  for (int di = 0; di < s->n_domains; ++di) {
    const insem_hip_domain& d = s->domains[di];
    if (d.medium != INSEM_MEDIUM_ACOUSTIC && d.medium != INSEM_MEDIUM_ELASTIC)
      return set_error(INSEM_HIP_EINVAL, "domain %d has an unknown medium", di);
    if (seen[d.medium]++) return set_error(INSEM_HIP_EINVAL, "two domains of the same medium");
    const int want = d.medium == INSEM_MEDIUM_ELASTIC ? (d.azimuthal_order == 1 ? 3 : 2) : 1;
    if (d.ncomp != want) return set_error(INSEM_HIP_EINVAL, "domain %d has %d components, expected %d", di, d.ncomp, want);
    // The kernels address fields and element arrays with 32-bit offsets.
    const int64_t npe = static_cast<int64_t>(s->ngll) * s->ngll;
    if (static_cast<int64_t>(d.ncomp) * d.n_global >= kMaxIndex32 || d.ncomp * npe * d.n_elements >= kMaxIndex32)
      return set_error(INSEM_HIP_EINVAL,
                       "domain %d is too large for one context (%lld points, %lld elements; at most %lld "
                       "values per field and per element array)",
                       di, static_cast<long long>(d.n_global), static_cast<long long>(d.n_elements),
                       static_cast<long long>(kMaxIndex32 - 1));
    int64_t covered = 0;
    for (int b = 0; b < d.n_blocks; ++b) {
      const insem_hip_block& B = d.blocks[b];
      if (!(B.azimuthal_order == 0 || (B.azimuthal_order == 1 && B.kind == INSEM_KIND_AXISYM)))
        return set_error(INSEM_HIP_EINVAL, "no kernel for azimuthal order %d on kind %d elements", B.azimuthal_order,
                         B.kind);
      if (B.kind < 0 || B.kind > 3) return set_error(INSEM_HIP_EINVAL, "unknown geometry kind %d", B.kind);
      if (B.kind == INSEM_KIND_AXISYM && (!B.wr || !B.hoop || !B.r || !B.axial))
        return set_error(INSEM_HIP_EINVAL, "axisymmetric block without wr/hoop/r/axial");
      for (int e = 0; e < B.n_elements; ++e)
        if (B.elements[e] < 0 || B.elements[e] >= d.n_elements)
          return set_error(INSEM_HIP_EINVAL, "block element index out of range");
      covered += B.n_elements;
    }
    if (covered != d.n_elements) return set_error(INSEM_HIP_EINVAL, "blocks of domain %d do not cover its elements", di);
  }
  for (int i = 0; i < s->n_sources; ++i)
    if (s->sources[i].domain < 0 || s->sources[i].domain >= s->n_domains)
      return set_error(INSEM_HIP_EINVAL, "source %d has an invalid domain", i);
  for (int i = 0; i < s->n_receivers; ++i)
    if (s->receivers[i].domain < 0 || s->receivers[i].domain >= s->n_domains)
      return set_error(INSEM_HIP_EINVAL, "receiver %d has an invalid domain", i);
  return INSEM_HIP_OK;
}

void build_domain(insem_hip_ctx* c, const insem_hip_setup* s, int di) {
  const insem_hip_domain& d = s->domains[di];
  DomainRT& R = c->doms[di];
  const int npe = s->ngll * s->ngll;
  const int C = d.ncomp;
  R.medium = d.medium;
  R.ncomp = C;
  R.am = d.azimuthal_order;
  R.npe = npe;
  R.ng = d.n_global;
  R.ne = d.n_elements;

  // Special points of this domain.
  std::vector<uint8_t> special(static_cast<size_t>(d.n_global), 0);
  auto mark = [&](const int32_t* p, int n) {
    for (int i = 0; i < n; ++i) special[p[i]] = 1;
  };
  mark(d.absorbing_points, d.n_absorbing);
  mark(d.dirichlet_points, d.n_dirichlet);
  mark(d.axis_points, d.n_axis);
  for (int i = 0; i < s->n_sources; ++i)
    if (s->sources[i].domain == di) mark(s->sources[i].points, s->sources[i].n_points);
  for (int i = 0; i < s->n_receivers; ++i)
    if (s->receivers[i].domain == di) mark(s->receivers[i].points, s->receivers[i].n_points);
  const bool is_solid = di == c->solid, is_fluid = di == c->fluid;
  if (s->coupling.n_points > 0) {
    if (is_solid) mark(s->coupling.solid_points, s->coupling.n_points);
    if (is_fluid) mark(s->coupling.fluid_points, s->coupling.n_points);
  }
  // Launch batches: the element at position p of block b belongs to batch
  // base_b + p / epb (one work-group's share of a force launch).
  const int epb = force_batch_elements(s->ngll);
  std::vector<int32_t> batch_of(static_cast<size_t>(d.n_elements), 0);
  int32_t n_batches = 0;
  for (int b = 0; b < d.n_blocks; ++b) {
    for (int p = 0; p < d.blocks[b].n_elements; ++p) batch_of[d.blocks[b].elements[p]] = n_batches + p / epb;
    n_batches += (d.blocks[b].n_elements + epb - 1) / epb;
  }
  // Point classes: 0 updated inside the force launch: owned (one
  // unit-weight slot of its own) or batch-local (2 .. kMaxGroups + 1
  // unit-weight slots of its own, all in one batch: the first slot's
  // work-item sums them); 1 grouped (2 .. kMaxGroups + 1 unit-weight slots of
  // its own in several batches: the slots write their forces into the group
  // layout); 2 other regular points (assembled from f_local through the
  // transpose CSR); 3 special.  Classes 0 and 1 are never special.
  std::vector<uint8_t> cls(static_cast<size_t>(d.n_global), 2);
  std::vector<uint8_t> cnt(static_cast<size_t>(d.n_global), 0);
  std::vector<uint8_t> local(static_cast<size_t>(d.n_global), 0);
  for (int32_t g = 0; g < d.n_global; ++g) {
    if (special[g]) {
      cls[g] = 3;
      continue;
    }
    const int32_t n = d.t_offsets[g + 1] - d.t_offsets[g];
    bool own = n >= 1 && n <= kMaxGroups + 1;
    for (int32_t t = d.t_offsets[g]; own && t < d.t_offsets[g + 1]; ++t)
      own = d.l2g[d.t_slots[t]] == g && (!d.t_weights || static_cast<float>(d.t_weights[t]) == 1.0f);
    if (own) {
      bool one_batch = c->local && n >= 2;
      const int32_t b0 = batch_of[d.t_slots[d.t_offsets[g]] / npe];
      for (int32_t t = d.t_offsets[g]; one_batch && t < d.t_offsets[g + 1]; ++t)
        one_batch = batch_of[d.t_slots[t] / npe] == b0;
      local[g] = one_batch;
      cls[g] = n == 1 || one_batch ? 0 : 1;
      cnt[g] = static_cast<uint8_t>(n);
    }
  }
  // Within a class, points are numbered in the order the element launches
  // first touch them (blocks, elements, slots), so that the fused update's
  // accesses and the tails' reads walk memory in step with the elements;
  // grouped points are ordered by contribution count first.
  std::vector<int32_t> touch;
  touch.reserve(d.n_global);
  {
    std::vector<uint8_t> seen(static_cast<size_t>(d.n_global), 0);
    for (int b = 0; b < d.n_blocks; ++b)
      for (int e = 0; e < d.blocks[b].n_elements; ++e) {
        const int32_t* row = d.l2g + static_cast<size_t>(d.blocks[b].elements[e]) * npe;
        for (int k = 0; k < npe; ++k)
          if (row[k] >= 0 && !seen[row[k]]) {
            seen[row[k]] = 1;
            touch.push_back(row[k]);
          }
      }
    for (int32_t g = 0; g < d.n_global; ++g)
      if (!seen[g]) touch.push_back(g);
  }
  R.perm_new.assign(d.n_global, -1);
  R.perm_old.clear();
  R.perm_old.reserve(d.n_global);
  auto take = [&](int32_t g) {
    R.perm_new[g] = static_cast<int32_t>(R.perm_old.size());
    R.perm_old.push_back(g);
  };
  for (int32_t g : touch)
    if (cls[g] == 0) take(g);
  R.n_int = static_cast<int64_t>(R.perm_old.size());
  Groups& G = R.dd.grp;
  G = Groups();
  G.begin = R.n_int;
  for (int n = 2; n <= kMaxGroups + 1; ++n) {
    const int64_t s0 = static_cast<int64_t>(R.perm_old.size());
    for (int32_t g : touch)
      if (cls[g] == 1 && cnt[g] == n) take(g);
    const int64_t sz = static_cast<int64_t>(R.perm_old.size()) - s0;
    if (sz == 0) continue;
    G.count[G.n] = n;
    G.start[G.n] = s0;
    G.size[G.n] = sz;
    G.rows = n;
    ++G.n;
  }
  G.end = static_cast<int64_t>(R.perm_old.size());
  R.contrib_len = static_cast<int64_t>(C) * G.rows * (G.end - G.begin);
  if (R.contrib_len >= kMaxIndex32) {
    set_error(INSEM_HIP_EINVAL, "domain %d: group rows of %lld values exceed the 32-bit index range", di,
              static_cast<long long>(R.contrib_len));
    throw HipFailure{INSEM_HIP_EINVAL};
  }
  for (int32_t g : touch)
    if (cls[g] == 2) take(g);
  for (int32_t g : touch)
    if (cls[g] == 3) take(g);
  R.n_sp = 0;
  for (uint8_t v : cls) R.n_sp += v == 3;
  R.n_reg = R.ng - R.n_sp;

  // Element positions: blocks concatenated.
  std::vector<int32_t> pos(d.n_elements, -1);
  {
    int32_t next = 0;
    for (int b = 0; b < d.n_blocks; ++b)
      for (int e = 0; e < d.blocks[b].n_elements; ++e) pos[d.blocks[b].elements[e]] = next++;
  }

  // Transpose CSR in the new numbering, slot order kept.
  std::vector<int32_t> off(d.n_global + 1, 0), addr;
  std::vector<float> tw;
  addr.reserve(d.t_offsets[d.n_global]);
  for (int64_t gn = 0; gn < R.ng; ++gn) {
    const int32_t go = R.perm_old[gn];
    for (int32_t t = d.t_offsets[go]; t < d.t_offsets[go + 1]; ++t) {
      const int32_t slot = d.t_slots[t];
      const int32_t e = slot / npe, k = slot - e * npe;
      addr.push_back((pos[e] * C) * npe + k);
      if (d.t_weights) tw.push_back(static_cast<float>(d.t_weights[t]));
    }
    off[gn + 1] = static_cast<int32_t>(addr.size());
  }
  // Fields, mass, CSR.
  R.dd.ncomp = C;
  R.dd.npe = npe;
  R.dd.ng = R.ng;
  R.dd.n_int = R.n_int;
  R.dd.n_reg = R.n_reg;
  const size_t len = static_cast<size_t>(C) * R.ng;
  R.dd.u = c->alloc<float>(len);
  R.dd.v = c->alloc<float>(len);
  R.dd.a = c->alloc<float>(len);
  R.dd.f = c->alloc<float>(len);
  R.tmp = c->alloc<float>(len);
  c->zero(R.dd.u, len * sizeof(float));
  c->zero(R.dd.v, len * sizeof(float));
  c->zero(R.dd.a, len * sizeof(float));
  c->zero(R.dd.f, len * sizeof(float));
  std::vector<float> im(R.ng), ms(R.ng);
  for (int64_t gn = 0; gn < R.ng; ++gn) {
    im[gn] = d.inv_mass[R.perm_old[gn]];
    ms[gn] = d.mass[R.perm_old[gn]];
  }
  R.dd.im = c->upload(im);
  R.dd.mass = c->upload(ms);
  R.dd.t_off = c->upload(off);
  R.dd.t_addr = c->upload(addr);
  R.dd.t_w = d.t_weights ? c->upload(tw) : nullptr;
  float* fl = c->alloc<float>(static_cast<size_t>(d.n_elements) * C * npe);
  c->zero(fl, static_cast<size_t>(d.n_elements) * C * npe * sizeof(float));
  R.dd.fl = fl;
  R.fl_tmp = c->alloc<float>(static_cast<size_t>(d.n_elements) * C * npe);
  // Group layout, and the code of each slot of a grouped point (its group
  // and its rank among the point's contributions) or of a batch-local point
  // (leader, count, place in the batch's force buffer).
  R.dd.contrib = R.contrib = c->alloc<float>(static_cast<size_t>(R.contrib_len));
  std::vector<uint16_t> slot_rank(static_cast<size_t>(d.n_elements) * npe, 0);
  std::vector<int32_t> batch_fill(static_cast<size_t>(n_batches), 0);
  for (int32_t g = 0; g < d.n_global; ++g) {
    const int32_t t0 = d.t_offsets[g], n = d.t_offsets[g + 1] - t0;
    if (local[g]) {
      const int32_t bt = batch_of[d.t_slots[t0] / npe];
      const int32_t z0 = batch_fill[bt];
      batch_fill[bt] += n;  // at most epb * npe <= work-group size <= 512 places per batch
      for (int32_t r = 0; r < n; ++r) slot_rank[d.t_slots[t0 + r]] = local_code(r == 0, n, z0 + r);
      continue;
    }
    const int64_t gn = R.perm_new[g];
    if (gn < R.dd.grp.begin || gn >= R.dd.grp.end) continue;
    for (int32_t r = 0; r < n; ++r) slot_rank[d.t_slots[t0 + r]] = slot_code(r);
  }

  // Constraint rows (new numbering, float weights as the reference's gather uses them).
  const int32_t* row_off = nullptr;
  const int32_t* row_dofs = nullptr;
  const float* row_w = nullptr;
  if (d.n_rows > 0) {
    const int32_t nnz = d.row_offsets[d.n_rows];
    std::vector<int32_t> rd(nnz);
    std::vector<float> rw(nnz);
    for (int32_t q = 0; q < nnz; ++q) {
      rd[q] = R.perm_new[d.row_dofs[q]];
      rw[q] = static_cast<float>(d.row_weights[q]);
    }
    row_off = c->upload(d.row_offsets, static_cast<size_t>(d.n_rows) + 1);
    row_dofs = c->upload(rd);
    row_w = c->upload(rw);
  }

  // Blocks.
  for (int bi = 0; bi < d.n_blocks; ++bi) {
    const insem_hip_block& B = d.blocks[bi];
    BlockRT br;
    br.nb = B.n_elements;
    br.pos0 = B.n_elements ? pos[B.elements[0]] : 0;
    br.phys = phys_of(d.medium, B.kind, B.azimuthal_order);
    br.rect = B.kind == INSEM_KIND_RECT;
    const int nb = B.n_elements;
    std::vector<int32_t> l2g(static_cast<size_t>(nb) * npe);
    std::vector<uint8_t> seen(static_cast<size_t>(d.n_global), 0);
    for (int b = 0; b < nb; ++b)
      for (int k = 0; k < npe; ++k) {
        const int32_t m = d.l2g[static_cast<size_t>(B.elements[b]) * npe + k];
        l2g[static_cast<size_t>(b) * npe + k] = m >= 0 ? R.perm_new[m] : m;
        const uint16_t code = slot_rank[static_cast<size_t>(B.elements[b]) * npe + k];
        const bool lead = !(code & kSlotLocal) || (code & kSlotLeader);
        br.n_owned += m >= 0 && R.perm_new[m] < R.n_int && lead;
        br.n_local += (code & kSlotLocal) && !(code & kSlotLeader);
        if (m >= 0) {
          br.n_read += !seen[m];
          seen[m] = 1;
        } else {
          for (int32_t q = d.row_offsets[-1 - m]; q < d.row_offsets[-m]; ++q) {
            br.n_read += !seen[d.row_dofs[q]];
            seen[d.row_dofs[q]] = 1;
          }
        }
      }
    ForceArgs& A = br.args;
    A.nb = nb;
    A.l2g = c->upload(l2g);
    A.row_off = row_off;
    A.row_dofs = row_dofs;
    A.row_w = row_w;
    A.u = R.dd.u;
    A.ng = R.ng;
    A.n_int = R.n_int;
    A.uw = R.dd.u;
    A.v = R.dd.v;
    A.a = R.dd.a;
    A.im = R.dd.im;
    A.mass = R.dd.mass;
    A.contrib = R.contrib;
    A.gbegin = static_cast<int>(R.dd.grp.begin);
    A.gstride = static_cast<int>(R.dd.grp.end - R.dd.grp.begin);
    A.grows = R.dd.grp.rows;
    {
      std::vector<uint16_t> rk(static_cast<size_t>(nb) * npe);
      for (int b = 0; b < nb; ++b)
        for (int k = 0; k < npe; ++k) {
          rk[static_cast<size_t>(b) * npe + k] = slot_rank[static_cast<size_t>(B.elements[b]) * npe + k];
          A.has_local = A.has_local || (rk[static_cast<size_t>(b) * npe + k] & kSlotLocal);
        }
      A.code = c->upload(rk);
    }
    br.d_elem = c->upload(B.elements, static_cast<size_t>(nb));
    A.elem = br.d_elem;
    A.fl = fl;
    A.out_pos0 = br.pos0;
    A.D = c->D;
    A.Dj = c->Dj;
    A.w2 = c->w2;
    const float* M = B.metric;
    if (B.kind == INSEM_KIND_RECT) {
      A.m[0] = upload_pf(c, nb, 1, [&](int b, int) { return M[3 * b]; });
      A.m[3] = upload_pf(c, nb, 1, [&](int b, int) { return M[3 * b + 1]; });
      A.m[4] = upload_pf(c, nb, 1, [&](int b, int) { return M[3 * b + 2]; });
      for (int q : {0, 3, 4}) A.m[q].sk = 0;
    } else if (B.kind == INSEM_KIND_AFFINE) {
      for (int q = 0; q < 5; ++q) {
        A.m[q] = upload_pf(c, nb, 1, [&](int b, int) { return M[5 * b + q]; });
        A.m[q].sk = 0;
      }
    } else {
      const int ncm = B.kind == INSEM_KIND_AXISYM ? 4 : 5;  // axisymmetric kernels use wr, not jac
      for (int q = 0; q < ncm; ++q)
        A.m[q] = upload_pf(c, nb, npe, [&](int b, int k) { return M[(static_cast<size_t>(b) * npe + k) * 5 + q]; });
    }
    A.c0 = upload_pf(c, nb, npe, [&](int b, int k) { return B.coef0[static_cast<size_t>(b) * npe + k]; });
    if (d.medium == INSEM_MEDIUM_ELASTIC)
      A.c1 = upload_pf(c, nb, npe, [&](int b, int k) { return B.coef1[static_cast<size_t>(b) * npe + k]; });
    if (B.kind == INSEM_KIND_AXISYM) {
      A.wr = upload_pf(c, nb, npe, [&](int b, int k) { return B.wr[static_cast<size_t>(b) * npe + k]; });
      const bool need_hoop = !(d.medium == INSEM_MEDIUM_ACOUSTIC && B.azimuthal_order == 0);
      if (need_hoop) A.hoop = upload_pf(c, nb, npe, [&](int b, int k) { return B.hoop[static_cast<size_t>(b) * npe + k]; });
      A.r = upload_pf(c, nb, npe, [&](int b, int k) { return B.r[static_cast<size_t>(b) * npe + k]; });
      A.axial = c->upload(B.axial, static_cast<size_t>(nb));
    }
    R.blocks.push_back(br);
  }

  // Dirichlet / axis.
  R.h_dir_mask.assign(d.n_global, 0);
  for (int i = 0; i < d.n_dirichlet; ++i) R.h_dir_mask[d.dirichlet_points[i]] |= 1;
  for (int i = 0; i < d.n_axis; ++i) R.h_dir_mask[d.axis_points[i]] |= 2;
  R.axis_mask = 0;
  for (int i = 0; i < d.n_axis_components; ++i) R.axis_mask |= 1 << d.axis_components[i];
  {
    std::vector<int32_t> p(d.n_dirichlet), q(d.n_axis);
    for (int i = 0; i < d.n_dirichlet; ++i) p[i] = R.perm_new[d.dirichlet_points[i]];
    for (int i = 0; i < d.n_axis; ++i) q[i] = R.perm_new[d.axis_points[i]];
    R.d_dir = c->upload(p);
    R.n_dir = d.n_dirichlet;
    R.d_axis = c->upload(q);
    R.n_axis = d.n_axis;
  }

  // Special-point tables.
  const int64_t nsp = R.n_sp;
  auto sidx = [&](int32_t g_old) { return static_cast<int64_t>(R.perm_new[g_old]) - R.n_reg; };
  std::vector<int32_t> abs_idx(nsp, -1), cpl_idx(nsp, -1);
  std::vector<uint8_t> flags(nsp, 0);
  for (int i = 0; i < d.n_absorbing; ++i) abs_idx[sidx(d.absorbing_points[i])] = i;
  for (int i = 0; i < d.n_dirichlet; ++i) flags[sidx(d.dirichlet_points[i])] |= 1;
  for (int i = 0; i < d.n_axis; ++i) flags[sidx(d.axis_points[i])] |= 2;
  std::vector<std::vector<std::pair<int32_t, std::vector<float>>>> src(nsp);
  for (int si = 0; si < s->n_sources; ++si) {
    const insem_hip_source& S = s->sources[si];
    if (S.domain != di) continue;
    for (int i = 0; i < S.n_points; ++i) {
      std::vector<float> cf(C);
      for (int cc = 0; cc < C; ++cc) cf[cc] = S.coef[static_cast<size_t>(cc) * S.n_points + i];
      src[sidx(S.points[i])].push_back({si, cf});
    }
  }
  std::vector<int32_t> soff(nsp + 1, 0), sid;
  std::vector<float> scoef;
  for (int64_t i = 0; i < nsp; ++i) {
    for (auto& e : src[i]) {
      sid.push_back(e.first);
      scoef.insert(scoef.end(), e.second.begin(), e.second.end());
    }
    soff[i + 1] = static_cast<int32_t>(sid.size());
  }
  SpecialDev& sp = R.sp;
  sp.n_sp = nsp;
  sp.abs_idx = c->upload(abs_idx);
  sp.abs_coef = c->upload(d.absorbing_coef, static_cast<size_t>(d.n_absorbing) * C * C);
  sp.flags = c->upload(flags);
  sp.axis_mask = R.axis_mask;
  sp.src_off = c->upload(soff);
  sp.src_id = c->upload(sid);
  sp.src_coef = c->upload(scoef);
  sp.pm = R.am == 1;
  sp.su = c->alloc<float>(static_cast<size_t>(C) * nsp);
  sp.sv = c->alloc<float>(static_cast<size_t>(C) * nsp);
  sp.sa = c->alloc<float>(static_cast<size_t>(C) * nsp);
  sp.cpl_idx = nullptr;
  sp.cpl_role = 0;
  R.n_cpl = 0;
  R.cpl_role = 0;
  (void)is_fluid;
  sp.cpl_idx = c->upload(cpl_idx);  // filled after both domains exist
}

// Coupling tables need the numbering of both domains.
void build_coupling(insem_hip_ctx* c, const insem_hip_setup* s) {
  const insem_hip_coupling& cp = s->coupling;
  if (cp.n_points == 0 || c->solid < 0 || c->fluid < 0) return;
  for (int role = 1; role <= 2; ++role) {
    DomainRT& R = c->doms[role == 1 ? c->solid : c->fluid];
    DomainRT& O = c->doms[role == 1 ? c->fluid : c->solid];
    const int32_t* mine = role == 1 ? cp.solid_points : cp.fluid_points;
    const int32_t* theirs = role == 1 ? cp.fluid_points : cp.solid_points;
    std::vector<int32_t> idx(R.n_sp, -1), pts(cp.n_points), other(cp.n_points);
    for (int j = 0; j < cp.n_points; ++j) {
      const int64_t si = static_cast<int64_t>(R.perm_new[mine[j]]) - R.n_reg;
      if (idx[si] >= 0) {
        set_error(INSEM_HIP_EINVAL, "coupling maps two interface entries to one %s point",
                  role == 1 ? "solid" : "fluid");
        throw HipFailure{INSEM_HIP_EINVAL};
      }
      idx[si] = j;
      pts[j] = R.perm_new[mine[j]];
      other[j] = O.perm_new[theirs[j]];
    }
    R.sp.cpl_idx = c->upload(idx);
    R.sp.cpl_other = c->upload(other);
    R.sp.cpl_cx = c->upload(cp.cx, cp.n_points);
    R.sp.cpl_cz = c->upload(cp.cz, cp.n_points);
    R.sp.cpl_role = role;
    R.sp.other = role == 1 ? O.dd.u : O.dd.a;
    R.sp.other_ng = O.ng;
    R.cpl_pts = c->upload(pts);
    R.cpl_other = const_cast<int32_t*>(R.sp.cpl_other);
    R.cpl_cx = const_cast<float*>(R.sp.cpl_cx);
    R.cpl_cz = const_cast<float*>(R.sp.cpl_cz);
    R.n_cpl = cp.n_points;
    R.cpl_role = role;
  }
}

void build_receivers(insem_hip_ctx* c, const insem_hip_setup* s) {
  std::vector<int32_t> dom, kind, nc, col, off{0}, pts;
  std::vector<float> val, ddx, ddz;
  std::vector<double> rho;
  int ncols = 0;
  for (int ri = 0; ri < s->n_receivers; ++ri) {
    const insem_hip_receiver& r = s->receivers[ri];
    const DomainRT& R = c->doms[r.domain];
    for (int q = 0; q < r.n_quantities; ++q) {
      const int k = r.quantities[q];
      const int raw = k == INSEM_Q_PARTICLE_VELOCITY ? 2 : R.ncomp;
      dom.push_back(r.domain);
      kind.push_back(k);
      nc.push_back(raw);
      col.push_back(ncols);
      c->cols.push_back({ri, q, raw, ncols, R.medium == INSEM_MEDIUM_ELASTIC && R.am == 1 ? 1 : 0});
      ncols += raw;
      if (k == INSEM_Q_PARTICLE_VELOCITY) c->has_pv = 1;
      for (int i = 0; i < r.n_points; ++i) {
        pts.push_back(static_cast<int32_t>(R.perm_new[r.points[i]] - R.n_reg));
        val.push_back(r.value[i]);
        ddx.push_back(r.d_dx[i]);
        ddz.push_back(r.d_dz[i]);
      }
      off.push_back(static_cast<int32_t>(pts.size()));
      rho.push_back(r.rho);
    }
  }
  ReceiverDev& rd = c->rec;
  rd.n_tasks = static_cast<int>(dom.size());
  c->n_traces = rd.n_tasks;
  rd.domain = c->upload(dom);
  rd.kind = c->upload(kind);
  rd.ncomp = c->upload(nc);
  rd.col = c->upload(col);
  rd.pt_off = c->upload(off);
  rd.pts = c->upload(pts);
  rd.value = c->upload(val);
  rd.d_dx = c->upload(ddx);
  rd.d_dz = c->upload(ddz);
  rd.rho = c->upload(rho);
  rd.pv = c->alloc<double>(4 * static_cast<size_t>(std::max(1, rd.n_tasks)));
  c->zero(rd.pv, 4 * static_cast<size_t>(std::max(1, rd.n_tasks)) * sizeof(double));
  rd.n_cols = ncols;
  for (size_t di = 0; di < c->doms.size(); ++di) {
    rd.su[di] = c->doms[di].sp.su;
    rd.sv[di] = c->doms[di].sp.sv;
    rd.sa[di] = c->doms[di].sp.sa;
    rd.n_sp[di] = c->doms[di].n_sp;
  }
}

// Model of every point, block element and receiver task of a batch, in
// the device numbering; the coefficient table is filled by set_time_batch.
void build_batch(insem_hip_ctx* c, const insem_hip_setup* s, const insem_hip_batch* b) {
  c->n_models = b->n_models;
  c->d_models = c->alloc<ModelCoef>(static_cast<size_t>(b->n_models));
  c->energy_every = 0;
  for (int di = 0; di < s->n_domains; ++di) {
    DomainRT& R = c->doms[di];
    const insem_hip_domain& d = s->domains[di];
    const int32_t* pm = b->point_model[di];
    std::vector<int32_t> pnew(static_cast<size_t>(R.ng));
    for (int64_t gn = 0; gn < R.ng; ++gn) pnew[gn] = pm[R.perm_old[gn]];
    R.dd.pmodel = c->upload(pnew);
    R.dd.models = c->d_models;
    R.h_pmodel = pnew;
    const int npe = R.npe;
    for (int bi = 0; bi < d.n_blocks; ++bi) {
      const insem_hip_block& B = d.blocks[bi];
      std::vector<int32_t> em(static_cast<size_t>(B.n_elements), 0);
      for (int e = 0; e < B.n_elements; ++e) {
        const int32_t* row = d.l2g + static_cast<size_t>(B.elements[e]) * npe;
        int k = 0;
        while (k < npe && row[k] < 0) ++k;
        em[e] = pm[k < npe ? row[k] : d.row_dofs[d.row_offsets[-1 - row[0]]]];
      }
      BlockRT& br = R.blocks[bi];
      br.args.emodel = c->upload(em);
      br.args.models = c->d_models;
      br.h_emodel = std::move(em);
    }
  }
  std::vector<int32_t> tm;
  for (const TraceCol& tc : c->cols) tm.push_back(b->receiver_model[tc.receiver]);
  c->rec.tmodel = c->upload(tm);
  c->rec.models = c->d_models;
}

int validate_batch(const insem_hip_setup* s, const insem_hip_batch* b) {
  if (b->n_models < 1) return set_error(INSEM_HIP_EINVAL, "a batch needs at least one model");
  for (int di = 0; di < s->n_domains; ++di) {
    if (!b->point_model[di]) return set_error(INSEM_HIP_EINVAL, "batch without the model of each point");
    for (int32_t g = 0; g < s->domains[di].n_global; ++g)
      if (b->point_model[di][g] < 0 || b->point_model[di][g] >= b->n_models)
        return set_error(INSEM_HIP_EINVAL, "point model out of range");
  }
  if (s->n_receivers > 0 && !b->receiver_model) return set_error(INSEM_HIP_EINVAL, "batch without receiver models");
  for (int i = 0; i < s->n_receivers; ++i)
    if (b->receiver_model[i] < 0 || b->receiver_model[i] >= b->n_models)
      return set_error(INSEM_HIP_EINVAL, "receiver model out of range");
  return INSEM_HIP_OK;
}

// ------------------------------------------------------------ step pieces

void timed_begin(insem_hip_ctx* c, hipEvent_t* e) {
  if (!c->timing) return;
  *e = c->event();
  HIP_TRY(hipEventRecord(*e, c->stream));
}

void timed_end(insem_hip_ctx* c, hipEvent_t b, bool force) {
  if (!c->timing) return;
  hipEvent_t e = c->event();
  HIP_TRY(hipEventRecord(e, c->stream));
  (force ? c->ev_force : c->ev_update).push_back({b, e});
  if (force) c->ev_force_tag.push_back(c->cur_tag);
}

// Parameters of the force launches of a time step (owned points updated,
// grouped forces written in the group rows).
StepParams fused_params(const StepCoef& sc, int corr, int pred, const StepClock& clk, double* partial) {
  StepParams p;
  p.fuse = 1;
  p.corr = corr;
  p.pred = pred;
  p.rdt = sc.rdt;
  p.hdt = sc.hdt;
  p.hdt2 = sc.hdt2;
  p.cc = sc.cc;
  p.clk = clk;
  p.partial = partial;
  return p;
}

// Element forces of every block of R from the field u, one launch per
// block, f_local for every slot.
void run_forces(insem_hip_ctx* c, DomainRT& R, const float* u, hipStream_t stream) {
  const int di = static_cast<int>(&R - c->doms.data());
  const StepParams p;
  for (size_t bi = 0; bi < R.blocks.size(); ++bi) {
    BlockRT& br = R.blocks[bi];
    c->cur_tag = di * 64 + static_cast<int>(bi);
    ForceArgs a = br.args;
    a.u = u;
    hipEvent_t e0 = nullptr;
    timed_begin(c, &e0);
    HIP_TRY(launch_forces(c->ngll, br.phys, br.rect, true, a, p, stream));
    timed_end(c, e0, true);
  }
}

StepCoef step_coef(const insem_hip_ctx* c, bool implicit) {
  StepCoef sc;
  sc.rdt = static_cast<float>(c->dt);
  sc.hdt = static_cast<float>(0.5 * c->dt);
  sc.hdt2 = static_cast<float>(0.5 * c->dt * c->dt);
  sc.h_damp = implicit ? 0.5 * c->dt : 0.0;
  sc.cc = c->dt * c->dt / 8.0;
  return sc;
}

StepClock clock_at(const insem_hip_ctx* c, int offset) {
  StepClock k;
  k.base = c->d_clock;
  k.offset = offset;
  k.energy_every = c->energy_every;
  k.record_every = c->record_every;
  k.max_samples = c->max_samples;
  return k;
}

// A block's force arguments for a time step of the current chunk: on a
// batch, only the leading elements that hold work at the chunk's level.
ForceArgs level_args(const insem_hip_ctx* c, const BlockRT& br) {
  ForceArgs a = br.args;
  if (c->n_models && static_cast<size_t>(c->level) < br.nb_at.size()) a.nb = br.nb_at[c->level];
  return a;
}

// The level of a chunk whose first step is s (0 for a single model).
int level_of(const insem_hip_ctx* c, int s) {
  int j = 0;
  while (j + 1 < static_cast<int>(c->floors.size()) && c->floors[j + 1] >= s) ++j;
  return j;
}

// One compute_acceleration + corrector (+ predictor) over both domains at
// step *d_clock + offset.  Energy partials are written on energy steps only
// (the kernels read the step on the device).
//
// In a time step the element forces of both domains go first (owned points
// updated in place), then the tail of the solid and that of the fluid: the
// fluid's coupling term reads the solid's new acceleration
// (solver.cpp:693-727).  Receivers: rec_prev samples the previous step in
// the first force launch's first work-groups (before this step's tails
// rewrite the side buffers they read), rec_now samples this step after its
// tails.
void step_body(insem_hip_ctx* c, int offset, bool implicit, int corr, int pred, int pv_mode, bool rec_prev,
               bool rec_now) {
  const StepCoef sc = step_coef(c, implicit);
  const StepClock clk = clock_at(c, offset);
  hipStream_t s0 = c->stream;
  int row = 0, nb = 0;
  bool rec_taken = !rec_prev || c->rec.n_tasks == 0;
  // The first elastic and the first acoustic block share one launch where a
  // paired kernel exists for their physics.
  BlockRT* pa = nullptr;
  BlockRT* pb = nullptr;
  if (c->fuse && c->pair && c->solid >= 0 && c->fluid >= 0 && !c->doms[c->solid].blocks.empty() &&
      !c->doms[c->fluid].blocks.empty()) {
    pa = &c->doms[c->solid].blocks[0];
    pb = &c->doms[c->fluid].blocks[0];
    if (!forces_pair_supported(pa->phys, pa->rect, pb->phys, pb->rect)) pa = pb = nullptr;
  }
  if (pa) {
    const StepParams p = fused_params(sc, corr, pred, clk, c->partial + 2 * row);
    const bool with_rec = !rec_taken;
    c->cur_tag = c->solid * 64;
    hipEvent_t e0 = nullptr;
    timed_begin(c, &e0);
    int grid = 0;
    HIP_TRY(launch_forces_pair(c->ngll, pa->phys, pa->rect, pb->phys, pb->rect, level_args(c, *pa),
                               level_args(c, *pb), p, s0, &grid, with_rec ? &c->rec : nullptr,
                               with_rec ? c->rec.n_tasks : 0, 0.5 * c->dt, c->force_variant));
    timed_end(c, e0, true);
    rec_taken = true;
    row += grid;
  }
  for (int medium : {INSEM_MEDIUM_ELASTIC, INSEM_MEDIUM_ACOUSTIC}) {
    const int di = medium == INSEM_MEDIUM_ELASTIC ? c->solid : c->fluid;
    if (di < 0) continue;
    DomainRT& R = c->doms[di];
    for (size_t bi = 0; bi < R.blocks.size(); ++bi) {
      BlockRT& br = R.blocks[bi];
      if (&br == pa || &br == pb) continue;
      StepParams p;
      if (c->fuse) p = fused_params(sc, corr, pred, clk, c->partial + 2 * row);
      const bool with_rec = c->fuse && !rec_taken;
      c->cur_tag = di * 64 + static_cast<int>(bi);
      hipEvent_t e0 = nullptr;
      timed_begin(c, &e0);
      int grid = 0;
      HIP_TRY(launch_forces(c->ngll, br.phys, br.rect, true, c->fuse ? level_args(c, br) : br.args, p, s0, &grid,
                            with_rec ? &c->rec : nullptr, with_rec ? c->rec.n_tasks : 0, 0.5 * c->dt));
      timed_end(c, e0, true);
      rec_taken = rec_taken || with_rec;
      if (c->fuse) row += grid;
    }
  }
  if (!rec_taken) {
    StepClock prev = clk;
    prev.offset -= 1;
    HIP_TRY(launch_receivers(c->rec, prev, 1, 0.5 * c->dt, s0));
  }
  if (c->fuse && c->pair && c->solid >= 0 && c->fluid >= 0) {
    DomainRT& S = c->doms[c->solid];
    DomainRT& F = c->doms[c->fluid];
    hipEvent_t e0 = nullptr;
    timed_begin(c, &e0);
    HIP_TRY(launch_tail_pair(S.dd, S.sp, F.dd, F.sp, sc, corr, pred, clk, c->partial + 2 * row, &nb, s0));
    timed_end(c, e0, false);
    row += nb;
  } else {
    for (int medium : {INSEM_MEDIUM_ELASTIC, INSEM_MEDIUM_ACOUSTIC}) {
      const int di = medium == INSEM_MEDIUM_ELASTIC ? c->solid : c->fluid;
      if (di < 0) continue;
      DomainRT& R = c->doms[di];
      hipEvent_t e0 = nullptr;
      timed_begin(c, &e0);
      HIP_TRY(launch_tail(R.dd, R.sp, sc, corr, pred, clk, c->fuse != 0, c->partial + 2 * row, &nb, s0));
      timed_end(c, e0, false);
      row += nb;
    }
  }
  if (c->energy_every > 0) HIP_TRY(launch_energy_row(c->partial, row, clk, c->energy_rows, s0));
  if (rec_now) HIP_TRY(launch_receivers(c->rec, clk, pv_mode, 0.5 * c->dt, s0));
}

void initialize(insem_hip_ctx* c) {
  c->level = level_of(c, 0);
  step_body(c, 0, false, 0, 0, 0, false, true);
  c->initialized = 1;
}

void predictor_all(insem_hip_ctx* c) {
  const StepCoef sc = step_coef(c, true);
  if (!c->n_models && c->doms.size() <= 2) {
    HIP_TRY(launch_predictor_flat(c->doms.empty() ? nullptr : &c->doms[0].dd,
                                  c->doms.size() > 1 ? &c->doms[1].dd : nullptr, sc, c->stream));
    return;
  }
  for (DomainRT& R : c->doms) HIP_TRY(launch_predictor(R.dd, sc, clock_at(c, 0), c->stream));
}

// The launches of k steps after *d_clock: a standalone predictor when the
// chunk opens an advance call, the predictor of each next step fused into
// the previous tail, none after the last step of the call.  The receivers
// of each step but the chunk's last ride on the next step's force launch.
void chunk_launches(insem_hip_ctx* c, int k, bool first, bool last) {
  if (first) predictor_all(c);
  const bool carry = c->fuse && c->rec.n_tasks > 0;
  for (int j = 1; j <= k; ++j)
    step_body(c, j, true, 1, (last && j == k) ? 0 : 1, 1, carry && j > 1, !carry || j == k);
  HIP_TRY(launch_clock_add(c->d_clock, k, c->stream));
}

// Runs a chunk, replaying a captured hipGraph of it when graphs are on (the
// step numbers come from the device clock, so one graph serves every chunk
// of the same shape).  Timing runs launch directly, with events.
hipGraphExec_t chunk_graph(insem_hip_ctx* c, int k, bool first, bool last, int level) {
  const int key = (level << 16) | (k << 2) | (first ? 2 : 0) | (last ? 1 : 0);
  auto it = c->graphs.find(key);
  if (it == c->graphs.end()) {
    c->level = level;
    std::unique_lock<std::mutex> lock(g_capture_mu);
    HIP_TRY(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    hipGraph_t g = nullptr;
    try {
      chunk_launches(c, k, first, last);
    } catch (...) {
      (void)hipStreamEndCapture(c->stream, &g);
      if (g) (void)hipGraphDestroy(g);
      throw;
    }
    HIP_TRY(hipStreamEndCapture(c->stream, &g));
    lock.unlock();
    hipGraphExec_t ge = nullptr;
    const hipError_t e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    HIP_TRY(e);
    it = c->graphs.emplace(key, ge).first;
    // Upload the graph's resources now, so that its first replay only
    // launches.
    HIP_TRY(hipGraphUpload(ge, c->stream));
  }
  return it->second;
}

void run_chunk(insem_hip_ctx* c, int k, bool first, bool last, int level) {
  if (!c->use_graphs || c->timing) {
    c->level = level;
    chunk_launches(c, k, first, last);
    return;
  }
  HIP_TRY(hipGraphLaunch(chunk_graph(c, k, first, last, level), c->stream));
}

// The chunks of an advance of n steps from the current step: (k, first,
// last, level) for each, in order.
template <class F>
void for_each_chunk(const insem_hip_ctx* c, int n, F&& f) {
  const int m = std::max(0, std::min(n, c->n_steps - c->step));
  for (int s = 0; s < m; s += c->chunk) {
    const int k = std::min(c->chunk, m - s);
    f(k, s == 0, s + k == m, level_of(c, c->step + s + 1));
  }
}

void require_time(insem_hip_ctx* c) {
  if (!c->have_time) {
    set_error(INSEM_HIP_ESTATE, "insem_hip_set_time has not been called");
    throw HipFailure{INSEM_HIP_ESTATE};
  }
}

template <class F>
int guarded(insem_hip_ctx* c, F&& f) {
  if (!c) return set_error(INSEM_HIP_EINVAL, "null context");
  if (c->staging) return set_error(INSEM_HIP_ESTATE, "the context is staged: insem_hip_realize it first");
  try {
    if (c->stream) HIP_TRY(hipSetDevice(c->device));
    f();
    return INSEM_HIP_OK;
  } catch (const HipFailure& e) {
    return e.code;
  } catch (const std::exception& e) {
    return set_error(INSEM_HIP_EDEVICE, "%s", e.what());
  }
}

// Queues n repetitions of the launches of f: f captured r times as one
// hipGraph (r the largest divisor of n up to 10: fewer graph launches on the
// host, a graph small enough to instantiate quickly), the graph launched
// n / r times, or f called n times when graphs are off.
template <class F>
void queue_repeated(insem_hip_ctx* c, int n, F&& f) {
  if (!c->use_graphs) {
    for (int k = 0; k < n; ++k) f();
    return;
  }
  int r = std::min(n, 10);
  while (r > 1 && n % r != 0) --r;
  std::unique_lock<std::mutex> lock(g_capture_mu);
  HIP_TRY(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  hipGraph_t g = nullptr;
  try {
    for (int k = 0; k < r; ++k) f();
  } catch (...) {
    (void)hipStreamEndCapture(c->stream, &g);
    if (g) (void)hipGraphDestroy(g);
    throw;
  }
  HIP_TRY(hipStreamEndCapture(c->stream, &g));
  lock.unlock();
  hipGraphExec_t ge = nullptr;
  const hipError_t e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  HIP_TRY(e);
  c->lam_execs.push_back(ge);
  for (int k = 0; k < n / r; ++k) HIP_TRY(hipGraphLaunch(ge, c->stream));
}

void permute_in(const DomainRT& R, const float* src, std::vector<float>& out) {
  out.resize(static_cast<size_t>(R.ncomp) * R.ng);
  for (int cc = 0; cc < R.ncomp; ++cc)
    for (int64_t gn = 0; gn < R.ng; ++gn) out[cc * R.ng + gn] = src[cc * R.ng + R.perm_old[gn]];
}

void permute_out(const DomainRT& R, const std::vector<float>& src, float* out) {
  for (int cc = 0; cc < R.ncomp; ++cc)
    for (int64_t gn = 0; gn < R.ng; ++gn) out[cc * R.ng + R.perm_old[gn]] = src[cc * R.ng + gn];
}

// Advance markers: the host may queue at most this many advance calls ahead
// of the device before an advance waits for the oldest.
constexpr std::size_t kMaxMarkers = 4096;

// Retires the markers of finished advance calls (the device step follows
// them); with wait, also waits for the oldest unfinished one.
void retire_markers(insem_hip_ctx* c, bool wait) {
  while (!c->progress.empty()) {
    hipEvent_t e = c->progress.front().first;
    hipError_t q = hipEventQuery(e);
    if (q == hipErrorNotReady) {
      if (!wait) break;
      q = hipEventSynchronize(e);
      wait = false;
    }
    HIP_TRY(q);
    c->device_step = c->progress.front().second;
    c->device_started = true;
    c->progress_pool.push_back(e);
    c->progress.pop_front();
  }
}

}  // namespace

// ===================================================================== ABI

extern "C" {

const char* insem_hip_last_error(void) { return g_error.c_str(); }

int insem_hip_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

int insem_hip_get_device(int32_t* device) {
  if (!device) return set_error(INSEM_HIP_EINVAL, "null argument");
  int d = 0;
  const hipError_t e = hipGetDevice(&d);
  if (e != hipSuccess) return set_error(INSEM_HIP_EDEVICE, "hipGetDevice failed: %s", hipGetErrorString(e));
  *device = d;
  return INSEM_HIP_OK;
}

int insem_hip_set_device(int32_t device) {
  const int n = insem_hip_device_count();
  if (device < 0 || device >= n) return set_error(INSEM_HIP_EINVAL, "no device %d (%d visible)", device, n);
  const hipError_t e = hipSetDevice(device);
  if (e != hipSuccess) return set_error(INSEM_HIP_EDEVICE, "hipSetDevice failed: %s", hipGetErrorString(e));
  return INSEM_HIP_OK;
}

namespace {
int create_impl(const insem_hip_setup* s, const insem_hip_batch* batch, insem_hip_ctx** out, int staging = 0,
                void* block = nullptr);
}

int insem_hip_create(const insem_hip_setup* s, insem_hip_ctx** out) { return create_impl(s, nullptr, out); }

int insem_hip_create_batch(const insem_hip_setup* s, const insem_hip_batch* batch, insem_hip_ctx** out) {
  if (!batch) return set_error(INSEM_HIP_EINVAL, "null batch");
  return create_impl(s, batch, out);
}

int insem_hip_stage_size(const insem_hip_setup* s, int64_t* bytes) {
  if (!bytes) return set_error(INSEM_HIP_EINVAL, "null argument");
  insem_hip_ctx* c = nullptr;
  const int rc = create_impl(s, nullptr, &c, 1);
  if (rc != INSEM_HIP_OK) return rc;
  *bytes = static_cast<int64_t>(c->stage_used);
  delete c;  // a measured context holds nothing on the device
  return INSEM_HIP_OK;
}

int insem_hip_reserve(int64_t bytes, void** block) {
  if (!block || bytes <= 0) return set_error(INSEM_HIP_EINVAL, "a reservation needs a positive size and an output");
  *block = nullptr;
  try {
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    size_t size = 0;
    bool reused = false;
    void* p = take_block(dev, static_cast<size_t>(bytes), &size, &reused);
    std::lock_guard<std::mutex> lock(g_capture_mu);
    g_reserved[p] = Reservation{dev, size, nullptr};
    *block = p;
    return INSEM_HIP_OK;
  } catch (const HipFailure& e) {
    return e.code;
  }
}

int insem_hip_unreserve(void* block) {
  Reservation r;
  {
    std::lock_guard<std::mutex> lock(g_capture_mu);
    const auto it = g_reserved.find(block);
    if (it == g_reserved.end()) return set_error(INSEM_HIP_EINVAL, "the block is not reserved");
    if (it->second.owner)
      return set_error(INSEM_HIP_ESTATE, "the block holds a staged context: insem_hip_destroy it instead");
    r = it->second;
    g_reserved.erase(it);
  }
  give_block(r.device, block, r.size);
  return INSEM_HIP_OK;
}

int insem_hip_stage(const insem_hip_setup* s, void* block, insem_hip_ctx** out) {
  return create_impl(s, nullptr, out, 2, block);
}

int insem_hip_realize(insem_hip_ctx* c) {
  if (!c) return set_error(INSEM_HIP_EINVAL, "null context");
  if (c->staging != 2) return set_error(INSEM_HIP_ESTATE, "the context is not a staged one");
  void* block = reinterpret_cast<void*>(c->stage_base);
  try {
    HIP_TRY(hipSetDevice(c->device));
    HIP_TRY(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    HIP_TRY(hipMemcpyAsync(block, c->stage_host.data(), c->stage_used, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
  } catch (const HipFailure& e) {
    if (c->stream) {
      (void)hipStreamDestroy(c->stream);
      c->stream = nullptr;
    }
    return e.code;  // still staged: insem_hip_destroy gives the block back
  }
  {
    std::lock_guard<std::mutex> lock(g_capture_mu);
    g_reserved.erase(block);  // the context owns the block now
  }
  c->allocs.emplace_back(block, c->stage_cap);
  c->staging = 0;
  std::vector<unsigned char>().swap(c->stage_host);
  return INSEM_HIP_OK;
}

}  // extern "C"

namespace {

// Launch order of one block's elements: batches of epb elements (one
// work-group's share) grown greedily from the first unplaced element in
// block order, each time adding the unplaced element of the block with the
// most slot incidences on points the batch already touches (ties: block
// order).  Batches come out compact (about square patches on a structured
// mesh instead of strips), so more points have all their slots in one
// batch.  Results do not depend on the order: every sum runs in the
// reference's slot order.
std::vector<int32_t> batch_order(const insem_hip_domain& d, const insem_hip_block& B, int npe, int epb) {
  const int nb = B.n_elements;
  std::vector<int32_t> loc(static_cast<size_t>(d.n_elements), -1);
  for (int i = 0; i < nb; ++i) loc[B.elements[i]] = i;
  std::vector<int32_t> order, score(static_cast<size_t>(nb), 0), cand;
  std::vector<uint8_t> placed(static_cast<size_t>(nb), 0);
  order.reserve(nb);
  int seed = 0;
  auto place = [&](int x) {
    placed[x] = 1;
    order.push_back(x);
    const int32_t* row = d.l2g + static_cast<size_t>(B.elements[x]) * npe;
    for (int k = 0; k < npe; ++k) {
      if (row[k] < 0) continue;
      for (int32_t t = d.t_offsets[row[k]]; t < d.t_offsets[row[k] + 1]; ++t) {
        const int32_t y = loc[d.t_slots[t] / npe];
        if (y < 0 || placed[y]) continue;
        if (score[y]++ == 0) cand.push_back(y);
      }
    }
  };
  while (static_cast<int>(order.size()) < nb) {
    for (int n = 0; n < epb && static_cast<int>(order.size()) < nb; ++n) {
      int best = -1;
      for (int y : cand)
        if (!placed[y] && (best < 0 || score[y] > score[best] || (score[y] == score[best] && y < best))) best = y;
      if (best < 0) {
        while (placed[seed]) ++seed;
        best = seed;
      }
      place(best);
    }
    for (int y : cand) score[y] = 0;
    cand.clear();
  }
  return order;
}

// A copy of a setup whose blocks list their elements (and per-element data)
// in batch_order.
struct BatchOrderedSetup {
  insem_hip_setup setup{};
  std::vector<insem_hip_domain> domains;
  std::vector<std::vector<insem_hip_block>> blocks;
  std::vector<std::unique_ptr<std::vector<int32_t>>> ints;
  std::vector<std::unique_ptr<std::vector<float>>> floats;
  std::vector<std::unique_ptr<std::vector<uint8_t>>> bytes;

  template <class T>
  const T* permuted(const T* src, const std::vector<int32_t>& order, size_t stride,
                    std::vector<std::unique_ptr<std::vector<T>>>& keep) {
    if (!src) return nullptr;
    auto v = std::make_unique<std::vector<T>>(order.size() * stride);
    for (size_t i = 0; i < order.size(); ++i)
      std::copy(src + order[i] * stride, src + (order[i] + 1) * stride, v->data() + i * stride);
    keep.push_back(std::move(v));
    return keep.back()->data();
  }

  explicit BatchOrderedSetup(const insem_hip_setup* s) : setup(*s) {
    const int npe = s->ngll * s->ngll, epb = force_batch_elements(s->ngll);
    domains.assign(s->domains, s->domains + s->n_domains);
    blocks.resize(s->n_domains);
    for (int di = 0; di < s->n_domains; ++di) {
      const insem_hip_domain& d = s->domains[di];
      blocks[di].assign(d.blocks, d.blocks + d.n_blocks);
      for (insem_hip_block& B : blocks[di]) {
        const std::vector<int32_t> order = batch_order(d, B, npe, epb);
        const size_t ms = B.kind == INSEM_KIND_RECT ? 3 : B.kind == INSEM_KIND_AFFINE ? 5 : 5 * npe;
        B.elements = permuted(B.elements, order, 1, ints);
        B.metric = permuted(B.metric, order, ms, floats);
        B.coef0 = permuted(B.coef0, order, npe, floats);
        B.coef1 = permuted(B.coef1, order, npe, floats);
        B.wr = permuted(B.wr, order, npe, floats);
        B.hoop = permuted(B.hoop, order, npe, floats);
        B.r = permuted(B.r, order, npe, floats);
        B.axial = permuted(B.axial, order, 1, bytes);
      }
      domains[di].blocks = blocks[di].data();
    }
    setup.domains = domains.data();
  }
};

// A staged context's claim on its reserved block, given up (the block stays
// reserved for its caller).
void release_claim(const insem_hip_ctx* c) {
  if (c->staging != 2) return;
  std::lock_guard<std::mutex> lock(g_capture_mu);
  const auto it = g_reserved.find(reinterpret_cast<void*>(c->stage_base));
  if (it != g_reserved.end() && it->second.owner == c) it->second.owner = nullptr;
}

// A context of the setup (and batch): on the current device (staging 0),
// or without any device call, measured (1) or laid out for the reserved
// block `block` (2) for insem_hip_realize.
int create_impl(const insem_hip_setup* s0, const insem_hip_batch* batch, insem_hip_ctx** out, int staging,
                void* block) {
  if (!out) return set_error(INSEM_HIP_EINVAL, "null output pointer");
  *out = nullptr;
  int v = validate_setup(s0);
  if (v == INSEM_HIP_OK && batch) v = validate_batch(s0, batch);
  if (v != INSEM_HIP_OK) return v;
  if (!staging && insem_hip_device_count() <= 0) return set_error(INSEM_HIP_EDEVICE, "no HIP device is visible");
  auto* c = new insem_hip_ctx();
  c->staging = staging;
  if (staging == 2) {
    std::lock_guard<std::mutex> lock(g_capture_mu);
    const auto it = g_reserved.find(block);
    if (it == g_reserved.end() || it->second.owner) {
      delete c;
      return set_error(INSEM_HIP_EINVAL, "the block was not reserved with insem_hip_reserve (or is in use)");
    }
    it->second.owner = c;
    c->device = it->second.device;
    c->stage_cap = it->second.size;
    c->stage_base = reinterpret_cast<uintptr_t>(block);
  }
  try {
    if (staging == 2) c->stage_host.assign(c->stage_cap, 0);
    // Blocks are launched in batch order unless INSEM_HIP_BATCH_ORDER=0.
    const char* bo = std::getenv("INSEM_HIP_BATCH_ORDER");
    std::unique_ptr<BatchOrderedSetup> ordered;
    if (!bo || std::atoi(bo) != 0) ordered = std::make_unique<BatchOrderedSetup>(s0);
    const insem_hip_setup* s = ordered ? &ordered->setup : s0;
    if (!staging) {
      HIP_TRY(hipGetDevice(&c->device));
      HIP_TRY(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    }
    c->ngll = s->ngll;
    const size_t npe = static_cast<size_t>(s->ngll) * s->ngll;
    c->D = c->upload(s->D, npe);
    c->Dj = c->upload(s->Dj, npe);
    c->w2 = c->upload(s->w2, npe);
    c->doms.resize(s->n_domains);
    for (int di = 0; di < s->n_domains; ++di) {
      if (s->domains[di].medium == INSEM_MEDIUM_ELASTIC) c->solid = di;
      if (s->domains[di].medium == INSEM_MEDIUM_ACOUSTIC) c->fluid = di;
    }
    if (const char* f = std::getenv("INSEM_HIP_LOCAL")) c->local = std::atoi(f) != 0;
    for (int di = 0; di < s->n_domains; ++di) build_domain(c, s, di);
    build_coupling(c, s);
    build_receivers(c, s);
    c->n_sources = s->n_sources;
    c->record_every = s->record_every;
    c->energy_every = s->energy_every;
    if (batch) build_batch(c, s, batch);
    int rows = 0;
    for (const DomainRT& R : c->doms) {
      rows += tail_blocks(R.ng) + tail_blocks(R.n_sp) + 2;  // grouped + other regular + special
      for (const BlockRT& br : R.blocks) rows += force_blocks(c->ngll, br.nb);
    }
    rows += c->rec.n_tasks + 16;  // receiver work-groups of a force launch, padding of a paired launch
    c->partial_rows = rows;
    int dot_rows = 0;
    for (const DomainRT& R : c->doms) dot_rows = std::max(dot_rows, tail_blocks(R.ncomp * R.ng));
    c->partial = c->alloc<double>(3 * static_cast<size_t>(std::max(rows, dot_rows)) + 16);
    c->red_out = c->alloc<double>(128);
    c->d_clock = c->alloc<int32_t>(1);
    c->zero(c->d_clock, sizeof(int32_t));
    if (const char* g = std::getenv("INSEM_HIP_GRAPHS")) c->use_graphs = std::atoi(g) != 0;
    if (const char* v = std::getenv("INSEM_HIP_FORCE_VARIANT")) c->force_variant = std::atoi(v);
    if (const char* k = std::getenv("INSEM_HIP_CHUNK")) {
      const int n = std::atoi(k);
      if (n >= 1 && n <= kMaxChunk) c->chunk = n, c->chunk_set = 1;
    }
    if (const char* f = std::getenv("INSEM_HIP_FUSE")) c->fuse = std::atoi(f) != 0;
    if (const char* f = std::getenv("INSEM_HIP_PAIR")) c->pair = std::atoi(f) != 0;
    if (!staging) HIP_TRY(hipStreamSynchronize(c->stream));
  } catch (const HipFailure& e) {
    release_claim(c);
    free_all(c);
    delete c;
    return e.code;
  } catch (const std::exception& e) {
    release_claim(c);
    free_all(c);
    delete c;
    return set_error(INSEM_HIP_EDEVICE, "%s", e.what());
  }
  *out = c;
  return INSEM_HIP_OK;
}

// dt, step count and amplitude table; trace and energy buffers sized for it.
// The paired launch's first blocks (the elastic and the acoustic domain's
// first), or nulls when the model has none.
std::pair<const BlockRT*, const BlockRT*> paired_blocks(const insem_hip_ctx* c) {
  if (c->solid < 0 || c->fluid < 0 || !c->pair || !c->fuse || c->doms[c->solid].blocks.empty() ||
      c->doms[c->fluid].blocks.empty())
    return {nullptr, nullptr};
  const BlockRT& a = c->doms[c->solid].blocks[0];
  const BlockRT& b = c->doms[c->fluid].blocks[0];
  if (!forces_pair_supported(a.phys, a.rect, b.phys, b.rect)) return {nullptr, nullptr};
  return {&a, &b};
}

void apply_time(insem_hip_ctx* c, double dt, int32_t n_steps, const float* amplitudes) {
  c->dt = dt;
  c->n_steps = n_steps;
  if (!c->chunk_set) {
    const auto [a, b] = paired_blocks(c);
    const int k = a ? tuned_chunk_steps(c->ngll, a->phys, a->rect, b->phys, b->rect) : 0;
    c->chunk = k >= 1 && k <= kMaxChunk ? k : 32;
  }
  // captured chunks hold the previous amplitude and trace buffers
  for (auto& g : c->graphs) (void)hipGraphExecDestroy(g.second);
  c->graphs.clear();
  const size_t na = static_cast<size_t>(c->n_sources) * (static_cast<size_t>(n_steps) + 1);
  c->amps = c->upload(amplitudes ? amplitudes : nullptr, amplitudes ? na : 0);
  for (DomainRT& R : c->doms) {
    R.sp.amps = c->amps;
    R.sp.amp_stride = n_steps + 1;
  }
  c->max_samples = n_steps / c->record_every + 1;
  c->traces = c->alloc<double>(static_cast<size_t>(c->max_samples) * std::max(1, c->rec.n_cols));
  c->zero(c->traces, static_cast<size_t>(c->max_samples) * std::max(1, c->rec.n_cols) * sizeof(double));
  c->rec.traces = c->traces;
  c->max_energy_rows = c->energy_every > 0 ? n_steps / c->energy_every + 1 : 0;
  c->energy_rows = c->alloc<double>(2 * static_cast<size_t>(std::max(1, c->max_energy_rows)));
  c->have_time = 1;
}

}  // namespace

extern "C" {

int insem_hip_destroy(insem_hip_ctx* c) {
  if (!c) return INSEM_HIP_OK;
  if (c->staging) {
    // never realized: no device resources but its reserved block
    if (c->staging == 2) {
      release_claim(c);
      (void)insem_hip_unreserve(reinterpret_cast<void*>(c->stage_base));
    }
    delete c;
    return INSEM_HIP_OK;
  }
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  free_all(c);
  delete c;
  return INSEM_HIP_OK;
}

int insem_hip_set_time(insem_hip_ctx* c, double dt, int32_t n_steps, const float* amplitudes) {
  if (c && c->initialized) return set_error(INSEM_HIP_ESTATE, "the time step must be set before the first step");
  if (c && c->n_models) return set_error(INSEM_HIP_ESTATE, "a batch takes its time steps from insem_hip_set_time_batch");
  if (!(dt > 0.0) || n_steps < 0) return set_error(INSEM_HIP_EINVAL, "invalid time step or step count");
  if (c && c->n_sources > 0 && !amplitudes) return set_error(INSEM_HIP_EINVAL, "source amplitudes are required");
  return guarded(c, [&] { apply_time(c, dt, n_steps, amplitudes); });
}

int insem_hip_set_time_batch(insem_hip_ctx* c, const double* dt, const int32_t* n_steps, const float* amplitudes) {
  if (!c) return set_error(INSEM_HIP_EINVAL, "null context");
  if (!c->n_models) return set_error(INSEM_HIP_ESTATE, "not a batch context");
  if (c->initialized) return set_error(INSEM_HIP_ESTATE, "the time step must be set before the first step");
  if (!dt || !n_steps) return set_error(INSEM_HIP_EINVAL, "null time steps or step counts");
  if (c->n_sources > 0 && !amplitudes) return set_error(INSEM_HIP_EINVAL, "source amplitudes are required");
  int32_t max_n = 0;
  std::vector<ModelCoef> mc(static_cast<size_t>(c->n_models));
  for (int m = 0; m < c->n_models; ++m) {
    if (!(dt[m] > 0.0) || n_steps[m] < 0) return set_error(INSEM_HIP_EINVAL, "invalid time step or step count of model %d", m);
    // the single-model expressions (step_coef)
    mc[m].rdt = static_cast<float>(dt[m]);
    mc[m].hdt = static_cast<float>(0.5 * dt[m]);
    mc[m].hdt2 = static_cast<float>(0.5 * dt[m] * dt[m]);
    mc[m].h_damp = 0.5 * dt[m];
    mc[m].cc = dt[m] * dt[m] / 8.0;
    mc[m].half_dt = 0.5 * dt[m];
    mc[m].n_steps = n_steps[m];
    max_n = std::max(max_n, n_steps[m]);
  }
  return guarded(c, [&] {
    c->h2d(c->d_models, mc.data(), mc.size() * sizeof(ModelCoef));
    // Last step and model of each force batch, so that work on models past
    // their last step is skipped by whole work-groups and one-model batches
    // read their coefficients uniformly; the model of each 256 points, for
    // the tails' uniform coefficients.
    const int epb = force_batch_elements(c->ngll);
    c->floors.assign(n_steps, n_steps + c->n_models);
    std::sort(c->floors.begin(), c->floors.end(), std::greater<int32_t>());
    c->floors.erase(std::unique(c->floors.begin(), c->floors.end()), c->floors.end());
    for (auto& g : c->graphs) (void)hipGraphExecDestroy(g.second);
    c->graphs.clear();
    for (DomainRT& R : c->doms) {
      for (BlockRT& br : R.blocks) {
        std::vector<int32_t> last(static_cast<size_t>((br.nb + epb - 1) / epb), -1);
        for (int e = 0; e < br.nb; ++e) last[e / epb] = std::max(last[e / epb], n_steps[br.h_emodel[e]]);
        // elements up to the end of the last force batch with work at each level
        br.nb_at.assign(c->floors.size(), 0);
        for (size_t j = 0; j < c->floors.size(); ++j)
          for (int t = static_cast<int>(last.size()) - 1; t >= 0; --t)
            if (last[t] >= c->floors[j]) {
              br.nb_at[j] = std::min(br.nb, (t + 1) * epb);
              break;
            }
        if (!br.args.batch_last) br.args.batch_last = c->upload(last);
        else c->h2d(const_cast<int32_t*>(br.args.batch_last), last.data(), last.size() * sizeof(int32_t));
        if (!br.args.batch_model) {
          // the one model of a force batch (-1: elements of several models)
          std::vector<int32_t> model(last.size(), -2);
          for (int e = 0; e < br.nb; ++e) {
            int32_t& x = model[e / epb];
            x = x == -2 || x == br.h_emodel[e] ? br.h_emodel[e] : -1;
          }
          br.args.batch_model = c->upload(model);
        }
      }
      if (!R.dd.chunk_model) {
        // the one model of each 256 points (-1: points of several models)
        std::vector<int32_t> model(static_cast<size_t>((R.ng + 255) / 256), -2);
        for (int64_t g = 0; g < R.ng; ++g) {
          int32_t& x = model[g >> 8];
          x = x == -2 || x == R.h_pmodel[g] ? R.h_pmodel[g] : -1;
        }
        R.dd.chunk_model = c->upload(model);
      }
    }
    apply_time(c, dt[0], max_n, amplitudes);
  });
}

int insem_hip_set_state(insem_hip_ctx* c, int32_t di, const float* u, const float* v) {
  if (!c || di < 0 || di >= static_cast<int>(c->doms.size())) return set_error(INSEM_HIP_EINVAL, "invalid domain");
  if (c->initialized) return set_error(INSEM_HIP_ESTATE, "initial conditions must be set before the first time step");
  if (!u || !v) return set_error(INSEM_HIP_EINVAL, "null state array");
  return guarded(c, [&] {
    DomainRT& R = c->doms[di];
    std::vector<float> hu, hv;
    permute_in(R, u, hu);
    permute_in(R, v, hv);
    // Dirichlet and axis points are forced to zero (solver.cpp:776-777).
    for (int64_t gn = 0; gn < R.ng; ++gn) {
      const uint8_t m = R.h_dir_mask[R.perm_old[gn]];
      for (int cc = 0; cc < R.ncomp; ++cc)
        if ((m & 1) || ((m & 2) && ((R.axis_mask >> cc) & 1))) hu[cc * R.ng + gn] = hv[cc * R.ng + gn] = 0.0f;
    }
    c->h2d(R.dd.u, hu.data(), hu.size() * sizeof(float));
    c->h2d(R.dd.v, hv.data(), hv.size() * sizeof(float));
  });
}

int insem_hip_get_field(insem_hip_ctx* c, int32_t di, int32_t field, float* out) {
  if (!c || di < 0 || di >= static_cast<int>(c->doms.size())) return set_error(INSEM_HIP_EINVAL, "invalid domain");
  if (!out || field < 0 || field > 3) return set_error(INSEM_HIP_EINVAL, "invalid field or output");
  return guarded(c, [&] {
    DomainRT& R = c->doms[di];
    HIP_TRY(hipStreamSynchronize(c->stream));
    const float* src = field == INSEM_FIELD_U ? R.dd.u : field == INSEM_FIELD_V ? R.dd.v : field == INSEM_FIELD_A ? R.dd.a : R.dd.f;
    std::vector<float> h(static_cast<size_t>(R.ncomp) * R.ng);
    c->d2h(h.data(), src, h.size() * sizeof(float));
    permute_out(R, h, out);
  });
}

int insem_hip_advance(insem_hip_ctx* c, int32_t n) {
  return guarded(c, [&] {
    require_time(c);
    if (!c->initialized) initialize(c);
    const int m = std::max(0, std::min(n, c->n_steps - c->step));
    for_each_chunk(c, m, [&](int k, bool first, bool last, int level) { run_chunk(c, k, first, last, level); });
    c->step += m;
    // a marker of this call's end, for insem_hip_progress; markers of
    // finished calls are retired here too, and a host that queues more than
    // kMaxMarkers calls ahead of the device waits for the oldest
    retire_markers(c, c->progress.size() >= kMaxMarkers);
    hipEvent_t e = nullptr;
    if (!c->progress_pool.empty()) {
      e = c->progress_pool.back();
      c->progress_pool.pop_back();
    } else {
      HIP_TRY(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    }
    HIP_TRY(hipEventRecord(e, c->stream));
    c->progress.emplace_back(e, c->step);
  });
}

int insem_hip_progress(insem_hip_ctx* c, int32_t* step) {
  if (!c || !step) return set_error(INSEM_HIP_EINVAL, "null argument");
  return guarded(c, [&] {
    retire_markers(c, false);
    *step = c->device_step;
  });
}

int insem_hip_set_force_variant(insem_hip_ctx* c, int32_t variant) {
  if (!c) return set_error(INSEM_HIP_EINVAL, "null context");
  if (variant < -1 || variant >= kForceVariants)
    return set_error(INSEM_HIP_EINVAL, "force variant %d is outside -1 .. %d", variant, kForceVariants - 1);
  return guarded(c, [&] {
    HIP_TRY(hipStreamSynchronize(c->stream));
    c->force_variant = variant;
    // captured chunks hold the previous variant's kernels
    for (auto& g : c->graphs) (void)hipGraphExecDestroy(g.second);
    c->graphs.clear();
  });
}

int insem_hip_force_variant(insem_hip_ctx* c, int32_t* variant, int32_t* paired) {
  if (!c || !variant || !paired) return set_error(INSEM_HIP_EINVAL, "null argument");
  *variant = 0;
  *paired = 0;
  const auto [a, b] = paired_blocks(c);
  if (a) {
    *paired = 1;
    *variant = c->force_variant >= 0 ? c->force_variant : tuned_force_variant(c->ngll, a->phys, a->rect, b->phys, b->rect);
    if (c->ngll < 5 || c->ngll > 7) *variant = 0;
  }
  return INSEM_HIP_OK;
}

int insem_hip_set_chunk(insem_hip_ctx* c, int32_t steps) {
  if (!c) return set_error(INSEM_HIP_EINVAL, "null context");
  if (steps < 0 || steps > kMaxChunk) return set_error(INSEM_HIP_EINVAL, "chunk of %d steps is outside 0 .. %d", steps, kMaxChunk);
  return guarded(c, [&] {
    HIP_TRY(hipStreamSynchronize(c->stream));
    c->chunk_set = steps > 0;
    if (steps > 0) {
      c->chunk = steps;
    } else {
      const auto [a, b] = paired_blocks(c);
      const int k = a ? tuned_chunk_steps(c->ngll, a->phys, a->rect, b->phys, b->rect) : 0;
      c->chunk = k >= 1 && k <= kMaxChunk ? k : 32;
    }
    for (auto& g : c->graphs) (void)hipGraphExecDestroy(g.second);
    c->graphs.clear();
  });
}

int insem_hip_chunk(insem_hip_ctx* c, int32_t* steps) {
  if (!c || !steps) return set_error(INSEM_HIP_EINVAL, "null argument");
  *steps = c->chunk;
  return INSEM_HIP_OK;
}

int insem_hip_prepare(insem_hip_ctx* c, int32_t n) {
  return guarded(c, [&] {
    require_time(c);
    if (!c->use_graphs) return;
    for_each_chunk(c, n, [&](int k, bool first, bool last, int level) { (void)chunk_graph(c, k, first, last, level); });
  });
}

int insem_hip_step(insem_hip_ctx* c, int32_t* step) {
  if (!c || !step) return set_error(INSEM_HIP_EINVAL, "null argument");
  *step = c->step;
  return INSEM_HIP_OK;
}

int insem_hip_n_samples(insem_hip_ctx* c, int32_t* n) {
  if (!c || !n) return set_error(INSEM_HIP_EINVAL, "null argument");
  *n = c->initialized ? c->step / c->record_every + 1 : 0;
  return INSEM_HIP_OK;
}

}  // extern "C"

namespace {

// The trace rows [0, ns) copied to the host h (on the context's stream, or
// on its copy stream without waiting for the context's stream), arranged as
// insem_hip_read_traces documents.
void format_traces(insem_hip_ctx* c, int ns, bool side, double* out) {
  const int nc = c->rec.n_cols;
  std::vector<double> h(static_cast<size_t>(ns) * std::max(1, nc));
  if (ns && nc) {
    if (side) {
      if (!c->copy_stream) HIP_TRY(hipStreamCreateWithFlags(&c->copy_stream, hipStreamNonBlocking));
      HIP_TRY(hipMemcpyAsync(h.data(), c->traces, h.size() * sizeof(double), hipMemcpyDeviceToHost, c->copy_stream));
      HIP_TRY(hipStreamSynchronize(c->copy_stream));
    } else {
      c->d2h(h.data(), c->traces, h.size() * sizeof(double));
    }
  }
  const double kh = 0.70710678118654752440;
  size_t o = 0;
  for (const TraceCol& tc : c->cols) {
    for (int s = 0; s < ns; ++s) {
      const double* row = h.data() + static_cast<size_t>(s) * nc + tc.col;
      if (tc.m1_elastic && tc.raw_ncomp == 3) {
        const double up = row[0], um = row[1];
        out[o++] = kh * (up + um);
        out[o++] = kh * (up - um);
        out[o++] = row[2];
      } else {
        for (int k = 0; k < tc.raw_ncomp; ++k) out[o++] = row[k];
      }
    }
  }
}

}  // namespace

extern "C" {

int insem_hip_read_traces(insem_hip_ctx* c, double* out) {
  if (!out) return set_error(INSEM_HIP_EINVAL, "null output");
  return guarded(c, [&] {
    HIP_TRY(hipStreamSynchronize(c->stream));
    format_traces(c, c->initialized ? c->step / c->record_every + 1 : 0, false, out);
  });
}

int insem_hip_read_traces_done(insem_hip_ctx* c, double* out, int32_t* n_samples) {
  if (!out || !n_samples) return set_error(INSEM_HIP_EINVAL, "null output");
  return guarded(c, [&] {
    retire_markers(c, false);
    // rows of finished advance calls are written and no queued launch
    // writes them again
    const int ns = c->device_started ? c->device_step / c->record_every + 1 : 0;
    format_traces(c, ns, true, out);
    *n_samples = ns;
  });
}

int insem_hip_n_energy(insem_hip_ctx* c, int32_t* n) {
  if (!c || !n) return set_error(INSEM_HIP_EINVAL, "null argument");
  *n = c->initialized && c->energy_every > 0 ? c->step / c->energy_every + 1 : 0;
  return INSEM_HIP_OK;
}

int insem_hip_read_energy(insem_hip_ctx* c, double* out) {
  if (!out) return set_error(INSEM_HIP_EINVAL, "null output");
  return guarded(c, [&] {
    HIP_TRY(hipStreamSynchronize(c->stream));
    const size_t n = c->initialized && c->energy_every > 0 ? c->step / c->energy_every + 1 : 0;
    std::vector<double> h(2 * std::max<size_t>(1, n));
    if (n) c->d2h(h.data(), c->energy_rows, 2 * n * sizeof(double));
    for (size_t r = 0; r < n; ++r) {
      out[3 * r] = static_cast<double>(r * c->energy_every) * c->dt;
      out[3 * r + 1] = h[2 * r + 1];
      out[3 * r + 2] = -0.5 * h[2 * r];
    }
  });
}

int insem_hip_internal_forces(insem_hip_ctx* c, int32_t di, const float* u, float* f, double* dot) {
  if (!c || di < 0 || di >= static_cast<int>(c->doms.size())) return set_error(INSEM_HIP_EINVAL, "invalid domain");
  if (!u || !f) return set_error(INSEM_HIP_EINVAL, "null field");
  return guarded(c, [&] {
    DomainRT& R = c->doms[di];
    std::vector<float> h;
    permute_in(R, u, h);
    HIP_TRY(hipMemcpyAsync(R.tmp, h.data(), h.size() * sizeof(float), hipMemcpyHostToDevice, c->stream));
    run_forces(c, R, R.tmp, c->stream);
    HIP_TRY(launch_assemble(R.dd, c->stream));
    if (dot) {
      int nb = 0;
      HIP_TRY(launch_dot(R.tmp, R.dd.f, static_cast<int64_t>(R.ncomp) * R.ng, c->partial, &nb, c->stream));
      HIP_TRY(launch_reduce(c->partial, nb, 1, c->red_out, 0, c->stream));
      HIP_TRY(hipMemcpyAsync(dot, c->red_out, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    }
    HIP_TRY(hipMemcpyAsync(h.data(), R.dd.f, h.size() * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    permute_out(R, h, f);
  });
}

int insem_hip_block_forces(insem_hip_ctx* c, int32_t di, int32_t bi, const float* ul, float* fl) {
  if (!c || di < 0 || di >= static_cast<int>(c->doms.size())) return set_error(INSEM_HIP_EINVAL, "invalid domain");
  if (bi < 0 || bi >= static_cast<int>(c->doms[di].blocks.size())) return set_error(INSEM_HIP_EINVAL, "invalid block");
  if (!ul || !fl) return set_error(INSEM_HIP_EINVAL, "null field");
  return guarded(c, [&] {
    DomainRT& R = c->doms[di];
    BlockRT& br = R.blocks[bi];
    const size_t n = static_cast<size_t>(R.ne) * R.ncomp * R.npe;
    float* din = R.fl_tmp;
    float* dout = c->alloc<float>(n);  // temporary, freed below
    HIP_TRY(hipMemcpyAsync(din, ul, n * sizeof(float), hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(dout, fl, n * sizeof(float), hipMemcpyHostToDevice, c->stream));
    ForceArgs a = br.args;
    a.ul = din;
    a.fl = dout;
    a.out_ref = 1;
    HIP_TRY(launch_forces(c->ngll, br.phys, br.rect, false, a, StepParams(), c->stream));
    HIP_TRY(hipMemcpyAsync(fl, dout, n * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    const auto blk = c->allocs.back();
    c->allocs.pop_back();
    c->footprint -= static_cast<int64_t>(n * sizeof(float));
    give_block(c->device, blk.first, blk.second);
  });
}

// Output staging of insem_hip_block_forces_noalloc: one array per GPU, held
// from the library's load until the process ends.  Host code uses only the
// device address hipGetSymbolAddress returns for it on each device (the
// array's name is a host-side stand-in, not a device address).
constexpr size_t kNoAllocFloats = size_t{8} << 20;  // 32 MiB
static __device__ __align__(256) float scratch[kNoAllocFloats];

namespace {

constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mu;
float* g_scratch_ptr[kMaxScratchDevices] = {};

// The device address of `scratch` on the current device, looked up on the
// first call on that device.
float* scratch_address(int device) {
  if (device < 0 || device >= kMaxScratchDevices) {
    set_error(INSEM_HIP_EDEVICE, "device %d is beyond the %d devices the no-allocation path supports", device,
              kMaxScratchDevices);
    throw HipFailure{INSEM_HIP_EDEVICE};
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    void* p = nullptr;
    HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_scratch_ptr[device] = static_cast<float*>(p);
  }
  return g_scratch_ptr[device];
}

}  // namespace

int insem_hip_block_forces_noalloc(insem_hip_ctx* c, int32_t di, int32_t bi, const float* ul, float* fl) {
  if (!c || di < 0 || di >= static_cast<int>(c->doms.size())) return set_error(INSEM_HIP_EINVAL, "invalid domain");
  if (bi < 0 || bi >= static_cast<int>(c->doms[di].blocks.size())) return set_error(INSEM_HIP_EINVAL, "invalid block");
  if (!ul || !fl) return set_error(INSEM_HIP_EINVAL, "null field");
  const DomainRT& R0 = c->doms[di];
  const size_t n = static_cast<size_t>(R0.ne) * R0.ncomp * R0.npe;
  if (n > kNoAllocFloats)
    return set_error(INSEM_HIP_EINVAL,
                     "element-local field of %zu values exceeds the no-allocation limit of %zu values (32 MiB)", n,
                     kNoAllocFloats);
  return guarded(c, [&] {
    DomainRT& R = c->doms[di];
    BlockRT& br = R.blocks[bi];
    float* dout = scratch_address(c->device);
    // The caller's f_local goes in first, so elements outside the block
    // keep their values whatever an earlier call left in the array.
    HIP_TRY(hipMemcpyAsync(R.fl_tmp, ul, n * sizeof(float), hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(dout, fl, n * sizeof(float), hipMemcpyHostToDevice, c->stream));
    ForceArgs a = br.args;
    a.ul = R.fl_tmp;
    a.fl = dout;
    a.out_ref = 1;
    HIP_TRY(launch_forces(c->ngll, br.phys, br.rect, false, a, StepParams(), c->stream));
    HIP_TRY(hipMemcpyAsync(fl, dout, n * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
  });
}

// Power iteration (solver.cpp:583-667) with FP64 reductions on the device.
int insem_hip_lambda_start(insem_hip_ctx* c, const float* start) {
  if (!start) return set_error(INSEM_HIP_EINVAL, "null argument");
  if (c && c->initialized) return set_error(INSEM_HIP_ESTATE, "the power iteration runs before the first step");
  if (c && c->n_models) return set_error(INSEM_HIP_ESTATE, "the power iteration is not available on a batch");
  return guarded(c, [&] {
    hipStream_t st = c->stream;
    // red_out: per run r (each domain, then the coupled pair) [32 r + 0..2]
    // = lambda, done, scale; [32 r + 8..10] reductions; [32 r + 16..17] nx, ny
    c->lam_runs = 0;
    c->lam_host.clear();
    size_t off = 0;
    for (size_t di = 0; di < c->doms.size(); ++di) {
      DomainRT& R = c->doms[di];
      double* state = c->red_out + 32 * c->lam_runs;
      double* out3 = state + 8;
      const size_t len = static_cast<size_t>(R.ncomp) * R.ng;
      c->lam_host.emplace_back();
      permute_in(R, start + off, c->lam_host.back());
      off += len;
      HIP_TRY(hipMemcpyAsync(R.tmp, c->lam_host.back().data(), len * sizeof(float), hipMemcpyHostToDevice, st));
      HIP_TRY(hipMemsetAsync(state, 0, 3 * sizeof(double), st));
      const int all = (1 << R.ncomp) - 1;
      queue_repeated(c, 100, [&] {
        HIP_TRY(launch_zero_points(R.tmp, R.ng, R.d_dir, R.n_dir, all, R.ncomp, st));
        HIP_TRY(launch_zero_points(R.tmp, R.ng, R.d_axis, R.n_axis, R.axis_mask, R.ncomp, st));
        run_forces(c, R, R.tmp, st);
        HIP_TRY(launch_assemble(R.dd, st));
        int nb = 0;
        HIP_TRY(launch_power_stats(R.dd, R.tmp, c->partial, state, &nb, st));
        HIP_TRY(launch_reduce(c->partial, nb, 3, out3, 0, st));
        HIP_TRY(launch_power_update(out3, state, st));
        HIP_TRY(launch_scale(R.tmp, static_cast<int64_t>(len), state, st));
      });
      ++c->lam_runs;
    }
    if (c->solid >= 0 && c->fluid >= 0 && c->doms[c->solid].n_cpl > 0) {
      DomainRT& S = c->doms[c->solid];
      DomainRT& F = c->doms[c->fluid];
      DomainRT* order[2] = {&S, &F};
      double* state = c->red_out + 32 * c->lam_runs;
      double* nxy = state + 16;
      // Keep u and a (device copies: tmp holds u, fl_tmp a); start vectors
      // solid u, then fluid u.
      for (DomainRT* R : order) {
        const size_t len = static_cast<size_t>(R->ncomp) * R->ng;
        HIP_TRY(hipMemcpyAsync(R->tmp, R->dd.u, len * sizeof(float), hipMemcpyDeviceToDevice, st));
        HIP_TRY(hipMemcpyAsync(R->fl_tmp, R->dd.a, len * sizeof(float), hipMemcpyDeviceToDevice, st));
      }
      for (DomainRT* R : order) {
        const size_t len = static_cast<size_t>(R->ncomp) * R->ng;
        c->lam_host.emplace_back();
        permute_in(*R, start + off, c->lam_host.back());
        off += len;
        HIP_TRY(hipMemcpyAsync(R->dd.u, c->lam_host.back().data(), len * sizeof(float), hipMemcpyHostToDevice, st));
      }
      HIP_TRY(hipMemsetAsync(state, 0, 3 * sizeof(double), st));
      // The growth ratio counts from iteration 60 on: the iterations before
      // and after differ only in that, so each half is one graph launched 60
      // times.
      auto iteration = [&](int it) {
        for (DomainRT* R : order) {
          HIP_TRY(launch_zero_points(R->dd.u, R->ng, R->d_dir, R->n_dir, (1 << R->ncomp) - 1, R->ncomp, st));
          HIP_TRY(launch_zero_points(R->dd.u, R->ng, R->d_axis, R->n_axis, R->axis_mask, R->ncomp, st));
        }
        for (int q = 0; q < 2; ++q) {
          int nb = 0;
          HIP_TRY(launch_mass_norm(order[q]->dd, order[q]->dd.u, c->partial, &nb, st));
          HIP_TRY(launch_reduce(c->partial, nb, 1, nxy, q, st));
        }
        for (DomainRT* R : order) {
          run_forces(c, *R, R->dd.u, st);
          HIP_TRY(launch_assemble(R->dd, st));
          const DomainRT& O = R == &S ? F : S;
          HIP_TRY(launch_coupling_power(R->dd, R->cpl_pts, R->cpl_other, R->cpl_cx, R->cpl_cz, R->n_cpl, R->cpl_role,
                                        R->am == 1, R->cpl_role == 1 ? O.dd.u : O.dd.a, O.ng, st));
          HIP_TRY(launch_accel(R->dd, st));
          HIP_TRY(launch_zero_points(R->dd.a, R->ng, R->d_dir, R->n_dir, (1 << R->ncomp) - 1, R->ncomp, st));
          HIP_TRY(launch_zero_points(R->dd.a, R->ng, R->d_axis, R->n_axis, R->axis_mask, R->ncomp, st));
        }
        for (int q = 0; q < 2; ++q) {
          int nb = 0;
          HIP_TRY(launch_mass_norm(order[q]->dd, order[q]->dd.a, c->partial, &nb, st));
          HIP_TRY(launch_reduce(c->partial, nb, 1, nxy + 1, q, st));
        }
        HIP_TRY(launch_coupled_update(nxy, nxy + 1, it, state, st));
        for (DomainRT* R : order)
          HIP_TRY(launch_scale_from(R->dd.u, R->dd.a, static_cast<int64_t>(R->ncomp) * R->ng, state, st));
      };
      queue_repeated(c, 60, [&] { iteration(0); });
      queue_repeated(c, 60, [&] { iteration(60); });
      for (DomainRT* R : order) {
        const size_t len = static_cast<size_t>(R->ncomp) * R->ng;
        HIP_TRY(hipMemcpyAsync(R->dd.u, R->tmp, len * sizeof(float), hipMemcpyDeviceToDevice, st));
        HIP_TRY(hipMemcpyAsync(R->dd.a, R->fl_tmp, len * sizeof(float), hipMemcpyDeviceToDevice, st));
      }
      ++c->lam_runs;
    }
  });
}

int insem_hip_lambda_result(insem_hip_ctx* c, double* lambda_out) {
  if (!lambda_out) return set_error(INSEM_HIP_EINVAL, "null argument");
  if (c && c->lam_runs == 0) return set_error(INSEM_HIP_ESTATE, "insem_hip_lambda_start has not been called");
  return guarded(c, [&] {
    std::vector<double> st(32 * static_cast<size_t>(c->lam_runs));
    HIP_TRY(hipMemcpyAsync(st.data(), c->red_out, st.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    double best = 0.0;
    for (int r = 0; r < c->lam_runs; ++r) best = std::max(best, st[32 * static_cast<size_t>(r)]);
    c->lam_host.clear();
    for (hipGraphExec_t g : c->lam_execs) (void)hipGraphExecDestroy(g);
    c->lam_execs.clear();
    *lambda_out = best;
  });
}

int insem_hip_lambda_max(insem_hip_ctx* c, const float* start, double* lambda_out) {
  if (!lambda_out) return set_error(INSEM_HIP_EINVAL, "null argument");
  const int rc = insem_hip_lambda_start(c, start);
  return rc != INSEM_HIP_OK ? rc : insem_hip_lambda_result(c, lambda_out);
}

int insem_hip_synchronize(insem_hip_ctx* c) {
  return guarded(c, [&] { HIP_TRY(hipStreamSynchronize(c->stream)); });
}

int insem_hip_query(insem_hip_ctx* c, int32_t* idle) {
  if (!idle) return set_error(INSEM_HIP_EINVAL, "null argument");
  return guarded(c, [&] {
    const hipError_t e = hipStreamQuery(c->stream);
    if (e != hipSuccess && e != hipErrorNotReady) HIP_TRY(e);
    *idle = e == hipSuccess ? 1 : 0;
  });
}

int insem_hip_device_memory(int64_t* free_bytes, int64_t* total_bytes) {
  if (!free_bytes || !total_bytes) return set_error(INSEM_HIP_EINVAL, "null argument");
  size_t f = 0, t = 0;
  const hipError_t e = hipMemGetInfo(&f, &t);
  if (e != hipSuccess) return set_error(INSEM_HIP_EDEVICE, "hipMemGetInfo failed: %s", hipGetErrorString(e));
  int dev = 0;
  if (hipGetDevice(&dev) == hipSuccess) {
    std::lock_guard<std::mutex> lock(g_capture_mu);
    if (const BlockCache* bc = cache_of(dev)) f += bc->bytes;  // cached blocks serve new contexts
  }
  *free_bytes = static_cast<int64_t>(f);
  *total_bytes = static_cast<int64_t>(t);
  return INSEM_HIP_OK;
}

int insem_hip_release_cache(int64_t* bytes) {
  int dev = 0;
  const hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return set_error(INSEM_HIP_EDEVICE, "hipGetDevice failed: %s", hipGetErrorString(e));
  std::lock_guard<std::mutex> lock(g_capture_mu);
  const BlockCache* bc = cache_of(dev);
  if (bytes) *bytes = bc ? static_cast<int64_t>(bc->bytes) : 0;
  release_cache_locked(dev);
  return INSEM_HIP_OK;
}

int insem_hip_set_timing(insem_hip_ctx* c, int32_t enabled) {
  return guarded(c, [&] {
    HIP_TRY(hipStreamSynchronize(c->stream));
    c->timing = enabled ? 1 : 0;
  });
}

int insem_hip_kernel_time(insem_hip_ctx* c, double* force_ms, int32_t* force_n, double* update_ms, int32_t* update_n) {
  return guarded(c, [&] {
    HIP_TRY(hipStreamSynchronize(c->stream));
    c->block_ms.clear();
    for (auto* v : {&c->ev_force, &c->ev_update}) {
      for (size_t q = 0; q < v->size(); ++q) {
        auto& e = (*v)[q];
        float ms = 0.0f;
        HIP_TRY(hipEventElapsedTime(&ms, e.first, e.second));
        if (v == &c->ev_force) {
          c->force_ms += ms;
          ++c->force_n;
          auto& bm = c->block_ms[c->ev_force_tag[q]];
          bm.first += ms;
          ++bm.second;
        } else {
          c->update_ms += ms;
          ++c->update_n;
        }
        c->ev_pool.push_back(e.first);
        c->ev_pool.push_back(e.second);
      }
      v->clear();
    }
    c->ev_force_tag.clear();
    if (force_ms) *force_ms = c->force_ms;
    if (force_n) *force_n = c->force_n;
    if (update_ms) *update_ms = c->update_ms;
    if (update_n) *update_n = c->update_n;
    c->force_ms = c->update_ms = 0.0;
    c->force_n = c->update_n = 0;
  });
}

int insem_hip_block_time(insem_hip_ctx* c, int32_t di, int32_t bi, double* ms, int32_t* n) {
  if (!c || !ms || !n) return set_error(INSEM_HIP_EINVAL, "null argument");
  const auto it = c->block_ms.find(di * 64 + bi);
  *ms = it == c->block_ms.end() ? 0.0 : it->second.first;
  *n = it == c->block_ms.end() ? 0 : it->second.second;
  return INSEM_HIP_OK;
}

int insem_hip_block_points(insem_hip_ctx* c, int32_t di, int32_t bi, int64_t* owned, int64_t* read) {
  if (!c || !owned || !read) return set_error(INSEM_HIP_EINVAL, "null argument");
  if (di < 0 || di >= static_cast<int32_t>(c->doms.size()) || bi < 0 ||
      bi >= static_cast<int32_t>(c->doms[di].blocks.size()))
    return set_error(INSEM_HIP_EINVAL, "no block %d of domain %d", bi, di);
  *owned = c->fuse ? c->doms[di].blocks[bi].n_owned : 0;
  *read = c->doms[di].blocks[bi].n_read;
  return INSEM_HIP_OK;
}

int insem_hip_block_local_slots(insem_hip_ctx* c, int32_t di, int32_t bi, int64_t* local_slots) {
  if (!c || !local_slots) return set_error(INSEM_HIP_EINVAL, "null argument");
  if (di < 0 || di >= static_cast<int32_t>(c->doms.size()) || bi < 0 ||
      bi >= static_cast<int32_t>(c->doms[di].blocks.size()))
    return set_error(INSEM_HIP_EINVAL, "no block %d of domain %d", bi, di);
  *local_slots = c->fuse ? c->doms[di].blocks[bi].n_local : 0;
  return INSEM_HIP_OK;
}

int insem_hip_footprint(insem_hip_ctx* c, int64_t* bytes) {
  if (!c || !bytes) return set_error(INSEM_HIP_EINVAL, "null argument");
  *bytes = c->footprint;
  return INSEM_HIP_OK;
}

}  // extern "C"
