// C-ABI of libmidsx_amd.so (declared in include/midsx_amd.h): context,
// table build and upload, domain, source, tallies, runs and read-back.
#include <dlfcn.h>
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <mutex>
#include <set>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "../../include/midsx_amd.h"
#include "kernels.h"
#include "mdx_types.h"
#include "scene.h"
#include "tables.h"

namespace {

thread_local std::string g_last_error;

struct DeviceError : std::runtime_error {
  using std::runtime_error::runtime_error;
};
struct NoMemError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

void hip_check(hipError_t e, const char* what) {
  if (e == hipSuccess) return;
  if (e == hipErrorOutOfMemory) throw NoMemError(std::string(what) + ": " + hipGetErrorString(e));
  throw DeviceError(std::string(what) + ": " + hipGetErrorString(e));
}

// Device blocks kept for reuse by exact size: jobs of the same shapes come
// and go, and a block a finished job gave back serves the next one without a
// hipFree (which waits for the device) and a hipMalloc.
struct Pool {
  std::multimap<size_t, void*> blocks;
  size_t cached = 0, limit = 0;  // bytes held for reuse, and their cap
  Pool() = default;
  Pool(const Pool&) = delete;
  Pool& operator=(const Pool&) = delete;
  ~Pool() { trim(); }
  void* get(size_t n) {
    auto it = blocks.find(n);
    if (it == blocks.end()) return nullptr;
    void* p = it->second;
    blocks.erase(it);
    cached -= n;
    return p;
  }
  void put(void* p, size_t n) {
    if (cached + n > limit) {
      (void)hipFree(p);
      return;
    }
    blocks.emplace(n, p);
    cached += n;
  }
  void trim() {
    for (auto& b : blocks) (void)hipFree(b.second);
    blocks.clear();
    cached = 0;
  }
};

void forget_clear(void* p, size_t bytes);

// Device allocation owned by the context (from its pool when one is set).
struct DBuf {
  void* p = nullptr;
  size_t bytes = 0;
  Pool* pool = nullptr;
  DBuf() = default;
  explicit DBuf(Pool* pl) : pool(pl) {}
  DBuf(const DBuf&) = delete;
  DBuf& operator=(const DBuf&) = delete;
  ~DBuf() { release(); }
  void release() {
    if (p) {
      forget_clear(p, bytes);
      if (pool) pool->put(p, bytes);
      else (void)hipFree(p);
    }
    p = nullptr;
    bytes = 0;
  }
  void alloc(size_t n) {
    release();
    if (n == 0) return;
    if (pool && (p = pool->get(n))) {
      bytes = n;
      return;
    }
    hipError_t e = hipMalloc(&p, n);
    if (e == hipErrorOutOfMemory && pool && pool->cached) {
      (void)hipGetLastError();
      pool->trim();
      e = hipMalloc(&p, n);
    }
    if (e != hipSuccess) p = nullptr;
    hip_check(e, "hipMalloc");
    bytes = n;
  }
  // Keeps an allocation of the same size (no free, which may wait for the device).
  void ensure(size_t n) {
    if (n != bytes || (n && !p)) alloc(n);
  }
  template <class T>
  void upload(const T* src, size_t count, hipStream_t st) {
    ensure(count * sizeof(T));
    if (count) hip_check(hipMemcpyAsync(p, src, count * sizeof(T), hipMemcpyHostToDevice, st), "upload");
  }
  template <class T>
  T* as() const {
    return static_cast<T*>(p);
  }
};

// The material IDs present in each macro cell of a grid (MDX_MACRO_EDGE^3
// voxels, x fastest; 256 bits per cell), read eight voxels at a time along x
// (a word of one repeated byte sets one bit); threads take the layers of
// cells along z.  Also the IDs present anywhere.
void scan_cells(const uint8_t* ids, const int dims[3], std::vector<uint64_t>& cells, int mc[3], bool* present) {
  static_assert(MDX_MACRO_EDGE == 8, "a cell's row of voxels is read as one 8-byte word");
  for (int a = 0; a < 3; ++a) mc[a] = (dims[a] + MDX_MACRO_EDGE - 1) >> MDX_MACRO_SHIFT;
  const size_t nc = (size_t)mc[0] * mc[1] * mc[2];
  cells.assign(nc * 4, 0);
  const int nx = dims[0], ny = dims[1], nz = dims[2];
  const int nth = std::max(1, std::min({16, (int)std::thread::hardware_concurrency(), mc[2]}));
  auto work = [&](int t) {  // a thread takes layers of cells along z and reads their voxel rows in memory order
    for (int ck = t; ck < mc[2]; ck += nth) {
      const int k1 = std::min(ck * MDX_MACRO_EDGE + MDX_MACRO_EDGE, nz);
      for (int k = ck * MDX_MACRO_EDGE; k < k1; ++k)
        for (int j = 0; j < ny; ++j) {
          const uint8_t* row = ids + (size_t)nx * ((size_t)j + (size_t)ny * k);
          uint64_t* m = &cells[4 * (size_t)mc[0] * ((size_t)(j >> MDX_MACRO_SHIFT) + (size_t)mc[1] * ck)];
          int i = 0;
          for (; i + MDX_MACRO_EDGE <= nx; i += MDX_MACRO_EDGE, m += 4) {
            uint64_t v;
            std::memcpy(&v, row + i, 8);
            const uint64_t b = v & 255u;
            if (v == b * 0x0101010101010101ull) {
              m[b >> 6] |= 1ull << (b & 63);
            } else {
              for (int q = 0; q < 8; ++q) m[row[i + q] >> 6] |= 1ull << (row[i + q] & 63);
            }
          }
          for (; i < nx; ++i) m[row[i] >> 6] |= 1ull << (row[i] & 63);
        }
    }
  };
  std::vector<std::thread> th;
  for (int t = 1; t < nth; ++t) th.emplace_back(work, t);
  work(0);
  for (auto& x : th) x.join();
  std::fill(present, present + 256, false);
  uint64_t all[4] = {0, 0, 0, 0};
  for (size_t c = 0; c < nc; ++c)
    for (int q = 0; q < 4; ++q) all[q] |= cells[4 * c + q];
  for (int id = 0; id < 256; ++id) present[id] = (all[id >> 6] >> (id & 63)) & 1u;
}

std::string default_data_dir() {
  if (const char* e = std::getenv("MIDSX_AMD_DATA")) return e;
  Dl_info info;
  if (dladdr((void*)&default_data_dir, &info) && info.dli_fname) {
    std::string so = info.dli_fname;
    size_t k = so.rfind('/');
    std::string dir = k == std::string::npos ? "." : so.substr(0, k);
    return dir + "/../data";  // midsx_amd/_lib/libmidsx_amd.so -> midsx_amd/data
  }
  return "midsx_amd/data";
}

struct GridHost {
  std::vector<uint8_t> ids;  // host copy (host-only contexts and NIfTI scenes)
  int dims[3];
  double spacing[3], origin[3];
  bool present[256] = {false};  // material IDs the grid holds
  DBuf d_ids, d_dose, d_dose_n;  // d_dose holds the sums in the job's MDX_DOSE_* layout
  int dose_mode = -1;            // layout d_dose was set up for
  std::vector<uint16_t> macro;   // local majorants per macro cell (see mdx_grid_dev::macro)
  DBuf d_macro;                  // its device copy
  int macro_gen = -1;            // tables generation the table was built for
  int mc[3] = {0, 0, 0};         // macro cells per axis
  std::vector<uint64_t> cell_ids;  // material IDs present per macro cell (256 bits each)
  explicit GridHost(Pool* pl) : d_ids(pl), d_dose(pl), d_dose_n(pl), d_macro(pl) {}
  size_t voxels() const { return (size_t)dims[0] * dims[1] * dims[2]; }
};

struct TallyHost {
  mdx_tally t;
  DBuf acc, hist, image;
  explicit TallyHost(Pool* pl) : acc(pl), hist(pl), image(pl) {}
};

}  // namespace

// One scene: domain and grids, source, tallies and their device copies.
// A context holds several (one per served job); the per-scene calls of the
// C-ABI act on the selected one.
struct Job {
  int domain_of = -1;  // >= 0: uses that job's domain (grids, dose) instead of its own
  bool domain_set = false;
  double dim_space[3] = {0, 0, 0};
  int background_id = -1;
  std::vector<std::unique_ptr<GridHost>> grids;
  bool voxel_dose = false;
  int dose_mode = MDX_DOSE_F64;  // accumulation precision of the voxel dose
  DBuf d_bg, d_domain;
  bool source_set = false;
  mdx_source_dev src{};
  DBuf d_cdf, d_energies, d_source;
  std::vector<std::unique_ptr<TallyHost>> tallies;
  DBuf d_tallies;
  DBuf d_counters;  // per-job event counters
  bool flights = true;  // background free flights when the tallies allow them
  // FP32 dose staged on the main stream (mdx_stage_voxel_dose_f32)
  DBuf d_stage;
  hipEvent_t ev_stage = nullptr;
  int staged_grid = -1;
  uint64_t staged_n = 0;
  hipEvent_t ev_copy = nullptr;  // end of the staged dose's copy to the host (mdx_read_staged_dose_f32_async)
  bool copy_pending = false;
  // Every buffer from the context's pool: a job's release then returns them
  // for reuse instead of calling hipFree, which waits for the whole device
  // (a round or another job's result copy in flight).
  explicit Job(Pool* pl)
      : d_bg(pl), d_domain(pl), d_cdf(pl), d_energies(pl), d_source(pl), d_tallies(pl), d_counters(pl), d_stage(pl) {}
  Job(const Job&) = delete;
  Job& operator=(const Job&) = delete;
  ~Job() {
    if (copy_pending && ev_copy) (void)hipEventSynchronize(ev_copy);
    if (ev_stage) (void)hipEventDestroy(ev_stage);
    if (ev_copy) (void)hipEventDestroy(ev_copy);
  }
  size_t device_bytes() const {
    size_t b = d_bg.bytes + d_domain.bytes + d_cdf.bytes + d_energies.bytes + d_source.bytes + d_tallies.bytes +
               d_counters.bytes;
    for (auto& g : grids) b += g->d_ids.bytes + g->d_dose.bytes + g->d_dose_n.bytes + g->d_macro.bytes;
    b += d_stage.bytes;  // the FP32 dose staged for reading, held until its copy is done
    for (auto& t : tallies) b += t->acc.bytes + t->hist.bytes + t->image.bytes;
    return b;
  }
};

struct mdx_ctx {
  Pool pool;  // declared first: destroyed after every job that returns blocks to it
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  int num_cu = 256;
  size_t lds_optin = 65536;  // LDS a block can take (opt-in limit)
  std::string data_dir;
  bool registry_loaded = false;
  std::map<int, mdx::MaterialDef> registry;

  // tables (shared by every job)
  std::vector<std::unique_ptr<mdx::MaterialTables>> mats;
  mdx::Table majorant;
  mdx::Packed packed;
  DBuf d_xs, d_nodes, d_bins, d_rE, d_rinv, d_roff, d_rrows, d_rhash, d_rguide, d_tables;
  bool tables_ready = false;
  int tables_gen = 0;  // bumped by every table build
  // per slot k: the slots whose total cross section k's bounds from above on
  // [kMacroElo, kMacroEhi] (local majorants)
  std::vector<uint32_t> dominated;
  bool local_majorants = true;
  bool rita_on_device = true;  // RITA tables built by k_rita_build (FP64) instead of on the host

  // scenes
  std::vector<std::unique_ptr<Job>> jobs;
  int cur = 0;
  Job& J() { return *jobs[cur]; }
  // The job owning the selected job's domain (the job itself unless it aliases another's).
  Job& DJ() {
    const int d = jobs[cur]->domain_of;
    if (d < 0) return *jobs[cur];
    if (d >= (int)jobs.size() || !jobs[d]) throw std::runtime_error("the job whose domain this job uses was released");
    return *jobs[d];
  }

  DBuf d_counters;
  // this device's static scratch (mdx_run_jobs_noalloc), looked up at creation
  void* scratch = nullptr;
  // queue-kernel launch parameters (mdx_set_queue_params)
  int queue_blocks_per_cu = 4;
  int refill_min = 32;
  int interact_min = 32;
  int event_kernel = MDX_EVENT_QUEUE;  // the event form mdx_run_jobs launches (mdx_set_event_kernel)
  int bank_size = 128;                 // photons per wave bank of MDX_EVENT_BANK
  DBuf d_work;  // queue kernel: histories claimed, one counter per job of a launch
  DBuf d_jobs, d_prefix;
  DBuf d_dose32;  // staging of mdx_get_voxel_dose_f32 (on the I/O stream)
  // Pinned host mailbox that kernels write small results into directly (the
  // round's counters, ROI sums) and read ROI requests from: no DMA copy, so
  // these reads never queue behind a large result copy on the copy engines.
  // Layout: [counters MDX_NCOUNTER u64][ROI requests][ROI sums].
  void* h_mail = nullptr;
  void* d_mail = nullptr;  // the mailbox's address on the device
  size_t mail_roi = 0;     // ROI boxes the mailbox holds

  // Round buffer of the graph path: [n_jobs, total, prefix[kRoundMax + 1], jobs[kRoundMax]]
  bool use_graphs = false;
  bool inflight = false;  // a round was launched by mdx_run_jobs_async and not yet waited for
  std::vector<int> round_jobs, round_owners;  // the jobs of that round and the owners of their domains
  hipStream_t io_stream = nullptr;
  hipStream_t copy_stream = nullptr;  // staged results to the host: large copies that set-up work must not queue behind
  hipStream_t js = nullptr;  // stream of the current call's job set-up work
  void* h_round = nullptr;  // pinned
  DBuf d_round;
  hipGraphExec_t graph = nullptr;
  int graph_params[5] = {0, 0, 0, 0, 0};

  void ensure_registry() {
    if (registry_loaded) return;
    registry = mdx::load_registry(data_dir + "/materials.csv");
    registry_loaded = true;
  }
  // mass-energy absorption coefficients of dry air (E eV, mu_en/rho cm^2/g)
  std::vector<double> muen_e, muen_v;
  int muen_table = 0;  // 0: NIST grid (air_muen_nist.csv), 1: TG-195 0.5 keV grid (air_muen_tg195.csv)
  void ensure_air_muen() {
    if (!muen_e.empty()) return;
    const std::string path = data_dir + (muen_table == 1 ? "/air_muen_tg195.csv" : "/air_muen_nist.csv");
    FILE* f = std::fopen(path.c_str(), "r");
    if (!f) throw std::runtime_error("air mu_en table not found: " + path);
    char line[256];
    while (std::fgets(line, sizeof line, f)) {
      double e, v;
      if (std::sscanf(line, "%lf,%lf", &e, &v) == 2) {
        muen_e.push_back(e * 1e3);
        muen_v.push_back(v);
      }
    }
    std::fclose(f);
    if (muen_e.size() < 2) throw std::runtime_error("air mu_en table is empty: " + path);
  }
  int slot_of(int id) const { return (id >= 0 && id < 256) ? packed.id_to_slot[id] : 255; }
};

namespace {

constexpr int kRoundMax = 1024;
constexpr size_t kRoundHeader = 2 * sizeof(uint64_t);
constexpr size_t kRoundPrefix = (kRoundMax + 1) * sizeof(uint64_t);
constexpr int kRoundClearMax = 1024;  // ranges a round's graph clears before its transport kernel
constexpr size_t kRoundClearOff = kRoundHeader + kRoundPrefix + kRoundMax * sizeof(mdx_job_dev);
constexpr size_t kRoundBytes = kRoundClearOff + (1 + 2 * kRoundClearMax) * sizeof(uint64_t);

// Deferred clearing.  While a graph round runs, its blocks hold every CU,
// so a kernel that clears a new job's dose or tally buffers on the set-up
// stream would only start when the round ends, and the set-up call would wait
// for it.  Such ranges are queued instead and cleared by the next round's
// graph, right before its transport kernel; anything that reads or launches
// on a queued range first clears it (flush_clears).  Ranges of a buffer that
// is freed or reallocated leave the queue.
struct ClearReq {
  const void* ctx;
  void* p;
  size_t bytes;
};
std::mutex g_clear_mu;
std::vector<ClearReq> g_clears;

void forget_clear(void* p, size_t bytes) {
  std::lock_guard<std::mutex> lk(g_clear_mu);
  const char* a = static_cast<const char*>(p);
  g_clears.erase(std::remove_if(g_clears.begin(), g_clears.end(),
                                [&](const ClearReq& r) {
                                  const char* b = static_cast<const char*>(r.p);
                                  return b < a + bytes && a < b + r.bytes;
                                }),
                 g_clears.end());
}

template <class F>
int guard(F&& f) {
  try {
    f();
    g_last_error.clear();
    return MDX_OK;
  } catch (const std::invalid_argument& e) {
    g_last_error = e.what();
    return MDX_ERR_INVALID;
  } catch (const std::out_of_range& e) {
    g_last_error = e.what();
    return MDX_ERR_RANGE;
  } catch (const DeviceError& e) {
    g_last_error = e.what();
    return MDX_ERR_DEVICE;
  } catch (const NoMemError& e) {
    g_last_error = e.what();
    return MDX_ERR_NOMEM;
  } catch (const std::bad_alloc&) {
    g_last_error = "host allocation failed";
    return MDX_ERR_NOMEM;
  } catch (const std::exception& e) {
    g_last_error = e.what();
    return MDX_ERR_RUNTIME;
  } catch (...) {
    g_last_error = "unknown error";
    return MDX_ERR_RUNTIME;
  }
}

void need(bool ok, const char* msg) {
  if (!ok) throw std::invalid_argument(msg);
}

// For reads that may overlap a round in flight (they copy on the I/O stream).
void set_device_nowait(mdx_ctx* c) {
  if (c->device < 0) throw std::invalid_argument("host-only context (device -1) cannot run device work");
  hip_check(hipSetDevice(c->device), "hipSetDevice");
}

// Every call that changes device state first waits for a round in flight.
void set_device(mdx_ctx* c) {
  set_device_nowait(c);
  if (c->inflight) hip_check(hipStreamSynchronize(c->stream), "round in flight");
  c->js = c->stream;
}

// Calls that change only the selected job's device state (domain, grids,
// source, tallies) run beside a round in flight, on the I/O stream, unless
// that round uses the job or the domain it owns.
void set_device_job(mdx_ctx* c) {
  set_device_nowait(c);
  c->js = c->stream;
  if (!c->inflight) return;
  const bool used = std::find(c->round_jobs.begin(), c->round_jobs.end(), c->cur) != c->round_jobs.end() ||
                    std::find(c->round_owners.begin(), c->round_owners.end(), c->cur) != c->round_owners.end();
  if (used) hip_check(hipStreamSynchronize(c->stream), "round in flight");
  else c->js = c->io_stream;
}

// Device-to-host copy on the I/O stream, which runs beside a round in flight.
// Clears bytes at p on stream st now (p holds doubles when bytes is a
// multiple of 8, 32-bit words otherwise).
void clear_now(void* p, size_t bytes, hipStream_t st) {
  if (bytes % 8 == 0) hip_check(mdx_launch::zero(static_cast<double*>(p), bytes / 8, st), "clear");
  else hip_check(mdx_launch::zero_u32(static_cast<uint32_t*>(p), bytes / 4, st), "clear");
}

// Clears a device buffer for a job's set-up: on the set-up stream, or queued
// for the next round while a graph round is in flight.
void clear_dev(mdx_ctx* c, void* p, size_t bytes) {
  if (!p || !bytes) return;
  if (c->use_graphs && c->inflight && (reinterpret_cast<uintptr_t>(p) & 15) == 0 && bytes % 4 == 0) {
    std::lock_guard<std::mutex> lk(g_clear_mu);
    int mine = 0;
    for (auto& r : g_clears) mine += r.ctx == c;
    if (mine < kRoundClearMax) {
      g_clears.push_back(ClearReq{c, p, bytes});
      return;
    }
  }
  clear_now(p, bytes, c->js);
}

// Takes this context's queued ranges off the queue.
std::vector<ClearReq> take_clears(mdx_ctx* c) {
  std::lock_guard<std::mutex> lk(g_clear_mu);
  std::vector<ClearReq> mine;
  for (auto& r : g_clears)
    if (r.ctx == c) mine.push_back(r);
  g_clears.erase(std::remove_if(g_clears.begin(), g_clears.end(), [&](const ClearReq& r) { return r.ctx == c; }),
                 g_clears.end());
  return mine;
}

// Clears this context's queued ranges on its transport stream (after a round
// in flight) and waits for them.
void flush_clears(mdx_ctx* c) {
  std::vector<ClearReq> mine = take_clears(c);
  if (mine.empty()) return;
  for (auto& r : mine) clear_now(r.p, r.bytes, c->stream);
  hip_check(hipStreamSynchronize(c->stream), "clear");
}

// flush_clears when [p, p + bytes) overlaps a queued range of this context.
void flush_clears_over(mdx_ctx* c, const void* p, size_t bytes) {
  bool hit = false;
  {
    std::lock_guard<std::mutex> lk(g_clear_mu);
    const char* a = static_cast<const char*>(p);
    for (auto& r : g_clears) {
      const char* b = static_cast<const char*>(r.p);
      if (r.ctx == c && b < a + bytes && a < b + r.bytes) hit = true;
    }
  }
  if (hit) flush_clears(c);
}

void read_back(mdx_ctx* c, void* dst, const void* src, size_t bytes, const char* what) {
  if (!bytes) return;
  flush_clears_over(c, src, bytes);
  hip_check(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->io_stream), what);
  hip_check(hipStreamSynchronize(c->io_stream), what);
}

// Local majorants.  Material k bounds material m when its total cross
// section is at least (1 + 1e-5) times m's at every node of either table in
// [kMacroElo, kMacroEhi] (both are log-log linear between their nodes, so the
// ratio's extremes lie on the nodes); the margin covers the FP32 evaluation.
constexpr double kMacroElo = 1.0e3, kMacroEhi = 5.0e5;

std::vector<uint32_t> dominance(const std::vector<std::unique_ptr<mdx::MaterialTables>>& mats) {
  const size_t n = std::min<size_t>(mats.size(), 32);
  std::vector<uint32_t> dom(n, 0);
  for (size_t a = 0; a < n; ++a)
    for (size_t b = 0; b < n; ++b) {
      std::vector<double> xs = {kMacroElo, kMacroEhi};
      for (const mdx::Table* t : {&mats[a]->total, &mats[b]->total})
        for (double x : t->x)
          if (x > kMacroElo && x < kMacroEhi) xs.push_back(x);
      bool ok = true;
      for (double x : xs)
        if (!(mats[a]->total_i(x) >= mats[b]->total_i(x) * (1.0 + 1e-5) || a == b)) {
          ok = false;
          break;
        }
      if (ok) dom[a] |= 1u << b;
    }
  return dom;
}

// Chebyshev distance (in cells; 65535 = none) from every cell to the nearest
// bad cell, and to the layer of cells around the grid when outside_bad: the
// two-pass chamfer transform with unit weights on the 26-neighbourhood, which
// is exact for this metric, on an array padded with that layer.
std::vector<uint16_t> cheb_distance(const std::vector<uint8_t>& bad, int nx, int ny, int nz, bool outside_bad) {
  constexpr uint16_t kNone = 65535;
  const int px = nx + 2, py = ny + 2, pz = nz + 2;
  const size_t sy = (size_t)px, sz = (size_t)px * py;
  std::vector<uint16_t> d((size_t)px * py * pz, outside_bad ? 0 : kNone);
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i)
        d[(i + 1) + sy * (j + 1) + sz * (k + 1)] = bad[(size_t)i + (size_t)nx * ((size_t)j + (size_t)ny * k)] ? 0 : kNone;
  std::vector<ptrdiff_t> back;  // the 13 neighbours before a cell in raster order
  for (int dk = -1; dk <= 1; ++dk)
    for (int dj = -1; dj <= 1; ++dj)
      for (int di = -1; di <= 1; ++di)
        if (dk * 9 + dj * 3 + di < 0) back.push_back(di + (ptrdiff_t)sy * dj + (ptrdiff_t)sz * dk);
  for (int k = 1; k <= nz; ++k)
    for (int j = 1; j <= ny; ++j)
      for (int i = 1; i <= nx; ++i) {
        const size_t c = i + sy * j + sz * k;
        int v = d[c];
        for (ptrdiff_t o : back) v = std::min(v, d[c + o] + 1);
        d[c] = (uint16_t)std::min(v, (int)kNone);
      }
  for (int k = nz; k >= 1; --k)
    for (int j = ny; j >= 1; --j)
      for (int i = nx; i >= 1; --i) {
        const size_t c = i + sy * j + sz * k;
        int v = d[c];
        for (ptrdiff_t o : back) v = std::min(v, d[c - o] + 1);
        d[c] = (uint16_t)std::min(v, (int)kNone);
      }
  std::vector<uint16_t> out((size_t)nx * ny * nz);
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) out[(size_t)i + (size_t)nx * ((size_t)j + (size_t)ny * k)] = d[(i + 1) + sy * (j + 1) + sz * (k + 1)];
  return out;
}

// The macro-cell table of grid G (one grid per domain): per cell the slot of
// the material that bounds every material present in the cell (the one with
// the smallest cross sections among those that do), and the Chebyshev radius
// in cells within which every cell holds only materials it bounds and, unless
// it bounds the domain's background material, lies inside the grid.
void build_macro(mdx_ctx* c, const Job& J, GridHost& G) {
  const size_t nc = (size_t)G.mc[0] * G.mc[1] * G.mc[2];
  std::vector<uint32_t> mask(nc, 0);
  for (size_t i = 0; i < nc; ++i)
    for (int q = 0; q < 4; ++q)
      for (uint64_t w = G.cell_ids[4 * i + q]; w; w &= w - 1) {
        const int id = q * 64 + __builtin_ctzll(w);
        const int slot = c->slot_of(id);
        mask[i] |= slot < 32 ? 1u << slot : 0x80000000u;
      }
  const std::vector<uint32_t>& dom = c->dominated;
  const int ns = (int)dom.size();
  std::vector<uint8_t> cls(nc, 255);
  for (size_t i = 0; i < nc; ++i) {
    const uint32_t m = mask[i];
    for (int k = 0; k < ns; ++k)
      if ((m >> k & 1u) && (m & ~dom[k]) == 0 && (cls[i] == 255 || (dom[cls[i]] >> k & 1u))) cls[i] = (uint8_t)k;
  }
  std::vector<uint16_t> word(nc, 255);
  std::vector<bool> used(ns, false);
  for (size_t i = 0; i < nc; ++i)
    if (cls[i] != 255) used[cls[i]] = true;
  // Outside the (only) grid lies the background material, or the outside of
  // the domain, where a photon escapes: a class that bounds the background
  // is not limited by the grid's faces.
  const int bg = c->slot_of(J.background_id);
  auto radius = [&](int k) {  // cells of class k only: one writer per entry
    std::vector<uint8_t> bad(nc);
    for (size_t i = 0; i < nc; ++i) bad[i] = (mask[i] & ~dom[k]) != 0;
    const bool outside_bad = bg >= 32 || !(dom[k] >> bg & 1u);
    const std::vector<uint16_t> d = cheb_distance(bad, G.mc[0], G.mc[1], G.mc[2], outside_bad);
    for (size_t i = 0; i < nc; ++i)  // radius = distance - 1, capped at 255
      if (cls[i] == k) word[i] = (uint16_t)(k | (std::min(std::max((int)d[i] - 1, 0), 255) << 8));
  };
  std::vector<std::thread> th;
  for (int k = 0; k < ns; ++k)
    if (used[k]) th.emplace_back(radius, k);
  for (auto& t : th) t.join();
  for (size_t i = 0; i < nc; ++i)  // one material in the cell: the voxel IDs need not be read
    if (cls[i] != 255 && __builtin_popcount(mask[i]) == 1 && mask[i] == 1u << cls[i]) word[i] |= MDX_MACRO_UNIFORM;
  G.macro = std::move(word);
  G.macro_gen = c->tables_gen;
  if (c->device >= 0) {
    G.d_macro.upload(G.macro.data(), nc, c->js);
    hip_check(hipStreamSynchronize(c->js), "macro upload");
  }
}

void upload_domain(mdx_ctx* c) {
  mdx_domain_dev D{};
  for (int k = 0; k < 3; ++k) D.dim_space[k] = (float)c->J().dim_space[k];
  D.bg_slot = c->tables_ready ? c->slot_of(c->J().background_id) : 0;
  D.n_grids = (int)c->J().grids.size();
  for (size_t g = 0; g < c->J().grids.size(); ++g) {
    GridHost& G = *c->J().grids[g];
    mdx_grid_dev& d = D.grids[g];
    size_t nv = (size_t)G.dims[0] * G.dims[1] * G.dims[2];
    for (int k = 0; k < 3; ++k) {
      d.origin[k] = (float)G.origin[k];
      d.spacing[k] = (float)G.spacing[k];
      d.inv_spacing[k] = 1.0f / d.spacing[k];
      d.dim[k] = (float)(G.dims[k] * G.spacing[k]);  // voxel_grid.cpp:175-177
      d.n[k] = G.dims[k];
    }
    d.ids = G.d_ids.as<uint8_t>();
    d.macro = nullptr;
    if (c->tables_ready && c->local_majorants && c->J().grids.size() == 1 && !c->dominated.empty()) {
      if (G.macro_gen != c->tables_gen) build_macro(c, c->J(), G);
      d.macro = G.d_macro.as<uint16_t>();
      d.mcx = G.mc[0];
      d.mcy = G.mc[1];
      d.mcell = (float)(MDX_MACRO_EDGE * std::min(G.spacing[0], std::min(G.spacing[1], G.spacing[2])));
    }
    if (c->J().voxel_dose) {
      // FP64: one record of MDX_DOSE_REC doubles per voxel (count inside);
      // FP32 layouts: the sums, and the counts in an array of their own
      const int mode = c->J().dose_mode;
      const size_t per = mode == MDX_DOSE_F64 ? MDX_DOSE_REC * sizeof(double) : (mode == MDX_DOSE_F32 ? 8 : 16);
      const size_t per_n = mode == MDX_DOSE_F64 ? 0 : sizeof(unsigned long long);
      if (G.d_dose.bytes != nv * per || G.dose_mode != mode || G.d_dose_n.bytes != nv * per_n) {
        G.d_dose.alloc(nv * per);
        G.d_dose_n.alloc(nv * per_n);
        G.dose_mode = mode;
        clear_dev(c, G.d_dose.p, nv * per);
        clear_dev(c, G.d_dose_n.p, nv * per_n);
      }
      d.dose_mode = mode;
      d.dose = mode == MDX_DOSE_F64 ? G.d_dose.as<double>() : nullptr;
      d.dose32 = mode == MDX_DOSE_F64 ? nullptr : G.d_dose.as<float>();
      d.dose_n = G.d_dose_n.as<unsigned long long>();
    } else {
      G.d_dose.release();
      G.d_dose_n.release();
      G.dose_mode = -1;
    }
  }
  if (!c->J().d_bg.p) {
    c->J().d_bg.alloc(3 * sizeof(double));
    clear_dev(c, c->J().d_bg.p, 3 * sizeof(double));
  }
  D.bg_dose = c->J().d_bg.as<double>();
  c->J().d_domain.upload(&D, 1, c->js);
  hip_check(hipStreamSynchronize(c->js), "upload domain");
}

void check_materials(mdx_ctx* c) {
  need(c->tables_ready, "physics tables not built (mdx_build_tables)");
  need(c->J().domain_set, "domain not set (mdx_set_domain / mdx_load_scene)");
  if (c->slot_of(c->J().background_id) == 255)
    throw std::out_of_range("background material " + std::to_string(c->J().background_id) + " has no tables");
  for (auto& g : c->J().grids) {
    const bool* seen = g->present;
    for (int v = 0; v < 256; ++v)
      if (seen[v] && c->slot_of(v) == 255)
        throw std::out_of_range("grid material " + std::to_string(v) + " has no tables");
  }
}

// The RITA tables of one material built on the device in FP64 (k_rita_build)
// and kept in its host tables, whose packing into FP32 rows follows.
void build_rita_device(mdx_ctx* c, mdx::MaterialTables& m) {
  mdx::Rita& R = m.rita;
  R.energies = m.incoh.x;  // material_data.cpp:70: the incoherent-CS grid
  const int ne = (int)R.energies.size();
  const int cap = mdx_launch::rita_cap();
  DBuf lx, ly, en, xs, ys, as, bs, cnt;
  lx.upload(m.ff_i.lx.data(), m.ff_i.lx.size(), c->stream);
  ly.upload(m.ff_i.ly.data(), m.ff_i.ly.size(), c->stream);
  en.upload(R.energies.data(), R.energies.size(), c->stream);
  const size_t cells = (size_t)ne * cap;
  xs.alloc(cells * 8);
  ys.alloc(cells * 8);
  as.alloc(cells * 8);
  bs.alloc(cells * 8);
  cnt.alloc((size_t)ne * 4);
  hip_check(mdx_launch::rita_build(lx.as<double>(), ly.as<double>(), (int)m.ff_i.lx.size(), mdx::alpha_inv_angstrom(),
                                   1.0 / mdx::kElectronRestMass, en.as<double>(), ne, xs.as<double>(), ys.as<double>(),
                                   as.as<double>(), bs.as<double>(), cnt.as<int>(), c->stream),
            "RITA build");
  std::vector<double> hx(cells), hy(cells), ha(cells), hb(cells);
  std::vector<int> hc(ne);
  hip_check(hipMemcpyAsync(hx.data(), xs.p, cells * 8, hipMemcpyDeviceToHost, c->stream), "read RITA");
  hip_check(hipMemcpyAsync(hy.data(), ys.p, cells * 8, hipMemcpyDeviceToHost, c->stream), "read RITA");
  hip_check(hipMemcpyAsync(ha.data(), as.p, cells * 8, hipMemcpyDeviceToHost, c->stream), "read RITA");
  hip_check(hipMemcpyAsync(hb.data(), bs.p, cells * 8, hipMemcpyDeviceToHost, c->stream), "read RITA");
  hip_check(hipMemcpyAsync(hc.data(), cnt.p, (size_t)ne * 4, hipMemcpyDeviceToHost, c->stream), "read RITA");
  hip_check(hipStreamSynchronize(c->stream), "RITA build");
  R.x.assign(ne, {});
  R.y.assign(ne, {});
  R.a.assign(ne, {});
  R.b.assign(ne, {});
  for (int ie = 0; ie < ne; ++ie) {
    const int n = hc[ie];
    if (n < 2) throw std::runtime_error("RITA table did not converge within " + std::to_string(cap) + " points");
    const size_t o = (size_t)ie * cap;
    R.x[ie].assign(hx.begin() + o, hx.begin() + o + n);
    R.y[ie].assign(hy.begin() + o, hy.begin() + o + n);
    R.a[ie].assign(ha.begin() + o, ha.begin() + o + n - 1);
    R.b[ie].assign(hb.begin() + o, hb.begin() + o + n - 1);
  }
}

void build_and_upload(mdx_ctx* c, const std::vector<int>& ids) {
  c->ensure_registry();
  need(!ids.empty(), "no materials given");
  need(ids.size() <= MDX_MAX_MATERIALS, "too many materials");
  std::vector<int> uniq;
  for (int id : ids)
    if (std::find(uniq.begin(), uniq.end(), id) == uniq.end()) uniq.push_back(id);
  int threads = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  std::vector<std::unique_ptr<mdx::MaterialTables>> mats;
  const bool rita_dev = c->device >= 0 && c->rita_on_device;
  if (rita_dev) set_device(c);
  for (int id : uniq) {
    auto it = c->registry.find(id);
    if (it == c->registry.end())
      throw std::runtime_error("material ID " + std::to_string(id) + " not in the registry");
    mats.push_back(std::make_unique<mdx::MaterialTables>(
        mdx::build_material(it->second, c->data_dir + "/epdl", !rita_dev, threads)));
    if (rita_dev) build_rita_device(c, *mats.back());
  }
  std::vector<const mdx::MaterialTables*> mp;
  for (auto& m : mats) mp.push_back(m.get());
  c->majorant = mdx::build_majorant(mp);
  mdx::pack_all(c->packed, mp, c->majorant, 1.0e3, 5.0e5);
  c->mats = std::move(mats);
  c->dominated = dominance(c->mats);
  ++c->tables_gen;
  if (c->device < 0) {  // host-only context: FP64 tables and packing, no upload
    c->tables_ready = true;
    return;
  }
  set_device(c);
  mdx::Packed& P = c->packed;
  c->d_xs.upload(P.xs.data(), P.xs.size(), c->stream);
  c->d_nodes.upload(P.nodes.data(), P.nodes.size(), c->stream);
  c->d_bins.upload(P.bins.data(), P.bins.size(), c->stream);
  c->d_rE.upload(P.rita_E.data(), P.rita_E.size(), c->stream);
  c->d_rinv.upload(P.rita_escale.data(), P.rita_escale.size(), c->stream);
  c->d_roff.upload(P.rita_row_off.data(), P.rita_row_off.size(), c->stream);
  c->d_rrows.upload(P.rita_rows.data(), P.rita_rows.size(), c->stream);
  c->d_rhash.upload(P.rita_ehash.data(), P.rita_ehash.size(), c->stream);
  c->d_rguide.upload(P.rita_guide.data(), P.rita_guide.size(), c->stream);
  mdx_tables_dev T{};
  T.xs = c->d_xs.as<float>();
  T.nodes = c->d_nodes.as<mdx_node>();
  T.bins = c->d_bins.as<mdx_node>();
  T.rita_E = c->d_rE.as<float>();
  T.rita_escale = c->d_rinv.as<float>();
  T.rita_row_off = c->d_roff.as<int32_t>();
  T.rita_rows = c->d_rrows.as<mdx_rita_row>();
  T.rita_ehash = c->d_rhash.as<int32_t>();
  T.rita_guide = c->d_rguide.as<uint8_t>();
  T.majorant = P.majorant;
  T.n_mat = (int)P.mats.size();
  for (size_t s = 0; s < P.mats.size(); ++s) T.mat[s] = P.mats[s];
  std::memcpy(T.id_to_slot, P.id_to_slot, 256);
  c->d_tables.upload(&T, 1, c->stream);
  hip_check(hipStreamSynchronize(c->stream), "upload tables");
  c->tables_ready = true;
  const int keep = c->cur;
  for (size_t j = 0; j < c->jobs.size(); ++j) {
    if (!c->jobs[j] || c->jobs[j]->domain_of >= 0) continue;
    c->cur = (int)j;
    if (c->J().domain_set) upload_domain(c);
  }
  c->cur = keep;
}

void upload_tallies(mdx_ctx* c) {
  mdx_tallies_dev TL{};
  TL.n = (int)c->J().tallies.size();
  for (size_t i = 0; i < c->J().tallies.size(); ++i) {
    TallyHost& H = *c->J().tallies[i];
    const mdx_tally& t = H.t;
    mdx_tally_dev& d = TL.t[i];
    d.kind = t.kind;
    for (int k = 0; k < 3; ++k) {
      d.p0[k] = (float)t.p0[k];
      d.p1[k] = (float)t.p1[k];
      d.p2[k] = (float)t.p2[k];
      d.nrm[k] = (float)t.normal[k];
    }
    d.radius = (float)t.radius;
    d.e1e1 = (float)(t.p1[0] * t.p1[0] + t.p1[1] * t.p1[1] + t.p1[2] * t.p1[2]);
    d.e2e2 = (float)(t.p2[0] * t.p2[0] + t.p2[1] * t.p2[1] + t.p2[2] * t.p2[2]);
    d.nx = t.nx;
    d.ny = t.ny;
    d.hist_bins = t.kind == MDX_TALLY_AACUBOID ? 0 : t.hist_bins;
    d.hist_lo = (float)t.hist_lo;
    d.hist_inv_w = t.hist_width > 0 ? (float)(1.0 / t.hist_width) : 0.0f;
    d.acc = H.acc.as<double>();
    d.hist = H.hist.as<double>();
    d.image = H.image.as<double>();
  }
  c->J().d_tallies.upload(&TL, 1, c->js);
  hip_check(hipStreamSynchronize(c->js), "upload tallies");
}

void zero_tallies(mdx_ctx* c) {
  for (auto& h : c->J().tallies) {
    clear_dev(c, h->acc.p, h->acc.bytes);
    clear_dev(c, h->hist.p, h->hist.bytes);
    clear_dev(c, h->image.p, h->image.bytes);
  }
  for (auto& g : c->J().grids) {
    clear_dev(c, g->d_dose.p, g->d_dose.bytes);
    clear_dev(c, g->d_dose_n.p, g->d_dose_n.bytes);
  }
  if (c->J().d_bg.p) clear_dev(c, c->J().d_bg.p, c->J().d_bg.bytes);
  hip_check(hipStreamSynchronize(c->js), "reset tallies");
}

const mdx::MaterialTables& host_material(mdx_ctx* c, int id) {
  need(c->tables_ready, "physics tables not built");
  for (auto& m : c->mats)
    if (m->def.id == id) return *m;
  throw std::out_of_range("material " + std::to_string(id) + " has no tables");
}

template <class T>
struct DevArray {
  DBuf b;
  DevArray(const T* src, size_t n, hipStream_t st) {
    if (src) b.upload(src, n, st);
    else b.alloc(n * sizeof(T));
  }
  T* p() { return b.as<T>(); }
};


constexpr size_t kMailCounters = MDX_NCOUNTER * sizeof(unsigned long long);
constexpr size_t kMailRoiOff = (kMailCounters + 255) & ~size_t(255);
// The mailbox's parts, at their host (dev = false) or device address.
mdx_roi_req* mail_roi_req(mdx_ctx* c, bool dev) {
  return reinterpret_cast<mdx_roi_req*>(static_cast<char*>(dev ? c->d_mail : c->h_mail) + kMailRoiOff);
}
double* mail_roi_out(mdx_ctx* c, bool dev) { return reinterpret_cast<double*>(mail_roi_req(c, dev) + c->mail_roi); }

// The mailbox with room for n ROI boxes (grown by reallocation: the device
// is idle then, as mdx_roi_stats requires).
void ensure_mail(mdx_ctx* c, size_t n) {
  if (c->h_mail && n <= c->mail_roi) return;
  const size_t cap = std::max<size_t>(n, 2 * c->mail_roi);
  const size_t bytes = kMailRoiOff + cap * (sizeof(mdx_roi_req) + 3 * sizeof(double));
  void* p = nullptr;
  void* dp = nullptr;
  hip_check(hipHostMalloc(&p, bytes, hipHostMallocDefault), "hipHostMalloc");
  const hipError_t e = hipHostGetDevicePointer(&dp, p, 0);
  if (e != hipSuccess) {
    (void)hipHostFree(p);
    hip_check(e, "hipHostGetDevicePointer");
  }
  if (c->h_mail) (void)hipHostFree(c->h_mail);
  c->h_mail = p;
  c->d_mail = dp;
  c->mail_roi = cap;
}

void read_stats(mdx_ctx* c, mdx_stats* stats) {
  if (!stats) return;
  std::memset(stats, 0, sizeof *stats);
  const unsigned long long* h = static_cast<const unsigned long long*>(c->h_mail);
  hip_check(mdx_launch::copy_u64(c->d_counters.as<unsigned long long>(), static_cast<unsigned long long*>(c->d_mail),
                                 MDX_NCOUNTER, c->stream),
            "counters");
  hip_check(hipStreamSynchronize(c->stream), "counters");
  for (int i = 0; i < MDX_NCOUNTER && i < MDX_STAT_NCOUNTERS; ++i) stats->counters[i] = h[i];
  float ms = 0;
  hip_check(hipEventElapsedTime(&ms, c->ev0, c->ev1), "hipEventElapsedTime");
  stats->kernel_ms = ms;
}

// Restores the selected job when a scope that switches it ends (also on throw).
struct KeepCur {
  mdx_ctx* c;
  int keep;
  explicit KeepCur(mdx_ctx* ctx) : c(ctx), keep(ctx->cur) {}
  ~KeepCur() { c->cur = keep; }
};

Job& job_at(mdx_ctx* c, int id) {
  if (id < 0 || id >= (int)c->jobs.size() || !c->jobs[id]) throw std::out_of_range("no such job");
  return *c->jobs[id];
}

// The job whose domain a job uses (itself unless it aliases another's).
int domain_owner(mdx_ctx* c, int id) {
  int d = job_at(c, id).domain_of;
  if (d < 0) return id;
  job_at(c, d);
  return d;
}

// Fills the device job records of one launch and validates every job.
// The bank kernel keeps 42 bits of a history index (MDX_BANK_MAX_HISTORY).
void check_bank_range(const mdx_ctx* c, uint64_t first, uint64_t n) {
  if (c->event_kernel == MDX_EVENT_BANK && (first > MDX_BANK_MAX_HISTORY || n > MDX_BANK_MAX_HISTORY - first))
    throw std::out_of_range("history indices of the bank event kernel end at 2^42 (first + n)");
}

uint64_t fill_round(mdx_ctx* c, int n_jobs, const int* ids, const uint64_t* first, const uint64_t* n,
                    const uint64_t* seed, mdx_job_dev* jd, uint64_t* prefix) {
  KeepCur kc(c);
  prefix[0] = 0;
  c->round_jobs.assign(ids, ids + n_jobs);
  c->round_owners.clear();
  for (int i = 0; i < n_jobs; ++i) {
    job_at(c, ids[i]);
    c->cur = domain_owner(c, ids[i]);
    c->round_owners.push_back(c->cur);
    check_materials(c);
    const mdx_domain_dev* D = c->J().d_domain.as<mdx_domain_dev>();
    c->cur = ids[i];
    need(c->J().source_set, "source not set (mdx_set_source)");
    if (!c->J().d_tallies.p) upload_tallies(c);
    jd[i].D = D;
    jd[i].S = c->J().d_source.as<mdx_source_dev>();
    jd[i].TL = c->J().d_tallies.as<mdx_tallies_dev>();
    check_bank_range(c, first[i], n[i]);
    jd[i].first = first[i];
    jd[i].n = n[i];
    jd[i].seed = seed[i];
    jd[i].stream = 0;
    bool volume = false;
    for (auto& t : c->J().tallies) volume = volume || t->t.kind == MDX_TALLY_AACUBOID;
    jd[i].flights = c->J().flights && !volume ? 1u : 0u;
    prefix[i + 1] = prefix[i] + n[i];
  }
  return prefix[n_jobs];
}

// Workgroups of a graph round: the queue kernel's persistent blocks fill
// every slot a CU has, so a few slots (one per XCD of the MI355X) are left
// free for what the next jobs' set-up puts on the I/O stream meanwhile
// (small copies and kernels), which would otherwise start only when the round
// ends.
int round_blocks(const mdx_ctx* c) { return std::max(1, c->num_cu * c->queue_blocks_per_cu - 8); }

// Captures the graph of one multi-job round: copy of the pinned round buffer,
// counter zeroing (kernels, no memset nodes) and the queue kernel, which reads
// the job count and total from the copied header.
void build_graph(mdx_ctx* c) {
  if (c->graph) {
    (void)hipGraphExecDestroy(c->graph);
    c->graph = nullptr;
  }
  if (!c->h_round) hip_check(hipHostMalloc(&c->h_round, kRoundBytes, hipHostMallocDefault), "hipHostMalloc");
  if (!c->d_round.p) c->d_round.alloc(kRoundBytes);
  char* d = c->d_round.as<char>();
  hipGraph_t g;
  hip_check(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal), "begin capture");
  hip_check(hipMemcpyAsync(d, c->h_round, kRoundBytes, hipMemcpyHostToDevice, c->stream), "capture copy");
  hip_check(mdx_launch::zero_list(reinterpret_cast<const uint64_t*>(d + kRoundClearOff), c->stream), "capture clear");
  hip_check(mdx_launch::zero_u64(c->d_counters.as<unsigned long long>(), MDX_NCOUNTER, c->stream), "capture zero");
  hip_check(mdx_launch::zero_u64(c->d_work.as<unsigned long long>(), kRoundMax, c->stream), "capture zero");
  if (c->event_kernel == MDX_EVENT_BANK)
    hip_check(mdx_launch::transport_bank_round(c->d_tables.as<mdx_tables_dev>(), d, c->d_work.as<unsigned long long>(),
                                               c->d_counters.as<unsigned long long>(), round_blocks(c), c->refill_min,
                                               c->interact_min, c->bank_size, c->stream),
              "capture kernel");
  else
    hip_check(mdx_launch::transport_queue_round(c->d_tables.as<mdx_tables_dev>(), d,
                                                c->d_work.as<unsigned long long>(),
                                                c->d_counters.as<unsigned long long>(), round_blocks(c),
                                                c->refill_min, c->interact_min, c->stream),
              "capture kernel");
  hip_check(hipStreamEndCapture(c->stream, &g), "end capture");
  hip_check(hipGraphInstantiate(&c->graph, g, nullptr, nullptr, 0), "hipGraphInstantiate");
  (void)hipGraphDestroy(g);
  c->graph_params[0] = c->queue_blocks_per_cu;
  c->graph_params[1] = c->refill_min;
  c->graph_params[2] = c->interact_min;
  c->graph_params[3] = c->event_kernel;
  c->graph_params[4] = c->bank_size;
}

void finish_round(mdx_ctx* c, mdx_stats* stats) {
  hip_check(hipStreamSynchronize(c->stream), "transport");
  c->inflight = false;
  read_stats(c, stats);
}

// The selected event kernel over job records already on the device (the
// lane-resident queue kernel for the per-block LDS dose mode, which the bank
// kernel does not keep).
hipError_t launch_events(mdx_ctx* c, const mdx_job_dev* jobs, const uint64_t* prefix, int n_jobs, uint64_t total,
                         unsigned long long* work, unsigned long long* cnt, uint32_t lds_voxels) {
  if (c->event_kernel == MDX_EVENT_BANK && lds_voxels == 0)
    return mdx_launch::transport_bank(c->d_tables.as<mdx_tables_dev>(), jobs, prefix, n_jobs, total, work, cnt,
                                      c->num_cu * c->queue_blocks_per_cu, c->refill_min, c->interact_min,
                                      c->bank_size, c->stream);
  return mdx_launch::transport_queue(c->d_tables.as<mdx_tables_dev>(), jobs, prefix, n_jobs, total, work, cnt,
                                     c->num_cu * c->queue_blocks_per_cu, c->refill_min, c->interact_min, lds_voxels,
                                     c->stream);
}

// One launch of the queue kernel over the listed jobs' history slices;
// with wait = false it returns once the round is queued (mdx_wait ends it).
void run_jobs(mdx_ctx* c, int n_jobs, const int* ids, const uint64_t* first, const uint64_t* n,
              const uint64_t* seed, mdx_stats* stats, bool wait = true) {
  need(c->tables_ready, "physics tables not built (mdx_build_tables)");
  need(!c->inflight, "a round is in flight (mdx_wait first)");
  set_device(c);
  unsigned long long* cnt = c->d_counters.as<unsigned long long>();
  if (c->use_graphs) {
    need(n_jobs <= kRoundMax, "too many jobs in one graph round (max 1024)");
    if (!c->graph || c->graph_params[0] != c->queue_blocks_per_cu || c->graph_params[1] != c->refill_min ||
        c->graph_params[2] != c->interact_min || c->graph_params[3] != c->event_kernel ||
        c->graph_params[4] != c->bank_size)
      build_graph(c);
    char* h = static_cast<char*>(c->h_round);
    uint64_t* hdr = reinterpret_cast<uint64_t*>(h);
    uint64_t* prefix = reinterpret_cast<uint64_t*>(h + kRoundHeader);
    mdx_job_dev* jd = reinterpret_cast<mdx_job_dev*>(h + kRoundHeader + kRoundPrefix);
    const uint64_t total = fill_round(c, n_jobs, ids, first, n, seed, jd, prefix);
    hdr[0] = (uint64_t)n_jobs;
    hdr[1] = total;
    uint64_t* cl = reinterpret_cast<uint64_t*>(h + kRoundClearOff);
    std::vector<ClearReq> clears = take_clears(c);
    cl[0] = clears.size();
    for (size_t k = 0; k < clears.size(); ++k) {
      cl[1 + 2 * k] = reinterpret_cast<uint64_t>(clears[k].p);
      cl[2 + 2 * k] = clears[k].bytes;
    }
    hip_check(hipEventRecord(c->ev0, c->stream), "hipEventRecord");
    hip_check(hipGraphLaunch(c->graph, c->stream), "hipGraphLaunch");
    hip_check(hipEventRecord(c->ev1, c->stream), "hipEventRecord");
    c->inflight = true;
    if (wait) finish_round(c, stats);
    return;
  }
  flush_clears(c);
  std::vector<mdx_job_dev> jd(n_jobs);
  std::vector<uint64_t> prefix(n_jobs + 1, 0);
  const uint64_t total = fill_round(c, n_jobs, ids, first, n, seed, jd.data(), prefix.data());
  uint32_t lds_voxels = 0;  // MDX_DOSE_F32C_LDS: grid 0 summed per block in LDS (single-job launches)
  if (n_jobs == 1) {
    const Job& O = job_at(c, domain_owner(c, ids[0]));
    if (O.voxel_dose && O.dose_mode == MDX_DOSE_F32C_LDS && !O.grids.empty()) lds_voxels = (uint32_t)O.grids[0]->voxels();
  }
  c->d_jobs.upload(jd.data(), jd.size(), c->stream);
  c->d_prefix.upload(prefix.data(), prefix.size(), c->stream);
  hip_check(mdx_launch::zero_u64(cnt, MDX_NCOUNTER, c->stream), "zero counters");
  if (c->d_work.bytes < (size_t)n_jobs * sizeof(unsigned long long)) c->d_work.alloc(n_jobs * sizeof(unsigned long long));
  hip_check(mdx_launch::zero_u64(c->d_work.as<unsigned long long>(), std::max(n_jobs, 1), c->stream), "zero work");
  hip_check(hipEventRecord(c->ev0, c->stream), "hipEventRecord");
  if (total > 0)
    hip_check(launch_events(c, c->d_jobs.as<mdx_job_dev>(), c->d_prefix.as<uint64_t>(), n_jobs, total,
                            c->d_work.as<unsigned long long>(), cnt, lds_voxels),
              "queue launch");
  hip_check(hipEventRecord(c->ev1, c->stream), "hipEventRecord");
  c->inflight = true;
  if (wait) finish_round(c, stats);
}

}  // namespace

extern "C" {

const char* mdx_last_error(void) { return g_last_error.c_str(); }

int mdx_create(int device, mdx_ctx** out) {
  return guard([&] {
    need(out != nullptr, "out is NULL");
    *out = nullptr;
    if (device == -1) {  // host-only context: registry and FP64 tables only
      auto c = std::make_unique<mdx_ctx>();
      c->device = -1;
      c->data_dir = default_data_dir();
      c->jobs.push_back(std::make_unique<Job>(&c->pool));
      *out = c.release();
      return;
    }
    int n = 0;
    hip_check(hipGetDeviceCount(&n), "hipGetDeviceCount");
    if (device < 0 || device >= n) throw std::out_of_range("device index out of range");
    auto c = std::make_unique<mdx_ctx>();
    c->device = device;
    c->data_dir = default_data_dir();
    c->jobs.push_back(std::make_unique<Job>(&c->pool));
    set_device(c.get());
    hip_check(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking), "hipStreamCreate");
    hip_check(hipStreamCreateWithFlags(&c->io_stream, hipStreamNonBlocking), "hipStreamCreate");
    hip_check(hipStreamCreateWithFlags(&c->copy_stream, hipStreamNonBlocking), "hipStreamCreate");
    c->js = c->stream;
    hip_check(hipEventCreate(&c->ev0), "hipEventCreate");
    hip_check(hipEventCreate(&c->ev1), "hipEventCreate");
    hipDeviceProp_t prop;
    hip_check(hipGetDeviceProperties(&prop, device), "hipGetDeviceProperties");
    c->num_cu = prop.multiProcessorCount;
    c->pool.limit = prop.totalGlobalMem / 4;
    c->lds_optin = std::max(prop.sharedMemPerBlockOptin, prop.sharedMemPerBlock);
    c->d_counters.alloc(MDX_NCOUNTER * sizeof(unsigned long long));
    c->d_work.alloc(kRoundMax * sizeof(unsigned long long));
    hip_check(mdx_launch::scratch_address(device, &c->scratch), "hipGetSymbolAddress(scratch)");
    ensure_mail(c.get(), 256);
    *out = c.release();
  });
}

int mdx_destroy(mdx_ctx* c) {
  return guard([&] {
    if (!c) return;
    if (c->device < 0) {
      delete c;
      return;
    }
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    (void)hipStreamSynchronize(c->io_stream);
    (void)hipStreamSynchronize(c->copy_stream);
    hipStream_t st = c->stream, io = c->io_stream, cp = c->copy_stream;
    hipEvent_t e0 = c->ev0, e1 = c->ev1;
    if (c->graph) (void)hipGraphExecDestroy(c->graph);
    void* pinned = c->h_round;
    void* mail = c->h_mail;
    delete c;
    if (pinned) (void)hipHostFree(pinned);
    if (mail) (void)hipHostFree(mail);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipStreamDestroy(st);
    (void)hipStreamDestroy(io);
    (void)hipStreamDestroy(cp);
  });
}

int mdx_set_data_dir(mdx_ctx* c, const char* dir) {
  return guard([&] {
    need(c && dir, "NULL argument");
    c->data_dir = dir;
    c->registry_loaded = false;
    c->registry.clear();
  });
}

int mdx_material_id(mdx_ctx* c, const char* name, int* id) {
  return guard([&] {
    need(c && name && id, "NULL argument");
    c->ensure_registry();
    for (auto& kv : c->registry)
      if (kv.second.name == name) {
        *id = kv.first;
        return;
      }
    throw std::runtime_error(std::string("unknown material name '") + name + "'");
  });
}

int mdx_air_muen(mdx_ctx* c, int n, const double* energy, double* out) {
  return guard([&] {
    need(c && (n == 0 || (energy && out)) && n >= 0, "bad argument");
    c->ensure_air_muen();
    const auto& X = c->muen_e;
    const auto& Y = c->muen_v;
    for (int i = 0; i < n; ++i) {
      const double e = energy[i];
      if (e <= X.front()) { out[i] = Y.front(); continue; }
      if (e >= X.back()) { out[i] = Y.back(); continue; }
      const size_t k = std::upper_bound(X.begin(), X.end(), e) - X.begin();  // above an edge's repeated energy
      const double x0 = X[k - 1], x1 = X[k], y0 = Y[k - 1], y1 = Y[k];
      if (e == x0) out[i] = y0;
      else if (y0 > 0 && y1 > 0)  // log-log linear, as the reference's tables
        out[i] = std::exp(std::log(y0) + (std::log(y1) - std::log(y0)) * (std::log(e) - std::log(x0)) /
                                             (std::log(x1) - std::log(x0)));
      else  // across the zeros below the first tabulated absorption
        out[i] = y0 + (y1 - y0) * (e - x0) / (x1 - x0);
    }
  });
}

int mdx_set_air_muen_table(mdx_ctx* c, int table) {
  return guard([&] {
    need(c, "NULL context");
    need(table == 0 || table == 1, "unknown air mu_en table");
    c->muen_table = table;
    c->muen_e.clear();
    c->muen_v.clear();
  });
}

int mdx_material_name(mdx_ctx* c, int id, char* buf, int buflen) {
  return guard([&] {
    need(c && buf && buflen > 0, "NULL argument");
    c->ensure_registry();
    auto it = c->registry.find(id);
    if (it == c->registry.end()) throw std::runtime_error("unknown material ID " + std::to_string(id));
    std::snprintf(buf, (size_t)buflen, "%s", it->second.name.c_str());
  });
}

int mdx_build_tables(mdx_ctx* c, const int* ids, int n) {
  return guard([&] {
    need(c && ids && n > 0, "bad material list");
    build_and_upload(c, std::vector<int>(ids, ids + n));
  });
}

int mdx_build_tables_for_domain(mdx_ctx* c) {
  return guard([&] {
    need(c && c->J().domain_set, "domain not set");
    std::vector<int> ids{c->J().background_id};
    for (auto& g : c->J().grids) {
      const bool* seen = g->present;
      for (int v = 0; v < 256; ++v)
        if (seen[v] && std::find(ids.begin(), ids.end(), v) == ids.end()) ids.push_back(v);
    }
    build_and_upload(c, ids);
  });
}

int mdx_set_domain(mdx_ctx* c, const double dim_space[3], int background_id) {
  return guard([&] {
    need(c && dim_space, "NULL argument");
    need(background_id >= 0 && background_id < 256, "background material ID must be in [0, 255]");
    if (c->device >= 0) set_device_job(c);  // before any buffer of the job is replaced
    for (int k = 0; k < 3; ++k) {
      need(dim_space[k] > 0, "dim_space must be positive");
      c->J().dim_space[k] = dim_space[k];
    }
    c->J().background_id = background_id;
    c->J().grids.clear();
    c->J().domain_of = -1;
    c->J().domain_set = true;
    if (c->device >= 0) upload_domain(c);
  });
}

int mdx_add_grid(mdx_ctx* c, const uint8_t* ids, const int dims[3], const double spacing[3],
                 const double origin[3]) {
  return guard([&] {
    need(c && ids && dims && spacing && origin, "NULL argument");
    need(c->J().domain_set, "mdx_set_domain must come first");
    need(c->J().domain_of < 0, "the selected job uses another job's domain");
    need((int)c->J().grids.size() < MDX_MAX_GRIDS, "too many grids");
    auto g = std::make_unique<GridHost>(&c->pool);
    size_t nv = 1;
    for (int k = 0; k < 3; ++k) {
      need(dims[k] > 0 && spacing[k] > 0, "grid dims and spacing must be positive");
      g->dims[k] = dims[k];
      g->spacing[k] = spacing[k];
      g->origin[k] = origin[k];
      nv *= (size_t)dims[k];
    }
    if (c->device >= 0) {  // straight from the caller's memory, which it may reuse on return
      set_device_job(c);
      g->d_ids.upload(ids, nv, c->js);
      scan_cells(ids, dims, g->cell_ids, g->mc, g->present);  // while the copy runs
      hip_check(hipStreamSynchronize(c->js), "upload grid");
    } else {
      scan_cells(ids, dims, g->cell_ids, g->mc, g->present);
      g->ids.assign(ids, ids + nv);
    }
    c->J().grids.push_back(std::move(g));
    if (c->device >= 0) upload_domain(c);
  });
}

int mdx_load_scene(mdx_ctx* c, const char* json_path) {
  return guard([&] {
    need(c && json_path, "NULL argument");
    mdx::Scene sc = mdx::read_scene(json_path);
    c->ensure_registry();
    int bg = -1;
    for (auto& kv : c->registry)
      if (kv.second.name == sc.background_material_name) bg = kv.first;
    if (bg < 0) throw std::runtime_error("unknown background material '" + sc.background_material_name + "'");
    std::vector<mdx::NiftiGrid> grids;
    for (auto& g : sc.grids) grids.push_back(mdx::read_nifti(g.path));
    double ds[3] = {sc.dim_space[0], sc.dim_space[1], sc.dim_space[2]};
    for (int k = 0; k < 3; ++k) need(ds[k] > 0, "dim_space must be positive");
    if (c->device >= 0) set_device_job(c);  // before any buffer of the job is replaced
    c->J().dim_space[0] = ds[0];
    c->J().dim_space[1] = ds[1];
    c->J().dim_space[2] = ds[2];
    c->J().background_id = bg;
    c->J().grids.clear();
    c->J().domain_of = -1;
    c->J().domain_set = true;
    for (size_t i = 0; i < grids.size(); ++i) {
      need(i < MDX_MAX_GRIDS, "too many grids");
      auto g = std::make_unique<GridHost>(&c->pool);
      for (int k = 0; k < 3; ++k) {
        g->dims[k] = grids[i].dims[k];
        g->spacing[k] = grids[i].spacing[k];
        g->origin[k] = sc.grids[i].origin[k];
      }
      g->ids = std::move(grids[i].ids);
      scan_cells(g->ids.data(), g->dims, g->cell_ids, g->mc, g->present);
      if (c->device >= 0) {
        set_device_job(c);
        g->d_ids.upload(g->ids.data(), g->ids.size(), c->js);
      }
      c->J().grids.push_back(std::move(g));
    }
    if (c->device >= 0) upload_domain(c);
  });
}

int mdx_num_grids(mdx_ctx* c, int* n) {
  return guard([&] {
    need(c && n, "NULL argument");
    *n = (int)c->DJ().grids.size();
  });
}

int mdx_grid_info(mdx_ctx* c, int grid, int dims[3], double spacing[3], double origin[3]) {
  return guard([&] {
    need(c, "NULL context");
    Job& D = c->DJ();
    if (grid < 0 || grid >= (int)D.grids.size()) throw std::out_of_range("grid index out of range");
    for (int k = 0; k < 3; ++k) {
      if (dims) dims[k] = D.grids[grid]->dims[k];
      if (spacing) spacing[k] = D.grids[grid]->spacing[k];
      if (origin) origin[k] = D.grids[grid]->origin[k];
    }
  });
}

int mdx_get_grid_ids(mdx_ctx* c, int grid, uint8_t* out) {
  return guard([&] {
    need(c && out, "NULL argument");
    Job& D = c->DJ();
    if (grid < 0 || grid >= (int)D.grids.size()) throw std::out_of_range("grid index out of range");
    const GridHost& g = *D.grids[grid];
    if (!g.ids.empty()) {
      std::memcpy(out, g.ids.data(), g.ids.size());
    } else {
      set_device_nowait(c);
      read_back(c, out, g.d_ids.p, g.voxels(), "read grid");
    }
  });
}

int mdx_set_source(mdx_ctx* c, const mdx_source* s) {
  return guard([&] {
    need(c && s, "NULL argument");
    mdx_source_dev d{};
    d.dir_kind = s->direction;
    d.spec_kind = s->spectrum;
    for (int k = 0; k < 3; ++k) d.pos[k] = (float)s->position[k];
    switch (s->direction) {
      case MDX_DIR_BEAM:
        for (int k = 0; k < 3; ++k) d.a[k] = (float)s->a[k];
        break;
      case MDX_DIR_ISOTROPIC:
        break;
      case MDX_DIR_RECT_ISOTROPIC: {
        double dot = s->b[0] * s->c[0] + s->b[1] * s->c[1] + s->b[2] * s->c[2];
        if (std::fabs(dot) >= 1e-9) throw std::invalid_argument("Edges are not orthogonal");  // source.cpp:134-143
        for (int k = 0; k < 3; ++k) {
          d.a[k] = (float)s->a[k];
          d.b[k] = (float)s->b[k];
          d.c[k] = (float)s->c[k];
        }
        break;
      }
      case MDX_DIR_DISC_ISOTROPIC: {
        // source.cpp:117-131: u = normal x x_hat (x_hat -> y_hat if normal _|_ x_hat), v = normal x u
        const double* n = s->b;
        double u[3], v[3];
        if (std::fabs(n[0]) < 1e-9) {
          u[0] = n[1] * 0 - n[2] * 1; u[1] = n[2] * 0 - n[0] * 0; u[2] = n[0] * 1 - n[1] * 0;
        } else {
          u[0] = 0; u[1] = n[2]; u[2] = -n[1];
        }
        double lu = std::sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
        need(lu > 0, "disc normal must be non-zero");
        for (double& x : u) x /= lu;
        v[0] = n[1] * u[2] - n[2] * u[1];
        v[1] = n[2] * u[0] - n[0] * u[2];
        v[2] = n[0] * u[1] - n[1] * u[0];
        double lv = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        for (double& x : v) x /= lv;
        for (int k = 0; k < 3; ++k) {
          d.a[k] = (float)s->a[k];
          d.b[k] = (float)u[k];
          d.c[k] = (float)v[k];
        }
        d.radius = (float)s->radius;
        break;
      }
      default:
        throw std::invalid_argument("unknown directionality");
    }
    const bool dev = c->device >= 0;  // a host-only context validates and keeps the source
    if (dev) set_device_job(c);
    if (s->spectrum == MDX_SPEC_MONO) {
      need(s->energy > 0, "energy must be positive");
      d.energy = (float)s->energy;
    } else if (s->spectrum == MDX_SPEC_POLY) {
      need(s->n_bins > 0 && s->bin_energy && s->bin_weight, "polyenergetic spectrum needs bins");
      need(s->n_bins <= 65535, "polyenergetic spectrum has more than 65535 bins");
      // probability_dist.cpp:11-60: normalise, cumulative sum
      double sum = 0;
      for (int i = 0; i < s->n_bins; ++i) sum += s->bin_weight[i];
      need(sum > 0, "spectrum weights sum to zero");
      std::vector<float> cdf(s->n_bins), en(s->n_bins);
      double acc = 0;
      for (int i = 0; i < s->n_bins; ++i) {
        acc += sum != 1.0 ? s->bin_weight[i] / sum : s->bin_weight[i];
        cdf[i] = (float)acc;
        en[i] = (float)s->bin_energy[i];
      }
      if (dev) {
        c->J().d_cdf.upload(cdf.data(), cdf.size(), c->js);
        c->J().d_energies.upload(en.data(), en.size(), c->js);
      }
      d.n_bins = s->n_bins;
      for (int j = 0; j < MDX_SPEC_GUIDE; ++j) {  // the device scales u by a power of two, exactly
        const float t = (float)j / (float)MDX_SPEC_GUIDE;
        int k = 0;
        while (k < s->n_bins && cdf[k] < t) ++k;
        d.guide[j] = (uint16_t)k;
      }
      d.cdf = c->J().d_cdf.as<float>();
      d.energies = c->J().d_energies.as<float>();
    } else {
      throw std::invalid_argument("unknown spectrum kind");
    }
    c->J().src = d;
    if (dev) {
      c->J().d_source.upload(&d, 1, c->js);
      hip_check(hipStreamSynchronize(c->js), "upload source");
    }
    c->J().source_set = true;
  });
}

int mdx_set_tallies(mdx_ctx* c, const mdx_tally* t, int n) {
  return guard([&] {
    need(c && (t || n == 0), "NULL argument");
    need(n >= 0 && n <= MDX_MAX_TALLIES, "too many tallies");
    std::vector<std::unique_ptr<TallyHost>> v;
    const bool dev = c->device >= 0;  // a host-only context validates and keeps the tallies
    if (dev) set_device_job(c);
    for (int i = 0; i < n; ++i) {
      auto h = std::make_unique<TallyHost>(&c->pool);
      h->t = t[i];
      mdx_tally& q = h->t;
      if (q.kind == MDX_TALLY_RECT || q.kind == MDX_TALLY_IMAGE) {
        double dot = q.p1[0] * q.p2[0] + q.p1[1] * q.p2[1] + q.p1[2] * q.p2[2];
        if (std::fabs(dot) >= 1e-9) throw std::invalid_argument("Edges are not orthogonal");  // surface_tally.cpp:80-84
        double nx = q.p1[1] * q.p2[2] - q.p1[2] * q.p2[1];
        double ny = q.p1[2] * q.p2[0] - q.p1[0] * q.p2[2];
        double nz = q.p1[0] * q.p2[1] - q.p1[1] * q.p2[0];
        double l = std::sqrt(nx * nx + ny * ny + nz * nz);
        need(l > 0, "degenerate rectangle");
        q.normal[0] = nx / l;
        q.normal[1] = ny / l;
        q.normal[2] = nz / l;
        if (q.kind == MDX_TALLY_IMAGE) need(q.nx > 0 && q.ny > 0, "image needs pixels");
      } else if (q.kind == MDX_TALLY_DISC) {
        need(q.radius > 0, "disc radius must be positive");
      } else if (q.kind != MDX_TALLY_AACUBOID) {
        throw std::invalid_argument("unknown tally kind");
      }
      if (q.kind != MDX_TALLY_AACUBOID && q.hist_bins > 0) need(q.hist_width > 0, "histogram width must be positive");
      if (dev) {
        h->acc.alloc((size_t)MDX_ACC_COPIES * MDX_TALLY_ACC_SIZE * sizeof(double));
        if (q.kind != MDX_TALLY_AACUBOID && q.hist_bins > 0) h->hist.alloc((size_t)4 * q.hist_bins * sizeof(double));
        if (q.kind == MDX_TALLY_IMAGE) h->image.alloc((size_t)3 * q.nx * q.ny * sizeof(double));
      }
      v.push_back(std::move(h));
    }
    c->J().tallies = std::move(v);
    if (dev) {
      upload_tallies(c);
      zero_tallies(c);
    }
  });
}

int mdx_enable_voxel_dose(mdx_ctx* c, int enable) {
  return guard([&] {
    need(c, "NULL context");
    need(c->J().domain_of < 0, "voxel dose belongs to the job that owns the domain");
    c->J().voxel_dose = enable != 0;
    if (c->device < 0) return;
    set_device_job(c);
    if (c->J().domain_set) upload_domain(c);
  });
}

int mdx_reset_tallies(mdx_ctx* c) {
  return guard([&] {
    need(c, "NULL context");
    set_device_job(c);
    zero_tallies(c);
  });
}

int mdx_run(mdx_ctx* c, uint64_t n, uint64_t seed, uint64_t first, int kernel, mdx_stats* stats) {
  return guard([&] {
    need(c, "NULL context");
    need(kernel == MDX_KERNEL_PLAIN || kernel == MDX_KERNEL_QUEUE, "unknown kernel");
    need(!c->inflight, "a round is in flight (mdx_wait first)");
    if (kernel == MDX_KERNEL_QUEUE) {
      int job = c->cur;
      run_jobs(c, 1, &job, &first, &n, &seed, stats);
      return;
    }
    const mdx_domain_dev* D;
    {
      KeepCur kc(c);
      c->cur = domain_owner(c, c->cur);
      check_materials(c);
      D = c->J().d_domain.as<mdx_domain_dev>();
    }
    need(c->J().source_set, "source not set (mdx_set_source)");
    set_device(c);
    flush_clears(c);
    if (!c->J().d_tallies.p) upload_tallies(c);
    unsigned long long* cnt = c->d_counters.as<unsigned long long>();
    hip_check(mdx_launch::zero_u64(cnt, MDX_NCOUNTER, c->stream), "zero counters");
    hip_check(hipEventRecord(c->ev0, c->stream), "hipEventRecord");
    if (n > 0) {
      hip_check(mdx_launch::transport_plain(c->d_tables.as<mdx_tables_dev>(), D,
                                            c->J().d_source.as<mdx_source_dev>(), c->J().d_tallies.as<mdx_tallies_dev>(),
                                            first, n, seed, 0u, cnt, c->num_cu * 8, c->stream),
                "transport launch");
    }
    hip_check(hipEventRecord(c->ev1, c->stream), "hipEventRecord");
    hip_check(hipStreamSynchronize(c->stream), "transport");
    read_stats(c, stats);
  });
}

int mdx_set_queue_params(mdx_ctx* c, int blocks_per_cu, int refill_min, int interact_min) {
  return guard([&] {
    need(c, "NULL context");
    need(blocks_per_cu >= 1 && blocks_per_cu <= 64, "blocks_per_cu must be in [1, 64]");
    need(refill_min >= 1 && refill_min <= 64, "refill_min must be in [1, 64]");
    need(interact_min >= 1 && interact_min <= 64, "interact_min must be in [1, 64]");
    c->queue_blocks_per_cu = blocks_per_cu;
    c->refill_min = refill_min;
    c->interact_min = interact_min;
  });
}

int mdx_set_event_kernel(mdx_ctx* c, int kind, int bank_size) {
  return guard([&] {
    need(c, "NULL context");
    need(kind == MDX_EVENT_QUEUE || kind == MDX_EVENT_BANK, "unknown event kernel");
    need(bank_size == 96 || bank_size == 128 || bank_size == 160, "bank_size must be 96, 128 or 160");
    c->event_kernel = kind;
    c->bank_size = bank_size;
  });
}

int mdx_get_event_kernel(mdx_ctx* c, int* kind, int* bank_size) {
  return guard([&] {
    need(c && kind && bank_size, "NULL argument");
    *kind = c->event_kernel;
    *bank_size = c->bank_size;
  });
}

int mdx_get_queue_params(mdx_ctx* c, int* blocks_per_cu, int* refill_min, int* interact_min) {
  return guard([&] {
    need(c && blocks_per_cu && refill_min && interact_min, "NULL argument");
    *blocks_per_cu = c->queue_blocks_per_cu;
    *refill_min = c->refill_min;
    *interact_min = c->interact_min;
  });
}

int mdx_autotune(mdx_ctx* c, uint64_t probe_histories, int flags, int* blocks_per_cu, int* refill_min,
                 int* interact_min, double* best_ms) {
  return guard([&] {
    need(c, "NULL context");
    need(probe_histories > 1, "probe_histories must be at least 2");
    need(!c->inflight, "a round is in flight (mdx_wait first)");
    need((flags & ~(MDX_TUNE_ROUNDS | MDX_TUNE_EVENTS)) == 0, "unknown autotune flags");
    const int job = c->cur;
    const bool rounds = (flags & MDX_TUNE_ROUNDS) != 0;
    const uint64_t seed = 0x7475u;
    uint64_t first = 1ull << 40;  // a stream region no production run uses
    // the probe launch: the job alone (single-job kernels) or as two slices
    // of itself in one launch (the multi-job kernels that rounds run, with
    // their own register and LDS budgets)
    auto launch = [&](mdx_stats* st) {
      if (rounds) {
        const int ids[2] = {job, job};
        const uint64_t h0 = probe_histories / 2;
        const uint64_t f[2] = {first, first + h0}, n[2] = {h0, probe_histories - h0}, s[2] = {seed, seed};
        run_jobs(c, 2, ids, f, n, s, st);
      } else {
        run_jobs(c, 1, &job, &first, &probe_histories, &seed, st);
      }
      first += probe_histories;
    };
    auto probe = [&](int kind, int bank, int b, int r, int i) {
      c->event_kernel = kind;
      c->bank_size = bank;
      c->queue_blocks_per_cu = b;
      c->refill_min = r;
      c->interact_min = i;
      double best = 1e300;
      for (int rep = 0; rep < 2; ++rep) {
        mdx_stats st;
        launch(&st);
        best = std::min(best, st.kernel_ms);
      }
      return best;
    };
    int kind = MDX_EVENT_QUEUE, bank = c->bank_size;
    int b = c->queue_blocks_per_cu, r = c->refill_min, i = c->interact_min;
    if (c->event_kernel != MDX_EVENT_QUEUE) b = 4, r = 32, i = 32;
    double cur = probe(kind, bank, b, r, i);
    // queue kernel, coordinate search: interaction batch, refill batch (up to
    // 62 of a wave's 64 lanes: a source far from the grid wants its photons'
    // free flights in large batches), both again, then occupancy (the
    // kernel's 128 registers allow four waves per SIMD, i.e. four blocks of
    // 256 lanes per CU; three is the only smaller choice worth trying, more
    // blocks than fit only queue behind the resident ones)
    for (int pass = 0; pass < 2; ++pass) {
      for (int v : {8, 16, 24, 32, 40, 48}) {
        if (v == i) continue;
        double t = probe(kind, bank, b, r, v);
        if (t < cur) { cur = t; i = v; }
      }
      for (int v : {8, 16, 24, 32, 40, 48, 56, 60, 62}) {
        if (v == r) continue;
        double t = probe(kind, bank, b, v, i);
        if (t < cur) { cur = t; r = v; }
      }
    }
    for (int v : {3, 4}) {
      if (v == b) continue;
      double t = probe(kind, bank, v, r, i);
      if (t < cur) { cur = t; b = v; }
    }
    // the bank kernel (not for the per-block LDS dose mode, which always runs
    // on the queue kernel): every bank size at two to four blocks per CU
    // (LDS and registers fit four blocks of each), then its refill and
    // interaction batches
    const Job& O = job_at(c, domain_owner(c, job));
    const bool lds_dose = O.voxel_dose && O.dose_mode == MDX_DOSE_F32C_LDS && !rounds;
    if ((flags & MDX_TUNE_EVENTS) && !lds_dose) {
      int kb = 96, bb = 2, rb = 64, ib = 32;
      double tb = 1e300;
      for (int bs : {96, 128, 160})
        for (int v : {2, 3, 4}) {
          double t = probe(MDX_EVENT_BANK, bs, v, rb, ib);
          if (t < tb) { tb = t; kb = bs; bb = v; }
        }
      for (int pass = 0; pass < 2; ++pass) {
        for (int v : {16, 24, 32, 48, 64}) {
          if (v == ib) continue;
          double t = probe(MDX_EVENT_BANK, kb, bb, rb, v);
          if (t < tb) { tb = t; ib = v; }
        }
        for (int v : {24, 32, 48, 64}) {
          if (v == rb) continue;
          double t = probe(MDX_EVENT_BANK, kb, bb, v, ib);
          if (t < tb) { tb = t; rb = v; }
        }
      }
      if (tb < cur) {
        cur = tb;
        kind = MDX_EVENT_BANK;
        bank = kb;
        b = bb;
        r = rb;
        i = ib;
      }
    }
    c->event_kernel = kind;
    c->bank_size = bank;
    c->queue_blocks_per_cu = b;
    c->refill_min = r;
    c->interact_min = i;
    zero_tallies(c);
    // a job that uses another job's domain deposited its probes' dose in
    // that job's grids and background voxel: those are cleared as well
    if (c->J().domain_of >= 0) {
      Job& O = c->DJ();
      for (auto& g : O.grids) {
        clear_dev(c, g->d_dose.p, g->d_dose.bytes);
        clear_dev(c, g->d_dose_n.p, g->d_dose_n.bytes);
      }
      if (O.d_bg.p) clear_dev(c, O.d_bg.p, O.d_bg.bytes);
      hip_check(hipStreamSynchronize(c->js), "reset tallies");
    }
    if (blocks_per_cu) *blocks_per_cu = b;
    if (refill_min) *refill_min = r;
    if (interact_min) *interact_min = i;
    if (best_ms) *best_ms = cur;
  });
}

int mdx_job_create(mdx_ctx* c, int* job) {
  return guard([&] {
    need(c && job, "NULL argument");
    int slot = -1;
    for (size_t i = 0; i < c->jobs.size(); ++i)
      if (!c->jobs[i]) slot = (int)i;
    if (slot < 0) {
      slot = (int)c->jobs.size();
      c->jobs.emplace_back();
    }
    c->jobs[slot] = std::make_unique<Job>(&c->pool);
    c->cur = slot;
    *job = slot;
  });
}

int mdx_job_alias_domain(mdx_ctx* c, int src) {
  return guard([&] {
    need(c, "NULL context");
    const int owner = domain_owner(c, src);
    need(owner != c->cur, "a job cannot alias its own domain");
    Job& J = c->J();
    need(J.grids.empty(), "the selected job already has grids");
    J.domain_of = owner;
    J.domain_set = true;
    J.background_id = job_at(c, owner).background_id;
    for (int k = 0; k < 3; ++k) J.dim_space[k] = job_at(c, owner).dim_space[k];
  });
}

int mdx_set_dose_precision(mdx_ctx* c, int mode) {
  return guard([&] {
    need(c, "NULL context");
    need(mode == MDX_DOSE_F64 || mode == MDX_DOSE_F32 || mode == MDX_DOSE_F32C || mode == MDX_DOSE_F32C_LDS,
         "unknown dose precision");
    need(c->J().domain_of < 0, "voxel dose belongs to the job that owns the domain");
    if (mode == MDX_DOSE_F32C_LDS) {
      need(!c->J().grids.empty(), "MDX_DOSE_F32C_LDS needs a grid");
      const size_t static_lds = mdx_launch::transport_queue_static_lds(c->device >= 0);
      const size_t need_b = c->J().grids[0]->voxels() * 8;
      if (need_b + static_lds > c->lds_optin)
        throw std::invalid_argument("grid 0 has " + std::to_string(c->J().grids[0]->voxels()) +
                                    " voxels; MDX_DOSE_F32C_LDS holds at most " +
                                    std::to_string((c->lds_optin - static_lds) / 8));
    }
    c->J().dose_mode = mode;
    if (c->device < 0) return;
    set_device_job(c);
    if (c->J().domain_set) upload_domain(c);
  });
}

int mdx_set_rita_builder(mdx_ctx* c, int on_device) {
  return guard([&] {
    need(c, "NULL context");
    c->rita_on_device = on_device != 0;
  });
}

int mdx_set_flights(mdx_ctx* c, int enable) {
  return guard([&] {
    need(c, "NULL context");
    c->J().flights = enable != 0;
  });
}

int mdx_set_local_majorants(mdx_ctx* c, int enable) {
  return guard([&] {
    need(c, "NULL context");
    c->local_majorants = enable != 0;
    if (c->device < 0 || !c->tables_ready) return;
    const int keep = c->cur;
    for (size_t j = 0; j < c->jobs.size(); ++j) {  // every job's domain record
      if (!c->jobs[j] || c->jobs[j]->domain_of >= 0 || !c->jobs[j]->domain_set) continue;
      c->cur = (int)j;
      set_device_job(c);
      upload_domain(c);
    }
    c->cur = keep;
  });
}

int mdx_get_macro_cells(mdx_ctx* c, int cells[3], uint16_t* out) {
  return guard([&] {
    need(c && cells, "NULL argument");
    Job& D = c->DJ();
    cells[0] = cells[1] = cells[2] = 0;
    if (D.grids.size() != 1 || !c->local_majorants || !c->tables_ready || c->dominated.empty()) return;
    GridHost& g = *D.grids[0];
    if (g.macro_gen != c->tables_gen) {
      if (c->device >= 0) return;  // a device context builds it with the domain's upload
      build_macro(c, D, g);        // a host-only context builds it on request
    }
    for (int a = 0; a < 3; ++a) cells[a] = g.mc[a];
    if (out) std::memcpy(out, g.macro.data(), g.macro.size() * sizeof(uint16_t));
  });
}

int mdx_set_graphs(mdx_ctx* c, int enable) {
  return guard([&] {
    need(c, "NULL context");
    c->use_graphs = enable != 0;
  });
}

int mdx_job_select(mdx_ctx* c, int job) {
  return guard([&] {
    need(c, "NULL context");
    if (job < 0 || job >= (int)c->jobs.size() || !c->jobs[job]) throw std::out_of_range("no such job");
    c->cur = job;
  });
}

int mdx_job_release(mdx_ctx* c, int job) {
  return guard([&] {
    need(c, "NULL context");
    if (job < 0 || job >= (int)c->jobs.size() || !c->jobs[job]) throw std::out_of_range("no such job");
    for (size_t j = 0; j < c->jobs.size(); ++j)
      if (c->jobs[j] && c->jobs[j]->domain_of == job)
        throw std::invalid_argument("job " + std::to_string(j) + " uses this job's domain; release it first");
    if (c->device >= 0) {
      set_device(c);
      hip_check(hipStreamSynchronize(c->stream), "release");
    }
    if (job == 0) c->jobs[0] = std::make_unique<Job>(&c->pool);
    else c->jobs[job].reset();
    if (c->cur == job) c->cur = 0;
  });
}

int mdx_job_device_bytes(mdx_ctx* c, int job, uint64_t* bytes) {
  return guard([&] {
    need(c && bytes, "NULL argument");
    if (job < 0 || job >= (int)c->jobs.size() || !c->jobs[job]) throw std::out_of_range("no such job");
    *bytes = c->jobs[job]->device_bytes();
  });
}

int mdx_tables_device_bytes(mdx_ctx* c, uint64_t* bytes) {
  return guard([&] {
    need(c && bytes, "NULL argument");
    *bytes = c->d_xs.bytes + c->d_nodes.bytes + c->d_bins.bytes + c->d_rE.bytes + c->d_rinv.bytes + c->d_roff.bytes +
             c->d_rrows.bytes + c->d_tables.bytes;
  });
}

int mdx_run_jobs(mdx_ctx* c, int n_jobs, const int* jobs, const uint64_t* first, const uint64_t* n,
                 const uint64_t* seed, mdx_stats* stats) {
  return guard([&] {
    need(c && n_jobs > 0 && jobs && first && n && seed, "bad argument");
    run_jobs(c, n_jobs, jobs, first, n, seed, stats);
  });
}

// mdx_run_jobs without allocating: the job records, prefix offsets, work
// counters and event counters of the launch go to this device's static
// scratch instead of context buffers; every job must be fully set up
// (domain, source, tallies) before.  No device or host heap allocation on
// this path: records are built on the stack, queued clears are flushed in
// place, statistics are read into a local array.
int mdx_run_jobs_noalloc(mdx_ctx* c, int n_jobs, const int* ids, const uint64_t* first, const uint64_t* n,
                         const uint64_t* seed, mdx_stats* stats) {
  return guard([&] {
    need(c && ids && first && n && seed, "NULL argument");
    need(n_jobs >= 1 && n_jobs <= MDX_NOALLOC_MAX_JOBS,
         "n_jobs must be in [1, 256]: the static scratch holds 256 job records");
    need(c->device >= 0 && c->scratch, "host-only context (device -1) cannot run device work");
    need(c->tables_ready, "physics tables not built (mdx_build_tables)");
    need(!c->inflight, "a round is in flight (mdx_wait first)");
    set_device(c);
    mdx_job_dev jd[MDX_NOALLOC_MAX_JOBS];
    uint64_t prefix[MDX_NOALLOC_MAX_JOBS + 1];
    prefix[0] = 0;
    uint32_t lds_voxels = 0;
    {
      KeepCur kc(c);
      for (int i = 0; i < n_jobs; ++i) {
        job_at(c, ids[i]);
        c->cur = domain_owner(c, ids[i]);
        check_materials(c);
        const Job& O = c->J();
        need(O.d_domain.p != nullptr, "domain not uploaded");
        if (n_jobs == 1 && O.voxel_dose && O.dose_mode == MDX_DOSE_F32C_LDS && !O.grids.empty())
          lds_voxels = (uint32_t)O.grids[0]->voxels();
        jd[i].D = O.d_domain.as<mdx_domain_dev>();
        c->cur = ids[i];
        need(c->J().source_set, "source not set (mdx_set_source)");
        need(c->J().d_tallies.p != nullptr, "tallies not set (mdx_set_tallies)");
        jd[i].S = c->J().d_source.as<mdx_source_dev>();
        jd[i].TL = c->J().d_tallies.as<mdx_tallies_dev>();
        check_bank_range(c, first[i], n[i]);
        jd[i].first = first[i];
        jd[i].n = n[i];
        jd[i].seed = seed[i];
        jd[i].stream = 0;
        bool volume = false;
        for (auto& t : c->J().tallies) volume = volume || t->t.kind == MDX_TALLY_AACUBOID;
        jd[i].flights = c->J().flights && !volume ? 1u : 0u;
        prefix[i + 1] = prefix[i] + n[i];
      }
    }
    // clears queued by job set-up while a graph round ran, done in place
    {
      std::lock_guard<std::mutex> lk(g_clear_mu);
      for (auto& r : g_clears)
        if (r.ctx == c) clear_now(r.p, r.bytes, c->stream);
      g_clears.erase(std::remove_if(g_clears.begin(), g_clears.end(), [&](const ClearReq& r) { return r.ctx == c; }),
                     g_clears.end());
    }
    const mdx_launch::ScratchLayout L = mdx_launch::scratch_layout();
    char* base = static_cast<char*>(c->scratch);
    auto* cnt = reinterpret_cast<unsigned long long*>(base + L.counters);
    auto* work = reinterpret_cast<unsigned long long*>(base + L.work);
    auto* dprefix = reinterpret_cast<uint64_t*>(base + L.prefix);
    auto* djobs = reinterpret_cast<mdx_job_dev*>(base + L.jobs);
    hip_check(hipMemcpyAsync(djobs, jd, n_jobs * sizeof(mdx_job_dev), hipMemcpyHostToDevice, c->stream), "jobs");
    hip_check(hipMemcpyAsync(dprefix, prefix, (n_jobs + 1) * sizeof(uint64_t), hipMemcpyHostToDevice, c->stream),
              "prefix");
    // the scratch keeps the last call's contents: the counters start at zero
    hip_check(mdx_launch::zero_u64(cnt, MDX_NCOUNTER, c->stream), "zero counters");
    hip_check(mdx_launch::zero_u64(work, n_jobs, c->stream), "zero work");
    const uint64_t total = prefix[n_jobs];
    hip_check(hipEventRecord(c->ev0, c->stream), "hipEventRecord");
    if (total > 0) hip_check(launch_events(c, djobs, dprefix, n_jobs, total, work, cnt, lds_voxels), "queue launch");
    hip_check(hipEventRecord(c->ev1, c->stream), "hipEventRecord");
    unsigned long long h[MDX_NCOUNTER];
    hip_check(hipMemcpyAsync(h, cnt, sizeof h, hipMemcpyDeviceToHost, c->stream), "counters");
    hip_check(hipStreamSynchronize(c->stream), "transport");
    if (stats) {
      std::memset(stats, 0, sizeof *stats);
      for (int i = 0; i < MDX_NCOUNTER && i < MDX_STAT_NCOUNTERS; ++i) stats->counters[i] = h[i];
      float ms = 0;
      hip_check(hipEventElapsedTime(&ms, c->ev0, c->ev1), "hipEventElapsedTime");
      stats->kernel_ms = ms;
    }
  });
}

int mdx_mem_info(mdx_ctx* c, uint64_t* free_bytes, uint64_t* total_bytes) {
  return guard([&] {
    need(c && free_bytes && total_bytes, "NULL argument");
    set_device_nowait(c);
    size_t f = 0, t = 0;
    hip_check(hipMemGetInfo(&f, &t), "hipMemGetInfo");
    *free_bytes = f;
    *total_bytes = t;
  });
}

int mdx_copy_probe(mdx_ctx* c, uint64_t bytes, double* ms) {
  return guard([&] {
    need(c && ms, "NULL argument");
    need(bytes >= 16 && bytes % 16 == 0, "bytes must be a positive multiple of 16");
    set_device(c);
    DBuf a, b;
    a.alloc(bytes);
    b.alloc(bytes);
    hip_check(mdx_launch::zero(a.as<double>(), bytes / 8, c->stream), "zero");
    const int blocks = c->num_cu * 8;
    hip_check(mdx_launch::copy16(a.p, b.p, bytes / 16, blocks, c->stream), "copy");  // warm
    hip_check(hipEventRecord(c->ev0, c->stream), "hipEventRecord");
    hip_check(mdx_launch::copy16(a.p, b.p, bytes / 16, blocks, c->stream), "copy");
    hip_check(hipEventRecord(c->ev1, c->stream), "hipEventRecord");
    hip_check(hipStreamSynchronize(c->stream), "copy");
    float t = 0;
    hip_check(hipEventElapsedTime(&t, c->ev0, c->ev1), "hipEventElapsedTime");
    *ms = t;
  });
}

int mdx_atomic_probe(mdx_ctx* c, uint64_t records, uint64_t deposits, double* ms, uint64_t* counted) {
  return guard([&] {
    need(c && ms, "NULL argument");
    need(records > 0 && deposits > 0, "records and deposits must be positive");
    need(records <= (1ull << 34), "records must be at most 2^34 (512 GiB of records)");
    set_device(c);
    DBuf r;
    r.alloc(records * 32);
    hip_check(mdx_launch::zero(r.as<double>(), records * 4, c->stream), "zero");
    const int blocks = c->num_cu * 8;
    hip_check(mdx_launch::atomic_probe(r.as<double>(), records, deposits, 1, blocks, c->stream), "probe");  // warm
    hip_check(hipEventRecord(c->ev0, c->stream), "hipEventRecord");
    hip_check(mdx_launch::atomic_probe(r.as<double>(), records, deposits, 2, blocks, c->stream), "probe");
    hip_check(hipEventRecord(c->ev1, c->stream), "hipEventRecord");
    hip_check(hipStreamSynchronize(c->stream), "probe");
    float t = 0;
    hip_check(hipEventElapsedTime(&t, c->ev0, c->ev1), "hipEventElapsedTime");
    *ms = t;
    if (counted) {  // the records' deposit counts summed on the host (both launches)
      std::vector<double> h(4 * (size_t)records);
      hip_check(hipMemcpy(h.data(), r.p, records * 32, hipMemcpyDeviceToHost), "read");
      uint64_t n = 0;
      for (uint64_t i = 0; i < records; ++i) {
        uint64_t k;
        std::memcpy(&k, &h[4 * i + 2], 8);
        n += k;
      }
      *counted = n;
    }
  });
}

int mdx_host_alloc(mdx_ctx* c, uint64_t bytes, void** out) {
  return guard([&] {
    need(c && out, "NULL argument");
    set_device_nowait(c);
    *out = nullptr;
    hip_check(hipHostMalloc(out, bytes ? bytes : 1, hipHostMallocDefault), "hipHostMalloc");
  });
}

int mdx_host_free(void* p) {
  return guard([&] {
    if (p) hip_check(hipHostFree(p), "hipHostFree");
  });
}

int mdx_run_jobs_async(mdx_ctx* c, int n_jobs, const int* jobs, const uint64_t* first, const uint64_t* n,
                       const uint64_t* seed) {
  return guard([&] {
    need(c && n_jobs > 0 && jobs && first && n && seed, "bad argument");
    run_jobs(c, n_jobs, jobs, first, n, seed, nullptr, false);
  });
}

int mdx_wait(mdx_ctx* c, mdx_stats* stats) {
  return guard([&] {
    need(c, "NULL context");
    need(c->inflight, "no round in flight");
    set_device_nowait(c);
    finish_round(c, stats);
  });
}

int mdx_get_tally(mdx_ctx* c, int i, double* acc, double* hist, double* image) {
  return guard([&] {
    need(c, "NULL context");
    if (i < 0 || i >= (int)c->J().tallies.size()) throw std::out_of_range("tally index out of range");
    set_device_nowait(c);
    TallyHost& h = *c->J().tallies[i];
    // beside a round on the I/O stream; otherwise on the transport stream, so
    // as not to queue behind dose copies in flight on the I/O stream
    hipStream_t keep = c->io_stream;
    if (!c->inflight) c->io_stream = c->stream;
    struct Restore {
      mdx_ctx* c;
      hipStream_t s;
      ~Restore() { c->io_stream = s; }
    } restore{c, keep};
    if (acc) {  // the sum of the copies
      std::vector<double> all((size_t)MDX_ACC_COPIES * MDX_TALLY_ACC_SIZE);
      read_back(c, all.data(), h.acc.p, h.acc.bytes, "read tally");
      for (int k = 0; k < MDX_TALLY_ACC_SIZE; ++k) {
        double s = 0.0;
        for (int r = 0; r < MDX_ACC_COPIES; ++r) s += all[(size_t)r * MDX_TALLY_ACC_SIZE + k];
        acc[k] = s;
      }
    }
    if (hist && h.hist.p) read_back(c, hist, h.hist.p, h.hist.bytes, "read hist");
    if (image && h.image.p) read_back(c, image, h.image.p, h.image.bytes, "read image");
  });
}

int mdx_get_voxel_dose(mdx_ctx* c, int grid, double* ss, uint64_t* count) {
  return guard([&] {
    need(c, "NULL context");
    Job& D = c->DJ();
    if (grid < 0 || grid >= (int)D.grids.size()) throw std::out_of_range("grid index out of range");
    need(D.voxel_dose, "voxel dose is not enabled");
    set_device_nowait(c);
    GridHost& g = *D.grids[grid];
    const size_t nv = g.voxels();
    if (g.dose_mode == MDX_DOSE_F64) {
      if (!ss && !count) return;
      // the records, in chunks, split into (sum, sum^2) pairs and counts
      const size_t chunk = std::min<size_t>(nv, (size_t)1 << 20);
      std::vector<double> rec(MDX_DOSE_REC * chunk);
      for (size_t v0 = 0; v0 < nv; v0 += chunk) {
        const size_t m = std::min(chunk, nv - v0);
        read_back(c, rec.data(), g.d_dose.as<double>() + MDX_DOSE_REC * v0, MDX_DOSE_REC * m * sizeof(double),
                  "read dose");
        for (size_t i = 0; i < m; ++i) {
          const double* r = &rec[MDX_DOSE_REC * i];
          if (ss) {
            ss[2 * (v0 + i)] = r[0];
            ss[2 * (v0 + i) + 1] = r[1];
          }
          if (count) std::memcpy(&count[v0 + i], &r[2], sizeof(uint64_t));
        }
      }
      return;
    }
    if (ss) {  // FP32 layouts, widened: (sum, sum^2) or (sum + error, sum^2 + error)
      const int w = g.dose_mode == MDX_DOSE_F32 ? 2 : 4;
      std::vector<float> f(nv * w);
      read_back(c, f.data(), g.d_dose.p, f.size() * sizeof(float), "read dose");
      for (size_t i = 0; i < nv; ++i) {
        if (w == 2) {
          ss[2 * i] = f[2 * i];
          ss[2 * i + 1] = f[2 * i + 1];
        } else {
          ss[2 * i] = (double)f[4 * i] + (double)f[4 * i + 1];
          ss[2 * i + 1] = (double)f[4 * i + 2] + (double)f[4 * i + 3];
        }
      }
    }
    if (count) read_back(c, count, g.d_dose_n.p, g.d_dose_n.bytes, "read dose n");
  });
}

int mdx_get_voxel_dose_f32(mdx_ctx* c, int grid, uint64_t n_histories, float* out) {
  return guard([&] {
    need(c && out, "NULL argument");
    need(n_histories > 0, "n_histories must be positive");
    Job& D = c->DJ();
    if (grid < 0 || grid >= (int)D.grids.size()) throw std::out_of_range("grid index out of range");
    need(D.voxel_dose, "voxel dose is not enabled");
    set_device_nowait(c);
    GridHost& g = *D.grids[grid];
    need(g.dose_mode == MDX_DOSE_F64, "FP32 dose reads need FP64 accumulation (MDX_DOSE_F64)");
    const size_t nv = g.voxels();
    if (D.staged_grid == grid && D.staged_n == n_histories) {
      hip_check(hipStreamWaitEvent(c->io_stream, D.ev_stage, 0), "wait staged dose");
      read_back(c, out, D.d_stage.p, nv * 8, "read dose f32");
      D.staged_grid = -1;
      D.d_stage.release();
      return;
    }
    c->d_dose32.ensure(nv * 8);
    flush_clears_over(c, g.d_dose.p, g.d_dose.bytes);
    hip_check(mdx_launch::dose_f32(g.d_dose.as<double>(), nv, (double)n_histories, c->d_dose32.as<float>(),
                                   c->io_stream),
              "dose f32");
    read_back(c, out, c->d_dose32.p, nv * 8, "read dose f32");
  });
}

int mdx_read_staged_dose_f32_async(mdx_ctx* c, float* out) {
  return guard([&] {
    need(c && out, "NULL argument");
    Job& D = c->DJ();
    need(D.staged_grid >= 0 && D.d_stage.p, "no staged dose (mdx_stage_voxel_dose_f32)");
    need(!D.copy_pending, "a copy of this job's dose is in flight");
    set_device_nowait(c);
    if (!D.ev_copy) hip_check(hipEventCreateWithFlags(&D.ev_copy, hipEventDisableTiming), "hipEventCreate");
    hip_check(hipStreamWaitEvent(c->copy_stream, D.ev_stage, 0), "wait staged dose");
    hip_check(hipMemcpyAsync(out, D.d_stage.p, D.d_stage.bytes, hipMemcpyDeviceToHost, c->copy_stream),
              "read dose f32");
    hip_check(hipEventRecord(D.ev_copy, c->copy_stream), "hipEventRecord");
    D.copy_pending = true;
  });
}

int mdx_read_done(mdx_ctx* c, int job, int wait, int* done) {
  return guard([&] {
    need(c && done, "NULL argument");
    Job& J = job_at(c, job);
    Job& D = J.domain_of >= 0 ? job_at(c, J.domain_of) : J;
    *done = 1;
    if (!D.copy_pending) return;
    set_device_nowait(c);
    if (wait) {
      hip_check(hipEventSynchronize(D.ev_copy), "read dose f32");
    } else {
      const hipError_t e = hipEventQuery(D.ev_copy);
      if (e == hipErrorNotReady) {
        (void)hipGetLastError();
        *done = 0;
        return;
      }
      hip_check(e, "read dose f32");
    }
    D.copy_pending = false;
    D.staged_grid = -1;
    D.d_stage.release();
  });
}

int mdx_stage_voxel_dose_f32(mdx_ctx* c, int grid, uint64_t n_histories) {
  return guard([&] {
    need(c, "NULL context");
    need(n_histories > 0, "n_histories must be positive");
    Job& D = c->DJ();
    if (grid < 0 || grid >= (int)D.grids.size()) throw std::out_of_range("grid index out of range");
    need(D.voxel_dose, "voxel dose is not enabled");
    need(!D.copy_pending, "a copy of this job's staged dose is in flight (mdx_read_done first)");
    set_device_nowait(c);
    GridHost& g = *D.grids[grid];
    need(g.dose_mode == MDX_DOSE_F64, "FP32 dose reads need FP64 accumulation (MDX_DOSE_F64)");
    const size_t nv = g.voxels();
    D.d_stage.ensure(nv * 8);
    if (!D.ev_stage) hip_check(hipEventCreateWithFlags(&D.ev_stage, hipEventDisableTiming), "hipEventCreate");
    flush_clears_over(c, g.d_dose.p, g.d_dose.bytes);
    hip_check(mdx_launch::dose_f32(g.d_dose.as<double>(), nv, (double)n_histories, D.d_stage.as<float>(), c->stream),
              "dose f32");
    hip_check(hipEventRecord(D.ev_stage, c->stream), "hipEventRecord");
    D.staged_grid = grid;
    D.staged_n = n_histories;
  });
}

int mdx_get_background_dose(mdx_ctx* c, double out[3]) {
  return guard([&] {
    need(c && out, "NULL argument");
    set_device_nowait(c);
    Job& D = c->DJ();
    need(D.d_bg.p, "domain not uploaded");
    read_back(c, out, D.d_bg.p, 3 * sizeof(double), "read bg");
  });
}

int mdx_interpolate(mdx_ctx* c, int id, int which, int n, const float* x, float* out) {
  return guard([&] {
    need(c && x && out && n >= 0, "bad argument");
    need(c->tables_ready, "physics tables not built");
    need(which >= 0 && which <= MDX_SF_INCOHERENT, "unknown table selector");
    int slot = -1;
    if (id >= 0) {
      slot = c->slot_of(id);
      if (slot == 255) throw std::out_of_range("material has no tables");
    }
    if (n == 0) return;
    set_device(c);
    DevArray<float> dx(x, n, c->stream), dout(nullptr, n, c->stream);
    hip_check(mdx_launch::eval_curve(c->d_tables.as<mdx_tables_dev>(), slot, which, n, dx.p(), dout.p(), c->stream),
              "eval");
    hip_check(hipMemcpyAsync(out, dout.p(), n * sizeof(float), hipMemcpyDeviceToHost, c->stream), "read");
    hip_check(hipStreamSynchronize(c->stream), "eval");
  });
}

int mdx_interpolate_host(mdx_ctx* c, int id, int which, int n, const double* x, double* out) {
  return guard([&] {
    need(c && x && out && n >= 0, "bad argument");
    if (id < 0) {
      need(c->tables_ready, "physics tables not built");
      mdx::LogLogLinear f(c->majorant.x, c->majorant.y);
      for (int i = 0; i < n; ++i) out[i] = f(x[i]);
      return;
    }
    const mdx::MaterialTables& m = host_material(c, id);
    for (int i = 0; i < n; ++i) {
      switch (which) {
        case MDX_XS_TOTAL: out[i] = m.total_i(x[i]); break;
        case MDX_XS_COHERENT: out[i] = m.coh_i(x[i]); break;
        case MDX_XS_INCOHERENT: out[i] = m.incoh_i(x[i]); break;
        case MDX_XS_PHOTOELECTRIC: out[i] = m.photo_i(x[i]); break;
        case MDX_SF_INCOHERENT: out[i] = m.sf_i(x[i]); break;
        case MDX_FF_COHERENT: out[i] = m.ff_i(x[i]); break;
        default: throw std::invalid_argument("unknown table selector");
      }
    }
  });
}

int mdx_get_table(mdx_ctx* c, int id, int which, int* n, double* x, double* y) {
  return guard([&] {
    need(c && n, "NULL argument");
    const mdx::Table* t = nullptr;
    if (id < 0) {
      need(c->tables_ready, "physics tables not built");
      t = &c->majorant;
    } else {
      const mdx::MaterialTables& m = host_material(c, id);
      switch (which) {
        case MDX_XS_TOTAL: t = &m.total; break;
        case MDX_XS_COHERENT: t = &m.coh; break;
        case MDX_XS_INCOHERENT: t = &m.incoh; break;
        case MDX_XS_PHOTOELECTRIC: t = &m.photo; break;
        case MDX_SF_INCOHERENT: t = &m.sf; break;
        case MDX_FF_COHERENT: t = &m.ff; break;
        default: throw std::invalid_argument("unknown table selector");
      }
    }
    if (!x) {
      *n = (int)t->x.size();
      return;
    }
    int m = std::min(*n, (int)t->x.size());
    std::memcpy(x, t->x.data(), m * sizeof(double));
    if (y) std::memcpy(y, t->y.data(), m * sizeof(double));
    *n = m;
  });
}

int mdx_sample_rayleigh(mdx_ctx* c, int id, int n, const float* E, const float* y, float* mu) {
  return guard([&] {
    need(c && E && y && mu && n >= 0, "bad argument");
    need(c->tables_ready, "physics tables not built");
    int slot = c->slot_of(id);
    if (slot == 255) throw std::out_of_range("material has no tables");
    if (n == 0) return;
    set_device(c);
    DevArray<float> dE(E, n, c->stream), dy(y, n, c->stream), dmu(nullptr, n, c->stream);
    hip_check(mdx_launch::rita(c->d_tables.as<mdx_tables_dev>(), slot, n, dE.p(), dy.p(), dmu.p(), c->stream), "rita");
    hip_check(hipMemcpyAsync(mu, dmu.p(), n * sizeof(float), hipMemcpyDeviceToHost, c->stream), "read");
    hip_check(hipStreamSynchronize(c->stream), "rita");
  });
}

int mdx_sample_rayleigh_host(mdx_ctx* c, int id, int n, const double* E, const double* y, double* mu) {
  return guard([&] {
    need(c && E && y && mu && n >= 0, "bad argument");
    const mdx::MaterialTables& m = host_material(c, id);
    for (int i = 0; i < n; ++i) mu[i] = m.rita.sample(E[i], y[i]);
  });
}

int mdx_get_rita_row(mdx_ctx* c, int id, int ie, int* rows, double* energy, double* x, double* y, double* a,
                     double* b) {
  return guard([&] {
    need(c && rows, "NULL argument");
    const mdx::MaterialTables& m = host_material(c, id);
    if (ie < 0 || ie >= (int)m.rita.energies.size()) throw std::out_of_range("energy index out of range");
    if (energy) *energy = m.rita.energies[ie];
    if (!x) {
      *rows = (int)m.rita.x[ie].size();
      return;
    }
    int r = std::min(*rows, (int)m.rita.x[ie].size());
    for (int i = 0; i < r; ++i) {
      x[i] = m.rita.x[ie][i];
      if (y) y[i] = m.rita.y[ie][i];
      if (a) a[i] = i < (int)m.rita.a[ie].size() ? m.rita.a[ie][i] : 0.0;
      if (b) b[i] = i < (int)m.rita.b[ie].size() ? m.rita.b[ie][i] : 0.0;
    }
    *rows = r;
  });
}

int mdx_sample_compton(mdx_ctx* c, int id, int n, const float* E, uint64_t seed, float* mu, float* Ep) {
  return guard([&] {
    need(c && E && mu && Ep && n >= 0, "bad argument");
    need(c->tables_ready, "physics tables not built");
    int slot = c->slot_of(id);
    if (slot == 255) throw std::out_of_range("material has no tables");
    if (n == 0) return;
    set_device(c);
    DevArray<float> dE(E, n, c->stream), dmu(nullptr, n, c->stream), dEp(nullptr, n, c->stream);
    hip_check(mdx_launch::kahn(c->d_tables.as<mdx_tables_dev>(), slot, n, dE.p(), seed, dmu.p(), dEp.p(), c->stream),
              "kahn");
    hip_check(hipMemcpyAsync(mu, dmu.p(), n * sizeof(float), hipMemcpyDeviceToHost, c->stream), "read");
    hip_check(hipMemcpyAsync(Ep, dEp.p(), n * sizeof(float), hipMemcpyDeviceToHost, c->stream), "read");
    hip_check(hipStreamSynchronize(c->stream), "kahn");
  });
}

int mdx_voxel_lookup(mdx_ctx* c, int n, const float* xyz, int* grid, long long* lin, int* mat) {
  return guard([&] {
    need(c && xyz && grid && lin && mat && n >= 0, "bad argument");
    check_materials(c);
    if (n == 0) return;
    set_device(c);
    DevArray<float> dp(xyz, 3 * (size_t)n, c->stream);
    DevArray<int> dg(nullptr, n, c->stream), dm(nullptr, n, c->stream);
    DevArray<long long> dl(nullptr, n, c->stream);
    hip_check(mdx_launch::voxel(c->d_tables.as<mdx_tables_dev>(), c->J().d_domain.as<mdx_domain_dev>(), n, dp.p(), dg.p(),
                                dl.p(), dm.p(), c->stream),
              "voxel");
    hip_check(hipMemcpyAsync(grid, dg.p(), n * sizeof(int), hipMemcpyDeviceToHost, c->stream), "read");
    hip_check(hipMemcpyAsync(lin, dl.p(), n * sizeof(long long), hipMemcpyDeviceToHost, c->stream), "read");
    hip_check(hipMemcpyAsync(mat, dm.p(), n * sizeof(int), hipMemcpyDeviceToHost, c->stream), "read");
    hip_check(hipStreamSynchronize(c->stream), "voxel");
  });
}

int mdx_select_interaction(mdx_ctx* c, int n, const float* sc, const float* si, const float* st, const float* u,
                           int* type) {
  return guard([&] {
    need(c && sc && si && st && u && type && n >= 0, "bad argument");
    if (n == 0) return;
    set_device(c);
    DevArray<float> a(sc, n, c->stream), b(si, n, c->stream), d(st, n, c->stream), e(u, n, c->stream);
    DevArray<int> o(nullptr, n, c->stream);
    hip_check(mdx_launch::select(n, a.p(), b.p(), d.p(), e.p(), o.p(), c->stream), "select");
    hip_check(hipMemcpyAsync(type, o.p(), n * sizeof(int), hipMemcpyDeviceToHost, c->stream), "read");
    hip_check(hipStreamSynchronize(c->stream), "select");
  });
}

int mdx_delta_decision(mdx_ctx* c, int n, const float* s_tot, const float* s_max, const float* u,
                       const int* limited, int* real) {
  return guard([&] {
    need(c && s_tot && s_max && u && limited && real && n >= 0, "bad argument");
    if (n == 0) return;
    set_device(c);
    DevArray<float> a(s_tot, n, c->stream), b(s_max, n, c->stream), d(u, n, c->stream);
    DevArray<int> l(limited, n, c->stream), o(nullptr, n, c->stream);
    hip_check(mdx_launch::delta(n, a.p(), b.p(), d.p(), l.p(), o.p(), c->stream), "delta");
    hip_check(hipMemcpyAsync(real, o.p(), n * sizeof(int), hipMemcpyDeviceToHost, c->stream), "read");
    hip_check(hipStreamSynchronize(c->stream), "delta");
  });
}

int mdx_rotate(mdx_ctx* c, int n, const float* dir, const float* mu, const float* phi, float* out) {
  return guard([&] {
    need(c && dir && mu && phi && out && n >= 0, "bad argument");
    if (n == 0) return;
    set_device(c);
    DevArray<float> d(dir, 3 * (size_t)n, c->stream), m(mu, n, c->stream), p(phi, n, c->stream);
    DevArray<float> o(nullptr, 3 * (size_t)n, c->stream);
    hip_check(mdx_launch::rotate(n, d.p(), m.p(), p.p(), o.p(), c->stream), "rotate");
    hip_check(hipMemcpyAsync(out, o.p(), 3 * (size_t)n * sizeof(float), hipMemcpyDeviceToHost, c->stream), "read");
    hip_check(hipStreamSynchronize(c->stream), "rotate");
  });
}

int mdx_sample_source(mdx_ctx* c, int n, uint64_t seed, float* dir, float* energy) {
  return guard([&] {
    need(c && dir && energy && n >= 0, "bad argument");
    need(c->J().source_set, "source not set (mdx_set_source)");
    if (n == 0) return;
    set_device(c);
    DevArray<float> d(nullptr, 3 * (size_t)n, c->stream), e(nullptr, n, c->stream);
    hip_check(mdx_launch::source(c->J().d_source.as<mdx_source_dev>(), n, seed, d.p(), e.p(), c->stream), "source");
    hip_check(hipMemcpyAsync(dir, d.p(), 3 * (size_t)n * sizeof(float), hipMemcpyDeviceToHost, c->stream), "read");
    hip_check(hipMemcpyAsync(energy, e.p(), n * sizeof(float), hipMemcpyDeviceToHost, c->stream), "read");
    hip_check(hipStreamSynchronize(c->stream), "source");
  });
}

int mdx_roi_stats(mdx_ctx* c, int n, const int* jobs, const int* boxes, double* out) {
  return guard([&] {
    need(c && (n == 0 || (jobs && boxes && out)) && n >= 0, "bad argument");
    need(!c->inflight, "a round is in flight (mdx_wait first)");
    if (n == 0) return;
    set_device(c);
    hip_check(hipStreamSynchronize(c->stream), "roi");  // the mailbox may be regrown only with the device idle
    ensure_mail(c, (size_t)n);
    mdx_roi_req* req = mail_roi_req(c, false);
    for (int b = 0; b < n; ++b) {
      Job& O = job_at(c, domain_owner(c, jobs[b]));
      need(O.voxel_dose && !O.grids.empty(), "the job's voxel dose is not enabled");
      GridHost& G = *O.grids[0];
      need(G.dose_mode == MDX_DOSE_F64 && G.d_dose.p, "ROI statistics need FP64 voxel dose");
      mdx_roi_req& q = req[b];
      q.rec = G.d_dose.as<double>();
      q.nx = G.dims[0];
      q.ny = G.dims[1];
      for (int k = 0; k < 3; ++k) {
        q.lo[k] = boxes[6 * b + 2 * k];
        q.hi[k] = boxes[6 * b + 2 * k + 1];
        if (!(0 <= q.lo[k] && q.lo[k] < q.hi[k] && q.hi[k] <= G.dims[k]))
          throw std::out_of_range("ROI box outside the grid or empty");
      }
      flush_clears_over(c, G.d_dose.p, G.d_dose.bytes);
    }
    // the kernel reads the requests from and writes the sums to the pinned
    // mailbox directly: no copy engine involved
    hip_check(mdx_launch::roi_stats(mail_roi_req(c, true), n, mail_roi_out(c, true), c->stream), "roi");
    hip_check(hipStreamSynchronize(c->stream), "roi");
    std::memcpy(out, mail_roi_out(c, false), (size_t)3 * n * sizeof(double));
  });
}

int mdx_philox(mdx_ctx* c, int n, int words, const uint32_t* counter, const uint32_t* key, uint32_t* out) {
  return guard([&] {
    need(c && counter && key && out && n >= 0, "bad argument");
    need(words == 2 || words == 4, "words must be 2 or 4");
    if (n == 0) return;
    set_device(c);
    DevArray<uint32_t> dc(counter, (size_t)n * words, c->stream), dk(key, (size_t)n * (words / 2), c->stream);
    DevArray<uint32_t> o(nullptr, (size_t)n * words, c->stream);
    hip_check(mdx_launch::philox(n, words, dc.p(), dk.p(), o.p(), c->stream), "philox");
    hip_check(hipMemcpyAsync(out, o.p(), (size_t)n * words * sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream),
              "read");
    hip_check(hipStreamSynchronize(c->stream), "philox");
  });
}

}  // extern "C"
