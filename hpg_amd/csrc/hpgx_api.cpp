// C ABI of the MI355X gridding device (include/hpgx.h): state management,
// argument validation, device memory, hipFFT, and the launch sequence of the
// kernels in hpgx_kernels.hip.
//
// One hpgx_state_s corresponds to one hpg::runtime::StateT<D>
// (src/hpg_runtime.hpp:720-1750): it owns the grid, weights, model, CF and
// per-call device buffers, and one HIP stream on which all its work is
// queued in order.

#include "hpgx.h"
#include "hpgx_internal.hpp"

#include <hip/hip_runtime.h>
#include <hipfft/hipfft.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

using hpgx::CallBuffers;
using hpgx::DeviceCF;
using hpgx::Geometry;
using hpgx::GroupInfo;
using hpgx::Mueller;

namespace {

constexpr int num_event_slots = 16;

std::mutex g_create_mtx;
std::string g_create_error;  // message of the last failed hpgx_create
int g_device = -1;

struct Plan {
  hipfftHandle h = 0;
  bool valid = false;
};

}  // namespace

// Fixed device scratch of hpgx_grid_visibilities_noalloc: the bin offset
// table, bin entries, entry record indexes and halo values of its gridding
// passes.  Each GPU the library runs on holds its own copy until the process
// ends.  Host code never uses the array's name as an address: the device
// address of each GPU's copy comes from hipGetSymbolAddress (noalloc_scratch).
constexpr size_t noalloc_scratch_bytes = size_t(HPGX_NOALLOC_SCRATCH_BYTES);
static __device__ __align__(256) unsigned char scratch[noalloc_scratch_bytes];

namespace {
constexpr int noalloc_max_devices = 64;
std::mutex g_noalloc_mtx;
void* g_noalloc_ptr[noalloc_max_devices] = {};

// device address of `device`'s scratch, looked up once per device with that
// device current
hipError_t noalloc_scratch(int device, void** out) {
  *out = nullptr;
  if (device < 0 || device >= noalloc_max_devices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lk(g_noalloc_mtx);
  if (!g_noalloc_ptr[device]) {
    int cur = 0;
    hipError_t e = hipGetDevice(&cur);
    if (e != hipSuccess) return e;
    if (cur != device && (e = hipSetDevice(device)) != hipSuccess) return e;
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (cur != device) (void)hipSetDevice(cur);
    if (e != hipSuccess) return e;
    if (!p) return hipErrorInvalidSymbol;
    g_noalloc_ptr[device] = p;
  }
  *out = g_noalloc_ptr[device];
  return hipSuccess;
}
}  // namespace

struct hpgx_state_s {
  int device = 0;
  hipStream_t stream = nullptr;
  unsigned max_added_tasks = 0;
  size_t max_batch = 0;
  unsigned gsize[4] = {0, 0, 0, 0};
  double gscale[2] = {0.0, 0.0};
  unsigned npol = 0;
  int mode = HPGX_GRID_FP64;
  Mueller mu{};
  std::vector<int> mueller_rows, conj_mueller_rows;  // as given, rows x npol

  void* grid = nullptr;       // double2 or float2, Layout::Left
  double* weights = nullptr;  // [cube][mrow]
  double2* model = nullptr;   // Layout::Left, optional
  void* scratch = nullptr;    // grid-sized, shift / out-of-place FFT / readback

  // convolution function
  DeviceCF cf{};
  std::vector<GroupInfo> groups;
  size_t taps_capacity = 0;   // float2 elements
  size_t wsum_capacity = 0;   // double2 elements
  float2* staging = nullptr;
  size_t staging_capacity = 0;

  CallBuffers cb{};
  size_t vis_capacity = 0;  // records
  // tasks (max_added_tasks + 1): record buffers that host batches are
  // uploaded into on the task's own copy stream, in rotation, so that the
  // upload of one call overlaps the kernels of the previous ones (all
  // kernels stay on `stream`: they update the grid in place).  Task 0's
  // buffer is cb.vis as allocated; cb.vis points to the current task's.
  unsigned tasks = 1;
  unsigned task = 0;
  std::vector<void*> task_vis;
  std::vector<hipStream_t> copy_streams;
  std::vector<hipEvent_t> copied, released;  // per task
  std::vector<bool> released_set;
  // gridding pipeline: 5 = as 4, but FP32 grids with (hi, lo) pair tiles
  // are added by one wave per tile quadrant with one lane per CF tap and no
  // LDS atomics (default; 4 where the quadrant regions do not fit LDS), 4 = block-local
  // bins of one entry per record and plane, haloed LDS tiles updated by LDS
  // atomics (falls back to 3 when a CF support reaches the tile edge; grids
  // with more (plane, tile) keys than one workgroup's LDS histogram holds
  // take global bins), 3 = block-local bins of one entry per touched tile,
  // 2 = compact entries binned through global counters
  int gridder = 5;
  size_t nkeys = 0;
  size_t pass_blocks = 0;  // bin blocks per pass of the block-local pipelines
  // tile accumulation of FP32 grids in the haloed-tile gridders: 2 = FP32
  // hi/lo pairs updated by two-sum (default; under compare-and-swap in
  // gridder 4), 0 = FP64 tile sums
  int fp32_tile = 2;
  // FP32-compensated grids with supports up to 8 x 8 on the block-local
  // haloed bins: quadrant-keyed bins and k_sub instead of k_quad
  // (HPGX_SUB_TILES=0 keeps k_quad)
  bool sub_tiles = true;
  // the finishing chain, deferred until an entry point touches the grid
  // (flush_finishing): normalize_by_weights, an in-place grid FFT and, for
  // even grid sizes, a shift directly after that FFT
  bool pend_norm = false;
  double pend_wfactor = 1.0;
  bool pend_fft = false;
  double pend_fft_norm = 1.0;
  int pend_fft_sign = 0;
  bool pend_shift = false;

  Plan plan_z2z, plan_c2c;
  hipEvent_t events[num_event_slots] = {};
  std::string err;

  size_t grid_elems() const {
    return size_t(gsize[0]) * gsize[1] * gsize[2] * gsize[3];
  }
  size_t grid_bytes() const {
    return grid_elems() * (mode == HPGX_GRID_FP64 ? sizeof(double2) : sizeof(float2));
  }
  Geometry geometry() const {
    Geometry g{};
    g.nx = int(gsize[0]);
    g.ny = int(gsize[1]);
    g.nr = int(gsize[2]);
    g.nc = int(gsize[3]);
    g.scale[0] = gscale[0];
    g.scale[1] = gscale[1];
    g.os = cf.os;
    g.F = cf.F;
    g.nf = cf.nf;
    g.ngroups = cf.ngroups;
    g.ntx = (g.nx + hpgx::tile - 1) / hpgx::tile;
    g.nty = (g.ny + hpgx::tile - 1) / hpgx::tile;
    return g;
  }
};

namespace {

int fail(hpgx_state_t st, int code, const std::string& msg) {
  if (st) st->err = msg;
  return code;
}

int hip_fail(hpgx_state_t st, hipError_t e, const char* what) {
  return fail(
    st, HPGX_ERR_OTHER, std::string(what) + ": " + hipGetErrorString(e));
}

#define HPGX_HIP(st, call)                                 \
  do {                                                     \
    hipError_t e_ = (call);                                \
    if (e_ != hipSuccess) return hip_fail((st), e_, #call); \
  } while (0)

template <typename T>
void dev_free(T*& p) {
  if (p) (void)hipFree(p);
  p = nullptr;
}

template <typename T>
hipError_t dev_alloc(T*& p, size_t bytes) {
  void* q = nullptr;
  hipError_t e = hipMalloc(&q, bytes ? bytes : 1);
  p = static_cast<T*>(q);
  return e;
}

size_t visdata_size(unsigned npol) {
  switch (npol) {
  case 1: return sizeof(hpgx_visdata1_t);
  case 2: return sizeof(hpgx_visdata2_t);
  case 3: return sizeof(hpgx_visdata3_t);
  case 4: return sizeof(hpgx_visdata4_t);
  default: return 0;
  }
}
static_assert(sizeof(hpgx_visdata1_t) == 88, "hpg::VisData<1> layout");
static_assert(sizeof(hpgx_visdata2_t) == 96, "hpg::VisData<2> layout");
static_assert(sizeof(hpgx_visdata3_t) == 112, "hpg::VisData<3> layout");
static_assert(sizeof(hpgx_visdata4_t) == 120, "hpg::VisData<4> layout");

// CFArray::radii (src/hpg.hpp:687-694)
void radii(const unsigned* ext, unsigned os, int& rx, int& ry) {
  rx = int(((ext[0] - 2 * hpgx::cf_padding * os) / os) / 2);
  ry = int(((ext[1] - 2 * hpgx::cf_padding * os) / os) / 2);
}

// device CF size in float2 elements for a CF shape (fine-offset slices)
size_t cf_taps_for(unsigned os, unsigned ngroups, const unsigned* extents,
                   size_t* wsum_elems = nullptr) {
  const size_t F = (os + 1) / 2, nf = 2 * F + 1;
  size_t taps = 0, ws = 0;
  for (unsigned g = 0; g < ngroups; ++g) {
    const unsigned* e = extents + 4 * g;
    int rx, ry;
    radii(e, os, rx, ry);
    const size_t slices = size_t(e[2]) * e[3] * nf * nf;
    taps += slices * size_t(2 * rx + 1) * size_t(2 * ry + 1);
    ws += slices;
  }
  if (wsum_elems) *wsum_elems = ws;
  return taps;
}

int ensure_cf_capacity(hpgx_state_t st, size_t taps, size_t ws) {
  if (taps > st->taps_capacity) {
    dev_free(st->cf.taps);
    st->taps_capacity = 0;
    HPGX_HIP(st, dev_alloc(st->cf.taps, taps * sizeof(float2)));
    st->taps_capacity = taps;
  }
  if (ws > st->wsum_capacity) {
    dev_free(st->cf.wsum);
    st->wsum_capacity = 0;
    HPGX_HIP(st, dev_alloc(st->cf.wsum, ws * sizeof(double2)));
    st->wsum_capacity = ws;
  }
  if (!st->cf.groups)
    HPGX_HIP(st, dev_alloc(st->cf.groups, hpgx::max_cf_groups * sizeof(GroupInfo)));
  return HPGX_SUCCESS;
}

// bin entries one record can produce: one per grid row per touched tile
// block-local bins: the pipelines of choice when one workgroup's LDS holds
// a histogram over all (plane, tile) keys
bool use_block_local(hpgx_state_t st) {
  return (st->gridder >= 3 && st->gridder <= 5)
    && st->nkeys <= size_t(hpgx::bin_keys_limit());
}

// halo widths of the haloed-tile gridder: largest CF support - 1 per axis
int halo_x(hpgx_state_t st) { return std::max(1, st->cf.max_sx) - 1; }
int halo_y(hpgx_state_t st) { return std::max(1, st->cf.max_sy) - 1; }

// haloed tiles: supports narrower than a tile, tile + halo within LDS
bool halo_ok(hpgx_state_t st) {
  if (st->gridder != 4 && st->gridder != 5) return false;
  const int hx = halo_x(st), hy = halo_y(st);
  return hx < hpgx::halo_tile_limit() && hy < hpgx::halo_tile_limit()
    && hpgx::halo_lds_bytes(hx, hy, 1) <= 65536;
}

// haloed tiles fed by block-local bins (grids of up to bin_keys_limit keys)
bool use_halo(hpgx_state_t st) { return halo_ok(st) && use_block_local(st); }

// haloed tiles fed by global bins (larger grids: counters are then
// rarely contended)
bool use_halo_global(hpgx_state_t st) {
  return halo_ok(st) && st->nkeys > size_t(hpgx::bin_keys_limit());
}

// the quadrant kernel (gridder 5) for FP32 grids with (hi, lo) pair tiles:
// its regions and tables fit LDS and CF tap indexes fit 32 bits.  FP64 tile
// sums stay with k_tile4, whose non-returning FP64 LDS atomics need no
// read-modify-write per tap.
bool use_quad(hpgx_state_t st) {
  return st->gridder == 5 && (use_halo(st) || use_halo_global(st))
    && st->mode != HPGX_GRID_FP64 && st->fp32_tile == 2
    && hpgx::quad_supported(halo_x(st), halo_y(st), int(st->npol))
    && st->taps_capacity < (size_t(1) << 32);
}

// the quadrant-key kernel (k_sub): FP32-compensated grids on block-local
// haloed bins with supports up to 8 x 8, and 4 bin keys per (plane, tile)
// within one workgroup's LDS histogram
bool use_sub(hpgx_state_t st) {
  return st->sub_tiles && use_quad(st) && use_halo(st)
    && hpgx::sub_supported(halo_x(st), halo_y(st))
    && st->taps_capacity * sizeof(float2) <= size_t(0xffffffffu)
    && 4 * st->nkeys <= size_t(hpgx::bin_quad_keys_limit());
}

// k_sub on global bins (quadrant keys counted and emitted through global
// counters): grids of more keys than the block-local histogram holds, one
// polarization
bool use_sub_global(hpgx_state_t st) {
  return st->sub_tiles && use_quad(st) && use_halo_global(st) && st->npol == 1
    && hpgx::sub_supported(halo_x(st), halo_y(st))
    && st->taps_capacity * sizeof(float2) <= size_t(0xffffffffu);
}

// every CF group has the same support
bool uniform_support(hpgx_state_t st) {
  for (const GroupInfo& g : st->groups)
    if (g.sx != st->groups[0].sx || g.sy != st->groups[0].sy) return false;
  return !st->groups.empty();
}

// bin keys of the block-local pipelines: (plane, tile), or (plane, tile,
// quadrant) for k_sub
size_t bin_keys(hpgx_state_t st) { return use_sub(st) ? 4 * st->nkeys : st->nkeys; }

// bin blocks per haloed-tile pass: the per-block segment tables share the
// workgroup's 64 KiB of LDS with the tile
size_t halo_bins_limit(hpgx_state_t st) {
  size_t nb = size_t(hpgx::tile_bins_limit());
  while (nb > 1 && hpgx::halo_lds_bytes(halo_x(st), halo_y(st), int(nb)) > 65536) nb /= 2;
  if (use_quad(st)) nb = std::min(nb, size_t(hpgx::quad_bins_limit()));
  if (use_sub(st)) nb = std::min(nb, size_t(hpgx::sub_bins_limit()));
  return nb;
}

size_t entries_per_record(hpgx_state_t st) {
  const size_t planes = size_t(std::max(1u, st->gsize[2]));
  if (use_halo(st) || use_halo_global(st)) return planes;
  return planes
    * hpgx::tiles_per_footprint(std::max(1, st->cf.max_sx), std::max(1, st->cf.max_sy));
}

// device memory for bin entries (HPGX_ENTRY_BUDGET_MB; default 16 GiB,
// under 6% of an MI355X's 288 GB: batches whose entries exceed it are
// gridded in passes, each with its own binning and kernel tail -- 4 GiB
// took the A/W-projection workload's call 12% longer, 64 GiB gained under
// 1% more)
constexpr long entry_budget_default_mb = 16384;
size_t entry_budget_bytes() {
  const char* e = std::getenv("HPGX_ENTRY_BUDGET_MB");
  const long mb = e ? std::atol(e) : entry_budget_default_mb;
  return size_t(mb > 0 ? mb : entry_budget_default_mb) << 20;
}

// per-call buffers sized for max_batch records and the current CF supports
int ensure_call_buffers(hpgx_state_t st) {
  CallBuffers& cb = st->cb;
  if (!cb.vis) {
    const size_t n = std::max<size_t>(st->max_batch, 1);
    HPGX_HIP(st, dev_alloc(cb.vis, n * visdata_size(st->npol)));
    HPGX_HIP(st, dev_alloc(cb.gvis, n * st->npol * sizeof(float2)));
    HPGX_HIP(st, dev_alloc(cb.invalid_count, sizeof(unsigned long long)));
    HPGX_HIP(st, hipMemsetAsync(cb.invalid_count, 0, sizeof(unsigned long long), st->stream));
    st->vis_capacity = n;
  }
  if (!cb.hist) {
    const Geometry geo = st->geometry();
    st->nkeys = size_t(geo.ntx) * geo.nty * geo.nr * geo.nc;
    // global bin counts for up to 4 keys per (plane, tile) (quadrant keys)
    const size_t gkeys = 4 * st->nkeys;
    HPGX_HIP(st, dev_alloc(cb.hist, (gkeys + 1) * sizeof(unsigned)));
    HPGX_HIP(st, dev_alloc(cb.offsets, (gkeys + 1) * sizeof(unsigned)));
    HPGX_HIP(st, dev_alloc(cb.cursor, (gkeys + 1) * sizeof(unsigned)));
    cb.scan_temp_bytes = hpgx::scan_temp_bytes(gkeys);
    HPGX_HIP(st, dev_alloc(cb.scan_temp, cb.scan_temp_bytes));
  }
  if (st->cf.ngroups > 0) {
    // bin entries for the whole batch, up to the entry memory budget; larger
    // batches are gridded in chunks (run_grid)
    const size_t epr = entries_per_record(st);
    const bool halo = use_halo(st) || use_halo_global(st);
    const size_t per_entry =
      (halo ? hpgx::halo_entry_bytes(int(st->npol)) : hpgx::entry_bytes(int(st->npol)))
      + sizeof(unsigned);
    // entries per pass: at least one record's, and below 2^32 (bin counts,
    // offsets and entry indexes are 32-bit)
    const size_t budget = std::min<size_t>(std::max(entry_budget_bytes() / per_entry, epr),
                                           0xffffffffull);
    size_t need = std::min(st->vis_capacity * epr, budget);
    if (use_block_local(st)) {
      // whole bin blocks of bin_block_records records each
      const size_t block_cap = hpgx::bin_block_records() * epr;
      const size_t blocks_needed =
        (st->vis_capacity + hpgx::bin_block_records() - 1) / hpgx::bin_block_records();
      const size_t bins_max = halo ? halo_bins_limit(st) : size_t(hpgx::tile_bins_limit());
      const size_t blocks =
        std::max<size_t>(1, std::min({blocks_needed, budget / block_cap, bins_max}));
      need = blocks * block_cap;
      st->pass_blocks = blocks;
      const size_t tab = blocks * (bin_keys(st) + 1);
      if (tab > cb.off_table_capacity) {
        dev_free(cb.off_table);
        cb.off_table_capacity = 0;
        HPGX_HIP(st, dev_alloc(cb.off_table, tab * sizeof(unsigned)));
        cb.off_table_capacity = tab;
      }
    }
    if (halo) {
      const size_t hp = st->nkeys * hpgx::halo_pixels_per_key(halo_x(st), halo_y(st));
      if (hp > cb.halo_capacity) {
        dev_free(cb.halo);
        cb.halo_capacity = 0;
        HPGX_HIP(st, dev_alloc(cb.halo, hp * sizeof(double2)));
        cb.halo_capacity = hp;
      }
      if (st->nkeys > cb.halo_set_capacity) {
        dev_free(cb.halo_set);
        cb.halo_set_capacity = 0;
        HPGX_HIP(st, dev_alloc(cb.halo_set, st->nkeys * sizeof(unsigned)));
        cb.halo_set_capacity = st->nkeys;
      }
    }
    if (use_sub_global(st) && need > cb.tally_capacity) {
      dev_free(cb.tally_entries);
      dev_free(cb.tally_keys);
      cb.tally_capacity = 0;
      HPGX_HIP(st, dev_alloc(cb.tally_entries, need * hpgx::halo_entry_bytes(int(st->npol))));
      HPGX_HIP(st, dev_alloc(cb.tally_keys, need * 2 * sizeof(unsigned)));
      cb.tally_capacity = need;
    }
    // the wave-per-record degridder's record order, once a model is set
    if (st->model && hpgx::degrid_sorts(st->cf) && st->vis_capacity > cb.order_capacity) {
      dev_free(cb.degrid_order);
      dev_free(cb.order_temp);
      cb.order_capacity = 0;
      cb.order_temp_bytes = hpgx::degrid_order_temp_bytes(st->vis_capacity);
      HPGX_HIP(st, dev_alloc(cb.degrid_order, 4 * st->vis_capacity * sizeof(unsigned)));
      HPGX_HIP(st, dev_alloc(cb.order_temp, cb.order_temp_bytes));
      cb.order_capacity = st->vis_capacity;
    }
    if (need > cb.entries_capacity) {
      dev_free(cb.stream_entries);
      dev_free(cb.entry_vis);
      cb.entries_capacity = 0;
      HPGX_HIP(st, dev_alloc(cb.stream_entries,
                             need * std::max(hpgx::entry_bytes(int(st->npol)),
                                             hpgx::halo_entry_bytes(int(st->npol)))));
      HPGX_HIP(st, dev_alloc(cb.entry_vis, need * sizeof(unsigned)));
      cb.entries_capacity = need;
    }
  }
  return HPGX_SUCCESS;
}

int ensure_scratch(hpgx_state_t st) {
  if (!st->scratch) HPGX_HIP(st, dev_alloc(st->scratch, st->grid_bytes()));
  return HPGX_SUCCESS;
}

void free_state(hpgx_state_t st) {
  if (st->stream) (void)hipStreamSynchronize(st->stream);
  dev_free(st->grid);
  dev_free(st->weights);
  dev_free(st->model);
  dev_free(st->scratch);
  dev_free(st->cf.groups);
  dev_free(st->cf.taps);
  dev_free(st->cf.wsum);
  dev_free(st->staging);
  CallBuffers& cb = st->cb;
  for (size_t k = 1; k < st->task_vis.size(); ++k) dev_free(st->task_vis[k]);
  if (!st->task_vis.empty()) cb.vis = st->task_vis[0];
  for (hipStream_t cs : st->copy_streams)
    if (cs) {
      (void)hipStreamSynchronize(cs);
      (void)hipStreamDestroy(cs);
    }
  for (hipEvent_t e : st->copied)
    if (e) (void)hipEventDestroy(e);
  for (hipEvent_t e : st->released)
    if (e) (void)hipEventDestroy(e);
  dev_free(cb.vis);
  dev_free(cb.gvis);
  dev_free(cb.tally_entries);
  dev_free(cb.tally_keys);
  dev_free(cb.hist);
  dev_free(cb.offsets);
  dev_free(cb.cursor);
  dev_free(cb.stream_entries);
  dev_free(cb.entry_vis);
  dev_free(cb.off_table);
  dev_free(cb.halo);
  dev_free(cb.halo_set);
  dev_free(cb.degrid_acc);
  dev_free(cb.scan_temp);
  dev_free(cb.degrid_order);
  dev_free(cb.order_temp);
  dev_free(cb.invalid_count);
  if (st->plan_z2z.valid) hipfftDestroy(st->plan_z2z.h);
  if (st->plan_c2c.valid) hipfftDestroy(st->plan_c2c.h);
  for (auto& e : st->events)
    if (e) (void)hipEventDestroy(e);
  if (st->stream) (void)hipStreamDestroy(st->stream);
}

int set_device(hpgx_state_t st) {
  HPGX_HIP(st, hipSetDevice(st->device));
  return HPGX_SUCCESS;
}

int zero_grid(hpgx_state_t st) {
  HPGX_HIP(st, hipMemsetAsync(st->grid, 0, st->grid_bytes(), st->stream));
  HPGX_HIP(st, hipMemsetAsync(
                 st->weights, 0, sizeof(double) * st->gsize[2] * st->gsize[3],
                 st->stream));
  return HPGX_SUCCESS;
}

// host-side transposes between Layout::Left (x fastest) and Layout::Right
// (cube fastest) of a (x, y, mrow, cube) complex<double> array
void left_to_right(const double2* src, double2* dst, const unsigned* g) {
  const size_t nx = g[0], ny = g[1], nr = g[2], nc = g[3];
  for (size_t c = 0; c < nc; ++c)
    for (size_t r = 0; r < nr; ++r)
      for (size_t y = 0; y < ny; ++y)
        for (size_t x = 0; x < nx; ++x)
          dst[((x * ny + y) * nr + r) * nc + c] = src[((c * nr + r) * ny + y) * nx + x];
}
void right_to_left(const double2* src, double2* dst, const unsigned* g) {
  const size_t nx = g[0], ny = g[1], nr = g[2], nc = g[3];
  for (size_t c = 0; c < nc; ++c)
    for (size_t r = 0; r < nr; ++r)
      for (size_t y = 0; y < ny; ++y)
        for (size_t x = 0; x < nx; ++x)
          dst[((c * nr + r) * ny + y) * nx + x] = src[((x * ny + y) * nr + r) * nc + c];
}

// copies a device Layout::Left grid-shaped array (double2 or float2) to host
// complex<double> in `layout`
int copy_grid_to_host(
  hpgx_state_t st, const void* dev, bool is_float, void* dst, int layout) {
  const size_t n = st->grid_elems();
  const void* src = dev;
  if (is_float) {
    // widen to complex<double> on the device in a temporary buffer
    double2* tmp = nullptr;
    HPGX_HIP(st, dev_alloc(tmp, n * sizeof(double2)));
    hipError_t e = hpgx::launch_widen(st->stream, static_cast<const float2*>(dev), tmp, n);
    if (e == hipSuccess) {
      if (layout == HPGX_LAYOUT_LEFT) {
        e = hipMemcpyAsync(dst, tmp, n * sizeof(double2), hipMemcpyDeviceToHost, st->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(st->stream);
      } else {
        std::vector<double2> h(n);
        e = hipMemcpyAsync(h.data(), tmp, n * sizeof(double2), hipMemcpyDeviceToHost, st->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(st->stream);
        if (e == hipSuccess) left_to_right(h.data(), static_cast<double2*>(dst), st->gsize);
      }
    }
    (void)hipStreamSynchronize(st->stream);
    dev_free(tmp);
    if (e != hipSuccess) return hip_fail(st, e, "grid readback");
    return HPGX_SUCCESS;
  }
  if (layout == HPGX_LAYOUT_LEFT) {
    HPGX_HIP(st, hipMemcpyAsync(dst, src, n * sizeof(double2), hipMemcpyDeviceToHost, st->stream));
    HPGX_HIP(st, hipStreamSynchronize(st->stream));
  } else {
    std::vector<double2> h(n);
    HPGX_HIP(st, hipMemcpyAsync(h.data(), src, n * sizeof(double2), hipMemcpyDeviceToHost, st->stream));
    HPGX_HIP(st, hipStreamSynchronize(st->stream));
    left_to_right(h.data(), static_cast<double2*>(dst), st->gsize);
  }
  return HPGX_SUCCESS;
}

int get_plan(hpgx_state_t st, bool fp64, hipfftHandle& out) {
  Plan& p = fp64 ? st->plan_z2z : st->plan_c2c;
  if (!p.valid) {
    int n[2] = {int(st->gsize[1]), int(st->gsize[0])};  // y slowest, x fastest
    const int batch = int(st->gsize[2] * st->gsize[3]);
    const int dist = n[0] * n[1];
    hipfftResult r = hipfftPlanMany(
      &p.h, 2, n, nullptr, 1, dist, nullptr, 1, dist,
      fp64 ? HIPFFT_Z2Z : HIPFFT_C2C, batch);
    if (r != HIPFFT_SUCCESS)
      return fail(st, HPGX_ERR_OTHER, "hipfftPlanMany failed (" + std::to_string(int(r)) + ")");
    p.valid = true;
  }
  if (hipfftSetStream(p.h, st->stream) != HIPFFT_SUCCESS)
    return fail(st, HPGX_ERR_OTHER, "hipfftSetStream failed");
  out = p.h;
  return HPGX_SUCCESS;
}

// unnormalized 2-D FFT of every plane of a Layout::Left grid-shaped array
int fft_planes(hpgx_state_t st, void* data, bool fp64, double norm, int sign,
               bool in_place, void** result) {
  hipfftHandle h;
  if (int rc = get_plan(st, fp64, h)) return rc;
  const int dir = sign == HPGX_FFT_NEGATIVE ? HIPFFT_FORWARD : HIPFFT_BACKWARD;
  void* out = data;
  if (!in_place) {
    const size_t bytes = st->grid_elems() * (fp64 ? sizeof(double2) : sizeof(float2));
    HPGX_HIP(st, dev_alloc(out, bytes));
  }
  hipfftResult r = fp64
    ? hipfftExecZ2Z(h, static_cast<hipfftDoubleComplex*>(data),
                    static_cast<hipfftDoubleComplex*>(out), dir)
    : hipfftExecC2C(h, static_cast<hipfftComplex*>(data),
                    static_cast<hipfftComplex*>(out), dir);
  if (r != HIPFFT_SUCCESS) {
    if (out != data) {
      (void)hipStreamSynchronize(st->stream);
      (void)hipFree(out);
    }
    return fail(st, HPGX_ERR_OTHER, "hipfftExec failed (" + std::to_string(int(r)) + ")");
  }
  if (out != data) {
    // the reference allocates a new grid for an out-of-place FFT
    // (src/hpg_runtime.hpp:1331-1347); the old one is released
    HPGX_HIP(st, hipStreamSynchronize(st->stream));
    (void)hipFree(data);
  }
  // GridNormalizer with norm (src/hpg_core.hpp:1863-1889)
  if (norm != 1) {
    HPGX_HIP(st, hpgx::launch_scale(
                   st->stream, fp64 ? 0 : 1, out, 1.0 / norm, st->grid_elems()));
  }
  *result = out;
  return HPGX_SUCCESS;
}

// Runs the deferred finishing chain.  With an FFT pending, normalize (the
// FP64 division by the weights), the FFT's 1 / norm and, with a shift
// pending, the (-1)^(x + y) modulation that stands for the shift of the
// FFT's output (even sizes: the FFT of x (-1)^(x+y) is the FFT of x shifted
// by half the grid on both axes) are one pass before the FFT
// (src/hpg_core.hpp:1849-1860, 1863-1889, 2329-2355); without one,
// normalize runs alone.
int flush_finishing(hpgx_state_t st) {
  if (!st->pend_norm && !st->pend_fft && !st->pend_shift) return HPGX_SUCCESS;
  const bool norm = st->pend_norm, fft = st->pend_fft, shift = st->pend_shift;
  st->pend_norm = st->pend_fft = st->pend_shift = false;
  const int mode = st->mode == HPGX_GRID_FP64 ? 0 : 1;
  if (!fft) {
    HPGX_HIP(st, hpgx::launch_normalize(st->stream, mode, st->grid, st->weights,
                                        st->pend_wfactor, st->geometry()));
    return HPGX_SUCCESS;
  }
  const Geometry geo = st->geometry();
  if (!hpgx::finish_pre_fits(geo)) {
    // the separate passes
    if (norm)
      HPGX_HIP(st, hpgx::launch_normalize(st->stream, mode, st->grid, st->weights,
                                          st->pend_wfactor, geo));
    void* out = nullptr;
    if (int rc = fft_planes(st, st->grid, mode == 0, st->pend_fft_norm, st->pend_fft_sign,
                            true, &out))
      return rc;
    st->grid = out;
    if (shift) {
      if (int rc = ensure_scratch(st)) return rc;
      HPGX_HIP(st, hpgx::launch_shift(st->stream, mode, st->grid, st->scratch,
                                      HPGX_SHIFT_FORWARD, geo));
    }
    return HPGX_SUCCESS;
  }
  const bool pre = norm || shift;
  if (pre) {
    HPGX_HIP(st, hpgx::launch_finish_pre(
                   st->stream, mode, st->grid, norm ? st->weights : nullptr, st->pend_wfactor,
                   1.0 / st->pend_fft_norm, shift, geo));
  }
  void* out = nullptr;
  if (int rc = fft_planes(st, st->grid, mode == 0, pre ? 1.0 : st->pend_fft_norm,
                          st->pend_fft_sign, true, &out))
    return rc;
  st->grid = out;
  return HPGX_SUCCESS;
}

int upload_model(hpgx_state_t st, const void* values, int layout, int memory) {
  const size_t n = st->grid_elems();
  if (!st->model) HPGX_HIP(st, dev_alloc(st->model, n * sizeof(double2)));
  if (layout == HPGX_LAYOUT_LEFT) {
    HPGX_HIP(st, hipMemcpyAsync(
                   st->model, values, n * sizeof(double2),
                   memory == HPGX_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice,
                   st->stream));
    HPGX_HIP(st, hipStreamSynchronize(st->stream));
    return HPGX_SUCCESS;
  }
  std::vector<double2> right(n), left(n);
  if (memory == HPGX_DEVICE) {
    HPGX_HIP(st, hipMemcpy(right.data(), values, n * sizeof(double2), hipMemcpyDeviceToHost));
  } else {
    std::memcpy(right.data(), values, n * sizeof(double2));
  }
  right_to_left(right.data(), left.data(), st->gsize);
  HPGX_HIP(st, hipMemcpyAsync(st->model, left.data(), n * sizeof(double2),
                              hipMemcpyHostToDevice, st->stream));
  HPGX_HIP(st, hipStreamSynchronize(st->stream));
  return HPGX_SUCCESS;
}

// gridding / degridding pipeline on records already in cb.vis
// gridding / degridding pipeline on records already in cb.vis; the records
// are gridded in chunks whose bin entries fit the entry buffer
// binned degridding of n records (cb.vis, or `records` read in place):
// FP64 predictions into pred64, or predicted values (residuals when
// do_grid) written into `out_records` (cb.vis when null and records is
// null; not written when only records is given), gvis when do_grid.
// Requires use_halo(st) and a model.
int run_degrid_binned(hpgx_state_t st, size_t n, bool do_grid, const void* records,
                      double2* pred64, void* out_records = nullptr) {
  const Geometry geo = st->geometry();
  CallBuffers& cb = st->cb;
  const int np = int(st->npol);
  const size_t acc_need = std::max<size_t>(st->max_batch, 1) * np;
  if (acc_need > cb.degrid_acc_capacity) {
    dev_free(cb.degrid_acc);
    cb.degrid_acc_capacity = 0;
    HPGX_HIP(st, dev_alloc(cb.degrid_acc, acc_need * sizeof(double2)));
    cb.degrid_acc_capacity = acc_need;
  }
  // one grid row whose Mueller rows (both w signs) index some CF element:
  // every record inside the grid has exactly one entry, and the tile kernel
  // finishes its prediction
  bool one_plane = st->gsize[2] == 1;
  if (one_plane) {
    bool m_any = false, mc_any = false;
    for (unsigned c = 0; c < st->npol; ++c) {
      m_any |= st->mu.M[0][c] >= 0;
      mc_any |= st->mu.Mc[0][c] >= 0;
    }
    one_plane = m_any && mc_any;
  }
  if (!one_plane)
    HPGX_HIP(st, hipMemsetAsync(cb.degrid_acc, 0, n * np * sizeof(double2), st->stream));
  const size_t epr = entries_per_record(st);
  const size_t block_cap = hpgx::bin_block_records() * epr;
  const size_t chunk = std::max<size_t>(1, st->pass_blocks) * hpgx::bin_block_records();
  const size_t rec_bytes = visdata_size(st->npol);
  const void* rec_base = records ? records : cb.vis;
  void* out_base = out_records ? out_records : (records ? nullptr : cb.vis);
  // records returned into another buffer: k_bin writes them there whole
  const bool copy_out = out_base && out_base != rec_base;
  for (size_t off = 0; off < n; off += chunk) {
    const size_t m = std::min(chunk, n - off);
    CallBuffers c = cb;
    c.vis = const_cast<char*>(static_cast<const char*>(rec_base)) + off * rec_bytes;
    c.gvis = cb.gvis + off * np;
    HPGX_HIP(st, hpgx::launch_bin(
                   st->stream, np, m, geo, st->cf, st->mu, false, false,
                   hpgx::GvisSource::Values, nullptr, c, block_cap, st->weights, true, true,
                   copy_out ? static_cast<char*>(out_base) + off * rec_bytes : nullptr));
    HPGX_HIP(st, hpgx::launch_degrid_tiles(
                   st->stream, np, m, st->nkeys, geo, st->cf, st->mu, c, block_cap,
                   halo_x(st), halo_y(st), st->model, cb.degrid_acc + off * np, one_plane,
                   do_grid, out_base ? static_cast<char*>(out_base) + off * rec_bytes : nullptr,
                   pred64 ? pred64 + off * np : nullptr));
  }
  if (!one_plane)
    HPGX_HIP(st, hpgx::launch_degrid_finish(
                   st->stream, np, n, geo, st->cf, st->mu, cb.degrid_acc, do_grid, rec_base,
                   out_base, cb, pred64));
  return HPGX_SUCCESS;
}

// records: device records read in place (nullptr: the staged copy cb.vis);
// out_records: where the binned degridder writes predicted / residual
// values when records are read in place (nullptr: not written)
// cb: the per-call buffers (the state's own, or views into the fixed
// scratch of hpgx_grid_visibilities_noalloc); pass_blocks: bin blocks per
// pass of the block-local pipelines
int run_grid_with(hpgx_state_t st, CallBuffers& cb, size_t pass_blocks, size_t n,
                  bool update_weights, bool do_degrid, bool do_grid, hpgx::GvisSource src,
                  const double2* residuals, const void* records, void* out_records) {
  const Geometry geo = st->geometry();
  void* rec_base = records ? const_cast<void*>(records) : cb.vis;
  const int np = int(st->npol);
  bool gvis_ready = false;
  if (do_degrid) {
    if (st->model && use_halo(st)) {
      if (int rc = run_degrid_binned(st, n, do_grid, records, nullptr, out_records)) return rc;
    } else {
      HPGX_HIP(st, hpgx::launch_degrid(
                     st->stream, np, n, geo, st->cf, st->mu, st->model, do_grid, cb));
    }
    gvis_ready = true;
  }
  if (!do_grid) return HPGX_SUCCESS;
  const int mode = st->mode == HPGX_GRID_FP64 ? 0 : 1;
  const size_t epr = entries_per_record(st);
  if (use_block_local(st)) {
    const bool halo = use_halo(st);
    const size_t block_cap = hpgx::bin_block_records() * epr;
    // bin blocks per pass as sized by ensure_call_buffers (entries, offset
    // table and the tile kernel's LDS segment tables)
    const size_t chunk = std::max<size_t>(1, pass_blocks) * hpgx::bin_block_records();
    const size_t rec_bytes = visdata_size(st->npol);
    for (size_t off = 0; off < n; off += chunk) {
      const size_t m = std::min(chunk, n - off);
      CallBuffers c = cb;
      c.vis = static_cast<char*>(rec_base) + off * rec_bytes;
      c.gvis = cb.gvis ? cb.gvis + off * np : nullptr;
      const double2* res = residuals ? residuals + off * np : nullptr;
      const bool sub = use_sub(st);
      HPGX_HIP(st, hpgx::launch_bin(
                     st->stream, np, m, geo, st->cf, st->mu, update_weights, gvis_ready, src,
                     res, c, block_cap, st->weights, halo, false, nullptr, sub));
      if (sub)
        HPGX_HIP(st, hpgx::launch_tile_sub(
                       st->stream, np, m, st->nkeys, geo, st->cf,
                       st->taps_capacity * sizeof(float2), uniform_support(st), st->mu, c,
                       block_cap, halo_x(st), halo_y(st), st->grid));
      else if (halo)
        HPGX_HIP(st, hpgx::launch_tile_halo(
                       st->stream, np, mode, st->fp32_tile, m, st->nkeys, geo, st->cf, st->mu, c,
                       block_cap, halo_x(st), halo_y(st), st->grid, false, use_quad(st)));
      else
        HPGX_HIP(st, hpgx::launch_tile_blocal(
                       st->stream, np, mode, m, st->nkeys, geo, st->cf, st->mu, c,
                       block_cap, st->grid));
    }
    return HPGX_SUCCESS;
  }
  // records per pass: their entries fit the entry buffer (and, through
  // launch_tally, its record-order slots)
  const size_t slots =
    use_sub_global(st) ? std::min(cb.entries_capacity, cb.tally_capacity) : cb.entries_capacity;
  if (slots < epr) return fail(st, HPGX_ERR_OTHER, "bin entry buffers not allocated");
  const size_t chunk = slots / epr;
  const size_t rec_bytes = visdata_size(st->npol);
  for (size_t off = 0; off < n; off += chunk) {
    const size_t m = std::min(chunk, n - off);
    CallBuffers c = cb;
    c.vis = static_cast<char*>(rec_base) + off * rec_bytes;
    c.gvis = cb.gvis + off * np;
    const double2* res = residuals ? residuals + off * np : nullptr;
    const bool halo = use_halo_global(st);
    const bool sub = use_sub_global(st);
    const size_t keys = sub ? 4 * st->nkeys : st->nkeys;
    HPGX_HIP(st, hipMemsetAsync(c.hist, 0, (keys + 1) * sizeof(unsigned), st->stream));
    if (sub) {
      HPGX_HIP(st, hpgx::launch_tally(
                     st->stream, np, m, geo, st->cf, st->mu, update_weights, gvis_ready, src,
                     res, c, st->weights));
      HPGX_HIP(st, hpgx::launch_scan(st->stream, keys, c));
      HPGX_HIP(st, hpgx::launch_place(st->stream, np, m, geo, c));
    } else {
      HPGX_HIP(st, hpgx::launch_count(
                     st->stream, np, m, geo, st->cf, st->mu, update_weights, true, c,
                     st->weights, halo));
      HPGX_HIP(st, hpgx::launch_scan(st->stream, keys, c));
      HPGX_HIP(st, hpgx::launch_emit(
                     st->stream, np, m, geo, st->cf, st->mu, gvis_ready, src, res, c, halo));
    }
    if (sub)
      HPGX_HIP(st, hpgx::launch_tile_sub(
                     st->stream, np, m, st->nkeys, geo, st->cf,
                     st->taps_capacity * sizeof(float2), uniform_support(st), st->mu, c, 0,
                     halo_x(st), halo_y(st), st->grid, true));
    else if (halo)
      HPGX_HIP(st, hpgx::launch_tile_halo(
                     st->stream, np, mode, st->fp32_tile, m, st->nkeys, geo, st->cf, st->mu,
                     c, 0, halo_x(st), halo_y(st), st->grid, true, use_quad(st)));
    else
      HPGX_HIP(st, hpgx::launch_tile_stream(
                     st->stream, np, mode, st->nkeys, geo, st->cf, st->mu, c, st->grid));
  }
  return HPGX_SUCCESS;
}

int run_grid(hpgx_state_t st, size_t n, bool update_weights, bool do_degrid,
             bool do_grid, hpgx::GvisSource src, const double2* residuals,
             const void* records = nullptr, void* out_records = nullptr) {
  return run_grid_with(st, st->cb, st->pass_blocks, n, update_weights, do_degrid, do_grid, src,
                       residuals, records, out_records);
}

// Views into the fixed scratch for gridding n records without allocating:
// bin blocks per pass as many as fit (at most what one pass of the
// pipeline takes), the records gridded in as many passes as needed.
// Nothing in these buffers has to start at zero: k_bin writes each offset
// table row and entry it produces before the tile kernel reads them, and
// the tile kernel writes every key's halo flag before k_halo reads it.
int noalloc_plan(hpgx_state_t st, void* base, size_t n, CallBuffers& c, size_t& pass_blocks) {
  if (!use_block_local(st))
    return fail(st, HPGX_ERR_OTHER,
                "hpgx_grid_visibilities_noalloc: grids of more than "
                  + std::to_string(hpgx::bin_keys_limit())
                  + " (plane, 32x32 tile) keys are not supported (this grid has "
                  + std::to_string(st->nkeys) + ")");
  const bool halo = use_halo(st);
  const size_t epr = entries_per_record(st);
  const size_t block_cap = hpgx::bin_block_records() * epr;
  const size_t esize =
    halo ? hpgx::halo_entry_bytes(int(st->npol)) : hpgx::entry_bytes(int(st->npol));
  auto up = [](size_t b) { return (b + 255) & ~size_t(255); };
  const size_t halo_bytes =
    halo ? up(st->nkeys * hpgx::halo_pixels_per_key(halo_x(st), halo_y(st)) * sizeof(double2))
    : 0;
  const size_t set_bytes = halo ? up(st->nkeys * sizeof(unsigned)) : 0;
  const size_t per_block =
    (bin_keys(st) + 1) * sizeof(unsigned) + block_cap * esize + block_cap * sizeof(unsigned);
  const size_t fixed = halo_bytes + set_bytes + 3 * 256;
  const size_t fit = noalloc_scratch_bytes > fixed ? (noalloc_scratch_bytes - fixed) / per_block : 0;
  const size_t bins_max = halo ? halo_bins_limit(st) : size_t(hpgx::tile_bins_limit());
  const size_t needed = (n + hpgx::bin_block_records() - 1) / hpgx::bin_block_records();
  const size_t blocks = std::min({needed, bins_max, fit});
  if (blocks == 0)
    return fail(st, HPGX_ERR_OTHER,
                "hpgx_grid_visibilities_noalloc: one bin block of this grid and CF needs "
                  + std::to_string(fixed + per_block) + " bytes of scratch; the fixed scratch holds "
                  + std::to_string(noalloc_scratch_bytes));
  char* p = static_cast<char*>(base);
  c = CallBuffers{};
  c.off_table = reinterpret_cast<unsigned*>(p);
  c.off_table_capacity = blocks * (bin_keys(st) + 1);
  p += up(c.off_table_capacity * sizeof(unsigned));
  c.stream_entries = p;
  c.entries_capacity = blocks * block_cap;
  p += up(c.entries_capacity * esize);
  c.entry_vis = reinterpret_cast<unsigned*>(p);
  p += up(c.entries_capacity * sizeof(unsigned));
  if (halo) {
    c.halo = reinterpret_cast<double2*>(p);
    c.halo_capacity = st->nkeys * hpgx::halo_pixels_per_key(halo_x(st), halo_y(st));
    p += halo_bytes;
    c.halo_set = reinterpret_cast<unsigned*>(p);
    c.halo_set_capacity = st->nkeys;
    p += set_bytes;
  }
  if (size_t(p - static_cast<char*>(base)) > noalloc_scratch_bytes)
    return fail(st, HPGX_ERR_OTHER, "hpgx_grid_visibilities_noalloc: scratch layout overflow");
  // skipped records are counted in the state's own counter (read by
  // hpgx_fence), as by hpgx_grid_visibilities
  c.invalid_count = st->cb.invalid_count;
  pass_blocks = blocks;
  return HPGX_SUCCESS;
}

// tasks a state keeps in rotation at most (max_added_tasks + 1 is clamped
// to it; the reference clamps to its device's active_task_limit,
// src/hpg_runtime.hpp:209-265, 758-770)
constexpr unsigned task_limit = 16;

int create_tasks(hpgx_state_t st, unsigned tasks) {
  st->tasks = tasks;
  st->task_vis.assign(tasks, nullptr);
  st->task_vis[0] = st->cb.vis;
  st->copy_streams.assign(tasks, nullptr);
  st->copied.assign(tasks, nullptr);
  st->released.assign(tasks, nullptr);
  st->released_set.assign(tasks, false);
  if (tasks < 2) return HPGX_SUCCESS;
  for (unsigned k = 0; k < tasks; ++k) {
    HPGX_HIP(st, hipStreamCreateWithFlags(&st->copy_streams[k], hipStreamNonBlocking));
    HPGX_HIP(st, hipEventCreateWithFlags(&st->copied[k], hipEventDisableTiming));
    HPGX_HIP(st, hipEventCreateWithFlags(&st->released[k], hipEventDisableTiming));
  }
  return HPGX_SUCCESS;
}

// the kernels queued so far may read the current task's record buffer: its
// next upload waits for them
int mark_task_use(hpgx_state_t st) {
  if (st->tasks < 2) return HPGX_SUCCESS;
  HPGX_HIP(st, hipEventRecord(st->released[st->task], st->stream));
  st->released_set[st->task] = true;
  return HPGX_SUCCESS;
}

// host records of the next call: uploaded into the next task's buffer on
// that task's copy stream once the kernels that last read the buffer are
// done; the state's stream waits for the upload.  The call returns when the
// upload is complete (the caller may then reuse its host memory, pinned or
// not), while the previous calls' kernels go on running.
int upload_task_records(hpgx_state_t st, const void* vis, size_t n) {
  const unsigned k = (st->task + 1) % st->tasks;
  const size_t rec_bytes = visdata_size(st->npol);
  if (!st->task_vis[k]) {
    if (k == 0) {
      st->task_vis[0] = st->cb.vis;
    } else {
      HPGX_HIP(st, dev_alloc(st->task_vis[k], std::max<size_t>(st->max_batch, 1) * rec_bytes));
    }
  }
  if (st->released_set[k]) HPGX_HIP(st, hipStreamWaitEvent(st->copy_streams[k], st->released[k], 0));
  if (n > 0)
    HPGX_HIP(st, hipMemcpyAsync(st->task_vis[k], vis, n * rec_bytes, hipMemcpyHostToDevice,
                                st->copy_streams[k]));
  HPGX_HIP(st, hipEventRecord(st->copied[k], st->copy_streams[k]));
  HPGX_HIP(st, hipStreamWaitEvent(st->stream, st->copied[k], 0));
  HPGX_HIP(st, hipEventSynchronize(st->copied[k]));
  st->task = k;
  st->cb.vis = st->task_vis[k];
  return HPGX_SUCCESS;
}

int stage_records(hpgx_state_t st, const void* vis, size_t n, int memory) {
  const size_t bytes = n * visdata_size(st->npol);
  if (bytes == 0) return HPGX_SUCCESS;
  HPGX_HIP(st, hipMemcpyAsync(
                 st->cb.vis, vis, bytes,
                 memory == HPGX_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice,
                 st->stream));
  // a host source is not kept alive by the caller beyond this call
  if (memory != HPGX_DEVICE) HPGX_HIP(st, hipStreamSynchronize(st->stream));
  return HPGX_SUCCESS;
}

}  // namespace

// ---------------------------------------------------------------------------

extern "C" {

int hpgx_initialize(int device_id) {
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count == 0) {
    std::lock_guard<std::mutex> lk(g_create_mtx);
    g_create_error = "no HIP device available";
    return HPGX_ERR_DISABLED_DEVICE;
  }
  if (device_id >= count) return HPGX_ERR_DISABLED_DEVICE;
  if (device_id >= 0) {
    if (hipSetDevice(device_id) != hipSuccess) return HPGX_ERR_DISABLED_DEVICE;
    g_device = device_id;
  } else {
    int d = 0;
    (void)hipGetDevice(&d);
    g_device = d;
  }
  void* p = nullptr;
  const hipError_t e = noalloc_scratch(g_device, &p);
  if (e != hipSuccess) {
    std::lock_guard<std::mutex> lk(g_create_mtx);
    g_create_error = std::string("scratch address of the allocation-free entry point: ")
      + hipGetErrorString(e);
    return HPGX_ERR_OTHER;
  }
  return HPGX_SUCCESS;
}

void hpgx_finalize(void) {}

int hpgx_device_count(void) {
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess) return 0;
  return count;
}

const char* hpgx_version(void) { return "hpgx 0.1.0 (gfx950)"; }

int hpgx_create(
  hpgx_state_t* out, unsigned max_added_tasks, size_t max_visibility_batch_size,
  const unsigned grid_size[4], const double grid_scale[2],
  unsigned num_polarizations, unsigned num_mueller_rows,
  const int* mueller_indexes, const int* conjugate_mueller_indexes,
  int grid_mode, unsigned init_cf_oversampling, unsigned init_cf_num_groups,
  const unsigned* init_cf_extents) {
  auto create_fail = [](int code, const std::string& msg) {
    std::lock_guard<std::mutex> lk(g_create_mtx);
    g_create_error = msg;
    return code;
  };
  if (!out) return create_fail(HPGX_ERR_OTHER, "null state pointer");
  *out = nullptr;
  // GridderState::create validation order (src/hpg.cc:238-263)
  if (num_mueller_rows != grid_size[2])
    return create_fail(
      HPGX_ERR_INVALID_NUMBER_MUELLER_INDEX_ROWS,
      "Number of rows of Mueller indexes does not equal grid mrow extent");
  if (hpgx_device_count() == 0)
    return create_fail(HPGX_ERR_DISABLED_DEVICE, "Requested device is not enabled");
  if (num_polarizations < 1 || num_polarizations > 4)
    return create_fail(HPGX_ERR_INVALID_NUMBER_POLARIZATIONS,
                       "number of polarizations must be 1..4");
  if (num_mueller_rows > 4)
    return create_fail(HPGX_ERR_INVALID_NUMBER_MUELLER_INDEX_ROWS,
                       "at most 4 Mueller index rows are supported");
  if (grid_mode != HPGX_GRID_FP64 && grid_mode != HPGX_GRID_FP32_COMPENSATED)
    return create_fail(HPGX_ERR_OTHER, "unknown grid mode");
  for (int d = 0; d < 4; ++d)
    if (grid_size[d] == 0 || grid_size[d] > (1u << 30))
      return create_fail(HPGX_ERR_OTHER, "grid extents must be in [1, 2^30]");

  auto* st = new (std::nothrow) hpgx_state_s();
  if (!st) return create_fail(HPGX_ERR_OTHER, "out of host memory");
  int dev = g_device;
  if (dev < 0) (void)hipGetDevice(&dev);
  st->device = dev;
  // max_added_tasks + 1 tasks, at most task_limit (reported back clamped)
  st->max_added_tasks = std::min(max_added_tasks, task_limit - 1);
  st->max_batch = max_visibility_batch_size;
  std::memcpy(st->gsize, grid_size, sizeof(st->gsize));
  st->gscale[0] = grid_scale[0];
  st->gscale[1] = grid_scale[1];
  st->npol = num_polarizations;
  st->mode = grid_mode;
  if (const char* a = std::getenv("HPGX_FP32_TILE")) {
    const std::string v(a);
    st->fp32_tile = v == "fp64" ? 0 : 2;
  }
  if (const char* a = std::getenv("HPGX_SUB_TILES")) st->sub_tiles = std::atoi(a) != 0;
  if (const char* g = std::getenv("HPGX_GRIDDER")) {
    const int v = std::atoi(g);
    st->gridder = (v >= 2 && v <= 5) ? v : 5;
  }
  // Mueller index matrices padded to 4x4 with -1
  // (copy_mueller_indexes, src/hpg_runtime.hpp:1590-1614)
  for (int r = 0; r < 4; ++r)
    for (int c = 0; c < 4; ++c) {
      const bool in = unsigned(r) < num_mueller_rows && unsigned(c) < num_polarizations;
      st->mu.M[r][c] = in ? mueller_indexes[r * num_polarizations + c] : -1;
      st->mu.Mc[r][c] = in ? conjugate_mueller_indexes[r * num_polarizations + c] : -1;
    }
  st->mueller_rows.assign(mueller_indexes, mueller_indexes + num_mueller_rows * num_polarizations);
  st->conj_mueller_rows.assign(
    conjugate_mueller_indexes, conjugate_mueller_indexes + num_mueller_rows * num_polarizations);

  int rc = HPGX_SUCCESS;
  hipError_t e = hipSetDevice(st->device);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&st->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = dev_alloc(st->grid, st->grid_bytes());
  if (e == hipSuccess)
    e = dev_alloc(st->weights, sizeof(double) * grid_size[2] * grid_size[3]);
  if (e != hipSuccess) rc = hip_fail(st, e, "state allocation");
  if (!rc) rc = zero_grid(st);
  if (!rc && init_cf_num_groups > 0 && init_cf_extents && init_cf_oversampling > 0)
    rc = hpgx_allocate_convolution_function_region(
      st, init_cf_oversampling, init_cf_num_groups, init_cf_extents);
  if (!rc) rc = ensure_call_buffers(st);
  if (!rc) rc = create_tasks(st, st->max_added_tasks + 1);
  if (!rc) {
    e = hipStreamSynchronize(st->stream);
    if (e != hipSuccess) rc = hip_fail(st, e, "state initialization");
  }
  if (rc) {
    std::string msg = st->err;
    free_state(st);
    delete st;
    return create_fail(rc, msg);
  }
  *out = st;
  return HPGX_SUCCESS;
}

int hpgx_clone(hpgx_state_t src, hpgx_state_t* dst) {
  if (!src || !dst) return HPGX_ERR_OTHER;
  *dst = nullptr;
  if (int rc = set_device(src)) return rc;
  if (int rc = flush_finishing(src)) return rc;
  HPGX_HIP(src, hipStreamSynchronize(src->stream));
  hpgx_state_t st = nullptr;
  std::vector<unsigned> ext;
  int rc = hpgx_create(
    &st, src->max_added_tasks, src->max_batch, src->gsize, src->gscale, src->npol,
    src->gsize[2], src->mueller_rows.data(), src->conj_mueller_rows.data(),
    src->mode, 0, 0, nullptr);
  if (rc) return fail(src, rc, "clone: " + std::string(g_create_error));
  st->gridder = src->gridder;
  st->fp32_tile = src->fp32_tile;
  auto copy = [&](void* d, const void* s, size_t bytes) {
    return hipMemcpyAsync(d, s, bytes, hipMemcpyDeviceToDevice, st->stream);
  };
  hipError_t e = copy(st->grid, src->grid, src->grid_bytes());
  if (e == hipSuccess)
    e = copy(st->weights, src->weights, sizeof(double) * src->gsize[2] * src->gsize[3]);
  if (e == hipSuccess && src->model) {
    e = dev_alloc(st->model, src->grid_elems() * sizeof(double2));
    if (e == hipSuccess) e = copy(st->model, src->model, src->grid_elems() * sizeof(double2));
  }
  if (e == hipSuccess && src->cf.ngroups > 0) {
    rc = ensure_cf_capacity(st, src->taps_capacity, src->wsum_capacity);
    if (rc) {
      const std::string msg = "clone: " + st->err;
      hpgx_destroy(st);
      return fail(src, rc, msg);
    }
    e = copy(st->cf.taps, src->cf.taps, src->taps_capacity * sizeof(float2));
    if (e == hipSuccess)
      e = copy(st->cf.wsum, src->cf.wsum, src->wsum_capacity * sizeof(double2));
    if (e == hipSuccess)
      e = copy(st->cf.groups, src->cf.groups, src->groups.size() * sizeof(GroupInfo));
    GroupInfo* g = st->cf.groups;
    float2* t = st->cf.taps;
    double2* w = st->cf.wsum;
    st->cf = src->cf;
    st->cf.groups = g;
    st->cf.taps = t;
    st->cf.wsum = w;
    st->groups = src->groups;
  }
  if (e == hipSuccess) e = hipStreamSynchronize(st->stream);
  if (e != hipSuccess) {
    hpgx_destroy(st);
    return hip_fail(src, e, "clone");
  }
  if ((rc = ensure_call_buffers(st))) {
    hpgx_destroy(st);
    return fail(src, rc, "clone: call buffers");
  }
  *dst = st;
  return HPGX_SUCCESS;
}

void hpgx_destroy(hpgx_state_t st) {
  if (!st) return;
  (void)hipSetDevice(st->device);
  free_state(st);
  delete st;
}

size_t hpgx_error_message(hpgx_state_t st, char* buf, size_t len) {
  std::string msg;
  if (st) {
    msg = st->err;
  } else {
    std::lock_guard<std::mutex> lk(g_create_mtx);
    msg = g_create_error;
  }
  if (buf && len > 0) {
    const size_t k = std::min(len - 1, msg.size());
    std::memcpy(buf, msg.data(), k);
    buf[k] = '\0';
  }
  return msg.size();
}

unsigned hpgx_max_added_tasks(hpgx_state_t st) { return st ? st->max_added_tasks : 0; }
size_t hpgx_max_visibility_batch_size(hpgx_state_t st) { return st ? st->max_batch : 0; }
void hpgx_grid_size(hpgx_state_t st, unsigned g[4]) {
  if (st) std::memcpy(g, st->gsize, sizeof(st->gsize));
}
void hpgx_grid_scale(hpgx_state_t st, double s[2]) {
  if (st) {
    s[0] = st->gscale[0];
    s[1] = st->gscale[1];
  }
}
unsigned hpgx_num_polarizations(hpgx_state_t st) { return st ? st->npol : 0; }
int hpgx_grid_mode(hpgx_state_t st) { return st ? st->mode : -1; }

size_t hpgx_convolution_function_region_size(
  hpgx_state_t st, unsigned oversampling, unsigned num_groups,
  const unsigned* extents) {
  if (!extents || num_groups == 0 || oversampling == 0) return st ? st->taps_capacity : 0;
  return cf_taps_for(oversampling, num_groups, extents);
}

int hpgx_allocate_convolution_function_region(
  hpgx_state_t st, unsigned oversampling, unsigned num_groups,
  const unsigned* extents) {
  if (!st) return HPGX_ERR_OTHER;
  if (int rc = set_device(st)) return rc;
  HPGX_HIP(st, hipStreamSynchronize(st->stream));
  if (num_groups == 0 || !extents || oversampling == 0) {
    dev_free(st->cf.taps);
    dev_free(st->cf.wsum);
    st->taps_capacity = st->wsum_capacity = 0;
    st->cf.ngroups = 0;
    st->groups.clear();
    return HPGX_SUCCESS;
  }
  size_t ws = 0;
  const size_t taps = cf_taps_for(oversampling, num_groups, extents, &ws);
  // force reallocation to the requested size (src/hpg_runtime.hpp:871-880)
  if (st->cf.ngroups == 0) {
    dev_free(st->cf.taps);
    dev_free(st->cf.wsum);
    st->taps_capacity = st->wsum_capacity = 0;
  }
  return ensure_cf_capacity(st, taps, ws);
}

int hpgx_set_convolution_function(
  hpgx_state_t st, unsigned oversampling, unsigned num_groups,
  const unsigned* extents, const void* const* values, int memory) {
  if (!st) return HPGX_ERR_OTHER;
  if (oversampling == 0 || oversampling > 253)
    return fail(st, HPGX_ERR_OTHER, "CF oversampling must be in [1, 253]");
  if (num_groups > unsigned(hpgx::max_cf_groups))
    return fail(st, HPGX_ERR_OTHER, "too many CF groups");
  // CFSupportExceedsGrid check (src/hpg_runtime.hpp:892-902)
  for (unsigned g = 0; g < num_groups; ++g) {
    const unsigned* e = extents + 4 * g;
    if (size_t(e[0]) > size_t(st->gsize[0]) * oversampling
        || size_t(e[1]) > size_t(st->gsize[1]) * oversampling)
      return fail(st, HPGX_ERR_CF_SUPPORT_EXCEEDS_GRID, "CF support size exceeds grid size");
    if (e[0] < 2 * hpgx::cf_padding * oversampling || e[1] < 2 * hpgx::cf_padding * oversampling)
      return fail(st, HPGX_ERR_OTHER, "CF extents smaller than the padding");
    if (e[3] > 65536)
      return fail(st, HPGX_ERR_OTHER, "at most 65536 CF cubes per group");
  }
  if (int rc = set_device(st)) return rc;
  // the previous CF may still be in use by queued kernels
  HPGX_HIP(st, hipStreamSynchronize(st->stream));

  const int os = int(oversampling);
  const int F = (os + 1) / 2;
  const int nf = 2 * F + 1;
  std::vector<GroupInfo> groups(num_groups);
  long long toff = 0, woff = 0;
  int max_sx = 1, max_sy = 1, max_ncube = 1;
  size_t max_stage = 0;
  for (unsigned g = 0; g < num_groups; ++g) {
    const unsigned* e = extents + 4 * g;
    GroupInfo& gi = groups[g];
    radii(e, oversampling, gi.rx, gi.ry);
    gi.sx = 2 * gi.rx + 1;
    gi.sy = 2 * gi.ry + 1;
    gi.nmueller = int(e[2]);
    gi.ncube = int(e[3]);
    gi.tap_offset = toff;
    gi.wsum_offset = woff;
    const long long slices = (long long)gi.ncube * gi.nmueller * nf * nf;
    toff += slices * gi.sx * gi.sy;
    woff += slices;
    max_sx = std::max(max_sx, gi.sx);
    max_sy = std::max(max_sy, gi.sy);
    max_ncube = std::max(max_ncube, gi.ncube);
    max_stage = std::max(max_stage, size_t(e[0]) * e[1] * e[2] * e[3]);
  }
  if (int rc = ensure_cf_capacity(st, size_t(toff), size_t(woff))) return rc;
  if (memory != HPGX_DEVICE && max_stage > st->staging_capacity) {
    dev_free(st->staging);
    st->staging_capacity = 0;
    HPGX_HIP(st, dev_alloc(st->staging, max_stage * sizeof(float2)));
    st->staging_capacity = max_stage;
  }
  for (unsigned g = 0; g < num_groups; ++g) {
    const unsigned* e = extents + 4 * g;
    const size_t count = size_t(e[0]) * e[1] * e[2] * e[3];
    const float2* src = static_cast<const float2*>(values[g]);
    if (memory != HPGX_DEVICE) {
      HPGX_HIP(st, hipMemcpyAsync(st->staging, values[g], count * sizeof(float2),
                                  hipMemcpyHostToDevice, st->stream));
      src = st->staging;
    }
    HPGX_HIP(st, hpgx::launch_cf_relayout(st->stream, src, e, groups[g], os, F, nf, st->cf.taps));
    HPGX_HIP(st, hpgx::launch_cf_wsum(st->stream, groups[g], nf, st->cf.taps, st->cf.wsum));
    if (memory != HPGX_DEVICE) HPGX_HIP(st, hipStreamSynchronize(st->stream));
  }
  if (num_groups > 0)
    HPGX_HIP(st, hipMemcpyAsync(st->cf.groups, groups.data(), num_groups * sizeof(GroupInfo),
                                hipMemcpyHostToDevice, st->stream));
  HPGX_HIP(st, hipStreamSynchronize(st->stream));
  st->groups = groups;
  st->cf.os = os;
  st->cf.F = F;
  st->cf.nf = nf;
  st->cf.ngroups = int(num_groups);
  st->cf.max_sx = max_sx;
  st->cf.max_sy = max_sy;
  st->cf.max_ncube = max_ncube;
  return ensure_call_buffers(st);
}

int hpgx_set_model(
  hpgx_state_t st, const unsigned model_size[4], const void* values, int layout,
  int memory) {
  if (!st) return HPGX_ERR_OTHER;
  if (std::memcmp(model_size, st->gsize, sizeof(st->gsize)) != 0) {
    char buf[256];
    std::snprintf(buf, sizeof(buf),
                  "model grid size [%u,%u,%u,%u] is different from visibility grid size [%u,%u,%u,%u]",
                  model_size[0], model_size[1], model_size[2], model_size[3],
                  st->gsize[0], st->gsize[1], st->gsize[2], st->gsize[3]);
    return fail(st, HPGX_ERR_INVALID_MODEL_GRID_SIZE, buf);
  }
  if (int rc = set_device(st)) return rc;
  HPGX_HIP(st, hipStreamSynchronize(st->stream));
  return upload_model(st, values, layout, memory);
}

int hpgx_reset_model(hpgx_state_t st) {
  if (!st) return HPGX_ERR_OTHER;
  if (int rc = set_device(st)) return rc;
  HPGX_HIP(st, hipStreamSynchronize(st->stream));
  dev_free(st->model);
  return HPGX_SUCCESS;
}

int hpgx_has_model(hpgx_state_t st) { return st && st->model ? 1 : 0; }

int hpgx_grid_visibilities(
  hpgx_state_t st, const void* vis, size_t n, unsigned npol,
  int update_grid_weights, int do_degrid, int return_visibilities, int do_grid,
  void* vis_out, int memory) {
  if (!st) return HPGX_ERR_OTHER;
  // validation order of runtime::GridderState::grid_visibilities
  // (src/hpg_runtime.hpp:1815-1827)
  if (n > st->max_batch)
    return fail(st, HPGX_ERR_EXCESSIVE_NUMBER_VISIBILITIES,
                "Number of visibilities exceeds maximum batch size");
  if (npol != st->npol)
    return fail(st, HPGX_ERR_INVALID_NUMBER_POLARIZATIONS,
                "Number of visibility polarizations does not match Mueller matrix");
  if (!do_grid && update_grid_weights)
    return fail(st, HPGX_ERR_UPDATE_WEIGHTS_WITHOUT_GRIDDING,
                "Grid weights cannot be updated without doing gridding");
  if (n == 0) return HPGX_SUCCESS;
  if (!vis) return fail(st, HPGX_ERR_OTHER, "null visibilities");
  if (return_visibilities && !vis_out)
    return fail(st, HPGX_ERR_OTHER, "null visibilities_out");
  if (int rc = set_device(st)) return rc;
  if (int rc = flush_finishing(st)) return rc;
  if (st->cf.ngroups == 0) {
    // no CF: every record refers to a missing CF group and is skipped
    if (return_visibilities && vis_out != vis) {
      HPGX_HIP(st, hipMemcpyAsync(
                     vis_out, vis, n * visdata_size(npol),
                     memory == HPGX_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyHostToHost,
                     st->stream));
    }
    return HPGX_SUCCESS;
  }
  if (int rc = ensure_call_buffers(st)) return rc;
  // Device records are read in place unless degridding rewrites them.  The
  // binned degridder reads them in place too and writes the returned
  // records straight into visibilities_out (its binning pass copies each
  // record there whole, then the predicted or residual values replace the
  // record's values), or nowhere when they are not returned; other
  // degridders use the staged copy.
  const bool dev = memory == HPGX_DEVICE;
  const bool binned_degrid = do_degrid && st->model && use_halo(st);
  const void* records = nullptr;  // nullptr: the staged copy
  void* out_records = nullptr;
  if (dev && !do_degrid) {
    records = vis;
  } else if (dev && binned_degrid) {
    records = vis;
    if (return_visibilities) out_records = vis_out;
  } else if (!dev && st->tasks > 1) {
    // uploaded on the next task's copy stream: overlaps the kernels of the
    // calls before
    if (int rc = upload_task_records(st, vis, n)) return rc;
  } else {
    if (int rc = stage_records(st, vis, n, memory)) return rc;
  }
  if (int rc = run_grid(st, n, update_grid_weights, do_degrid, do_grid,
                        hpgx::GvisSource::Values, nullptr, records, out_records))
    return rc;
  const void* rec_base = records ? records : st->cb.vis;
  if (return_visibilities && !out_records && vis_out != rec_base) {
    HPGX_HIP(st, hipMemcpyAsync(
                   vis_out, rec_base, n * visdata_size(npol),
                   memory == HPGX_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost,
                   st->stream));
  }
  return mark_task_use(st);
}

int hpgx_grid_visibilities_noalloc(
  hpgx_state_t st, const void* vis, size_t n, unsigned npol, int update_grid_weights) {
  if (!st) return HPGX_ERR_OTHER;
  // validation of hpgx_grid_visibilities, in the same order
  if (n > st->max_batch)
    return fail(st, HPGX_ERR_EXCESSIVE_NUMBER_VISIBILITIES,
                "Number of visibilities exceeds maximum batch size");
  if (npol != st->npol)
    return fail(st, HPGX_ERR_INVALID_NUMBER_POLARIZATIONS,
                "Number of visibility polarizations does not match Mueller matrix");
  if (n == 0) return HPGX_SUCCESS;
  if (!vis) return fail(st, HPGX_ERR_OTHER, "null visibilities");
  if (int rc = set_device(st)) return rc;
  if (int rc = flush_finishing(st)) return rc;
  // no CF: every record refers to a missing CF group and is skipped
  if (st->cf.ngroups == 0) return HPGX_SUCCESS;
  void* base = nullptr;
  HPGX_HIP(st, noalloc_scratch(st->device, &base));
  CallBuffers c{};
  size_t pass_blocks = 0;
  if (int rc = noalloc_plan(st, base, n, c, pass_blocks)) return rc;
  return run_grid_with(st, c, pass_blocks, n, update_grid_weights != 0, false, true,
                       hpgx::GvisSource::Values, nullptr, vis, nullptr);
}

int hpgx_degrid_adjoint(
  hpgx_state_t st, const void* vis, const void* residuals, size_t n,
  unsigned npol, int memory) {
  if (!st) return HPGX_ERR_OTHER;
  if (n > st->max_batch)
    return fail(st, HPGX_ERR_EXCESSIVE_NUMBER_VISIBILITIES,
                "Number of visibilities exceeds maximum batch size");
  if (npol != st->npol)
    return fail(st, HPGX_ERR_INVALID_NUMBER_POLARIZATIONS,
                "Number of visibility polarizations does not match Mueller matrix");
  if (n == 0 || st->cf.ngroups == 0) return HPGX_SUCCESS;
  if (!vis || !residuals) return fail(st, HPGX_ERR_OTHER, "null argument");
  if (int rc = set_device(st)) return rc;
  if (int rc = flush_finishing(st)) return rc;
  if (int rc = ensure_call_buffers(st)) return rc;
  const bool in_place = memory == HPGX_DEVICE;
  if (!in_place)
    if (int rc = stage_records(st, vis, n, memory)) return rc;
  const double2* res = static_cast<const double2*>(residuals);
  double2* tmp = nullptr;
  if (memory != HPGX_DEVICE) {
    HPGX_HIP(st, dev_alloc(tmp, n * npol * sizeof(double2)));
    HPGX_HIP(st, hipMemcpyAsync(tmp, residuals, n * npol * sizeof(double2),
                                hipMemcpyHostToDevice, st->stream));
    res = tmp;
  }
  int rc = run_grid(st, n, false, false, true, hpgx::GvisSource::Adjoint, res,
                    in_place ? vis : nullptr);
  if (tmp) {
    (void)hipStreamSynchronize(st->stream);
    dev_free(tmp);
  }
  if (!rc) rc = mark_task_use(st);
  return rc;
}

int hpgx_degrid_forward(
  hpgx_state_t st, const void* vis, void* predictions, size_t n, unsigned npol,
  int memory) {
  if (!st) return HPGX_ERR_OTHER;
  if (n > st->max_batch)
    return fail(st, HPGX_ERR_EXCESSIVE_NUMBER_VISIBILITIES,
                "Number of visibilities exceeds maximum batch size");
  if (npol != st->npol)
    return fail(st, HPGX_ERR_INVALID_NUMBER_POLARIZATIONS,
                "Number of visibility polarizations does not match Mueller matrix");
  if (n == 0) return HPGX_SUCCESS;
  if (!vis || !predictions) return fail(st, HPGX_ERR_OTHER, "null argument");
  if (int rc = set_device(st)) return rc;
  if (int rc = flush_finishing(st)) return rc;
  const size_t bytes = n * npol * sizeof(double2);
  double2* out = static_cast<double2*>(predictions);
  double2* tmp = nullptr;
  if (memory != HPGX_DEVICE) {
    HPGX_HIP(st, dev_alloc(tmp, bytes));
    out = tmp;
  }
  // records outside the grid (or with no CF) predict zero
  HPGX_HIP(st, hipMemsetAsync(out, 0, bytes, st->stream));
  if (st->cf.ngroups > 0) {
    if (int rc = ensure_call_buffers(st)) return rc;
    // device records are read in place by the lane-per-record degridder
    const bool in_place = memory == HPGX_DEVICE
      && size_t(st->cf.max_sx) * size_t(st->cf.max_sy) <= hpgx::degrid_lane_max_taps();
    if (!in_place)
      if (int rc = stage_records(st, vis, n, memory)) return rc;
    if (st->model && use_halo(st)) {
      if (int rc = run_degrid_binned(st, n, false, in_place ? vis : nullptr, out)) return rc;
    } else {
      HPGX_HIP(st, hpgx::launch_degrid(st->stream, int(npol), n, st->geometry(), st->cf,
                                       st->mu, st->model, false, st->cb, out,
                                       in_place ? vis : nullptr));
    }
  }
  if (tmp) {
    HPGX_HIP(st, hipMemcpyAsync(predictions, tmp, bytes, hipMemcpyDeviceToHost, st->stream));
    HPGX_HIP(st, hipStreamSynchronize(st->stream));
    dev_free(tmp);
  }
  return mark_task_use(st);
}

int hpgx_fence(hpgx_state_t st) {
  if (!st) return HPGX_ERR_OTHER;
  if (int rc = set_device(st)) return rc;
  if (int rc = flush_finishing(st)) return rc;
  HPGX_HIP(st, hipStreamSynchronize(st->stream));
  for (hipStream_t cs : st->copy_streams)
    if (cs) HPGX_HIP(st, hipStreamSynchronize(cs));
  unsigned long long invalid = 0;
  if (st->cb.invalid_count) {
    HPGX_HIP(st, hipMemcpy(&invalid, st->cb.invalid_count, sizeof(invalid), hipMemcpyDeviceToHost));
    if (invalid) {
      HPGX_HIP(st, hipMemset(st->cb.invalid_count, 0, sizeof(invalid)));
      return fail(st, HPGX_ERR_OUT_OF_BOUNDS_CF_INDEX,
                  std::to_string(invalid)
                    + " visibilities skipped: cf_index or grid_cube out of bounds");
    }
  }
  return HPGX_SUCCESS;
}

int hpgx_normalize_by_weights(hpgx_state_t st, double wfactor) {
  if (!st) return HPGX_ERR_OTHER;
  if (int rc = set_device(st)) return rc;
  // deferred: it joins the pass before a following FFT
  if (st->pend_norm || st->pend_fft)
    if (int rc = flush_finishing(st)) return rc;
  st->pend_norm = true;
  st->pend_wfactor = wfactor;
  return HPGX_SUCCESS;
}

int hpgx_apply_grid_fft(hpgx_state_t st, double norm, int sign, int in_place) {
  if (!st) return HPGX_ERR_OTHER;
  if (int rc = set_device(st)) return rc;
  if (in_place) {
    // deferred (its plan made now, so that a plan failure is reported here)
    if (st->pend_fft)
      if (int rc = flush_finishing(st)) return rc;
    hipfftHandle h;
    if (int rc = get_plan(st, st->mode == HPGX_GRID_FP64, h)) return rc;
    st->pend_fft = true;
    st->pend_fft_norm = norm;
    st->pend_fft_sign = sign;
    return HPGX_SUCCESS;
  }
  if (int rc = flush_finishing(st)) return rc;
  void* out = nullptr;
  if (int rc = fft_planes(st, st->grid, st->mode == HPGX_GRID_FP64, norm, sign,
                          in_place != 0, &out))
    return rc;
  st->grid = out;
  return HPGX_SUCCESS;
}

int hpgx_apply_model_fft(hpgx_state_t st, double norm, int sign, int in_place) {
  if (!st) return HPGX_ERR_OTHER;
  if (!st->model) return HPGX_SUCCESS;
  if (int rc = set_device(st)) return rc;
  void* out = nullptr;
  if (int rc = fft_planes(st, st->model, true, norm, sign, in_place != 0, &out)) return rc;
  st->model = static_cast<double2*>(out);
  return HPGX_SUCCESS;
}

int hpgx_shift_grid(hpgx_state_t st, int direction) {
  if (!st) return HPGX_ERR_OTHER;
  if (int rc = set_device(st)) return rc;
  // directly after a deferred FFT of an even grid the shift is the FFT
  // input's modulation (both directions shift by half the grid)
  if (st->pend_fft && !st->pend_shift && st->gsize[0] % 2 == 0 && st->gsize[1] % 2 == 0) {
    st->pend_shift = true;
    return HPGX_SUCCESS;
  }
  if (int rc = flush_finishing(st)) return rc;
  if (int rc = ensure_scratch(st)) return rc;
  HPGX_HIP(st, hpgx::launch_shift(
                 st->stream, st->mode == HPGX_GRID_FP64 ? 0 : 1, st->grid, st->scratch,
                 direction, st->geometry()));
  return HPGX_SUCCESS;
}

int hpgx_shift_model(hpgx_state_t st, int direction) {
  if (!st) return HPGX_ERR_OTHER;
  if (!st->model) return HPGX_SUCCESS;
  if (int rc = set_device(st)) return rc;
  const Geometry geo = st->geometry();
  void* scratch = nullptr;
  const bool even = geo.nx % 2 == 0 && geo.ny % 2 == 0;
  if (!even) HPGX_HIP(st, dev_alloc(scratch, st->grid_elems() * sizeof(double2)));
  hipError_t e = hpgx::launch_shift(st->stream, 0, st->model, scratch, direction, geo);
  if (scratch) {
    (void)hipStreamSynchronize(st->stream);
    dev_free(scratch);
  }
  if (e != hipSuccess) return hip_fail(st, e, "shift_model");
  return HPGX_SUCCESS;
}

int hpgx_reset_grid(hpgx_state_t st) {
  if (!st) return HPGX_ERR_OTHER;
  if (int rc = set_device(st)) return rc;
  // the deferred finishing chain of a grid that is cleared has no effect
  st->pend_norm = st->pend_fft = st->pend_shift = false;
  HPGX_HIP(st, hipStreamSynchronize(st->stream));
  if (int rc = zero_grid(st)) return rc;
  HPGX_HIP(st, hipStreamSynchronize(st->stream));
  return HPGX_SUCCESS;
}

int hpgx_grid_values(hpgx_state_t st, void* dst, int layout) {
  if (!st || !dst) return HPGX_ERR_OTHER;
  if (int rc = set_device(st)) return rc;
  if (int rc = flush_finishing(st)) return rc;
  HPGX_HIP(st, hipStreamSynchronize(st->stream));
  return copy_grid_to_host(st, st->grid, st->mode != HPGX_GRID_FP64, dst, layout);
}

int hpgx_grid_weights(hpgx_state_t st, double* dst, int layout) {
  if (!st || !dst) return HPGX_ERR_OTHER;
  if (int rc = set_device(st)) return rc;
  HPGX_HIP(st, hipStreamSynchronize(st->stream));
  const size_t nr = st->gsize[2], nc = st->gsize[3];
  std::vector<double> h(nr * nc);
  HPGX_HIP(st, hipMemcpy(h.data(), st->weights, h.size() * sizeof(double), hipMemcpyDeviceToHost));
  if (layout == HPGX_LAYOUT_LEFT) {
    std::memcpy(dst, h.data(), h.size() * sizeof(double));
  } else {
    for (size_t c = 0; c < nc; ++c)
      for (size_t r = 0; r < nr; ++r) dst[r * nc + c] = h[c * nr + r];
  }
  return HPGX_SUCCESS;
}

int hpgx_model_values(hpgx_state_t st, void* dst, int layout) {
  if (!st || !dst) return HPGX_ERR_OTHER;
  if (int rc = set_device(st)) return rc;
  HPGX_HIP(st, hipStreamSynchronize(st->stream));
  if (!st->model) {
    // an unset model reads as zero (UnallocatedModelValueArray,
    // src/hpg_runtime.hpp:1230-1240)
    std::memset(dst, 0, st->grid_elems() * sizeof(double2));
    return HPGX_SUCCESS;
  }
  return copy_grid_to_host(st, st->model, false, dst, layout);
}

int hpgx_grid_window(
  hpgx_state_t st, unsigned x0, unsigned y0, unsigned nx, unsigned ny,
  unsigned mrow, unsigned cube, void* dst) {
  if (!st || !dst) return HPGX_ERR_OTHER;
  if (size_t(x0) + nx > st->gsize[0] || size_t(y0) + ny > st->gsize[1]
      || mrow >= st->gsize[2] || cube >= st->gsize[3])
    return fail(st, HPGX_ERR_OTHER, "grid window out of bounds");
  if (int rc = set_device(st)) return rc;
  if (int rc = flush_finishing(st)) return rc;
  HPGX_HIP(st, hipStreamSynchronize(st->stream));
  const size_t plane = size_t(cube) * st->gsize[2] + mrow;
  const size_t ps = size_t(st->gsize[0]) * st->gsize[1];
  std::vector<float2> row32;
  for (unsigned y = 0; y < ny; ++y) {
    const size_t off = plane * ps + size_t(y0 + y) * st->gsize[0] + x0;
    double2* d = static_cast<double2*>(dst) + size_t(y) * nx;
    if (st->mode == HPGX_GRID_FP64) {
      HPGX_HIP(st, hipMemcpy(d, static_cast<double2*>(st->grid) + off,
                             nx * sizeof(double2), hipMemcpyDeviceToHost));
    } else {
      row32.resize(nx);
      HPGX_HIP(st, hipMemcpy(row32.data(), static_cast<float2*>(st->grid) + off,
                             nx * sizeof(float2), hipMemcpyDeviceToHost));
      for (unsigned x = 0; x < nx; ++x)
        d[x] = make_double2(double(row32[x].x), double(row32[x].y));
    }
  }
  return HPGX_SUCCESS;
}

int hpgx_grid_device_ptr(hpgx_state_t st, void** ptr, size_t* count) {
  if (!st) return HPGX_ERR_OTHER;
  if (int rc = set_device(st)) return rc;
  if (int rc = flush_finishing(st)) return rc;
  if (ptr) *ptr = st->grid;
  if (count) *count = st->grid_elems();
  return HPGX_SUCCESS;
}

int hpgx_event_record(hpgx_state_t st, int slot) {
  if (!st || slot < 0 || slot >= num_event_slots) return HPGX_ERR_OTHER;
  if (int rc = set_device(st)) return rc;
  if (int rc = flush_finishing(st)) return rc;
  if (!st->events[slot]) HPGX_HIP(st, hipEventCreate(&st->events[slot]));
  HPGX_HIP(st, hipEventRecord(st->events[slot], st->stream));
  return HPGX_SUCCESS;
}

int hpgx_event_elapsed(hpgx_state_t st, int slot0, int slot1, float* ms) {
  if (!st || !ms || slot0 < 0 || slot1 < 0 || slot0 >= num_event_slots
      || slot1 >= num_event_slots || !st->events[slot0] || !st->events[slot1])
    return HPGX_ERR_OTHER;
  HPGX_HIP(st, hipEventSynchronize(st->events[slot1]));
  HPGX_HIP(st, hipEventElapsedTime(ms, st->events[slot0], st->events[slot1]));
  return HPGX_SUCCESS;
}

}  // extern "C"
