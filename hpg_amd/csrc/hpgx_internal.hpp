// Internal declarations shared by the C ABI (hpgx_api.cpp) and the HIP
// kernels (hpgx_kernels.hip).  Not installed; the public interface is
// include/hpgx.h.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace hpgx {

// CFArray::padding (src/hpg.hpp:655)
constexpr int cf_padding = 2;
// maximum number of CF groups (HPG_MAX_NUM_CF_GROUPS, src/CMakeLists.txt:37)
constexpr int max_cf_groups = 1000;
// grid tile edge (pixels) owned by one workgroup of the gridding kernel
constexpr int tile = 32;
// waves per gridding workgroup; each wave accumulates into its own tile copy
constexpr int tile_waves = 4;

// Per-group geometry of the device CF.  The device layout of group g is
// "fine-offset sliced": for every (cube, mueller, fx, fy) -- fx, fy in
// [-F, F] the visibility's fine offset on each axis, F = (os + 1) / 2 -- the
// sx x sy taps of the support are contiguous, x fastest:
//   taps[tap_offset + (((cube * nmueller + m) * nf + fx + F) * nf + fy + F)
//        * sx * sy + Y * sx + X]
//   = CFArray(ix, iy, m, cube, g), ix = (X + padding) * os + fx, iy likewise,
// which is the element the reference reaches through its 6-D layout
// (src/hpg_impl.hpp:221-281, src/hpg_core.hpp:422-446: major * os + minor).
// wsum[wsum_offset + ((cube * nmueller + m) * nf + fx + F) * nf + fy + F]
// is the FP64 sum of that slice's taps, X outer and Y inner.
struct GroupInfo {
  long long tap_offset;
  long long wsum_offset;
  int sx, sy;          // support 2r + 1 per axis
  int rx, ry;          // radii (src/hpg.hpp:687-694)
  int nmueller, ncube;
};

struct Mueller {
  int M[4][4];   // mueller_indexes, padded with -1
  int Mc[4][4];  // conjugate_mueller_indexes, padded with -1
};

// grid / CF geometry passed by value to kernels
struct Geometry {
  int nx, ny, nr, nc;  // grid_size: x, y, mrow, cube
  double scale[2];     // grid_scale
  int os;              // CF oversampling
  int F;               // max |fine offset| = (os + 1) / 2
  int nf;              // 2F + 1
  int ngroups;
  int ntx, nty;        // tiles per plane along x and y
};

// flags of a bin entry
constexpr unsigned short flag_pos_w = 1;
constexpr unsigned short flag_gradient = 4;

enum class GvisSource { Values, Adjoint };

// Device buffers for one gridding call (owned by the state).
struct CallBuffers {
  void* vis;            // raw VisData<N> records
  float2* gvis;         // per visibility x N
  unsigned* hist;       // per (plane, tile) key, nkeys + 1
  unsigned* offsets;    // exclusive scan of hist, nkeys + 1
  unsigned* cursor;     // scatter positions, nkeys
  size_t entries_capacity;  // entries of stream_entries
  void* scan_temp;
  size_t scan_temp_bytes;
  void* stream_entries;  // compact binned entries (Entry<N>), sorted by key
  unsigned* entry_vis;   // record index of entries with a phase gradient
  unsigned* off_table;   // block-local bins: per bin block, nkeys + 1 offsets
  size_t off_table_capacity;
  double2* halo;         // haloed tiles: per key, the halo's FP64 values
  unsigned* halo_set;    // per key: 1 if its halo was written by this pass
  size_t halo_capacity;  // elements of halo
  size_t halo_set_capacity;
  unsigned long long* invalid_count;
  double2* degrid_acc;   // binned degridding: per record x polarization sums
  size_t degrid_acc_capacity;
  void* tally_entries;   // launch_tally: entries in record order (Entry4<N>)
  void* tally_keys;      // launch_tally: (key, place within the key) per slot
  size_t tally_capacity; // slots of both
  unsigned* degrid_order; // launch_degrid (wave form): keys, sorted keys, record
                          // indexes, sorted indexes, order_capacity each
  size_t order_capacity;
  void* order_temp;       // the sort's temporary storage
  size_t order_temp_bytes;
};

struct DeviceCF {
  GroupInfo* groups;  // device array
  float2* taps;
  double2* wsum;
  int os, F, nf, ngroups;
  int max_sx, max_sy;
  int max_ncube;  // CF cubes of the group with the most
};

// ---- launchers (hpgx_kernels.hip) -------------------------------------

hipError_t launch_cf_relayout(
  hipStream_t s, const float2* staging, const unsigned ext[4],
  const GroupInfo& gi, int os, int F, int nf, float2* taps);
hipError_t launch_cf_wsum(
  hipStream_t s, const GroupInfo& gi, int nf, const float2* taps,
  double2* wsum);

size_t tiles_per_footprint(int max_sx, int max_sy);

// exclusive scan of the global bin counts (hist -> offsets, cursor)
hipError_t launch_scan(hipStream_t s, size_t nkeys, CallBuffers& cb);
size_t scan_temp_bytes(size_t nkeys);
// global bins; halo_entries: one Entry4 per record and plane in the bin of
// the tile holding the support corner (for launch_tile_halo with
// global_bins; with quad_keys the bin of the corner's quadrant, key 4 x tile
// key + quadrant, for launch_tile_sub with global_bins), else one Entry per
// touched tile
hipError_t launch_count(
  hipStream_t s, int npol, size_t n, const Geometry& geo, const DeviceCF& cf,
  const Mueller& mu, bool update_weights, bool do_grid, CallBuffers& cb,
  double* weights, bool halo_entries = false, bool quad_keys = false);
size_t entry_bytes(int npol);
hipError_t launch_emit(
  hipStream_t s, int npol, size_t n, const Geometry& geo, const DeviceCF& cf,
  const Mueller& mu, bool gvis_ready, GvisSource src, const double2* residuals,
  CallBuffers& cb, bool halo_entries = false, bool quad_keys = false);
// global quadrant bins (one polarization, haloed entries) in two passes:
// launch_tally (counts in cb.hist, entries and their (key, place) in
// record order), launch_scan, then launch_place (entries into their keys'
// ranges); n x nr slots of cb.tally_entries / cb.tally_keys
hipError_t launch_tally(
  hipStream_t s, int npol, size_t n, const Geometry& geo, const DeviceCF& cf,
  const Mueller& mu, bool update_weights, bool gvis_ready, GvisSource src,
  const double2* residuals, CallBuffers& cb, double* weights);
hipError_t launch_place(hipStream_t s, int npol, size_t n, const Geometry& geo,
                        CallBuffers& cb);
hipError_t launch_tile_stream(
  hipStream_t s, int npol, int mode, size_t nkeys, const Geometry& geo,
  const DeviceCF& cf, const Mueller& mu, const CallBuffers& cb, void* grid);
// block-local binning (small key counts)
int bin_blocks_for(size_t n);
size_t bin_block_records();
int bin_keys_limit();
int bin_quad_keys_limit();  // with quad_keys: 4 x (plane, tile) keys
int tile_bins_limit();
hipError_t launch_bin(
  hipStream_t s, int npol, size_t n, const Geometry& geo, const DeviceCF& cf,
  const Mueller& mu, bool update_weights, bool gvis_ready, GvisSource src,
  const double2* residuals, CallBuffers& cb, size_t block_cap, double* weights,
  bool halo_entries, bool degrid_pass = false, void* copy_records = nullptr,
  bool quad_keys = false);
// haloed tiles (one entry per record and plane; supports below tile)
size_t halo_entry_bytes(int npol);
size_t halo_pixels_per_key(int hx, int hy);
int halo_tile_limit();
size_t halo_lds_bytes(int hx, int hy, int nbins);
// the quadrant kernel (k_quad): its LDS, bin blocks per pass, and whether it
// takes halo widths hx, hy (its regions fit LDS, the tile + halo its flush)
size_t quad_lds_bytes(int hx, int hy, int npol);
int quad_bins_limit();
bool quad_supported(int hx, int hy, int npol);
// acc (FP32 grids): 0 FP64 tile sums, 1 FP32 two-sum via returning atomics,
// 2 FP32 two-sum of (hi, lo) pairs (compare-and-swap in k_tile4, plain
// read-modify-write in k_quad).  quad: k_quad (acc 0 or 2) instead of k_tile4
hipError_t launch_tile_halo(
  hipStream_t s, int npol, int mode, int acc, size_t n, size_t nkeys, const Geometry& geo,
  const DeviceCF& cf, const Mueller& mu, const CallBuffers& cb, size_t block_cap,
  int hx, int hy, void* grid, bool global_bins = false, bool quad = false);
// the quadrant-key kernel (k_sub): FP32-compensated grids, supports of at
// most 8 x 8 (hx, hy <= 7), bins of launch_bin with quad_keys; ntile_keys
// (plane, tile) keys, whose offset-table rows hold 4 ntile_keys + 1 values;
// CF taps of at most 2^32 - 1 bytes
size_t sub_lds_bytes(int hx, int hy);
int sub_bins_limit();
bool sub_supported(int hx, int hy);
// uni: every CF group has the same support; global_bins (one polarization):
// the entries of launch_emit with quad_keys, cb.offsets their key offsets
hipError_t launch_tile_sub(
  hipStream_t s, int npol, size_t n, size_t ntile_keys, const Geometry& geo, const DeviceCF& cf,
  size_t taps_bytes, bool uni, const Mueller& mu, const CallBuffers& cb, size_t block_cap,
  int hx, int hy, void* grid, bool global_bins = false);
hipError_t launch_tile_blocal(
  hipStream_t s, int npol, int mode, size_t n, size_t nkeys, const Geometry& geo,
  const DeviceCF& cf, const Mueller& mu, const CallBuffers& cb, size_t block_cap,
  void* grid);
// binned degridding from the model's LDS tiles (entries of launch_bin with
// degrid_pass).  one_plane (every record has exactly one entry): the
// predictions are finished in the same kernel -- FP64 values to pred64, or
// values written into out_recs (residuals, and gvis in cb, when do_grid).
// Otherwise per-record sums are added into acc (zeroed beforehand) for
// launch_degrid_finish.
hipError_t launch_degrid_tiles(
  hipStream_t s, int npol, size_t n, size_t nkeys, const Geometry& geo,
  const DeviceCF& cf, const Mueller& mu, const CallBuffers& cb, size_t block_cap,
  int hx, int hy, const double2* model, double2* acc, bool one_plane, bool do_grid,
  void* out_recs, double2* pred64);
// predictions from those sums for the records at `records`: FP64 values
// to pred64, or values written into out_records when not null (residuals,
// and gvis in cb, when do_grid)
hipError_t launch_degrid_finish(
  hipStream_t s, int npol, size_t n, const Geometry& geo, const DeviceCF& cf,
  const Mueller& mu, const double2* acc, bool do_grid, const void* records,
  void* out_records, CallBuffers& cb, double2* pred64);
// degridding; with pred64 the FP64 predictions are written there and the
// records are only read, from `records` when given (else cb.vis; `records`
// requires supports of at most degrid_lane_max_taps() taps)
size_t degrid_lane_max_taps();
// the wave-per-record degridder runs (supports above degrid_lane_max_taps()
// taps) and can take its records in CF-slice order: with cb.degrid_order
// for at least n records it sorts them by (CF group, cube, fine offset)
// first, so that records reading the same CF slices run together
bool degrid_sorts(const DeviceCF& cf);
size_t degrid_order_temp_bytes(size_t n);
hipError_t launch_degrid(
  hipStream_t s, int npol, size_t n, const Geometry& geo, const DeviceCF& cf,
  const Mueller& mu, const double2* model, bool do_grid, CallBuffers& cb,
  double2* pred64 = nullptr, const void* records = nullptr);

hipError_t launch_normalize(
  hipStream_t s, int mode, void* grid, const double* weights, double wfactor,
  const Geometry& geo);
hipError_t launch_scale(
  hipStream_t s, int mode, void* grid, double inv_norm, size_t count);
// the deferred finishing chain's pass before its FFT: divided by wfactor *
// the plane's weight (weights not null), times scale, and with modulate
// multiplied by (-1)^(x + y); for grids with finish_pre_fits (rows and
// planes within a launch's y and z extents)
bool finish_pre_fits(const Geometry& geo);
hipError_t launch_finish_pre(
  hipStream_t s, int mode, void* grid, const double* weights, double wfactor, double scale,
  bool modulate, const Geometry& geo);
hipError_t launch_shift(
  hipStream_t s, int mode, void* grid, void* scratch, int direction,
  const Geometry& geo);
hipError_t launch_widen(hipStream_t s, const float2* src, double2* dst, size_t count);
hipError_t launch_narrow(hipStream_t s, const double2* src, float2* dst, size_t count);

}  // namespace hpgx
