// HIP kernels of the MI355X gridding device (gfx950, wave64).
//
// Build with -ffp-contract=off: every FP32 product and FP64 sum below is
// rounded where the reference's source rounds it (no fused multiply-add), so
// that the FP64-grid mode reproduces the reference's per-tap values bit for
// bit and differs from it only by the order of FP64 accumulation.
//
// Gridding (VisibilityGridder::kernel, src/hpg_core.hpp:1082-1253) is
// re-organized for the device in two stages: binning, then one workgroup per
// (grid plane, 32 x 32 tile) key accumulating the tile in LDS and adding it
// into the grid with one coalesced read-modify-write, so that no global
// atomic touches the grid.
//   k_bin + k_quad / k_tile4 + k_halo   records binned by the tile holding
//        their support corner (block-local bins, no global atomics); the
//        tile kernels accumulate the tile plus a halo of (support - 1)
//        pixels, and k_halo folds each tile's halo into its neighbours.
//        Supports narrower than a tile on grids of up to 8192 keys.
//   k_bin<N,false> + k_tile3            one entry per touched tile, clipped
//        footprints (supports reaching the tile edge, small grids).
//   k_count + scan + k_emit + k_tile2   the same entries binned through
//        global counters (grids with more keys than one workgroup's LDS
//        histogram holds).
// Degridding (degrid_vis, src/hpg_core.hpp:579-682): k_degrid4 reads the
// model's tile and halo from LDS for binned records; k_degrid_lane (one lane
// per record) and k_degrid (one wave per record, cross-lane reduction)
// gather the model directly.

#include "hpgx_internal.hpp"

#include <hipcub/hipcub.hpp>

#include <type_traits>

namespace hpgx {
namespace {

constexpr double c_light = 299792458.0;  // src/hpg_core.hpp:527
constexpr int wave = 64;
constexpr int block = 256;

// hpg::VisData<N> as seen by the device (src/hpg.hpp:236-311)
template <int N>
struct VisRec {
  float2 vis[N];
  float w[N];
  double freq;
  double phase;
  double uvw[3];
  unsigned grid_cube;
  unsigned cf_index[2];  // {cube, grp}
  double grad[2];
};
static_assert(sizeof(VisRec<1>) == 88, "VisData<1>");
static_assert(sizeof(VisRec<2>) == 96, "VisData<2>");
static_assert(sizeof(VisRec<3>) == 112, "VisData<3>");
static_assert(sizeof(VisRec<4>) == 120, "VisData<4>");

typedef int v2i_t __attribute__((ext_vector_type(2)));

__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// complex<float> * complex<double> -> complex<double> (Kokkos mixed product)
__device__ __forceinline__ double2 cmul(float2 a, double2 b) {
  return make_double2(
    double(a.x) * b.x - double(a.y) * b.y, double(a.x) * b.y + double(a.y) * b.x);
}
// Kokkos 4.x complex division: scale by the divisor's 1-norm
__device__ __forceinline__ double2 kdiv(double2 x, double2 y) {
  const double s = fabs(y.x) + fabs(y.y);
  if (s == 0.0) return make_double2(x.x / s, x.y / s);
  const double2 xs = make_double2(x.x / s, x.y / s);
  const double2 ycs = make_double2(y.x / s, -y.y / s);
  const double yabs = ycs.x * ycs.x + ycs.y * ycs.y;
  double2 r = cmul(xs, ycs);
  r.x /= yabs;
  r.y /= yabs;
  return r;
}
__device__ __forceinline__ double2 cphase(double ph) {
  double s, c;
  sincos(ph, &s, &c);
  return make_double2(c, s);
}
__device__ __forceinline__ float2 narrow(double2 a) {
  return make_float2(float(a.x), float(a.y));
}
// gvis = fl32(vis * phasor * weight) (src/hpg_core.hpp:1172-1182)
__device__ __forceinline__ float2 prephase(float2 v, double2 phasor, float w) {
  const double2 t = cmul(v, phasor);
  return make_float2(float(t.x * double(w)), float(t.y * double(w)));
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int m = wave / 2; m > 0; m >>= 1) v += __shfl_xor(v, m, wave);
  return v;
}

// q / d for 0 <= q < 2^32 / d, with m = magic(d) (exact in that range)
__device__ __forceinline__ unsigned magic(unsigned d) {
  return d == 1u ? 0u : 0xFFFFFFFFu / d + 1u;
}
__device__ __forceinline__ unsigned div_magic(unsigned q, unsigned d, unsigned m) {
  return d == 1u ? q : __umulhi(q, m);
}

__device__ __forceinline__ int uniform(int v) {
  return __builtin_amdgcn_readfirstlane(v);
}

// compute_vis_coord + Vis<N> ctor + all_within_grid
// (src/hpg_core.hpp:422-446, 493-550, 1070-1080)
struct Coord {
  int gx, gy;    // grid coordinate of the support corner
  int fx, fy;    // fine offsets
  bool inside;
};

__device__ __forceinline__ bool axis_coord(
  int g_size, int os, int radius, int support, double coord, double inv_lambda,
  double scale, int& corner, int& fine) {
  corner = 0;
  fine = 0;
  const double position = scale * coord * inv_lambda + g_size / 2.0;
  if (!(fabs(position) < 1.0e9)) return false;
  const double G = rint(position);
  fine = int(rint((G - position) * os));
  corner = int(G) - radius;
  return 0 <= corner && corner + support <= g_size;
}

template <int N>
__device__ __forceinline__ Coord vis_coord(
  const VisRec<N>& r, const Geometry& geo, const GroupInfo& g) {
  Coord c;
  const double inv_lambda = r.freq / c_light;
  const bool ix = axis_coord(
    geo.nx, geo.os, g.rx, g.sx, r.uvw[0], inv_lambda, geo.scale[0], c.gx, c.fx);
  const bool iy = axis_coord(
    geo.ny, geo.os, g.ry, g.sy, r.uvw[1], inv_lambda, geo.scale[1], c.gy, c.fy);
  c.inside = ix && iy;
  return c;
}

__device__ __forceinline__ int tiles_begin(int corner) { return corner / tile; }
__device__ __forceinline__ int tiles_end(int corner, int support) {
  return (corner + support - 1) / tile + 1;
}

__device__ __forceinline__ bool row_has_index(const int* row, int npol) {
  bool any = false;
  for (int v = 0; v < npol; ++v) any |= row[v] >= 0;
  return any;
}

// W_v = sum over grid rows of the degridding CF slice sums
// (degrid_vis weights, src/hpg_core.hpp:643-668): with the im factor applied
// to every tap, the sum is T[m] (w > 0) or conj(T[m]) (w <= 0).
template <int N>
__device__ __forceinline__ void degrid_norms(
  const Geometry& geo, const GroupInfo& g, const double2* wsum,
  const Mueller& mu, bool pos_w, unsigned cfcube, int fxi, int fyi,
  double2* W) {
  for (int v = 0; v < N; ++v) W[v] = make_double2(0.0, 0.0);
  for (int gpol = 0; gpol < geo.nr; ++gpol) {
    const int* row = pos_w ? mu.Mc[gpol] : mu.M[gpol];
    for (int v = 0; v < N; ++v) {
      const int m = row[v];
      if (m >= 0 && m < g.nmueller) {
        const double2 t = wsum[g.wsum_offset
                               + ((long long)(cfcube * g.nmueller + m) * geo.nf + fxi)
                                 * geo.nf + fyi];
        W[v].x += t.x;
        W[v].y += pos_w ? t.y : -t.y;
      }
    }
  }
}

// ---- CF ingestion ----------------------------------------------------

__global__ void k_cf_relayout(
  const float2* __restrict__ staging, unsigned ex, unsigned ey, GroupInfo gi,
  int os, int F, int nf, float2* __restrict__ taps, long long total) {
  const long long o = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (o >= total) return;
  long long q = o;
  const int X = int(q % gi.sx);
  q /= gi.sx;
  const int Y = int(q % gi.sy);
  q /= gi.sy;
  const int fyi = int(q % nf);
  q /= nf;
  const int fxi = int(q % nf);
  q /= nf;
  const int m = int(q % gi.nmueller);
  const int c = int(q / gi.nmueller);
  const long long ix = (long long)(X + cf_padding) * os + (fxi - F);
  const long long iy = (long long)(Y + cf_padding) * os + (fyi - F);
  float2 v = make_float2(0.0f, 0.0f);
  if (ix >= 0 && iy >= 0 && ix < ex && iy < ey)
    v = staging[((long long)(c * gi.nmueller + m) * ey + iy) * ex + ix];
  taps[gi.tap_offset + o] = v;
}

// the FP64 sum of each CF slice's taps (the weights' normalization), one
// wave per slice: lane l sums taps l, l + 64, ... (coalesced reads), then a
// fixed-order wave reduction.  FP64 sums of FP32 taps: exact while the
// taps' exponents span fewer than about 29 bits, else within a few FP64
// ulps of any other order
constexpr int wsum_block = 256;
__global__ __launch_bounds__(wsum_block) void k_cf_wsum(
  GroupInfo gi, int nf, const float2* __restrict__ taps,
  double2* __restrict__ wsum, long long nslices) {
  const long long s = blockIdx.x * (long long)(wsum_block / wave) + threadIdx.x / wave;
  const int lane = int(threadIdx.x % wave);
  if (s >= nslices) return;  // the whole wave
  const int ntap = gi.sx * gi.sy;
  const float2* t = taps + gi.tap_offset + s * ntap;
  double re = 0.0, im = 0.0;
  for (int k = lane; k < ntap; k += wave) {
    const float2 c = t[k];
    re += double(c.x);
    im += double(c.y);
  }
  re = wave_sum(re);
  im = wave_sum(im);
  if (lane == 0) wsum[gi.wsum_offset + s] = make_double2(re, im);
}

// ---- gridding ----------------------------------------------------------

// s + e = a + b exactly, s = fl(a + b) (Knuth's two-sum)
__device__ __forceinline__ void two_sum(float a, float b, float& s, float& e) {
  s = a + b;
  const float bb = s - a;
  e = (a - (s - bb)) + (b - bb);
}

// An FP32-compensated tile pixel component is a pair (hi, lo) of floats
// packed into 64 bits (hi in the low word), updated as one unit by
// compare-and-swap where several lanes may add to it.
__device__ __forceinline__ unsigned long long pack_pair(float hi, float lo) {
  return (unsigned long long)__float_as_uint(hi)
    | ((unsigned long long)__float_as_uint(lo) << 32);
}

// the pair (hi, lo) after adding h (+ the low-order part l, when HAS_L)
// to it: hi' = fl(hi + h), lo' = fl(lo + (err + l)) with err the exact
// rounding error of the first sum (two-sum)
template <bool HAS_L>
__device__ __forceinline__ unsigned long long pair_plus(
  unsigned long long pair, float h, float l) {
  const float hi = __uint_as_float(unsigned(pair));
  const float lo = __uint_as_float(unsigned(pair >> 32));
  float sum, err;
  two_sum(hi, h, sum, err);
  return pack_pair(sum, HAS_L ? lo + (err + l) : lo + err);
}

// compare-and-swap of the pair at a from `expect` to `desired`; on failure
// `expect` receives the value found
__device__ __forceinline__ bool pair_swap(
  unsigned long long* a, unsigned long long& expect, unsigned long long desired) {
  return __hip_atomic_compare_exchange_strong(
    a, &expect, desired, __ATOMIC_RELAXED, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// retries until the update of the pair at a lands; every retry follows a
// successful update of the same pair by another lane or wave
template <bool HAS_L>
__device__ __forceinline__ void pair_add(
  unsigned long long* a, unsigned long long expect, float h, float l) {
  while (!pair_swap(a, expect, pair_plus<HAS_L>(expect, h, l))) {
  }
}

// ---- gridding, binned-stream form ------------------------------------------
//
// k_count   one thread per visibility: skip rule, weights, bin counts
// k_emit    one thread per visibility: writes one compact Entry per touched
//           (plane, tile) bin -- support corner, fine offsets, CF group and
//           cube, pre-phased values -- so the tile kernel reads its bin as one
//           contiguous stream instead of gathering per-visibility records
// k_tile2   one workgroup per bin, one shared LDS tile: waves load 64 entries
//           at a time (one per lane), broadcast each to the whole wave with
//           readlane, and accumulate with LDS atomics (FP64 adds), or
//           (FP32-compensated grids) with compare-and-swap updates of
//           (hi, lo) pairs by two-sum

struct EntryHead {
  int gx, gy;
  unsigned fpk;  // fxi | fyi << 8 | flags << 16
  unsigned gpk;  // grp | cfcube << 16
};
template <int N>
struct Entry {
  EntryHead h;
  float2 g[N];
};

// entries of the haloed-tile pipeline (k_tile4): one per (record, plane),
// position relative to the tile that holds the support corner
struct Entry4Head {
  unsigned pos;  // lx | ly << 6 | fxi << 12 | fyi << 20 | flags << 28
  unsigned gpk;  // grp | cfcube << 16
};
template <int N>
struct Entry4 {
  Entry4Head h;
  float2 g[N];
};
constexpr unsigned e4_pos_w = 1u << 28;
constexpr unsigned e4_gradient = 1u << 29;
// k_emit<N, true> writes this layout as 2 + 2N 32-bit words
static_assert(sizeof(Entry4<1>) == 16 && sizeof(Entry4<4>) == 40, "Entry4 layout");

// bin key of tile key k for a support corner (gx, gy): k itself, or with
// quad_keys 4 k + the quadrant of the corner within its tile
__device__ __forceinline__ int bin_key(int k, int gx, int gy, int quad_keys) {
  if (!quad_keys) return k;
  const int qx = (gx % tile) >= tile / 2 ? 1 : 0;
  const int qy = (gy % tile) >= tile / 2 ? 1 : 0;
  return 4 * k + (qx | (qy << 1));
}

// HALO: one entry per plane, in the bin of the tile holding the support
// corner (the haloed-tile pipeline; with quad_keys in the bin of the
// corner's quadrant, bin_key); else one per touched tile
// Weights: each wave sums its records' weights per (grid cube, row), and
// for grids of at most count_planes_lds planes adds the sums into the
// workgroup's LDS, which adds into the grid weights once at its end (the
// workgroups stride over the records, so a few thousand global atomics per
// plane instead of one per wave).
constexpr int count_planes_lds = 64;
constexpr unsigned count_blocks_max = 2048;

template <int N, bool HALO>
__global__ __launch_bounds__(block) void k_count(
  const VisRec<N>* __restrict__ recs, size_t n, Geometry geo,
  const GroupInfo* __restrict__ groups, const double2* __restrict__ wsum,
  Mueller mu, int update_weights, int do_grid, unsigned* __restrict__ hist,
  double* __restrict__ weights, unsigned long long* __restrict__ invalid, int quad_keys) {

  __shared__ double wblk[count_planes_lds];
  const int nplanes = geo.nr * geo.nc;
  const bool lds_w = update_weights && nplanes <= count_planes_lds;
  if (lds_w)
    for (int k = threadIdx.x; k < nplanes; k += blockDim.x) wblk[k] = 0.0;
  __syncthreads();

  const size_t stride = size_t(gridDim.x) * blockDim.x;
  const size_t i0 = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  // every thread runs the same number of iterations (the wave reductions
  // below need all lanes)
  const size_t iters = (n + stride - 1) / stride;
  for (size_t it = 0; it < iters; ++it) {
    const size_t i = i0 + it * stride;
    bool counted = false;
    unsigned gcube = 0;
    double tw[4] = {0.0, 0.0, 0.0, 0.0};
    if (i < n) {
      const VisRec<N>& r = recs[i];
      const unsigned grp = r.cf_index[1];
      const unsigned cfcube = r.cf_index[0];
      gcube = r.grid_cube;
      bool valid = grp < unsigned(geo.ngroups) && gcube < unsigned(geo.nc);
      if (valid) valid = cfcube < unsigned(groups[grp].ncube);
      if (!valid) {
        atomicAdd(invalid, 1ull);
      } else {
        const GroupInfo g = groups[grp];
        const Coord c = vis_coord(r, geo, g);
        const bool pos_w = r.uvw[2] > 0;
        if (c.inside && update_weights) {
          counted = true;
          const int fxi = c.fx + geo.F, fyi = c.fy + geo.F;
          for (int gpol = 0; gpol < geo.nr; ++gpol) {
            const int* row = pos_w ? mu.M[gpol] : mu.Mc[gpol];
            double t = 0.0;
            for (int v = 0; v < N; ++v) {
              const int m = row[v];
              double mag = 0.0;
              if (m >= 0 && m < g.nmueller) {
                const double2 sw = wsum[g.wsum_offset
                                        + ((long long)(cfcube * g.nmueller + m) * geo.nf
                                           + fxi) * geo.nf + fyi];
                mag = hypot(sw.x, sw.y);
              }
              t += mag * double(r.w[v]);
            }
            tw[gpol] = t;
          }
        }
        if (c.inside && do_grid) {
          const int ntiles = geo.ntx * geo.nty;
          const int tx0 = tiles_begin(c.gx), tx1 = HALO ? tx0 + 1 : tiles_end(c.gx, g.sx);
          const int ty0 = tiles_begin(c.gy), ty1 = HALO ? ty0 + 1 : tiles_end(c.gy, g.sy);
          for (int gpol = 0; gpol < geo.nr; ++gpol) {
            if (!row_has_index(pos_w ? mu.M[gpol] : mu.Mc[gpol], N)) continue;
            const size_t base = size_t(gcube * geo.nr + gpol) * ntiles;
            for (int ty = ty0; ty < ty1; ++ty)
              for (int tx = tx0; tx < tx1; ++tx)
                atomicAdd(hist + bin_key(int(base) + ty * geo.ntx + tx, c.gx, c.gy, quad_keys),
                          1u);
          }
        }
      }
    }
    if (update_weights) {
      unsigned long long active = __ballot(counted);
      const int lane = threadIdx.x % wave;
      while (active) {
        const int leader = __ffsll((long long)active) - 1;
        const unsigned cube = __shfl(gcube, leader, wave);
        const bool mine = counted && gcube == cube;
        active &= ~__ballot(mine);
        for (int gpol = 0; gpol < geo.nr; ++gpol) {
          const double sw = wave_sum(mine ? tw[gpol] : 0.0);
          if (lane == leader) {
            const size_t k = size_t(cube) * geo.nr + gpol;
            if (lds_w)
              __hip_atomic_fetch_add(wblk + k, sw, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            else
              atomicAdd(weights + k, sw);
          }
        }
      }
    }
  }
  if (lds_w) {
    __syncthreads();
    for (int k = threadIdx.x; k < nplanes; k += blockDim.x)
      if (wblk[k] != 0.0) atomicAdd(weights + k, wblk[k]);
  }
}

template <int N, bool HALO>
__global__ __launch_bounds__(block) void k_emit(
  const VisRec<N>* __restrict__ recs, size_t n, Geometry geo,
  const GroupInfo* __restrict__ groups, const double2* __restrict__ wsum,
  Mueller mu, int gvis_ready, int adjoint, const float2* __restrict__ gvis_in,
  const double2* __restrict__ residuals, unsigned* __restrict__ cursor,
  void* __restrict__ entries_out, unsigned* __restrict__ entry_vis, int quad_keys) {

  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const VisRec<N>& r = recs[i];
  const unsigned grp = r.cf_index[1];
  const unsigned cfcube = r.cf_index[0];
  const unsigned gcube = r.grid_cube;
  if (!(grp < unsigned(geo.ngroups) && gcube < unsigned(geo.nc))) return;
  const GroupInfo g = groups[grp];
  if (!(cfcube < unsigned(g.ncube))) return;
  const Coord c = vis_coord(r, geo, g);
  if (!c.inside) return;
  const bool pos_w = r.uvw[2] > 0;
  const bool grad = r.grad[0] != 0.0 || r.grad[1] != 0.0;
  const int fxi = c.fx + geo.F, fyi = c.fy + geo.F;

  Entry<N> e;
  e.h.gx = c.gx;
  e.h.gy = c.gy;
  e.h.fpk = unsigned(fxi) | (unsigned(fyi) << 8)
    | (unsigned((pos_w ? flag_pos_w : 0) | (grad ? flag_gradient : 0)) << 16);
  e.h.gpk = grp | (cfcube << 16);
  if (gvis_ready) {
    for (int v = 0; v < N; ++v) e.g[v] = gvis_in[i * N + v];
  } else {
    const double2 phasor = cphase(r.phase);
    if (!adjoint) {
      for (int v = 0; v < N; ++v) e.g[v] = prephase(r.vis[v], phasor, r.w[v]);
    } else {
      double2 W[N];
      degrid_norms<N>(geo, g, wsum, mu, pos_w, cfcube, fxi, fyi, W);
      for (int v = 0; v < N; ++v) {
        double2 d = W[v];
        if (d.x == 0.0 && d.y == 0.0) d = make_double2(1.0, 0.0);
        d.y = -d.y;
        e.g[v] = narrow(kdiv(cmul(residuals[i * N + v], phasor), d));
      }
    }
  }
  const int ntiles = geo.ntx * geo.nty;
  const int tx0 = tiles_begin(c.gx), tx1 = HALO ? tx0 + 1 : tiles_end(c.gx, g.sx);
  const int ty0 = tiles_begin(c.gy), ty1 = HALO ? ty0 + 1 : tiles_end(c.gy, g.sy);
  for (int gpol = 0; gpol < geo.nr; ++gpol) {
    if (!row_has_index(pos_w ? mu.M[gpol] : mu.Mc[gpol], N)) continue;
    const size_t base = size_t(gcube * geo.nr + gpol) * ntiles;
    for (int ty = ty0; ty < ty1; ++ty)
      for (int tx = tx0; tx < tx1; ++tx) {
        const unsigned pos =
          atomicAdd(cursor + bin_key(int(base) + ty * geo.ntx + tx, c.gx, c.gy, quad_keys), 1u);
        if constexpr (HALO) {
          // the haloed-tile entry (layout of Entry4 below): corner relative
          // to the tile, fine offsets, flags
          const unsigned hpos = unsigned(c.gx - tx0 * tile)
            | (unsigned(c.gy - ty0 * tile) << 6) | ((e.h.fpk & 0xffffu) << 12)
            | (pos_w ? (1u << 28) : 0u) | (grad ? (1u << 29) : 0u);
          unsigned* q = static_cast<unsigned*>(entries_out)
            + size_t(pos) * (2 + 2 * N);
          q[0] = hpos;
          q[1] = e.h.gpk;
          float2* qg = reinterpret_cast<float2*>(q + 2);
          for (int v = 0; v < N; ++v) qg[v] = e.g[v];
        } else {
          static_cast<Entry<N>*>(entries_out)[pos] = e;
        }
        if (grad) entry_vis[pos] = unsigned(i);
      }
  }
}

// Global quadrant bins in two passes over the records instead of the three
// of k_count + k_emit: k_tally makes each record's entries (one per grid row
// with a Mueller index, keyed by the quadrant of the support corner) once,
// takes each entry's place within its key from the returning count, and
// writes the entry and its (key, place) in record order (slot
// record * nr + row); after the scan of the counts, k_place moves each
// entry into its key's contiguous range.  Weights as k_count.
template <int N>
__global__ __launch_bounds__(block) void k_tally(
  const VisRec<N>* __restrict__ recs, size_t n, Geometry geo,
  const GroupInfo* __restrict__ groups, const double2* __restrict__ wsum,
  Mueller mu, int update_weights, int gvis_ready, int adjoint,
  const float2* __restrict__ gvis_in, const double2* __restrict__ residuals,
  unsigned* __restrict__ hist, double* __restrict__ weights,
  unsigned long long* __restrict__ invalid, Entry4<N>* __restrict__ tally,
  uint2* __restrict__ tally_keys) {

  __shared__ double wblk[count_planes_lds];
  const int nplanes = geo.nr * geo.nc;
  const bool lds_w = update_weights && nplanes <= count_planes_lds;
  if (lds_w)
    for (int k = threadIdx.x; k < nplanes; k += blockDim.x) wblk[k] = 0.0;
  __syncthreads();

  const size_t stride = size_t(gridDim.x) * blockDim.x;
  const size_t i0 = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  const size_t iters = (n + stride - 1) / stride;
  const int ntiles = geo.ntx * geo.nty;
  for (size_t it = 0; it < iters; ++it) {
    const size_t i = i0 + it * stride;
    bool counted = false;
    unsigned gcube = 0;
    double tw[4] = {0.0, 0.0, 0.0, 0.0};
    if (i < n) {
      // slots of rows without an entry say so
      for (int gpol = 0; gpol < geo.nr; ++gpol)
        tally_keys[i * geo.nr + gpol] = make_uint2(~0u, 0u);
      const VisRec<N>& r = recs[i];
      const unsigned grp = r.cf_index[1];
      const unsigned cfcube = r.cf_index[0];
      gcube = r.grid_cube;
      bool valid = grp < unsigned(geo.ngroups) && gcube < unsigned(geo.nc);
      if (valid) valid = cfcube < unsigned(groups[grp].ncube);
      if (!valid) {
        atomicAdd(invalid, 1ull);
      } else {
        const GroupInfo g = groups[grp];
        const Coord c = vis_coord(r, geo, g);
        const bool pos_w = r.uvw[2] > 0;
        const int fxi = c.fx + geo.F, fyi = c.fy + geo.F;
        if (c.inside && update_weights) {
          counted = true;
          for (int gpol = 0; gpol < geo.nr; ++gpol) {
            const int* row = pos_w ? mu.M[gpol] : mu.Mc[gpol];
            double t = 0.0;
            for (int v = 0; v < N; ++v) {
              const int m = row[v];
              double mag = 0.0;
              if (m >= 0 && m < g.nmueller) {
                const double2 sw = wsum[g.wsum_offset
                                        + ((long long)(cfcube * g.nmueller + m) * geo.nf
                                           + fxi) * geo.nf + fyi];
                mag = hypot(sw.x, sw.y);
              }
              t += mag * double(r.w[v]);
            }
            tw[gpol] = t;
          }
        }
        if (c.inside) {
          const bool grad = r.grad[0] != 0.0 || r.grad[1] != 0.0;
          Entry4<N> e;
          const int tx = tiles_begin(c.gx), ty = tiles_begin(c.gy);
          e.h.pos = unsigned(c.gx - tx * tile) | (unsigned(c.gy - ty * tile) << 6)
            | (unsigned(fxi) << 12) | (unsigned(fyi) << 20) | (pos_w ? e4_pos_w : 0u)
            | (grad ? e4_gradient : 0u);
          e.h.gpk = grp | (cfcube << 16);
          if (gvis_ready) {
            for (int v = 0; v < N; ++v) e.g[v] = gvis_in[i * N + v];
          } else {
            const double2 phasor = cphase(r.phase);
            if (!adjoint) {
              for (int v = 0; v < N; ++v) e.g[v] = prephase(r.vis[v], phasor, r.w[v]);
            } else {
              double2 W[N];
              degrid_norms<N>(geo, g, wsum, mu, pos_w, cfcube, fxi, fyi, W);
              for (int v = 0; v < N; ++v) {
                double2 d = W[v];
                if (d.x == 0.0 && d.y == 0.0) d = make_double2(1.0, 0.0);
                d.y = -d.y;
                e.g[v] = narrow(kdiv(cmul(residuals[i * N + v], phasor), d));
              }
            }
          }
          for (int gpol = 0; gpol < geo.nr; ++gpol) {
            if (!row_has_index(pos_w ? mu.M[gpol] : mu.Mc[gpol], N)) continue;
            const int key =
              bin_key((int(gcube) * geo.nr + gpol) * ntiles + ty * geo.ntx + tx, c.gx, c.gy, 1);
            const unsigned place = atomicAdd(hist + key, 1u);
            tally[i * geo.nr + gpol] = e;
            tally_keys[i * geo.nr + gpol] = make_uint2(unsigned(key), place);
          }
        }
      }
    }
    if (update_weights) {
      unsigned long long active = __ballot(counted);
      const int lane = threadIdx.x % wave;
      while (active) {
        const int leader = __ffsll((long long)active) - 1;
        const unsigned cube = __shfl(gcube, leader, wave);
        const bool mine = counted && gcube == cube;
        active &= ~__ballot(mine);
        for (int gpol = 0; gpol < geo.nr; ++gpol) {
          const double sw = wave_sum(mine ? tw[gpol] : 0.0);
          if (lane == leader) {
            const size_t k = size_t(cube) * geo.nr + gpol;
            if (lds_w)
              __hip_atomic_fetch_add(wblk + k, sw, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            else
              atomicAdd(weights + k, sw);
          }
        }
      }
    }
  }
  if (lds_w) {
    __syncthreads();
    for (int k = threadIdx.x; k < nplanes; k += blockDim.x)
      if (wblk[k] != 0.0) atomicAdd(weights + k, wblk[k]);
  }
}

template <int N>
__global__ __launch_bounds__(block) void k_place(
  size_t slots, int nr, const unsigned* __restrict__ offsets,
  const Entry4<N>* __restrict__ tally, const uint2* __restrict__ tally_keys,
  Entry4<N>* __restrict__ entries, unsigned* __restrict__ entry_vis) {
  const size_t sl = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (sl >= slots) return;
  const uint2 kp = tally_keys[sl];
  if (kp.x == ~0u) return;
  const unsigned pos = offsets[kp.x] + kp.y;
  const Entry4<N> e = tally[sl];
  entries[pos] = e;
  if (e.h.pos & e4_gradient) entry_vis[pos] = unsigned(sl / size_t(nr));
}

__device__ __forceinline__ int bcast(int v, int k) {
  return __builtin_amdgcn_readlane(v, k);
}
__device__ __forceinline__ float bcast(float v, int k) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), k));
}

// k_tile2 with UNR: tap iterations whose loads are issued together (two for
// four polarizations, whose loads would otherwise take the registers of
// several waves).  The unrolled kernel holds fewer waves, which pays for
// large supports only (unroll_min_taps and more taps).
template <int N>
constexpr int tap_unroll() { return N <= 2 ? 4 : 2; }
constexpr int unroll_min_taps = 2048;

template <int N, int MODE, bool UNR>
__global__ __launch_bounds__(block) void k_tile2(
  Geometry geo, const GroupInfo* __restrict__ groups,
  const float2* __restrict__ taps, Mueller mu,
  const unsigned* __restrict__ offsets, const Entry<N>* __restrict__ entries,
  const unsigned* __restrict__ entry_vis, const VisRec<N>* __restrict__ recs,
  void* __restrict__ grid) {

  constexpr int TT = tile * tile;
  // MODE 0: re[TT], im[TT] (double); MODE 1: hi_re, hi_im, lo_re, lo_im (float)
  __shared__ double acc64[MODE == 0 ? 2 * TT : 1];
  // MODE 1: (hi, lo) pairs of re[TT], then im[TT]
  __shared__ unsigned long long pairs[MODE == 1 ? 2 * TT : 1];

  const unsigned key = blockIdx.x;
  const unsigned beg = offsets[key];
  const unsigned end = offsets[key + 1];
  if (beg == end) return;

  const int ntiles = geo.ntx * geo.nty;
  const unsigned plane = key / ntiles;
  const int t = int(key % ntiles);
  const int x0 = (t % geo.ntx) * tile;
  const int y0 = (t / geo.ntx) * tile;
  const int gpol = int(plane % geo.nr);

  if constexpr (MODE == 0) {
    for (int k = threadIdx.x; k < 2 * TT; k += block) acc64[k] = 0.0;
  } else {
    for (int k = threadIdx.x; k < 2 * TT; k += block) pairs[k] = 0ull;
  }
  __syncthreads();

  const int w = threadIdx.x / wave;
  const int lane = threadIdx.x % wave;

  for (unsigned base = beg + w * wave; base < end; base += tile_waves * wave) {
    const int cnt = int(min(unsigned(wave), end - base));
    Entry<N> mine;
    if (lane < cnt) mine = entries[base + lane];
    for (int k = 0; k < cnt; ++k) {
      const int gx = bcast(mine.h.gx, k);
      const int gy = bcast(mine.h.gy, k);
      const unsigned fpk = unsigned(bcast(int(mine.h.fpk), k));
      const unsigned gpk = unsigned(bcast(int(mine.h.gpk), k));
      float2 gv[N];
      for (int v = 0; v < N; ++v)
        gv[v] = make_float2(bcast(mine.g[v].x, k), bcast(mine.g[v].y, k));
      const int fxi = int(fpk & 0xffu);
      const int fyi = int((fpk >> 8) & 0xffu);
      const unsigned flags = fpk >> 16;
      const int grp = int(gpk & 0xffffu);
      const int cfcube = int(gpk >> 16);
      const GroupInfo g = groups[grp];
      const bool pos_w = flags & flag_pos_w;
      const int* row = pos_w ? mu.M[gpol] : mu.Mc[gpol];
      const float imf = pos_w ? -1.0f : 1.0f;

      const int xa = max(gx, x0) - gx;
      const int xb = min(gx + g.sx, x0 + tile) - gx;
      const int ya = max(gy, y0) - gy;
      const int yb = min(gy + g.sy, y0 + tile) - gy;
      const int wdt = xb - xa;
      const int ntap = wdt * (yb - ya);
      const unsigned mg = magic(unsigned(wdt));

      const float2* slice[N];
      bool use[N];
      const long long slice_taps = (long long)g.sx * g.sy;
      for (int v = 0; v < N; ++v) {
        const int m = row[v];
        use[v] = m >= 0 && m < g.nmueller;
        slice[v] = taps + g.tap_offset
          + (((long long)(cfcube * g.nmueller + (use[v] ? m : 0)) * geo.nf + fxi)
               * geo.nf + fyi) * slice_taps;
      }
      const bool grad = flags & flag_gradient;
      double phi0x = 0.0, phi0y = 0.0, dphix = 0.0, dphiy = 0.0;
      if (grad) {
        const unsigned i = entry_vis[base + k];
        const double gxr = recs[i].grad[0];
        const double gyr = recs[i].grad[1];
        phi0x = -gxr * ((g.sx / 2) * geo.os - (fxi - geo.F));
        phi0y = -gyr * ((g.sy / 2) * geo.os - (fyi - geo.F));
        dphix = gxr * geo.os;
        dphiy = gyr * geo.os;
      }

      if constexpr (!UNR) {
        for (int q = lane; q < ntap; q += wave) {
          const int yl = int(div_magic(unsigned(q), unsigned(wdt), mg));
          const int X = xa + (q - yl * wdt);
          const int Y = ya + yl;
          float2 scr = make_float2(1.0f, 0.0f);
          if (grad) {
            const double phX = phi0x + X * dphix;
            const double phY = phi0y + Y * dphiy;
            scr = narrow(cphase(phX + phY));
          }
          const int li = (gy + Y - y0) * tile + (gx + X - x0);
          double sr = 0.0, si = 0.0;
          for (int v = 0; v < N; ++v) {
            if (!use[v]) continue;
            float2 c = slice[v][Y * g.sx + X];
            c.y *= imf;
            if (grad) c = cmul(c, scr);
            const float2 pr = cmul(c, gv[v]);
            sr += double(pr.x);
            si += double(pr.y);
          }
          if constexpr (MODE == 0) {
            __hip_atomic_fetch_add(acc64 + li, sr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            __hip_atomic_fetch_add(acc64 + TT + li, si, __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_WORKGROUP);
          } else {
            // split the FP64 tap sum into float hi + lo (exact for N == 1)
            const float hr = float(sr), hi = float(si);
            const float lr = float(sr - double(hr)), lim = float(si - double(hi));
            pair_add<(N > 1)>(pairs + li, pairs[li], hr, lr);
            pair_add<(N > 1)>(pairs + TT + li, pairs[TT + li], hi, lim);
          }
        }
        continue;
      }
      // the footprint tile_unroll x 64 taps at a time: the CF loads of all
      // of them issued before the first is used (one step for footprints of
      // at most 64 taps, and for records with a phase gradient: their FP64
      // screen would not fit the registers of the unrolled loop)
      constexpr int tile_unroll = tap_unroll<N>();
      const int unroll = grad || ntap <= wave ? 1 : tile_unroll;
      for (int q0 = lane; q0 < ntap; q0 += unroll * wave) {
        float2 cv[tile_unroll][N];
        int tx[tile_unroll], ty[tile_unroll];
#pragma unroll
        for (int u = 0; u < tile_unroll; ++u) {
          const int q = q0 + u * wave;
          const bool ok = u < unroll && q < ntap;
          const int qq = ok ? q : 0;
          const int yl = int(div_magic(unsigned(qq), unsigned(wdt), mg));
          tx[u] = xa + (qq - yl * wdt);
          ty[u] = ya + yl;
#pragma unroll
          for (int v = 0; v < N; ++v)
            cv[u][v] = use[v] && ok ? slice[v][ty[u] * g.sx + tx[u]] : make_float2(0.0f, 0.0f);
        }
        if (grad) {
          // one tap (u = 0) with its phase screen
          const int X = tx[0], Y = ty[0];
          const double phX = phi0x + X * dphix;
          const double phY = phi0y + Y * dphiy;
          const float2 scr = narrow(cphase(phX + phY));
          const int li = (gy + Y - y0) * tile + (gx + X - x0);
          double sr = 0.0, si = 0.0;
          for (int v = 0; v < N; ++v) {
            if (!use[v]) continue;
            float2 c = cv[0][v];
            c.y *= imf;
            c = cmul(c, scr);
            const float2 pr = cmul(c, gv[v]);
            sr += double(pr.x);
            si += double(pr.y);
          }
          if constexpr (MODE == 0) {
            __hip_atomic_fetch_add(acc64 + li, sr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            __hip_atomic_fetch_add(acc64 + TT + li, si, __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_WORKGROUP);
          } else {
            const float hr = float(sr), hi = float(si);
            const float lr = float(sr - double(hr)), lim = float(si - double(hi));
            pair_add<(N > 1)>(pairs + li, pairs[li], hr, lr);
            pair_add<(N > 1)>(pairs + TT + li, pairs[TT + li], hi, lim);
          }
          continue;
        }
#pragma unroll
        for (int u = 0; u < tile_unroll; ++u) {
          if (q0 + u * wave >= ntap) break;
          const int li = (gy + ty[u] - y0) * tile + (gx + tx[u] - x0);
          double sr = 0.0, si = 0.0;
          for (int v = 0; v < N; ++v) {
            if (!use[v]) continue;
            float2 c = cv[u][v];
            c.y *= imf;
            const float2 pr = cmul(c, gv[v]);
            sr += double(pr.x);
            si += double(pr.y);
          }
          if constexpr (MODE == 0) {
            __hip_atomic_fetch_add(acc64 + li, sr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            __hip_atomic_fetch_add(acc64 + TT + li, si, __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_WORKGROUP);
          } else {
            // split the FP64 tap sum into float hi + lo (exact for N == 1)
            const float hr = float(sr), hi = float(si);
            const float lr = float(sr - double(hr)), lim = float(si - double(hi));
            pair_add<(N > 1)>(pairs + li, pairs[li], hr, lr);
            pair_add<(N > 1)>(pairs + TT + li, pairs[TT + li], hi, lim);
          }
        }
      }
    }
  }
  __syncthreads();

  for (int k = threadIdx.x; k < TT; k += block) {
    const int xx = x0 + k % tile;
    const int yy = y0 + k / tile;
    if (xx >= geo.nx || yy >= geo.ny) continue;
    const size_t gi = (size_t(plane) * geo.ny + yy) * geo.nx + xx;
    if constexpr (MODE == 0) {
      double2* gp = static_cast<double2*>(grid) + gi;
      double2 gvl = *gp;
      gvl.x += acc64[k];
      gvl.y += acc64[TT + k];
      *gp = gvl;
    } else {
      float2* gp = static_cast<float2*>(grid) + gi;
      const float2 g0 = *gp;
      float sr, er, si, ei;
      const unsigned long long pr = pairs[k], pi = pairs[TT + k];
      two_sum(g0.x, __uint_as_float(unsigned(pr)), sr, er);
      two_sum(g0.y, __uint_as_float(unsigned(pi)), si, ei);
      *gp = make_float2(sr + (er + __uint_as_float(unsigned(pr >> 32))),
                        si + (ei + __uint_as_float(unsigned(pi >> 32))));
    }
  }
}

// ---- gridding, block-local bins ----------------------------------------------
//
// For grids with few (plane, tile) keys (<= bin_keys_max) the bins are built
// without any contended global atomic:
// k_bin     one workgroup per 4096 records: a histogram of the block's
//           entries over all keys in LDS, its exclusive scan, the per-block
//           offset table off[block][key] (nkeys + 1 values), then each entry
//           written at its block-local position -- one pass over the records.
//           Weights are reduced per grid plane in LDS first.
// k_tile3   one workgroup per key: gathers the key's segment from every bin
//           block (offset table + LDS scan), then processes the virtual
//           entry list 64 entries per wave; entries with at most 64 taps are
//           handled four at a time with all their CF-tap loads issued
//           together (latency overlap), the others tap-chunk by tap-chunk.

constexpr int bin_threads = 1024;
constexpr int bin_recs_per_thread = 4;
constexpr int bin_records = bin_threads * bin_recs_per_thread;  // per block
constexpr int bin_keys_max = 8192;        // (plane, tile) keys of the block-local bins
constexpr int bin_quad_keys_max = 16384;  // quadrant keys (4 per tile key): LDS histogram
constexpr int bin_planes_max = 1024;
constexpr int tile_bins_max = 2048;  // bin blocks per pass (LDS of k_tile3)
constexpr int bin_stage_bytes = 65536;  // LDS staging of k_bin's haloed entries (one
                                        // chunk for a block of one-polarization entries)

// block-wide exclusive scan of a[0, n) in LDS (n <= bin_quad_keys_max), returns
// the total; `part` holds blockDim.x / wave + 1 values
__device__ unsigned block_exclusive_scan(unsigned* a, int n, unsigned* part) {
  const int nt = blockDim.x;
  const int per = (n + nt - 1) / nt;
  const int beg = min(n, int(threadIdx.x) * per);
  const int end = min(n, beg + per);
  unsigned sum = 0;
  for (int k = beg; k < end; ++k) sum += a[k];
  // inclusive wave scan of the per-thread sums
  const int lane = threadIdx.x % wave;
  unsigned v = sum;
#pragma unroll
  for (int d = 1; d < wave; d <<= 1) {
    const unsigned t = __shfl_up(v, d, wave);
    if (lane >= d) v += t;
  }
  const int w = threadIdx.x / wave;
  if (lane == wave - 1) part[w] = v;
  __syncthreads();
  if (w == 0) {
    // exclusive scan of the wave totals (at most 64 waves) by one wave
    const int nw = nt / wave;
    const unsigned t = lane < nw ? part[lane] : 0u;
    unsigned incl = t;
#pragma unroll
    for (int d = 1; d < wave; d <<= 1) {
      const unsigned u = __shfl_up(incl, d, wave);
      if (lane >= d) incl += u;
    }
    if (lane < nw) part[lane] = incl - t;
    if (lane == nw - 1) part[nw] = incl;
  }
  __syncthreads();
  unsigned run = part[w] + v - sum;  // exclusive prefix of this thread
  for (int k = beg; k < end; ++k) {
    const unsigned t = a[k];
    a[k] = run;
    run += t;
  }
  const unsigned total = part[nt / wave];
  __syncthreads();
  return total;
}


// HALO = false: one entry per (plane, tile) the support touches (Entry<N>,
// for k_tile3); HALO = true: one entry per plane, keyed by the tile of the
// support corner (Entry4<N>, for k_tile4)
template <int N, bool HALO>
__global__ __launch_bounds__(bin_threads) void k_bin(
  const VisRec<N>* __restrict__ recs, size_t n, Geometry geo,
  const GroupInfo* __restrict__ groups, const double2* __restrict__ wsum,
  Mueller mu, int update_weights, int gvis_ready, int adjoint, int degrid_pass,
  const float2* __restrict__ gvis_in, const double2* __restrict__ residuals,
  unsigned* __restrict__ off_table, void* __restrict__ entries_out,
  unsigned* __restrict__ entry_vis, size_t block_cap, double* __restrict__ weights,
  unsigned long long* __restrict__ invalid, VisRec<N>* __restrict__ copy_recs, int quad_keys) {

  __shared__ unsigned cnt[bin_quad_keys_max];
  __shared__ double wl[bin_planes_max];
  __shared__ unsigned part[bin_threads / wave + 1];

  // quad_keys (HALO only): a key per 16 x 16 quadrant of each (plane, tile),
  // key = 4 * tile key + quadrant of the support corner, so that the
  // entries of one quadrant are contiguous within each bin block
  const int nkeys = geo.ntx * geo.nty * geo.nr * geo.nc * (quad_keys ? 4 : 1);
  const int nplanes = geo.nr * geo.nc;
  const bool local_w = nplanes <= bin_planes_max;
  for (int k = threadIdx.x; k < nkeys; k += bin_threads) cnt[k] = 0;
  if (local_w)
    for (int k = threadIdx.x; k < nplanes; k += bin_threads) wl[k] = 0.0;
  __syncthreads();

  const int ntiles = geo.ntx * geo.nty;
  const size_t b = blockIdx.x;
  // weights of this thread's records, per grid row, for one grid cube at a
  // time (records of another cube flush the sums first)
  int wcube = -1;
  double wacc[4] = {0.0, 0.0, 0.0, 0.0};
  auto flush_weights = [&](int cube, const double* acc) {
    for (int gpol = 0; gpol < geo.nr; ++gpol) {
      const int k = cube * geo.nr + gpol;
      if (local_w)
        __hip_atomic_fetch_add(wl + k, acc[gpol], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
      else
        atomicAdd(weights + k, acc[gpol]);
    }
  };
  Entry<N> e[bin_recs_per_thread];
  bool keep[bin_recs_per_thread];
  unsigned short sxy[bin_recs_per_thread][2];
  unsigned gc[bin_recs_per_thread];
  unsigned ninvalid = 0;

  // grid rows with an index, per sign of w (bit gpol): the Mueller rows read
  // at constant indices (kernel-argument scalars), not per record
  unsigned rows_m = 0u, rows_mc = 0u;
#pragma unroll
  for (int gp = 0; gp < 4; ++gp) {
    if (gp >= geo.nr) continue;
    rows_m |= row_has_index(mu.M[gp], N) ? 1u << gp : 0u;
    rows_mc |= row_has_index(mu.Mc[gp], N) ? 1u << gp : 0u;
  }
  const bool one_group = geo.ngroups == 1;
  const GroupInfo group0 = groups[0];

  // N == 1: the thread's records are loaded whole, two at a time, so that
  // two records' loads are in flight at once rather than one record (and its
  // chain of dependent loads) at a time; four would spill the 128 registers
  // a 1024-thread block allows, larger records load per record
  constexpr bool hoist = N == 1;
  constexpr int words = int(sizeof(VisRec<N>) / sizeof(unsigned long long));
  static_assert(sizeof(VisRec<N>) % sizeof(unsigned long long) == 0, "8-byte record words");
  static_assert(bin_recs_per_thread % 2 == 0, "records loaded in pairs");
  VisRec<N> rc[hoist ? 2 : 1];
  // HALO: the place of each of this thread's entries (one per record and
  // grid row with an index; ~0u for none) within its key, from the count's
  // returned value, then (after the scan) its block-local position; the
  // entries are then written through LDS in position order
  unsigned hpos[bin_recs_per_thread][4];
#pragma unroll
  for (int j = 0; j < bin_recs_per_thread; ++j)
#pragma unroll
    for (int r = 0; r < 4; ++r) hpos[j][r] = ~0u;

#pragma unroll
  for (int j = 0; j < bin_recs_per_thread; ++j) {
    if constexpr (hoist) {
      if (j % 2 == 0) {
#pragma unroll
        for (int jj = 0; jj < 2; ++jj) {
          const size_t i = b * bin_records + size_t(j + jj) * bin_threads + threadIdx.x;
          if (i >= n) continue;
          const unsigned long long* src = reinterpret_cast<const unsigned long long*>(recs + i);
          unsigned long long wds[words];
#pragma unroll
          for (int k = 0; k < words; ++k) wds[k] = src[k];
          __builtin_memcpy(&rc[jj], wds, sizeof(VisRec<N>));
        }
      }
    }
    keep[j] = false;
    const size_t i = b * bin_records + size_t(j) * bin_threads + threadIdx.x;
    if (i >= n) continue;
    // N == 2: the whole record in registers first, so its fields load in one
    // round instead of one round per dependent use
    VisRec<N> r_copy;
    if constexpr (N == 2) {
      const unsigned long long* src = reinterpret_cast<const unsigned long long*>(recs + i);
      unsigned long long wds[words];
#pragma unroll
      for (int k = 0; k < words; ++k) wds[k] = src[k];
      __builtin_memcpy(&r_copy, wds, sizeof(r_copy));
    }
    const VisRec<N>& r = hoist ? rc[j % 2] : N == 2 ? r_copy : recs[i];
    // degridding that returns the batch: every record is written to the
    // returned batch here, whose values the degridding kernels then replace
    if (copy_recs) {
      constexpr int words = int(sizeof(VisRec<N>) / sizeof(unsigned long long));
      const unsigned long long* src = reinterpret_cast<const unsigned long long*>(&r);
      unsigned long long* dst = reinterpret_cast<unsigned long long*>(copy_recs + i);
#pragma unroll
      for (int k = 0; k < words; ++k) dst[k] = src[k];
    }
    const unsigned grp = r.cf_index[1];
    const unsigned cfcube = r.cf_index[0];
    const unsigned gcube = r.grid_cube;
    bool valid = grp < unsigned(geo.ngroups) && gcube < unsigned(geo.nc);
    const GroupInfo g = !valid || one_group ? group0 : groups[grp];
    if (valid) valid = cfcube < unsigned(g.ncube);
    if (!valid) {
      // counted by the gridding pass only
      if (!degrid_pass) ++ninvalid;
      continue;
    }
    const Coord c = vis_coord(r, geo, g);
    if (!c.inside) continue;
    const bool pos_w = r.uvw[2] > 0;
    const bool grad = r.grad[0] != 0.0 || r.grad[1] != 0.0;
    const int fxi = c.fx + geo.F, fyi = c.fy + geo.F;
    keep[j] = true;
    gc[j] = gcube;
    sxy[j][0] = (unsigned short)g.sx;
    sxy[j][1] = (unsigned short)g.sy;
    e[j].h.gx = c.gx;
    e[j].h.gy = c.gy;
    e[j].h.fpk = unsigned(fxi) | (unsigned(fyi) << 8)
      | (unsigned((pos_w ? flag_pos_w : 0) | (grad ? flag_gradient : 0)) << 16);
    e[j].h.gpk = grp | (cfcube << 16);
    if (degrid_pass) {
      // degridding reads the model: entries carry no values
      for (int v = 0; v < N; ++v) e[j].g[v] = make_float2(0.0f, 0.0f);
    } else if (gvis_ready) {
      for (int v = 0; v < N; ++v) e[j].g[v] = gvis_in[i * N + v];
    } else {
      const double2 phasor = cphase(r.phase);
      if (!adjoint) {
        for (int v = 0; v < N; ++v) e[j].g[v] = prephase(r.vis[v], phasor, r.w[v]);
      } else {
        double2 W[N];
        degrid_norms<N>(geo, g, wsum, mu, pos_w, cfcube, fxi, fyi, W);
        for (int v = 0; v < N; ++v) {
          double2 d = W[v];
          if (d.x == 0.0 && d.y == 0.0) d = make_double2(1.0, 0.0);
          d.y = -d.y;
          e[j].g[v] = narrow(kdiv(cmul(residuals[i * N + v], phasor), d));
        }
      }
    }
    if (update_weights) {
      // twgt = sum_v |sum_XY cf'| * w_v (src/hpg_core.hpp:877-885); grid
      // rows and polarizations unrolled, so the Mueller indices are scalars
      // and the slice-sum loads of all rows issue together
      if (int(gcube) != wcube) {
        if (wcube >= 0) flush_weights(wcube, wacc);
        wcube = int(gcube);
#pragma unroll
        for (int k = 0; k < 4; ++k) wacc[k] = 0.0;
      }
#pragma unroll
      for (int gp = 0; gp < 4; ++gp) {
        if (gp >= geo.nr) continue;
        double t = 0.0;
#pragma unroll
        for (int v = 0; v < N; ++v) {
          const int m = pos_w ? mu.M[gp][v] : mu.Mc[gp][v];
          double mag = 0.0;
          if (m >= 0 && m < g.nmueller) {
            const double2 sw = wsum[g.wsum_offset
                                    + ((long long)(cfcube * g.nmueller + m) * geo.nf + fxi)
                                      * geo.nf + fyi];
            mag = hypot(sw.x, sw.y);
          }
          t += mag * double(r.w[v]);
        }
        wacc[gp] += t;
      }
    }
    const int tx0 = tiles_begin(c.gx), tx1 = HALO ? tx0 + 1 : tiles_end(c.gx, g.sx);
    const int ty0 = tiles_begin(c.gy), ty1 = HALO ? ty0 + 1 : tiles_end(c.gy, g.sy);
    const unsigned rows = (pos_w != bool(degrid_pass)) ? rows_m : rows_mc;
    for (int gpol = 0; gpol < geo.nr; ++gpol) {
      if (!((rows >> gpol) & 1u)) continue;
      const int base = (int(gcube) * geo.nr + gpol) * ntiles;
      for (int ty = ty0; ty < ty1; ++ty)
        for (int tx = tx0; tx < tx1; ++tx) {
          const unsigned place = __hip_atomic_fetch_add(
            cnt + bin_key(base + ty * geo.ntx + tx, c.gx, c.gy, quad_keys), 1u,
            __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
          if constexpr (HALO) {
#pragma unroll
            for (int r = 0; r < 4; ++r)
              if (r == gpol) hpos[j][r] = place;
          } else {
            (void)place;
          }
        }
    }
  }
  // the threads' weight sums: one sum per (wave, grid cube) present, the
  // wave's lanes all active here
  if (update_weights) {
    const int lane = threadIdx.x % wave;
    int key = wcube;
    for (;;) {
      const unsigned long long act = __ballot(key >= 0);
      if (act == 0) break;
      const int leader = __ffsll((long long)act) - 1;
      const int lk = __shfl(key, leader, wave);
      double part_w[4];
#pragma unroll
      for (int k = 0; k < 4; ++k)
        part_w[k] = k < geo.nr ? wave_sum(key == lk ? wacc[k] : 0.0) : 0.0;
      if (lane == leader) flush_weights(lk, part_w);
      if (key == lk) key = -1;
    }
  }
  // invalid records: one global atomic per wave at most
  {
    unsigned long long inv = ninvalid;
#pragma unroll
    for (int m = wave / 2; m > 0; m >>= 1) inv += __shfl_xor(inv, m, wave);
    if (threadIdx.x % wave == 0 && inv) atomicAdd(invalid, inv);
  }
  __syncthreads();
  const unsigned btotal = block_exclusive_scan(cnt, nkeys, part);
  unsigned* tab = off_table + b * size_t(nkeys + 1);
  for (int k = threadIdx.x; k < nkeys; k += bin_threads) tab[k] = cnt[k];
  if (threadIdx.x == 0) tab[nkeys] = btotal;
  if (update_weights && local_w)
    for (int k = threadIdx.x; k < nplanes; k += bin_threads)
      if (wl[k] != 0.0) atomicAdd(weights + k, wl[k]);
  __syncthreads();

  unsigned* outv = entry_vis + b * block_cap;
#pragma unroll
  for (int j = 0; j < bin_recs_per_thread; ++j) {
    if (!keep[j]) continue;
    const size_t i = size_t(j) * bin_threads + threadIdx.x;
    const bool pos_w = (e[j].h.fpk >> 16) & flag_pos_w;
    const bool grad = (e[j].h.fpk >> 16) & flag_gradient;
    const int tx0 = tiles_begin(e[j].h.gx);
    const int tx1 = HALO ? tx0 + 1 : tiles_end(e[j].h.gx, sxy[j][0]);
    const int ty0 = tiles_begin(e[j].h.gy);
    const int ty1 = HALO ? ty0 + 1 : tiles_end(e[j].h.gy, sxy[j][1]);
    const unsigned rows = (pos_w != bool(degrid_pass)) ? rows_m : rows_mc;
    for (int gpol = 0; gpol < geo.nr; ++gpol) {
      if (!((rows >> gpol) & 1u)) continue;
      const int base = (int(gc[j]) * geo.nr + gpol) * ntiles;
      for (int ty = ty0; ty < ty1; ++ty)
        for (int tx = tx0; tx < tx1; ++tx) {
          const int key = bin_key(base + ty * geo.ntx + tx, e[j].h.gx, e[j].h.gy, quad_keys);
          unsigned pos;
          if constexpr (HALO) {
            // the key's offset (cnt holds the scan now) + the entry's place
            unsigned place = 0;
#pragma unroll
            for (int r = 0; r < 4; ++r)
              if (r == gpol) place = hpos[j][r];
            pos = cnt[key] + place;
#pragma unroll
            for (int r = 0; r < 4; ++r)
              if (r == gpol) hpos[j][r] = pos;
          } else {
            pos = __hip_atomic_fetch_add(cnt + key, 1u, __ATOMIC_RELAXED,
                                         __HIP_MEMORY_SCOPE_WORKGROUP);
            static_cast<Entry<N>*>(entries_out)[b * block_cap + pos] = e[j];
          }
          if (grad || degrid_pass) outv[pos] = unsigned(b * bin_records + i);
        }
    }
  }
  if constexpr (HALO) {
    // chunks of the block's entry range: each thread puts its entries of
    // the chunk into LDS, then the block copies the chunk out contiguously
    // (scattered 16 to 40-byte stores would each write a partial line)
    constexpr unsigned words = unsigned(sizeof(Entry4<N>) / sizeof(unsigned long long));
    constexpr unsigned chunk = unsigned(bin_stage_bytes / sizeof(Entry4<N>));
    __shared__ unsigned long long stage[chunk * words];
    Entry4<N>* staged = reinterpret_cast<Entry4<N>*>(stage);
    unsigned long long* dst =
      reinterpret_cast<unsigned long long*>(static_cast<Entry4<N>*>(entries_out) + b * block_cap);
    for (unsigned c0 = 0; c0 < btotal; c0 += chunk) {
#pragma unroll
      for (int j = 0; j < bin_recs_per_thread; ++j) {
        if (!keep[j]) continue;
        const bool pos_w = (e[j].h.fpk >> 16) & flag_pos_w;
        const bool grad = (e[j].h.fpk >> 16) & flag_gradient;
        Entry4<N> q;
        q.h.pos = unsigned(e[j].h.gx - tiles_begin(e[j].h.gx) * tile)
          | (unsigned(e[j].h.gy - tiles_begin(e[j].h.gy) * tile) << 6)
          | ((e[j].h.fpk & 0xffffu) << 12) | (pos_w ? e4_pos_w : 0u)
          | (grad ? e4_gradient : 0u);
        q.h.gpk = e[j].h.gpk;
        for (int v = 0; v < N; ++v) q.g[v] = e[j].g[v];
#pragma unroll
        for (int r = 0; r < 4; ++r)
          if (hpos[j][r] - c0 < chunk) staged[hpos[j][r] - c0] = q;
      }
      __syncthreads();
      const unsigned n = min(chunk, btotal - c0) * words;
      for (unsigned k = threadIdx.x; k < n; k += bin_threads) dst[size_t(c0) * words + k] = stage[k];
      __syncthreads();
    }
  }
}

// parameters of one entry, uniform across the wave
struct EntryParams {
  int gx, gy, fxi, fyi, grp, cfcube, sx;
  unsigned flags;
};

template <int N, int MODE, int U>
__global__ __launch_bounds__(block) void k_tile3(
  Geometry geo, const GroupInfo* __restrict__ groups,
  const float2* __restrict__ taps, Mueller mu, const unsigned* __restrict__ off_table,
  int nbins, size_t block_cap, const Entry<N>* __restrict__ entries,
  const unsigned* __restrict__ entry_vis, const VisRec<N>* __restrict__ recs,
  void* __restrict__ grid) {

  constexpr int TT = tile * tile;
  __shared__ double acc64[MODE == 0 ? 2 * TT : 1];
  // MODE 1: (hi, lo) pairs of re[TT], then im[TT]
  __shared__ unsigned long long pairs[MODE == 1 ? 2 * TT : 1];
  __shared__ unsigned seg_pref[tile_bins_max + 1];
  __shared__ unsigned seg_beg[tile_bins_max];
  __shared__ unsigned part[block / wave + 1];

  const unsigned key = blockIdx.x;
  const int nkeys = geo.ntx * geo.nty * geo.nr * geo.nc;
  // segment of this key in every bin block
  for (int bb = threadIdx.x; bb < nbins; bb += block) {
    const unsigned* tab = off_table + size_t(bb) * (nkeys + 1);
    const unsigned s0 = tab[key], s1 = tab[key + 1];
    seg_pref[bb] = s1 - s0;
    seg_beg[bb] = s0;
  }
  __syncthreads();
  const unsigned total = block_exclusive_scan(seg_pref, nbins, part);
  if (total == 0) return;
  if (threadIdx.x == 0) seg_pref[nbins] = total;

  const int ntiles = geo.ntx * geo.nty;
  const unsigned plane = key / ntiles;
  const int t = int(key % ntiles);
  const int x0 = (t % geo.ntx) * tile;
  const int y0 = (t / geo.ntx) * tile;
  const int gpol = int(plane % geo.nr);

  if constexpr (MODE == 0) {
    for (int k = threadIdx.x; k < 2 * TT; k += block) acc64[k] = 0.0;
  } else {
    for (int k = threadIdx.x; k < 2 * TT; k += block) pairs[k] = 0ull;
  }
  __syncthreads();

  const int w = threadIdx.x / wave;
  const int lane = threadIdx.x % wave;

  auto accumulate = [&](int li, double sr, double si) {
    if constexpr (MODE == 0) {
      __hip_atomic_fetch_add(acc64 + li, sr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
      __hip_atomic_fetch_add(acc64 + TT + li, si, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    } else {
      const float hr = float(sr), hi = float(si);
      const float lr = float(sr - double(hr)), lim = float(si - double(hi));
      pair_add<(N > 1)>(pairs + li, pairs[li], hr, lr);
      pair_add<(N > 1)>(pairs + TT + li, pairs[TT + li], hi, lim);
    }
  };

  for (unsigned base = w * wave; base < total; base += tile_waves * wave) {
    const int cnt = int(min(unsigned(wave), total - base));
    Entry<N> mine;
    unsigned mine_idx = 0;
    if (lane < cnt) {
      // locate entry base + lane in the concatenated segments
      const unsigned q = base + lane;
      int lo = 0, hi = nbins;  // seg_pref[lo] <= q < seg_pref[hi]
      while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (seg_pref[mid] <= q) lo = mid; else hi = mid;
      }
      mine_idx = unsigned(lo) * unsigned(block_cap) + seg_beg[lo] + (q - seg_pref[lo]);
      mine = entries[mine_idx];
    }
    for (int k0 = 0; k0 < cnt; k0 += U) {
      EntryParams P[U];
      float2 gv[U][N];
      int xa[U], wdt[U], ya[U], ntap[U];
      const float2* slice[U][N];
      bool use[U][N];
      bool fast = true;
      for (int u = 0; u < U; ++u) {
        const int k = min(k0 + u, cnt - 1);
        P[u].gx = bcast(mine.h.gx, k);
        P[u].gy = bcast(mine.h.gy, k);
        const unsigned fpk = unsigned(bcast(int(mine.h.fpk), k));
        const unsigned gpk = unsigned(bcast(int(mine.h.gpk), k));
        P[u].fxi = int(fpk & 0xffu);
        P[u].fyi = int((fpk >> 8) & 0xffu);
        P[u].flags = fpk >> 16;
        P[u].grp = int(gpk & 0xffffu);
        P[u].cfcube = int(gpk >> 16);
        for (int v = 0; v < N; ++v)
          gv[u][v] = make_float2(bcast(mine.g[v].x, k), bcast(mine.g[v].y, k));
        const GroupInfo g = groups[P[u].grp];
        xa[u] = max(P[u].gx, x0) - P[u].gx;
        const int xb = min(P[u].gx + g.sx, x0 + tile) - P[u].gx;
        ya[u] = max(P[u].gy, y0) - P[u].gy;
        const int yb = min(P[u].gy + g.sy, y0 + tile) - P[u].gy;
        wdt[u] = xb - xa[u];
        ntap[u] = k0 + u < cnt ? wdt[u] * (yb - ya[u]) : 0;
        const bool pos_w = P[u].flags & flag_pos_w;
        const int* row = pos_w ? mu.M[gpol] : mu.Mc[gpol];
        const long long slice_taps = (long long)g.sx * g.sy;
        for (int v = 0; v < N; ++v) {
          const int m = row[v];
          use[u][v] = m >= 0 && m < g.nmueller;
          slice[u][v] = taps + g.tap_offset
            + (((long long)(P[u].cfcube * g.nmueller + (use[u][v] ? m : 0)) * geo.nf + P[u].fxi)
                 * geo.nf + P[u].fyi) * slice_taps;
        }
        fast = fast && ntap[u] <= wave && !(P[u].flags & flag_gradient);
        P[u].sx = g.sx;
      }
      if (fast) {
        // all loads first, then the products and the LDS accumulation
        float2 c[U][N];
        int li[U];
        for (int u = 0; u < U; ++u) {
          const unsigned mg = magic(unsigned(max(wdt[u], 1)));
          const int yl = int(div_magic(unsigned(lane), unsigned(max(wdt[u], 1)), mg));
          const int X = xa[u] + (lane - yl * wdt[u]);
          const int Y = ya[u] + yl;
          li[u] = (P[u].gy + Y - y0) * tile + (P[u].gx + X - x0);
          for (int v = 0; v < N; ++v)
            c[u][v] = (lane < ntap[u] && use[u][v]) ? slice[u][v][Y * P[u].sx + X]
                                                    : make_float2(0.0f, 0.0f);
        }
        for (int u = 0; u < U; ++u) {
          if (lane >= ntap[u]) continue;
          const float imf = (P[u].flags & flag_pos_w) ? -1.0f : 1.0f;
          double sr = 0.0, si = 0.0;
          for (int v = 0; v < N; ++v) {
            if (!use[u][v]) continue;
            float2 cc = c[u][v];
            cc.y *= imf;
            const float2 pr = cmul(cc, gv[u][v]);
            sr += double(pr.x);
            si += double(pr.y);
          }
          accumulate(li[u], sr, si);
        }
      } else {
        for (int u = 0; u < U; ++u) {
          if (ntap[u] == 0) continue;
          const int k = k0 + u;
          const bool grad = P[u].flags & flag_gradient;
          const float imf = (P[u].flags & flag_pos_w) ? -1.0f : 1.0f;
          const int sx = P[u].sx;
          double phi0x = 0.0, phi0y = 0.0, dphix = 0.0, dphiy = 0.0;
          if (grad) {
            const unsigned idx = unsigned(bcast(int(mine_idx), k));
            const unsigned i = entry_vis[idx];
            const GroupInfo g = groups[P[u].grp];
            const double gxr = recs[i].grad[0];
            const double gyr = recs[i].grad[1];
            phi0x = -gxr * ((g.sx / 2) * geo.os - (P[u].fxi - geo.F));
            phi0y = -gyr * ((g.sy / 2) * geo.os - (P[u].fyi - geo.F));
            dphix = gxr * geo.os;
            dphiy = gyr * geo.os;
          }
          const unsigned mg = magic(unsigned(wdt[u]));
          for (int q = lane; q < ntap[u]; q += wave) {
            const int yl = int(div_magic(unsigned(q), unsigned(wdt[u]), mg));
            const int X = xa[u] + (q - yl * wdt[u]);
            const int Y = ya[u] + yl;
            float2 scr = make_float2(1.0f, 0.0f);
            if (grad) {
              const double phX = phi0x + X * dphix;
              const double phY = phi0y + Y * dphiy;
              scr = narrow(cphase(phX + phY));
            }
            double sr = 0.0, si = 0.0;
            for (int v = 0; v < N; ++v) {
              if (!use[u][v]) continue;
              float2 cc = slice[u][v][Y * sx + X];
              cc.y *= imf;
              if (grad) cc = cmul(cc, scr);
              const float2 pr = cmul(cc, gv[u][v]);
              sr += double(pr.x);
              si += double(pr.y);
            }
            accumulate((P[u].gy + Y - y0) * tile + (P[u].gx + X - x0), sr, si);
          }
        }
      }
    }
  }
  __syncthreads();

  for (int k = threadIdx.x; k < TT; k += block) {
    const int xx = x0 + k % tile;
    const int yy = y0 + k / tile;
    if (xx >= geo.nx || yy >= geo.ny) continue;
    const size_t gi = (size_t(plane) * geo.ny + yy) * geo.nx + xx;
    if constexpr (MODE == 0) {
      double2* gp = static_cast<double2*>(grid) + gi;
      double2 gvl = *gp;
      gvl.x += acc64[k];
      gvl.y += acc64[TT + k];
      *gp = gvl;
    } else {
      float2* gp = static_cast<float2*>(grid) + gi;
      const float2 g0 = *gp;
      float sr, er, si, ei;
      const unsigned long long pr = pairs[k], pi = pairs[TT + k];
      two_sum(g0.x, __uint_as_float(unsigned(pr)), sr, er);
      two_sum(g0.y, __uint_as_float(unsigned(pi)), si, ei);
      *gp = make_float2(sr + (er + __uint_as_float(unsigned(pr >> 32))),
                        si + (ei + __uint_as_float(unsigned(pi >> 32))));
    }
  }
}

// ---- gridding, haloed tiles ---------------------------------------------
//
// k_bin<N, true> writes one Entry4 per (record, plane) into the bin of the
// tile that holds the support corner.  The workgroup of a key accumulates
// the tile plus a halo of hx = max_sx - 1 columns and hy = max_sy - 1 rows on
// its high sides (R = (tile + hx) x (tile + hy) pixels in LDS), so every tap
// of an entry lands in LDS without clipping.  One lane carries one entry
// through all of its taps: per tap a CF load, the FP32 product and the LDS
// accumulation, with no cross-lane traffic.  The workgroup then adds its own
// tile into the grid (it is the only writer of those pixels during the
// kernel) and stores the halo, as FP64 values, into a per-key side buffer;
// k_halo adds each tile's neighbours' halos into it afterwards, in a fixed
// order.  Requires hx, hy < tile.

constexpr int halo_block = 256;

// The dispatcher hands workgroups to the 8 XCDs in turn (workgroup b to
// XCD b mod 8).  The keyed kernels take key xcd_key(b) instead of b, so
// that each XCD works through one contiguous run of keys: neighbouring
// keys' bin segments and offset-table entries share cache lines, which
// are then fetched into one XCD's L2 instead of several.  A bijection on
// [0, nkeys).
constexpr unsigned xcd_count = 8;
__device__ __forceinline__ unsigned xcd_key(unsigned b, unsigned nkeys) {
  const unsigned x = b % xcd_count, j = b / xcd_count;
  const unsigned q = nkeys / xcd_count, r = nkeys % xcd_count;
  return x * q + min(x, r) + j;
}
constexpr int halo_chunk = 8;     // CF taps loaded together per lane
constexpr int halo_classes = 32;  // LDS bank classes of the entry schedule
constexpr int halo_sched = 2048;  // entries per scheduling round

// pixels of the halo of one key: columns [tile, tile + hx) of rows
// [0, tile), then all columns of rows [tile, tile + hy)
__host__ __device__ inline int halo_pixels(int hx, int hy) {
  return hx * tile + (tile + hx) * hy;
}

// tile accumulation schemes of k_tile4
//   acc_f64       FP64 LDS atomics (FP64 grids; also selectable for FP32 grids)
//   acc_f32_cas   FP32 hi/lo per pixel packed into 64 bits and updated by
//                 compare-and-swap: hi' = fl(hi + p), lo' = fl(lo + err) with
//                 err the exact two-sum error, so each update is one
//                 error-free transformation on one coherent pair
constexpr int acc_f64 = 0;
constexpr int acc_f32_cas = 2;

template <int N, int MODE, int ACC>
__global__ __launch_bounds__(halo_block) void k_tile4(
  Geometry geo, const GroupInfo* __restrict__ groups,
  const float2* __restrict__ taps, Mueller mu, const unsigned* __restrict__ off_table,
  int nbins, size_t block_cap, const Entry4<N>* __restrict__ entries,
  const unsigned* __restrict__ entry_vis, const VisRec<N>* __restrict__ recs,
  int hx, int hy, void* __restrict__ grid, double2* __restrict__ halo,
  unsigned* __restrict__ halo_set) {
  static_assert(MODE == 1 || ACC == acc_f64, "FP64 grids accumulate in FP64");

  extern __shared__ double smem[];
  const int RX = tile + hx;
  const int RY = tile + hy;
  const int R = RX * RY;
  // LDS: accumulators (16 bytes per region pixel), then the segment tables
  double* acc64 = smem;                                // acc_f64: re[R], im[R]
  unsigned long long* pairs = reinterpret_cast<unsigned long long*>(smem);  // cas: re[R], im[R]
  unsigned* seg_pref = reinterpret_cast<unsigned*>(smem + 2 * R);
  unsigned* seg_beg = seg_pref + nbins + 1;
  unsigned* part = seg_beg + nbins;
  unsigned* cls = part + (halo_block / wave + 1);  // bank classes: counts, then cursors
  unsigned* sched = cls + halo_classes;            // entry indexes of one round

  const unsigned key = xcd_key(blockIdx.x, gridDim.x);  // one block per key
  const int nkeys = geo.ntx * geo.nty * geo.nr * geo.nc;
  for (int bb = threadIdx.x; bb < nbins; bb += halo_block) {
    const unsigned* tab = off_table + size_t(bb) * (nkeys + 1);
    const unsigned s0 = tab[key], s1 = tab[key + 1];
    seg_pref[bb] = s1 - s0;
    seg_beg[bb] = s0;
  }
  __syncthreads();
  const unsigned total = block_exclusive_scan(seg_pref, nbins, part);
  if (total == 0) {
    if (threadIdx.x == 0) halo_set[key] = 0u;
    return;
  }
  if (threadIdx.x == 0) seg_pref[nbins] = total;

  const int ntiles = geo.ntx * geo.nty;
  const unsigned plane = key / ntiles;
  const int t = int(key % ntiles);
  const int x0 = (t % geo.ntx) * tile;
  const int y0 = (t / geo.ntx) * tile;
  const int gpol = int(plane % geo.nr);

  for (int k = threadIdx.x; k < 2 * R; k += halo_block) acc64[k] = 0.0;
  __syncthreads();

  const int w = threadIdx.x / wave;
  const int lane = threadIdx.x % wave;
  constexpr int nw = halo_block / wave;

  // entry q of the key (0 <= q < total): its index in the entry array
  auto locate = [&](unsigned q) {
    int lo = 0, hi = nbins;  // seg_pref[lo] <= q < seg_pref[hi]
    while (hi - lo > 1) {
      const int mid = (lo + hi) >> 1;
      if (seg_pref[mid] <= q) lo = mid; else hi = mid;
    }
    return unsigned(lo) * unsigned(block_cap) + seg_beg[lo] + (q - seg_pref[lo]);
  };
  // LDS bank class of an entry's first tap; every tap of every lane adds
  // the same offset to it, so a wave whose lanes hold consecutive classes
  // spreads each of its LDS updates evenly over the banks
  auto bank_class = [&](unsigned pos) {
    return unsigned(int((pos >> 6) & 63u) * RX + int(pos & 63u)) % unsigned(halo_classes);
  };
  // rank of this lane in a wave's sample of the class-sorted round: lanes
  // l and l + 32 take neighbouring samples, so lanes 0-31 (and 32-63) hold
  // ascending classes
  const unsigned rank = unsigned((lane & 31) * 2 + (lane >> 5));

  for (unsigned r0 = 0; r0 < total; r0 += halo_sched) {
    const unsigned rn = min(unsigned(halo_sched), total - r0);
    // counting sort of the round's entries by bank class
    if (threadIdx.x < halo_classes) cls[threadIdx.x] = 0u;
    __syncthreads();
    for (unsigned q = threadIdx.x; q < rn; q += halo_block)
      __hip_atomic_fetch_add(cls + bank_class(entries[locate(r0 + q)].h.pos), 1u,
                             __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    __syncthreads();
    if (threadIdx.x == 0) {
      unsigned run = 0;
      for (int c = 0; c < halo_classes; ++c) {
        const unsigned t = cls[c];
        cls[c] = run;
        run += t;
      }
    }
    __syncthreads();
    for (unsigned q = threadIdx.x; q < rn; q += halo_block) {
      const unsigned idx = locate(r0 + q);
      const unsigned slot = __hip_atomic_fetch_add(
        cls + bank_class(entries[idx].h.pos), 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
      sched[slot] = idx;
    }
    __syncthreads();
    // batch j of the round takes sorted entries rank * nbatch + j
    const unsigned nbatch = (rn + wave - 1) / wave;
    for (unsigned j = w; j < nbatch; j += nw) {
      const unsigned sp = rank * nbatch + j;
      if (sp >= rn) continue;
      const unsigned idx = sched[sp];
      const Entry4<N> e = entries[idx];
      const unsigned pos = e.h.pos;
      const int lx = int(pos & 63u);
      const int ly = int((pos >> 6) & 63u);
      const int fxi = int((pos >> 12) & 255u);
      const int fyi = int((pos >> 20) & 255u);
      const bool pos_w = pos & e4_pos_w;
      const bool grad = pos & e4_gradient;
      const int grp = int(e.h.gpk & 0xffffu);
      const int cfcube = int(e.h.gpk >> 16);
      const GroupInfo g = groups[grp];
      const int sx = g.sx, sy = g.sy;
      const int* row = pos_w ? mu.M[gpol] : mu.Mc[gpol];
      const float imf = pos_w ? -1.0f : 1.0f;
      const float2* slice[N];
      bool use[N];
      const long long slice_taps = (long long)sx * sy;
      for (int v = 0; v < N; ++v) {
        const int m = row[v];
        use[v] = m >= 0 && m < g.nmueller;
        slice[v] = taps + g.tap_offset
          + (((long long)(cfcube * g.nmueller + (use[v] ? m : 0)) * geo.nf + fxi) * geo.nf
             + fyi) * slice_taps;
      }
      double phi0x = 0.0, phi0y = 0.0, dphix = 0.0, dphiy = 0.0;
      if (grad) {
        const unsigned i = entry_vis[idx];
        const double gxr = recs[i].grad[0];
        const double gyr = recs[i].grad[1];
        phi0x = -gxr * ((sx / 2) * geo.os - (fxi - geo.F));
        phi0y = -gyr * ((sy / 2) * geo.os - (fyi - geo.F));
        dphix = gxr * geo.os;
        dphiy = gyr * geo.os;
      }
      // the value of one tap: hi (FP32) and, for N > 1, the low-order part
      // of the FP64 sum over the polarizations; d = the FP64 value
      auto tap_value = [&](const float2* cv, int X, int Y, float& h_r, float& h_i,
                           float& l_r, float& l_i, double& d_r, double& d_i) {
        float2 scr = make_float2(1.0f, 0.0f);
        if (grad) scr = narrow(cphase((phi0x + X * dphix) + (phi0y + Y * dphiy)));
        if constexpr (N == 1) {
          float2 cc = cv[0];
          cc.y *= imf;
          if (grad) cc = cmul(cc, scr);
          const float2 pr = cmul(cc, e.g[0]);
          h_r = pr.x;
          h_i = pr.y;
          l_r = 0.0f;
          l_i = 0.0f;
          d_r = double(pr.x);
          d_i = double(pr.y);
        } else {
          double sr = 0.0, si = 0.0;
#pragma unroll
          for (int v = 0; v < N; ++v) {
            if (!use[v]) continue;
            float2 cc = cv[v];
            cc.y *= imf;
            if (grad) cc = cmul(cc, scr);
            const float2 pr = cmul(cc, e.g[v]);
            sr += double(pr.x);
            si += double(pr.y);
          }
          d_r = sr;
          d_i = si;
          h_r = float(sr);
          h_i = float(si);
          l_r = float(sr - double(h_r));
          l_i = float(si - double(h_i));
        }
      };
      // adds the M taps held in the arrays to LDS pixels li[0..M) (ok[m]:
      // tap present).  The compare-and-swap form reads each pixel just
      // before its swap: the wave's LDS operations execute in order, so two
      // lanes of one wave only race on a pixel when their corners coincide;
      // failed swaps are retried after the last tap.
      auto accumulate = [&](auto m_c, const int* li, const bool* ok, const float* h_r,
                            const float* h_i, const float* l_r, const float* l_i,
                            const double* d_r, const double* d_i) {
        constexpr int M = decltype(m_c)::value;
        if constexpr (ACC == acc_f64) {
#pragma unroll
          for (int m = 0; m < M; ++m) {
            if (!ok[m]) continue;
            __hip_atomic_fetch_add(acc64 + li[m], d_r[m], __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_WORKGROUP);
            __hip_atomic_fetch_add(acc64 + R + li[m], d_i[m], __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_WORKGROUP);
          }
        } else {
          // pixel li holds (H_re, L_re) at pairs[li], (H_im, L_im) at pairs[R + li]
          constexpr bool has_l = N > 1;
          unsigned long long er[M], ei[M];
          bool okr[M], oki[M];
#pragma unroll
          for (int m = 0; m < M; ++m) {
            if (!ok[m]) continue;
            er[m] = pairs[li[m]];
            ei[m] = pairs[R + li[m]];
            okr[m] = pair_swap(pairs + li[m], er[m], pair_plus<has_l>(er[m], h_r[m], l_r[m]));
            oki[m] = pair_swap(pairs + R + li[m], ei[m],
                               pair_plus<has_l>(ei[m], h_i[m], l_i[m]));
          }
#pragma unroll
          for (int m = 0; m < M; ++m) {
            if (!ok[m]) continue;
            if (!okr[m]) pair_add<has_l>(pairs + li[m], er[m], h_r[m], l_r[m]);
            if (!oki[m]) pair_add<has_l>(pairs + R + li[m], ei[m], h_i[m], l_i[m]);
          }
        }
      };
      // any support: taps in slice order (t = Y * sx + X), halo_chunk at a
      // time; full chunks are straight-line code
      auto run_chunks = [&]() {
        const int ntap = sx * sy;
        int cx = 0, cy = 0;          // (X, Y) of the chunk's first tap
        int lc = ly * RX + lx;       // its LDS pixel
        const int wrap = RX - sx;    // pixel step from a row end to the next row
        auto do_chunk = [&](auto full, int t0) {
          constexpr bool FULL = decltype(full)::value;
          bool ok[halo_chunk];
          int li[halo_chunk];
          float h_r[halo_chunk], h_i[halo_chunk], l_r[halo_chunk], l_i[halo_chunk];
          double d_r[halo_chunk], d_i[halo_chunk];
          float2 c[halo_chunk][N];
          int Xk[halo_chunk], Yk[halo_chunk];
#pragma unroll
          for (int k = 0; k < halo_chunk; ++k) {
            ok[k] = FULL || t0 + k < ntap;
            Xk[k] = cx;
            Yk[k] = cy;
            li[k] = lc;
            const bool eol = ++cx == sx;
            cx = eol ? 0 : cx;
            cy += eol ? 1 : 0;
            lc += eol ? 1 + wrap : 1;
#pragma unroll
            for (int v = 0; v < N; ++v) {
              float2 tv = make_float2(0.0f, 0.0f);
              if (ok[k]) tv = slice[v][t0 + k];
              c[k][v] = use[v] ? tv : make_float2(0.0f, 0.0f);
            }
          }
#pragma unroll
          for (int k = 0; k < halo_chunk; ++k)
            tap_value(c[k], Xk[k], Yk[k], h_r[k], h_i[k], l_r[k], l_i[k], d_r[k], d_i[k]);
          accumulate(std::integral_constant<int, halo_chunk>{}, li, ok, h_r, h_i, l_r, l_i,
                     d_r, d_i);
        };
        int t0 = 0;
        for (; t0 + halo_chunk <= ntap; t0 += halo_chunk) do_chunk(std::true_type{}, t0);
        if (t0 < ntap) do_chunk(std::false_type{}, t0);
      };
      run_chunks();
    }
    __syncthreads();  // the next round reuses cls and sched
  }

  const int hs = halo_pixels(hx, hy);
  double2* hk = halo + size_t(key) * hs;
  for (int k = threadIdx.x; k < R; k += halo_block) {
    const int ry = k / RX;
    const int rx = k - ry * RX;
    // tile value: FP64 sum (acc_f64) or FP32 hi + lo
    double vr, vi;
    float fr = 0.0f, fi = 0.0f, flr = 0.0f, fli = 0.0f;
    if constexpr (ACC == acc_f64) {
      vr = acc64[k];
      vi = acc64[R + k];
    } else {
      const unsigned long long pr = pairs[k], pi = pairs[R + k];
      fr = __uint_as_float(unsigned(pr));
      flr = __uint_as_float(unsigned(pr >> 32));
      fi = __uint_as_float(unsigned(pi));
      fli = __uint_as_float(unsigned(pi >> 32));
      vr = double(fr) + double(flr);
      vi = double(fi) + double(fli);
    }
    if (rx < tile && ry < tile) {
      const int xx = x0 + rx, yy = y0 + ry;
      if (xx >= geo.nx || yy >= geo.ny) continue;
      const size_t gi = (size_t(plane) * geo.ny + yy) * geo.nx + xx;
      if constexpr (MODE == 0) {
        double2* gp = static_cast<double2*>(grid) + gi;
        double2 gvl = *gp;
        gvl.x += vr;
        gvl.y += vi;
        *gp = gvl;
      } else if constexpr (ACC == acc_f64) {
        // one rounding of grid value + FP64 tile sum
        float2* gp = static_cast<float2*>(grid) + gi;
        const float2 g0 = *gp;
        *gp = make_float2(float(double(g0.x) + vr), float(double(g0.y) + vi));
      } else {
        float2* gp = static_cast<float2*>(grid) + gi;
        const float2 g0 = *gp;
        float sr, er, si, ei;
        two_sum(g0.x, fr, sr, er);
        two_sum(g0.y, fi, si, ei);
        *gp = make_float2(sr + (er + flr), si + (ei + fli));
      }
    } else {
      const int h = ry < tile ? ry * hx + (rx - tile) : tile * hx + (ry - tile) * RX + rx;
      hk[h] = make_double2(vr, vi);
    }
  }
  if (threadIdx.x == 0) halo_set[key] = 1u;
}

// adds the halos of each tile's left, upper and upper-left neighbours into
// the tile (pixels (u, v) of the tile with u < hx or v < hy), one thread per
// pixel, in that fixed order
template <typename T>
__global__ void k_halo(
  Geometry geo, int hx, int hy, const double2* __restrict__ halo,
  const unsigned* __restrict__ halo_set, size_t nkeys, T* __restrict__ grid) {
  const int zone = hy * tile + (tile - hy) * hx;
  const size_t tid = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (zone == 0 || tid >= nkeys * zone) return;
  const size_t key = tid / zone;
  const int z = int(tid - key * zone);
  int u, v;
  if (z < hy * tile) {
    v = z / tile;
    u = z - v * tile;
  } else {
    const int z2 = z - hy * tile;
    v = hy + z2 / hx;
    u = z2 - (v - hy) * hx;
  }
  const int ntiles = geo.ntx * geo.nty;
  const size_t plane = key / ntiles;
  const int t = int(key - plane * ntiles);
  const int TX = t % geo.ntx;
  const int TY = t / geo.ntx;
  const int x = TX * tile + u;
  const int y = TY * tile + v;
  if (x >= geo.nx || y >= geo.ny) return;
  const int hs = halo_pixels(hx, hy);
  const int RX = tile + hx;
  // the neighbours' flags, their halo values and the grid value load
  // together (a halo not written by this pass is read but not added)
  const bool has_l = TX > 0 && u < hx;
  const bool has_u = TY > 0 && v < hy;
  const bool has_ul = has_l && has_u;
  const unsigned set_l = has_l ? halo_set[key - 1] : 0u;
  const unsigned set_u = has_u ? halo_set[key - geo.ntx] : 0u;
  const unsigned set_ul = has_ul ? halo_set[key - geo.ntx - 1] : 0u;
  double2 h_l = make_double2(0.0, 0.0), h_u = h_l, h_ul = h_l;
  if (has_l) h_l = halo[(key - 1) * hs + v * hx + u];
  if (has_u) h_u = halo[(key - geo.ntx) * hs + tile * hx + v * RX + u];
  if (has_ul) h_ul = halo[(key - geo.ntx - 1) * hs + tile * hx + v * RX + tile + u];
  T* gp = grid + (plane * geo.ny + y) * geo.nx + x;
  T g = *gp;
  if (!(set_l | set_u | set_ul)) return;
  // added in the order left, upper, upper-left
  double sr = 0.0, si = 0.0;
  if (set_l) {
    sr += h_l.x;
    si += h_l.y;
  }
  if (set_u) {
    sr += h_u.x;
    si += h_u.y;
  }
  if (set_ul) {
    sr += h_ul.x;
    si += h_ul.y;
  }
  g.x = double(g.x) + sr;
  g.y = double(g.y) + si;
  *gp = g;
}

// ---- gridding: haloed tiles, one wave per tile quadrant, one lane per tap --
//
// k_quad accumulates the keys and entries of k_tile4 into an FP32 grid,
// (hi, lo) pairs per pixel updated by two-sum as in k_tile4's
// compare-and-swap scheme, without LDS atomics.
// The workgroup of a (plane, tile) key has four waves; wave w owns quadrant
// (w & 1, w >> 1) of the 32 x 32 tile -- the entries whose support corner
// lies in that 16 x 16 quadrant -- and a private LDS region of
// (16 + hx) x (16 + hy) pixels, which holds those supports whole.  The wave
// adds one entry at a time, one lane per CF tap: an entry's taps are
// distinct pixels, a wave's LDS operations execute in order and no other wave
// writes the region, so each tap is a plain 16-byte read-modify-write.  Per
// round each thread prepares one entry (CF slice, values, region pixel) and
// the round's entries are sorted by quadrant into LDS.  The flush sums the
// four regions over the tile and its halo and writes both as k_tile4 does,
// so k_halo folds the halos unchanged.
constexpr int quad_block = 4 * wave;
constexpr int quad_edge = tile / 2;
constexpr int quad_bins_max = 256;  // bin blocks per pass (LDS segment tables)
constexpr int quad_flush = 8;       // tile + halo pixels per thread in the flush
// entries whose first CF taps are loaded ahead of their turn (2 measured
// faster than 1, 3 or 4 for one polarization: the pipeline refills at
// every round)
template <int N>
constexpr int quad_depth() { return 2; }

// an entry prepared for its quadrant's wave; gc = (pos_w ? -1 : 1) * g with
// re and im swapped, so that cmul(conj_w(c), g) = (c.x g.x - c.y gc.x,
// c.x g.y + c.y gc.y) with the reference's roundings.  Entries with a phase
// gradient do not use gc: gc[0].x holds the bits of their entry index.
template <int N>
struct QuadRec {
  unsigned w0;        // region pixel | sx << 12 | sy << 18 | use << 24 | flags
  unsigned slice[N];  // first tap of each polarization's CF slice
  float2 g[N];
  float2 gc[N];
};
constexpr unsigned quad_gradient = 1u << 28;
constexpr unsigned quad_pos_w = 1u << 29;
constexpr unsigned quad_skip = 1u << 30;  // not added by the pipelined loop
                                          // (phase gradient, or no used
                                          // polarization)
template <int N>
constexpr int quad_round() { return N == 1 ? quad_block : quad_block / 2; }

__device__ __forceinline__ unsigned uniform_u(unsigned v) {
  return unsigned(__builtin_amdgcn_readfirstlane(int(v)));
}

// adds the tap value h + l (FP32 high and low parts) to the (hi, lo) pairs
// of a region pixel: hi' = fl(hi + h), lo' = fl(lo + (err + l)), err the
// exact rounding error of the first sum (two-sum)
template <int N>
__device__ __forceinline__ void quad_add(
  float4* region, int px, float h_r, float h_i, float l_r, float l_i) {
  float4 a = region[px];
  float sum, err;
  two_sum(a.x, h_r, sum, err);
  a.x = sum;
  a.z = N > 1 ? a.z + (err + l_r) : a.z + err;
  two_sum(a.y, h_i, sum, err);
  a.y = sum;
  a.w = N > 1 ? a.w + (err + l_i) : a.w + err;
  region[px] = a;
}

// an entry with a phase gradient: its CF taps times the phase screen
// exp(i phi(X, Y)) (FP64 phase, the reference's cphase) added by one wave,
// one lane per tap.  Not inlined: the FP64 sine and cosine would otherwise
// raise the register count of k_quad's pipelined loop.
template <int N>
__device__ __noinline__ void quad_gradient_entry(
  const QuadRec<N>& p, unsigned ei, const Geometry& geo, const float2* __restrict__ taps,
  const Entry4<N>* __restrict__ entries, const unsigned* __restrict__ entry_vis,
  const VisRec<N>* __restrict__ recs, int QX, float4* region) {
  const int lane = int(threadIdx.x % wave);
  const unsigned w0 = uniform_u(p.w0);
  const int base = int(w0 & 4095u);
  const int sx = int((w0 >> 12) & 63u);
  const int sy = int((w0 >> 18) & 63u);
  const unsigned use = (w0 >> 24) & 15u;
  const float imf = (w0 & quad_pos_w) ? -1.0f : 1.0f;
  ei = uniform_u(ei);
  const unsigned pos = entries[ei].h.pos;
  const int fxi = int((pos >> 12) & 255u);
  const int fyi = int((pos >> 20) & 255u);
  const unsigned ri = entry_vis[ei];
  const double gxr = recs[ri].grad[0];
  const double gyr = recs[ri].grad[1];
  const double phi0x = -gxr * ((sx / 2) * geo.os - (fxi - geo.F));
  const double phi0y = -gyr * ((sy / 2) * geo.os - (fyi - geo.F));
  const double dphix = gxr * geo.os;
  const double dphiy = gyr * geo.os;
  for (int t0 = 0; t0 < sx * sy; t0 += wave) {
    const int tp = t0 + lane;
    if (tp >= sx * sy) continue;
    const int Y = tp / sx;
    const int X = tp - Y * sx;
    const float2 scr = narrow(cphase((phi0x + X * dphix) + (phi0y + Y * dphiy)));
    double sr = 0.0, si = 0.0;
    float h_r = 0.0f, h_i = 0.0f, l_r = 0.0f, l_i = 0.0f;
    for (int v = 0; v < N; ++v) {
      if (!((use >> v) & 1u)) continue;
      float2 cc = taps[uniform_u(p.slice[v]) + size_t(tp)];
      cc.y *= imf;
      const float2 pr = cmul(cmul(cc, scr), p.g[v]);
      if constexpr (N == 1) {
        h_r = pr.x;
        h_i = pr.y;
      }
      sr += double(pr.x);
      si += double(pr.y);
    }
    if constexpr (N > 1) {
      h_r = float(sr);
      h_i = float(si);
      l_r = float(sr - double(h_r));
      l_i = float(si - double(h_i));
    }
    quad_add<N>(region, base + Y * QX + X, h_r, h_i, l_r, l_i);
  }
}

// The flush of the quadrant kernels: the sum of the four quadrant regions
// of (hi, lo) pairs (LDS, float4 per pixel, quadrant qd at qd * QX * QY)
// over the tile and its halo.  Tile pixels are added to the FP32 grid with
// one rounding (a two-sum of the grid value and hi, then hi + (err + lo));
// halo pixels are stored as FP64 hi + lo into the key's halo values hk.
// BLOCK threads, PER pixels per thread: BLOCK * PER >= (tile + hx) * (tile + hy).
template <int BLOCK, int PER>
__device__ __forceinline__ void quad_flush_tile(
  const float4* regions, int QX, int QY, int hx, int hy, const Geometry& geo,
  unsigned plane, int x0, int y0, float2* __restrict__ grid, double2* __restrict__ hk) {
  const int Q = QX * QY;
  const int RX = tile + hx;
  const int RY = tile + hy;
  const int R = RX * RY;
  // every pixel's value first, with the grid loads of the tile's pixels
  // issued together, then the stores
  float4 val[PER];
  float2 g0[PER];
#pragma unroll
  for (int j = 0; j < PER; ++j) {
    const int k = int(threadIdx.x) + j * BLOCK;
    val[j] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    g0[j] = make_float2(0.0f, 0.0f);
    if (k >= R) continue;
    const int ry = k / RX;
    const int rx = k - ry * RX;
    // (hi, lo) of the pixel: the quadrants' pairs combined by two-sum
    float fr = 0.0f, fi = 0.0f, flr = 0.0f, fli = 0.0f;
    bool first = true;
#pragma unroll
    for (int qd = 0; qd < 4; ++qd) {
      const int ux = rx - (qd & 1) * quad_edge;
      const int uy = ry - (qd >> 1) * quad_edge;
      if (ux < 0 || uy < 0 || ux >= QX || uy >= QY) continue;
      const float4 a = regions[qd * Q + uy * QX + ux];
      if (first) {
        fr = a.x;
        fi = a.y;
        flr = a.z;
        fli = a.w;
      } else {
        float sum, err;
        two_sum(fr, a.x, sum, err);
        fr = sum;
        flr = flr + (err + a.z);
        two_sum(fi, a.y, sum, err);
        fi = sum;
        fli = fli + (err + a.w);
      }
      first = false;
    }
    val[j] = make_float4(fr, fi, flr, fli);
    const int xx = x0 + rx, yy = y0 + ry;
    if (rx < tile && ry < tile && xx < geo.nx && yy < geo.ny)
      g0[j] = grid[(size_t(plane) * geo.ny + yy) * geo.nx + xx];
  }
#pragma unroll
  for (int j = 0; j < PER; ++j) {
    const int k = int(threadIdx.x) + j * BLOCK;
    if (k >= R) continue;
    const int ry = k / RX;
    const int rx = k - ry * RX;
    const float4 v = val[j];
    if (rx < tile && ry < tile) {
      const int xx = x0 + rx, yy = y0 + ry;
      if (xx >= geo.nx || yy >= geo.ny) continue;
      float sr, er, si, ei;
      two_sum(g0[j].x, v.x, sr, er);
      two_sum(g0[j].y, v.y, si, ei);
      grid[(size_t(plane) * geo.ny + yy) * geo.nx + xx] =
        make_float2(sr + (er + v.z), si + (ei + v.w));
    } else {
      const int h = ry < tile ? ry * hx + (rx - tile) : tile * hx + (ry - tile) * RX + rx;
      hk[h] = make_double2(double(v.x) + double(v.z), double(v.y) + double(v.w));
    }
  }
}

// ONE: every support fits one step (sx, sy <= 8); the pipelined loop is
// then free of branches around vector memory loads, so each entry waits for
// its own CF taps only and not for the loads issued after them
template <int N, bool ONE>
__global__ __launch_bounds__(quad_block) void k_quad(
  Geometry geo, const GroupInfo* __restrict__ groups,
  const float2* __restrict__ taps, Mueller mu, const unsigned* __restrict__ off_table,
  int nbins, size_t block_cap, const Entry4<N>* __restrict__ entries,
  const unsigned* __restrict__ entry_vis, const VisRec<N>* __restrict__ recs,
  int hx, int hy, void* __restrict__ grid, double2* __restrict__ halo,
  unsigned* __restrict__ halo_set) {
  constexpr int RB = quad_round<N>();
  // quadrant regions, 16 bytes per pixel: FP32 (hi_re, hi_im, lo_re, lo_im)
  extern __shared__ float4 regions[];
  __shared__ QuadRec<N> list[RB];
  __shared__ unsigned seg_pref[quad_bins_max + 1];
  __shared__ unsigned seg_beg[quad_bins_max];
  __shared__ unsigned part[quad_block / wave + 1];
  // per round: entries per quadrant, then a gradient flag; two sets, used
  // by alternate rounds, so that a round needs no barrier before its counts
  __shared__ unsigned qcnt2[2][5];

  const int QX = quad_edge + hx;
  const int QY = quad_edge + hy;
  const int Q = QX * QY;
  const unsigned key = xcd_key(blockIdx.x, gridDim.x);  // one block per key
  const int nkeys = geo.ntx * geo.nty * geo.nr * geo.nc;
  for (int bb = threadIdx.x; bb < nbins; bb += quad_block) {
    const unsigned* tab = off_table + size_t(bb) * (nkeys + 1);
    const unsigned s0 = tab[key], s1 = tab[key + 1];
    seg_pref[bb] = s1 - s0;
    seg_beg[bb] = s0;
  }
  __syncthreads();
  const unsigned total = block_exclusive_scan(seg_pref, nbins, part);
  if (total == 0) {
    if (threadIdx.x == 0) halo_set[key] = 0u;
    return;
  }
  if (threadIdx.x == 0) seg_pref[nbins] = total;

  const int ntiles = geo.ntx * geo.nty;
  const unsigned plane = key / ntiles;
  const int t = int(key % ntiles);
  const int x0 = (t % geo.ntx) * tile;
  const int y0 = (t / geo.ntx) * tile;
  const int gpol = int(plane % geo.nr);

  for (int k = threadIdx.x; k < 4 * Q; k += quad_block)
    regions[k] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  if (threadIdx.x < 10) qcnt2[threadIdx.x / 5][threadIdx.x % 5] = 0u;
  // a single CF group is read once (the common case), not once per entry
  const bool one_group = geo.ngroups == 1;
  const GroupInfo group0 = groups[0];
  __syncthreads();

  const unsigned tid = threadIdx.x;
  const int w = int(tid) / wave;
  const int lane = int(tid) % wave;
  float4* reg32 = regions + size_t(w) * Q;

  auto locate = [&](unsigned q) {
    int lo = 0, hi = nbins;  // seg_pref[lo] <= q < seg_pref[hi]
    while (hi - lo > 1) {
      const int mid = (lo + hi) >> 1;
      if (seg_pref[mid] <= q) lo = mid; else hi = mid;
    }
    return unsigned(lo) * unsigned(block_cap) + seg_beg[lo] + (q - seg_pref[lo]);
  };
  // this thread's raw entry of the next round
  Entry4<N> ne{};
  unsigned nidx = 0;
  auto load_round = [&](unsigned r0) {
    if (tid < unsigned(RB) && r0 + tid < total) {
      nidx = locate(r0 + tid);
      ne = entries[nidx];
    }
  };
  load_round(0);

  // Lane layout of an entry's taps, step st of its steps: supports up to 8
  // wide put tap (X, Y) = (lane & 7, lane >> 3 + 8 st) on this lane, so that
  // the rows of one 16-lane LDS phase use distinct banks when the region row
  // is 8 modulo 16 pixels; wider supports take the taps t = 64 st + lane
  // contiguously, Y = t / sx from a float reciprocal (exact here: t < 4096
  // and (t + 1/2) / sx stays 1/(2 sx) > 0.0079 away from an integer, far
  // beyond the 4e-4 error of the product).
  const int lane_x8 = lane & 7;
  const int lane_y8 = lane >> 3;
  auto tap_xy = [&](int sx, float rsx, int st, int& X, int& Y) {
    if (ONE || sx <= 8) {
      X = lane_x8;
      Y = lane_y8 + st * (wave / 8);
    } else {
      const int tp = st * wave + lane;
      Y = int((float(tp) + 0.5f) * rsx);
      X = tp - Y * sx;
    }
  };
  auto num_steps = [](int sx, int sy) {
    return sx <= 8 ? (sy + wave / 8 - 1) / (wave / 8) : (sx * sy + wave - 1) / wave;
  };

  // an entry in flight: its prepared values and its first step's CF taps
  struct Slot {
    unsigned i, w0;
    const float2* slice[N];
    float2 c[N];
    unsigned long long lanes;  // ONE: lanes holding a tap of the support
  };
  // Byte offset of this lane's first-step tap in a slice, for the support
  // last seen (lanes outside the support read tap 0).  The first step's
  // taps load unconditionally: a load under a condition would need its
  // register cleared first, which waits for all loads in flight.
  int tap_key = -1;
  unsigned tap_off = 0;
  unsigned long long tap_lanes = 0ull;
  auto fetch = [&](unsigned i, Slot& s) {
    const QuadRec<N>& p = list[i];
    const unsigned w0v = p.w0;
    unsigned slv[N];
#pragma unroll
    for (int v = 0; v < N; ++v) slv[v] = p.slice[v];
    s.i = i;
    s.w0 = uniform_u(w0v);
    const int key_s = int((s.w0 >> 12) & 4095u);  // sx | sy << 6
    if (key_s != __builtin_amdgcn_readfirstlane(tap_key)) {
      tap_key = key_s;
      const int sx = key_s & 63, sy = key_s >> 6;
      int X, Y;
      tap_xy(sx, __builtin_amdgcn_rcpf(float(sx)), 0, X, Y);
      const bool ok = unsigned(X) < unsigned(sx) && unsigned(Y) < unsigned(sy);
      tap_off = unsigned(ok ? Y * sx + X : 0) * unsigned(sizeof(float2));
    }
#pragma unroll
    for (int v = 0; v < N; ++v) {
      s.slice[v] = taps + uniform_u(slv[v]);
      s.c[v] = *reinterpret_cast<const float2*>(
        reinterpret_cast<const char*>(s.slice[v]) + tap_off);
    }
  };
  auto add_tap = [&](int px, float h_r, float h_i, float l_r, float l_i) {
    quad_add<N>(reg32, px, h_r, h_i, l_r, l_i);
  };
  const int lane_px = lane_y8 * QX + lane_x8;
  // ONE: the raw record of an entry, read from LDS one pipeline stage before
  // its fetch, so that the fetch's scalar reads find it there
  struct RawRec {
    unsigned w0;
    unsigned sl[N];
    float2 g[N], gc[N];
  };
  auto read_rec = [&](unsigned i, RawRec& r) {
    const QuadRec<N>& p = list[i];
    r.w0 = p.w0;
#pragma unroll
    for (int v = 0; v < N; ++v) {
      r.sl[v] = p.slice[v];
      r.g[v] = p.g[v];
      r.gc[v] = p.gc[v];
    }
  };
  auto fetch_raw = [&](const RawRec& r, Slot& s) {
    s.w0 = uniform_u(r.w0);
    const int key_s = int((s.w0 >> 12) & 4095u);  // sx | sy << 6
    if (key_s != __builtin_amdgcn_readfirstlane(tap_key)) {
      tap_key = key_s;
      const int sx = key_s & 63, sy = key_s >> 6;
      const bool ok = lane_x8 < sx && lane_y8 < sy;
      tap_off = unsigned(ok ? lane_y8 * sx + lane_x8 : 0) * unsigned(sizeof(float2));
      tap_lanes = __ballot(ok);
    }
    s.lanes = (s.w0 & quad_skip) ? 0ull : tap_lanes;
#pragma unroll
    for (int v = 0; v < N; ++v) {
      s.slice[v] = taps + uniform_u(r.sl[v]);
      s.c[v] = *reinterpret_cast<const float2*>(
        reinterpret_cast<const char*>(s.slice[v]) + tap_off);
    }
  };
  // an entry of supports up to 8 x 8 (ONE): a single step, no branch when
  // the entry is live; `live` false (past the wave's entries) adds nothing
  auto process_one = [&](Slot& s, const RawRec& r, bool live) {
    const unsigned use = (s.w0 >> 24) & 15u;
    if (__builtin_amdgcn_inverse_ballot_w64(live ? s.lanes : 0ull)) {
      const int px = int(s.w0 & 4095u) + lane_px;
      if constexpr (N == 1) {
        const float2 c = s.c[0];
        const float2 g = r.g[0], gc = r.gc[0];
        const float pr = c.x * g.x - c.y * gc.x;
        const float pi = c.x * g.y + c.y * gc.y;
        add_tap(px, pr, pi, 0.0f, 0.0f);
      } else {
        double sr = 0.0, si = 0.0;
#pragma unroll
        for (int v = 0; v < N; ++v) {
          if (!((use >> v) & 1u)) continue;
          const float2 c = s.c[v];
          const float2 g = r.g[v], gc = r.gc[v];
          sr += double(c.x * g.x - c.y * gc.x);
          si += double(c.x * g.y + c.y * gc.y);
        }
        const float h_r = float(sr), h_i = float(si);
        add_tap(px, h_r, h_i, float(sr - double(h_r)), float(si - double(h_i)));
      }
    }
  };
  auto process = [&](Slot& s) {
    if ((s.w0 & quad_gradient) || !((s.w0 >> 24) & 15u)) return;
    const int base = int(s.w0 & 4095u);
    const int sx = int((s.w0 >> 12) & 63u);
    const int sy = int((s.w0 >> 18) & 63u);
    const unsigned use = (s.w0 >> 24) & 15u;
    const float rsx = __builtin_amdgcn_rcpf(float(sx));
    const int steps = num_steps(sx, sy);
    for (int st = 0; st < steps; ++st) {
      int X, Y;
      tap_xy(sx, rsx, st, X, Y);
      if (st > 0) {
        const bool okn = unsigned(X) < unsigned(sx) && unsigned(Y) < unsigned(sy);
        const int tp = okn ? Y * sx + X : 0;
#pragma unroll
        for (int v = 0; v < N; ++v) s.c[v] = s.slice[v][tp];
      }
      if (!(unsigned(X) < unsigned(sx) && unsigned(Y) < unsigned(sy))) continue;
      // FP32 products c * g (conjugated c for w > 0), their FP64 sum over
      // the polarizations
      const QuadRec<N>& p = list[s.i];
      if constexpr (N == 1) {
        const float2 c = s.c[0];
        const float2 g = p.g[0], gc = p.gc[0];
        const float pr = c.x * g.x - c.y * gc.x;
        const float pi = c.x * g.y + c.y * gc.y;
        add_tap(base + Y * QX + X, pr, pi, 0.0f, 0.0f);
      } else {
        double sr = 0.0, si = 0.0;
#pragma unroll
        for (int v = 0; v < N; ++v) {
          if (!((use >> v) & 1u)) continue;
          const float2 c = s.c[v];
          const float2 g = p.g[v], gc = p.gc[v];
          sr += double(c.x * g.x - c.y * gc.x);
          si += double(c.x * g.y + c.y * gc.y);
        }
        const float h_r = float(sr), h_i = float(si);
        add_tap(base + Y * QX + X, h_r, h_i, float(sr - double(h_r)), float(si - double(h_i)));
      }
    }
  };
  unsigned round = 0;
  for (unsigned r0 = 0; r0 < total; r0 += RB, ++round) {
    __builtin_amdgcn_s_setprio(0);
    const unsigned rn = min(unsigned(RB), total - r0);
    const bool mine = tid < rn;
    const Entry4<N> e = ne;
    const unsigned idx = nidx;
    unsigned* qcnt = qcnt2[round & 1u];
    unsigned* qgrad = qcnt + 4;
    QuadRec<N> rec{};
    int qd = -1;
    if (mine) {
      const unsigned pos = e.h.pos;
      const int lx = int(pos & 63u);
      const int ly = int((pos >> 6) & 63u);
      const int fxi = int((pos >> 12) & 255u);
      const int fyi = int((pos >> 20) & 255u);
      const bool pos_w = pos & e4_pos_w;
      const bool grad = pos & e4_gradient;
      const int grp = int(e.h.gpk & 0xffffu);
      const int cfcube = int(e.h.gpk >> 16);
      const GroupInfo g = one_group ? group0 : groups[grp];
      const int* row = pos_w ? mu.M[gpol] : mu.Mc[gpol];
      const long long slice_taps = (long long)g.sx * g.sy;
      const float imf = pos_w ? -1.0f : 1.0f;
      unsigned use = 0;
#pragma unroll
      for (int v = 0; v < N; ++v) {
        const int m = row[v];
        const bool u = m >= 0 && m < g.nmueller;
        use |= u ? 1u << v : 0u;
        rec.slice[v] = unsigned(
          g.tap_offset
          + (((long long)(cfcube * g.nmueller + (u ? m : 0)) * geo.nf + fxi) * geo.nf + fyi)
            * slice_taps);
        rec.g[v] = e.g[v];
        rec.gc[v] = make_float2(imf * e.g[v].y, imf * e.g[v].x);
      }
      const int qx = lx >= quad_edge ? 1 : 0;
      const int qy = ly >= quad_edge ? 1 : 0;
      qd = qx | (qy << 1);
      const unsigned base = unsigned((ly - qy * quad_edge) * QX + (lx - qx * quad_edge));
      rec.w0 = base | (unsigned(g.sx) << 12) | (unsigned(g.sy) << 18) | (use << 24)
        | (grad ? quad_gradient : 0u) | (pos_w ? quad_pos_w : 0u)
        | ((grad || use == 0u) ? quad_skip : 0u);
      if (grad) rec.gc[0].x = __uint_as_float(idx);
    }
    // counting sort by quadrant: one LDS atomic per wave and quadrant
    if (__ballot(mine && (rec.w0 & quad_gradient)) != 0ull && lane == 0)
      __hip_atomic_fetch_or(qgrad, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    unsigned slot = 0;
    const unsigned long long below = (1ull << lane) - 1ull;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const unsigned long long m = __ballot(qd == k);
      unsigned wbase = 0;
      if (lane == 0 && m != 0ull)
        wbase = __hip_atomic_fetch_add(qcnt + k, unsigned(__popcll(m)), __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_WORKGROUP);
      wbase = __shfl(wbase, 0, wave);
      if (qd == k) slot = wbase + unsigned(__popcll(m & below));
    }
    // (the list is free: every wave added the previous round's entries
    // before its counts above)
    __syncthreads();
    const unsigned c0 = qcnt[0], c1 = qcnt[1], c2 = qcnt[2], c3 = qcnt[3];
    const bool any_gradient = *qgrad != 0u;
    // the other set was last read in the previous round, before its list
    // barrier; cleared before this round's list barrier, it is ready for
    // the next round's counts
    if (tid < 5) qcnt2[(round + 1u) & 1u][tid] = 0u;
    if (mine) {
      const unsigned off = (qd > 0 ? c0 : 0u) + (qd > 1 ? c1 : 0u) + (qd > 2 ? c2 : 0u) + slot;
      list[off] = rec;
    }
    __syncthreads();
    // this wave's entries of the round, D in flight
    // (wave-uniform: kept in scalar registers, so that the loops below index
    // the list with scalar arithmetic)
    const unsigned beg = uniform_u((w > 0 ? c0 : 0u) + (w > 1 ? c1 : 0u) + (w > 2 ? c2 : 0u));
    const unsigned end = uniform_u(beg + (w == 0 ? c0 : w == 1 ? c1 : w == 2 ? c2 : c3));
    // issue priority by the share of the round: the wave with the most
    // entries sets the round's length, so it goes first on its SIMD (whose
    // other waves belong to other workgroups)
    {
      const unsigned cnt[4] = {c0, c1, c2, c3};
      const unsigned own = end - beg;
      int rank = 0;
#pragma unroll
      for (int k = 0; k < 4; ++k) rank += (cnt[k] < own || (cnt[k] == own && k < w)) ? 1 : 0;
      if (rank == 3)
        __builtin_amdgcn_s_setprio(3);
      else if (rank == 2)
        __builtin_amdgcn_s_setprio(2);
      else if (rank == 1)
        __builtin_amdgcn_s_setprio(1);
    }
    constexpr int D = quad_depth<N>();
    Slot sl[D];
    if constexpr (ONE) {
      // records read 2D entries ahead into 2D banks (entry beg + j in bank
      // j mod 2D, where it stays until it is added), CF taps loaded D
      // entries ahead; clamped reads past the end repeat the last entry and
      // are not added
      if (beg < end) {
        RawRec rr[2 * D];
#pragma unroll
        for (int b = 0; b < 2 * D; ++b) read_rec(min(beg + b, end - 1), rr[b]);
#pragma unroll
        for (int k = 0; k < D; ++k) fetch_raw(rr[k], sl[k]);
        // the next round's entries load while the first taps are in flight
        load_round(r0 + RB);
        unsigned i = beg;
        // every entry added and every read in range: no checks
        for (; i + 4 * D <= end; i += 2 * D) {
#pragma unroll
          for (int k = 0; k < 2 * D; ++k) {
            process_one(sl[k % D], rr[k], true);
            fetch_raw(rr[(k + D) % (2 * D)], sl[k % D]);
            read_rec(i + k + 2 * D, rr[k]);
          }
        }
        for (; i < end; i += 2 * D) {
#pragma unroll
          for (int k = 0; k < 2 * D; ++k) {
            process_one(sl[k % D], rr[k], i + k < end);
            fetch_raw(rr[(k + D) % (2 * D)], sl[k % D]);
            read_rec(min(i + k + 2 * D, end - 1), rr[k]);
          }
        }
      } else {
        load_round(r0 + RB);
      }
    } else {
      // the next round's entries load while this round's are added
      load_round(r0 + RB);
#pragma unroll
      for (int k = 0; k < D; ++k)
        if (beg + k < end) fetch(beg + k, sl[k]);
      for (unsigned i = beg; i < end; i += D) {
#pragma unroll
        for (int k = 0; k < D; ++k) {
          if (i + k < end) {
            process(sl[k]);
            if (i + k + D < end) fetch(i + k + D, sl[k]);
          }
        }
      }
    }
    if (any_gradient)
#pragma nounroll
      for (unsigned i = beg; i < end; ++i)
        if (uniform_u(list[i].w0) & quad_gradient)
          quad_gradient_entry<N>(list[i], __float_as_uint(list[i].gc[0].x), geo, taps, entries,
                                 entry_vis, recs, QX, reg32);
  }
  __builtin_amdgcn_s_setprio(0);
  __syncthreads();

  quad_flush_tile<quad_block, quad_flush>(regions, QX, QY, hx, hy, geo, plane, x0, y0,
                                         static_cast<float2*>(grid),
                                         halo + size_t(key) * halo_pixels(hx, hy));
  if (threadIdx.x == 0) halo_set[key] = 1u;
}

// ---- gridding: quadrant keys, two waves per quadrant, FP32 block sums ----
//
// k_sub grids FP32-compensated tiles for supports of at most 8 x 8 pixels
// from quadrant-keyed bins (k_bin with quad_keys: quadrant q of tile key t
// is bin key 4 t + q in every bin block).  The workgroup of a (plane, tile)
// has eight waves; waves q and q + 4 own quadrant q (the entries whose
// support corner lies in it) and take its entry list alternately, 64
// entries at a time: lane j locates, loads and prepares entry j of the
// batch (CF slices, region pixel, w-sign, flags).  The wave then adds the
// batch's entries one at a time, one lane per CF tap, to its own FP32
// region of (16 + hx) x (16 + hy) pixels in LDS by a plain read-modify-write
// (an entry's taps are distinct pixels, the region is the wave's own and a
// wave's LDS operations execute in order).  The batches run in rounds,
// in step across the workgroup (two halves of 32 slots, each followed by
// its entries with a phase gradient); after every sub_fold_rounds-th round
// and the last, every wave folds (between two barriers): wave q + 4 h holds
// the
// (hi, lo) pairs of half h of quadrant q's region pixels (lane l: pixels
// l + 64 k), adds both of the quadrant's regions' FP32 sums of those pixels
// to its pairs by two-sum and clears them.  So the products of a pixel are
// summed in FP32 over at most sub_fold entries of a region, those sums are
// accumulated with compensation, and a wave keeps the pairs of half a
// region in registers.  The flush sums the four quadrants' pairs over the
// tile and halo (quad_flush_tile), as k_quad does, so k_halo folds the
// halos unchanged.
constexpr int sub_waves = 8;
constexpr int sub_block = sub_waves * wave;
constexpr int sub_bins_max = 256;  // bin blocks per pass (LDS segment tables)
constexpr int sub_fold_rounds = 4; // rounds between folds
constexpr int sub_fold = 64 * sub_fold_rounds;  // entries of a region between folds
constexpr int sub_ring = 8;        // slots whose CF taps load ahead
constexpr int sub_flush = 3;       // tile + halo pixels per thread in the flush

// lane -> CF tap (X, Y) for supports up to 8 x 8: X = lane & 7, and the
// 8-lane groups take rows 0, 4, 1, 5, 2, 6, 3, 7, so that each 16 lanes
// read two rows 4 region rows apart
__device__ __forceinline__ int sub_lane_y(int lane) {
  const int g = lane >> 3;
  return (g >> 1) + 4 * (g & 1);
}

// PK (one polarization, CF of fewer than 2^21 taps): an entry's region
// pixel, w-sign and CF slice travel in one word, read once per slot
constexpr unsigned sub_pk_flip = 1u << 10;
constexpr int sub_pk_shift = 11;
constexpr size_t sub_pk_taps = size_t(1) << (32 - sub_pk_shift);

// exp(i phi) rounded to FP32 for an FP64 phase: phi is reduced to
// [-pi, pi] in FP64 (2 pi split into a 53-bit head and its tail, exact for
// the multiples used for |phi| < 2^20), then the FP32 sine and cosine of the
// remainder.  Within 2 ulp of fl32(cos phi), fl32(sin phi) (the reference's
// cphase in FP64, rounded) -- far inside the FP32-compensated grid's 1e-6.
__device__ __forceinline__ float2 screen_f32(double phi) {
  constexpr double two_pi_hi = 6.283185307179586;       // fl64(2 pi)
  constexpr double two_pi_lo = 2.4492935982947064e-16;  // 2 pi - fl64(2 pi)
  const double k = rint(phi * 0.15915494309189535);      // phi / (2 pi)
  const double r = (phi - k * two_pi_hi) - k * two_pi_lo;
  float sn, cs;
  __sincosf(float(r), &sn, &cs);
  return make_float2(cs, sn);
}

// an entry with a phase gradient: its CF taps times the phase screen
// exp(i phi(X, Y)) (FP64 phase as the reference's, screen_f32), added by
// one wave, one lane per tap, to the wave's FP32 region
template <int N>
__device__ __forceinline__ void sub_gradient_entry(
  unsigned w0, const unsigned* slice, const float2* g, unsigned ri, unsigned pos,
  const Geometry& geo, const float2* __restrict__ taps, const VisRec<N>* __restrict__ recs,
  int QX, float2* region) {
  const int lane = int(threadIdx.x % wave);
  const int base = int(w0 & 4095u);
  const int sx = int((w0 >> 12) & 63u);
  const int sy = int((w0 >> 18) & 63u);
  const unsigned use = (w0 >> 24) & 15u;
  const float imf = (w0 & quad_pos_w) ? -1.0f : 1.0f;
  const int fxi = int((pos >> 12) & 255u);
  const int fyi = int((pos >> 20) & 255u);
  const double gxr = recs[ri].grad[0];
  const double gyr = recs[ri].grad[1];
  const double phi0x = -gxr * ((sx / 2) * geo.os - (fxi - geo.F));
  const double phi0y = -gyr * ((sy / 2) * geo.os - (fyi - geo.F));
  const double dphix = gxr * geo.os;
  const double dphiy = gyr * geo.os;
  const int X = lane & 7, Y = sub_lane_y(lane);
  if (X >= sx || Y >= sy) return;
  const int tp = Y * sx + X;
  const float2 scr = screen_f32((phi0x + X * dphix) + (phi0y + Y * dphiy));
  double sr = 0.0, si = 0.0;
  float h_r = 0.0f, h_i = 0.0f;
  for (int v = 0; v < N; ++v) {
    if (!((use >> v) & 1u)) continue;
    float2 cc = taps[size_t(slice[v]) + size_t(tp)];
    cc.y *= imf;
    const float2 pr = cmul(cmul(cc, scr), g[v]);
    if constexpr (N == 1) {
      h_r = pr.x;
      h_i = pr.y;
    }
    sr += double(pr.x);
    si += double(pr.y);
  }
  if constexpr (N > 1) {
    h_r = float(sr);
    h_i = float(si);
  }
  float2* a = region + base + Y * QX + X;
  float2 t = *a;
  t.x += h_r;
  t.y += h_i;
  *a = t;
}

// The entry lists of tile key tkey's four quadrant keys: seg_pref[q][b] =
// entries of quadrant q in the bin blocks before b (seg_pref[q][nbins] the
// quadrant's total), seg_beg[q][b] = its first entry within block b.  GB
// (global bins, grids of more keys than one workgroup's LDS histogram
// holds): off_table is the exclusive scan of the quadrant keys' entry
// counts, each key one contiguous range from gb_beg[q] (nbins = 1).  Ends
// with a barrier; the block has at least four waves.
template <int BLOCK, bool GB>
__device__ __forceinline__ void sub_tables(
  const unsigned* __restrict__ off_table, int nbins, unsigned ntk, unsigned tkey,
  unsigned (*seg_pref)[sub_bins_max + 1], unsigned short (*seg_beg)[sub_bins_max],
  unsigned* gb_beg) {
  const unsigned tid = threadIdx.x;
  if constexpr (GB) {
    if (tid < 4) {
      const unsigned s0 = off_table[4 * size_t(tkey) + tid];
      const unsigned s1 = off_table[4 * size_t(tkey) + tid + 1];
      gb_beg[tid] = s0;
      seg_pref[tid][0] = 0u;
      seg_pref[tid][1] = s1 - s0;
    }
    __syncthreads();
  } else {
    // segment tables of the four quadrant keys, then their exclusive scans
    const size_t row = size_t(4) * ntk + 1;  // offset-table row
    for (int i = int(tid); i < 4 * nbins; i += BLOCK) {
      const int bb = i >> 2, qq = i & 3;
      const unsigned* tab = off_table + size_t(bb) * row + 4 * size_t(tkey) + qq;
      const unsigned s0 = tab[0], s1 = tab[1];
      seg_pref[qq][bb] = s1 - s0;
      seg_beg[qq][bb] = (unsigned short)s0;
    }
    __syncthreads();
    const int w = int(tid) / wave;
    const int lane = int(tid) % wave;
    if (w < 4) {
      unsigned* a = seg_pref[w];
      unsigned v[4];
      unsigned sum = 0;
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const int idx = lane * 4 + k;
        v[k] = idx < nbins ? a[idx] : 0u;
        sum += v[k];
      }
      unsigned incl = sum;
#pragma unroll
      for (int d = 1; d < wave; d <<= 1) {
        const unsigned t = __shfl_up(incl, d, wave);
        if (lane >= d) incl += t;
      }
      unsigned run = incl - sum;
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const int idx = lane * 4 + k;
        if (idx < nbins) a[idx] = run;
        run += v[k];
      }
      if (lane == wave - 1) a[nbins] = incl;
    }
    __syncthreads();
  }
}

// entry i of quadrant q's list (sub_tables): its index in the entry array
template <bool GB>
__device__ __forceinline__ unsigned sub_locate(
  unsigned i, int q, int nbins, size_t block_cap, const unsigned (*seg_pref)[sub_bins_max + 1],
  const unsigned short (*seg_beg)[sub_bins_max], const unsigned* gb_beg) {
  if constexpr (GB) return gb_beg[q] + i;
  int lo = 0, hi = nbins;  // seg_pref[q][lo] <= i < seg_pref[q][hi]
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (seg_pref[q][mid] <= i) lo = mid; else hi = mid;
  }
  return unsigned(lo) * unsigned(block_cap) + seg_beg[q][lo] + (i - seg_pref[q][lo]);
}

// PX: region pixels per lane; each wave's region is padded to PX * 64
// pixels (>= (16 + hx) * (16 + hy)), so a lane folds pixels lane + 64 k with
// no bounds test.  UNI: every CF group has the same support, whose lane mask
// then holds for the whole call.  PK: see sub_pk_flip.  GB: global bins
// (sub_tables).
template <int N, int PX, bool UNI, bool PK, bool GB>
__global__ __launch_bounds__(sub_block) __attribute__((amdgpu_waves_per_eu(N == 1 ? 8 : 4)))
void k_sub(
  Geometry geo, const GroupInfo* __restrict__ groups, const float2* __restrict__ taps,
  unsigned taps_bytes, Mueller mu, const unsigned* __restrict__ off_table, int nbins,
  size_t block_cap,
  const Entry4<N>* __restrict__ entries, const unsigned* __restrict__ entry_vis,
  const VisRec<N>* __restrict__ recs, int hx, int hy, float2* __restrict__ grid,
  double2* __restrict__ halo, unsigned* __restrict__ halo_set) {
  // eight FP32 regions of QP = PX * 64 pixels (float2); in the flush the
  // same bytes hold the four quadrants' (hi, lo) pairs (float4, Q each)
  extern __shared__ float2 sregions[];
  __shared__ unsigned seg_pref[4][sub_bins_max + 1];
  __shared__ unsigned short seg_beg[4][sub_bins_max];
  __shared__ unsigned gb_beg[4];  // GB: first entry of each quadrant key
  constexpr int QP = PX * wave;

  const int QX = quad_edge + hx;
  const int QY = quad_edge + hy;
  const int Q = QX * QY;
  const unsigned ntk = unsigned(geo.ntx) * geo.nty * geo.nr * geo.nc;
  const unsigned tkey = xcd_key(blockIdx.x, gridDim.x);  // one block per tile key
  const unsigned tid = threadIdx.x;
  const int w = int(tid) / wave;
  const int lane = int(tid) % wave;
  const int q = w & 3;
  const int half = w >> 2;

  sub_tables<sub_block, GB>(off_table, nbins, ntk, tkey, seg_pref, seg_beg, gb_beg);
  if (seg_pref[0][nbins] + seg_pref[1][nbins] + seg_pref[2][nbins] + seg_pref[3][nbins] == 0u) {
    if (tid == 0) halo_set[tkey] = 0u;
    return;
  }

  const unsigned ntiles = unsigned(geo.ntx) * geo.nty;
  const unsigned plane = tkey / ntiles;
  const int t = int(tkey % ntiles);
  const int x0 = (t % geo.ntx) * tile;
  const int y0 = (t / geo.ntx) * tile;
  const int gpol = int(plane % geo.nr);
  const int qx = q & 1, qy = q >> 1;

  float2* region = sregions + size_t(w) * QP;
#pragma unroll
  for (int k = 0; k < PX; ++k) region[lane + k * wave] = make_float2(0.0f, 0.0f);
  // the (hi, lo) pairs of this wave's share of its quadrant's pixels: wave
  // half 0 holds pixels lane + 64 k for k < PXA, half 1 the rest
  constexpr int PXA = (PX + 1) / 2;
  const int k0 = half ? PXA : 0;
  const int nk = half ? PX - PXA : PXA;
  float4 acc[PXA];
#pragma unroll
  for (int k = 0; k < PXA; ++k) acc[k] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  const float2* reg_a = sregions + size_t(q) * QP;      // the quadrant's two regions
  const float2* reg_b = sregions + size_t(q + 4) * QP;
  // the FP32 sums of both regions' pixels of this wave's share into its
  // (hi, lo) pairs, by two-sum, clearing the pixels (all waves at once,
  // between barriers)
  auto fold = [&]() {
#pragma unroll
    for (int k = 0; k < PXA; ++k) {
      if (PX % 2 == 1 && k >= nk) continue;
      const int px = lane + (k0 + k) * wave;
      const float2 pa = reg_a[px];
      const float2 pb = reg_b[px];
      const_cast<float2*>(reg_a)[px] = make_float2(0.0f, 0.0f);
      const_cast<float2*>(reg_b)[px] = make_float2(0.0f, 0.0f);
      float sum, err;
      two_sum(acc[k].x, pa.x, sum, err);
      acc[k].x = sum;
      acc[k].z += err;
      two_sum(acc[k].y, pa.y, sum, err);
      acc[k].y = sum;
      acc[k].w += err;
      two_sum(acc[k].x, pb.x, sum, err);
      acc[k].x = sum;
      acc[k].z += err;
      two_sum(acc[k].y, pb.y, sum, err);
      acc[k].y = sum;
      acc[k].w += err;
    }
  };

  const bool one_group = geo.ngroups == 1;
  const GroupInfo group0 = groups[0];
  // this plane's Mueller rows for both signs of w
  int row_m[N], row_mc[N];
#pragma unroll
  for (int v = 0; v < N; ++v) {
    row_m[v] = mu.M[gpol][v];
    row_mc[v] = mu.Mc[gpol][v];
  }
  // the quadrant's list, taken by its two waves alternately 64 entries at
  // a time (fewer partly filled batches than an even split, whose per-batch
  // costs outweigh its balance)
  const unsigned total = seg_pref[q][nbins];
  auto locate = [&](unsigned i) {
    return sub_locate<GB>(i, q, nbins, block_cap, seg_pref, seg_beg, gb_beg);
  };
  const int lane_x = lane & 7;
  const int lane_y = sub_lane_y(lane);
  // LDS byte offset of this lane's tap pixel in the wave's region, relative
  // to an entry's support corner
  const unsigned lane_addr =
    unsigned((size_t(w) * QP + size_t(lane_y * QX + lane_x)) * sizeof(float2));
  // CF taps through a buffer resource: the slice's byte offset goes in a
  // scalar register and the lane's tap offset in a 32-bit vector register;
  // reads past taps_bytes return 0
  const __amdgpu_buffer_rsrc_t tap_rsrc = __builtin_amdgcn_make_buffer_rsrc(
    const_cast<float2*>(taps), 0, int(taps_bytes), 0x00020000);
  // lanes holding a tap of a support (sx, sy), and their tap byte offset
  auto support_lanes = [&](int sx, int sy, unsigned& off) {
    const bool ok = lane_x < sx && lane_y < sy;
    off = unsigned(ok ? lane_y * sx + lane_x : 0) * unsigned(sizeof(float2));
    return ok;
  };
  unsigned tap_off = 0;
  const bool uni_ok = UNI ? support_lanes(group0.sx, group0.sy, tap_off) : true;
  int tap_key = -1;
  unsigned long long tap_lanes = 0ull;

  // rounds: in round r, wave half h takes batch 2 r + h of its quadrant's
  // list; all eight waves run every round (the folds are shared)
  const unsigned tq[4] = {seg_pref[0][nbins], seg_pref[1][nbins], seg_pref[2][nbins],
                          seg_pref[3][nbins]};
  const unsigned tmax = max(max(tq[0], tq[1]), max(tq[2], tq[3]));
  const unsigned nrounds = (tmax + 2u * wave - 1u) / (2u * wave);
  for (unsigned rd = 0; rd < nrounds; ++rd) {
    const unsigned b0 = (2u * rd + unsigned(half)) * wave;
    // entries in the round's largest batch over the workgroup (uniform)
    unsigned rmax = 0;
#pragma unroll
    for (int qq = 0; qq < 4; ++qq)
#pragma unroll
      for (unsigned hh = 0; hh < 2; ++hh) {
        const unsigned bb = (2u * rd + hh) * wave;
        rmax = max(rmax, tq[qq] > bb ? min(unsigned(wave), tq[qq] - bb) : 0u);
      }
    // lane j: entry b0 + j of the quadrant's list, prepared.  Slots that
    // add nothing here (past the list, phase gradient, no CF index) keep
    // zero values and the region's first pixel, and read a valid CF slice.
    const unsigned i = b0 + unsigned(lane);
    const bool valid = i < total;
    unsigned ei = 0;
    Entry4<N> e{};
    if (valid) {
      ei = locate(i);
      e = entries[ei];
    }
    // p_addr: the entry's LDS byte offset from the region origin, with the
    // w-sign flip (bit 31: conj(c) for w > 0) -- 0 for slots that add
    // nothing.  PK: the region pixel | sub_pk_flip | slice << sub_pk_shift
    unsigned p_addr = 0, p_key = 0;
    unsigned p_sl[N];
    float p_gx[N], p_gy[N];
#pragma unroll
    for (int v = 0; v < N; ++v) {
      p_sl[v] = 0u;
      p_gx[v] = p_gy[v] = 0.0f;
    }
    bool grad = false, normal = false;
    if (valid) {
      const unsigned pos = e.h.pos;
      const int lx = int(pos & 63u);
      const int ly = int((pos >> 6) & 63u);
      const int fxi = int((pos >> 12) & 255u);
      const int fyi = int((pos >> 20) & 255u);
      const bool pos_w = pos & e4_pos_w;
      grad = pos & e4_gradient;
      const int grp = int(e.h.gpk & 0xffffu);
      const int cfcube = int(e.h.gpk >> 16);
      const GroupInfo g = one_group ? group0 : groups[grp];
      const long long slice_taps = (long long)g.sx * g.sy;
      unsigned use = 0;
#pragma unroll
      for (int v = 0; v < N; ++v) {
        const int m = pos_w ? row_m[v] : row_mc[v];
        const bool u = m >= 0 && m < g.nmueller;
        use |= u ? 1u << v : 0u;
        p_sl[v] = unsigned(
          g.tap_offset
          + (((long long)(cfcube * g.nmueller + (u ? m : 0)) * geo.nf + fxi) * geo.nf + fyi)
            * slice_taps);
        if (u && !grad) {
          p_gx[v] = e.g[v].x;
          p_gy[v] = e.g[v].y;
        }
      }
      const unsigned base = unsigned((ly - qy * quad_edge) * QX + (lx - qx * quad_edge));
      p_key = unsigned(g.sx) | (unsigned(g.sy) << 6);
      normal = !grad && use != 0u;
      if constexpr (PK) {
        p_addr = (p_sl[0] << sub_pk_shift) | (normal ? base | (pos_w ? sub_pk_flip : 0u) : 0u);
      } else {
        if (normal) p_addr = base * unsigned(sizeof(float2)) | (pos_w ? 0x80000000u : 0u);
      }
      grad = grad && use != 0u;
    }
    const unsigned long long nmask = __ballot(normal);
    const unsigned long long gtodo = __ballot(valid && grad);
    // slots up to the last entry added here
    const int nslots = nmask ? 64 - __builtin_clzll(nmask) : 0;

    // CF taps of slot j: loaded sub_ring slots ahead into ring[j % sub_ring]
    // (the fetches run on the whole wave: a lane's prepared values are read
    // from it while it is active); the slot's other values are read from
    // its lane when it is added
    float2 ring[sub_ring][N];
    unsigned ring_pk[sub_ring];               // PK: the slot's word
    unsigned long long ring_lanes[sub_ring];  // !UNI: lanes of the slot's support
    auto fetch = [&](int j) {
      const int r = j % sub_ring;
      if constexpr (!UNI) {
        const int key_s = __builtin_amdgcn_readlane(int(p_key), j);
        if (key_s != tap_key) {
          tap_key = key_s;
          tap_lanes = __ballot(support_lanes(key_s & 63, key_s >> 6, tap_off));
        }
        ring_lanes[r] = tap_lanes;
      }
      if constexpr (PK) {
        const unsigned pk = unsigned(__builtin_amdgcn_readlane(int(p_addr), j));
        ring_pk[r] = pk;
        const v2i_t raw = __builtin_amdgcn_raw_buffer_load_b64(
          tap_rsrc, int(tap_off), int((pk >> sub_pk_shift) * unsigned(sizeof(float2))), 0);
        ring[r][0] = make_float2(__int_as_float(raw.x), __int_as_float(raw.y));
      } else {
#pragma unroll
        for (int v = 0; v < N; ++v) {
          const int slv = __builtin_amdgcn_readlane(int(p_sl[v]), j);
          const v2i_t raw = __builtin_amdgcn_raw_buffer_load_b64(
            tap_rsrc, int(tap_off), slv * int(sizeof(float2)), 0);
          ring[r][v] = make_float2(__int_as_float(raw.x), __int_as_float(raw.y));
        }
      }
    };
    auto add = [&](int j) {
      const int r = j % sub_ring;
      unsigned flip, addr;
      if constexpr (PK) {
        const unsigned pk = ring_pk[r];
        flip = (pk << (31 - 10)) & 0x80000000u;
        addr = (pk & (sub_pk_flip - 1u)) * unsigned(sizeof(float2));
      } else {
        const unsigned aw = unsigned(__builtin_amdgcn_readlane(int(p_addr), j));
        flip = aw & 0x80000000u;
        addr = aw & 0x7fffffffu;
      }
      // UNI: the slot loops run on the support's lanes only
      bool on;
      if constexpr (UNI) on = true; else on = __builtin_amdgcn_inverse_ballot_w64(ring_lanes[r]);
      float h_r = 0.0f, h_i = 0.0f;
      double sr = 0.0, si = 0.0;
#pragma unroll
      for (int v = 0; v < N; ++v) {
        const unsigned ux = unsigned(__builtin_amdgcn_readlane(__float_as_int(p_gx[v]), j));
        const unsigned uy = unsigned(__builtin_amdgcn_readlane(__float_as_int(p_gy[v]), j));
        const float gx = __uint_as_float(ux), gy = __uint_as_float(uy);
        // conj(c) g for w > 0: c.y g.y and c.y g.x enter with the sign flipped
        const float gcx = __uint_as_float(uy ^ flip), gcy = __uint_as_float(ux ^ flip);
        // the FP32 complex product, each part as one product and one fused
        // multiply-add
        const float2 c = ring[r][v];
        const float pr = __builtin_fmaf(c.x, gx, -(c.y * gcx));
        const float pi = __builtin_fmaf(c.x, gy, c.y * gcy);
        if constexpr (N == 1) {
          h_r = pr;
          h_i = pi;
        } else {
          sr += double(pr);
          si += double(pi);
        }
      }
      if constexpr (N > 1) {
        h_r = float(sr);
        h_i = float(si);
      }
      if (on) {
        float2* a =
          reinterpret_cast<float2*>(reinterpret_cast<char*>(sregions) + addr + lane_addr);
        float2 tv = *a;
        tv.x += h_r;
        tv.y += h_i;
        *a = tv;
      }
    };
    // two halves of 32 slots, each followed by the half's entries with a
    // phase gradient; the shared fold after the last half of every
    // sub_fold_rounds-th round and of the last, so at most sub_fold entries
    // reach a region between folds.  A half without entries in any wave of
    // the workgroup is skipped
#pragma unroll
    for (int hb = 0; hb < 2; ++hb) {
      if (unsigned(hb * 32) >= rmax) continue;
      if (hb * 32 < nslots && uni_ok) {
#pragma unroll
        for (int j = 0; j < sub_ring; ++j) fetch(hb * 32 + j);
#pragma unroll
        for (int j = 0; j < 32; ++j) {
          add(hb * 32 + j);
          if (j + sub_ring < 32) fetch(hb * 32 + j + sub_ring);
        }
      }
      unsigned long long gl = gtodo & (hb ? 0xffffffff00000000ull : 0xffffffffull);
      const bool last_half = hb == 1 || 32u >= rmax;
      while (gl) {
        const int j = __builtin_ctzll(gl);
        gl &= gl - 1ull;
        // the entry again (a scalar load), so that the batch's raw entries
        // need not stay in registers through the loop above
        const unsigned ej = unsigned(__builtin_amdgcn_readlane(int(ei), j));
        const Entry4<N> ge = entries[ej];
        const unsigned gpos = ge.h.pos;
        const bool gpos_w = gpos & e4_pos_w;
        const GroupInfo gg = one_group ? group0 : groups[ge.h.gpk & 0xffffu];
        unsigned guse = 0;
        unsigned sl[N];
        float2 gv[N];
#pragma unroll
        for (int v = 0; v < N; ++v) {
          const int m = gpos_w ? row_m[v] : row_mc[v];
          guse |= (m >= 0 && m < gg.nmueller) ? 1u << v : 0u;
          sl[v] = PK ? unsigned(__builtin_amdgcn_readlane(int(p_addr), j)) >> sub_pk_shift
                     : unsigned(__builtin_amdgcn_readlane(int(p_sl[v]), j));
          gv[v] = ge.g[v];
        }
        const unsigned gbase = ((gpos >> 6) & 63u) - unsigned(qy * quad_edge);
        const unsigned w0 = (gbase * unsigned(QX) + ((gpos & 63u) - unsigned(qx * quad_edge)))
          | (unsigned(gg.sx) << 12) | (unsigned(gg.sy) << 18) | (guse << 24)
          | (gpos_w ? quad_pos_w : 0u);
        sub_gradient_entry<N>(w0, sl, gv, entry_vis[ej], ge.h.pos, geo, taps, recs, QX, region);
      }
      if (last_half && ((rd + 1u) % unsigned(sub_fold_rounds) == 0u || rd + 1u == nrounds)) {
        __syncthreads();
        fold();
        __syncthreads();
      }
    }
  }

  // each wave's pairs into LDS (the regions are no longer read), then flush
  float4* pairs = reinterpret_cast<float4*>(sregions);
#pragma unroll
  for (int k = 0; k < PXA; ++k) {
    if (PX % 2 == 1 && k >= nk) continue;
    const int px = lane + (k0 + k) * wave;
    if (px < Q) pairs[q * Q + px] = acc[k];
  }
  __syncthreads();
  quad_flush_tile<sub_block, sub_flush>(pairs, QX, QY, hx, hy, geo, plane, x0, y0, grid,
                                        halo + size_t(tkey) * halo_pixels(hx, hy));
  if (tid == 0) halo_set[tkey] = 1u;
}

// ---- degridding --------------------------------------------------------

// The wave-per-record degridder's record order: key = the record's CF slice
// (group, CF cube, fine offsets: grp * stride + (cube * nf + fxi) * nf +
// fyi), kmax for records it skips (no CF, outside the grid).  Sorted by key,
// the records that read the same 16 Mueller slices run together and find
// them in L2 (a record's taps of all Mueller elements are read once per
// slice set and batch rather than once per record).
template <int N>
__global__ void k_degrid_keys(
  size_t n, Geometry geo, const GroupInfo* __restrict__ groups,
  const VisRec<N>* __restrict__ recs, unsigned stride, unsigned kmax,
  unsigned* __restrict__ keys, unsigned* __restrict__ idx) {
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const VisRec<N>& r = recs[i];
  const unsigned grp = r.cf_index[1];
  const unsigned cfcube = r.cf_index[0];
  const unsigned gcube = r.grid_cube;
  unsigned key = kmax;
  if (grp < unsigned(geo.ngroups) && gcube < unsigned(geo.nc)) {
    const GroupInfo g = groups[grp];
    if (cfcube < unsigned(g.ncube)) {
      const Coord c = vis_coord(r, geo, g);
      if (c.inside)
        key = grp * stride
          + (cfcube * unsigned(geo.nf) + unsigned(c.fx + geo.F)) * unsigned(geo.nf)
          + unsigned(c.fy + geo.F);
    }
  }
  keys[i] = key;
  idx[i] = unsigned(i);
}

// order: the records' processing order (k_degrid_keys, sorted), or null
template <int N>
__global__ __launch_bounds__(block) void k_degrid(
  size_t n, Geometry geo, const GroupInfo* __restrict__ groups,
  const float2* __restrict__ taps, const double2* __restrict__ wsum, Mueller mu,
  const double2* __restrict__ model, int do_grid, VisRec<N>* __restrict__ recs,
  float2* __restrict__ gvis, double2* __restrict__ pred64,
  const unsigned* __restrict__ order) {

  const size_t wi = blockIdx.x * (size_t)(block / wave) + threadIdx.x / wave;
  const int lane = threadIdx.x % wave;
  if (wi >= n) return;
  const size_t i = order ? size_t(order[wi]) : wi;
  VisRec<N>& r = recs[i];
  const unsigned grp = r.cf_index[1];
  const unsigned cfcube = r.cf_index[0];
  const unsigned gcube = r.grid_cube;
  if (!(grp < unsigned(geo.ngroups) && gcube < unsigned(geo.nc))) return;
  if (!(cfcube < unsigned(groups[grp].ncube))) return;
  const GroupInfo g = groups[grp];
  const Coord c = vis_coord(r, geo, g);
  if (!c.inside) return;  // values left unchanged (src/hpg_core.hpp:1136-1169)
  const bool pos_w = r.uvw[2] > 0;
  const double2 phasor = cphase(r.phase);
  const int fxi = c.fx + geo.F;
  const int fyi = c.fy + geo.F;

  float2 pred[N];
  double2 pred_d[N];
  for (int v = 0; v < N; ++v) {
    pred[v] = make_float2(0.0f, 0.0f);
    pred_d[v] = make_double2(0.0, 0.0);
  }
  if (model) {
    const float imf = pos_w ? 1.0f : -1.0f;
    const bool grad = r.grad[0] != 0.0 || r.grad[1] != 0.0;
    const double phi0x = -r.grad[0] * ((g.sx / 2) * geo.os - c.fx);
    const double phi0y = -r.grad[1] * ((g.sy / 2) * geo.os - c.fy);
    const double dphix = r.grad[0] * geo.os;
    const double dphiy = r.grad[1] * geo.os;
    const int ntap = g.sx * g.sy;
    const unsigned mg = magic(unsigned(g.sx));
    const long long slice_taps = ntap;
    double2 a[N];
    for (int v = 0; v < N; ++v) a[v] = make_double2(0.0, 0.0);
    for (int gpol = 0; gpol < geo.nr; ++gpol) {
      const int* row = pos_w ? mu.Mc[gpol] : mu.M[gpol];
      const float2* slice[N];
      bool use[N];
      for (int v = 0; v < N; ++v) {
        const int m = row[v];
        use[v] = m >= 0 && m < g.nmueller;
        slice[v] = taps + g.tap_offset
          + (((long long)(cfcube * g.nmueller + (use[v] ? m : 0)) * geo.nf + fxi)
               * geo.nf + fyi) * slice_taps;
      }
      const double2* mplane =
        model + (size_t(gcube) * geo.nr + gpol) * size_t(geo.ny) * geo.nx;
      for (int q = lane; q < ntap; q += wave) {
        const int Y = int(div_magic(unsigned(q), unsigned(g.sx), mg));
        const int X = q - Y * g.sx;
        double2 mv = mplane[size_t(c.gy + Y) * geo.nx + (c.gx + X)];
        if (grad) {
          const double phX = phi0x + X * dphix;
          const double phY = phi0y + Y * dphiy;
          mv = cmul(mv, cphase(-phX - phY));
        }
        for (int v = 0; v < N; ++v) {
          if (!use[v]) continue;
          float2 cf = slice[v][q];
          cf.y *= imf;
          const double2 p = cmul(cf, mv);
          a[v].x += p.x;
          a[v].y += p.y;
        }
      }
    }
    double2 W[N];
    degrid_norms<N>(geo, g, wsum, mu, pos_w, cfcube, fxi, fyi, W);
    const double2 conj_phasor = make_double2(phasor.x, -phasor.y);
    for (int v = 0; v < N; ++v) {
      const double2 s = make_double2(wave_sum(a[v].x), wave_sum(a[v].y));
      const double2 d = (W[v].x != 0.0 || W[v].y != 0.0) ? W[v] : make_double2(1.0, 0.0);
      pred_d[v] = cmul(kdiv(s, d), conj_phasor);
      pred[v] = narrow(pred_d[v]);
    }
  }
  if (pred64) {
    // forward operator of the differentiable form: FP64 predictions
    if (lane == 0)
      for (int v = 0; v < N; ++v) pred64[i * N + v] = pred_d[v];
    return;
  }
  if (lane == 0) {
    for (int v = 0; v < N; ++v) {
      if (do_grid) {
        float2 val = r.vis[v];
        val.x -= pred[v].x;
        val.y -= pred[v].y;
        r.vis[v] = val;
        gvis[i * N + v] = prephase(val, phasor, r.w[v]);
      } else {
        r.vis[v] = pred[v];
      }
    }
  }
}

// One lane per visibility (supports up to degrid_lane_taps taps): the lane
// walks the footprint in chunks of halo_chunk taps, issuing the chunk's
// model-pixel and CF-tap loads together, and accumulates in FP64 registers;
// no cross-lane reduction.  Same per-tap arithmetic as k_degrid.
constexpr int degrid_lane_taps = 1024;

template <int N>
__global__ __launch_bounds__(block) void k_degrid_lane(
  size_t n, Geometry geo, const GroupInfo* __restrict__ groups,
  const float2* __restrict__ taps, const double2* __restrict__ wsum, Mueller mu,
  const double2* __restrict__ model, int do_grid, const VisRec<N>* recs_in,
  VisRec<N>* recs, float2* __restrict__ gvis, double2* __restrict__ pred64) {
  // recs_in and recs may be the same records: each lane reads its record's
  // fields before writing them back

  const size_t i = blockIdx.x * (size_t)block + threadIdx.x;
  if (i >= n) return;
  const VisRec<N>& r = recs_in[i];
  const unsigned grp = r.cf_index[1];
  const unsigned cfcube = r.cf_index[0];
  const unsigned gcube = r.grid_cube;
  if (!(grp < unsigned(geo.ngroups) && gcube < unsigned(geo.nc))) return;
  const GroupInfo g = groups[grp];
  if (!(cfcube < unsigned(g.ncube))) return;
  const Coord c = vis_coord(r, geo, g);
  if (!c.inside) return;  // values left unchanged (src/hpg_core.hpp:1136-1169)
  const bool pos_w = r.uvw[2] > 0;
  const double2 phasor = cphase(r.phase);
  const int fxi = c.fx + geo.F;
  const int fyi = c.fy + geo.F;

  float2 pred[N];
  double2 pred_d[N];
  for (int v = 0; v < N; ++v) {
    pred[v] = make_float2(0.0f, 0.0f);
    pred_d[v] = make_double2(0.0, 0.0);
  }
  if (model) {
    const float imf = pos_w ? 1.0f : -1.0f;
    const bool grad = r.grad[0] != 0.0 || r.grad[1] != 0.0;
    const double phi0x = -r.grad[0] * ((g.sx / 2) * geo.os - c.fx);
    const double phi0y = -r.grad[1] * ((g.sy / 2) * geo.os - c.fy);
    const double dphix = r.grad[0] * geo.os;
    const double dphiy = r.grad[1] * geo.os;
    const int sx = g.sx, sy = g.sy;
    const long long slice_taps = (long long)sx * sy;
    double2 a[N];
    for (int v = 0; v < N; ++v) a[v] = make_double2(0.0, 0.0);
    for (int gpol = 0; gpol < geo.nr; ++gpol) {
      const int* row = pos_w ? mu.Mc[gpol] : mu.M[gpol];
      const float2* slice[N];
      bool use[N];
      bool any = false;
      for (int v = 0; v < N; ++v) {
        const int m = row[v];
        use[v] = m >= 0 && m < g.nmueller;
        any |= use[v];
        slice[v] = taps + g.tap_offset
          + (((long long)(cfcube * g.nmueller + (use[v] ? m : 0)) * geo.nf + fxi) * geo.nf
             + fyi) * slice_taps;
      }
      if (!any) continue;
      const double2* mplane =
        model + (size_t(gcube) * geo.nr + gpol) * size_t(geo.ny) * geo.nx;
      for (int Y = 0; Y < sy; ++Y) {
        const double2* mrow = mplane + size_t(c.gy + Y) * geo.nx + c.gx;
        for (int X0 = 0; X0 < sx; X0 += halo_chunk) {
          double2 mv[halo_chunk];
          float2 cf[N][halo_chunk];
#pragma unroll
          for (int k = 0; k < halo_chunk; ++k) {
            const bool in = X0 + k < sx;
            mv[k] = in ? mrow[X0 + k] : make_double2(0.0, 0.0);
#pragma unroll
            for (int v = 0; v < N; ++v)
              cf[v][k] = (in && use[v]) ? slice[v][Y * sx + X0 + k] : make_float2(0.0f, 0.0f);
          }
#pragma unroll
          for (int k = 0; k < halo_chunk; ++k) {
            if (X0 + k >= sx) break;
            const int X = X0 + k;
            double2 m = mv[k];
            if (grad) {
              const double phX = phi0x + X * dphix;
              const double phY = phi0y + Y * dphiy;
              m = cmul(m, cphase(-phX - phY));
            }
#pragma unroll
            for (int v = 0; v < N; ++v) {
              if (!use[v]) continue;
              float2 cc = cf[v][k];
              cc.y *= imf;
              const double2 p = cmul(cc, m);
              a[v].x += p.x;
              a[v].y += p.y;
            }
          }
        }
      }
    }
    double2 W[N];
    degrid_norms<N>(geo, g, wsum, mu, pos_w, cfcube, fxi, fyi, W);
    const double2 conj_phasor = make_double2(phasor.x, -phasor.y);
    for (int v = 0; v < N; ++v) {
      const double2 d = (W[v].x != 0.0 || W[v].y != 0.0) ? W[v] : make_double2(1.0, 0.0);
      pred_d[v] = cmul(kdiv(a[v], d), conj_phasor);
      pred[v] = narrow(pred_d[v]);
    }
  }
  if (pred64) {
    for (int v = 0; v < N; ++v) pred64[i * N + v] = pred_d[v];
    return;
  }
  VisRec<N>& o = recs[i];
  for (int v = 0; v < N; ++v) {
    if (do_grid) {
      float2 val = r.vis[v];
      val.x -= pred[v].x;
      val.y -= pred[v].y;
      o.vis[v] = val;
      gvis[i * N + v] = prephase(val, phasor, r.w[v]);
    } else {
      o.vis[v] = pred[v];
    }
  }
}

// Binned degridding (supports below a tile): the entries of k_bin<N, true>
// in degrid mode (record index beside every entry) are processed one
// workgroup per (plane, tile) key.  The workgroup loads the model's tile +
// halo region of its plane into LDS with coalesced reads, then one lane per
// entry sums c' x model over the footprint from LDS (FP64), in slice-order
// chunks whose CF loads issue together.  The per-plane sums go to acc
// (stored when one plane per record, else added atomically); k_degrid_finish
// turns them into predictions.
// model value times the phase screen exp(i ph) (FP64, the reference's
// cphase).  Not inlined: the FP64 sine and cosine would otherwise raise the
// register count of k_degrid4's tap loop, which runs them only for records
// with a phase gradient.
__device__ __noinline__ double2 degrid_screen(double2 m, double ph) {
  return cmul(m, cphase(ph));
}
// cphase out of line, for the same reason
__device__ __noinline__ double2 cphase_call(double ph) { return cphase(ph); }

template <int N>
__global__ __launch_bounds__(halo_block) __attribute__((amdgpu_waves_per_eu(N == 1 ? 4 : 1))) void k_degrid4(
  Geometry geo, const GroupInfo* __restrict__ groups,
  const float2* __restrict__ taps, Mueller mu, const unsigned* __restrict__ off_table,
  int nbins, size_t block_cap, const Entry4<N>* __restrict__ entries,
  const unsigned* __restrict__ entry_vis, const VisRec<N>* recs,
  int hx, int hy, const double2* __restrict__ model, double2* __restrict__ acc,
  int one_plane, const double2* __restrict__ wsum, int do_grid, VisRec<N>* out_recs,
  float2* __restrict__ gvis, double2* __restrict__ pred64) {
  // one_plane: each record has exactly one entry, so the lane finishes the
  // prediction itself (the end of degrid_vis and the degridding branch of
  // VisibilityGridder::kernel, src/hpg_core.hpp:643-682, 1136-1169): FP64
  // predictions to pred64, or values written into out_recs when given
  // (residuals, and gvis, when do_grid); recs and out_recs may be the same
  // records, each lane reading its record's fields before writing them.
  // Otherwise the per-plane sums are added into acc for k_degrid_finish.

  extern __shared__ double smem[];
  const int RX = tile + hx;
  const int RY = tile + hy;
  const int R = RX * RY;
  double2* mreg = reinterpret_cast<double2*>(smem);
  unsigned* seg_pref = reinterpret_cast<unsigned*>(smem + 2 * R);
  unsigned* seg_beg = seg_pref + nbins + 1;
  unsigned* part = seg_beg + nbins;

  const unsigned key = xcd_key(blockIdx.x, gridDim.x);  // one block per key
  const int nkeys = geo.ntx * geo.nty * geo.nr * geo.nc;
  for (int bb = threadIdx.x; bb < nbins; bb += halo_block) {
    const unsigned* tab = off_table + size_t(bb) * (nkeys + 1);
    const unsigned s0 = tab[key], s1 = tab[key + 1];
    seg_pref[bb] = s1 - s0;
    seg_beg[bb] = s0;
  }
  __syncthreads();
  const unsigned total = block_exclusive_scan(seg_pref, nbins, part);
  if (total == 0) return;
  if (threadIdx.x == 0) seg_pref[nbins] = total;

  const int ntiles = geo.ntx * geo.nty;
  const unsigned plane = key / ntiles;
  const int t = int(key % ntiles);
  const int x0 = (t % geo.ntx) * tile;
  const int y0 = (t / geo.ntx) * tile;
  const int gpol = int(plane % geo.nr);
  const double2* mplane = model + size_t(plane) * geo.ny * size_t(geo.nx);
  for (int k = threadIdx.x; k < R; k += halo_block) {
    const int ry = k / RX;
    const int rx = k - ry * RX;
    const int xx = x0 + rx, yy = y0 + ry;
    mreg[k] = (xx < geo.nx && yy < geo.ny) ? mplane[size_t(yy) * geo.nx + xx]
                                           : make_double2(0.0, 0.0);
  }
  __syncthreads();

  const int w = threadIdx.x / wave;
  const int lane = threadIdx.x % wave;
  constexpr int nw = halo_block / wave;
  for (unsigned base = w * wave; base < total; base += nw * wave) {
    // issue priority by the passes this wave has left: near the end the
    // waves of the keys with the most entries, which set the kernel's
    // length, go first on their SIMD
    {
      const unsigned left = (total - base + unsigned(nw * wave) - 1u) / unsigned(nw * wave);
      if (left >= 3u)
        __builtin_amdgcn_s_setprio(3);
      else if (left == 2u)
        __builtin_amdgcn_s_setprio(2);
      else
        __builtin_amdgcn_s_setprio(1);
    }
    const unsigned q = base + lane;
    if (q >= total) continue;
    int lo = 0, hi = nbins;  // seg_pref[lo] <= q < seg_pref[hi]
    while (hi - lo > 1) {
      const int mid = (lo + hi) >> 1;
      if (seg_pref[mid] <= q) lo = mid; else hi = mid;
    }
    const unsigned idx = unsigned(lo) * unsigned(block_cap) + seg_beg[lo] + (q - seg_pref[lo]);
    const Entry4Head eh = entries[idx].h;
    const unsigned i = entry_vis[idx];
    const unsigned pos = eh.pos;
    const int lx = int(pos & 63u);
    const int ly = int((pos >> 6) & 63u);
    const int fxi = int((pos >> 12) & 255u);
    const int fyi = int((pos >> 20) & 255u);
    const bool pos_w = pos & e4_pos_w;
    const bool grad = pos & e4_gradient;
    const int grp = int(eh.gpk & 0xffffu);
    const int cfcube = int(eh.gpk >> 16);
    const GroupInfo g = groups[grp];
    const int sx = g.sx, sy = g.sy;
    // degridding uses the other Mueller matrix and the opposite conjugation
    // of gridding (src/hpg_core.hpp:598-600)
    const int* row = pos_w ? mu.Mc[gpol] : mu.M[gpol];
    const float imf = pos_w ? 1.0f : -1.0f;
    const float2* slice[N];
    bool use[N];
    const long long slice_taps = (long long)sx * sy;
    for (int v = 0; v < N; ++v) {
      const int m = row[v];
      use[v] = m >= 0 && m < g.nmueller;
      slice[v] = taps + g.tap_offset
        + (((long long)(cfcube * g.nmueller + (use[v] ? m : 0)) * geo.nf + fxi) * geo.nf
           + fyi) * slice_taps;
    }
    double phi0x = 0.0, phi0y = 0.0, dphix = 0.0, dphiy = 0.0;
    if (grad) {
      const double gxr = recs[i].grad[0];
      const double gyr = recs[i].grad[1];
      phi0x = -gxr * ((sx / 2) * geo.os - (fxi - geo.F));
      phi0y = -gyr * ((sy / 2) * geo.os - (fyi - geo.F));
      dphix = gxr * geo.os;
      dphiy = gyr * geo.os;
    }
    double2 a[N];
    for (int v = 0; v < N; ++v) a[v] = make_double2(0.0, 0.0);
    // c' x model for one tap, added to a[]
    auto add_tap = [&](const float2* cv, double2 m, int X, int Y) {
      if (grad) {
        const double phX = phi0x + X * dphix;
        const double phY = phi0y + Y * dphiy;
        m = degrid_screen(m, -phX - phY);
      }
#pragma unroll
      for (int v = 0; v < N; ++v) {
        if (!use[v]) continue;
        float2 cc = cv[v];
        cc.y *= imf;
        const double2 p = cmul(cc, m);
        a[v].x += p.x;
        a[v].y += p.y;
      }
    };
    // any support: taps in slice order (t = Y * sx + X), halo_chunk at a
    // time, the LDS pixel carried from tap to tap
    auto run_chunks = [&]() {
      const int ntap = sx * sy;
      int cx = 0, cy = 0;          // (X, Y) of the chunk's first tap
      int lc = ly * RX + lx;       // its LDS pixel
      const int wrap = RX - sx;    // pixel step from a row end to the next row
      auto do_chunk = [&](auto full, int t0) {
        constexpr bool FULL = decltype(full)::value;
        bool ok[halo_chunk];
        int Xk[halo_chunk], Yk[halo_chunk];
        double2 mv[halo_chunk];
        float2 c[halo_chunk][N];
#pragma unroll
        for (int k = 0; k < halo_chunk; ++k) {
          ok[k] = FULL || t0 + k < ntap;
          Xk[k] = cx;
          Yk[k] = cy;
          mv[k] = ok[k] ? mreg[lc] : make_double2(0.0, 0.0);
          const bool eol = ++cx == sx;
          cx = eol ? 0 : cx;
          cy += eol ? 1 : 0;
          lc += eol ? 1 + wrap : 1;
#pragma unroll
          for (int v = 0; v < N; ++v) {
            float2 tv = make_float2(0.0f, 0.0f);
            if (ok[k]) tv = slice[v][t0 + k];
            c[k][v] = tv;
          }
        }
#pragma unroll
        for (int k = 0; k < halo_chunk; ++k)
          if (ok[k]) add_tap(c[k], mv[k], Xk[k], Yk[k]);
      };
      int t0 = 0;
      for (; t0 + halo_chunk <= ntap; t0 += halo_chunk) do_chunk(std::true_type{}, t0);
      if (t0 < ntap) do_chunk(std::false_type{}, t0);
    };
    run_chunks();
    if (!one_plane) {
      for (int v = 0; v < N; ++v) {
        double2* dst = acc + size_t(i) * N + v;
        atomicAdd(&dst->x, a[v].x);
        atomicAdd(&dst->y, a[v].y);
      }
      continue;
    }
    double2 W[N];
    degrid_norms<N>(geo, g, wsum, mu, pos_w, unsigned(cfcube), fxi, fyi, W);
    const double2 phasor = cphase_call(recs[i].phase);
    const double2 conj_phasor = make_double2(phasor.x, -phasor.y);
    double2 pred_d[N];
    for (int v = 0; v < N; ++v) {
      const double2 d = (W[v].x != 0.0 || W[v].y != 0.0) ? W[v] : make_double2(1.0, 0.0);
      pred_d[v] = cmul(kdiv(a[v], d), conj_phasor);
    }
    if (pred64) {
      for (int v = 0; v < N; ++v) pred64[size_t(i) * N + v] = pred_d[v];
      continue;
    }
    for (int v = 0; v < N; ++v) {
      const float2 pred = narrow(pred_d[v]);
      if (do_grid) {
        float2 val = recs[i].vis[v];
        val.x -= pred.x;
        val.y -= pred.y;
        const float wv = recs[i].w[v];
        if (out_recs) out_recs[i].vis[v] = val;
        gvis[size_t(i) * N + v] = prephase(val, phasor, wv);
      } else if (out_recs) {
        out_recs[i].vis[v] = pred;
      }
    }
  }
}

// predictions from the per-record sums of k_degrid4 (the end of degrid_vis
// and of the degridding branch of VisibilityGridder::kernel,
// src/hpg_core.hpp:643-682, 1136-1169), one thread per record
template <int N>
__global__ __launch_bounds__(block) void k_degrid_finish(
  size_t n, Geometry geo, const GroupInfo* __restrict__ groups,
  const double2* __restrict__ wsum, Mueller mu, const double2* __restrict__ acc,
  int do_grid, const VisRec<N>* recs_in, VisRec<N>* recs, float2* __restrict__ gvis,
  double2* __restrict__ pred64) {
  // recs_in and recs may be the same records: each thread reads its
  // record's fields before writing them back; recs null: values not written
  const size_t i = blockIdx.x * (size_t)block + threadIdx.x;
  if (i >= n) return;
  const VisRec<N>& r = recs_in[i];
  const unsigned grp = r.cf_index[1];
  const unsigned cfcube = r.cf_index[0];
  const unsigned gcube = r.grid_cube;
  if (!(grp < unsigned(geo.ngroups) && gcube < unsigned(geo.nc))) return;
  const GroupInfo g = groups[grp];
  if (!(cfcube < unsigned(g.ncube))) return;
  const Coord c = vis_coord(r, geo, g);
  if (!c.inside) return;
  const bool pos_w = r.uvw[2] > 0;
  const double2 phasor = cphase(r.phase);
  double2 W[N];
  degrid_norms<N>(geo, g, wsum, mu, pos_w, cfcube, c.fx + geo.F, c.fy + geo.F, W);
  const double2 conj_phasor = make_double2(phasor.x, -phasor.y);
  float2 pred[N];
  double2 pred_d[N];
  for (int v = 0; v < N; ++v) {
    const double2 d = (W[v].x != 0.0 || W[v].y != 0.0) ? W[v] : make_double2(1.0, 0.0);
    pred_d[v] = cmul(kdiv(acc[i * N + v], d), conj_phasor);
    pred[v] = narrow(pred_d[v]);
  }
  if (pred64) {
    for (int v = 0; v < N; ++v) pred64[i * N + v] = pred_d[v];
    return;
  }
  for (int v = 0; v < N; ++v) {
    if (do_grid) {
      float2 val = r.vis[v];
      val.x -= pred[v].x;
      val.y -= pred[v].y;
      if (recs) recs[i].vis[v] = val;
      gvis[i * N + v] = prephase(val, phasor, r.w[v]);
    } else if (recs) {
      recs[i].vis[v] = pred[v];
    }
  }
}

// ---- finishing passes ------------------------------------------------------

template <typename T>
__global__ void k_normalize(
  T* __restrict__ grid, const double* __restrict__ weights, double wfactor,
  size_t plane_size, size_t count) {
  const size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (k >= count) return;
  const double d = wfactor * weights[k / plane_size];
  T g = grid[k];
  g.x = double(g.x) / d;
  g.y = double(g.y) / d;
  grid[k] = g;
}

// the finishing chain's one elementwise pass before its FFT: each value
// divided by wfactor * its plane's weight (when weights is given: the
// GridNormalizer's FP64 division, src/hpg_core.hpp:1849-1860), times s,
// and negated where x + y is odd (modulate: for even nx, ny the FFT of
// x (-1)^(x+y) is the FFT of x shifted by half the grid on both axes, so
// the shift that follows the FFT needs no pass of its own).  The launch
// is (x blocks, row y, plane), so no index needs a division.
template <typename T>
__global__ void k_finish_pre(
  T* __restrict__ grid, const double* __restrict__ weights, double wfactor, double s,
  int modulate, int nx, int ny) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= nx) return;
  const int y = blockIdx.y;
  const size_t plane = blockIdx.z;
  T* p = grid + (plane * size_t(ny) + size_t(y)) * size_t(nx) + x;
  T g = *p;
  double vx = double(g.x), vy = double(g.y);
  if (weights) {
    const double d = wfactor * weights[plane];
    vx /= d;
    vy /= d;
  }
  const double f = modulate && ((x + y) & 1) ? -s : s;
  g.x = vx * f;
  g.y = vy * f;
  *p = g;
}

template <typename T>
__global__ void k_scale(T* __restrict__ grid, double s, size_t count) {
  const size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (k >= count) return;
  T g = grid[k];
  g.x = double(g.x) * s;
  g.y = double(g.y) * s;
  grid[k] = g;
}

// even sizes: FORWARD and BACKWARD both swap (x, y) with
// ((x + nx/2) mod nx, y + ny/2) -- one thread per pair
template <typename T>
__global__ void k_shift_even(T* __restrict__ grid, int nx, int ny, size_t nplanes) {
  const size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  const size_t half = size_t(nx) * (ny / 2);
  if (k >= half * nplanes) return;
  const size_t p = k / half;
  const int r = int(k % half);
  const int x = r % nx;
  const int y = r / nx;
  T* plane = grid + p * size_t(nx) * ny;
  const size_t a = size_t(y) * nx + x;
  const size_t b = size_t(y + ny / 2) * nx + (x + nx / 2) % nx;
  const T t = plane[a];
  plane[a] = plane[b];
  plane[b] = t;
}

template <typename T>
__global__ void k_shift_general(
  const T* __restrict__ src, T* __restrict__ dst, int nx, int ny, int sx, int sy,
  size_t count) {
  const size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (k >= count) return;
  const size_t ps = size_t(nx) * ny;
  const size_t p = k / ps;
  const int r = int(k % ps);
  const int x = r % nx;
  const int y = r / nx;
  dst[p * ps + size_t((y + sy) % ny) * nx + (x + sx) % nx] = src[k];
}

__global__ void k_widen(const float2* __restrict__ s, double2* __restrict__ d, size_t n) {
  const size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (k < n) d[k] = make_double2(double(s[k].x), double(s[k].y));
}
__global__ void k_narrow(const double2* __restrict__ s, float2* __restrict__ d, size_t n) {
  const size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (k < n) d[k] = make_float2(float(s[k].x), float(s[k].y));
}

inline unsigned blocks_for(size_t n, int b = block) {
  return unsigned((n + b - 1) / b);
}

}  // namespace

// ---- launchers -------------------------------------------------------------

hipError_t launch_cf_relayout(
  hipStream_t s, const float2* staging, const unsigned ext[4],
  const GroupInfo& gi, int os, int F, int nf, float2* taps) {
  const long long total =
    (long long)gi.ncube * gi.nmueller * nf * nf * gi.sx * gi.sy;
  if (total == 0) return hipSuccess;
  hipLaunchKernelGGL(
    k_cf_relayout, dim3(blocks_for(size_t(total))), dim3(block), 0, s, staging,
    ext[0], ext[1], gi, os, F, nf, taps, total);
  return hipGetLastError();
}

hipError_t launch_cf_wsum(
  hipStream_t s, const GroupInfo& gi, int nf, const float2* taps,
  double2* wsum) {
  const long long ns = (long long)gi.ncube * gi.nmueller * nf * nf;
  if (ns == 0) return hipSuccess;
  hipLaunchKernelGGL(
    k_cf_wsum, dim3(blocks_for(size_t(ns), wsum_block / wave)), dim3(wsum_block), 0, s, gi,
    nf, taps, wsum, ns);
  return hipGetLastError();
}

size_t tiles_per_footprint(int max_sx, int max_sy) {
  const size_t tx = size_t((max_sx - 1 + tile - 1) / tile + 1);
  const size_t ty = size_t((max_sy - 1 + tile - 1) / tile + 1);
  return tx * ty;
}

size_t scan_temp_bytes(size_t nkeys) {
  size_t bytes = 0;
  (void)hipcub::DeviceScan::ExclusiveSum(
    nullptr, bytes, static_cast<unsigned*>(nullptr),
    static_cast<unsigned*>(nullptr), int(nkeys + 1));
  return bytes;
}

hipError_t launch_scan(hipStream_t s, size_t nkeys, CallBuffers& cb) {
  size_t bytes = cb.scan_temp_bytes;
  hipError_t e = hipcub::DeviceScan::ExclusiveSum(
    cb.scan_temp, bytes, cb.hist, cb.offsets, int(nkeys + 1), s);
  if (e != hipSuccess) return e;
  return hipMemcpyAsync(
    cb.cursor, cb.offsets, nkeys * sizeof(unsigned), hipMemcpyDeviceToDevice, s);
}

hipError_t launch_count(
  hipStream_t s, int npol, size_t n, const Geometry& geo, const DeviceCF& cf,
  const Mueller& mu, bool update_weights, bool do_grid, CallBuffers& cb,
  double* weights, bool halo_entries, bool quad_keys) {
  if (n == 0) return hipSuccess;
#define HPGX_COUNT(NP, H)                                                      \
  hipLaunchKernelGGL(                                                          \
    (k_count<NP, H>), dim3(std::min(unsigned(blocks_for(n)), count_blocks_max)), \
    dim3(block), 0, s,                                                         \
    static_cast<const VisRec<NP>*>(cb.vis), n, geo, cf.groups, cf.wsum, mu,    \
    int(update_weights), int(do_grid), cb.hist, weights, cb.invalid_count,     \
    int(halo_entries && quad_keys))
  if (halo_entries) {
    switch (npol) {
    case 1: HPGX_COUNT(1, true); break;
    case 2: HPGX_COUNT(2, true); break;
    case 3: HPGX_COUNT(3, true); break;
    case 4: HPGX_COUNT(4, true); break;
    default: return hipErrorInvalidValue;
    }
  } else {
    switch (npol) {
    case 1: HPGX_COUNT(1, false); break;
    case 2: HPGX_COUNT(2, false); break;
    case 3: HPGX_COUNT(3, false); break;
    case 4: HPGX_COUNT(4, false); break;
    default: return hipErrorInvalidValue;
    }
  }
#undef HPGX_COUNT
  return hipGetLastError();
}

size_t entry_bytes(int npol) { return sizeof(EntryHead) + sizeof(float2) * size_t(npol); }

hipError_t launch_tally(
  hipStream_t s, int npol, size_t n, const Geometry& geo, const DeviceCF& cf,
  const Mueller& mu, bool update_weights, bool gvis_ready, GvisSource src,
  const double2* residuals, CallBuffers& cb, double* weights) {
  if (n == 0) return hipSuccess;
  if (npol != 1) return hipErrorInvalidValue;
  hipLaunchKernelGGL(
    k_tally<1>, dim3(std::min(blocks_for(n), count_blocks_max)), dim3(block), 0, s,
    static_cast<const VisRec<1>*>(cb.vis), n, geo, cf.groups, cf.wsum, mu,
    int(update_weights), int(gvis_ready), int(src == GvisSource::Adjoint), cb.gvis, residuals,
    cb.hist, weights, cb.invalid_count, static_cast<Entry4<1>*>(cb.tally_entries),
    static_cast<uint2*>(cb.tally_keys));
  return hipGetLastError();
}

hipError_t launch_place(hipStream_t s, int npol, size_t n, const Geometry& geo,
                        CallBuffers& cb) {
  const size_t slots = n * size_t(geo.nr);
  if (slots == 0) return hipSuccess;
  if (npol != 1) return hipErrorInvalidValue;
  hipLaunchKernelGGL(
    k_place<1>, dim3(blocks_for(slots)), dim3(block), 0, s, slots, geo.nr, cb.offsets,
    static_cast<const Entry4<1>*>(cb.tally_entries), static_cast<const uint2*>(cb.tally_keys),
    static_cast<Entry4<1>*>(cb.stream_entries), cb.entry_vis);
  return hipGetLastError();
}

hipError_t launch_emit(
  hipStream_t s, int npol, size_t n, const Geometry& geo, const DeviceCF& cf,
  const Mueller& mu, bool gvis_ready, GvisSource src, const double2* residuals,
  CallBuffers& cb, bool halo_entries, bool quad_keys) {
  if (n == 0) return hipSuccess;
  const int adj = src == GvisSource::Adjoint;
#define HPGX_EMIT(NP, H)                                                       \
  hipLaunchKernelGGL(                                                          \
    (k_emit<NP, H>), dim3(blocks_for(n)), dim3(block), 0, s,                   \
    static_cast<const VisRec<NP>*>(cb.vis), n, geo, cf.groups, cf.wsum, mu,    \
    int(gvis_ready), adj, cb.gvis, residuals, cb.cursor, cb.stream_entries,    \
    cb.entry_vis, int(halo_entries && quad_keys))
  if (halo_entries) {
    switch (npol) {
    case 1: HPGX_EMIT(1, true); break;
    case 2: HPGX_EMIT(2, true); break;
    case 3: HPGX_EMIT(3, true); break;
    case 4: HPGX_EMIT(4, true); break;
    default: return hipErrorInvalidValue;
    }
  } else {
    switch (npol) {
    case 1: HPGX_EMIT(1, false); break;
    case 2: HPGX_EMIT(2, false); break;
    case 3: HPGX_EMIT(3, false); break;
    case 4: HPGX_EMIT(4, false); break;
    default: return hipErrorInvalidValue;
    }
  }
#undef HPGX_EMIT
  return hipGetLastError();
}

hipError_t launch_tile_stream(
  hipStream_t s, int npol, int mode, size_t nkeys, const Geometry& geo,
  const DeviceCF& cf, const Mueller& mu, const CallBuffers& cb, void* grid) {
  if (nkeys == 0) return hipSuccess;
  const bool unr = size_t(cf.max_sx) * size_t(cf.max_sy) >= size_t(unroll_min_taps);
#define HPGX_TILE2(NP, MD)                                                     \
  hipLaunchKernelGGL(                                                          \
    unr ? (k_tile2<NP, MD, true>) : (k_tile2<NP, MD, false>),                  \
    dim3(unsigned(nkeys)), dim3(block), 0, s, geo,                             \
    cf.groups, cf.taps, mu, cb.offsets,                                        \
    static_cast<const Entry<NP>*>(cb.stream_entries), cb.entry_vis,            \
    static_cast<const VisRec<NP>*>(cb.vis), grid)
  if (mode == 0) {
    switch (npol) {
    case 1: HPGX_TILE2(1, 0); break;
    case 2: HPGX_TILE2(2, 0); break;
    case 3: HPGX_TILE2(3, 0); break;
    case 4: HPGX_TILE2(4, 0); break;
    default: return hipErrorInvalidValue;
    }
  } else {
    switch (npol) {
    case 1: HPGX_TILE2(1, 1); break;
    case 2: HPGX_TILE2(2, 1); break;
    case 3: HPGX_TILE2(3, 1); break;
    case 4: HPGX_TILE2(4, 1); break;
    default: return hipErrorInvalidValue;
    }
  }
#undef HPGX_TILE2
  return hipGetLastError();
}

int bin_blocks_for(size_t n) { return int((n + bin_records - 1) / bin_records); }
size_t bin_block_records() { return bin_records; }
int bin_keys_limit() { return bin_keys_max; }
int bin_quad_keys_limit() { return bin_quad_keys_max; }
int tile_bins_limit() { return tile_bins_max; }

hipError_t launch_bin(
  hipStream_t s, int npol, size_t n, const Geometry& geo, const DeviceCF& cf,
  const Mueller& mu, bool update_weights, bool gvis_ready, GvisSource src,
  const double2* residuals, CallBuffers& cb, size_t block_cap, double* weights,
  bool halo_entries, bool degrid_pass, void* copy_records, bool quad_keys) {
  if (n == 0) return hipSuccess;
  if (copy_records && !degrid_pass) return hipErrorInvalidValue;
  if (quad_keys && (!halo_entries || degrid_pass)) return hipErrorInvalidValue;
  const int adj = src == GvisSource::Adjoint;
  const int dg = degrid_pass;
  const unsigned nb = unsigned(bin_blocks_for(n));
#define HPGX_BIN(NP, H)                                                        \
  hipLaunchKernelGGL(                                                          \
    (k_bin<NP, H>), dim3(nb), dim3(bin_threads), 0, s,                         \
    static_cast<const VisRec<NP>*>(cb.vis), n, geo, cf.groups, cf.wsum, mu,    \
    int(update_weights), int(gvis_ready), adj, dg, cb.gvis, residuals,         \
    cb.off_table, cb.stream_entries, cb.entry_vis, block_cap, weights,         \
    cb.invalid_count, static_cast<VisRec<NP>*>(copy_records), int(quad_keys))
  if (halo_entries) {
    switch (npol) {
    case 1: HPGX_BIN(1, true); break;
    case 2: HPGX_BIN(2, true); break;
    case 3: HPGX_BIN(3, true); break;
    case 4: HPGX_BIN(4, true); break;
    default: return hipErrorInvalidValue;
    }
  } else {
    switch (npol) {
    case 1: HPGX_BIN(1, false); break;
    case 2: HPGX_BIN(2, false); break;
    case 3: HPGX_BIN(3, false); break;
    case 4: HPGX_BIN(4, false); break;
    default: return hipErrorInvalidValue;
    }
  }
#undef HPGX_BIN
  return hipGetLastError();
}

size_t halo_entry_bytes(int npol) {
  return sizeof(Entry4Head) + sizeof(float2) * size_t(npol);
}
size_t halo_pixels_per_key(int hx, int hy) { return size_t(halo_pixels(hx, hy)); }
int halo_tile_limit() { return tile; }
size_t halo_lds_bytes(int hx, int hy, int nbins) {
  const size_t r = size_t(tile + hx) * (tile + hy);
  return r * 16
    + sizeof(unsigned)
      * (2 * size_t(nbins) + 1 + halo_block / wave + 1 + halo_classes + halo_sched);
}

// the halo pass after a haloed-tile kernel
hipError_t launch_halo(
  hipStream_t s, int mode, size_t nkeys, const Geometry& geo, const CallBuffers& cb,
  int hx, int hy, void* grid) {
  const size_t zone = size_t(hy) * tile + size_t(tile - hy) * hx;
  const size_t count = nkeys * zone;
  if (count == 0) return hipSuccess;
  if (mode == 0)
    hipLaunchKernelGGL(
      k_halo<double2>, dim3(blocks_for(count)), dim3(block), 0, s, geo, hx, hy,
      cb.halo, cb.halo_set, nkeys, static_cast<double2*>(grid));
  else
    hipLaunchKernelGGL(
      k_halo<float2>, dim3(blocks_for(count)), dim3(block), 0, s, geo, hx, hy,
      cb.halo, cb.halo_set, nkeys, static_cast<float2*>(grid));
  return hipGetLastError();
}

size_t quad_lds_bytes(int hx, int hy, int npol) {
  const size_t regions = 4 * size_t(quad_edge + hx) * (quad_edge + hy) * sizeof(float4);
  size_t rec = 0, rb = 0;
  switch (npol) {
  case 1: rec = sizeof(QuadRec<1>); rb = quad_round<1>(); break;
  case 2: rec = sizeof(QuadRec<2>); rb = quad_round<2>(); break;
  case 3: rec = sizeof(QuadRec<3>); rb = quad_round<3>(); break;
  default: rec = sizeof(QuadRec<4>); rb = quad_round<4>(); break;
  }
  return regions + rb * rec
    + sizeof(unsigned) * (2 * size_t(quad_bins_max) + 1 + quad_block / wave + 1 + 5);
}
int quad_bins_limit() { return quad_bins_max; }
bool quad_supported(int hx, int hy, int npol) {
  return hx >= 0 && hy >= 0 && hx < tile && hy < tile && quad_lds_bytes(hx, hy, npol) <= 65536
    && (tile + hx) * (tile + hy) <= quad_flush * quad_block;
}

size_t sub_lds_bytes(int hx, int hy) {
  const int px = ((quad_edge + hx) * (quad_edge + hy) + wave - 1) / wave;
  const int pxc = px <= 4 ? 4 : px <= 6 ? 6 : px <= 8 ? 8 : 9;
  return size_t(sub_waves) * pxc * wave * sizeof(float2);
}
int sub_bins_limit() { return sub_bins_max; }
bool sub_supported(int hx, int hy) { return hx >= 0 && hy >= 0 && hx < 8 && hy < 8; }

hipError_t launch_tile_sub(
  hipStream_t s, int npol, size_t n, size_t ntile_keys, const Geometry& geo, const DeviceCF& cf,
  size_t taps_bytes, bool uni, const Mueller& mu, const CallBuffers& cb, size_t block_cap,
  int hx, int hy, void* grid, bool global_bins) {
  if (ntile_keys == 0 || n == 0) return hipSuccess;
  const int nb = global_bins ? 1 : bin_blocks_for(n);
  // segment starts are 16-bit offsets within a bin block; CF taps are
  // addressed by 32-bit byte offsets; global bins for one polarization
  if (!sub_supported(hx, hy) || taps_bytes > 0xffffffffull) return hipErrorInvalidValue;
  if (global_bins ? npol != 1 : (nb > sub_bins_max || block_cap > 65536))
    return hipErrorInvalidValue;
  const unsigned* offs = global_bins ? cb.offsets : cb.off_table;
  const int Q = (quad_edge + hx) * (quad_edge + hy);
  const int px = (Q + wave - 1) / wave;  // 4 .. 9
  const int pxc = px <= 4 ? 4 : px <= 6 ? 6 : px <= 8 ? 8 : 9;
  const size_t dyn = size_t(sub_waves) * pxc * wave * sizeof(float2);
#define HPGX_SUB_G(NP, PX, PK, GB)                                             \
  hipLaunchKernelGGL(                                                          \
    uni ? (k_sub<NP, PX, true, PK, GB>) : (k_sub<NP, PX, false, PK, GB>),      \
    dim3(unsigned(ntile_keys)), dim3(sub_block), dyn, s,                       \
    geo, cf.groups, cf.taps, unsigned(taps_bytes), mu, offs, nb,               \
    block_cap,                                                                 \
    static_cast<const Entry4<NP>*>(cb.stream_entries), cb.entry_vis,           \
    static_cast<const VisRec<NP>*>(cb.vis), hx, hy, static_cast<float2*>(grid),\
    cb.halo, cb.halo_set)
#define HPGX_SUB_N(NP, PX, PK) HPGX_SUB_G(NP, PX, PK, false)
#define HPGX_SUB_1(NP, PX, PK)                                                 \
  do {                                                                         \
    if (global_bins) HPGX_SUB_G(1, PX, PK, true);                              \
    else HPGX_SUB_G(1, PX, PK, false);                                         \
  } while (0)
#define HPGX_SUB_PX(HPGX_SUB, NP, PK)                                          \
  switch (pxc) {                                                               \
  case 4: HPGX_SUB(NP, 4, PK); break;                                          \
  case 6: HPGX_SUB(NP, 6, PK); break;                                          \
  case 8: HPGX_SUB(NP, 8, PK); break;                                          \
  default: HPGX_SUB(NP, 9, PK); break;                                         \
  }
  // one word per slot: region pixels below sub_pk_flip, slices addressed
  // by tap index below sub_pk_taps
  static_assert(9 * wave < int(sub_pk_flip), "region pixel in the packed word");
  const bool pk = npol == 1 && taps_bytes / sizeof(float2) < sub_pk_taps;
  switch (npol) {
  case 1:
    if (pk) HPGX_SUB_PX(HPGX_SUB_1, 1, true) else HPGX_SUB_PX(HPGX_SUB_1, 1, false)
    break;
  case 2: HPGX_SUB_PX(HPGX_SUB_N, 2, false); break;
  case 3: HPGX_SUB_PX(HPGX_SUB_N, 3, false); break;
  case 4: HPGX_SUB_PX(HPGX_SUB_N, 4, false); break;
  default: return hipErrorInvalidValue;
  }
#undef HPGX_SUB_PX
#undef HPGX_SUB_1
#undef HPGX_SUB_N
#undef HPGX_SUB_G
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  return launch_halo(s, 1, ntile_keys, geo, cb, hx, hy, grid);
}

hipError_t launch_tile_halo(
  hipStream_t s, int npol, int mode, int acc, size_t n, size_t nkeys, const Geometry& geo,
  const DeviceCF& cf, const Mueller& mu, const CallBuffers& cb, size_t block_cap,
  int hx, int hy, void* grid, bool global_bins, bool quad) {
  if (nkeys == 0 || n == 0) return hipSuccess;
  if (hx < 0 || hy < 0 || hx >= tile || hy >= tile) return hipErrorInvalidValue;
  // global bins: one contiguous range per key, read as a single bin block
  // whose offset-table row is the scanned offsets
  const int nb = global_bins ? 1 : bin_blocks_for(n);
  const unsigned* table = global_bins ? cb.offsets : cb.off_table;
  if (quad) {
    if (nb > quad_bins_max || !quad_supported(hx, hy, npol) || mode != 1 || acc != acc_f32_cas)
      return hipErrorInvalidValue;
    const size_t dyn = 4 * size_t(quad_edge + hx) * (quad_edge + hy) * sizeof(float4);
    // every support within one lane step (hx, hy = largest support - 1)
    const bool one = hx < 8 && hy < 8;
#define HPGX_QUAD(NP)                                                          \
  hipLaunchKernelGGL(                                                          \
    one ? (k_quad<NP, true>) : (k_quad<NP, false>),                            \
    dim3(unsigned(nkeys)), dim3(quad_block), dyn, s,                           \
    geo, cf.groups, cf.taps, mu, table, nb, block_cap,                         \
    static_cast<const Entry4<NP>*>(cb.stream_entries), cb.entry_vis,           \
    static_cast<const VisRec<NP>*>(cb.vis), hx, hy, grid, cb.halo, cb.halo_set)
    switch (npol) {
    case 1: HPGX_QUAD(1); break;
    case 2: HPGX_QUAD(2); break;
    case 3: HPGX_QUAD(3); break;
    case 4: HPGX_QUAD(4); break;
    default: return hipErrorInvalidValue;
    }
#undef HPGX_QUAD
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    return launch_halo(s, mode, nkeys, geo, cb, hx, hy, grid);
  }
  const size_t lds = halo_lds_bytes(hx, hy, nb);
  if (lds > 65536) return hipErrorInvalidValue;
#define HPGX_TILE4(NP, MD, AC)                                                 \
  hipLaunchKernelGGL(                                                          \
    (k_tile4<NP, MD, AC>), dim3(unsigned(nkeys)), dim3(halo_block), lds, s,    \
    geo, cf.groups, cf.taps, mu, table, nb, block_cap,                         \
    static_cast<const Entry4<NP>*>(cb.stream_entries), cb.entry_vis,           \
    static_cast<const VisRec<NP>*>(cb.vis), hx, hy, grid, cb.halo, cb.halo_set)
#define HPGX_TILE4_NP(MD, AC)                                                  \
  switch (npol) {                                                              \
  case 1: HPGX_TILE4(1, MD, AC); break;                                        \
  case 2: HPGX_TILE4(2, MD, AC); break;                                        \
  case 3: HPGX_TILE4(3, MD, AC); break;                                        \
  case 4: HPGX_TILE4(4, MD, AC); break;                                        \
  default: return hipErrorInvalidValue;                                        \
  }
  if (mode == 0) {
    HPGX_TILE4_NP(0, acc_f64)
  } else if (acc == 0) {
    HPGX_TILE4_NP(1, acc_f64)
  } else {
    HPGX_TILE4_NP(1, acc_f32_cas)
  }
#undef HPGX_TILE4_NP
#undef HPGX_TILE4
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  return launch_halo(s, mode, nkeys, geo, cb, hx, hy, grid);
}

hipError_t launch_tile_blocal(
  hipStream_t s, int npol, int mode, size_t n, size_t nkeys, const Geometry& geo,
  const DeviceCF& cf, const Mueller& mu, const CallBuffers& cb, size_t block_cap,
  void* grid) {
  if (nkeys == 0 || n == 0) return hipSuccess;
  const int nb = bin_blocks_for(n);
#define HPGX_TILE3(NP, MD, UU)                                                 \
  hipLaunchKernelGGL(                                                          \
    (k_tile3<NP, MD, UU>), dim3(unsigned(nkeys)), dim3(block), 0, s, geo,      \
    cf.groups, cf.taps, mu, cb.off_table, nb, block_cap,                       \
    static_cast<const Entry<NP>*>(cb.stream_entries), cb.entry_vis,            \
    static_cast<const VisRec<NP>*>(cb.vis), grid)
  if (mode == 0) {
    switch (npol) {
    case 1: HPGX_TILE3(1, 0, 4); break;
    case 2: HPGX_TILE3(2, 0, 4); break;
    case 3: HPGX_TILE3(3, 0, 2); break;
    case 4: HPGX_TILE3(4, 0, 2); break;
    default: return hipErrorInvalidValue;
    }
  } else {
    switch (npol) {
    case 1: HPGX_TILE3(1, 1, 4); break;
    case 2: HPGX_TILE3(2, 1, 4); break;
    case 3: HPGX_TILE3(3, 1, 2); break;
    case 4: HPGX_TILE3(4, 1, 2); break;
    default: return hipErrorInvalidValue;
    }
  }
#undef HPGX_TILE3
  return hipGetLastError();
}

size_t degrid_lane_max_taps() { return size_t(degrid_lane_taps); }

hipError_t launch_degrid_tiles(
  hipStream_t s, int npol, size_t n, size_t nkeys, const Geometry& geo,
  const DeviceCF& cf, const Mueller& mu, const CallBuffers& cb, size_t block_cap,
  int hx, int hy, const double2* model, double2* acc, bool one_plane, bool do_grid,
  void* out_recs, double2* pred64) {
  if (nkeys == 0 || n == 0) return hipSuccess;
  if (hx < 0 || hy < 0 || hx >= tile || hy >= tile) return hipErrorInvalidValue;
  const int nb = bin_blocks_for(n);
  const size_t lds = halo_lds_bytes(hx, hy, nb);
  if (lds > 65536) return hipErrorInvalidValue;
#define HPGX_DEGRID4(NP)                                                       \
  hipLaunchKernelGGL(                                                          \
    k_degrid4<NP>, dim3(unsigned(nkeys)), dim3(halo_block), lds, s, geo,       \
    cf.groups, cf.taps, mu, cb.off_table, nb, block_cap,                       \
    static_cast<const Entry4<NP>*>(cb.stream_entries), cb.entry_vis,           \
    static_cast<const VisRec<NP>*>(cb.vis), hx, hy, model, acc, int(one_plane),  \
    cf.wsum, int(do_grid), static_cast<VisRec<NP>*>(out_recs), cb.gvis, pred64)
  switch (npol) {
  case 1: HPGX_DEGRID4(1); break;
  case 2: HPGX_DEGRID4(2); break;
  case 3: HPGX_DEGRID4(3); break;
  case 4: HPGX_DEGRID4(4); break;
  default: return hipErrorInvalidValue;
  }
#undef HPGX_DEGRID4
  return hipGetLastError();
}

hipError_t launch_degrid_finish(
  hipStream_t s, int npol, size_t n, const Geometry& geo, const DeviceCF& cf,
  const Mueller& mu, const double2* acc, bool do_grid, const void* records,
  void* out_records, CallBuffers& cb, double2* pred64) {
  if (n == 0) return hipSuccess;
  const void* in = records;
#define HPGX_DFIN(NP)                                                          \
  hipLaunchKernelGGL(                                                          \
    k_degrid_finish<NP>, dim3(blocks_for(n)), dim3(block), 0, s, n, geo,       \
    cf.groups, cf.wsum, mu, acc, int(do_grid),                                 \
    static_cast<const VisRec<NP>*>(in), static_cast<VisRec<NP>*>(out_records), \
    cb.gvis, pred64)
  switch (npol) {
  case 1: HPGX_DFIN(1); break;
  case 2: HPGX_DFIN(2); break;
  case 3: HPGX_DFIN(3); break;
  case 4: HPGX_DFIN(4); break;
  default: return hipErrorInvalidValue;
  }
#undef HPGX_DFIN
  return hipGetLastError();
}

bool degrid_sorts(const DeviceCF& cf) {
  return size_t(cf.max_sx) * size_t(cf.max_sy) > size_t(degrid_lane_taps);
}

size_t degrid_order_temp_bytes(size_t n) {
  size_t bytes = 0;
  (void)hipcub::DeviceRadixSort::SortPairs(
    nullptr, bytes, static_cast<unsigned*>(nullptr), static_cast<unsigned*>(nullptr),
    static_cast<unsigned*>(nullptr), static_cast<unsigned*>(nullptr),
    int(std::min<size_t>(n, 0x7fffffffull)), 0, 32);
  return bytes;
}

hipError_t launch_degrid(
  hipStream_t s, int npol, size_t n, const Geometry& geo, const DeviceCF& cf,
  const Mueller& mu, const double2* model, bool do_grid, CallBuffers& cb,
  double2* pred64, const void* records) {
  if (n == 0) return hipSuccess;
  const void* in = records ? records : cb.vis;
  if (size_t(cf.max_sx) * size_t(cf.max_sy) <= size_t(degrid_lane_taps)) {
#define HPGX_DEGRID_LANE(NP)                                                   \
  hipLaunchKernelGGL(                                                          \
    k_degrid_lane<NP>, dim3(blocks_for(n)), dim3(block), 0, s, n, geo,         \
    cf.groups, cf.taps, cf.wsum, mu, model, int(do_grid),                      \
    static_cast<const VisRec<NP>*>(in), static_cast<VisRec<NP>*>(cb.vis),      \
    cb.gvis, pred64)
    switch (npol) {
    case 1: HPGX_DEGRID_LANE(1); break;
    case 2: HPGX_DEGRID_LANE(2); break;
    case 3: HPGX_DEGRID_LANE(3); break;
    case 4: HPGX_DEGRID_LANE(4); break;
    default: return hipErrorInvalidValue;
    }
#undef HPGX_DEGRID_LANE
    return hipGetLastError();
  }
  if (records) return hipErrorInvalidValue;  // the wave form reads cb.vis
  // (callers stage records for supports above degrid_lane_max_taps())
  // the records in CF-slice order where the order buffers hold the batch
  // and the keys fit 32 bits
  const unsigned* order = nullptr;
  const unsigned long long stride = (unsigned long long)std::max(cf.max_ncube, 1) * cf.nf * cf.nf;
  const unsigned long long kmax = stride * (unsigned long long)cf.ngroups;
  if (model && cb.degrid_order && cb.order_capacity >= n && n <= 0x7fffffffull
      && kmax > 0 && kmax < 0xffffffffull) {
    unsigned* keys = cb.degrid_order;
    unsigned* keys_sorted = keys + cb.order_capacity;
    unsigned* idx = keys + 2 * cb.order_capacity;
    unsigned* idx_sorted = keys + 3 * cb.order_capacity;
#define HPGX_DKEYS(NP)                                                         \
  hipLaunchKernelGGL(                                                          \
    k_degrid_keys<NP>, dim3(blocks_for(n)), dim3(block), 0, s, n, geo,         \
    cf.groups, static_cast<const VisRec<NP>*>(cb.vis), unsigned(stride),       \
    unsigned(kmax), keys, idx)
    switch (npol) {
    case 1: HPGX_DKEYS(1); break;
    case 2: HPGX_DKEYS(2); break;
    case 3: HPGX_DKEYS(3); break;
    case 4: HPGX_DKEYS(4); break;
    default: return hipErrorInvalidValue;
    }
#undef HPGX_DKEYS
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const int bits = 32 - __builtin_clz(unsigned(kmax));
    size_t bytes = cb.order_temp_bytes;
    e = hipcub::DeviceRadixSort::SortPairs(cb.order_temp, bytes, keys, keys_sorted, idx,
                                           idx_sorted, int(n), 0, bits, s);
    if (e != hipSuccess) return e;
    order = idx_sorted;
  }
  const unsigned nb = unsigned((n + (block / wave) - 1) / (block / wave));
#define HPGX_DEGRID(NP)                                                        \
  hipLaunchKernelGGL(                                                          \
    k_degrid<NP>, dim3(nb), dim3(block), 0, s, n, geo, cf.groups, cf.taps,     \
    cf.wsum, mu, model, int(do_grid), static_cast<VisRec<NP>*>(cb.vis),        \
    cb.gvis, pred64, order)
  switch (npol) {
  case 1: HPGX_DEGRID(1); break;
  case 2: HPGX_DEGRID(2); break;
  case 3: HPGX_DEGRID(3); break;
  case 4: HPGX_DEGRID(4); break;
  default: return hipErrorInvalidValue;
  }
#undef HPGX_DEGRID
  return hipGetLastError();
}

bool finish_pre_fits(const Geometry& geo) {
  return geo.ny <= 65535 && size_t(geo.nr) * geo.nc <= 65535;
}

hipError_t launch_normalize(
  hipStream_t s, int mode, void* grid, const double* weights, double wfactor,
  const Geometry& geo) {
  const size_t ps = size_t(geo.nx) * geo.ny;
  const size_t count = ps * geo.nr * geo.nc;
  if (count == 0) return hipSuccess;
  // the same FP64 division, times 1
  if (finish_pre_fits(geo))
    return launch_finish_pre(s, mode, grid, weights, wfactor, 1.0, false, geo);
  if (mode == 0)
    hipLaunchKernelGGL(
      k_normalize<double2>, dim3(blocks_for(count)), dim3(block), 0, s,
      static_cast<double2*>(grid), weights, wfactor, ps, count);
  else
    hipLaunchKernelGGL(
      k_normalize<float2>, dim3(blocks_for(count)), dim3(block), 0, s,
      static_cast<float2*>(grid), weights, wfactor, ps, count);
  return hipGetLastError();
}

hipError_t launch_finish_pre(
  hipStream_t s, int mode, void* grid, const double* weights, double wfactor, double scale,
  bool modulate, const Geometry& geo) {
  const size_t planes = size_t(geo.nr) * geo.nc;
  if (size_t(geo.nx) * geo.ny * planes == 0) return hipSuccess;
  // grid rows and planes index the launch's y and z
  if (!finish_pre_fits(geo)) return hipErrorInvalidValue;
  const dim3 grid_dim(unsigned((geo.nx + block - 1) / block), unsigned(geo.ny), unsigned(planes));
  if (mode == 0)
    hipLaunchKernelGGL(
      k_finish_pre<double2>, grid_dim, dim3(block), 0, s,
      static_cast<double2*>(grid), weights, wfactor, scale, int(modulate), geo.nx, geo.ny);
  else
    hipLaunchKernelGGL(
      k_finish_pre<float2>, grid_dim, dim3(block), 0, s,
      static_cast<float2*>(grid), weights, wfactor, scale, int(modulate), geo.nx, geo.ny);
  return hipGetLastError();
}

hipError_t launch_scale(
  hipStream_t s, int mode, void* grid, double inv_norm, size_t count) {
  if (count == 0) return hipSuccess;
  if (mode == 0)
    hipLaunchKernelGGL(
      k_scale<double2>, dim3(blocks_for(count)), dim3(block), 0, s,
      static_cast<double2*>(grid), inv_norm, count);
  else
    hipLaunchKernelGGL(
      k_scale<float2>, dim3(blocks_for(count)), dim3(block), 0, s,
      static_cast<float2*>(grid), inv_norm, count);
  return hipGetLastError();
}

hipError_t launch_shift(
  hipStream_t s, int mode, void* grid, void* scratch, int direction,
  const Geometry& geo) {
  const size_t ps = size_t(geo.nx) * geo.ny;
  const size_t nplanes = size_t(geo.nr) * geo.nc;
  const size_t count = ps * nplanes;
  if (count == 0) return hipSuccess;
  const size_t esize = mode == 0 ? sizeof(double2) : sizeof(float2);
  if (geo.nx % 2 == 0 && geo.ny % 2 == 0) {
    const size_t pairs = size_t(geo.nx) * (geo.ny / 2) * nplanes;
    if (mode == 0)
      hipLaunchKernelGGL(
        k_shift_even<double2>, dim3(blocks_for(pairs)), dim3(block), 0, s,
        static_cast<double2*>(grid), geo.nx, geo.ny, nplanes);
    else
      hipLaunchKernelGGL(
        k_shift_even<float2>, dim3(blocks_for(pairs)), dim3(block), 0, s,
        static_cast<float2*>(grid), geo.nx, geo.ny, nplanes);
    return hipGetLastError();
  }
  const int sx = direction == 0 ? geo.nx / 2 : geo.nx - geo.nx / 2;
  const int sy = direction == 0 ? geo.ny / 2 : geo.ny - geo.ny / 2;
  if (mode == 0)
    hipLaunchKernelGGL(
      k_shift_general<double2>, dim3(blocks_for(count)), dim3(block), 0, s,
      static_cast<const double2*>(grid), static_cast<double2*>(scratch), geo.nx,
      geo.ny, sx, sy, count);
  else
    hipLaunchKernelGGL(
      k_shift_general<float2>, dim3(blocks_for(count)), dim3(block), 0, s,
      static_cast<const float2*>(grid), static_cast<float2*>(scratch), geo.nx,
      geo.ny, sx, sy, count);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  return hipMemcpyAsync(grid, scratch, count * esize, hipMemcpyDeviceToDevice, s);
}

hipError_t launch_widen(hipStream_t s, const float2* src, double2* dst, size_t count) {
  if (count == 0) return hipSuccess;
  hipLaunchKernelGGL(k_widen, dim3(blocks_for(count)), dim3(block), 0, s, src, dst, count);
  return hipGetLastError();
}

hipError_t launch_narrow(hipStream_t s, const double2* src, float2* dst, size_t count) {
  if (count == 0) return hipSuccess;
  hipLaunchKernelGGL(k_narrow, dim3(blocks_for(count)), dim3(block), 0, s, src, dst, count);
  return hipGetLastError();
}

}  // namespace hpgx
