// TEST INFRASTRUCTURE ONLY -- the CPU oracle for the HPG gridding hot path.
//
// This file restates, in plain single-threaded FP64 C++, the semantics of the
// reference's Serial device for one gridding state (hannajonathan/hpg v2.0.1):
// visibility coordinates, visibility pre-phasing, gridding with and without
// weights, degridding with residual/predicted write-back, weight
// normalization, the per-plane FFT and the grid shift.  A second entry point
// runs the gridding stage the way the reference's OpenMP device does (one
// task per (visibility, grid polarization), per-tap FP64 sin/cos, atomic FP64
// adds) and is the CPU baseline that the GPU path is timed against.
//
// Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
// load this library, and only as the checker (or, for the OpenMP entry point,
// as the timed CPU baseline).  The product (hpg_amd/) never calls it.
//
// Parity pins for each function are listed in the header of
// tests/test_oracle.py; functions whose values only the reference itself could
// pin are marked "parity unpinned" below.
//
// Floating point: build with -O2 -ffp-contract=off and without -ffast-math, so
// that every product and sum is rounded exactly where the reference's source
// rounds it (no fused multiply-add).

#include <algorithm>
#include <array>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

// CFArray::padding (src/hpg.hpp:655)
constexpr int cf_padding = 2;
// speed of light used by Vis<N> (src/hpg_core.hpp:527)
constexpr double c_light = 299792458.0;

struct cf32 {
  float re, im;
};
struct cf64 {
  double re, im;
};

// hpg::VisData<N> record (src/hpg.hpp:236-311): same member order and
// therefore the same memory layout (88/96/112/120 bytes for N = 1..4).
template <unsigned N>
struct VisData {
  std::array<std::complex<float>, N> vis;
  std::array<float, N> weights;
  double frequency;
  double phase;
  std::array<double, 3> uvw;
  unsigned grid_cube;
  std::array<unsigned, 2> cf_index;  // {cube, grp}
  std::array<double, 2> cf_phase_gradient;
};
static_assert(sizeof(VisData<1>) == 88, "VisData<1> layout");
static_assert(sizeof(VisData<2>) == 96, "VisData<2> layout");
static_assert(sizeof(VisData<3>) == 112, "VisData<3> layout");
static_assert(sizeof(VisData<4>) == 120, "VisData<4> layout");

// Kokkos complex arithmetic as used by the kernels (Kokkos 4.x,
// Kokkos_Complex.hpp): textbook product, no scaling.
inline cf32 mul(cf32 a, cf32 b) {
  return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
inline cf64 mul(cf64 a, cf64 b) {
  return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
// complex<float> * complex<double> promotes to double (Kokkos mixed operator*)
inline cf64 mul(cf32 a, cf64 b) {
  return {double(a.re) * b.re - double(a.im) * b.im,
          double(a.re) * b.im + double(a.im) * b.re};
}
inline cf64 widen(cf32 a) { return {double(a.re), double(a.im)}; }
inline cf32 narrow(cf64 a) { return {float(a.re), float(a.im)}; }

// Kokkos 4.x operator/(complex<T1>, complex<T2>): scale both operands by the
// 1-norm of the divisor, multiply by the scaled conjugate and divide by the
// scaled squared magnitude.  Third-party arithmetic (Kokkos >= 4.0.00, not
// vendored by the reference): parity unpinned by the reference's fixtures.
inline cf64 kokkos_div(cf64 x, cf64 y) {
  const double s = std::fabs(y.re) + std::fabs(y.im);
  if (s == 0.0) return {x.re / s, x.im / s};
  const cf64 xs{x.re / s, x.im / s};
  const cf64 ycs{y.re / s, -y.im / s};
  const double yabs = ycs.re * ycs.re + ycs.im * ycs.im;
  cf64 r = mul(xs, ycs);
  r.re /= yabs;
  r.im /= yabs;
  return r;
}

inline cf64 cphase(double ph) { return {std::cos(ph), std::sin(ph)}; }

// compute_vis_coord (src/hpg_core.hpp:422-446)
struct AxisCoord {
  long grid_coord, cf_major, cf_minor, fine_offset;
};
inline AxisCoord compute_vis_coord(
  int g_size, int oversampling, int cf_radius, double coord,
  double inv_lambda, double grid_scale) {
  const double position = grid_scale * coord * inv_lambda + g_size / 2.0;
  long grid_coord = std::lrint(position);
  const long fine_offset = std::lrint((grid_coord - position) * oversampling);
  grid_coord -= cf_radius;
  AxisCoord r;
  r.grid_coord = grid_coord;
  r.fine_offset = fine_offset;
  if (fine_offset >= 0) {
    r.cf_minor = fine_offset;
    r.cf_major = cf_padding;
  } else {
    r.cf_minor = oversampling + fine_offset;
    r.cf_major = cf_padding - 1;
  }
  return r;
}

struct Oracle {
  // grid_size in logical order x, y, mrow, cube
  unsigned gsize[4];
  double gscale[2];
  unsigned npol;
  int M[4][4], Mc[4][4];
  // grid, weights and model; grid and model Layout::Left (x fastest)
  std::vector<cf64> grid;
  std::vector<double> weights;  // [cube][mrow] (mrow fastest)
  std::vector<cf64> model;
  // convolution function: per group, Layout::Left over (x, y, mueller, cube)
  unsigned os = 0;
  std::vector<std::array<unsigned, 4>> cf_ext;
  std::vector<std::array<int, 2>> cf_radii;
  std::vector<std::vector<cf32>> cf;

  size_t plane_size() const { return size_t(gsize[0]) * gsize[1]; }
  size_t gidx(long x, long y, unsigned mrow, unsigned cube) const {
    return ((size_t(cube) * gsize[2] + mrow) * gsize[1] + size_t(y)) * gsize[0]
      + size_t(x);
  }
  // CF element through the reference's 6-D device layout
  // cf(X_major, Y_major, mueller, cube, x_minor, y_minor)
  // (src/hpg_impl.hpp:1035-1046): index (major * os + minor) on each axis.
  // Elements beyond the CFArray extents read as zero.
  cf32 cf_at(unsigned grp, long i, long j, int mueller, unsigned cube) const {
    const auto& e = cf_ext[grp];
    if (i < 0 || j < 0 || i >= long(e[0]) || j >= long(e[1])
        || mueller < 0 || unsigned(mueller) >= e[2] || cube >= e[3])
      return {0.0f, 0.0f};
    return cf[grp][((size_t(cube) * e[2] + unsigned(mueller)) * e[1] + j) * e[0]
                   + i];
  }
};

// Vis<N> constructor (src/hpg_core.hpp:493-550)
struct Vis {
  long grid_coord[2];
  long cf_major[2];
  long cf_minor[2];
  long fine_offset[2];
  int cf_size[2];
  cf64 phasor;
  unsigned grid_cube, cf_cube, cf_grp;
  bool pos_w;
  double phi0[2];
  double dphi[2];
};

template <unsigned N>
inline Vis make_vis(const Oracle& o, const VisData<N>& v) {
  Vis r;
  r.phasor = cphase(v.phase);
  r.grid_cube = v.grid_cube;
  r.cf_cube = v.cf_index[0];
  r.cf_grp = v.cf_index[1];
  r.pos_w = v.uvw[2] > 0;
  const double inv_lambda = v.frequency / c_light;
  for (int d = 0; d < 2; ++d) {
    r.cf_size[d] = 2 * o.cf_radii[r.cf_grp][d] + 1;
    auto c = compute_vis_coord(
      int(o.gsize[d]), int(o.os), r.cf_size[d] / 2, v.uvw[d], inv_lambda,
      o.gscale[d]);
    r.grid_coord[d] = int(c.grid_coord);
    r.cf_major[d] = int(c.cf_major);
    r.cf_minor[d] = int(c.cf_minor);
    r.fine_offset[d] = int(c.fine_offset);
    r.phi0[d] =
      -v.cf_phase_gradient[d]
      * ((r.cf_size[d] / 2) * int(o.os) - int(r.fine_offset[d]));
    r.dphi[d] = v.cf_phase_gradient[d] * int(o.os);
  }
  return r;
}

// all_within_grid (src/hpg_core.hpp:1070-1080)
inline bool all_within_grid(const Oracle& o, const Vis& v) {
  return 0 <= v.grid_coord[0] && v.grid_coord[0] + v.cf_size[0] <= long(o.gsize[0])
    && 0 <= v.grid_coord[1] && v.grid_coord[1] + v.cf_size[1] <= long(o.gsize[1]);
}

// records with an out-of-range CF group / cube or grid cube are undefined in
// the reference; the oracle skips them, as the HIP device does.
template <unsigned N>
inline bool valid_indexes(const Oracle& o, const VisData<N>& v) {
  return v.cf_index[1] < o.cf_ext.size()
    && v.cf_index[0] < o.cf_ext[v.cf_index[1]][3]
    && v.grid_cube < o.gsize[3];
}

// grid_vis / grid_vis_no_weights for one (vis, gpol) in the Serial order
// (src/hpg_core.hpp:791-886, 995-1068); with `sincos_always` false this is
// the oracle, with atomics it is the OpenMP baseline's task body.
template <unsigned N, bool Atomic>
void grid_one(
  Oracle& o, const Vis& vis, const cf32* gvis, const float* wts,
  unsigned gpol, bool update_weights) {

  const int NX = vis.cf_size[0];
  const int NY = vis.cf_size[1];
  const int* mindex = vis.pos_w ? o.M[gpol] : o.Mc[gpol];
  const float im_factor = vis.pos_w ? -1.0f : 1.0f;
  std::vector<double> phi_Y(NY);
  for (int Y = 0; Y < NY; ++Y) phi_Y[Y] = vis.phi0[1] + Y * vis.dphi[1];

  cf64 grid_wgt[N];
  for (unsigned v = 0; v < N; ++v) grid_wgt[v] = {0.0, 0.0};
  for (int X = 0; X < NX; ++X) {
    const double phi_X = vis.phi0[0] + X * vis.dphi[0];
    for (int Y = 0; Y < NY; ++Y) {
      const cf32 screen = narrow(cphase(phi_X + phi_Y[Y]));
      cf64 gv{0.0, 0.0};
      for (unsigned v = 0; v < N; ++v) {
        const int m = mindex[v];
        if (m >= 0) {
          cf32 cfv = o.cf_at(
            vis.cf_grp,
            (vis.cf_major[0] + X) * long(o.os) + vis.cf_minor[0],
            (vis.cf_major[1] + Y) * long(o.os) + vis.cf_minor[1],
            m, vis.cf_cube);
          cfv.im *= im_factor;
          const cf64 p = widen(mul(mul(cfv, screen), gvis[v]));
          gv.re += p.re;
          gv.im += p.im;
          if (update_weights) {
            grid_wgt[v].re += double(cfv.re);
            grid_wgt[v].im += double(cfv.im);
          }
        }
      }
      cf64& g = o.grid[o.gidx(
        vis.grid_coord[0] + X, vis.grid_coord[1] + Y, gpol, vis.grid_cube)];
      if (Atomic) {
#pragma omp atomic
        g.re += gv.re;
#pragma omp atomic
        g.im += gv.im;
      } else {
        g.re += gv.re;
        g.im += gv.im;
      }
    }
  }
  if (update_weights) {
    double twgt = 0;
    for (unsigned v = 0; v < N; ++v)
      twgt += std::hypot(grid_wgt[v].re, grid_wgt[v].im) * double(wts[v]);
    double& w = o.weights[size_t(vis.grid_cube) * o.gsize[2] + gpol];
    if (Atomic) {
#pragma omp atomic
      w += twgt;
    } else {
      w += twgt;
    }
  }
}

// degrid_vis (src/hpg_core.hpp:579-682), Serial summation order: for each
// gpol an X-outer/Y-inner partial sum, then added into the running total.
template <unsigned N>
void degrid_one(const Oracle& o, const Vis& vis, cf32* result) {
  for (unsigned v = 0; v < N; ++v) result[v] = {0.0f, 0.0f};
  if (o.model.empty()) return;
  const int NX = vis.cf_size[0];
  const int NY = vis.cf_size[1];
  const unsigned NR = o.gsize[2];
  const float im_factor = vis.pos_w ? 1.0f : -1.0f;
  std::vector<double> phi_Y(NY);
  for (int Y = 0; Y < NY; ++Y) phi_Y[Y] = vis.phi0[1] + Y * vis.dphi[1];

  cf64 tot_vis[N], tot_wgt[N];
  for (unsigned v = 0; v < N; ++v) tot_vis[v] = tot_wgt[v] = {0.0, 0.0};
  for (unsigned gpol = 0; gpol < NR; ++gpol) {
    const int* mindex = vis.pos_w ? o.Mc[gpol] : o.M[gpol];
    cf64 va_vis[N], va_wgt[N];
    for (unsigned v = 0; v < N; ++v) va_vis[v] = va_wgt[v] = {0.0, 0.0};
    for (int X = 0; X < NX; ++X) {
      const double phi_X = vis.phi0[0] + X * vis.dphi[0];
      for (int Y = 0; Y < NY; ++Y) {
        const cf64 screen = cphase(-phi_X - phi_Y[Y]);
        const cf64 mv = mul(
          o.model[o.gidx(
            vis.grid_coord[0] + X, vis.grid_coord[1] + Y, gpol, vis.grid_cube)],
          screen);
        for (unsigned v = 0; v < N; ++v) {
          const int m = mindex[v];
          if (m >= 0) {
            cf32 cfv = o.cf_at(
              vis.cf_grp,
              (vis.cf_major[0] + X) * long(o.os) + vis.cf_minor[0],
              (vis.cf_major[1] + Y) * long(o.os) + vis.cf_minor[1],
              m, vis.cf_cube);
            cfv.im *= im_factor;
            const cf64 p = mul(cfv, mv);
            va_vis[v].re += p.re;
            va_vis[v].im += p.im;
            va_wgt[v].re += double(cfv.re);
            va_wgt[v].im += double(cfv.im);
          }
        }
      }
    }
    for (unsigned v = 0; v < N; ++v) {
      tot_vis[v].re += va_vis[v].re;
      tot_vis[v].im += va_vis[v].im;
      tot_wgt[v].re += va_wgt[v].re;
      tot_wgt[v].im += va_wgt[v].im;
    }
  }
  const cf64 conj_phasor{vis.phasor.re, -vis.phasor.im};
  for (unsigned v = 0; v < N; ++v) {
    const cf64 w =
      (tot_wgt[v].re != 0.0 || tot_wgt[v].im != 0.0) ? tot_wgt[v] : cf64{1.0, 0.0};
    result[v] = narrow(mul(kokkos_div(tot_vis[v], w), conj_phasor));
  }
}

// gvis = fl32(vis * phasor * weight) (src/hpg_core.hpp:1172-1182, 1153-1159)
inline cf32 prephase(cf32 vis, cf64 phasor, float weight) {
  cf64 t = mul(vis, phasor);
  return narrow({t.re * double(weight), t.im * double(weight)});
}

// VisibilityGridder<N, exec, 0>::kernel orchestration
// (src/hpg_core.hpp:1082-1253).  `vis` is modified in place exactly as the
// reference modifies its device copy of the records (residual or predicted
// values for records inside the grid; others untouched).
template <unsigned N>
int grid_visibilities(
  Oracle& o, VisData<N>* vis, size_t n, bool update_weights, bool do_degrid,
  bool do_grid, bool openmp) {

  std::vector<std::array<cf32, N>> gvis(n);
  if (do_degrid) {
    for (size_t i = 0; i < n; ++i) {
      if (!valid_indexes(o, vis[i])) continue;
      Vis v = make_vis(o, vis[i]);
      if (!all_within_grid(o, v)) continue;
      cf32 pred[N];
      degrid_one<N>(o, v, pred);
      for (unsigned p = 0; p < N; ++p) {
        cf32 val{vis[i].vis[p].real(), vis[i].vis[p].imag()};
        if (do_grid) {
          val.re -= pred[p].re;
          val.im -= pred[p].im;
          vis[i].vis[p] = std::complex<float>(val.re, val.im);
          gvis[i][p] = prephase(val, v.phasor, vis[i].weights[p]);
        } else {
          vis[i].vis[p] = std::complex<float>(pred[p].re, pred[p].im);
        }
      }
    }
  } else {
    for (size_t i = 0; i < n; ++i) {
      const cf64 phasor = cphase(vis[i].phase);
      for (unsigned p = 0; p < N; ++p)
        gvis[i][p] = prephase(
          {vis[i].vis[p].real(), vis[i].vis[p].imag()}, phasor,
          vis[i].weights[p]);
    }
  }
  if (!do_grid) return 0;
  const long NR = o.gsize[2];
  const long ntasks = long(n) * NR;
  if (openmp) {
#pragma omp parallel for schedule(dynamic, 16)
    for (long rank = 0; rank < ntasks; ++rank) {
      const size_t i = size_t(rank / NR);
      const unsigned gpol = unsigned(rank % NR);
      if (!valid_indexes(o, vis[i])) continue;
      Vis v = make_vis(o, vis[i]);
      if (all_within_grid(o, v))
        grid_one<N, true>(
          o, v, gvis[i].data(), vis[i].weights.data(), gpol, update_weights);
    }
  } else {
    for (long rank = 0; rank < ntasks; ++rank) {
      const size_t i = size_t(rank / NR);
      const unsigned gpol = unsigned(rank % NR);
      if (!valid_indexes(o, vis[i])) continue;
      Vis v = make_vis(o, vis[i]);
      if (all_within_grid(o, v))
        grid_one<N, false>(
          o, v, gvis[i].data(), vis[i].weights.data(), gpol, update_weights);
    }
  }
  return 0;
}

// ---- FFT (src/hpg_core.hpp:2029-2131, 2192-2276): an unnormalized 2-D DFT of
// every (mrow, cube) plane over (x, y); NEGATIVE = exp(-i...), POSITIVE =
// exp(+i...).  Radix-2 for power-of-two lengths, direct DFT otherwise.
// (The reference's host FFT addresses x with stride Ny for Nmrow > 1; this
// follows the mathematics -- independent planes -- like its CUDA path.)

void fft1d(cf64* a, size_t n, size_t stride, int sign, std::vector<cf64>& tmp) {
  tmp.resize(n);
  for (size_t k = 0; k < n; ++k) tmp[k] = a[k * stride];
  if (n > 1 && (n & (n - 1)) == 0) {
    // bit reversal
    for (size_t i = 1, j = 0; i < n; ++i) {
      size_t bit = n >> 1;
      for (; j & bit; bit >>= 1) j ^= bit;
      j ^= bit;
      if (i < j) std::swap(tmp[i], tmp[j]);
    }
    for (size_t len = 2; len <= n; len <<= 1) {
      const size_t half = len / 2;
      for (size_t k = 0; k < half; ++k) {
        const double ang = sign * 2.0 * M_PI * double(k) / double(len);
        const cf64 w{std::cos(ang), std::sin(ang)};
        for (size_t s = 0; s < n; s += len) {
          const cf64 u = tmp[s + k];
          const cf64 t = mul(tmp[s + k + half], w);
          tmp[s + k] = {u.re + t.re, u.im + t.im};
          tmp[s + k + half] = {u.re - t.re, u.im - t.im};
        }
      }
    }
    for (size_t k = 0; k < n; ++k) a[k * stride] = tmp[k];
  } else {
    for (size_t k = 0; k < n; ++k) {
      cf64 s{0.0, 0.0};
      for (size_t j = 0; j < n; ++j) {
        const double ang =
          sign * 2.0 * M_PI * double((k * j) % n) / double(n);
        const cf64 t = mul(tmp[j], cf64{std::cos(ang), std::sin(ang)});
        s.re += t.re;
        s.im += t.im;
      }
      a[k * stride] = s;
    }
  }
}

void fft_planes(Oracle& o, std::vector<cf64>& g, double norm, int sign) {
  const size_t nx = o.gsize[0], ny = o.gsize[1];
  const size_t nplanes = size_t(o.gsize[2]) * o.gsize[3];
  std::vector<cf64> tmp;
  for (size_t p = 0; p < nplanes; ++p) {
    cf64* plane = g.data() + p * nx * ny;
    for (size_t y = 0; y < ny; ++y) fft1d(plane + y * nx, nx, 1, sign, tmp);
    for (size_t x = 0; x < nx; ++x) fft1d(plane + x, ny, nx, sign, tmp);
  }
  // GridNormalizer with norm (src/hpg_core.hpp:1863-1889)
  if (norm != 1) {
    const double inv_norm = 1.0 / norm;
    for (auto& v : g) {
      v.re *= inv_norm;
      v.im *= inv_norm;
    }
  }
}

// GridShifter (src/hpg_core.hpp:2303-2542): FORWARD moves element (x, y) to
// ((x + Nx/2) mod Nx, (y + Ny/2) mod Ny), i.e. numpy.fft.fftshift on (x, y);
// BACKWARD is the inverse (numpy.fft.ifftshift).
void shift_planes(Oracle& o, std::vector<cf64>& g, int backward) {
  const size_t nx = o.gsize[0], ny = o.gsize[1];
  const size_t nplanes = size_t(o.gsize[2]) * o.gsize[3];
  const size_t sx = backward ? nx - nx / 2 : nx / 2;
  const size_t sy = backward ? ny - ny / 2 : ny / 2;
  std::vector<cf64> tmp(nx * ny);
  for (size_t p = 0; p < nplanes; ++p) {
    cf64* plane = g.data() + p * nx * ny;
    for (size_t y = 0; y < ny; ++y)
      for (size_t x = 0; x < nx; ++x)
        tmp[((y + sy) % ny) * nx + (x + sx) % nx] = plane[y * nx + x];
    std::memcpy(plane, tmp.data(), sizeof(cf64) * nx * ny);
  }
}

}  // namespace

// ---------------------------------------------------------------------------
// C ABI used through ctypes by the tests and the CPU-baseline leg of bench.py

extern "C" {

void* hpgo_create(
  const unsigned* grid_size, const double* grid_scale, unsigned npol,
  const int* mueller, const int* conj_mueller) {
  auto* o = new Oracle();
  std::memcpy(o->gsize, grid_size, sizeof(o->gsize));
  o->gscale[0] = grid_scale[0];
  o->gscale[1] = grid_scale[1];
  o->npol = npol;
  // Mueller index matrices padded to 4x4 with -1
  // (copy_mueller_indexes, src/hpg_runtime.hpp:1590-1614)
  for (int r = 0; r < 4; ++r)
    for (int c = 0; c < 4; ++c) {
      const bool in = unsigned(r) < grid_size[2] && unsigned(c) < npol;
      o->M[r][c] = in ? mueller[r * npol + c] : -1;
      o->Mc[r][c] = in ? conj_mueller[r * npol + c] : -1;
    }
  o->grid.assign(size_t(grid_size[0]) * grid_size[1] * grid_size[2] * grid_size[3],
                 cf64{0.0, 0.0});
  o->weights.assign(size_t(grid_size[2]) * grid_size[3], 0.0);
  return o;
}

void hpgo_destroy(void* h) { delete static_cast<Oracle*>(h); }

// values[g]: complex<float> array, Layout::Left over (x, y, mueller, cube)
void hpgo_set_cf(
  void* h, unsigned oversampling, unsigned ngroups, const unsigned* extents,
  const float* const* values) {
  auto* o = static_cast<Oracle*>(h);
  o->os = oversampling;
  o->cf_ext.clear();
  o->cf_radii.clear();
  o->cf.clear();
  for (unsigned g = 0; g < ngroups; ++g) {
    std::array<unsigned, 4> e{extents[4 * g], extents[4 * g + 1],
                              extents[4 * g + 2], extents[4 * g + 3]};
    o->cf_ext.push_back(e);
    // CFArray::radii (src/hpg.hpp:687-694)
    o->cf_radii.push_back(
      {int(((e[0] - 2 * cf_padding * oversampling) / oversampling) / 2),
       int(((e[1] - 2 * cf_padding * oversampling) / oversampling) / 2)});
    const size_t n = size_t(e[0]) * e[1] * e[2] * e[3];
    std::vector<cf32> v(n);
    std::memcpy(v.data(), values[g], n * sizeof(cf32));
    o->cf.push_back(std::move(v));
  }
}

// model: complex<double>, Layout::Left over (x, y, mrow, cube); null clears it
void hpgo_set_model(void* h, const double* values) {
  auto* o = static_cast<Oracle*>(h);
  if (!values) {
    o->model.clear();
    return;
  }
  o->model.resize(o->grid.size());
  std::memcpy(o->model.data(), values, o->model.size() * sizeof(cf64));
}

// Gridding / degridding of `n` VisData<npol> records, modified in place.
// openmp != 0 runs the OpenMP baseline form (atomics; order-dependent).
int hpgo_grid_visibilities(
  void* h, void* records, size_t n, unsigned npol, int update_weights,
  int do_degrid, int do_grid, int openmp) {
  auto* o = static_cast<Oracle*>(h);
  if (npol != o->npol) return -1;
  switch (npol) {
  case 1:
    return grid_visibilities<1>(*o, static_cast<VisData<1>*>(records), n,
                                update_weights, do_degrid, do_grid, openmp);
  case 2:
    return grid_visibilities<2>(*o, static_cast<VisData<2>*>(records), n,
                                update_weights, do_degrid, do_grid, openmp);
  case 3:
    return grid_visibilities<3>(*o, static_cast<VisData<3>*>(records), n,
                                update_weights, do_degrid, do_grid, openmp);
  case 4:
    return grid_visibilities<4>(*o, static_cast<VisData<4>*>(records), n,
                                update_weights, do_degrid, do_grid, openmp);
  default:
    return -1;
  }
}

// GridNormalizer with weights (src/hpg_core.hpp:1832-1861)
void hpgo_normalize_by_weights(void* h, double wfactor) {
  auto* o = static_cast<Oracle*>(h);
  const size_t ps = o->plane_size();
  for (unsigned cube = 0; cube < o->gsize[3]; ++cube)
    for (unsigned mrow = 0; mrow < o->gsize[2]; ++mrow) {
      const double d = wfactor * o->weights[size_t(cube) * o->gsize[2] + mrow];
      cf64* p = o->grid.data() + (size_t(cube) * o->gsize[2] + mrow) * ps;
      for (size_t k = 0; k < ps; ++k) {
        p[k].re /= d;
        p[k].im /= d;
      }
    }
}

// which: 0 = grid, 1 = model; sign: 0 = POSITIVE, 1 = NEGATIVE
void hpgo_fft(void* h, int which, double norm, int sign) {
  auto* o = static_cast<Oracle*>(h);
  auto& g = which ? o->model : o->grid;
  if (g.empty()) return;
  fft_planes(*o, g, norm, sign == 1 ? -1 : +1);
}

// direction: 0 = FORWARD, 1 = BACKWARD
void hpgo_shift(void* h, int which, int direction) {
  auto* o = static_cast<Oracle*>(h);
  auto& g = which ? o->model : o->grid;
  if (g.empty()) return;
  shift_planes(*o, g, direction);
}

void hpgo_reset_grid(void* h) {
  auto* o = static_cast<Oracle*>(h);
  std::fill(o->grid.begin(), o->grid.end(), cf64{0.0, 0.0});
  std::fill(o->weights.begin(), o->weights.end(), 0.0);
}

// copies the grid (Layout::Left, complex<double>) and weights ([cube][mrow])
void hpgo_grid_values(void* h, double* dst) {
  auto* o = static_cast<Oracle*>(h);
  std::memcpy(dst, o->grid.data(), o->grid.size() * sizeof(cf64));
}
void hpgo_grid_weights(void* h, double* dst) {
  auto* o = static_cast<Oracle*>(h);
  std::memcpy(dst, o->weights.data(), o->weights.size() * sizeof(double));
}
// copies the window [x0, x0+nx) x [y0, y0+ny) of one grid plane, x fastest
void hpgo_grid_window(
  void* h, unsigned x0, unsigned y0, unsigned nx, unsigned ny, unsigned mrow,
  unsigned cube, double* dst) {
  auto* o = static_cast<Oracle*>(h);
  for (unsigned y = 0; y < ny; ++y)
    std::memcpy(dst + 2 * size_t(y) * nx,
                o->grid.data() + o->gidx(x0, y0 + y, mrow, cube),
                size_t(nx) * sizeof(cf64));
}
void hpgo_model_values(void* h, double* dst) {
  auto* o = static_cast<Oracle*>(h);
  if (!o->model.empty())
    std::memcpy(dst, o->model.data(), o->model.size() * sizeof(cf64));
}

// ConeCFArray::operator() of the reference's tests
// (tests/GridderStateUnitTests.cc:198-253), restated with the same float
// arithmetic: polar((mueller+1) * max(R - |p|, 0), arg(|p.re| + i|p.im|)),
// p = (x - R + 0.5, y - R + 0.5), R = (radius + padding) * oversampling.
// Fills a Layout::Left (x, y, mueller, cube=1) array of extent 2R+1.
void hpgo_cone_cf(unsigned nmueller, unsigned oversampling, unsigned radius,
                  float* dst) {
  const int R = int((radius + cf_padding) * oversampling);
  const unsigned w = unsigned(2 * R + 1);
  for (unsigned m = 0; m < nmueller; ++m)
    for (unsigned y = 0; y < w; ++y)
      for (unsigned x = 0; x < w; ++x) {
        const std::complex<float> p(
          (-R + int(x)) + 0.5f, (-R + int(y)) + 0.5f);
        const std::complex<float> q =
          std::abs(p.real()) + std::complex<float>(0.0f, 1.0f) * std::abs(p.imag());
        const std::complex<float> v = std::polar(
          (m + 1) * std::max(R - std::abs(p), 0.0f), std::arg(q));
        const size_t k = (size_t(m) * w + y) * w + x;
        dst[2 * k] = v.real();
        dst[2 * k + 1] = v.imag();
      }
}

int hpgo_num_threads() {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

}  // extern "C"
