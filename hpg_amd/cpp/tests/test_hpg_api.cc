// Tests of the hpg C++ API (hpg_amd/cpp) on the HIP device, restating the
// reference's GridderState / Gridder unit tests
// (tests/GridderStateUnitTests.cc, tests/GridderUnitTests.cc) with a
// minimal harness.  Usage: test_hpg_api <path of tests/golden/gridone.bin>
#include "hpg/hpg.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <functional>
#include <iostream>
#include <random>
#include <string>
#include <vector>

namespace {

int g_failures = 0;
int g_checks = 0;

#define CHECK(cond)                                                        \
  do {                                                                     \
    ++g_checks;                                                            \
    if (!(cond)) {                                                         \
      ++g_failures;                                                        \
      std::fprintf(stderr, "  FAILED %s:%d: %s\n", __FILE__, __LINE__, #cond); \
    }                                                                      \
  } while (0)

const hpg::Device dev = hpg::Device::Hip;
const hpg::Device host = hpg::Device::Serial;

// the reference tests' analytic cone CF (GridderStateUnitTests.cc:198-253):
// polar((m+1) * max(R - |p|, 0), arg(|p.re| + i|p.im|)), R = (r + pad) * os
struct ConeCF final : public hpg::CFArray {
  int m_nmueller, m_os, m_radius, m_R;
  ConeCF(unsigned nmueller, unsigned os, unsigned radius)
    : m_nmueller(nmueller), m_os(os), m_radius(radius), m_R((radius + padding) * os) {}
  unsigned oversampling() const override { return m_os; }
  unsigned num_groups() const override { return 1; }
  std::array<unsigned, 4> extents(unsigned) const override {
    unsigned w = 2 * m_R + 1;
    return {w, w, unsigned(m_nmueller), 1};
  }
  std::complex<float> operator()(unsigned x, unsigned y, unsigned mueller, unsigned,
                                 unsigned) const override {
    const std::complex<float> p((-m_R + int(x)) + 0.5f, (-m_R + int(y)) + 0.5f);
    const std::complex<float> q =
      std::abs(p.real()) + std::complex<float>(0.0f, 1.0f) * std::abs(p.imag());
    return std::polar((mueller + 1) * std::max(m_R - std::abs(p), 0.0f), std::arg(q));
  }
};

// random CF, uniform taps in [-1, 1]^2 (MyCFArray, GridderStateUnitTests.cc:148-196)
struct RandomCF final : public hpg::CFArray {
  unsigned m_os;
  std::vector<std::array<unsigned, 4>> m_ext;
  std::vector<std::vector<std::complex<float>>> m_v;
  RandomCF(unsigned os, const std::vector<std::array<unsigned, 4>>& sizes, unsigned seed)
    : m_os(os) {
    std::mt19937 rng(seed);
    std::uniform_real_distribution<float> d(-1, 1);
    for (auto& s : sizes) {
      m_ext.push_back({s[0] * os, s[1] * os, s[2], s[3]});
      std::vector<std::complex<float>> v(size_t(s[0]) * os * s[1] * os * s[2] * s[3]);
      for (auto& x : v) x = {d(rng), d(rng)};
      m_v.push_back(std::move(v));
    }
  }
  unsigned oversampling() const override { return m_os; }
  unsigned num_groups() const override { return unsigned(m_ext.size()); }
  std::array<unsigned, 4> extents(unsigned g) const override { return m_ext[g]; }
  std::complex<float> operator()(unsigned x, unsigned y, unsigned m, unsigned c,
                                 unsigned g) const override {
    const auto& e = m_ext[g];
    return m_v[g][((size_t(c) * e[2] + m) * e[1] + y) * e[0] + x];
  }
};

struct ConstGrid final : public hpg::GridValueArray {
  std::array<unsigned, 4> m_e;
  value_type m_value{0};
  explicit ConstGrid(const std::array<unsigned, 4>& e) : m_e(e) {}
  unsigned extent(unsigned d) const override { return m_e[d]; }
  const value_type& operator()(unsigned, unsigned, unsigned, unsigned) const override {
    return m_value;
  }
  value_type& operator()(unsigned, unsigned, unsigned, unsigned) override { return m_value; }
};

void run(const char* name, const std::function<void()>& f) {
  const int before = g_failures;
  std::printf("%s ...\n", name);
  std::fflush(stdout);
  f();
  std::printf("%s %s\n", name, g_failures == before ? "ok" : "FAILED");
  std::fflush(stdout);
}

bool any_nonzero(const hpg::GridValueArray& g) {
  for (unsigned c = 0; c < g.extent(3); ++c)
    for (unsigned r = 0; r < g.extent(2); ++r)
      for (unsigned y = 0; y < g.extent(1); ++y)
        for (unsigned x = 0; x < g.extent(0); ++x)
          if (g(x, y, r, c) != hpg::GridValueArray::value_type(0)) return true;
  return false;
}

void test_constructor_args() {
  const std::array<unsigned, 4> gs{10, 12, 4, 3};
  const std::array<double, 2> sc{0.12, -0.34};
  RandomCF cf(10, {{3 + 4, 3 + 4, 4, 1}}, 1);
  auto rv = hpg::GridderState::create<1>(dev, 0, 15, &cf, gs, sc, {{0}, {1}, {2}, {3}},
                                         {{3}, {2}, {1}, {0}});
  CHECK(hpg::is_value(rv));
  if (!hpg::is_value(rv)) return;
  auto g = hpg::get_value(std::move(rv));
  CHECK(g.device() == dev);
  CHECK(g.grid_size() == gs);
  CHECK(g.grid_scale() == sc);
  CHECK(g.num_polarizations() == 1);
  CHECK(g.max_visibility_batch_size() == 15);
  CHECK(!g.is_null());
  // InvalidNumberMuellerIndexRows (GridderStateUnitTests.cc:582-645)
  auto bad = hpg::GridderState::create<1>(dev, 0, 15, &cf, gs, sc, {{0}, {1}}, {{0}, {1}});
  CHECK(hpg::is_error(bad) &&
        hpg::get_error(bad).type() == hpg::ErrorType::InvalidNumberMuellerIndexRows);
  auto nodev = hpg::GridderState::create<1>(hpg::Device::Cuda, 0, 15, &cf, {10, 12, 1, 1}, sc,
                                            {{0}}, {{0}});
  CHECK(hpg::is_error(nodev) && hpg::get_error(nodev).type() == hpg::ErrorType::DisabledDevice);
  // initial values are zero (InitValues, 739-767)
  auto [g1, v] = std::move(g).grid_values();
  auto [g2, w] = std::move(g1).grid_weights();
  CHECK(!any_nonzero(*v));
  CHECK(w->extent(0) == 4 && w->extent(1) == 3);
}

void test_gridone(const std::string& golden_path) {
  // GridOne (GridderStateUnitTests.cc:1343-1463)
  const std::array<unsigned, 4> gsz{16384, 16384, 1, 1};
  const std::array<double, 2> sc{0.0476591, 0.0476591};
  ConeCF cf(1, 20, 45);
  const hpg::vis_uvw_t uvw{2344.1, 638.066, -1826.55};
  auto result =
    hpg::RvalM<void, hpg::GridderState>::pure([&]() {
      return hpg::GridderState::create<1>(dev, 0, 1, &cf, gsz, sc, {{0}}, {{0}});
    })
      .and_then([&](hpg::GridderState&& gs) {
        return std::move(gs).set_convolution_function(host, ConeCF(cf));
      })
      .and_then([&](hpg::GridderState&& gs) {
        return std::move(gs).grid_visibilities(
          host, std::vector<hpg::VisData<1>>{hpg::VisData<1>(
                  {std::complex<float>(1, -1)}, {1.0f}, 3.693e9, 0.0, uvw, 0, {0, 0})},
          true);
      })
      .map([](hpg::GridderState&& gs) {
        auto [gs1, gv] = std::move(gs).grid_values();
        return std::move(gv);
      })();
  CHECK(hpg::is_value(result));
  if (!hpg::is_value(result)) return;
  auto g = hpg::get_value(std::move(result));
  std::vector<double> gold(91 * 91 * 2);
  std::ifstream f(golden_path, std::ios::binary);
  f.read(reinterpret_cast<char*>(gold.data()), gold.size() * sizeof(double));
  CHECK(bool(f));
  int exact = 0, close = 0, nz_ok = 0;
  for (int x = 0; x < 91; ++x)
    for (int y = 0; y < 91; ++y) {
      const std::complex<double> e(gold[2 * (x * 91 + y)], gold[2 * (x * 91 + y) + 1]);
      const auto v = (*g)(9523 + x, 8522 + y, 0, 0);
      nz_ok += (v != 0.0) == (e != 0.0);
      if (x < 90 && y < 90) {
        exact += v == e;
        close += std::abs(v - e) <= 3e-7 * std::abs(e);
      }
    }
  CHECK(nz_ok == 91 * 91);
  CHECK(exact > 0.98 * 8100);
  CHECK(close == 8100);
  // nothing outside the footprint
  CHECK((*g)(9522, 8560, 0, 0) == 0.0 && (*g)(9614, 8560, 0, 0) == 0.0);
  CHECK((*g)(9560, 8521, 0, 0) == 0.0 && (*g)(9560, 8613, 0, 0) == 0.0);
}

void test_gridding_sweep() {
  // Gridding (GridderStateUnitTests.cc:1273-1341): exact float products at
  // the closed-form CF taps, including the lrint tie 7.5 -> 8
  const std::array<unsigned, 4> gsz{14, 14, 1, 1};
  const std::array<double, 2> sc{0.1, 0.1};
  const int os = 10, r = 3;
  ConeCF cf(1, os, r);
  auto gs = hpg::get_value(hpg::GridderState::create<1>(dev, 0, 1, &cf, gsz, sc, {{0}}, {{0}}));
  gs = hpg::get_value(std::move(gs).set_convolution_function(host, ConeCF(cf)));
  const std::complex<float> vis(1, -1);
  int bad = 0;
  for (int i = 0; i <= os; ++i)
    for (int j = 0; j <= os; ++j) {
      const hpg::vis_uvw_t uvw{i * 1.0, j * 1.0, 0.0};
      auto g1 = hpg::get_value(
        std::move(gs).reset_grid().grid_visibilities(
          host, std::vector<hpg::VisData<1>>{
                  hpg::VisData<1>({vis}, {1.0f}, 299792458.0, 0.0, uvw, 0, {0, 0})}));
      auto [g2, v] = std::move(g1).grid_values();
      gs = std::move(g2);
      long gc[2], maj[2], mn[2];
      for (int d = 0; d < 2; ++d) {
        const double pos = sc[d] * uvw[d] * 1.0 + gsz[d] / 2.0;
        const long G = std::lrint(pos);
        const long fo = std::lrint((G - pos) * os);
        gc[d] = G - r;
        maj[d] = fo >= 0 ? 2 : 1;
        mn[d] = fo >= 0 ? fo : os + fo;
      }
      for (int x = 0; x < 14; ++x)
        for (int y = 0; y < 14; ++y) {
          std::complex<double> e = 0;
          if (gc[0] <= x && x < gc[0] + 2 * r + 1 && gc[1] <= y && y < gc[1] + 2 * r + 1) {
            const auto c = cf((x - gc[0] + maj[0]) * os + mn[0], (y - gc[1] + maj[1]) * os + mn[1],
                              0, 0, 0);
            const float re = c.real() * vis.real() - c.imag() * vis.imag();
            const float im = c.real() * vis.imag() + c.imag() * vis.real();
            e = {re, im};
          }
          bad += (*v)(x, y, 0, 0) != e;
        }
    }
  CHECK(bad == 0);
}

void test_residual_visibilities() {
  // ResidualVisibilities (GridderStateUnitTests.cc:1743-1815)
  ConeCF cf(1, 10, 3);
  auto gs = hpg::get_value(hpg::GridderState::create<2>(
    dev, 0, 1, &cf, {20, 20, 2, 1}, {0.0476591, 0.0476591}, {{0, -1}, {-1, 0}},
    {{0, -1}, {-1, 0}}));
  gs = hpg::get_value(std::move(gs).set_convolution_function(host, ConeCF(cf)));
  const std::complex<float> vis(1, -1);
  std::vector<hpg::VisData<2>> v{hpg::VisData<2>(
    {vis, -vis}, {1.0f, 0.5f}, 3.693e9, 0.0, {0.0, 0.0, 0.1}, 0, {0, 0})};
  auto rv = gs.degrid_grid_get_residual_visibilities(host, std::vector<hpg::VisData<2>>(v));
  CHECK(hpg::is_value(rv));
  if (!hpg::is_value(rv)) return;
  auto [gs1, fv] = hpg::get_value(std::move(rv));
  CHECK(!fv.get());  // not resolved before a fence
  auto gs2 = std::move(gs1).fence();
  auto& orv = fv.get();
  CHECK(bool(orv));
  if (orv) CHECK(orv.value() == v);
  // the Gridder facade (GridderUnitTests.cc:490-563)
  auto gr = hpg::get_value(hpg::Gridder::create<2>(
    dev, 0, 1, &cf, {20, 20, 2, 1}, {0.0476591, 0.0476591}, {{0, -1}, {-1, 0}},
    {{0, -1}, {-1, 0}}));
  CHECK(!gr.set_convolution_function(host, ConeCF(cf)));
  auto fr = gr.degrid_grid_get_residual_visibilities(host, hpg::VisDataVector(v));
  CHECK(hpg::is_value(fr));
  gr.fence();
  if (hpg::is_value(fr)) {
    auto& o = hpg::get_value(fr).get();
    CHECK(bool(o) && o->get<2>() == v);
  }
}

void test_errors() {
  ConeCF cf(1, 10, 3);
  auto gs = hpg::get_value(
    hpg::GridderState::create<1>(dev, 0, 2, &cf, {20, 20, 1, 1}, {1.0, 1.0}, {{0}}, {{0}}));
  gs = hpg::get_value(std::move(gs).set_convolution_function(host, ConeCF(cf)));
  std::vector<hpg::VisData<1>> three(3);
  auto e1 = gs.grid_visibilities(host, std::vector<hpg::VisData<1>>(three));
  CHECK(hpg::is_error(e1) &&
        hpg::get_error(e1).type() == hpg::ErrorType::ExcessiveNumberVisibilities);
  auto e2 = gs.grid_visibilities(host, std::vector<hpg::VisData<2>>(1));
  CHECK(hpg::is_error(e2) &&
        hpg::get_error(e2).type() == hpg::ErrorType::InvalidNumberPolarizations);
  auto e3 = gs.grid_visibilities_base(host, hpg::VisDataVector(std::vector<hpg::VisData<1>>(1)),
                                      true, true, false, false);
  CHECK(hpg::is_error(e3) &&
        hpg::get_error(e3).type() == hpg::ErrorType::UpdateWeightsWithoutGridding);
  auto e4 = gs.grid_visibilities(hpg::Device::Cuda, std::vector<hpg::VisData<1>>(1));
  CHECK(hpg::is_error(e4) && hpg::get_error(e4).type() == hpg::ErrorType::DisabledHostDevice);
  // CFError (GridderStateUnitTests.cc:901-959)
  auto e5 = gs.set_convolution_function(host, ConeCF(1, 10, 12));
  CHECK(hpg::is_error(e5) && hpg::get_error(e5).type() == hpg::ErrorType::CFSupportExceedsGrid);
  // SetModel (772-808)
  auto e6 = gs.set_model(host, ConstGrid({20, 21, 1, 1}));
  CHECK(hpg::is_error(e6) && hpg::get_error(e6).type() == hpg::ErrorType::InvalidModelGridSize);
  // the target is unchanged by failed calls
  CHECK(!gs.is_null());
  auto ok = gs.grid_visibilities(host, std::vector<hpg::VisData<1>>(1));
  CHECK(hpg::is_value(ok));
}

void test_copy_semantics_and_serialization() {
  // Copies (650-734) and Serialization (1126-1218)
  RandomCF cf1(10, {{3 + 4, 3 + 4, 1, 1}}, 7), cf2(10, {{2 + 4, 2 + 4, 1, 1}}, 8);
  std::mt19937 rng(3);
  std::uniform_real_distribution<double> du(-30, 30);
  std::uniform_real_distribution<float> dv(-1, 1);
  auto vis = [&](unsigned n) {
    std::vector<hpg::VisData<1>> v;
    for (unsigned i = 0; i < n; ++i)
      v.emplace_back(std::array<std::complex<float>, 1>{std::complex<float>(dv(rng), dv(rng))},
                     std::array<float, 1>{0.5f + 0.5f * dv(rng)}, 299792458.0, 0.0,
                     hpg::vis_uvw_t{du(rng), du(rng), 0.0}, 0u, std::array<unsigned, 2>{0, 0});
    return v;
  };
  auto v1 = vis(500), v2 = vis(500);
  auto make = [&]() {
    return hpg::get_value(
      hpg::GridderState::create<1>(dev, 0, 500, &cf1, {80, 80, 1, 1}, {1.0, 1.0}, {{0}}, {{0}}));
  };
  auto a = hpg::get_value(
    hpg::get_value(
      hpg::get_value(
        hpg::get_value(make().set_convolution_function(host, RandomCF(cf1)))
          .grid_visibilities(host, std::vector<hpg::VisData<1>>(v1)))
        .set_convolution_function(host, RandomCF(cf2)))
      .grid_visibilities(host, std::vector<hpg::VisData<1>>(v2)));
  auto b = hpg::get_value(
    hpg::get_value(
      hpg::get_value(
        hpg::get_value(make().set_convolution_function(host, RandomCF(cf2)))
          .grid_visibilities(host, std::vector<hpg::VisData<1>>(v2)))
        .set_convolution_function(host, RandomCF(cf1)))
      .grid_visibilities(host, std::vector<hpg::VisData<1>>(v1)));
  auto [a1, ga] = a.grid_values();
  auto [b1, gb] = b.grid_values();
  double mx = 0, d = 0;
  for (unsigned x = 0; x < 80; ++x)
    for (unsigned y = 0; y < 80; ++y) {
      mx = std::max(mx, std::abs((*ga)(x, y, 0, 0)));
      d = std::max(d, std::abs((*ga)(x, y, 0, 0) - (*gb)(x, y, 0, 0)));
    }
  CHECK(mx > 0 && d <= 1e-12 * mx);
  // a copy does not see later changes of the original
  hpg::GridderState c = a;
  a = hpg::get_value(std::move(a).grid_visibilities(host, std::vector<hpg::VisData<1>>(v1)));
  auto [c1, gc] = c.grid_values();
  double dd = 0;
  for (unsigned x = 0; x < 80; ++x)
    for (unsigned y = 0; y < 80; ++y) dd = std::max(dd, std::abs((*gc)(x, y, 0, 0) - (*ga)(x, y, 0, 0)));
  CHECK(dd == 0);
}

void test_shift_cycle_and_fft() {
  // ShiftCycle (1860-1924) through the Gridder facade, OneFFT (1704-1740)
  for (auto sz : std::vector<std::array<unsigned, 2>>{{20, 20}, {19, 19}, {18, 19}, {19, 18}}) {
    RandomCF cf(10, {{3 + 4, 3 + 4, 1, 1}}, 42);
    auto g = hpg::get_value(hpg::Gridder::create<1>(dev, 0, 100, &cf, {sz[0], sz[1], 1, 2},
                                                    {0.0476591, 0.0476591}, {{0}}, {{0}}));
    CHECK(!g.set_convolution_function(host, RandomCF(cf)));
    std::mt19937 rng(5);
    std::uniform_real_distribution<double> du(-60, 60);
    std::vector<hpg::VisData<1>> v;
    for (int i = 0; i < 100; ++i)
      v.emplace_back(std::array<std::complex<float>, 1>{std::complex<float>(1, 2)},
                     std::array<float, 1>{1.0f}, 299792458.0 * 9.75719, 0.0,
                     hpg::vis_uvw_t{du(rng), du(rng), 0.0}, unsigned(i % 2),
                     std::array<unsigned, 2>{0, 0});
    CHECK(!g.grid_visibilities(host, std::move(v), false));
    auto g0 = g.grid_values();
    g.shift_grid(hpg::ShiftDirection::FORWARD);
    g.shift_grid(hpg::ShiftDirection::BACKWARD);
    auto g1 = g.grid_values();
    bool same = true;
    for (unsigned c = 0; c < 2; ++c)
      for (unsigned x = 0; x < sz[0]; ++x)
        for (unsigned y = 0; y < sz[1]; ++y) same &= (*g0)(x, y, 0, c) == (*g1)(x, y, 0, c);
    CHECK(same);
  }
  RandomCF cf(10, {{3 + 4, 3 + 4, 3, 3}}, 1);
  auto gs = hpg::get_value(hpg::GridderState::create<1>(dev, 0, 15, &cf, {2048, 2048, 1, 1},
                                                        {0.12, -0.34}, {{0}}, {{0}}));
  auto gs1 = hpg::get_value(std::move(gs).set_model(host, ConstGrid({2048, 2048, 1, 1})));
  auto [gs2, model] = std::move(gs1).model_values();
  (*model)(1024, 1108, 0, 0) = 0.825;
  auto gs3 = hpg::get_value(std::move(gs2).set_model(host, std::move(*model)));
  auto gs4 = hpg::get_value(std::move(gs3).apply_model_fft());
  auto gs5 = std::move(gs4).shift_model(hpg::ShiftDirection::FORWARD);
  auto [gs6, modelf] = std::move(gs5).model_values();
  CHECK((*modelf)(1070, 10, 0, 0).imag() < 0);
}

// CF whose tap (x, y, mueller) holds the value x + 1000 y + i mueller
// (EncodedCFArray, DeviceCFArrayUnitTests.cc:34-90)
struct EncodedCF final : public hpg::CFArray {
  unsigned m_nm, m_os, m_radius;
  EncodedCF(unsigned nm, unsigned os, unsigned radius) : m_nm(nm), m_os(os), m_radius(radius) {}
  unsigned oversampling() const override { return m_os; }
  unsigned num_groups() const override { return 1; }
  std::array<unsigned, 4> extents(unsigned) const override {
    const unsigned w = (2 * m_radius + 1 + 2 * padding) * m_os;
    return {w, w, m_nm, 1};
  }
  std::complex<float> operator()(unsigned x, unsigned y, unsigned m, unsigned,
                                 unsigned) const override {
    return {float(x + 1000 * y), float(m)};
  }
};

// records uniform over the part of a grid where supports of radius r fit
std::vector<hpg::VisData<1>> random_vis(unsigned n, const std::array<unsigned, 4>& gsz,
                                        const std::array<double, 2>& scale, unsigned r,
                                        unsigned ncube, unsigned seed) {
  std::mt19937 rng(seed);
  std::uniform_real_distribution<double> ux(-(gsz[0] / 2.0 - r - 2), gsz[0] / 2.0 - r - 2);
  std::uniform_real_distribution<double> uy(-(gsz[1] / 2.0 - r - 2), gsz[1] / 2.0 - r - 2);
  std::uniform_real_distribution<float> dv(-1, 1);
  std::vector<hpg::VisData<1>> v;
  for (unsigned i = 0; i < n; ++i)
    v.emplace_back(std::array<std::complex<float>, 1>{std::complex<float>(dv(rng), dv(rng))},
                   std::array<float, 1>{0.5f + 0.5f * dv(rng)}, 299792458.0, 0.0,
                   hpg::vis_uvw_t{ux(rng) / scale[0], uy(rng) / scale[1], 0.0},
                   unsigned(i % ncube), std::array<unsigned, 2>{0, 0});
  return v;
}

// max |a - b| / max |a| over every element
double rel_diff(const hpg::GridValueArray& a, const hpg::GridValueArray& b) {
  double mx = 0, d = 0;
  for (unsigned c = 0; c < a.extent(3); ++c)
    for (unsigned r = 0; r < a.extent(2); ++r)
      for (unsigned y = 0; y < a.extent(1); ++y)
        for (unsigned x = 0; x < a.extent(0); ++x) {
          mx = std::max(mx, std::abs(a(x, y, r, c)));
          d = std::max(d, std::abs(a(x, y, r, c) - b(x, y, r, c)));
        }
  return mx > 0 ? d / mx : (d == 0 ? 0.0 : 1.0);
}

void test_device_cf_layout() {
  // DeviceCFArrayUnitTests.cc:384-428 for this device's layout: the buffer
  // of copy_to() holds tap (x, y, mueller) at x + ext_x (y + ext_y mueller)
  EncodedCF cf(4, 3, 2);
  auto sz = hpg::get_value(cf.min_buffer_size(dev, 0));
  const auto e = cf.extents(0);
  CHECK(sz == size_t(e[0]) * e[1] * e[2] * e[3]);
  std::vector<hpg::CFArray::value_type> a(sz);
  auto vsn = cf.copy_to(dev, host, 0, a.data());
  CHECK(hpg::is_value(vsn));
  bool ok = true;
  for (size_t i = 0; i < a.size(); ++i) {
    const unsigned x = unsigned(i % e[0]);
    const unsigned y = unsigned((i / e[0]) % e[1]);
    const unsigned m = unsigned(i / (size_t(e[0]) * e[1]));
    ok &= a[i] == std::complex<float>(float(x + 1000 * y), float(m));
  }
  CHECK(ok);
  // OpenMP host device: the same buffer
  std::vector<hpg::CFArray::value_type> b(sz);
  CHECK(hpg::is_value(cf.copy_to(dev, hpg::Device::OpenMP, 0, b.data())));
  CHECK(a == b);
}

void test_device_cf_layout_version() {
  EncodedCF cf(2, 4, 1);
  std::vector<hpg::CFArray::value_type> a(hpg::get_value(cf.min_buffer_size(dev, 0)));
  auto vsn = hpg::get_value(cf.copy_to(dev, host, 0, a.data()));
  CHECK(vsn == "1-Hip");
  auto arrays = [&]() {
    std::vector<std::tuple<std::array<unsigned, 4>, std::vector<hpg::CFArray::value_type>>> v;
    v.emplace_back(cf.extents(0), a);
    return v;
  };
  auto ok = hpg::DeviceCFArray::create(vsn, 4, arrays());
  CHECK(hpg::is_value(ok));
  if (hpg::is_value(ok)) {
    auto d = hpg::get_value(std::move(ok));
    CHECK(d->device() == dev);
    CHECK((*d)(5, 7, 1, 0, 0) == cf(5, 7, 1, 0, 0));
  }
  auto bad = hpg::DeviceCFArray::create("Hip", 4, arrays());
  CHECK(hpg::is_error(bad) && hpg::get_error(bad).type() == hpg::ErrorType::InvalidCFLayout);
  auto other = hpg::DeviceCFArray::create("1-Cuda", 4, arrays());
  CHECK(hpg::is_error(other) && hpg::get_error(other).type() == hpg::ErrorType::DisabledDevice);
  auto newer = hpg::DeviceCFArray::create("2-Hip", 4, arrays());
  CHECK(hpg::is_error(newer) && hpg::get_error(newer).type() == hpg::ErrorType::InvalidCFLayout);
  auto host_dev = cf.min_buffer_size(hpg::Device::Serial, 0);
  CHECK(hpg::is_error(host_dev) && hpg::get_error(host_dev).type() == hpg::ErrorType::DisabledDevice);
  auto no_host = cf.copy_to(dev, dev, 0, a.data());
  CHECK(hpg::is_error(no_host)
        && hpg::get_error(no_host).type() == hpg::ErrorType::DisabledHostDevice);
  auto rw_host = hpg::RWDeviceCFArray::create(hpg::Device::Serial, cf);
  CHECK(hpg::is_error(rw_host) && hpg::get_error(rw_host).type() == hpg::ErrorType::DisabledDevice);
}

// the grid of 1000 records on a 50^2 grid (scale 0.1, -0.1) with a CF given
// as a CFArray and as a DeviceCFArray / RWDeviceCFArray of the same values
// (DeviceCFArrayUnitTests.cc:458-547, RWGridding): FP64 grids, equal to
// 1e-12 of max|grid| (the tile kernels add with LDS atomics, whose order
// varies from run to run)
void test_device_cf_gridding() {
  const std::array<unsigned, 4> gsz{50, 50, 1, 1};
  const std::array<double, 2> sc{0.1, -0.1};
  ConeCF cf(1, 20, 10);
  auto vis = random_vis(1000, gsz, sc, 10, 1, 42);
  std::vector<hpg::CFArray::value_type> a(hpg::get_value(cf.min_buffer_size(dev, 0)));
  auto vsn = hpg::get_value(cf.copy_to(dev, host, 0, a.data()));
  std::vector<std::tuple<std::array<unsigned, 4>, std::vector<hpg::CFArray::value_type>>> arr;
  arr.emplace_back(cf.extents(0), std::move(a));
  auto devcf = hpg::get_value(hpg::DeviceCFArray::create(vsn, 20, std::move(arr)));
  auto rw = hpg::get_value(hpg::RWDeviceCFArray::create(dev, cf));
  const auto e = cf.extents(0);
  for (unsigned y = 0; y < e[1]; ++y)
    for (unsigned x = 0; x < e[0]; ++x) (*rw)(x, y, 0, 0, 0) = cf(x, y, 0, 0, 0);
  auto grid_with = [&](hpg::CFArray&& c) {
    auto gs = hpg::get_value(hpg::GridderState::create<1>(dev, 0, 1000, &c, gsz, sc, {{0}}, {{0}}));
    gs = hpg::get_value(std::move(gs).set_convolution_function(host, std::move(c)));
    gs = hpg::get_value(
      std::move(gs).grid_visibilities(host, std::vector<hpg::VisData<1>>(vis)));
    return std::get<1>(std::move(gs).grid_values());
  };
  auto g_cf = grid_with(ConeCF(cf));
  CHECK(any_nonzero(*g_cf));
  auto g_dev = grid_with(std::move(*devcf));
  CHECK(rel_diff(*g_cf, *g_dev) <= 1e-12);
  auto g_rw = grid_with(std::move(*rw));
  CHECK(rel_diff(*g_cf, *g_rw) <= 1e-12);
}

// a CF with costly per-tap access (LargeCFArray, DeviceCFArrayUnitTests.cc:
// 92-145): two groups of 4 Mueller elements on 1330^2 oversampled taps
struct LargeCF final : public hpg::CFArray {
  unsigned m_groups;
  explicit LargeCF(unsigned groups) : m_groups(groups) {}
  unsigned oversampling() const override { return 10; }
  unsigned num_groups() const override { return m_groups; }
  std::array<unsigned, 4> extents(unsigned) const override { return {1330, 1330, 4, 1}; }
  std::complex<float> operator()(unsigned x, unsigned y, unsigned m, unsigned,
                                 unsigned g) const override {
    const float fx = float(x) - 665.0f, fy = float(y) - 665.0f;
    return std::polar(1.0f / (1.0f + 1e-4f * (fx * fx + fy * fy)), 1e-3f * fx * fy * (m + g + 1));
  }
};

void test_device_cf_efficiency() {
  // DeviceCFArrayUnitTests.cc:549-659: setting ten copies of a CF is faster
  // from DeviceCFArrays (values handed over) than from CFArrays (one call
  // per tap)
  LargeCF cf(2);
  std::vector<std::tuple<std::array<unsigned, 4>, std::vector<hpg::CFArray::value_type>>> arr;
  std::string vsn;
  for (unsigned g = 0; g < cf.num_groups(); ++g) {
    std::vector<hpg::CFArray::value_type> a(hpg::get_value(cf.min_buffer_size(dev, g)));
    vsn = hpg::get_value(cf.copy_to(dev, host, g, a.data()));
    arr.emplace_back(cf.extents(g), std::move(a));
  }
  auto time_set_cf = [&](std::vector<std::unique_ptr<hpg::CFArray>>& cfs) {
    auto gs = hpg::get_value(hpg::GridderState::create<1>(dev, 0, 10, cfs[0].get(), {1000, 1000, 1, 1},
                                                          {0.1, 0.1}, {{0}}, {{0}}));
    gs = std::move(gs).fence();
    const auto t0 = std::chrono::steady_clock::now();
    for (auto& c : cfs)
      gs = hpg::get_value(std::move(gs).set_convolution_function(host, std::move(*c)));
    gs = std::move(gs).fence();
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  };
  double t_cf, t_devcf;
  {
    std::vector<std::unique_ptr<hpg::CFArray>> cfs;
    for (int i = 0; i < 10; ++i) cfs.emplace_back(new LargeCF(cf));
    t_cf = time_set_cf(cfs);
  }
  {
    std::vector<std::unique_ptr<hpg::CFArray>> cfs;
    for (int i = 0; i < 10; ++i)
      cfs.push_back(hpg::get_value(hpg::DeviceCFArray::create(vsn, 10, decltype(arr)(arr))));
    t_devcf = time_set_cf(cfs);
  }
  std::printf("  cf %.3f s; devcf %.3f s\n", t_cf, t_devcf);
  CHECK(t_devcf < t_cf);
}

// ---- grid value and weight arrays (GridArraysUnitTests.cc:266-709) ---------

hpg::GridValueArray::value_type grid_value_encode(unsigned x, unsigned y, unsigned mr,
                                                  unsigned cb) {
  return {double(x * 1000 + y), double(mr * 1000 + cb)};
}
double grid_weight_encode(unsigned mr, unsigned cb) { return double(1000 * mr) + double(cb); }

void test_grid_arrays() {
  const std::array<unsigned, 4> gsz{8, 9, 2, 3};
  RandomCF cf(10, {{3 + 4, 3 + 4, 2, 3}, {2 + 4, 2 + 4, 2, 2}}, 42);
  auto gs = hpg::get_value(
    hpg::GridderState::create<1>(dev, 0, 4, &cf, gsz, {0.1, -0.1}, {{0}, {0}}, {{0}, {0}}));
  // GridValueReadWrite, CopyToValuesLayouts
  auto gvals = std::get<1>(gs.grid_values());
  for (unsigned x = 0; x < gsz[0]; ++x)
    for (unsigned y = 0; y < gsz[1]; ++y)
      for (unsigned mr = 0; mr < gsz[2]; ++mr)
        for (unsigned cb = 0; cb < gsz[3]; ++cb) (*gvals)(x, y, mr, cb) = grid_value_encode(x, y, mr, cb);
  bool rw = true, left = true, right = true;
  const size_t n = gvals->min_buffer_size();
  CHECK(n == size_t(gsz[0]) * gsz[1] * gsz[2] * gsz[3]);
  std::vector<hpg::GridValueArray::value_type> vl(n), vr(n);
  CHECK(!gvals->copy_to(host, vl.data(), hpg::Layout::Left));
  CHECK(!gvals->copy_to(host, vr.data(), hpg::Layout::Right));
  for (unsigned x = 0; x < gsz[0]; ++x)
    for (unsigned y = 0; y < gsz[1]; ++y)
      for (unsigned mr = 0; mr < gsz[2]; ++mr)
        for (unsigned cb = 0; cb < gsz[3]; ++cb) {
          const auto val = grid_value_encode(x, y, mr, cb);
          rw &= (*gvals)(x, y, mr, cb) == val;
          left &= vl[x + gsz[0] * (y + gsz[1] * (mr + gsz[2] * cb))] == val;
          right &= vr[cb + gsz[3] * (mr + gsz[2] * (y + gsz[1] * x))] == val;
        }
  CHECK(rw && left && right);
  // CopyFromValuesLayouts: from each layout, then back through the state
  const std::array<unsigned, 4> g2{8, 9, 4, 3};
  for (auto layout : {hpg::Layout::Left, hpg::Layout::Right}) {
    std::vector<hpg::GridValueArray::value_type> src(size_t(g2[0]) * g2[1] * g2[2] * g2[3]);
    for (unsigned x = 0; x < g2[0]; ++x)
      for (unsigned y = 0; y < g2[1]; ++y)
        for (unsigned mr = 0; mr < g2[2]; ++mr)
          for (unsigned cb = 0; cb < g2[3]; ++cb)
            src[layout == hpg::Layout::Left ? x + g2[0] * (y + g2[1] * (mr + g2[2] * cb))
                                            : cb + g2[3] * (mr + g2[2] * (y + g2[1] * x))] =
              grid_value_encode(x, y, mr, cb);
    auto a = hpg::GridValueArray::copy_from("v", dev, host, src.data(), g2, layout);
    bool eq = true;
    for (unsigned x = 0; x < g2[0]; ++x)
      for (unsigned y = 0; y < g2[1]; ++y)
        for (unsigned mr = 0; mr < g2[2]; ++mr)
          for (unsigned cb = 0; cb < g2[3]; ++cb) eq &= (*a)(x, y, mr, cb) == grid_value_encode(x, y, mr, cb);
    CHECK(eq);
    // the model of a state set from it reads back the same values
    RandomCF cf4(10, {{3 + 4, 3 + 4, 4, 3}}, 43);
    auto s4 = hpg::get_value(hpg::GridderState::create<1>(
      dev, 0, 4, &cf4, g2, {0.1, -0.1}, {{0}, {0}, {0}, {0}}, {{0}, {0}, {0}, {0}}));
    s4 = hpg::get_value(std::move(s4).set_model(host, std::move(*a)));
    auto m = std::get<1>(std::move(s4).model_values());
    bool meq = true;
    for (unsigned x = 0; x < g2[0]; ++x)
      for (unsigned y = 0; y < g2[1]; ++y)
        for (unsigned mr = 0; mr < g2[2]; ++mr)
          for (unsigned cb = 0; cb < g2[3]; ++cb) meq &= (*m)(x, y, mr, cb) == grid_value_encode(x, y, mr, cb);
    CHECK(meq);
  }
  // CopyFromWeightsLayouts
  const std::array<unsigned, 2> w2{4, 3};
  for (auto layout : {hpg::Layout::Left, hpg::Layout::Right}) {
    std::vector<double> src(size_t(w2[0]) * w2[1]);
    for (unsigned mr = 0; mr < w2[0]; ++mr)
      for (unsigned cb = 0; cb < w2[1]; ++cb)
        src[layout == hpg::Layout::Left ? mr + w2[0] * cb : cb + w2[1] * mr] = grid_weight_encode(mr, cb);
    auto a = hpg::GridWeightArray::copy_from("w", dev, host, src.data(), w2, layout);
    bool eq = a->extent(0) == w2[0] && a->extent(1) == w2[1];
    for (unsigned mr = 0; mr < w2[0]; ++mr)
      for (unsigned cb = 0; cb < w2[1]; ++cb) eq &= (*a)(mr, cb) == grid_weight_encode(mr, cb);
    CHECK(eq);
  }
  // CopyToWeightsLayouts
  auto gw = std::get<1>(gs.grid_weights());
  for (unsigned mr = 0; mr < gsz[2]; ++mr)
    for (unsigned cb = 0; cb < gsz[3]; ++cb) (*gw)(mr, cb) = grid_weight_encode(mr, cb);
  std::vector<double> wl(gw->min_buffer_size()), wr(gw->min_buffer_size());
  CHECK(!gw->copy_to(host, wl.data(), hpg::Layout::Left));
  CHECK(!gw->copy_to(host, wr.data(), hpg::Layout::Right));
  bool weq = true;
  for (unsigned mr = 0; mr < gsz[2]; ++mr)
    for (unsigned cb = 0; cb < gsz[3]; ++cb) {
      weq &= wl[mr + gsz[2] * cb] == grid_weight_encode(mr, cb);
      weq &= wr[cb + gsz[3] * mr] == grid_weight_encode(mr, cb);
    }
  CHECK(weq);
  CHECK(gvals->copy_to(dev, vl.data(), hpg::Layout::Left).has_value());  // not a host device
}

void test_compare_layouts() {
  // GridArraysUnitTests.cc:580-709: the grid and weights of a gridding run,
  // read in both layouts, agree element by element with each other and
  // with the arrays
  const std::array<unsigned, 4> gsz{20, 21, 1, 3};
  const std::array<double, 2> sc{0.1, -0.1};
  RandomCF cf(10, {{3 + 4, 3 + 4, 1, 3}, {2 + 4, 2 + 4, 1, 2}}, 42);
  auto vis = random_vis(100, gsz, sc, 3, 3, 7);
  for (size_t i = 0; i < vis.size(); ++i) vis[i].m_cf_index = {unsigned(i % 2), unsigned(i % 2)};
  auto gs = hpg::get_value(hpg::GridderState::create<1>(dev, 0, 100, &cf, gsz, sc, {{0}}, {{0}}));
  gs = hpg::get_value(std::move(gs).set_convolution_function(host, RandomCF(cf)));
  gs = hpg::get_value(std::move(gs).grid_visibilities(host, std::move(vis), true));
  auto [gs1, gvals] = std::move(gs).grid_values();
  auto [gs2, gwgts] = std::move(gs1).grid_weights();
  CHECK(any_nonzero(*gvals));
  std::vector<hpg::GridValueArray::value_type> vl(gvals->min_buffer_size()), vr(vl.size());
  std::vector<double> wl(gwgts->min_buffer_size()), wr(wl.size());
  CHECK(!gvals->copy_to(host, vl.data(), hpg::Layout::Left));
  CHECK(!gvals->copy_to(host, vr.data(), hpg::Layout::Right));
  CHECK(!gwgts->copy_to(host, wl.data(), hpg::Layout::Left));
  CHECK(!gwgts->copy_to(host, wr.data(), hpg::Layout::Right));
  bool eq = true, wnz = false;
  for (unsigned mr = 0; mr < gsz[2]; ++mr)
    for (unsigned cb = 0; cb < gsz[3]; ++cb) {
      const double a = wl[mr + gsz[2] * cb], b = wr[cb + gsz[3] * mr];
      eq &= a == b && a == (*gwgts)(mr, cb);
      wnz |= a != 0;
      for (unsigned x = 0; x < gsz[0]; ++x)
        for (unsigned y = 0; y < gsz[1]; ++y) {
          const auto& l = vl[x + gsz[0] * (y + gsz[1] * (mr + gsz[2] * cb))];
          const auto& r = vr[cb + gsz[3] * (mr + gsz[2] * (y + gsz[1] * x))];
          eq &= l == r && l == (*gvals)(x, y, mr, cb);
        }
    }
  CHECK(eq && wnz);
}

}  // namespace

int main(int argc, char** argv) {
  hpg::ScopeGuard guard;
  if (hpg::devices().count(hpg::Device::Hip) == 0) {
    std::printf("no HIP device\n");
    return 2;
  }
  const std::string golden = argc > 1 ? argv[1] : "tests/golden/gridone.bin";
  run("ConstructorArgs", test_constructor_args);
  run("GridOne", [&]() { test_gridone(golden); });
  run("Gridding", test_gridding_sweep);
  run("ResidualVisibilities", test_residual_visibilities);
  run("Errors", test_errors);
  run("CopiesAndSerialization", test_copy_semantics_and_serialization);
  run("ShiftCycleAndOneFFT", test_shift_cycle_and_fft);
  run("DeviceCFArrayLayout", test_device_cf_layout);
  run("DeviceCFArrayLayoutVersion", test_device_cf_layout_version);
  run("DeviceCFArrayGridding", test_device_cf_gridding);
  run("DeviceCFArrayEfficiency", test_device_cf_efficiency);
  run("GridArrays", test_grid_arrays);
  run("GridArraysCompareLayouts", test_compare_layouts);
  std::printf("%d checks, %d failures\n", g_checks, g_failures);
  return g_failures == 0 ? 0 : 1;
}
