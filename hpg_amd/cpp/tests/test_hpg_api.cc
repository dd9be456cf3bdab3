// Tests of the hpg C++ API (hpg_amd/cpp) on the HIP device, restating the
// reference's GridderState / Gridder unit tests
// (tests/GridderStateUnitTests.cc, tests/GridderUnitTests.cc) with a
// minimal harness.  Usage: test_hpg_api <path of tests/golden/gridone.bin>
#include "hpg/hpg.hpp"

#include <algorithm>
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
  std::printf("%d checks, %d failures\n", g_checks, g_failures);
  return g_failures == 0 ? 0 : 1;
}
