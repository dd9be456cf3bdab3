// hpg.hpp -- C++ API of HPG's gridding path on the MI355X device.
//
// The reference's public API (hannajonathan/hpg v2.0.1, src/hpg.hpp) for the
// gridding hot path, with the same names, argument meanings, value semantics
// and error values (rval_t = std::variant<Error, T>, as with Hpg_API 17), so
// that code written against hpg::GridderState / hpg::Gridder compiles and
// behaves the same.  The only device is Device::Hip, implemented by the C ABI
// of include/hpgx.h (hpg_amd/libhpgx.so).  Serial and OpenMP are accepted as
// host devices (for CF and model layout conversion on the host).
#pragma once

#include <array>
#include <cassert>
#include <complex>
#include <functional>
#include <memory>
#include <optional>
#include <set>
#include <string>
#include <tuple>
#include <type_traits>
#include <variant>
#include <vector>

namespace hpg {

// ---- errors (src/hpg_error.hpp:26-66) ----------------------------------

enum class ErrorType {
  DisabledDevice,
  DisabledHostDevice,
  OutOfBoundsCFIndex,
  InvalidNumberMuellerIndexRows,
  InvalidNumberPolarizations,
  InvalidCFLayout,
  InvalidModelGridSize,
  ExcessiveNumberVisibilities,
  UpdateWeightsWithoutGridding,
  CFSupportExceedsGrid,
  Other
};

class Error {
  ErrorType m_type = ErrorType::Other;
  std::string m_msg;

public:
  Error() {}
  Error(const std::string& msg, ErrorType err = ErrorType::Other)
    : m_type(err), m_msg(msg) {}
  const std::string& message() const { return m_msg; }
  ErrorType type() const { return m_type; }
  virtual ~Error() {}
};

// ---- rval_t and helpers (src/hpg_rval.hpp:33-225) -------------------------

template <typename T>
using rval_t = std::variant<Error, T>;

template <typename T>
struct rval_value {
  using type = void;
};
template <typename T>
struct rval_value<rval_t<T>> {
  using type = T;
};

template <typename T>
bool is_error(const rval_t<T>& rv) { return rv.index() == 0; }
template <typename T>
bool is_value(const rval_t<T>& rv) { return rv.index() == 1; }

template <typename RV>
decltype(auto) get_value(RV&& rv) { return std::get<1>(std::forward<RV>(rv)); }
template <typename RV>
decltype(auto) get_error(RV&& rv) { return std::get<0>(std::forward<RV>(rv)); }

template <typename T>
rval_t<std::decay_t<T>> rval(T&& t) {
  return rval_t<std::decay_t<T>>(std::in_place_index<1>, std::forward<T>(t));
}
template <typename T>
rval_t<T> rval(const Error& err) {
  return rval_t<T>(std::in_place_index<0>, err);
}

// rval_t<T> -> rval_t<U> with f: T -> U
template <typename RV, typename F>
auto map(RV&& rv, F f) {
  using T = typename rval_value<std::decay_t<RV>>::type;
  using U = std::invoke_result_t<F, T&&>;
  if (is_value(rv)) return rval_t<U>(std::in_place_index<1>, f(get_value(std::forward<RV>(rv))));
  return rval_t<U>(std::in_place_index<0>, get_error(std::forward<RV>(rv)));
}
// rval_t<T> -> rval_t<U> with f: T -> rval_t<U>
template <typename RV, typename F>
auto flatmap(RV&& rv, F f) {
  using T = typename rval_value<std::decay_t<RV>>::type;
  using R = std::invoke_result_t<F, T&&>;
  if (is_value(rv)) return f(get_value(std::forward<RV>(rv)));
  return R(std::in_place_index<0>, get_error(std::forward<RV>(rv)));
}
template <typename RV, typename ValF, typename ErrF>
auto fold(RV&& rv, ValF vf, ErrF ef) {
  if (is_value(rv)) return vf(get_value(std::forward<RV>(rv)));
  return ef(get_error(std::forward<RV>(rv)));
}

// Composable rval_t-valued functions (src/hpg_rval.hpp:228-416): RvalM<A, B>
// ::pure(f) wraps f: A -> rval_t<B> (or () -> rval_t<B> for A = void) in a
// value with and_then, and_then_loop, and_then_repeat and map.
template <typename A, typename B, typename F>
struct RvalChain;

template <typename A, typename B>
struct RvalM {
  template <typename F>
  static RvalChain<A, B, std::decay_t<F>> pure(F&& f) {
    return RvalChain<A, B, std::decay_t<F>>(std::forward<F>(f));
  }
};

template <typename A, typename B, typename F>
struct RvalChain : public RvalM<A, B> {
  F m_f;
  explicit RvalChain(F f) : m_f(std::move(f)) {}

  template <typename... Args>
  rval_t<B> operator()(Args&&... a) const { return m_f(std::forward<Args>(a)...); }

  template <typename G>
  auto and_then(const G& g) const {
    using C = typename rval_value<std::invoke_result_t<G, B&&>>::type;
    auto f = m_f;
    if constexpr (std::is_void_v<A>)
      return RvalM<A, C>::pure([g, f]() { return flatmap(f(), g); });
    else
      return RvalM<A, C>::pure([g, f](A&& a) { return flatmap(f(std::forward<A>(a)), g); });
  }
  template <typename G>
  auto and_then_loop(unsigned n, const G& g) const {
    auto f = m_f;
    auto step = [n, g](rval_t<B>&& r0) {
      rval_t<B> r = std::move(r0);
      for (unsigned i = 0; i < n; ++i)
        r = flatmap(std::move(r), [i, &g](B&& b) { return g(i, std::move(b)); });
      return r;
    };
    if constexpr (std::is_void_v<A>)
      return RvalM<A, B>::pure([f, step]() { return step(f()); });
    else
      return RvalM<A, B>::pure([f, step](A&& a) { return step(f(std::forward<A>(a))); });
  }
  template <typename G>
  auto and_then_repeat(unsigned n, const G& g) const {
    return and_then_loop(n, [g](unsigned, B&& b) { return g(std::move(b)); });
  }
  template <typename G>
  auto map(const G& g) const {
    using C = std::invoke_result_t<G, B&&>;
    auto f = m_f;
    if constexpr (std::is_void_v<A>)
      return RvalM<A, C>::pure([g, f]() { return ::hpg::map(f(), g); });
    else
      return RvalM<A, C>::pure([g, f](A&& a) { return ::hpg::map(f(std::forward<A>(a)), g); });
  }
};

template <typename T>
using opt_t = std::optional<T>;
using opt_error_t = opt_t<Error>;

// ---- library (src/hpg.hpp:59-190) ----------------------------------------

struct InitArguments {
  int num_threads = -1;
  int device_id = -1;
  bool disable_warnings = false;
  bool print_configuration = false;
  bool cleanup_fftw = false;
};

bool initialize();
bool initialize(const InitArguments& args);
void finalize();
bool is_initialized() noexcept;

enum class Device {
  Serial,
  OpenMP,
  Cuda,
  Hip, /**< AMD Instinct MI355X (gfx950) */
};

const std::set<Device>& devices() noexcept;
const std::set<Device>& host_devices() noexcept;

struct ScopeGuard {
private:
  bool init;

public:
  ScopeGuard();
  ScopeGuard(const InitArguments& args);
  ~ScopeGuard();
  ScopeGuard(const ScopeGuard&) = delete;
  ScopeGuard& operator=(const ScopeGuard&) = delete;
};

// ---- precision and value types (src/hpg.hpp:192-311) -----------------------

using cf_fp = float;
using visibility_fp = float;
using grid_value_fp = double;
using grid_scale_fp = double;
using vis_uvw_fp = double;
using vis_weight_fp = float;
using vis_frequency_fp = double;
using vis_phase_fp = double;
using cf_phase_gradient_fp = double;
using vis_uvw_t = std::array<vis_uvw_fp, 3>;
using cf_phase_gradient_t = std::array<cf_phase_gradient_fp, 2>;

template <unsigned N>
struct VisData {
  static constexpr unsigned npol = N;
  std::array<std::complex<visibility_fp>, N> m_visibilities;
  std::array<vis_weight_fp, N> m_weights;
  vis_frequency_fp m_frequency;
  vis_phase_fp m_phase;
  vis_uvw_t m_uvw;
  unsigned m_grid_cube;
  std::array<unsigned, 2> m_cf_index;
  cf_phase_gradient_t m_cf_phase_gradient;

  VisData(
    const std::array<std::complex<visibility_fp>, N>& visibilities,
    const std::array<vis_weight_fp, N>& weights, const vis_frequency_fp& frequency,
    const vis_phase_fp& phase, const vis_uvw_t& uvw, const unsigned& grid_cube,
    const std::array<unsigned, 2>& cf_index,
    const cf_phase_gradient_t& cf_phase_gradient = {0, 0})
    : m_visibilities(visibilities), m_weights(weights), m_frequency(frequency),
      m_phase(phase), m_uvw(uvw), m_grid_cube(grid_cube), m_cf_index(cf_index),
      m_cf_phase_gradient(cf_phase_gradient) {}
  VisData() {}
  bool operator==(const VisData& rhs) const {
    return m_visibilities == rhs.m_visibilities && m_weights == rhs.m_weights
      && m_frequency == rhs.m_frequency && m_phase == rhs.m_phase
      && m_uvw == rhs.m_uvw && m_grid_cube == rhs.m_grid_cube
      && m_cf_index == rhs.m_cf_index
      && m_cf_phase_gradient == rhs.m_cf_phase_gradient;
  }
};

// Vector of E<N> elements with N (1..4) erased (src/hpg.hpp:322-608): one
// alternative per polarization count, held in a std::variant.
template <template <unsigned> typename E>
struct VectorNPol {
  using V = std::variant<std::monostate, std::vector<E<1>>, std::vector<E<2>>,
                         std::vector<E<3>>, std::vector<E<4>>>;
  unsigned m_npol = 0;
  V m_v;

  VectorNPol() {}
  // one pair of constructors per polarization count (a template on N cannot
  // deduce N through std::array<int, N> for IArrayVector)
  VectorNPol(std::vector<E<1>>&& v) noexcept : m_npol(1), m_v(std::move(v)) {}
  VectorNPol(const std::vector<E<1>>& v) : m_npol(1), m_v(v) {}
  VectorNPol(std::vector<E<2>>&& v) noexcept : m_npol(2), m_v(std::move(v)) {}
  VectorNPol(const std::vector<E<2>>& v) : m_npol(2), m_v(v) {}
  VectorNPol(std::vector<E<3>>&& v) noexcept : m_npol(3), m_v(std::move(v)) {}
  VectorNPol(const std::vector<E<3>>& v) : m_npol(3), m_v(v) {}
  VectorNPol(std::vector<E<4>>&& v) noexcept : m_npol(4), m_v(std::move(v)) {}
  VectorNPol(const std::vector<E<4>>& v) : m_npol(4), m_v(v) {}

  size_t size() const {
    return std::visit(
      [](const auto& v) -> size_t {
        if constexpr (std::is_same_v<std::decay_t<decltype(v)>, std::monostate>)
          return 0;
        else
          return v.size();
      },
      m_v);
  }
  size_t num_elements() const { return m_npol * size(); }
  template <unsigned N>
  std::vector<E<N>>& get() { return std::get<N>(m_v); }
  template <unsigned N>
  const std::vector<E<N>>& get() const { return std::get<N>(m_v); }
  // address of the first element (null for an empty vector)
  const void* data() const {
    return std::visit(
      [](const auto& v) -> const void* {
        if constexpr (std::is_same_v<std::decay_t<decltype(v)>, std::monostate>)
          return nullptr;
        else
          return v.empty() ? nullptr : static_cast<const void*>(v.data());
      },
      m_v);
  }
  void* data() { return const_cast<void*>(static_cast<const VectorNPol*>(this)->data()); }
  virtual ~VectorNPol() {}
};

using VisDataVector = VectorNPol<VisData>;
template <unsigned N>
using iarray = std::array<int, N>;
using IArrayVector = VectorNPol<iarray>;

enum class Layout {
  Right, /**< C order: rightmost index has smallest stride */
  Left,  /**< FORTRAN order: leftmost index has smallest stride */
};

// ---- convolution function (src/hpg.hpp:623-721) ----------------------------

class CFArrayShape {
public:
  static constexpr unsigned rank = 5;
  virtual unsigned oversampling() const = 0;
  virtual unsigned num_groups() const = 0;
  virtual std::array<unsigned, rank - 1> extents(unsigned grp) const = 0;
  virtual ~CFArrayShape() {}
};

class CFArray : public CFArrayShape {
public:
  using value_type = std::complex<cf_fp>;
  static constexpr unsigned padding = 2;
  enum Axis { x, y, mueller, cube, group };

  virtual std::complex<cf_fp> operator()(
    unsigned x, unsigned y, unsigned mueller, unsigned cube, unsigned group) const = 0;

  std::array<unsigned, 2> radii(unsigned grp) const {
    auto os = oversampling();
    auto ext = extents(grp);
    return {((ext[0] - 2 * padding * os) / os) / 2,
            ((ext[1] - 2 * padding * os) / os) / 2};
  }
  virtual ~CFArray() {}

  // copy the values of group grp into dst, in the layout `device` takes
  // them without a host-side re-layout, returning that layout's version
  // string (src/hpg.hpp:697-720, src/hpg.cc:1566-1612).  For Device::Hip
  // the layout is the CFArray order, x fastest, then y, mueller, cube
  // (Layout::Left): the device lays its CF out for gridding itself, on the
  // GPU; a DeviceCFArray saves the per-tap calls of operator().
  rval_t<std::string> copy_to(Device device, Device host_device, unsigned grp,
                              value_type* dst) const;
  // elements of the destination of copy_to()
  rval_t<size_t> min_buffer_size(Device device, unsigned grp) const;
};

// version string of CF layouts with no device (src/hpg.cc:1562-1563)
extern const char* const cf_layout_unspecified_version;

// CF values stored in a device's layout (src/hpg.hpp:723-751): created from
// copy_to() buffers; set_convolution_function() hands them over unchanged
class DeviceCFArray : public CFArray {
public:
  static rval_t<std::unique_ptr<DeviceCFArray>> create(
    const std::string& version, unsigned oversampling,
    std::vector<std::tuple<std::array<unsigned, rank - 1>, std::vector<value_type>>>&& arrays);
  virtual Device device() const = 0;
  virtual ~DeviceCFArray() {}
};

// writable DeviceCFArray (src/hpg.hpp:753-790), zero-initialized to a shape
class RWDeviceCFArray : public DeviceCFArray {
public:
  static rval_t<std::unique_ptr<RWDeviceCFArray>> create(Device device,
                                                         const CFArrayShape& shape);
  using CFArray::operator();
  virtual std::complex<cf_fp>& operator()(unsigned x, unsigned y, unsigned mueller,
                                          unsigned cube, unsigned group) = 0;
  virtual ~RWDeviceCFArray() {}
};

// ---- grid values and weights (src/hpg.hpp:796-960) ---------------------------

class GridValueArray {
public:
  static constexpr unsigned rank = 4;
  using value_type = std::complex<grid_value_fp>;
  enum Axis { x, y, mrow, cube };

  virtual unsigned extent(unsigned dim) const = 0;
  virtual const value_type& operator()(unsigned x, unsigned y, unsigned mrow, unsigned cube) const = 0;
  virtual value_type& operator()(unsigned x, unsigned y, unsigned mrow, unsigned cube) = 0;

  opt_t<Error> copy_to(Device host_device, value_type* dst, Layout layout = Layout::Left) const;
  virtual size_t min_buffer_size() const {
    return size_t(extent(0)) * extent(1) * extent(2) * extent(3);
  }
  static std::unique_ptr<GridValueArray> copy_from(
    const std::string& name, Device target_device, Device host_device,
    const value_type* src, const std::array<unsigned, rank>& extents,
    Layout layout = Layout::Left);
  virtual ~GridValueArray() {}

protected:
  virtual void unsafe_copy_to(Device host_device, value_type* dst, Layout layout) const;
};

class GridWeightArray {
public:
  static constexpr unsigned rank = 2;
  using value_type = grid_value_fp;
  enum Axis { mrow, cube };

  virtual unsigned extent(unsigned dim) const = 0;
  virtual const value_type& operator()(unsigned mrow, unsigned cube) const = 0;
  virtual value_type& operator()(unsigned mrow, unsigned cube) = 0;

  opt_t<Error> copy_to(Device host_device, value_type* dst, Layout layout = Layout::Left) const;
  virtual size_t min_buffer_size() const { return size_t(extent(0)) * extent(1); }
  // src/hpg.hpp:940-950: weights [mrow][cube] from a buffer in `layout`
  static std::unique_ptr<GridWeightArray> copy_from(
    const std::string& name, Device target_device, Device host_device,
    const value_type* src, const std::array<unsigned, rank>& extents,
    Layout layout = Layout::Left);
  virtual ~GridWeightArray() {}
};

enum class FFTSign { POSITIVE, NEGATIVE };
constexpr FFTSign grid_fft_sign_dflt = FFTSign::POSITIVE;
constexpr FFTSign model_fft_sign_dflt = FFTSign::NEGATIVE;

enum class ShiftDirection { FORWARD, BACKWARD };

// future resolved only inside GridderState calls (src/hpg.hpp:1004-1066)
template <typename T>
class future final {
public:
  future() {}
  future(const std::function<opt_t<T>&()>& f) : m_f(f) {}
  opt_t<T>& get() noexcept {
    if (!m_f) {
      static opt_t<T> none;
      none.reset();
      return none;
    }
    return m_f();
  }
  template <typename U>
  future<U> map(const std::function<U(T&&)>& f) && {
    return future<U>(
      [f, mf = std::move(m_f), result = opt_t<U>()]() mutable -> opt_t<U>& {
        if (!result) {
          auto& ot = mf();
          if (ot) result = f(std::move(ot).value());
        }
        return result;
      });
  }

private:
  std::function<opt_t<T>&()> m_f;
};

namespace runtime {
struct State;
}

class Gridder;

// ---- GridderState (src/hpg.hpp:1103-2162) -------------------------------------

class GridderState {
protected:
  friend class Gridder;
  std::shared_ptr<runtime::State> impl;

public:
  GridderState();

  static rval_t<GridderState> create(
    Device device, unsigned max_added_tasks, size_t max_visibility_batch_size,
    const CFArrayShape* init_cf_shape, const std::array<unsigned, 4>& grid_size,
    const std::array<grid_scale_fp, 2>& grid_scale, IArrayVector&& mueller_indexes,
    IArrayVector&& conjugate_mueller_indexes) noexcept;

  template <unsigned N>
  static rval_t<GridderState> create(
    Device device, unsigned max_added_tasks, size_t max_visibility_batch_size,
    const CFArrayShape* init_cf_shape, const std::array<unsigned, 4>& grid_size,
    const std::array<grid_scale_fp, 2>& grid_scale,
    const std::vector<std::array<int, size_t(N)>>& mueller_indexes,
    const std::vector<std::array<int, size_t(N)>>& conjugate_mueller_indexes) noexcept {
    return create(device, max_added_tasks, max_visibility_batch_size, init_cf_shape,
                  grid_size, grid_scale, IArrayVector(mueller_indexes),
                  IArrayVector(conjugate_mueller_indexes));
  }

  GridderState(const GridderState&);
  GridderState(GridderState&&) noexcept = default;
  virtual ~GridderState();
  GridderState& operator=(const GridderState&);
  GridderState& operator=(GridderState&&) noexcept = default;

  Device device() const noexcept;
  unsigned max_added_tasks() const noexcept;
  size_t max_visibility_batch_size() const noexcept;
  const std::array<unsigned, 4>& grid_size() const noexcept;
  std::array<grid_scale_fp, 2> grid_scale() const noexcept;
  unsigned num_polarizations() const noexcept;
  bool is_null() const noexcept;

  size_t convolution_function_region_size(const CFArrayShape* shape) const noexcept;
  rval_t<GridderState> allocate_convolution_function_region(const CFArrayShape* shape) const &;
  rval_t<GridderState> allocate_convolution_function_region(const CFArrayShape* shape) &&;

  rval_t<GridderState> set_convolution_function(Device host_device, CFArray&& cf) const &;
  rval_t<GridderState> set_convolution_function(Device host_device, CFArray&& cf) &&;

  rval_t<GridderState> set_model(Device host_device, GridValueArray&& gv) const &;
  rval_t<GridderState> set_model(Device host_device, GridValueArray&& gv) &&;

  // the implementation's positional meaning of the flags (src/hpg.cc:434-478)
  rval_t<std::tuple<GridderState, future<VisDataVector>>> grid_visibilities_base(
    Device host_device, VisDataVector&& visibilities, bool update_grid_weights,
    bool do_degrid, bool return_visibilities, bool do_grid) const &;
  rval_t<std::tuple<GridderState, future<VisDataVector>>> grid_visibilities_base(
    Device host_device, VisDataVector&& visibilities, bool update_grid_weights,
    bool do_degrid, bool return_visibilities, bool do_grid) &&;

  rval_t<GridderState> grid_visibilities(
    Device host_device, VisDataVector&& visibilities, bool update_grid_weights = true) const &;
  rval_t<GridderState> grid_visibilities(
    Device host_device, VisDataVector&& visibilities, bool update_grid_weights = true) &&;
  template <unsigned N>
  rval_t<GridderState> grid_visibilities(
    Device host_device, std::vector<VisData<N>>&& visibilities,
    bool update_grid_weights = true) const & {
    return grid_visibilities(host_device, VisDataVector(std::move(visibilities)),
                             update_grid_weights);
  }
  template <unsigned N>
  rval_t<GridderState> grid_visibilities(
    Device host_device, std::vector<VisData<N>>&& visibilities,
    bool update_grid_weights = true) && {
    return std::move(*this).grid_visibilities(
      host_device, VisDataVector(std::move(visibilities)), update_grid_weights);
  }
  template <unsigned N>
  rval_t<GridderState> grid_visibilities(
    Device host_device, const std::vector<VisData<N>>& visibilities,
    bool update_grid_weights = true) const & {
    return grid_visibilities(host_device, VisDataVector(visibilities), update_grid_weights);
  }

  rval_t<GridderState> degrid_grid_visibilities(
    Device host_device, VisDataVector&& visibilities, bool update_grid_weights = true) const &;
  rval_t<GridderState> degrid_grid_visibilities(
    Device host_device, VisDataVector&& visibilities, bool update_grid_weights = true) &&;
  template <unsigned N>
  rval_t<GridderState> degrid_grid_visibilities(
    Device host_device, std::vector<VisData<N>>&& visibilities,
    bool update_grid_weights = true) const & {
    return degrid_grid_visibilities(host_device, VisDataVector(std::move(visibilities)),
                                    update_grid_weights);
  }

  rval_t<std::tuple<GridderState, future<VisDataVector>>> degrid_get_predicted_visibilities(
    Device host_device, VisDataVector&& visibilities) const &;
  rval_t<std::tuple<GridderState, future<VisDataVector>>> degrid_get_predicted_visibilities(
    Device host_device, VisDataVector&& visibilities) &&;
  template <unsigned N>
  rval_t<std::tuple<GridderState, future<std::vector<VisData<N>>>>>
  degrid_get_predicted_visibilities(
    Device host_device, std::vector<VisData<N>>&& visibilities) const & {
    return narrow_future<N>(
      degrid_get_predicted_visibilities(host_device, VisDataVector(std::move(visibilities))));
  }

  rval_t<std::tuple<GridderState, future<VisDataVector>>>
  degrid_grid_get_residual_visibilities(
    Device host_device, VisDataVector&& visibilities, bool update_grid_weights = true) const &;
  rval_t<std::tuple<GridderState, future<VisDataVector>>>
  degrid_grid_get_residual_visibilities(
    Device host_device, VisDataVector&& visibilities, bool update_grid_weights = true) &&;
  template <unsigned N>
  rval_t<std::tuple<GridderState, future<std::vector<VisData<N>>>>>
  degrid_grid_get_residual_visibilities(
    Device host_device, std::vector<VisData<N>>&& visibilities,
    bool update_grid_weights = true) const & {
    return narrow_future<N>(degrid_grid_get_residual_visibilities(
      host_device, VisDataVector(std::move(visibilities)), update_grid_weights));
  }

  GridderState fence() const &;
  GridderState fence() &&;

  std::tuple<GridderState, std::unique_ptr<GridWeightArray>> grid_weights() const &;
  std::tuple<GridderState, std::unique_ptr<GridWeightArray>> grid_weights() &&;
  std::tuple<GridderState, std::unique_ptr<GridValueArray>> grid_values() const &;
  std::tuple<GridderState, std::unique_ptr<GridValueArray>> grid_values() &&;
  std::tuple<GridderState, std::unique_ptr<GridValueArray>> model_values() const &;
  std::tuple<GridderState, std::unique_ptr<GridValueArray>> model_values() &&;

  GridderState reset_grid() const &;
  GridderState reset_grid() &&;
  GridderState reset_model() const &;
  GridderState reset_model() &&;
  GridderState normalize_by_weights(grid_value_fp wgt_factor = 1) const &;
  GridderState normalize_by_weights(grid_value_fp wgt_factor = 1) &&;

  rval_t<GridderState> apply_grid_fft(
    grid_value_fp norm = 1, FFTSign sign = grid_fft_sign_dflt, bool in_place = true) const &;
  rval_t<GridderState> apply_grid_fft(
    grid_value_fp norm = 1, FFTSign sign = grid_fft_sign_dflt, bool in_place = true) &&;
  rval_t<GridderState> apply_model_fft(
    grid_value_fp norm = 1, FFTSign sign = model_fft_sign_dflt, bool in_place = true) const &;
  rval_t<GridderState> apply_model_fft(
    grid_value_fp norm = 1, FFTSign sign = model_fft_sign_dflt, bool in_place = true) &&;

  GridderState shift_grid(ShiftDirection direction) const &;
  GridderState shift_grid(ShiftDirection direction) &&;
  GridderState shift_model(ShiftDirection direction) const &;
  GridderState shift_model(ShiftDirection direction) &&;

  void swap(GridderState& other) noexcept;

private:
  template <unsigned N>
  static rval_t<std::tuple<GridderState, future<std::vector<VisData<N>>>>> narrow_future(
    rval_t<std::tuple<GridderState, future<VisDataVector>>>&& r) {
    return map(std::move(r), [](std::tuple<GridderState, future<VisDataVector>>&& t) {
      auto& [gs, fv] = t;
      return std::make_tuple(
        std::move(gs),
        std::move(fv).template map<std::vector<VisData<N>>>(
          [](VisDataVector&& v) { return std::move(v.template get<N>()); }));
    });
  }
};

// ---- Gridder (src/hpg.hpp:2206-2792) -------------------------------------------

class Gridder {
protected:
  mutable GridderState state;

public:
  Gridder();
  static rval_t<Gridder> create(
    Device device, unsigned max_added_tasks, size_t max_visibility_batch_size,
    const CFArrayShape* init_cf_shape, const std::array<unsigned, 4>& grid_size,
    const std::array<grid_scale_fp, 2>& grid_scale, IArrayVector&& mueller_indexes,
    IArrayVector&& conjugate_mueller_indexes) noexcept;
  template <unsigned N>
  static rval_t<Gridder> create(
    Device device, unsigned max_added_tasks, size_t max_visibility_batch_size,
    const CFArrayShape* init_cf_shape, const std::array<unsigned, 4>& grid_size,
    const std::array<grid_scale_fp, 2>& grid_scale,
    const std::vector<std::array<int, size_t(N)>>& mueller_indexes,
    const std::vector<std::array<int, size_t(N)>>& conjugate_mueller_indexes) noexcept {
    return create(device, max_added_tasks, max_visibility_batch_size, init_cf_shape,
                  grid_size, grid_scale, IArrayVector(mueller_indexes),
                  IArrayVector(conjugate_mueller_indexes));
  }
  Gridder(const Gridder&) = default;
  Gridder(Gridder&&) noexcept = default;
  Gridder& operator=(const Gridder&) = default;
  Gridder& operator=(Gridder&&) noexcept = default;
  virtual ~Gridder();

  Device device() const noexcept;
  unsigned max_added_tasks() const noexcept;
  size_t max_visibility_batch_size() const noexcept;
  const std::array<unsigned, 4>& grid_size() const noexcept;
  std::array<grid_scale_fp, 2> grid_scale() const noexcept;
  unsigned num_polarizations() const noexcept;
  bool is_null() const noexcept;

  size_t convolution_function_region_size(const CFArrayShape* shape) const noexcept;
  opt_error_t allocate_convolution_function_region(const CFArrayShape* shape);
  opt_error_t set_convolution_function(Device host_device, CFArray&& cf);
  opt_error_t set_model(Device host_device, GridValueArray&& gv);
  opt_error_t grid_visibilities(
    Device host_device, VisDataVector&& visibilities, bool update_grid_weights = true);
  template <unsigned N>
  opt_error_t grid_visibilities(
    Device host_device, std::vector<VisData<N>>&& visibilities,
    bool update_grid_weights = true) {
    return grid_visibilities(host_device, VisDataVector(std::move(visibilities)),
                             update_grid_weights);
  }
  opt_error_t degrid_grid_visibilities(
    Device host_device, VisDataVector&& visibilities, bool update_grid_weights = true);
  rval_t<future<VisDataVector>> degrid_get_predicted_visibilities(
    Device host_device, VisDataVector&& visibilities);
  rval_t<future<VisDataVector>> degrid_grid_get_residual_visibilities(
    Device host_device, VisDataVector&& visibilities, bool update_grid_weights = true);

  void fence() const;
  std::unique_ptr<GridWeightArray> grid_weights() const;
  std::unique_ptr<GridValueArray> grid_values() const;
  std::unique_ptr<GridValueArray> model_values() const;
  void reset_grid();
  void reset_model();
  void normalize_by_weights(grid_value_fp wgt_factor = 1);
  opt_error_t apply_grid_fft(
    grid_value_fp norm = 1, FFTSign sign = grid_fft_sign_dflt, bool in_place = true);
  opt_error_t apply_model_fft(
    grid_value_fp norm = 1, FFTSign sign = model_fft_sign_dflt, bool in_place = true);
  void shift_grid(ShiftDirection direction);
  void shift_model(ShiftDirection direction);

protected:
  Gridder(GridderState&& st) : state(std::move(st)) {}
};

}  // namespace hpg
