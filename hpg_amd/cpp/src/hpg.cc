// Implementation of the hpg C++ API (include/hpg/hpg.hpp) over the C ABI of
// the MI355X device (include/hpgx.h).
//
// runtime::State owns one hpgx_state_t.  GridderState methods follow the
// reference's value semantics (src/hpg.cc:345-919): a `const &` method acts
// on a deep copy (hpgx_clone after a fence), a `&&` method on the target
// itself; validation errors leave the target untouched.

#include "hpg/hpg.hpp"

#include "hpgx.h"

#include <algorithm>
#include <atomic>
#include <cstddef>
#include <cstring>
#include <mutex>
#include <stdexcept>

namespace hpg {

// the C ABI takes hpg::VisData<N> records as they lie in memory
static_assert(sizeof(VisData<1>) == sizeof(hpgx_visdata1_t), "VisData<1> layout");
static_assert(sizeof(VisData<2>) == sizeof(hpgx_visdata2_t), "VisData<2> layout");
static_assert(sizeof(VisData<3>) == sizeof(hpgx_visdata3_t), "VisData<3> layout");
static_assert(sizeof(VisData<4>) == sizeof(hpgx_visdata4_t), "VisData<4> layout");
static_assert(offsetof(VisData<4>, m_frequency) == offsetof(hpgx_visdata4_t, frequency),
              "VisData frequency offset");
static_assert(offsetof(VisData<4>, m_grid_cube) == offsetof(hpgx_visdata4_t, grid_cube),
              "VisData grid_cube offset");
static_assert(offsetof(VisData<4>, m_cf_phase_gradient)
                == offsetof(hpgx_visdata4_t, cf_phase_gradient),
              "VisData cf_phase_gradient offset");

namespace {

Error hpgx_error(hpgx_state_t h, int rc) {
  std::string msg(hpgx_error_message(h, nullptr, 0), '\0');
  if (!msg.empty()) hpgx_error_message(h, msg.data(), msg.size() + 1);
  return Error(msg, ErrorType(rc - 1));
}

struct DisabledDeviceError : public Error {
  DisabledDeviceError() : Error("Requested device is not enabled", ErrorType::DisabledDevice) {}
};
struct DisabledHostDeviceError : public Error {
  DisabledHostDeviceError()
    : Error("Requested host device is not enabled", ErrorType::DisabledHostDevice) {}
};

std::atomic<bool> g_initialized{false};

// Layout::Left index of (x, y, mrow, cube)
inline size_t left_index(const std::array<unsigned, 4>& e, unsigned x, unsigned y,
                         unsigned mrow, unsigned cube) {
  return ((size_t(cube) * e[2] + mrow) * e[1] + y) * e[0] + x;
}
inline size_t right_index(const std::array<unsigned, 4>& e, unsigned x, unsigned y,
                          unsigned mrow, unsigned cube) {
  return ((size_t(x) * e[1] + y) * e[2] + mrow) * e[3] + cube;
}

// grid values held on the host in Layout::Left
class HostGridValueArray final : public GridValueArray {
public:
  std::array<unsigned, 4> m_ext;
  std::vector<value_type> m_v;

  explicit HostGridValueArray(const std::array<unsigned, 4>& ext)
    : m_ext(ext), m_v(size_t(ext[0]) * ext[1] * ext[2] * ext[3]) {}
  unsigned extent(unsigned dim) const override { return m_ext[dim]; }
  const value_type& operator()(unsigned x, unsigned y, unsigned mrow, unsigned cube) const override {
    return m_v[left_index(m_ext, x, y, mrow, cube)];
  }
  value_type& operator()(unsigned x, unsigned y, unsigned mrow, unsigned cube) override {
    return m_v[left_index(m_ext, x, y, mrow, cube)];
  }

protected:
  void unsafe_copy_to(Device, value_type* dst, Layout layout) const override {
    if (layout == Layout::Left) {
      std::copy(m_v.begin(), m_v.end(), dst);
      return;
    }
    for (unsigned c = 0; c < m_ext[3]; ++c)
      for (unsigned r = 0; r < m_ext[2]; ++r)
        for (unsigned y = 0; y < m_ext[1]; ++y)
          for (unsigned x = 0; x < m_ext[0]; ++x)
            dst[right_index(m_ext, x, y, r, c)] = m_v[left_index(m_ext, x, y, r, c)];
  }
};

class HostGridWeightArray final : public GridWeightArray {
public:
  std::array<unsigned, 2> m_ext;  // mrow, cube
  std::vector<value_type> m_v;    // [cube][mrow]

  explicit HostGridWeightArray(const std::array<unsigned, 2>& ext)
    : m_ext(ext), m_v(size_t(ext[0]) * ext[1]) {}
  unsigned extent(unsigned dim) const override { return m_ext[dim]; }
  const value_type& operator()(unsigned mrow, unsigned cube) const override {
    return m_v[size_t(cube) * m_ext[0] + mrow];
  }
  value_type& operator()(unsigned mrow, unsigned cube) override {
    return m_v[size_t(cube) * m_ext[0] + mrow];
  }
};

// ---- CF layouts (src/hpg_impl.cc:77-141) -----------------------------------

// version number of the Hip device's CF layout: the CFArray order, x
// fastest, then y, mueller, cube (Layout::Left over the CFArray axes), which
// hpgx_set_convolution_function takes and lays out for gridding on the GPU
constexpr unsigned cf_layout_version_number = 1;
const char* const hip_device_name = "Hip";

std::string hip_cf_layout_version() {
  return std::to_string(cf_layout_version_number) + "-" + hip_device_name;
}

// "<version number>-<device name>" -> (number, device name)
std::optional<std::pair<unsigned, std::string>> parse_cf_layout(const std::string& layout) {
  const auto dash = layout.find('-');
  if (dash == std::string::npos || dash == 0) return std::nullopt;
  unsigned vn = 0;
  for (size_t i = 0; i < dash; ++i) {
    if (layout[i] < '0' || layout[i] > '9') return std::nullopt;
    vn = vn * 10 + unsigned(layout[i] - '0');
  }
  return std::make_pair(vn, layout.substr(dash + 1));
}

// CF values in the Hip layout, one array per group
class HipDeviceCFArray final : public RWDeviceCFArray {
public:
  unsigned m_os;
  std::vector<std::array<unsigned, 4>> m_ext;
  std::vector<std::vector<value_type>> m_v;

  HipDeviceCFArray(unsigned os,
                   std::vector<std::tuple<std::array<unsigned, 4>, std::vector<value_type>>>&& a)
    : m_os(os) {
    for (auto& [e, v] : a) {
      m_ext.push_back(e);
      m_v.push_back(std::move(v));
    }
  }
  explicit HipDeviceCFArray(const CFArrayShape& shape) : m_os(shape.oversampling()) {
    for (unsigned g = 0; g < shape.num_groups(); ++g) {
      const auto e = shape.extents(g);
      m_ext.push_back(e);
      m_v.emplace_back(size_t(e[0]) * e[1] * e[2] * e[3], value_type(0));
    }
  }
  unsigned oversampling() const override { return m_os; }
  unsigned num_groups() const override { return unsigned(m_ext.size()); }
  std::array<unsigned, 4> extents(unsigned g) const override { return m_ext[g]; }
  Device device() const override { return Device::Hip; }
  std::complex<cf_fp> operator()(unsigned x, unsigned y, unsigned m, unsigned c,
                                 unsigned g) const override {
    return m_v[g][left_index(m_ext[g], x, y, m, c)];
  }
  std::complex<cf_fp>& operator()(unsigned x, unsigned y, unsigned m, unsigned c,
                                  unsigned g) override {
    return m_v[g][left_index(m_ext[g], x, y, m, c)];
  }
  const value_type* data(unsigned g) const { return m_v[g].data(); }
};

}  // namespace

const char* const cf_layout_unspecified_version = "";

rval_t<std::string> CFArray::copy_to(Device device, Device host_device, unsigned grp,
                                     value_type* dst) const {
  // src/hpg.cc:1566-1606
  if (host_devices().count(host_device) == 0) return rval<std::string>(DisabledHostDeviceError());
  if (devices().count(device) == 0) return rval<std::string>(DisabledDeviceError());
  const auto e = extents(grp);
  const long planes = long(e[2]) * e[3];
#pragma omp parallel for if (host_device == Device::OpenMP) schedule(static)
  for (long pc = 0; pc < planes; ++pc) {
    const unsigned m = unsigned(pc % e[2]);
    const unsigned c = unsigned(pc / e[2]);
    for (unsigned y = 0; y < e[1]; ++y)
      for (unsigned x = 0; x < e[0]; ++x) dst[left_index(e, x, y, m, c)] = (*this)(x, y, m, c, grp);
  }
  return rval(hip_cf_layout_version());
}

rval_t<size_t> CFArray::min_buffer_size(Device device, unsigned grp) const {
  // src/hpg.cc:1608-1612
  if (devices().count(device) == 0) return rval<size_t>(DisabledDeviceError());
  const auto e = extents(grp);
  return rval(size_t(e[0]) * e[1] * e[2] * e[3]);
}

rval_t<std::unique_ptr<DeviceCFArray>> DeviceCFArray::create(
  const std::string& version, unsigned oversampling,
  std::vector<std::tuple<std::array<unsigned, rank - 1>, std::vector<value_type>>>&& arrays) {
  // src/hpg.cc:1614-1660: the layout must name an enabled device exactly
  using R = rval_t<std::unique_ptr<DeviceCFArray>>;
  const auto parsed = parse_cf_layout(version);
  if (!parsed) return R(Error("Provided layout is invalid", ErrorType::InvalidCFLayout));
  if (parsed->second != hip_device_name || devices().count(Device::Hip) == 0)
    return R(DisabledDeviceError());
  if (parsed->first != cf_layout_version_number)
    return R(Error("Provided layout version is not supported", ErrorType::InvalidCFLayout));
  for (auto& [e, v] : arrays)
    if (v.size() != size_t(e[0]) * e[1] * e[2] * e[3])
      return R(Error("CF values do not match their extents", ErrorType::InvalidCFLayout));
  return R(std::unique_ptr<DeviceCFArray>(new HipDeviceCFArray(oversampling, std::move(arrays))));
}

rval_t<std::unique_ptr<RWDeviceCFArray>> RWDeviceCFArray::create(Device device,
                                                                 const CFArrayShape& shape) {
  // src/hpg.cc:1662-1701
  using R = rval_t<std::unique_ptr<RWDeviceCFArray>>;
  if (device != Device::Hip || devices().count(device) == 0) return R(DisabledDeviceError());
  return R(std::unique_ptr<RWDeviceCFArray>(new HipDeviceCFArray(shape)));
}

// ---- runtime::State: one device state ---------------------------------

namespace runtime {

struct Pending {
  VisDataVector vis;
  bool ready = false;
};

struct State {
  hpgx_state_t h = nullptr;
  std::array<unsigned, 4> grid_size{};
  std::array<grid_scale_fp, 2> grid_scale{};
  unsigned npol = 0;
  size_t max_batch = 0;
  mutable std::mutex mtx;
  mutable std::vector<std::shared_ptr<Pending>> pending;

  State() {}
  State(const State&) = delete;
  ~State() {
    if (h) hpgx_destroy(h);
  }

  void fence() const {
    std::lock_guard<std::mutex> lk(mtx);
    hpgx_fence(h);
    for (auto& p : pending) p->ready = true;
    pending.clear();
  }

  std::shared_ptr<State> clone() const {
    fence();
    auto s = std::make_shared<State>();
    int rc = hpgx_clone(h, &s->h);
    if (rc) throw std::runtime_error(hpgx_error(h, rc).message());
    s->grid_size = grid_size;
    s->grid_scale = grid_scale;
    s->npol = npol;
    s->max_batch = max_batch;
    return s;
  }
};

}  // namespace runtime

// ---- library -------------------------------------------------------------

bool initialize() { return initialize(InitArguments()); }

bool initialize(const InitArguments& args) {
  const int rc = hpgx_initialize(args.device_id);
  g_initialized = true;
  return rc == HPGX_SUCCESS || rc == HPGX_ERR_DISABLED_DEVICE;
}

void finalize() { hpgx_finalize(); }

bool is_initialized() noexcept { return g_initialized; }

const std::set<Device>& devices() noexcept {
  static const std::set<Device> with_gpu{Device::Hip};
  static const std::set<Device> none{};
  return hpgx_device_count() > 0 ? with_gpu : none;
}

const std::set<Device>& host_devices() noexcept {
  static const std::set<Device> hosts{Device::Serial, Device::OpenMP};
  return hosts;
}

ScopeGuard::ScopeGuard() : init(false) {
  if (!is_initialized()) {
    initialize();
    init = true;
  }
}
ScopeGuard::ScopeGuard(const InitArguments& args) : init(false) {
  if (!is_initialized()) {
    initialize(args);
    init = true;
  }
}
ScopeGuard::~ScopeGuard() {
  if (init) finalize();
}

// ---- arrays --------------------------------------------------------------

opt_t<Error> GridValueArray::copy_to(Device host_device, value_type* dst, Layout layout) const {
  if (host_devices().count(host_device) == 0) return DisabledHostDeviceError();
  unsafe_copy_to(host_device, dst, layout);
  return std::nullopt;
}

void GridValueArray::unsafe_copy_to(Device, value_type* dst, Layout layout) const {
  const std::array<unsigned, 4> e{extent(0), extent(1), extent(2), extent(3)};
  for (unsigned c = 0; c < e[3]; ++c)
    for (unsigned r = 0; r < e[2]; ++r)
      for (unsigned y = 0; y < e[1]; ++y)
        for (unsigned x = 0; x < e[0]; ++x)
          dst[layout == Layout::Left ? left_index(e, x, y, r, c) : right_index(e, x, y, r, c)] =
            (*this)(x, y, r, c);
}

std::unique_ptr<GridValueArray> GridValueArray::copy_from(
  const std::string&, Device, Device, const value_type* src,
  const std::array<unsigned, rank>& extents, Layout layout) {
  auto a = std::make_unique<HostGridValueArray>(extents);
  if (layout == Layout::Left) {
    std::copy(src, src + a->m_v.size(), a->m_v.begin());
  } else {
    for (unsigned c = 0; c < extents[3]; ++c)
      for (unsigned r = 0; r < extents[2]; ++r)
        for (unsigned y = 0; y < extents[1]; ++y)
          for (unsigned x = 0; x < extents[0]; ++x)
            a->m_v[left_index(extents, x, y, r, c)] = src[right_index(extents, x, y, r, c)];
  }
  return a;
}

opt_t<Error> GridWeightArray::copy_to(Device host_device, value_type* dst, Layout layout) const {
  if (host_devices().count(host_device) == 0) return DisabledHostDeviceError();
  for (unsigned c = 0; c < extent(1); ++c)
    for (unsigned r = 0; r < extent(0); ++r)
      dst[layout == Layout::Left ? size_t(c) * extent(0) + r : size_t(r) * extent(1) + c] =
        (*this)(r, c);
  return std::nullopt;
}

std::unique_ptr<GridWeightArray> GridWeightArray::copy_from(
  const std::string&, Device, Device, const value_type* src,
  const std::array<unsigned, rank>& extents, Layout layout) {
  auto a = std::make_unique<HostGridWeightArray>(extents);
  for (unsigned c = 0; c < extents[1]; ++c)
    for (unsigned r = 0; r < extents[0]; ++r)
      (*a)(r, c) =
        src[layout == Layout::Left ? size_t(c) * extents[0] + r : size_t(r) * extents[1] + c];
  return a;
}

// ---- GridderState ----------------------------------------------------------

GridderState::GridderState() {}

GridderState::GridderState(const GridderState& h) { *this = h; }

GridderState& GridderState::operator=(const GridderState& rhs) {
  if (this != &rhs) impl = rhs.impl ? rhs.impl->clone() : nullptr;
  return *this;
}

GridderState::~GridderState() {}

rval_t<GridderState> GridderState::create(
  Device device, unsigned max_added_tasks, size_t max_visibility_batch_size,
  const CFArrayShape* init_cf_shape, const std::array<unsigned, 4>& grid_size,
  const std::array<grid_scale_fp, 2>& grid_scale, IArrayVector&& mueller_indexes,
  IArrayVector&& conjugate_mueller_indexes) noexcept {
  // src/hpg.cc:238-263
  if (grid_size[2] != mueller_indexes.size() || grid_size[2] != conjugate_mueller_indexes.size())
    return rval<GridderState>(Error(
      "Number of rows of Mueller indexes does not equal grid mrow extent",
      ErrorType::InvalidNumberMuellerIndexRows));
  if (device != Device::Hip || devices().count(device) == 0)
    return rval<GridderState>(DisabledDeviceError());
  if (mueller_indexes.m_npol != conjugate_mueller_indexes.m_npol)
    return rval<GridderState>(
      Error("Mueller index vectors differ in polarizations", ErrorType::InvalidNumberPolarizations));
  try {
    const unsigned npol = mueller_indexes.m_npol;
    const unsigned rows = unsigned(mueller_indexes.size());
    std::vector<int> m(size_t(rows) * npol), mc(size_t(rows) * npol);
    if (rows > 0) {
      std::memcpy(m.data(), mueller_indexes.data(), m.size() * sizeof(int));
      std::memcpy(mc.data(), conjugate_mueller_indexes.data(), mc.size() * sizeof(int));
    }
    std::vector<unsigned> ext;
    unsigned os = 0, ng = 0;
    if (init_cf_shape) {
      os = init_cf_shape->oversampling();
      ng = init_cf_shape->num_groups();
      for (unsigned g = 0; g < ng; ++g)
        for (auto e : init_cf_shape->extents(g)) ext.push_back(e);
    }
    auto st = std::make_shared<runtime::State>();
    int rc = hpgx_create(
      &st->h, max_added_tasks, max_visibility_batch_size, grid_size.data(), grid_scale.data(),
      npol, rows, m.data(), mc.data(), HPGX_GRID_FP64, os, ng, ext.empty() ? nullptr : ext.data());
    if (rc) return rval<GridderState>(hpgx_error(nullptr, rc));
    st->grid_size = grid_size;
    st->grid_scale = grid_scale;
    st->npol = npol;
    st->max_batch = max_visibility_batch_size;
    GridderState gs;
    gs.impl = st;
    return rval(std::move(gs));
  } catch (const std::exception& e) {
    return rval<GridderState>(Error(e.what(), ErrorType::Other));
  }
}

Device GridderState::device() const noexcept { return Device::Hip; }
unsigned GridderState::max_added_tasks() const noexcept { return hpgx_max_added_tasks(impl->h); }
size_t GridderState::max_visibility_batch_size() const noexcept { return impl->max_batch; }
const std::array<unsigned, 4>& GridderState::grid_size() const noexcept { return impl->grid_size; }
std::array<grid_scale_fp, 2> GridderState::grid_scale() const noexcept { return impl->grid_scale; }
unsigned GridderState::num_polarizations() const noexcept { return impl->npol; }
bool GridderState::is_null() const noexcept { return !impl; }

void GridderState::swap(GridderState& other) noexcept { std::swap(impl, other.impl); }

namespace {

std::vector<unsigned> shape_extents(const CFArrayShape* shape) {
  std::vector<unsigned> ext;
  for (unsigned g = 0; shape && g < shape->num_groups(); ++g)
    for (auto e : shape->extents(g)) ext.push_back(e);
  return ext;
}

opt_error_t do_set_cf(runtime::State& st, Device host_device, const CFArray& cf) {
  const unsigned ng = cf.num_groups();
  std::vector<unsigned> ext = shape_extents(&cf);
  // fail early on CFSupportExceedsGrid, before the per-tap host loop
  for (unsigned g = 0; g < ng; ++g)
    if (size_t(ext[4 * g]) > size_t(st.grid_size[0]) * cf.oversampling()
        || size_t(ext[4 * g + 1]) > size_t(st.grid_size[1]) * cf.oversampling())
      return Error("CF support size exceeds grid size", ErrorType::CFSupportExceedsGrid);
  std::vector<std::vector<std::complex<float>>> vals(ng);
  std::vector<const void*> ptrs(ng);
  // a DeviceCFArray in this device's layout is handed over as it is
  // (CFPool::add_device_cfs, src/hpg_runtime.hpp:513-551, taken through the
  // dynamic_cast of src/hpg_runtime.hpp:907-913)
  if (auto* d = dynamic_cast<const HipDeviceCFArray*>(&cf)) {
    for (unsigned g = 0; g < ng; ++g) ptrs[g] = d->data(g);
    int rc = hpgx_set_convolution_function(
      st.h, cf.oversampling(), ng, ext.empty() ? nullptr : ext.data(),
      ptrs.empty() ? nullptr : ptrs.data(), HPGX_HOST);
    if (rc) return hpgx_error(st.h, rc);
    return std::nullopt;
  }
  // CFArray -> Layout::Left (x, y, mueller, cube), one call per tap, as
  // init_cf_host does (src/hpg_impl.hpp:1035-1046)
  for (unsigned g = 0; g < ng; ++g) {
    const unsigned* e = &ext[4 * g];
    auto& v = vals[g];
    v.resize(size_t(e[0]) * e[1] * e[2] * e[3]);
    const long planes = long(e[2]) * e[3];
#pragma omp parallel for if (host_device == Device::OpenMP) schedule(static)
    for (long pc = 0; pc < planes; ++pc) {
      const unsigned m = unsigned(pc % e[2]);
      const unsigned c = unsigned(pc / e[2]);
      for (unsigned y = 0; y < e[1]; ++y)
        for (unsigned x = 0; x < e[0]; ++x)
          v[((size_t(c) * e[2] + m) * e[1] + y) * e[0] + x] = cf(x, y, m, c, g);
    }
    ptrs[g] = v.data();
  }
  int rc = hpgx_set_convolution_function(
    st.h, cf.oversampling(), ng, ext.empty() ? nullptr : ext.data(),
    ptrs.empty() ? nullptr : ptrs.data(), HPGX_HOST);
  if (rc) return hpgx_error(st.h, rc);
  return std::nullopt;
}

opt_error_t do_set_model(runtime::State& st, const GridValueArray& gv) {
  const std::array<unsigned, 4> msz{gv.extent(0), gv.extent(1), gv.extent(2), gv.extent(3)};
  if (msz != st.grid_size) {
    auto s = [](const std::array<unsigned, 4>& a) {
      return "[" + std::to_string(a[0]) + "," + std::to_string(a[1]) + "," +
        std::to_string(a[2]) + "," + std::to_string(a[3]) + "]";
    };
    return Error("model grid size " + s(msz) + " is different from visibility grid size " +
                   s(st.grid_size),
                 ErrorType::InvalidModelGridSize);
  }
  std::vector<GridValueArray::value_type> buf;
  const GridValueArray::value_type* p = nullptr;
  if (auto* h = dynamic_cast<const HostGridValueArray*>(&gv)) {
    p = h->m_v.data();
  } else {
    buf.resize(gv.min_buffer_size());
    gv.copy_to(Device::Serial, buf.data(), Layout::Left);
    p = buf.data();
  }
  int rc = hpgx_set_model(st.h, msz.data(), p, HPGX_LAYOUT_LEFT, HPGX_HOST);
  if (rc) return hpgx_error(st.h, rc);
  return std::nullopt;
}

}  // namespace

size_t GridderState::convolution_function_region_size(const CFArrayShape* shape) const noexcept {
  std::vector<unsigned> ext = shape_extents(shape);
  return hpgx_convolution_function_region_size(
    impl->h, shape ? shape->oversampling() : 0, shape ? shape->num_groups() : 0,
    ext.empty() ? nullptr : ext.data());
}

rval_t<GridderState> GridderState::allocate_convolution_function_region(
  const CFArrayShape* shape) const & {
  GridderState r(*this);
  return std::move(r).allocate_convolution_function_region(shape);
}
rval_t<GridderState> GridderState::allocate_convolution_function_region(
  const CFArrayShape* shape) && {
  std::vector<unsigned> ext = shape_extents(shape);
  int rc = hpgx_allocate_convolution_function_region(
    impl->h, shape ? shape->oversampling() : 0, shape ? shape->num_groups() : 0,
    ext.empty() ? nullptr : ext.data());
  if (rc) return rval<GridderState>(hpgx_error(impl->h, rc));
  return rval(std::move(*this));
}

rval_t<GridderState> GridderState::set_convolution_function(Device host_device, CFArray&& cf) const & {
  if (host_devices().count(host_device) == 0) return rval<GridderState>(DisabledHostDeviceError());
  GridderState r(*this);
  return std::move(r).set_convolution_function(host_device, std::move(cf));
}
rval_t<GridderState> GridderState::set_convolution_function(Device host_device, CFArray&& cf) && {
  if (host_devices().count(host_device) == 0) return rval<GridderState>(DisabledHostDeviceError());
  if (auto err = do_set_cf(*impl, host_device, cf)) return rval<GridderState>(*err);
  return rval(std::move(*this));
}

rval_t<GridderState> GridderState::set_model(Device host_device, GridValueArray&& gv) const & {
  if (host_devices().count(host_device) == 0) return rval<GridderState>(DisabledHostDeviceError());
  GridderState r(*this);
  return std::move(r).set_model(host_device, std::move(gv));
}
rval_t<GridderState> GridderState::set_model(Device host_device, GridValueArray&& gv) && {
  if (host_devices().count(host_device) == 0) return rval<GridderState>(DisabledHostDeviceError());
  if (auto err = do_set_model(*impl, gv)) return rval<GridderState>(*err);
  return rval(std::move(*this));
}

rval_t<std::tuple<GridderState, future<VisDataVector>>> GridderState::grid_visibilities_base(
  Device host_device, VisDataVector&& visibilities, bool update_grid_weights, bool do_degrid,
  bool return_visibilities, bool do_grid) const & {
  using R = std::tuple<GridderState, future<VisDataVector>>;
  // validation before the copy (src/hpg_runtime.hpp:1815-1827)
  if (host_devices().count(host_device) == 0) return rval<R>(DisabledHostDeviceError());
  if (visibilities.size() > impl->max_batch)
    return rval<R>(Error("Number of visibilities exceeds maximum batch size",
                         ErrorType::ExcessiveNumberVisibilities));
  if (visibilities.m_npol != impl->npol)
    return rval<R>(Error("Number of visibility polarizations does not match Mueller matrix",
                         ErrorType::InvalidNumberPolarizations));
  if (!do_grid && update_grid_weights)
    return rval<R>(Error("Grid weights cannot be updated without doing gridding",
                         ErrorType::UpdateWeightsWithoutGridding));
  GridderState r(*this);
  return std::move(r).grid_visibilities_base(host_device, std::move(visibilities),
                                             update_grid_weights, do_degrid,
                                             return_visibilities, do_grid);
}

rval_t<std::tuple<GridderState, future<VisDataVector>>> GridderState::grid_visibilities_base(
  Device host_device, VisDataVector&& visibilities, bool update_grid_weights, bool do_degrid,
  bool return_visibilities, bool do_grid) && {
  using R = std::tuple<GridderState, future<VisDataVector>>;
  if (host_devices().count(host_device) == 0) return rval<R>(DisabledHostDeviceError());
  if (visibilities.size() > impl->max_batch)
    return rval<R>(Error("Number of visibilities exceeds maximum batch size",
                         ErrorType::ExcessiveNumberVisibilities));
  if (visibilities.m_npol != impl->npol)
    return rval<R>(Error("Number of visibility polarizations does not match Mueller matrix",
                         ErrorType::InvalidNumberPolarizations));
  if (!do_grid && update_grid_weights)
    return rval<R>(Error("Grid weights cannot be updated without doing gridding",
                         ErrorType::UpdateWeightsWithoutGridding));
  auto pend = std::make_shared<runtime::Pending>();
  pend->vis = std::move(visibilities);
  void* p = pend->vis.data();
  int rc = hpgx_grid_visibilities(
    impl->h, p, pend->vis.size(), impl->npol, update_grid_weights, do_degrid,
    return_visibilities, do_grid, return_visibilities ? p : nullptr, HPGX_HOST);
  if (rc) return rval<R>(hpgx_error(impl->h, rc));
  future<VisDataVector> fv;
  if (return_visibilities) {
    {
      std::lock_guard<std::mutex> lk(impl->mtx);
      impl->pending.push_back(pend);
    }
    // resolved by the next fence of this state (src/hpg_runtime.hpp:699-717)
    fv = future<VisDataVector>(
      [pend, result = opt_t<VisDataVector>()]() mutable -> opt_t<VisDataVector>& {
        if (!result && pend->ready) result = std::move(pend->vis);
        return result;
      });
  }
  return rval(std::make_tuple(std::move(*this), std::move(fv)));
}

rval_t<GridderState> GridderState::grid_visibilities(
  Device host_device, VisDataVector&& visibilities, bool update_grid_weights) const & {
  return map(grid_visibilities_base(host_device, std::move(visibilities), update_grid_weights,
                                    false, false, true),
             [](std::tuple<GridderState, future<VisDataVector>>&& t) {
               return std::move(std::get<0>(t));
             });
}
rval_t<GridderState> GridderState::grid_visibilities(
  Device host_device, VisDataVector&& visibilities, bool update_grid_weights) && {
  return map(std::move(*this).grid_visibilities_base(
               host_device, std::move(visibilities), update_grid_weights, false, false, true),
             [](std::tuple<GridderState, future<VisDataVector>>&& t) {
               return std::move(std::get<0>(t));
             });
}

rval_t<GridderState> GridderState::degrid_grid_visibilities(
  Device host_device, VisDataVector&& visibilities, bool update_grid_weights) const & {
  return map(grid_visibilities_base(host_device, std::move(visibilities), update_grid_weights,
                                    true, false, true),
             [](std::tuple<GridderState, future<VisDataVector>>&& t) {
               return std::move(std::get<0>(t));
             });
}
rval_t<GridderState> GridderState::degrid_grid_visibilities(
  Device host_device, VisDataVector&& visibilities, bool update_grid_weights) && {
  return map(std::move(*this).grid_visibilities_base(
               host_device, std::move(visibilities), update_grid_weights, true, false, true),
             [](std::tuple<GridderState, future<VisDataVector>>&& t) {
               return std::move(std::get<0>(t));
             });
}

rval_t<std::tuple<GridderState, future<VisDataVector>>>
GridderState::degrid_get_predicted_visibilities(Device host_device, VisDataVector&& visibilities) const & {
  return grid_visibilities_base(host_device, std::move(visibilities), false, true, true, false);
}
rval_t<std::tuple<GridderState, future<VisDataVector>>>
GridderState::degrid_get_predicted_visibilities(Device host_device, VisDataVector&& visibilities) && {
  return std::move(*this).grid_visibilities_base(host_device, std::move(visibilities), false,
                                                 true, true, false);
}

rval_t<std::tuple<GridderState, future<VisDataVector>>>
GridderState::degrid_grid_get_residual_visibilities(
  Device host_device, VisDataVector&& visibilities, bool update_grid_weights) const & {
  return grid_visibilities_base(host_device, std::move(visibilities), update_grid_weights, true,
                                true, true);
}
rval_t<std::tuple<GridderState, future<VisDataVector>>>
GridderState::degrid_grid_get_residual_visibilities(
  Device host_device, VisDataVector&& visibilities, bool update_grid_weights) && {
  return std::move(*this).grid_visibilities_base(host_device, std::move(visibilities),
                                                 update_grid_weights, true, true, true);
}

GridderState GridderState::fence() const & {
  GridderState r(*this);  // the copy fences the target
  return r;
}
GridderState GridderState::fence() && {
  impl->fence();
  return std::move(*this);
}

std::tuple<GridderState, std::unique_ptr<GridWeightArray>> GridderState::grid_weights() const & {
  GridderState r(*this);
  return std::move(r).grid_weights();
}
std::tuple<GridderState, std::unique_ptr<GridWeightArray>> GridderState::grid_weights() && {
  impl->fence();
  auto w = std::make_unique<HostGridWeightArray>(
    std::array<unsigned, 2>{impl->grid_size[2], impl->grid_size[3]});
  hpgx_grid_weights(impl->h, w->m_v.data(), HPGX_LAYOUT_LEFT);
  return std::make_tuple(std::move(*this), std::unique_ptr<GridWeightArray>(std::move(w)));
}

std::tuple<GridderState, std::unique_ptr<GridValueArray>> GridderState::grid_values() const & {
  GridderState r(*this);
  return std::move(r).grid_values();
}
std::tuple<GridderState, std::unique_ptr<GridValueArray>> GridderState::grid_values() && {
  impl->fence();
  auto g = std::make_unique<HostGridValueArray>(impl->grid_size);
  hpgx_grid_values(impl->h, g->m_v.data(), HPGX_LAYOUT_LEFT);
  return std::make_tuple(std::move(*this), std::unique_ptr<GridValueArray>(std::move(g)));
}

std::tuple<GridderState, std::unique_ptr<GridValueArray>> GridderState::model_values() const & {
  GridderState r(*this);
  return std::move(r).model_values();
}
std::tuple<GridderState, std::unique_ptr<GridValueArray>> GridderState::model_values() && {
  impl->fence();
  auto g = std::make_unique<HostGridValueArray>(impl->grid_size);
  hpgx_model_values(impl->h, g->m_v.data(), HPGX_LAYOUT_LEFT);
  return std::make_tuple(std::move(*this), std::unique_ptr<GridValueArray>(std::move(g)));
}

GridderState GridderState::reset_grid() const & {
  GridderState r(*this);
  return std::move(r).reset_grid();
}
GridderState GridderState::reset_grid() && {
  impl->fence();
  hpgx_reset_grid(impl->h);
  return std::move(*this);
}
GridderState GridderState::reset_model() const & {
  GridderState r(*this);
  return std::move(r).reset_model();
}
GridderState GridderState::reset_model() && {
  impl->fence();
  hpgx_reset_model(impl->h);
  return std::move(*this);
}

GridderState GridderState::normalize_by_weights(grid_value_fp wgt_factor) const & {
  GridderState r(*this);
  return std::move(r).normalize_by_weights(wgt_factor);
}
GridderState GridderState::normalize_by_weights(grid_value_fp wgt_factor) && {
  hpgx_normalize_by_weights(impl->h, wgt_factor);
  return std::move(*this);
}

rval_t<GridderState> GridderState::apply_grid_fft(grid_value_fp norm, FFTSign sign, bool in_place) const & {
  GridderState r(*this);
  return std::move(r).apply_grid_fft(norm, sign, in_place);
}
rval_t<GridderState> GridderState::apply_grid_fft(grid_value_fp norm, FFTSign sign, bool in_place) && {
  int rc = hpgx_apply_grid_fft(impl->h, norm, sign == FFTSign::NEGATIVE ? HPGX_FFT_NEGATIVE
                                                                         : HPGX_FFT_POSITIVE,
                               in_place);
  if (rc) return rval<GridderState>(hpgx_error(impl->h, rc));
  return rval(std::move(*this));
}
rval_t<GridderState> GridderState::apply_model_fft(grid_value_fp norm, FFTSign sign, bool in_place) const & {
  GridderState r(*this);
  return std::move(r).apply_model_fft(norm, sign, in_place);
}
rval_t<GridderState> GridderState::apply_model_fft(grid_value_fp norm, FFTSign sign, bool in_place) && {
  int rc = hpgx_apply_model_fft(impl->h, norm, sign == FFTSign::NEGATIVE ? HPGX_FFT_NEGATIVE
                                                                          : HPGX_FFT_POSITIVE,
                                in_place);
  if (rc) return rval<GridderState>(hpgx_error(impl->h, rc));
  return rval(std::move(*this));
}

GridderState GridderState::shift_grid(ShiftDirection direction) const & {
  GridderState r(*this);
  return std::move(r).shift_grid(direction);
}
GridderState GridderState::shift_grid(ShiftDirection direction) && {
  hpgx_shift_grid(impl->h, direction == ShiftDirection::FORWARD ? HPGX_SHIFT_FORWARD
                                                                : HPGX_SHIFT_BACKWARD);
  return std::move(*this);
}
GridderState GridderState::shift_model(ShiftDirection direction) const & {
  GridderState r(*this);
  return std::move(r).shift_model(direction);
}
GridderState GridderState::shift_model(ShiftDirection direction) && {
  hpgx_shift_model(impl->h, direction == ShiftDirection::FORWARD ? HPGX_SHIFT_FORWARD
                                                                 : HPGX_SHIFT_BACKWARD);
  return std::move(*this);
}

// ---- Gridder (src/hpg.cc:974-1400) -----------------------------------------

Gridder::Gridder() {}
Gridder::~Gridder() {}

rval_t<Gridder> Gridder::create(
  Device device, unsigned max_added_tasks, size_t max_visibility_batch_size,
  const CFArrayShape* init_cf_shape, const std::array<unsigned, 4>& grid_size,
  const std::array<grid_scale_fp, 2>& grid_scale, IArrayVector&& mueller_indexes,
  IArrayVector&& conjugate_mueller_indexes) noexcept {
  return map(GridderState::create(device, max_added_tasks, max_visibility_batch_size,
                                  init_cf_shape, grid_size, grid_scale,
                                  std::move(mueller_indexes),
                                  std::move(conjugate_mueller_indexes)),
             [](GridderState&& gs) { return Gridder(std::move(gs)); });
}

Device Gridder::device() const noexcept { return state.device(); }
unsigned Gridder::max_added_tasks() const noexcept { return state.max_added_tasks(); }
size_t Gridder::max_visibility_batch_size() const noexcept { return state.max_visibility_batch_size(); }
const std::array<unsigned, 4>& Gridder::grid_size() const noexcept { return state.grid_size(); }
std::array<grid_scale_fp, 2> Gridder::grid_scale() const noexcept { return state.grid_scale(); }
unsigned Gridder::num_polarizations() const noexcept { return state.num_polarizations(); }
bool Gridder::is_null() const noexcept { return state.is_null(); }

size_t Gridder::convolution_function_region_size(const CFArrayShape* shape) const noexcept {
  return state.convolution_function_region_size(shape);
}

namespace {
template <typename RV>
opt_error_t absorb(GridderState& state, RV&& rv) {
  return fold(
    std::forward<RV>(rv),
    [&state](GridderState&& gs) -> opt_error_t {
      state = std::move(gs);
      return std::nullopt;
    },
    [](const Error& err) -> opt_error_t { return err; });
}
}  // namespace

opt_error_t Gridder::allocate_convolution_function_region(const CFArrayShape* shape) {
  return absorb(state, std::move(state).allocate_convolution_function_region(shape));
}
opt_error_t Gridder::set_convolution_function(Device host_device, CFArray&& cf) {
  return absorb(state, std::move(state).set_convolution_function(host_device, std::move(cf)));
}
opt_error_t Gridder::set_model(Device host_device, GridValueArray&& gv) {
  return absorb(state, std::move(state).set_model(host_device, std::move(gv)));
}
opt_error_t Gridder::grid_visibilities(
  Device host_device, VisDataVector&& visibilities, bool update_grid_weights) {
  return absorb(state, std::move(state).grid_visibilities(host_device, std::move(visibilities),
                                                          update_grid_weights));
}
opt_error_t Gridder::degrid_grid_visibilities(
  Device host_device, VisDataVector&& visibilities, bool update_grid_weights) {
  return absorb(state, std::move(state).degrid_grid_visibilities(
                         host_device, std::move(visibilities), update_grid_weights));
}

namespace {
rval_t<future<VisDataVector>> absorb_future(
  GridderState& state, rval_t<std::tuple<GridderState, future<VisDataVector>>>&& rv) {
  return fold(
    std::move(rv),
    [&state](std::tuple<GridderState, future<VisDataVector>>&& t) {
      state = std::move(std::get<0>(t));
      return rval(std::move(std::get<1>(t)));
    },
    [](const Error& err) { return rval<future<VisDataVector>>(err); });
}
}  // namespace

rval_t<future<VisDataVector>> Gridder::degrid_get_predicted_visibilities(
  Device host_device, VisDataVector&& visibilities) {
  return absorb_future(
    state, std::move(state).degrid_get_predicted_visibilities(host_device, std::move(visibilities)));
}
rval_t<future<VisDataVector>> Gridder::degrid_grid_get_residual_visibilities(
  Device host_device, VisDataVector&& visibilities, bool update_grid_weights) {
  return absorb_future(state, std::move(state).degrid_grid_get_residual_visibilities(
                                host_device, std::move(visibilities), update_grid_weights));
}

void Gridder::fence() const { state = std::move(state).fence(); }
std::unique_ptr<GridWeightArray> Gridder::grid_weights() const {
  auto [gs, w] = std::move(state).grid_weights();
  state = std::move(gs);
  return std::move(w);
}
std::unique_ptr<GridValueArray> Gridder::grid_values() const {
  auto [gs, v] = std::move(state).grid_values();
  state = std::move(gs);
  return std::move(v);
}
std::unique_ptr<GridValueArray> Gridder::model_values() const {
  auto [gs, v] = std::move(state).model_values();
  state = std::move(gs);
  return std::move(v);
}
void Gridder::reset_grid() { state = std::move(state).reset_grid(); }
void Gridder::reset_model() { state = std::move(state).reset_model(); }
void Gridder::normalize_by_weights(grid_value_fp wgt_factor) {
  state = std::move(state).normalize_by_weights(wgt_factor);
}
opt_error_t Gridder::apply_grid_fft(grid_value_fp norm, FFTSign sign, bool in_place) {
  return absorb(state, std::move(state).apply_grid_fft(norm, sign, in_place));
}
opt_error_t Gridder::apply_model_fft(grid_value_fp norm, FFTSign sign, bool in_place) {
  return absorb(state, std::move(state).apply_model_fft(norm, sign, in_place));
}
void Gridder::shift_grid(ShiftDirection direction) {
  state = std::move(state).shift_grid(direction);
}
void Gridder::shift_model(ShiftDirection direction) {
  state = std::move(state).shift_model(direction);
}

}  // namespace hpg
