/* hpgx.h -- C ABI of the MI355X gridding device for HPG.
 *
 * This is the drop-in boundary: a new hpg::runtime::State implementation
 * ("StateHip", one per GridderState) calls these functions in place of the
 * Kokkos StateT<Device::Cuda> methods.  Every entry point names the reference
 * interface it replaces (paths relative to the hannajonathan/hpg v2.0.1 tree).
 *
 * Conventions
 *   - Every function returns an hpgx_status (0 on success) unless stated.
 *     Nonzero values are 1 + the hpg::ErrorType enumerator
 *     (src/hpg_error.hpp:26-38); the message of the last failure on a state
 *     is available from hpgx_error_message().
 *   - All work is queued on the state's HIP stream and returns without
 *     waiting, like the reference's device methods; functions documented as
 *     fencing wait for the stream first.
 *   - No torch or C++ types cross this boundary: plain pointers and sizes.
 *   - Visibilities are arrays of hpg::VisData<N> records
 *     (src/hpg.hpp:236-311), laid out as hpgx_visdataN_t below.
 *   - Grid and model values are complex<double> (re, im interleaved),
 *     Layout::Left (x fastest, then y, mrow, cube) unless a layout argument
 *     says otherwise; weights are double [cube][mrow] (mrow fastest) for
 *     Layout::Left and [mrow][cube] for Layout::Right.
 */
#ifndef HPGX_H
#define HPGX_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HPGX_API __attribute__((visibility("default")))

typedef enum hpgx_status {
  HPGX_SUCCESS = 0,
  HPGX_ERR_DISABLED_DEVICE = 1,
  HPGX_ERR_DISABLED_HOST_DEVICE = 2,
  HPGX_ERR_OUT_OF_BOUNDS_CF_INDEX = 3,
  HPGX_ERR_INVALID_NUMBER_MUELLER_INDEX_ROWS = 4,
  HPGX_ERR_INVALID_NUMBER_POLARIZATIONS = 5,
  HPGX_ERR_INVALID_CF_LAYOUT = 6,
  HPGX_ERR_INVALID_MODEL_GRID_SIZE = 7,
  HPGX_ERR_EXCESSIVE_NUMBER_VISIBILITIES = 8,
  HPGX_ERR_UPDATE_WEIGHTS_WITHOUT_GRIDDING = 9,
  HPGX_ERR_CF_SUPPORT_EXCEEDS_GRID = 10,
  HPGX_ERR_OTHER = 11
} hpgx_status_t;

/* grid value representation on the device */
typedef enum hpgx_grid_mode {
  /* complex<double> grid, FP32 CF x visibility products, FP64 accumulation:
   * the reference's grid_value_fp (src/hpg.hpp:197) */
  HPGX_GRID_FP64 = 0,
  /* complex<float> grid; FP32 products with compensated (two-sum) FP32 tile
   * accumulation; FP64 weights and FP64 normalization division */
  HPGX_GRID_FP32_COMPENSATED = 1
} hpgx_grid_mode_t;

/* hpg::Layout (src/hpg.hpp:617-620) */
typedef enum hpgx_layout { HPGX_LAYOUT_RIGHT = 0, HPGX_LAYOUT_LEFT = 1 } hpgx_layout_t;
/* hpg::FFTSign (src/hpg.hpp:965-968) */
typedef enum hpgx_fft_sign { HPGX_FFT_POSITIVE = 0, HPGX_FFT_NEGATIVE = 1 } hpgx_fft_sign_t;
/* hpg::ShiftDirection (src/hpg.hpp:989-992) */
typedef enum hpgx_shift { HPGX_SHIFT_FORWARD = 0, HPGX_SHIFT_BACKWARD = 1 } hpgx_shift_t;
/* where a buffer argument lives */
typedef enum hpgx_memory { HPGX_HOST = 0, HPGX_DEVICE = 1 } hpgx_memory_t;

/* hpg::VisData<N> (src/hpg.hpp:236-311): 88, 96, 112, 120 bytes */
#define HPGX_VISDATA(N)                                                   \
  struct {                                                                \
    float visibilities[N][2]; /* complex<float> per polarization */       \
    float weights[N];                                                     \
    double frequency;                                                     \
    double phase;                                                         \
    double uvw[3];                                                        \
    unsigned grid_cube;                                                   \
    unsigned cf_index[2]; /* {cube, group} */                             \
    double cf_phase_gradient[2];                                          \
  }
typedef HPGX_VISDATA(1) hpgx_visdata1_t;
typedef HPGX_VISDATA(2) hpgx_visdata2_t;
typedef HPGX_VISDATA(3) hpgx_visdata3_t;
typedef HPGX_VISDATA(4) hpgx_visdata4_t;

typedef struct hpgx_state_s* hpgx_state_t;

/* ---- library ---------------------------------------------------------- */

/* hpg::initialize(const InitArguments&) (src/hpg.cc:62-75,
 * src/hpg_impl.cc:34-62): selects the HIP device (device_id < 0: current). */
HPGX_API int hpgx_initialize(int device_id);
/* hpg::finalize() (src/hpg.cc:77-80) */
HPGX_API void hpgx_finalize(void);
/* number of HIP devices visible; hpg::devices() (src/hpg.cc:82-95) */
HPGX_API int hpgx_device_count(void);
HPGX_API const char* hpgx_version(void);

/* ---- state lifetime ---------------------------------------------------- */

/* hpg::GridderState::create (src/hpg.cc:226-264) +
 * StateT ctor (src/hpg_runtime.hpp:757-781): allocates the zeroed grid and
 * weights.  max_added_tasks + 1 tasks (clamped to 16; hpgx_max_added_tasks
 * reports the clamped value) rotate over host batches: each uploads its
 * records into its own device buffer on its own stream, so that the upload
 * of one grid_visibilities call overlaps the kernels of the calls before
 * (the rotation of src/hpg_runtime.hpp:1643-1681); all kernels run in call
 * order on the state's stream.  mueller_indexes /
 * conjugate_mueller_indexes: num_mueller_rows x num_polarizations ints,
 * row-major; num_mueller_rows must equal grid_size[2].  init_cf_extents (may be null): init_cf_num_groups x 4
 * extents used to size the first CF region. */
HPGX_API int hpgx_create(
  hpgx_state_t* state, unsigned max_added_tasks,
  size_t max_visibility_batch_size, const unsigned grid_size[4],
  const double grid_scale[2], unsigned num_polarizations,
  unsigned num_mueller_rows, const int* mueller_indexes,
  const int* conjugate_mueller_indexes, int grid_mode,
  unsigned init_cf_oversampling, unsigned init_cf_num_groups,
  const unsigned* init_cf_extents);

/* GridderState copy constructor: deep copy after a fence
 * (src/hpg.cc:266-306, src/hpg_runtime.hpp:783-799) */
HPGX_API int hpgx_clone(hpgx_state_t src, hpgx_state_t* dst);
/* StateT destructor (src/hpg_runtime.hpp:825-841) */
HPGX_API void hpgx_destroy(hpgx_state_t state);
/* copies the message of the state's last error (or of the last failed
 * hpgx_create when state is null); returns its full length */
HPGX_API size_t hpgx_error_message(hpgx_state_t state, char* buf, size_t len);

/* accessors: GridderState::device/max_added_tasks/max_visibility_batch_size/
 * grid_size/grid_scale/num_polarizations (src/hpg.cc:308-345) */
HPGX_API unsigned hpgx_max_added_tasks(hpgx_state_t state);
HPGX_API size_t hpgx_max_visibility_batch_size(hpgx_state_t state);
HPGX_API void hpgx_grid_size(hpgx_state_t state, unsigned grid_size[4]);
HPGX_API void hpgx_grid_scale(hpgx_state_t state, double grid_scale[2]);
HPGX_API unsigned hpgx_num_polarizations(hpgx_state_t state);
HPGX_API int hpgx_grid_mode(hpgx_state_t state);

/* ---- convolution function ---------------------------------------------- */

/* StateT::convolution_function_region_size (src/hpg_runtime.hpp:862-869):
 * number of complex<float> elements of device CF memory for a CF of the
 * given shape (num_groups x 4 extents). */
HPGX_API size_t hpgx_convolution_function_region_size(
  hpgx_state_t state, unsigned oversampling, unsigned num_groups,
  const unsigned* extents);
/* StateT::allocate_convolution_function_region (src/hpg_runtime.hpp:871-880);
 * num_groups == 0 releases the region.  Fences. */
HPGX_API int hpgx_allocate_convolution_function_region(
  hpgx_state_t state, unsigned oversampling, unsigned num_groups,
  const unsigned* extents);
/* StateT::set_convolution_function (src/hpg_runtime.hpp:890-915) with the
 * CFPool::add_host_cfs re-layout (src/hpg_runtime.hpp:458-511,
 * src/hpg_impl.hpp:1012-1047).  values[g]: complex<float> values of group g,
 * Layout::Left over the CFArray axes (x, y, mueller, cube), i.e. element
 * (x, y, m, c) at ((c * ext[2] + m) * ext[1] + y) * ext[0] + x; memory:
 * HPGX_HOST or HPGX_DEVICE.  Returns HPGX_ERR_CF_SUPPORT_EXCEEDS_GRID when
 * ext[0] > grid_size[0] * oversampling or ext[1] > grid_size[1] *
 * oversampling for any group. */
HPGX_API int hpgx_set_convolution_function(
  hpgx_state_t state, unsigned oversampling, unsigned num_groups,
  const unsigned* extents, const void* const* values, int memory);

/* ---- model ----------------------------------------------------------- */

/* StateT::set_model (src/hpg_runtime.hpp:917-963): model values
 * complex<double> in the given layout; HPGX_ERR_INVALID_MODEL_GRID_SIZE if
 * model_size differs from the grid size.  Fences. */
HPGX_API int hpgx_set_model(
  hpgx_state_t state, const unsigned model_size[4], const void* values,
  int layout, int memory);
/* StateT::reset_model (src/hpg_runtime.hpp:1249-1253) */
HPGX_API int hpgx_reset_model(hpgx_state_t state);
/* 1 if a model is set */
HPGX_API int hpgx_has_model(hpgx_state_t state);

/* ---- gridding / degridding ------------------------------------------- */

/* runtime::GridderState::grid_visibilities validation
 * (src/hpg_runtime.hpp:1804-1828) + StateT::grid_visibilities /
 * default_grid_visibilities (src/hpg_runtime.hpp:965-1001, 1132-1188) +
 * VisibilityGridder<N, exec, 0>::kernel (src/hpg_core.hpp:1082-1253).
 * Flags keep the reference's positional meaning (src/hpg.cc:434-478).
 * visibilities: num_visibilities records of hpgx_visdata<num_polarizations>
 * in `memory`.  With return_visibilities, the records -- with residual
 * values (do_degrid && do_grid) or predicted values (do_degrid && !do_grid)
 * for records inside the grid -- are written to visibilities_out (same
 * memory kind; may equal visibilities) once the stream reaches that point;
 * they are valid after the next hpgx_fence.  HPGX_HOST records are copied
 * before the call returns.  HPGX_DEVICE records may be read in place by the
 * queued kernels, so they must be ready when the call is made and stay
 * unchanged until the next hpgx_fence; they are never written (rewritten
 * values go to visibilities_out or to the state's own copy). */
HPGX_API int hpgx_grid_visibilities(
  hpgx_state_t state, const void* visibilities, size_t num_visibilities,
  unsigned num_polarizations, int update_grid_weights, int do_degrid,
  int return_visibilities, int do_grid, void* visibilities_out, int memory);

/* Bytes of the fixed device scratch of hpgx_grid_visibilities_noalloc, held
 * by every GPU the library runs on from the library's load until the process
 * ends. */
#define HPGX_NOALLOC_SCRATCH_BYTES (64u << 20)

/* Allocation-free form of hpgx_grid_visibilities for device-resident records
 * (HPGX_DEVICE memory): gridding only, i.e. do_degrid = 0,
 * return_visibilities = 0, do_grid = 1, with the same validation, results and
 * skipped-record reporting (hpgx_fence).  No counterpart in the reference: it
 * is for callers that cannot allocate device memory -- code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc.  It calls no
 * hipMalloc, hipFree, hipMemset or allocating library routine: its bin
 * entries, offset table and halo values live in a fixed
 * HPGX_NOALLOC_SCRATCH_BYTES device array of the library, and a batch whose
 * bins exceed it is gridded in several passes.  Grids of more than 8192
 * (plane, 32x32 tile) keys, or a grid and CF for which one bin block of 4096
 * records does not fit the scratch, are rejected with HPGX_ERR_OTHER and a
 * message.  The records must be ready when the call is made and stay
 * unchanged until the next hpgx_fence.  Every call on one GPU uses that GPU's
 * one scratch array: calls that may overlap in time on the same GPU (states
 * on different streams, or several host threads) race and must be ordered by
 * the caller; calls on one state are ordered by its stream. */
HPGX_API int hpgx_grid_visibilities_noalloc(
  hpgx_state_t state, const void* visibilities, size_t num_visibilities,
  unsigned num_polarizations, int update_grid_weights);

/* Forward operator of the differentiable form: the degridded (predicted)
 * visibilities of the current model for the records' coordinates, before
 * their rounding to complex<float>: predictions are num_visibilities x
 * num_polarizations complex<double>, in `memory`.  The records are not
 * modified and the grid is untouched.  Same arithmetic as the degridding of
 * hpgx_grid_visibilities (degrid_vis, src/hpg_core.hpp:579-682).  HPGX_DEVICE
 * records are read in place: unchanged until the next hpgx_fence.  No
 * reference counterpart (the reference has no differentiable form). */
HPGX_API int hpgx_degrid_forward(
  hpgx_state_t state, const void* visibilities, void* predictions,
  size_t num_visibilities, unsigned num_polarizations, int memory);

/* Adjoint of hpgx_degrid_forward (model -> predicted visibilities): grids
 * r_v * phasor / conj(W_v) with unit weight, where r are `residuals`
 * (num_visibilities x num_polarizations complex<double>, in `memory`) and
 * W_v the degridder's CF normalization, for the records' coordinates.  Adds
 * into the grid; grid weights unchanged.  Exact adjoint when the Mueller
 * index matrices equal their conjugates.  HPGX_DEVICE records and residuals
 * are read in place: unchanged until the next hpgx_fence.  No reference
 * counterpart. */
HPGX_API int hpgx_degrid_adjoint(
  hpgx_state_t state, const void* visibilities, const void* residuals,
  size_t num_visibilities, unsigned num_polarizations, int memory);

/* StateT::fence (src/hpg_runtime.hpp:1190-1194): waits for all queued work.
 * Also reports (and clears) records skipped since the last fence because
 * their cf_index or grid_cube was out of range:
 * HPGX_ERR_OUT_OF_BOUNDS_CF_INDEX. */
HPGX_API int hpgx_fence(hpgx_state_t state);

/* ---- finishing passes ------------------------------------------------ */

/* hpgx_normalize_by_weights, hpgx_apply_grid_fft with in_place and, for a
 * grid of even x and y sizes, hpgx_shift_grid directly after such an FFT are
 * queued on the state and run when the next entry point that reads or
 * writes the grid is called (gridding, degridding, fence, readback, window,
 * device pointer, event record, clone): a normalize and a shift around an
 * FFT then form one elementwise pass before the FFT (the shift as the
 * (-1)^(x+y) modulation of the FFT's input), with the results of the
 * separate passes.  hpgx_reset_grid drops a queued chain.  An error of the
 * queued work is returned by the entry point that runs it. */

/* StateT::normalize_by_weights + GridNormalizer
 * (src/hpg_runtime.hpp:1293-1310, src/hpg_core.hpp:1832-1861) */
HPGX_API int hpgx_normalize_by_weights(hpgx_state_t state, double wfactor);
/* StateT::apply_grid_fft / apply_model_fft (src/hpg_runtime.hpp:1312-1426)
 * with FFT<exec, 0> (src/hpg_core.hpp:2029-2276): unnormalized 2-D FFT of
 * every (mrow, cube) plane, then scaling by 1/norm when norm != 1. */
HPGX_API int hpgx_apply_grid_fft(
  hpgx_state_t state, double norm, int sign, int in_place);
HPGX_API int hpgx_apply_model_fft(
  hpgx_state_t state, double norm, int sign, int in_place);
/* StateT::shift_grid / shift_model + GridShifter
 * (src/hpg_runtime.hpp:1428-1456, src/hpg_core.hpp:2303-2542) */
HPGX_API int hpgx_shift_grid(hpgx_state_t state, int direction);
HPGX_API int hpgx_shift_model(hpgx_state_t state, int direction);
/* StateT::reset_grid (src/hpg_runtime.hpp:1243-1247): zero grid and
 * weights.  Fences. */
HPGX_API int hpgx_reset_grid(hpgx_state_t state);

/* ---- readback -------------------------------------------------------- */

/* StateT::grid_values / grid_weights / model_values
 * (src/hpg_runtime.hpp:1196-1241) + GridValueViewArray::copy_to
 * (src/hpg_impl.hpp:393-432).  Fence, then copy into host memory dst:
 * grid/model as complex<double> in `layout`, weights as double. */
HPGX_API int hpgx_grid_values(hpgx_state_t state, void* dst, int layout);
HPGX_API int hpgx_grid_weights(hpgx_state_t state, double* dst, int layout);
HPGX_API int hpgx_model_values(hpgx_state_t state, void* dst, int layout);
/* Fence, then copy the window [x0, x0+nx) x [y0, y0+ny) of plane
 * (mrow, cube) of the grid into host memory dst, complex<double>, x fastest.
 * No reference counterpart: a narrow form of grid_values() for large grids. */
HPGX_API int hpgx_grid_window(
  hpgx_state_t state, unsigned x0, unsigned y0, unsigned nx, unsigned ny,
  unsigned mrow, unsigned cube, void* dst);
/* GridderState::grid_values_ptr / grid_values_span
 * (src/hpg_runtime.hpp:1216-1228): device address and element count of the
 * grid (complex<double> in HPGX_GRID_FP64 mode, complex<float> in
 * HPGX_GRID_FP32_COMPENSATED mode), Layout::Left. */
HPGX_API int hpgx_grid_device_ptr(hpgx_state_t state, void** ptr, size_t* count);

/* ---- timing --------------------------------------------------------- */

/* Records a HIP event on the state's stream (slot 0..15); the elapsed time
 * between two recorded slots, in milliseconds, is returned by
 * hpgx_event_elapsed after both have completed.  No reference counterpart:
 * used to time the device work of a call on the stream it runs on. */
HPGX_API int hpgx_event_record(hpgx_state_t state, int slot);
HPGX_API int hpgx_event_elapsed(hpgx_state_t state, int slot0, int slot1, float* ms);

#ifdef __cplusplus
}
#endif

#endif /* HPGX_H */
