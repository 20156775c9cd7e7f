/*
 * rexlib_hip.h — C ABI of the MI355X (gfx950) backend for rexlib's Fourier
 * slice insertion path.
 *
 * Every entry point takes plain device pointers, extents and strides, and a
 * HIP stream (NULL: the null stream). Work is enqueued on the stream and the
 * call returns without waiting, like a program launched by a command queue of
 * the reference's CUDA plugin (plugins/cuda/src/hardware/cuda_command_queue.cpp:102-145).
 *
 * Errors follow the reference's conventions, as status codes:
 *   REXLIB_HIP_INVALID_ARGUMENT  std::invalid_argument (bad operands or parameters,
 *                                e.g. src/em/ops/reconstruction/insert_ctf_slices_operation.cpp:17-27)
 *   REXLIB_HIP_UNSUPPORTED       unsupported_operation_error (no program for the
 *                                operands, src/core/dispatch/program_service.cpp:84-89)
 *   REXLIB_HIP_RUNTIME_ERROR     a HIP runtime failure (REXLIB_CUDA_CHECK family,
 *                                docs/agents/cuda-conventions.md:26-28)
 * rexlib_hip_last_error() returns the message of the calling thread's last error.
 */
#ifndef REXLIB_HIP_H
#define REXLIB_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define REXLIB_HIP_MAX_RANK 8

typedef enum rexlib_hip_status {
	REXLIB_HIP_OK = 0,
	REXLIB_HIP_INVALID_ARGUMENT = 1,
	REXLIB_HIP_UNSUPPORTED = 2,
	REXLIB_HIP_RUNTIME_ERROR = 3
} rexlib_hip_status;

/* Element types (include/rexlib/core/numerical/numerical_type.hpp:18-46, plus
 * bfloat16, which the reference lacks and the staging forms below use). */
typedef enum rexlib_hip_dtype {
	REXLIB_HIP_FLOAT32 = 0,
	REXLIB_HIP_FLOAT64 = 1,
	REXLIB_HIP_COMPLEX64 = 2,
	REXLIB_HIP_COMPLEX128 = 3,
	REXLIB_HIP_FLOAT16 = 4,
	REXLIB_HIP_BFLOAT16 = 5,
	REXLIB_HIP_BOOL = 6, /* one byte, 0 or 1 */
	REXLIB_HIP_UINT64 = 7
} rexlib_hip_dtype;

/* A strided view of device memory: the address of element (0, ..., 0),
 * extents, and strides in elements (0 broadcasts an axis), as rexlib's
 * strided_layout (include/rexlib/core/layout/strided_layout.hpp:48-205)
 * with its offset already applied to `data`. */
typedef struct rexlib_hip_array {
	void *data;
	int32_t dtype;
	int32_t rank;
	int64_t extents[REXLIB_HIP_MAX_RANK];
	int64_t strides[REXLIB_HIP_MAX_RANK];
} rexlib_hip_array;

/* gridding_window (include/rexlib/em/ops/parameters/gridding_window.hpp:23-90,
 * gridding_profile.hpp:14-44). */
typedef enum rexlib_hip_profile {
	REXLIB_HIP_BOX = 0,
	REXLIB_HIP_TRIANGLE = 1,
	REXLIB_HIP_KAISER_BESSEL = 2
} rexlib_hip_profile;

typedef struct rexlib_hip_window {
	int32_t profile;
	double half_width; /* 0.5 for box, 1 for triangle */
	double shape;      /* beta of Kaiser-Bessel, 0 otherwise */
} rexlib_hip_window;

/* ctf_optics (include/rexlib/em/ops/parameters/ctf_optics.hpp): kV, mm, fraction. */
typedef struct rexlib_hip_ctf_optics {
	double voltage;
	double spherical_aberration;
	double amplitude_contrast;
} rexlib_hip_ctf_optics;

/* ------------------------------------------------------------------------- */
/* Library state                                                             */
/* ------------------------------------------------------------------------- */

/* The message of the calling thread's last error ("" if none). */
const char *rexlib_hip_last_error(void);

/* Version of the rexlib contract this backend implements: 0.1.x
 * (VERSION:1 of the reference). */
void rexlib_hip_version(int32_t *major, int32_t *minor, int32_t *patch);

/* Number of visible HIP devices; 0 without a GPU, which is not an error
 * (plugins/cuda/src/hardware/cuda_device_backend.cpp:37-42). */
int32_t rexlib_hip_device_count(void);

/* Tuning of the brick schedule of the insertions, per volume extent N
 * (extent 0: every extent without an entry of its own): the first-tap cells
 * along x, y and z of the brick a wavefront owns in LDS -- (16, 8, 6),
 * (16, 8, 8), (16, 16, 8), (8, 8, 8), (8, 8, 6), (8, 8, 4) or (8, 6, 6) --
 * and the expected plane hits per task, in [16, 2^20], that splits busy
 * bricks into shares of their slices. Each call reads its extent's entry
 * once; without entries the default is (16, 8, 6) and 256. Returns
 * REXLIB_HIP_INVALID_ARGUMENT for other values.
 * No reference counterpart (the reference's CPU tuning is
 * slice_insertion_tuning, src/backends/cpu/builders/em/reconstruction/slice_insertion_tuning.cpp:10-13);
 * set by the autotuner (rexlib_amd/tune.py). */
int rexlib_hip_set_insertion_tuning(int32_t extent, int32_t extent_x, int32_t extent_y, int32_t extent_z,
	int32_t share);
int rexlib_hip_get_insertion_tuning(int32_t extent, int32_t *extent_x, int32_t *extent_y, int32_t *extent_z,
	int32_t *share);
/* The same for one window family: `taps` per axis of the brick layout, 1
 * (nearest), 2 (trilinear) or 4 (every Kaiser-Bessel window). A call with
 * that window reads this entry before the extent's. No reference
 * counterpart; set by the autotuner (rexlib_amd/tune.py). */
int rexlib_hip_set_insertion_tuning_for_window(int32_t extent, int32_t taps, int32_t extent_x, int32_t extent_y,
	int32_t extent_z, int32_t share);
/* Drops every entry (the defaults apply again). */
void rexlib_hip_clear_insertion_tuning(void);

/*
 * Event timing of single kernels, for measurement. When switched on, the
 * library brackets each launch of the named kernels with two HIP events on
 * the launch stream (not while the stream is being captured); switching on or
 * off discards earlier records. rexlib_hip_kernel_timing waits for the
 * recorded launches of `kernel` and returns their total time and count.
 * No reference counterpart (the reference times whole calls from outside).
 */
typedef enum rexlib_hip_kernel
{
    REXLIB_HIP_KERNEL_BRICK = 0,   /* brick insertion kernel */
    REXLIB_HIP_KERNEL_PREPARE = 1, /* per-sample preparation for the brick kernel; for images
                                      through rexlib_hip_insert_ctf_images, the column
                                      transforms fused with it */
    REXLIB_HIP_KERNEL_SCATTER = 2  /* atomic scatter insertion kernel */
} rexlib_hip_kernel;

int rexlib_hip_set_kernel_timing(int32_t enable);
int rexlib_hip_kernel_timing(int32_t kernel, double *total_ms, int64_t *launches);

/* ------------------------------------------------------------------------- */
/* Slice insertion — the hot path                                            */
/* ------------------------------------------------------------------------- */

/*
 * Replaces the program behind rexlib::em::insert_slices
 * (include/rexlib/em/functional/reconstruction.hpp:47-57;
 *  operation include/rexlib/em/ops/reconstruction/insert_slices_operation.hpp:16-141;
 *  CPU program src/backends/cpu/builders/em/reconstruction/insert_slices_program_builder.cpp:15-18).
 *
 * Operands, in the operation's order: numerator (..., N, N, N/2+1) complex64,
 * denominator (..., N, N, N/2+1) float32, slices (..., n, n/2+1) complex64,
 * transforms (..., 3, 4) float32, radial_weights (..., n/2+1) float32. The
 * outputs numerator_out / denominator_out have the accumulators' shapes; pass
 * the accumulators themselves to update them in place. Batch axes broadcast;
 * those the accumulators lack, or hold at extent 1, are summed.
 */
int rexlib_hip_insert_slices(
	const rexlib_hip_array *numerator,
	const rexlib_hip_array *denominator,
	const rexlib_hip_array *slices,
	const rexlib_hip_array *transforms,
	const rexlib_hip_array *radial_weights,
	const rexlib_hip_window *window,
	const rexlib_hip_array *numerator_out,
	const rexlib_hip_array *denominator_out,
	void *stream);

/*
 * Replaces the program behind rexlib::em::insert_weighted_slices
 * (include/rexlib/em/functional/reconstruction.hpp:88-100;
 *  operation insert_weighted_slices_operation.hpp:17-78): one more input,
 * weights (..., n, n/2+1) float32, after the slices.
 */
int rexlib_hip_insert_weighted_slices(
	const rexlib_hip_array *numerator,
	const rexlib_hip_array *denominator,
	const rexlib_hip_array *slices,
	const rexlib_hip_array *weights,
	const rexlib_hip_array *transforms,
	const rexlib_hip_array *radial_weights,
	const rexlib_hip_window *window,
	const rexlib_hip_array *numerator_out,
	const rexlib_hip_array *denominator_out,
	void *stream);

/*
 * Replaces the program behind rexlib::em::insert_ctf_slices
 * (include/rexlib/em/functional/reconstruction.hpp:136-149;
 *  operation insert_ctf_slices_operation.hpp:17-116): one more input,
 * ctf (..., 4) float32 = [D_u A, D_v A, angle rad, phase rad], after the
 * transforms; pixel_size must be finite and positive.
 */
int rexlib_hip_insert_ctf_slices(
	const rexlib_hip_array *numerator,
	const rexlib_hip_array *denominator,
	const rexlib_hip_array *slices,
	const rexlib_hip_array *transforms,
	const rexlib_hip_array *ctf,
	const rexlib_hip_array *radial_weights,
	const rexlib_hip_window *window,
	const rexlib_hip_ctf_optics *optics,
	double pixel_size,
	const rexlib_hip_array *numerator_out,
	const rexlib_hip_array *denominator_out,
	void *stream);

/*
 * rexlib_hip_insert_ctf_slices that allocates nothing. No counterpart in
 * the reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc). Same operands and results as rexlib_hip_insert_ctf_slices; its
 * workspace (slice entries, phase tables, the brick task list, prepared
 * samples) lies in one fixed 128 MiB device array of the library, one per
 * GPU, held until the process ends. A call whose entries and tables (3.2 KB
 * per 256^2 slice) with one slice's prepared samples do not fit returns
 * REXLIB_HIP_INVALID_ARGUMENT; larger calls go through in chunks of slices.
 * Every call on one GPU shares that GPU's array: calls that overlap on one
 * GPU (other streams, other threads) race.
 */
int rexlib_hip_insert_ctf_slices_no_alloc(
	const rexlib_hip_array *numerator,
	const rexlib_hip_array *denominator,
	const rexlib_hip_array *slices,
	const rexlib_hip_array *transforms,
	const rexlib_hip_array *ctf,
	const rexlib_hip_array *radial_weights,
	const rexlib_hip_window *window,
	const rexlib_hip_ctf_optics *optics,
	double pixel_size,
	const rexlib_hip_array *numerator_out,
	const rexlib_hip_array *denominator_out,
	void *stream);

/*
 * Fused form of rfft2 followed by insert_ctf_slices for real images
 * (..., n, n) stored as float32, bfloat16, or float16 (float16 optionally with
 * a per-image scale (..., ) float32 that multiplies each image). The images are
 * zero-padded on chip to `padded` (>= n; 0 means n) before the transform, and
 * the radial weights then have padded/2+1 entries. All arithmetic is float32.
 * No reference counterpart: it composes rexlib::rfft2
 * (include/rexlib/functional/fourier.hpp:483) and em::insert_ctf_slices, with
 * the same result as the two calls on the float32 images.
 */
int rexlib_hip_insert_ctf_images(
	const rexlib_hip_array *numerator,
	const rexlib_hip_array *denominator,
	const rexlib_hip_array *images,
	const rexlib_hip_array *image_scales,
	int64_t padded,
	const rexlib_hip_array *transforms,
	const rexlib_hip_array *ctf,
	const rexlib_hip_array *radial_weights,
	const rexlib_hip_window *window,
	const rexlib_hip_ctf_optics *optics,
	double pixel_size,
	void *stream);

/* ------------------------------------------------------------------------- */
/* Fourier transforms of the path                                            */
/* ------------------------------------------------------------------------- */

/*
 * Replaces the rfft program behind rexlib::rfft2
 * (include/rexlib/functional/fourier.hpp:483; ops::rfft_operation
 *  include/rexlib/ops/fourier/rfft_operation.hpp:20-72; CPU
 *  src/backends/cpu/kernels/pocketfft_fourier_transform.cpp:119-140):
 * images (..., n, n) float32 -> spectra (..., n, n/2+1) complex64, unscaled.
 */
int rexlib_hip_rfft2(const rexlib_hip_array *images, const rexlib_hip_array *spectra, void *stream);

/*
 * Replaces the irfft program behind rexlib::irfft3
 * (include/rexlib/functional/fourier.hpp:723; ops::irfft_operation
 *  include/rexlib/ops/fourier/irfft_operation.hpp:22-75; CPU
 *  pocketfft_fourier_transform.cpp:142-163): spectrum (..., N, N, N/2+1)
 * complex64 -> volume (..., N, N, N) float32, scaled by 1/N^3. The imaginary
 * parts of the x = 0 and (even N) x = N/2 planes' self-conjugate elements are
 * ignored, as a c2r transform does. The input is not modified.
 */
int rexlib_hip_irfft3(const rexlib_hip_array *spectrum, const rexlib_hip_array *volume, void *stream);

/* Inverse of rfft2: spectra (..., n, n/2+1) complex64 -> images (..., n, n)
 * float32, scaled by 1/n^2 (rexlib::irfft2, include/rexlib/functional/fourier.hpp). */
int rexlib_hip_irfft2(const rexlib_hip_array *spectra, const rexlib_hip_array *images, void *stream);

/* ------------------------------------------------------------------------- */
/* Reconstruction steps after the insertion (no reference function; the      */
/* quotient follows tests/integration/src/em/functional/test_cpu_reconstruction.cpp:457-464) */
/* ------------------------------------------------------------------------- */

/* out = numerator / denominator where denominator > threshold, else 0.
 * Contiguous (..., N, N, N/2+1) of one shape; out may alias numerator. */
int rexlib_hip_normalize(const rexlib_hip_array *numerator, const rexlib_hip_array *denominator,
	double threshold, const rexlib_hip_array *out, void *stream);

/* Divide a real volume (..., N, N, N) with its origin at N/2 by the separable
 * transform of the window: sinc (box), sinc^2 (triangle), or the Kaiser-Bessel
 * closed form, each normalised to 1 at the origin. out is (..., M, M, M) with
 * M <= N: for M < N it receives the central M^3 (starting at N/2 - M/2, the
 * crop of a reconstruction inserted with padding), else the whole volume, and
 * it may then alias volume. Contiguous operands. */
int rexlib_hip_gridding_correct(const rexlib_hip_array *volume, const rexlib_hip_window *window,
	const rexlib_hip_array *out, void *stream);

/* ------------------------------------------------------------------------- */
/* General transforms: any axes (in the given order; the last one listed is  */
/* the halved axis of rfft / irfft), any strides, lengths whose prime factors */
/* are 2, 3 and 5 up to 4096 directly and other lengths up to 2048 by        */
/* Bluestein's chirp-z algorithm. normalization: 0 backward (inverse scaled  */
/* by 1/count), 1 ortho (both by 1/sqrt(count)), 2 forward (forward scaled). */
/* float32 / complex64; float64 operands are UNSUPPORTED (the reference's    */
/* CPU backend keeps them). out may alias in.                                 */
/* Replace the programs behind ops::fft_operation, ifft_operation,            */
/* rfft_operation and irfft_operation (include/rexlib/ops/fourier/; CPU     */
/* src/backends/cpu/kernels/pocketfft_fourier_transform.cpp:74-163).          */
/* ------------------------------------------------------------------------- */

/* ------------------------------------------------------------------------- */
/* Background standardization of images: replaces the program behind          */
/* em::standardize_background_operation                                       */
/* (include/rexlib/em/ops/preprocessing/standardize_background_operation.hpp; */
/* CPU src/backends/cpu/builders/em/preprocessing/).                          */
/* ------------------------------------------------------------------------- */

/* An image is the value's part along the axes whose bits are set in
 * image_axes (bit a for axis a); the other axes are batch axes. Per image:
 * mean and population deviation, in double, of the finite background
 * elements (squared distance from the centre, index n / 2 along each image
 * axis, above radius^2); again over those whose value standardized with the
 * first statistics lies in [lower_bound, upper_bound]; then result =
 * (value - mean) * (1 / deviation) in the element type, and elements not
 * finite or outside the bounds take the standard normal draw of their
 * counter: the Box-Muller pair of the Philox4x32-10 block (c / 2, 0, 0, 0)
 * keyed by the image's key, cosine branch for even c, sine for odd. The
 * counter is the element's row-major position among the elements sharing
 * its key. value and result: float32 or float64, same extents (any
 * strides; result may alias value only element for element); key: uint64,
 * broadcasting against the batch axes from the right. radius >= 0,
 * lower_bound <= 0 <= upper_bound. */
int rexlib_hip_standardize_background(const rexlib_hip_array *value, const rexlib_hip_array *key,
	uint32_t image_axes, double radius, double lower_bound, double upper_bound, const rexlib_hip_array *result,
	void *stream);

/* Complex-to-complex DFT over `axes`: forward (exp(-i...)) or inverse. */
int rexlib_hip_fft(const rexlib_hip_array *in, const rexlib_hip_array *out, const int32_t *axes,
	int32_t axis_count, int32_t inverse, int32_t normalization, void *stream);

/* Real-to-complex DFT over `axes`; out holds n/2 + 1 along the last axis listed. */
int rexlib_hip_rfft(const rexlib_hip_array *in, const rexlib_hip_array *out, const int32_t *axes,
	int32_t axis_count, int32_t normalization, void *stream);

/* Complex-to-real inverse DFT over `axes`; in holds n/2 + 1 along the last axis
 * listed, out's extent n there; the imaginary parts of self-conjugate
 * coefficients are ignored. */
int rexlib_hip_irfft(const rexlib_hip_array *in, const rexlib_hip_array *out, const int32_t *axes,
	int32_t axis_count, int32_t normalization, void *stream);

/* ------------------------------------------------------------------------- */
/* Elementwise programs the quotient is written with. Operands broadcast     */
/* against the result as numpy's (an axis of extent 1, or a missing leading  */
/* axis, repeats); any strides.                                               */
/* ------------------------------------------------------------------------- */

/* Every element of out = real (+ i imaginary for complex types); float32,
 * float64, complex64, complex128, float16, bfloat16 and bool (real != 0).
 * Replaces the program behind ops::fill_operation (src/functional/creation.cpp:301-349). */
int rexlib_hip_fill(const rexlib_hip_array *out, double real, double imaginary, void *stream);

/* result = dividend / divisor, one element type: float32 or complex64
 * (complex quotients scale by the divisor's larger part first).
 * Replaces the program behind ops::divide_operation
 * (include/rexlib/ops/arithmetic/divide_operation.hpp; CPU
 * src/backends/cpu/builders/arithmetic/divide_program_builder.cpp). */
int rexlib_hip_divide(const rexlib_hip_array *dividend, const rexlib_hip_array *divisor,
	const rexlib_hip_array *result, void *stream);

/* result (bool) = a > b, float32 operands.
 * Replaces the program behind ops::greater_operation
 * (include/rexlib/ops/compare/greater_operation.hpp). */
int rexlib_hip_greater(const rexlib_hip_array *a, const rexlib_hip_array *b, const rexlib_hip_array *result,
	void *stream);

/* result = condition ? x : y; condition bool, x, y and result one type,
 * float32 or complex64. Replaces the program behind ops::where_operation
 * (include/rexlib/ops/selection/where_operation.hpp). */
int rexlib_hip_where(const rexlib_hip_array *condition, const rexlib_hip_array *x, const rexlib_hip_array *y,
	const rexlib_hip_array *result, void *stream);

/* Per-shell sums, in float64, of two half spectra a, b (N, N, N/2+1) complex64:
 * sums has 4 * (N/2+1) doubles = [sum Re(a conj b), sum |a|^2, sum |b|^2, count]
 * per shell round-half-up(|g|), each element of the stored half counted once.
 * The caller zeroes `sums` (device memory) before the first call. */
int rexlib_hip_shell_sums(const rexlib_hip_array *a, const rexlib_hip_array *b, double *sums,
	void *stream);

/* ------------------------------------------------------------------------- */
/* Density compensation: the iterative estimate and its learned predictor.   */
/* No reference counterpart (the reference divides the numerator by the      */
/* denominator in a test, test_cpu_reconstruction.cpp:457-464).              */
/* ------------------------------------------------------------------------- */

/* weights (N, N, N/2+1) float32 = the iterative density compensation of the
 * denominator: from W = 1, `passes` times W <- W / max(K3 * (W den), floor),
 * K3 the trilinear window's autocorrelation (1/6, 2/3, 1/6) per axis;
 * W = 0 where den <= 0. One contiguous grid. */
int rexlib_hip_density_compensation(const rexlib_hip_array *denominator, int32_t passes, double floor_value,
	const rexlib_hip_array *weights, void *stream);

/* table (8, 33, 64, 8) float32: the angular coverage of the slice normals
 * of `transforms` (..., 3, 4) for a volume of extent big. Entry
 * (r, t, p) is the radius (r + 0.5) / 8 * N/2 and the direction
 * u = (sin T cos P, sin T sin P, cos T), T = t pi / 32 (poles and equator
 * included), P = p 2 pi / 64 (x, y, z order, as the volume's frequencies);
 * its 8 values are the fractions of the slices whose plane passes within one
 * voxel of that point, widened by half the grid spacing
 * (|n . u| < 1 / radius + 0.05), falling in each eighth of
 * [0, pi) of the normal's angle around u, measured from d u / d T towards
 * d u / d P; zeros when no plane passes. No reference counterpart. */
int rexlib_hip_coverage_table(const rexlib_hip_array *transforms, int32_t big, const rexlib_hip_array *table,
	void *stream);

/* features (N, N, N/2+1, 16) float16: per voxel |g|/(N/2),
 * log(1 + den * density_scale), and the 8 coverage fractions of the table
 * entry of the voxel's radius bin floor(8 |g| / (N/2)) (at most 7) and of the
 * grid direction nearest in T and in P (P wraps around), then 6 zeros. */
int rexlib_hip_compensation_features(const rexlib_hip_array *denominator, const rexlib_hip_array *table,
	double density_scale, const rexlib_hip_array *features,
	void *stream);

/* The perceptron 10 -> 64 -> 64 -> 64 -> 1 (ReLU) on v_mfma_f32_16x16x16_f16:
 * parameters are rexlib_hip_mlp_parameter_count() elements, row-major
 * W1 (16x64), W2 (64x64), W3 (64x64), W4 (64x16, column 0), b1, b2, b3 (64)
 * and b4 (16, element 0), as float16 (the weights used) and float32 (the
 * master copy, whose biases are used). */
int32_t rexlib_hip_mlp_parameter_count(void);

/* out (rows) float32 = MLP(features (rows, 16) float16). */
int rexlib_hip_mlp_forward(const rexlib_hip_array *features, const rexlib_hip_array *parameters_f16,
	const rexlib_hip_array *parameters_f32, const rexlib_hip_array *out, void *stream);

/* Adds loss_scale * d(mean squared error)/d(parameter) to gradients and the
 * mean squared error to loss[0]. */
int rexlib_hip_mlp_backward(const rexlib_hip_array *features, const rexlib_hip_array *target,
	const rexlib_hip_array *parameters_f16, const rexlib_hip_array *parameters_f32,
	const rexlib_hip_array *gradients, const rexlib_hip_array *loss, double loss_scale, void *stream);

/* One Adam step (step counts from 1) on the float32 parameters from
 * gradients / loss_scale; refreshes the float16 copy and clears the gradients. */
int rexlib_hip_mlp_adam(const rexlib_hip_array *parameters_f32, const rexlib_hip_array *parameters_f16,
	const rexlib_hip_array *gradients, const rexlib_hip_array *first_moment,
	const rexlib_hip_array *second_moment, double learning_rate, double beta1, double beta2, double epsilon,
	int32_t step, double loss_scale, void *stream);

#ifdef __cplusplus
} /* extern "C" */
#endif


#endif /* REXLIB_HIP_H */
