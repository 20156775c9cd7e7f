// Host launchers of the kernels, called by the C ABI.
#pragma once

#include "fft.hpp"
#include "insertion.hpp"

namespace rexlib_hip
{

// Workspace areas, one per use, per device and stream.
enum workspace_area : int
{
	area_insertion = 0,
	area_spectra = 1,
	area_fft = 2,
	area_prepared = 3,
	area_compensation = 4,
	area_count = 5
};

void *workspace_area_bytes(int area, std::size_t bytes, hipStream_t stream);

// Where an insertion takes its workspace: the per-stream areas above (grown
// on demand), or one fixed device buffer that the whole call must fit in
// (taken from its start, 256-byte aligned, nothing allocated).
struct workspace_source
{
	char *arena = nullptr; // null: the per-stream areas
	std::size_t arena_bytes = 0;
	std::size_t used = 0;

	void *take(int area, std::size_t bytes, hipStream_t stream)
	{
		if (arena == nullptr)
		{
			return workspace_area_bytes(area, bytes, stream);
		}
		const std::size_t at = (used + 255) / 256 * 256;
		if (at > arena_bytes || bytes > arena_bytes - at)
		{
			fail(REXLIB_HIP_INVALID_ARGUMENT, "workspace: the call needs " + std::to_string(at + bytes) +
				" bytes of workspace or more, beyond the " + std::to_string(arena_bytes) +
				" bytes of the fixed scratch buffer");
		}
		used = at + bytes;
		return arena + at;
	}

	// Bytes a take() can still get (the per-stream areas: no limit).
	std::size_t remaining() const
	{
		if (arena == nullptr)
		{
			return ~std::size_t(0);
		}
		const std::size_t at = (used + 255) / 256 * 256;
		return at >= arena_bytes ? 0 : arena_bytes - at;
	}
};

// Insertion schedules.
enum insertion_schedule : int
{
	schedule_auto = 0,
	schedule_scatter = 1,
	schedule_brick = 2
};

// Real images whose row transforms are still to run (the fused image
// insertion): image b has `m` rows of `in_rows` from element b * s_outer of
// `images`, and its n rows of half spectra go to `out_rows` of the spectra.
struct pending_rows
{
	int kind; // real_kind
	const void *images;
	std::size_t element_bytes;
	line_set in_rows;
	int m;
	const float *scales; // per-image scales of float16 images, or null
	int64_t scale_stride;
	line_set out_rows;
	int64_t n;
};

// Slices whose rows are transformed but whose columns are not yet (the
// fused image insertion): the insertion runs the column transforms, on the
// brick schedule together with the preparation of the samples. With `rows`,
// the row transforms are pending too, and the insertion runs them first
// (per chunk of slices, on the brick schedule).
struct pending_columns
{
	float2 *spectra;
	line_set lines;
	const pending_rows *rows = nullptr;
};

// `columns`, when not null, describes the slices operand (transformed in
// place where the transforms run on their own).
void run_insertion(insertion_params p, int schedule, const float2 *num_in, const int64_t *num_in_strides,
	const float *den_in, const int64_t *den_in_strides, bool in_place, double shape, const pending_columns *columns,
	workspace_source &workspace, hipStream_t stream);

// rexlib_hip_insert_ctf_slices with its workspace taken from `workspace`.
void insert_ctf_slices_within(const rexlib_hip_array *numerator, const rexlib_hip_array *denominator,
	const rexlib_hip_array *slices, const rexlib_hip_array *transforms, const rexlib_hip_array *ctf,
	const rexlib_hip_array *radial_weights, const rexlib_hip_window *window, const rexlib_hip_ctf_optics *optics,
	double pixel_size, const rexlib_hip_array *numerator_out, const rexlib_hip_array *denominator_out,
	workspace_source &workspace, void *stream);

// Brick extents along x, y and z and the plane hits per task of the brick
// schedule, per volume extent (0: every extent without an entry).
void set_insertion_tuning(int extent, int taps, int extent_x, int extent_y, int extent_z, int share);
void get_insertion_tuning(int extent, int *extent_x, int *extent_y, int *extent_z, int *share);
void clear_insertion_tuning();

// Event timing of single kernels (ids as REXLIB_HIP_KERNEL_*).
void set_kernel_timing(bool enable);
int timing_begin(int kernel, hipStream_t stream);
void timing_end(int token, hipStream_t stream);
void kernel_timing(int kernel, double *total_ms, int64_t *launches);

void launch_rfft_rows(int kind, const void *in, const line_set &in_rows, int m, const float *scales,
	int64_t scale_stride, float2 *out, const line_set &out_rows, int64_t n, hipStream_t stream);

void launch_fft_lines(const float2 *src, const line_set &src_lines, float2 *dst, const line_set &dst_lines,
	int64_t n, float sign, float scale, hipStream_t stream);

void launch_irfft_rows(const float2 *src, const line_set &src_rows, float *dst, const line_set &dst_rows,
	int64_t n, float scale, hipStream_t stream);

void launch_normalize(const float2 *num, const float *den, float2 *out, int64_t count, double threshold,
	hipStream_t stream);

void launch_gridding(const float *in, float *out, int big, int crop, int64_t volumes, int profile,
	double half_width, double shape, hipStream_t stream);

int mlp_parameter_count();
void launch_compensation(const float *den, float *w, float *scratch, int big, int passes, float floor_value,
	hipStream_t stream);
// Coverage table (8 radii x 32 theta x 64 phi x 8 angle bins, float32) of
// the slice normals of `slices` transforms (strides in elements).
void launch_coverage(const float *transforms, int64_t slices, int64_t ts0, int64_t ts1, int64_t ts2, int big,
	float *table, hipStream_t stream);
void launch_features(const float *den, const float *table, float density_scale, void *features, int big,
	hipStream_t stream);
void launch_mlp_forward(const void *features, const void *params_h, const float *params_f, float *out, int64_t count,
	hipStream_t stream);
void launch_mlp_backward(const void *features, const float *target, const void *params_h, const float *params_f,
	float *grads, float *loss, int64_t count, float loss_scale, hipStream_t stream);
void launch_adam(float *params, void *params_h, float *grads, float *m, float *v, float lr, float beta1, float beta2,
	float eps, int step, float loss_scale, hipStream_t stream);

void launch_shell_sums(const float2 *a, const float2 *b, double *sums, int big, hipStream_t stream);

} // namespace rexlib_hip
