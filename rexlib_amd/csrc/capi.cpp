// The C ABI of the gfx950 backend (include/rexlib_hip.h).
#include "launch.hpp"

#include <cmath>
#include <map>
#include <mutex>
#include <sstream>
#include <string>
#include <tuple>
#include <vector>

namespace rexlib_hip
{

namespace
{

thread_local std::string last_error;

struct area_entry
{
	void *pointer = nullptr;
	std::size_t bytes = 0;
};

std::mutex area_mutex;
std::map<std::tuple<int, hipStream_t, int>, area_entry> areas;
std::vector<void *> retired_areas;

} // anonymous namespace

void set_last_error(const std::string &message)
{
	last_error = message;
}

void *workspace_area_bytes(int area, std::size_t bytes, hipStream_t stream)
{
	int device = 0;
	REXLIB_HIP_CHECK(hipGetDevice(&device));
	std::lock_guard<std::mutex> lock(area_mutex);
	auto &entry = areas[std::make_tuple(device, stream, area)];
	if (entry.bytes < bytes)
	{
		// A graph captured on this stream holds the current area's address:
		// growing while a capture is under way would leave it pointing at a
		// smaller area, so the areas are sized by an eager call first.
		hipStreamCaptureStatus capture = hipStreamCaptureStatusNone;
		REXLIB_HIP_CHECK(hipStreamIsCapturing(stream, &capture));
		if (capture != hipStreamCaptureStatusNone)
		{
			fail(REXLIB_HIP_INVALID_ARGUMENT, "workspace: an insertion captured into a graph needs more workspace "
				"than an eager call of the same shape has sized on this stream; run it once eagerly before capture");
		}
		const std::size_t grown = std::max(bytes, entry.bytes + entry.bytes / 2);
		void *fresh = nullptr;
		REXLIB_HIP_CHECK(hipMalloc(&fresh, grown));
		if (entry.pointer != nullptr)
		{
			// Enqueued work and graphs captured earlier may still read the old
			// area, so it is kept until the process ends (the areas grow
			// geometrically, so the retired ones hold at most twice the last).
			retired_areas.push_back(entry.pointer);
		}
		entry.pointer = fresh;
		entry.bytes = grown;
	}
	return entry.pointer;
}

namespace
{

hipStream_t as_stream(void *stream)
{
	return static_cast<hipStream_t>(stream);
}

bool same_view(const rexlib_hip_array &a, const rexlib_hip_array &b)
{
	if (a.data != b.data || a.rank != b.rank)
	{
		return false;
	}
	for (int i = 0; i < a.rank; ++i)
	{
		if (a.extents[i] != b.extents[i] || (a.extents[i] > 1 && a.strides[i] != b.strides[i]))
		{
			return false;
		}
	}
	return true;
}

void check_same_shape(const rexlib_hip_array &a, const rexlib_hip_array &b, const char *what)
{
	bool same = a.rank == b.rank;
	for (int i = 0; same && i < a.rank; ++i)
	{
		same = a.extents[i] == b.extents[i];
	}
	if (!same)
	{
		fail(REXLIB_HIP_INVALID_ARGUMENT, std::string(what) + ": shape " + shape_string(b) +
			" differs from the accumulator's " + shape_string(a));
	}
}

// Collapse the leading `rank - core` axes into one (count, stride); false if
// they are not one arithmetic progression.
bool collapse_batch(const rexlib_hip_array &a, int core, int64_t &count, int64_t &stride)
{
	count = 1;
	stride = 0;
	int64_t expected = -1;
	for (int i = a.rank - core - 1; i >= 0; --i)
	{
		const int64_t e = a.extents[i];
		if (e == 1)
		{
			continue;
		}
		if (expected < 0)
		{
			stride = a.strides[i];
		}
		else if (a.strides[i] != expected)
		{
			return false;
		}
		count *= e;
		expected = a.strides[i] * e;
	}
	for (int i = 0; i < a.rank - core; ++i)
	{
		if (a.extents[i] == 0)
		{
			count = 0;
		}
	}
	return true;
}

void check_positive(double value, const char *message)
{
	if (!(std::isfinite(value) && value > 0.0))
	{
		std::ostringstream oss;
		oss << message << value << ".";
		fail(REXLIB_HIP_INVALID_ARGUMENT, oss.str());
	}
}

struct window_choice
{
	int profile;
	double half_width;
	double shape;
	int taps;
	int kb_count;
	float kb_scale;
};

// gridding_window's checks (src/em/ops/parameters/gridding_window.cpp:14-42) and
// the tap count of each profile.
window_choice choose_window(const rexlib_hip_window *window)
{
	if (window == nullptr)
	{
		fail(REXLIB_HIP_INVALID_ARGUMENT, "insertion: missing gridding window");
	}
	window_choice c{};
	c.profile = window->profile;
	switch (window->profile)
	{
	case REXLIB_HIP_BOX:
		c.half_width = 0.5;
		c.taps = 1;
		return c;
	case REXLIB_HIP_TRIANGLE:
		c.half_width = 1.0;
		c.taps = 2;
		return c;
	case REXLIB_HIP_KAISER_BESSEL:
		break;
	default:
		fail(REXLIB_HIP_INVALID_ARGUMENT, "insertion: unknown gridding profile");
	}
	if (!(std::isfinite(window->half_width) && window->half_width > 0.0))
	{
		std::ostringstream oss;
		oss << "gridding_window: the half-width of a Kaiser-Bessel window must be finite and positive, but it is "
			<< window->half_width << ".";
		fail(REXLIB_HIP_INVALID_ARGUMENT, oss.str());
	}
	if (!(std::isfinite(window->shape) && window->shape >= 0.0))
	{
		std::ostringstream oss;
		oss << "gridding_window: the shape of a Kaiser-Bessel window must be finite and 0 or more, but it is "
			<< window->shape << ".";
		fail(REXLIB_HIP_INVALID_ARGUMENT, oss.str());
	}
	c.half_width = window->half_width;
	c.shape = window->shape;
	c.taps = static_cast<int>(std::ceil(2.0 * window->half_width));
	if (c.taps > 8)
	{
		fail(REXLIB_HIP_UNSUPPORTED, "insertion: the gfx950 program takes Kaiser-Bessel windows of at most 8 taps");
	}
	// 1024 samples per voxel over the half-width (kaiser_bessel_window.hpp:37, .inl:31-35).
	c.kb_count = static_cast<int>(std::ceil(window->half_width * 1024.0));
	c.kb_scale = static_cast<float>(static_cast<double>(c.kb_count) / window->half_width);
	return c;
}

// ctf_optics's checks (src/em/ops/parameters/ctf_optics.cpp:15-56).
void check_optics(const rexlib_hip_ctf_optics *optics)
{
	if (optics == nullptr)
	{
		fail(REXLIB_HIP_INVALID_ARGUMENT, "insert_ctf_slices: missing optics");
	}
	std::ostringstream oss;
	if (!(std::isfinite(optics->voltage) && optics->voltage > 0.0))
	{
		oss << "ctf_optics: the voltage must be finite and positive, but it is " << optics->voltage << ".";
		fail(REXLIB_HIP_INVALID_ARGUMENT, oss.str());
	}
	if (!std::isfinite(optics->spherical_aberration))
	{
		oss << "ctf_optics: the spherical aberration must be finite, but it is " << optics->spherical_aberration
			<< ".";
		fail(REXLIB_HIP_INVALID_ARGUMENT, oss.str());
	}
	if (!(optics->amplitude_contrast >= 0.0 && optics->amplitude_contrast < 1.0))
	{
		oss << "ctf_optics: the amplitude contrast must lie in [0, 1), but it is " << optics->amplitude_contrast
			<< ".";
		fail(REXLIB_HIP_INVALID_ARGUMENT, oss.str());
	}
}

struct insertion_request
{
	int variant;
	const rexlib_hip_array *ops[slot_count];
	const rexlib_hip_window *window;
	const rexlib_hip_ctf_optics *optics;
	double pixel_size;
	int schedule;
	hipStream_t stream;
	const pending_columns *columns = nullptr;
	workspace_source *workspace = nullptr; // null: the per-stream areas
};

// The shared body of the three insertions.
void insert(const insertion_request &r)
{
	if (r.variant == variant_ctf)
	{
		check_positive(r.pixel_size, "insert_ctf_slices: the pixel size must be finite and positive, but it is ");
		check_optics(r.optics);
	}
	const window_choice w = choose_window(r.window);

	static const char *const names[slot_count] = {"numerator", "denominator", "slice", "transform",
		"radial_weight", "extra", "numerator_out", "denominator_out"};
	const char *const extra_name = r.variant == variant_ctf ? "ctf" : "weight";
	const int core_ranks[slot_count] = {3, 3, 2, 2, 1, r.variant == variant_ctf ? 1 : 2, 3, 3};
	const rexlib_hip_array *ops[slot_count];
	for (int s = 0; s < slot_count; ++s)
	{
		ops[s] = r.ops[s];
	}
	if (ops[slot_numerator_out] == nullptr)
	{
		ops[slot_numerator_out] = ops[slot_numerator];
	}
	if (ops[slot_denominator_out] == nullptr)
	{
		ops[slot_denominator_out] = ops[slot_denominator];
	}
	for (int s = 0; s < slot_count; ++s)
	{
		if (s == slot_extra && r.variant == variant_plain)
		{
			continue;
		}
		check_array(ops[s], s == slot_extra ? extra_name : names[s], core_ranks[s]);
	}

	const bool complex_slot[slot_count] = {true, false, true, false, false, false, true, false};
	const char *type_names[slot_count];
	for (int s = 0; s < slot_count; ++s)
	{
		type_names[s] = s == slot_extra ? extra_name : names[s];
	}
	check_types(ops, complex_slot, slot_count, type_names);

	const rexlib_hip_array &num = *ops[slot_numerator];
	const rexlib_hip_array &sl = *ops[slot_slices];
	const int64_t big = core_extent(num, 3);
	const int64_t n = core_extent(sl, 2);
	if (big <= 0 || n <= 0)
	{
		fail(REXLIB_HIP_INVALID_ARGUMENT, "insertion: the volume and slice extents must be positive");
	}
	if (big > 16384 || n > 16384)
	{
		fail(REXLIB_HIP_UNSUPPORTED, "insertion: extents above 16384 have no gfx950 program");
	}
	const int64_t half = big / 2 + 1;
	const int64_t h = n / 2 + 1;
	expect_core(num, names[slot_numerator], {big, big, half});
	expect_core(*ops[slot_denominator], names[slot_denominator], {big, big, half});
	expect_core(sl, names[slot_slices], {n, h});
	expect_core(*ops[slot_transforms], names[slot_transforms], {3, 4});
	expect_core(*ops[slot_radial], names[slot_radial], {h});
	if (r.variant == variant_ctf)
	{
		expect_core(*ops[slot_extra], extra_name, {4});
	}
	else if (r.variant == variant_weighted)
	{
		expect_core(*ops[slot_extra], extra_name, {n, h});
	}
	check_same_shape(num, *ops[slot_denominator], names[slot_denominator]);
	check_same_shape(num, *ops[slot_numerator_out], names[slot_numerator_out]);
	check_same_shape(num, *ops[slot_denominator_out], names[slot_denominator_out]);

	insertion_params p{};
	const int accumulator_slot_of[slot_count] = {};
	p.map = make_batch_map(ops, core_ranks, accumulator_slot_of);
	p.n = static_cast<int32_t>(n);
	p.h = static_cast<int32_t>(h);
	p.big = static_cast<int32_t>(big);
	p.half = static_cast<int32_t>(half);
	p.variant = r.variant;
	p.profile = w.profile;
	p.taps = w.taps;
	p.kb_count = w.kb_count;
	p.kb_half_width = static_cast<float>(w.half_width);
	p.kb_scale = w.kb_scale;
	p.slice_count = p.map.kept_count * p.map.reduced_count;
	p.slices = static_cast<const float2 *>(sl.data);
	p.slice_s0 = core_stride(sl, 2);
	p.slice_s1 = core_stride(sl, 1);
	p.transforms = static_cast<const float *>(ops[slot_transforms]->data);
	p.transform_s0 = core_stride(*ops[slot_transforms], 2);
	p.transform_s1 = core_stride(*ops[slot_transforms], 1);
	p.radial = static_cast<const float *>(ops[slot_radial]->data);
	p.radial_s0 = core_stride(*ops[slot_radial], 1);
	if (r.variant != variant_plain)
	{
		p.extra = static_cast<const float *>(ops[slot_extra]->data);
		if (r.variant == variant_ctf)
		{
			p.extra_s0 = core_stride(*ops[slot_extra], 1);
		}
		else
		{
			p.extra_s0 = core_stride(*ops[slot_extra], 2);
			p.extra_s1 = core_stride(*ops[slot_extra], 1);
		}
	}
	const rexlib_hip_array &num_out = *ops[slot_numerator_out];
	const rexlib_hip_array &den_out = *ops[slot_denominator_out];
	p.numerator = static_cast<float2 *>(num_out.data);
	p.num_s0 = core_stride(num_out, 3);
	p.num_s1 = core_stride(num_out, 2);
	p.num_s2 = core_stride(num_out, 1);
	p.denominator = static_cast<float *>(den_out.data);
	p.den_s0 = core_stride(den_out, 3);
	p.den_s1 = core_stride(den_out, 2);
	p.den_s2 = core_stride(den_out, 1);
	if (r.variant == variant_ctf)
	{
		// contrast_transfer_function.inl:25-70, in double.
		const double pi = 3.14159265358979323846;
		const double volts = r.optics->voltage * 1e3;
		const double wavelength = 12.2643247 / std::sqrt(volts * (1.0 + 0.978466e-6 * volts));
		const double aberration = r.optics->spherical_aberration * 1e7;
		const double frequency_scale = 1.0 / (static_cast<double>(n) * r.pixel_size);
		const double squared_scale = frequency_scale * frequency_scale;
		p.ctf_defocus_scale = pi * wavelength * squared_scale;
		p.ctf_aberration = 0.5 * pi * aberration * wavelength * wavelength * wavelength * squared_scale *
			squared_scale;
		p.ctf_phase0 = std::asin(r.optics->amplitude_contrast);
	}

	const bool in_place = same_view(num, num_out) && same_view(*ops[slot_denominator], den_out);
	const int64_t num_in_strides[3] = {core_stride(num, 3), core_stride(num, 2), core_stride(num, 1)};
	const rexlib_hip_array &den = *ops[slot_denominator];
	const int64_t den_in_strides[3] = {core_stride(den, 3), core_stride(den, 2), core_stride(den, 1)};
	workspace_source areas;
	run_insertion(p, r.schedule, static_cast<const float2 *>(num.data), num_in_strides,
		static_cast<const float *>(den.data), den_in_strides, in_place, w.shape, r.columns,
		r.workspace != nullptr ? *r.workspace : areas, r.stream);
}

int insertion_schedule_from_env()
{
	const char *value = std::getenv("REXLIB_HIP_INSERTION_SCHEDULE");
	if (value == nullptr)
	{
		return schedule_auto;
	}
	const std::string s(value);
	if (s == "scatter")
	{
		return schedule_scatter;
	}
	if (s == "brick")
	{
		return schedule_brick;
	}
	return schedule_auto;
}

void expect_dtype(const rexlib_hip_array &a, int dtype, const char *what)
{
	if (a.dtype != dtype)
	{
		fail(REXLIB_HIP_INVALID_ARGUMENT, std::string(what) + ": expected " + dtype_name(dtype) + ", got " +
			dtype_name(a.dtype));
	}
}

bool contiguous(const rexlib_hip_array &a)
{
	int64_t expected = 1;
	for (int i = a.rank - 1; i >= 0; --i)
	{
		if (a.extents[i] != 1 && a.strides[i] != expected)
		{
			return false;
		}
		expected *= a.extents[i];
	}
	return true;
}

int64_t check_mlp_operands(const rexlib_hip_array *features, const rexlib_hip_array *params_h,
	const rexlib_hip_array *params_f)
{
	check_array(features, "mlp features", 2);
	check_array(params_h, "mlp float16 parameters", 1);
	check_array(params_f, "mlp float32 parameters", 1);
	expect_dtype(*features, REXLIB_HIP_FLOAT16, "mlp features");
	expect_dtype(*params_h, REXLIB_HIP_FLOAT16, "mlp float16 parameters");
	expect_dtype(*params_f, REXLIB_HIP_FLOAT32, "mlp float32 parameters");
	const int64_t p = rexlib_hip::mlp_parameter_count();
	if (features->rank != 2 || features->extents[1] != 16 || !contiguous(*features))
	{
		fail(REXLIB_HIP_INVALID_ARGUMENT, "mlp: features are a contiguous (rows, 16) float16 array");
	}
	if (params_h->extents[0] != p || params_f->extents[0] != p || !contiguous(*params_h) || !contiguous(*params_f))
	{
		fail(REXLIB_HIP_INVALID_ARGUMENT, "mlp: parameters hold rexlib_hip_mlp_parameter_count() elements");
	}
	return features->extents[0];
}

} // anonymous namespace

} // namespace rexlib_hip

using namespace rexlib_hip;

extern "C" {

const char *rexlib_hip_last_error(void)
{
	return rexlib_hip::last_error.c_str();
}

void rexlib_hip_version(int32_t *major, int32_t *minor, int32_t *patch)
{
	if (major)
	{
		*major = 0;
	}
	if (minor)
	{
		*minor = 1;
	}
	if (patch)
	{
		*patch = 0;
	}
}

int rexlib_hip_set_insertion_tuning(int32_t extent, int32_t extent_x, int32_t extent_y, int32_t extent_z,
	int32_t share)
{
	return guarded([&] { rexlib_hip::set_insertion_tuning(extent, 0, extent_x, extent_y, extent_z, share); });
}

int rexlib_hip_set_insertion_tuning_for_window(int32_t extent, int32_t taps, int32_t extent_x, int32_t extent_y,
	int32_t extent_z, int32_t share)
{
	return guarded([&] { rexlib_hip::set_insertion_tuning(extent, taps, extent_x, extent_y, extent_z, share); });
}

int rexlib_hip_get_insertion_tuning(int32_t extent, int32_t *extent_x, int32_t *extent_y, int32_t *extent_z,
	int32_t *share)
{
	return guarded([&]
		{
			if (!extent_x || !extent_y || !extent_z || !share)
			{
				throw rexlib_hip::error(REXLIB_HIP_INVALID_ARGUMENT, "get_insertion_tuning: null output");
			}
			int ex = 0, ey = 0, ez = 0, sh = 0;
			rexlib_hip::get_insertion_tuning(extent, &ex, &ey, &ez, &sh);
			*extent_x = ex;
			*extent_y = ey;
			*extent_z = ez;
			*share = sh;
		});
}

void rexlib_hip_clear_insertion_tuning(void)
{
	rexlib_hip::clear_insertion_tuning();
}

int rexlib_hip_set_kernel_timing(int32_t enable)
{
	return guarded([&] { rexlib_hip::set_kernel_timing(enable != 0); });
}

int rexlib_hip_kernel_timing(int32_t kernel, double *total_ms, int64_t *launches)
{
	return guarded([&]
		{
			if (!total_ms || !launches)
			{
				throw rexlib_hip::error(REXLIB_HIP_INVALID_ARGUMENT, "kernel_timing: null output");
			}
			if (kernel < REXLIB_HIP_KERNEL_BRICK || kernel > REXLIB_HIP_KERNEL_SCATTER)
			{
				throw rexlib_hip::error(REXLIB_HIP_INVALID_ARGUMENT, "kernel_timing: unknown kernel id");
			}
			rexlib_hip::kernel_timing(kernel, total_ms, launches);
		});
}

int32_t rexlib_hip_device_count(void)
{
	int count = 0;
	if (hipGetDeviceCount(&count) != hipSuccess)
	{
		return 0;
	}
	return count;
}

int rexlib_hip_insert_slices(const rexlib_hip_array *numerator, const rexlib_hip_array *denominator,
	const rexlib_hip_array *slices, const rexlib_hip_array *transforms, const rexlib_hip_array *radial_weights,
	const rexlib_hip_window *window, const rexlib_hip_array *numerator_out, const rexlib_hip_array *denominator_out,
	void *stream)
{
	return guarded([&]
	{
		insertion_request r{variant_plain,
			{numerator, denominator, slices, transforms, radial_weights, nullptr, numerator_out, denominator_out},
			window, nullptr, 1.0, insertion_schedule_from_env(), as_stream(stream)};
		insert(r);
	});
}

int rexlib_hip_insert_weighted_slices(const rexlib_hip_array *numerator, const rexlib_hip_array *denominator,
	const rexlib_hip_array *slices, const rexlib_hip_array *weights, const rexlib_hip_array *transforms,
	const rexlib_hip_array *radial_weights, const rexlib_hip_window *window, const rexlib_hip_array *numerator_out,
	const rexlib_hip_array *denominator_out, void *stream)
{
	return guarded([&]
	{
		insertion_request r{variant_weighted,
			{numerator, denominator, slices, transforms, radial_weights, weights, numerator_out, denominator_out},
			window, nullptr, 1.0, insertion_schedule_from_env(), as_stream(stream)};
		insert(r);
	});
}

int rexlib_hip_insert_ctf_slices(const rexlib_hip_array *numerator, const rexlib_hip_array *denominator,
	const rexlib_hip_array *slices, const rexlib_hip_array *transforms, const rexlib_hip_array *ctf,
	const rexlib_hip_array *radial_weights, const rexlib_hip_window *window, const rexlib_hip_ctf_optics *optics,
	double pixel_size, const rexlib_hip_array *numerator_out, const rexlib_hip_array *denominator_out, void *stream)
{
	return guarded([&]
	{
		insertion_request r{variant_ctf,
			{numerator, denominator, slices, transforms, radial_weights, ctf, numerator_out, denominator_out},
			window, optics, pixel_size, insertion_schedule_from_env(), as_stream(stream)};
		insert(r);
	});
}

} // extern "C"

namespace rexlib_hip
{

void insert_ctf_slices_within(const rexlib_hip_array *numerator, const rexlib_hip_array *denominator,
	const rexlib_hip_array *slices, const rexlib_hip_array *transforms, const rexlib_hip_array *ctf,
	const rexlib_hip_array *radial_weights, const rexlib_hip_window *window, const rexlib_hip_ctf_optics *optics,
	double pixel_size, const rexlib_hip_array *numerator_out, const rexlib_hip_array *denominator_out,
	workspace_source &workspace, void *stream)
{
	insertion_request r{variant_ctf,
		{numerator, denominator, slices, transforms, radial_weights, ctf, numerator_out, denominator_out},
		window, optics, pixel_size, insertion_schedule_from_env(), as_stream(stream)};
	r.workspace = &workspace;
	insert(r);
}

} // namespace rexlib_hip

extern "C" {

int rexlib_hip_rfft2(const rexlib_hip_array *images, const rexlib_hip_array *spectra, void *stream)
{
	return guarded([&]
	{
		check_array(images, "rfft2 input", 2);
		check_array(spectra, "rfft2 output", 2);
		expect_dtype(*images, REXLIB_HIP_FLOAT32, "rfft2 input");
		expect_dtype(*spectra, REXLIB_HIP_COMPLEX64, "rfft2 output");
		const int64_t rows = core_extent(*images, 2);
		const int64_t n = core_extent(*images, 1);
		if (rows != n)
		{
			fail(REXLIB_HIP_UNSUPPORTED, "rfft2: the gfx950 program takes square images");
		}
		expect_core(*spectra, "rfft2 output", {n, n / 2 + 1});
		int64_t batch = 0, in_stride = 0, out_batch = 0, out_stride = 0;
		if (!collapse_batch(*images, 2, batch, in_stride) || !collapse_batch(*spectra, 2, out_batch, out_stride) ||
			batch != out_batch)
		{
			fail(REXLIB_HIP_UNSUPPORTED, "rfft2: batch axes must collapse to one strided axis");
		}
		if (batch == 0 || n == 0)
		{
			return;
		}
		const hipStream_t s = as_stream(stream);
		float2 *out = static_cast<float2 *>(spectra->data);
		const line_set in_rows{batch * n, 1, n, in_stride, 0, core_stride(*images, 2), core_stride(*images, 1)};
		const line_set out_rows{batch * n, 1, n, out_stride, 0, core_stride(*spectra, 2), core_stride(*spectra, 1)};
		launch_rfft_rows(real_f32, images->data, in_rows, static_cast<int>(n), nullptr, 0, out, out_rows, n, s);
		const int64_t h = n / 2 + 1;
		const line_set columns{batch * h, 1, h, out_stride, 0, core_stride(*spectra, 1), core_stride(*spectra, 2)};
		launch_fft_lines(out, columns, out, columns, n, -1.0f, 1.0f, s);
	});
}

int rexlib_hip_irfft2(const rexlib_hip_array *spectra, const rexlib_hip_array *images, void *stream)
{
	return guarded([&]
	{
		check_array(spectra, "irfft2 input", 2);
		check_array(images, "irfft2 output", 2);
		expect_dtype(*spectra, REXLIB_HIP_COMPLEX64, "irfft2 input");
		expect_dtype(*images, REXLIB_HIP_FLOAT32, "irfft2 output");
		const int64_t n = core_extent(*images, 1);
		if (core_extent(*images, 2) != n)
		{
			fail(REXLIB_HIP_UNSUPPORTED, "irfft2: the gfx950 program takes square images");
		}
		const int64_t h = n / 2 + 1;
		expect_core(*spectra, "irfft2 input", {n, h});
		int64_t batch = 0, in_stride = 0, out_batch = 0, out_stride = 0;
		if (!collapse_batch(*spectra, 2, batch, in_stride) || !collapse_batch(*images, 2, out_batch, out_stride) ||
			batch != out_batch)
		{
			fail(REXLIB_HIP_UNSUPPORTED, "irfft2: batch axes must collapse to one strided axis");
		}
		if (batch == 0 || n == 0)
		{
			return;
		}
		const hipStream_t s = as_stream(stream);
		float2 *w = static_cast<float2 *>(
			workspace_area_bytes(area_fft, sizeof(float2) * static_cast<std::size_t>(batch * n * h), s));
		const line_set columns_in{batch * h, 1, h, in_stride, 0, core_stride(*spectra, 1), core_stride(*spectra, 2)};
		const line_set columns_w{batch * h, 1, h, n * h, 0, 1, h};
		launch_fft_lines(static_cast<const float2 *>(spectra->data), columns_in, w, columns_w, n, 1.0f, 1.0f, s);
		const line_set rows_w{batch * n, 1, n, n * h, 0, h, 1};
		const line_set rows_out{batch * n, 1, n, out_stride, 0, core_stride(*images, 2), core_stride(*images, 1)};
		launch_irfft_rows(w, rows_w, static_cast<float *>(images->data), rows_out, n,
			static_cast<float>(1.0 / (static_cast<double>(n) * static_cast<double>(n))), s);
	});
}

int rexlib_hip_irfft3(const rexlib_hip_array *spectrum, const rexlib_hip_array *volume, void *stream)
{
	return guarded([&]
	{
		check_array(spectrum, "irfft3 input", 3);
		check_array(volume, "irfft3 output", 3);
		expect_dtype(*spectrum, REXLIB_HIP_COMPLEX64, "irfft3 input");
		expect_dtype(*volume, REXLIB_HIP_FLOAT32, "irfft3 output");
		const int64_t big = core_extent(*volume, 1);
		if (core_extent(*volume, 2) != big || core_extent(*volume, 3) != big)
		{
			fail(REXLIB_HIP_UNSUPPORTED, "irfft3: the gfx950 program takes cubic volumes");
		}
		const int64_t half = big / 2 + 1;
		expect_core(*spectrum, "irfft3 input", {big, big, half});
		int64_t batch = 0, in_stride = 0, out_batch = 0, out_stride = 0;
		if (!collapse_batch(*spectrum, 3, batch, in_stride) || !collapse_batch(*volume, 3, out_batch, out_stride) ||
			batch != out_batch)
		{
			fail(REXLIB_HIP_UNSUPPORTED, "irfft3: batch axes must collapse to one strided axis");
		}
		if (batch == 0 || big == 0)
		{
			return;
		}
		const hipStream_t s = as_stream(stream);
		const int64_t plane = big * half;
		const int64_t vol = big * plane;
		float2 *w = static_cast<float2 *>(
			workspace_area_bytes(area_fft, sizeof(float2) * static_cast<std::size_t>(batch * vol), s));
		// z: lines over (batch, y, x), from the input into the workspace.
		const line_set z_in{batch * plane, big, half, in_stride, core_stride(*spectrum, 2), core_stride(*spectrum, 1),
			core_stride(*spectrum, 3)};
		const line_set z_w{batch * plane, big, half, vol, half, 1, plane};
		launch_fft_lines(static_cast<const float2 *>(spectrum->data), z_in, w, z_w, big, 1.0f, 1.0f, s);
		// y: lines over (batch, z, x), in place.
		const line_set y_w{batch * plane, big, half, vol, plane, 1, half};
		launch_fft_lines(w, y_w, w, y_w, big, 1.0f, 1.0f, s);
		// x: rows over (batch, z, y), complex to real.
		const line_set x_w{batch * big * big, big, big, vol, plane, half, 1};
		const line_set x_out{batch * big * big, big, big, out_stride, core_stride(*volume, 3),
			core_stride(*volume, 2), core_stride(*volume, 1)};
		const double scale = 1.0 / (static_cast<double>(big) * static_cast<double>(big) * static_cast<double>(big));
		launch_irfft_rows(w, x_w, static_cast<float *>(volume->data), x_out, big, static_cast<float>(scale), s);
	});
}

int rexlib_hip_normalize(const rexlib_hip_array *numerator, const rexlib_hip_array *denominator, double threshold,
	const rexlib_hip_array *out, void *stream)
{
	return guarded([&]
	{
		check_array(numerator, "normalize numerator", 3);
		check_array(denominator, "normalize denominator", 3);
		check_array(out, "normalize output", 3);
		expect_dtype(*numerator, REXLIB_HIP_COMPLEX64, "normalize numerator");
		expect_dtype(*denominator, REXLIB_HIP_FLOAT32, "normalize denominator");
		expect_dtype(*out, REXLIB_HIP_COMPLEX64, "normalize output");
		const int64_t big = core_extent(*numerator, 3);
		expect_core(*numerator, "normalize numerator", {big, big, big / 2 + 1});
		check_same_shape(*numerator, *denominator, "normalize denominator");
		check_same_shape(*numerator, *out, "normalize output");
		if (!contiguous(*numerator) || !contiguous(*denominator) || !contiguous(*out))
		{
			fail(REXLIB_HIP_UNSUPPORTED, "normalize: the gfx950 program takes contiguous operands");
		}
		if (!std::isfinite(threshold))
		{
			fail(REXLIB_HIP_INVALID_ARGUMENT, "normalize: the threshold must be finite");
		}
		int64_t count = 1;
		for (int i = 0; i < numerator->rank; ++i)
		{
			count *= numerator->extents[i];
		}
		launch_normalize(static_cast<const float2 *>(numerator->data), static_cast<const float *>(denominator->data),
			static_cast<float2 *>(out->data), count, threshold, as_stream(stream));
	});
}

int rexlib_hip_gridding_correct(const rexlib_hip_array *volume, const rexlib_hip_window *window,
	const rexlib_hip_array *out, void *stream)
{
	return guarded([&]
	{
		check_array(volume, "gridding_correct input", 3);
		check_array(out, "gridding_correct output", 3);
		expect_dtype(*volume, REXLIB_HIP_FLOAT32, "gridding_correct input");
		expect_dtype(*out, REXLIB_HIP_FLOAT32, "gridding_correct output");
		const window_choice w = choose_window(window);
		const int64_t big = core_extent(*volume, 1);
		const int64_t crop = core_extent(*out, 1);
		expect_core(*volume, "gridding_correct input", {big, big, big});
		expect_core(*out, "gridding_correct output", {crop, crop, crop});
		if (crop > big)
		{
			fail(REXLIB_HIP_INVALID_ARGUMENT, "gridding_correct: the output extent exceeds the volume's");
		}
		if (volume->rank != out->rank)
		{
			fail(REXLIB_HIP_INVALID_ARGUMENT, "gridding_correct: input and output batch shapes differ");
		}
		int64_t volumes = 1;
		for (int i = 0; i < volume->rank - 3; ++i)
		{
			if (volume->extents[i] != out->extents[i])
			{
				fail(REXLIB_HIP_INVALID_ARGUMENT, "gridding_correct: input and output batch shapes differ");
			}
			volumes *= volume->extents[i];
		}
		if (!contiguous(*volume) || !contiguous(*out))
		{
			fail(REXLIB_HIP_UNSUPPORTED, "gridding_correct: the gfx950 program takes contiguous operands");
		}
		if (crop < big && volume->data == out->data)
		{
			fail(REXLIB_HIP_INVALID_ARGUMENT, "gridding_correct: a cropped output cannot alias the input");
		}
		if (volumes == 0 || crop == 0)
		{
			return;
		}
		launch_gridding(static_cast<const float *>(volume->data), static_cast<float *>(out->data),
			static_cast<int>(big), static_cast<int>(crop), volumes, w.profile, w.half_width, w.shape,
			as_stream(stream));
	});
}

int rexlib_hip_shell_sums(const rexlib_hip_array *a, const rexlib_hip_array *b, double *sums, void *stream)
{
	return guarded([&]
	{
		check_array(a, "shell_sums a", 3);
		check_array(b, "shell_sums b", 3);
		expect_dtype(*a, REXLIB_HIP_COMPLEX64, "shell_sums a");
		expect_dtype(*b, REXLIB_HIP_COMPLEX64, "shell_sums b");
		check_same_shape(*a, *b, "shell_sums b");
		const int64_t big = core_extent(*a, 3);
		expect_core(*a, "shell_sums a", {big, big, big / 2 + 1});
		if (a->rank != 3 || !contiguous(*a) || !contiguous(*b))
		{
			fail(REXLIB_HIP_UNSUPPORTED, "shell_sums: one contiguous (N, N, N/2+1) grid per operand");
		}
		if (sums == nullptr)
		{
			fail(REXLIB_HIP_INVALID_ARGUMENT, "shell_sums: null output");
		}
		launch_shell_sums(static_cast<const float2 *>(a->data), static_cast<const float2 *>(b->data), sums,
			static_cast<int>(big), as_stream(stream));
	});
}

int rexlib_hip_insert_ctf_images(const rexlib_hip_array *numerator, const rexlib_hip_array *denominator,
	const rexlib_hip_array *images, const rexlib_hip_array *image_scales, int64_t padded,
	const rexlib_hip_array *transforms, const rexlib_hip_array *ctf, const rexlib_hip_array *radial_weights,
	const rexlib_hip_window *window, const rexlib_hip_ctf_optics *optics, double pixel_size, void *stream)
{
	return guarded([&]
	{
		check_array(images, "images", 2);
		const int64_t m = core_extent(*images, 1);
		if (core_extent(*images, 2) != m)
		{
			fail(REXLIB_HIP_UNSUPPORTED, "insert_ctf_images: images must be square");
		}
		const int64_t n = padded == 0 ? m : padded;
		if (n < m)
		{
			fail(REXLIB_HIP_INVALID_ARGUMENT, "insert_ctf_images: the padded extent is below the image extent");
		}
		int kind = real_f32;
		switch (images->dtype)
		{
		case REXLIB_HIP_FLOAT32: kind = real_f32; break;
		case REXLIB_HIP_BFLOAT16: kind = real_bf16; break;
		case REXLIB_HIP_FLOAT16: kind = real_f16; break;
		default: fail(REXLIB_HIP_INVALID_ARGUMENT, "insert_ctf_images: images must be float32, bfloat16 or float16");
		}
		int64_t batch = 0, in_stride = 0;
		if (!collapse_batch(*images, 2, batch, in_stride))
		{
			fail(REXLIB_HIP_UNSUPPORTED, "insert_ctf_images: batch axes must collapse to one strided axis");
		}
		const float *scales = nullptr;
		int64_t scale_stride = 0;
		if (image_scales != nullptr)
		{
			expect_dtype(*image_scales, REXLIB_HIP_FLOAT32, "image scales");
			int64_t scale_count = 0;
			if (!collapse_batch(*image_scales, 0, scale_count, scale_stride) || scale_count != batch)
			{
				fail(REXLIB_HIP_INVALID_ARGUMENT, "insert_ctf_images: one scale per image");
			}
			scales = static_cast<const float *>(image_scales->data);
		}
		const hipStream_t s = as_stream(stream);
		const int64_t h = n / 2 + 1;
		// Rows of the half spectra start every 16 elements (128 bytes), so
		// the column passes read whole cache lines.
		const int64_t pitch = (h + 15) / 16 * 16;
		float2 *spectra = static_cast<float2 *>(workspace_area_bytes(area_spectra,
			sizeof(float2) * static_cast<std::size_t>(std::max<int64_t>(batch, 1) * n * pitch), s));
		// The row and column transforms are left to the insertion, which runs
		// them chunk by chunk beside the insertion of the chunk before.
		const pending_rows rows{kind, images->data, kind == real_f32 ? sizeof(float) : sizeof(uint16_t),
			line_set{batch * m, 1, m, in_stride, 0, core_stride(*images, 2), core_stride(*images, 1)},
			static_cast<int>(m), scales, scale_stride, line_set{batch * n, 1, n, n * pitch, 0, pitch, 1}, n};
		const pending_columns columns{spectra, line_set{batch * h, 1, h, n * pitch, 0, 1, pitch}, &rows};
		rexlib_hip_array slices{};
		slices.data = spectra;
		slices.dtype = REXLIB_HIP_COMPLEX64;
		slices.rank = 3;
		slices.extents[0] = batch;
		slices.extents[1] = n;
		slices.extents[2] = h;
		slices.strides[0] = n * pitch;
		slices.strides[1] = pitch;
		slices.strides[2] = 1;
		insertion_request r{variant_ctf,
			{numerator, denominator, &slices, transforms, radial_weights, ctf, numerator, denominator},
			window, optics, pixel_size, insertion_schedule_from_env(), s, &columns};
		insert(r);
	});
}

int rexlib_hip_density_compensation(const rexlib_hip_array *denominator, int32_t passes, double floor_value,
	const rexlib_hip_array *weights, void *stream)
{
	return guarded([&]
	{
		check_array(denominator, "density_compensation denominator", 3);
		check_array(weights, "density_compensation weights", 3);
		expect_dtype(*denominator, REXLIB_HIP_FLOAT32, "density_compensation denominator");
		expect_dtype(*weights, REXLIB_HIP_FLOAT32, "density_compensation weights");
		check_same_shape(*denominator, *weights, "density_compensation weights");
		const int64_t big = core_extent(*denominator, 3);
		expect_core(*denominator, "density_compensation denominator", {big, big, big / 2 + 1});
		if (denominator->rank != 3 || !contiguous(*denominator) || !contiguous(*weights))
		{
			fail(REXLIB_HIP_UNSUPPORTED, "density_compensation: one contiguous (N, N, N/2+1) grid");
		}
		if (passes < 0)
		{
			fail(REXLIB_HIP_INVALID_ARGUMENT, "density_compensation: the number of passes must be 0 or more");
		}
		if (!(floor_value > 0.0) || !std::isfinite(floor_value))
		{
			fail(REXLIB_HIP_INVALID_ARGUMENT, "density_compensation: the floor must be finite and positive");
		}
		const hipStream_t s = as_stream(stream);
		const int64_t count = big * big * (big / 2 + 1);
		float *scratch = static_cast<float *>(workspace_area_bytes(area_compensation,
			sizeof(float) * static_cast<std::size_t>(count), s));
		launch_compensation(static_cast<const float *>(denominator->data), static_cast<float *>(weights->data),
			scratch, static_cast<int>(big), passes, static_cast<float>(floor_value), s);
	});
}

int rexlib_hip_coverage_table(const rexlib_hip_array *transforms, int32_t big, const rexlib_hip_array *table,
	void *stream)
{
	return guarded([&]
	{
		check_array(transforms, "coverage_table transforms", 2);
		check_array(table, "coverage_table output", 4);
		expect_dtype(*transforms, REXLIB_HIP_FLOAT32, "coverage_table transforms");
		expect_dtype(*table, REXLIB_HIP_FLOAT32, "coverage_table output");
		expect_core(*transforms, "coverage_table transforms", {3, 4});
		expect_core(*table, "coverage_table output", {8, 33, 64, 8});
		if (table->rank != 4 || !contiguous(*table))
		{
			fail(REXLIB_HIP_UNSUPPORTED, "coverage_table: a contiguous (8, 33, 64, 8) table");
		}
		if (big < 2)
		{
			fail(REXLIB_HIP_INVALID_ARGUMENT, "coverage_table: the volume extent must be at least 2");
		}
		int64_t slices = 0, stride = 0;
		if (!collapse_batch(*transforms, 2, slices, stride))
		{
			fail(REXLIB_HIP_UNSUPPORTED, "coverage_table: batch axes must collapse to one strided axis");
		}
		launch_coverage(static_cast<const float *>(transforms->data), slices, stride, core_stride(*transforms, 2),
			core_stride(*transforms, 1), big, static_cast<float *>(table->data), as_stream(stream));
	});
}

int rexlib_hip_compensation_features(const rexlib_hip_array *denominator, const rexlib_hip_array *table,
	double density_scale, const rexlib_hip_array *features, void *stream)
{
	return guarded([&]
	{
		check_array(denominator, "compensation_features denominator", 3);
		check_array(table, "compensation_features table", 4);
		check_array(features, "compensation_features output", 4);
		expect_dtype(*denominator, REXLIB_HIP_FLOAT32, "compensation_features denominator");
		expect_dtype(*table, REXLIB_HIP_FLOAT32, "compensation_features table");
		expect_dtype(*features, REXLIB_HIP_FLOAT16, "compensation_features output");
		const int64_t big = core_extent(*denominator, 3);
		expect_core(*denominator, "compensation_features denominator", {big, big, big / 2 + 1});
		expect_core(*table, "compensation_features table", {8, 33, 64, 8});
		expect_core(*features, "compensation_features output", {big, big, big / 2 + 1, 16});
		if (denominator->rank != 3 || !contiguous(*denominator) || !contiguous(*features) || !contiguous(*table))
		{
			fail(REXLIB_HIP_UNSUPPORTED, "compensation_features: contiguous operands of one grid");
		}
		if (!(density_scale > 0.0) || !std::isfinite(density_scale))
		{
			fail(REXLIB_HIP_INVALID_ARGUMENT, "compensation_features: the density scale must be finite and positive");
		}
		launch_features(static_cast<const float *>(denominator->data), static_cast<const float *>(table->data),
			static_cast<float>(density_scale), features->data, static_cast<int>(big), as_stream(stream));
	});
}

int32_t rexlib_hip_mlp_parameter_count(void)
{
	return rexlib_hip::mlp_parameter_count();
}

int rexlib_hip_mlp_forward(const rexlib_hip_array *features, const rexlib_hip_array *parameters_f16,
	const rexlib_hip_array *parameters_f32, const rexlib_hip_array *out, void *stream)
{
	return guarded([&]
	{
		const int64_t rows = check_mlp_operands(features, parameters_f16, parameters_f32);
		check_array(out, "mlp output", 1);
		expect_dtype(*out, REXLIB_HIP_FLOAT32, "mlp output");
		if (out->rank != 1 || out->extents[0] != rows || !contiguous(*out))
		{
			fail(REXLIB_HIP_INVALID_ARGUMENT, "mlp_forward: the output holds one float32 per feature row");
		}
		launch_mlp_forward(features->data, parameters_f16->data, static_cast<const float *>(parameters_f32->data),
			static_cast<float *>(out->data), rows, as_stream(stream));
	});
}

int rexlib_hip_mlp_backward(const rexlib_hip_array *features, const rexlib_hip_array *target,
	const rexlib_hip_array *parameters_f16, const rexlib_hip_array *parameters_f32, const rexlib_hip_array *gradients,
	const rexlib_hip_array *loss, double loss_scale, void *stream)
{
	return guarded([&]
	{
		const int64_t rows = check_mlp_operands(features, parameters_f16, parameters_f32);
		check_array(target, "mlp target", 1);
		check_array(gradients, "mlp gradients", 1);
		check_array(loss, "mlp loss", 1);
		expect_dtype(*target, REXLIB_HIP_FLOAT32, "mlp target");
		expect_dtype(*gradients, REXLIB_HIP_FLOAT32, "mlp gradients");
		expect_dtype(*loss, REXLIB_HIP_FLOAT32, "mlp loss");
		if (target->extents[0] != rows || gradients->extents[0] != rexlib_hip::mlp_parameter_count() ||
			loss->extents[0] < 1 || !contiguous(*target) || !contiguous(*gradients))
		{
			fail(REXLIB_HIP_INVALID_ARGUMENT, "mlp_backward: target per row, gradients per parameter, one loss");
		}
		launch_mlp_backward(features->data, static_cast<const float *>(target->data), parameters_f16->data,
			static_cast<const float *>(parameters_f32->data), static_cast<float *>(gradients->data),
			static_cast<float *>(loss->data), rows, static_cast<float>(loss_scale), as_stream(stream));
	});
}

int rexlib_hip_mlp_adam(const rexlib_hip_array *parameters_f32, const rexlib_hip_array *parameters_f16,
	const rexlib_hip_array *gradients, const rexlib_hip_array *first_moment, const rexlib_hip_array *second_moment,
	double learning_rate, double beta1, double beta2, double epsilon, int32_t step, double loss_scale, void *stream)
{
	return guarded([&]
	{
		const int64_t p = rexlib_hip::mlp_parameter_count();
		const rexlib_hip_array *f32[] = {parameters_f32, gradients, first_moment, second_moment};
		for (const auto *a : f32)
		{
			check_array(a, "mlp_adam operand", 1);
			expect_dtype(*a, REXLIB_HIP_FLOAT32, "mlp_adam operand");
			if (a->extents[0] != p || !contiguous(*a))
			{
				fail(REXLIB_HIP_INVALID_ARGUMENT, "mlp_adam: one contiguous float32 per parameter");
			}
		}
		check_array(parameters_f16, "mlp_adam float16 parameters", 1);
		expect_dtype(*parameters_f16, REXLIB_HIP_FLOAT16, "mlp_adam float16 parameters");
		if (step < 1)
		{
			fail(REXLIB_HIP_INVALID_ARGUMENT, "mlp_adam: steps count from 1");
		}
		launch_adam(static_cast<float *>(parameters_f32->data), parameters_f16->data,
			static_cast<float *>(gradients->data), static_cast<float *>(first_moment->data),
			static_cast<float *>(second_moment->data), static_cast<float>(learning_rate), static_cast<float>(beta1),
			static_cast<float>(beta2), static_cast<float>(epsilon), step, static_cast<float>(loss_scale),
			as_stream(stream));
	});
}

} // extern "C"
