// TEST INFRASTRUCTURE — drives the reference's own CPU insertion kernels.
//
// This file is compiled together with the reference's sources where they lie
// (see oracle/Makefile) into oracle/_ref/librexlib_ref.so. It contains no
// reference source: it instantiates the reference's templates
// (row_insertion_kernel, the windows, the sample weights, the volume phases)
// and calls its plan helpers (volume_tiling, bound_rows, clip_row,
// fill_axis_phase, slice_frame) over contiguous operands.
//
// The driver restates the two member functions of slice_insertion_functor that
// cannot be compiled here, because the functor's layout plan needs Boost:
//   - settle_slices: src/backends/cpu/builders/em/reconstruction/slice_insertion_functor.inl:158-269
//   - insert_tile:   src/backends/cpu/builders/em/reconstruction/slice_insertion_functor.inl:271-465
//   - run (tiling and claimed tasks): slice_insertion_functor.inl:105-156
// Only the batch addressing differs: the operands are K accumulators, each
// receiving B contiguous slices, instead of a joint layout.
#include <backends/cpu/builders/em/reconstruction/alternating_sign.hpp>
#include <backends/cpu/builders/em/reconstruction/ctf_weight.hpp>
#include <backends/cpu/builders/em/reconstruction/row_insertion_kernel.hpp>
#include <backends/cpu/builders/em/reconstruction/spectrum_weight.hpp>
#include <backends/cpu/builders/em/reconstruction/tabulated_phase.hpp>
#include <backends/cpu/builders/em/reconstruction/unit_weight.hpp>
#include <backends/cpu/builders/em/reconstruction/virtual_row.hpp>
#include <backends/cpu/kernels/axis_phase.hpp>
#include <backends/cpu/kernels/bessel_i0.hpp>
#include <backends/cpu/kernels/box_window.hpp>
#include <backends/cpu/kernels/contrast_transfer_function.hpp>
#include <backends/cpu/kernels/kaiser_bessel_window.hpp>
#include <backends/cpu/kernels/row_interval.hpp>
#include <backends/cpu/kernels/slice_frame.hpp>
#include <backends/cpu/kernels/triangle_window.hpp>
#include <backends/cpu/plans/volume_tiling.hpp>

#include <rexlib/em/ops/parameters/ctf_optics.hpp>
#include <rexlib/em/ops/parameters/gridding_window.hpp>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <complex>
#include <cstdint>
#include <thread>
#include <vector>

using namespace rexlib;
using namespace rexlib::cpu;

namespace
{

struct insertion_args
{
	int variant; // 0 plain, 1 weighted, 2 ctf
	std::int64_t n, big, K, B;
	const void *slices, *transforms, *radial, *extra;
	void *num, *den;
	double voltage, cs, q, pixel_size;
	int threads;
	std::int64_t tile_budget;
	double half_width;
};

template <typename SampleWeight>
struct weight_parameters;

template <>
struct weight_parameters<unit_weight>
{
	static unit_weight::parameters_type make(const insertion_args&)
	{
		return unit_weight::parameters_type();
	}
};

template <>
struct weight_parameters<spectrum_weight>
{
	static spectrum_weight::parameters_type make(const insertion_args&)
	{
		return spectrum_weight::parameters_type();
	}
};

template <>
struct weight_parameters<ctf_weight>
{
	static ctf_weight::parameters_type make(const insertion_args &a)
	{
		return ctf_weight::parameters_type(
			em::ctf_optics(a.voltage, a.cs, a.q), a.pixel_size);
	}
};

struct slice_record
{
	slice_frame frame{{{0, 0, 0}}, {{0, 0, 0}}, 1.0};
	std::ptrdiff_t last_shell = -1;
};

template <typename T, typename SampleWeight, typename Window, typename Phase>
void run(const insertion_args &a, const Window &window, const Phase &phase)
{
	using complex_type = std::complex<T>;
	using state_type = typename SampleWeight::template slice_state<T>;
	const std::size_t n = static_cast<std::size_t>(a.n);
	const std::size_t h = n / 2 + 1;
	const std::size_t big = static_cast<std::size_t>(a.big);
	const std::size_t half = big / 2 + 1;
	const std::size_t count = static_cast<std::size_t>(a.K * a.B);
	const std::size_t table_size = h + n;
	const double scale = static_cast<double>(big) / static_cast<double>(n);
	const double centre = static_cast<double>(n / 2);
	const auto parameters = weight_parameters<SampleWeight>::make(a);

	const auto *slices = static_cast<const complex_type*>(a.slices);
	const auto *transforms = static_cast<const T*>(a.transforms);
	const auto *radials = static_cast<const T*>(a.radial);
	const auto *extras = static_cast<const T*>(a.extra);
	auto *numerator = static_cast<complex_type*>(a.num);
	auto *denominator = static_cast<T*>(a.den);

	// Settle every slice: frame, last shell, shift phases, sample state.
	std::vector<slice_record> records(count);
	std::vector<state_type> states(count);
	std::vector<complex_type> phases(count * table_size);
	const operand_core extra_core = a.variant == 2
		? operand_core(1, {{4, 0, 0}}, {{1, 0, 0}})
		: operand_core(2, {{n, h, 0}}, {{static_cast<std::ptrdiff_t>(h), 1, 0}});
	for (std::size_t s = 0; s < count; ++s)
	{
		const T *t = transforms + s * 12;
		const auto at = [&](std::size_t r, std::size_t c)
		{
			return static_cast<double>(t[r * 4 + c]);
		};
		records[s].frame = slice_frame(
			{{at(0, 0), at(0, 1), at(0, 2)}},
			{{at(1, 0), at(1, 1), at(1, 2)}},
			scale);
		const T *radial = radials + s * h;
		for (std::size_t shell = 0; shell < h; ++shell)
		{
			if (radial[shell] != T(0))
			{
				records[s].last_shell = static_cast<std::ptrdiff_t>(shell);
			}
		}
		fill_axis_phase(&phases[s * table_size], h, n, centre + at(0, 3));
		fill_axis_phase(&phases[s * table_size + h], n, n, centre + at(1, 3));
		const std::size_t extra_size = a.variant == 2 ? 4 : n * h;
		states[s] = a.variant == 0
			? state_type()
			: state_type(parameters, n, extras + s * extra_size, extra_core);
	}

	const std::size_t threads = static_cast<std::size_t>(std::max(1, a.threads));
	const std::size_t row_bytes = half * (sizeof(complex_type) + sizeof(T));
	const volume_tiling tiling(
		big,
		choose_tile_extent(big, row_bytes,
			static_cast<std::size_t>(a.tile_budget), 8 * threads, a.half_width));
	const std::size_t tile_count = tiling.get_tile_count();
	const operand_core volume_core(3, {{big, big, half}},
		{{static_cast<std::ptrdiff_t>(big * half), static_cast<std::ptrdiff_t>(half), 1}});
	const operand_core slice_core(2, {{n, h, 0}}, {{static_cast<std::ptrdiff_t>(h), 1, 0}});
	const std::ptrdiff_t signed_n = static_cast<std::ptrdiff_t>(n);
	const std::ptrdiff_t last_column = static_cast<std::ptrdiff_t>(h) - (n % 2 == 0 ? 2 : 1);
	const std::ptrdiff_t highest_row = n % 2 == 0 ? signed_n / 2 - 1 : signed_n / 2;

	const auto insert_tile = [&](std::size_t kept, const volume_tile &tile)
	{
		row_insertion_kernel<T, Window, Phase, SampleWeight> kernel(
			window, phase, big, volume_core, volume_core);
		kernel.bind(numerator + kept * big * big * half, denominator + kept * big * big * half);
		volume_tile::interval_list_type z_intervals;
		volume_tile::interval_list_type y_intervals;
		const auto z_count = tile.list_z_frequencies(z_intervals);
		const auto y_count = tile.list_y_frequencies(y_intervals);
		for (std::size_t b = 0; b < static_cast<std::size_t>(a.B); ++b)
		{
			const std::size_t s = kept * static_cast<std::size_t>(a.B) + b;
			const auto last_shell = records[s].last_shell;
			if (last_shell < 0)
			{
				continue;
			}
			const auto &frame = records[s].frame;
			const complex_type *slice = slices + s * n * h;
			const T *radial = radials + s * h;
			const complex_type *x_phases = &phases[s * table_size];
			const complex_type *y_phases = x_phases + h;
			const double radius = static_cast<double>(last_shell) + 0.5;
			const auto widest = std::min(last_column, static_cast<std::ptrdiff_t>(radius));
			const auto first_row = std::max(-highest_row, -widest);
			const auto last_row = std::min(highest_row, widest);
			for (int sign = 1; sign >= -1; sign -= 2)
			{
				const bool negated = sign < 0;
				const std::ptrdiff_t first = negated ? 1 : 0;
				const double ss = static_cast<double>(sign);
				const auto &xf = frame.get_x_step();
				const auto &yf = frame.get_y_step();
				const volume_frequency step = {{ss * xf[0], ss * xf[1], ss * xf[2]}};
				const volume_frequency y_step = {{ss * yf[0], ss * yf[1], ss * yf[2]}};
				for (std::size_t zi = 0; zi < z_count; ++zi)
				for (std::size_t yi = 0; yi < y_count; ++yi)
				{
					kernel.set_box(z_intervals[zi], y_intervals[yi]);
					const auto rows = bound_rows(step, y_step, kernel.get_reach(),
						first, widest, first_row, last_row);
					for (auto k_y = rows.get_begin(); k_y < rows.get_end(); ++k_y)
					{
						const double ky = static_cast<double>(k_y);
						const double reach = radius * radius - ky * ky;
						if (reach < 0.0)
						{
							continue;
						}
						const auto last = std::min(widest,
							static_cast<std::ptrdiff_t>(std::sqrt(reach)));
						if (first > last)
						{
							continue;
						}
						const volume_frequency base = {{ky * y_step[0], ky * y_step[1], ky * y_step[2]}};
						const auto interval = clip_row(base, step, kernel.get_reach(), first, last);
						if (interval.is_empty())
						{
							continue;
						}
						const auto stored_row = k_y < 0 ? k_y + signed_n : k_y;
						const virtual_row<T> row(negated, stored_row, ky, base, step,
							slice + stored_row * slice_core.get_stride(0),
							slice_core.get_stride(1), x_phases, y_phases[stored_row],
							radial, 1, last_shell);
						kernel.insert(row, states[s], interval.get_begin(), interval.get_end());
					}
				}
			}
		}
	};

	const std::size_t tasks = static_cast<std::size_t>(a.K) * tile_count;
	std::atomic<std::size_t> next(0);
	const auto worker = [&]()
	{
		for (std::size_t task = next++; task < tasks; task = next++)
		{
			insert_tile(task / tile_count, tiling.get_tile(task % tile_count));
		}
	};
	std::vector<std::thread> pool;
	for (std::size_t i = 1; i < threads; ++i)
	{
		pool.emplace_back(worker);
	}
	worker();
	for (auto &thread : pool)
	{
		thread.join();
	}
}

template <typename T, typename SampleWeight, typename Window>
void run_with_window(const insertion_args &a, const Window &window)
{
	// Even extents use the alternating sign, odd ones the phase table
	// (slice_insertion_functor.inl:75-103).
	if (a.big % 2 == 0)
	{
		run<T, SampleWeight>(a, window, alternating_sign());
	}
	else
	{
		run<T, SampleWeight>(a, window, tabulated_phase<T>(static_cast<std::size_t>(a.big)));
	}
}

template <typename T, typename SampleWeight>
void run_with_profile(const insertion_args &a, int profile, double half_width, double shape)
{
	switch (profile)
	{
	case 0:
		run_with_window<T, SampleWeight>(a, box_window());
		break;
	case 1:
		run_with_window<T, SampleWeight>(a, triangle_window());
		break;
	default:
		run_with_window<T, SampleWeight>(a,
			kaiser_bessel_window<T>(em::gridding_window::kaiser_bessel(half_width, shape)));
		break;
	}
}

template <typename T>
void run_typed(const insertion_args &a, int profile, double half_width, double shape)
{
	switch (a.variant)
	{
	case 0:
		run_with_profile<T, unit_weight>(a, profile, half_width, shape);
		break;
	case 1:
		run_with_profile<T, spectrum_weight>(a, profile, half_width, shape);
		break;
	default:
		run_with_profile<T, ctf_weight>(a, profile, half_width, shape);
		break;
	}
}

} // anonymous namespace

extern "C" {

// Insert K*B contiguous slices into K contiguous accumulators, in place, with
// the reference's own row kernel. dtype 0 is float32, 1 is float64.
int ref_insert(int dtype, int variant, int profile, double half_width, double shape,
	std::int64_t n, std::int64_t big, std::int64_t K, std::int64_t B,
	void *num, void *den, const void *slices, const void *transforms,
	const void *radial, const void *extra,
	double voltage, double cs, double q, double pixel_size,
	int threads, std::int64_t tile_budget)
{
	try
	{
		insertion_args a{variant, n, big, K, B, slices, transforms, radial, extra,
			num, den, voltage, cs, q, pixel_size, threads, tile_budget,
			profile == 0 ? 0.5 : (profile == 1 ? 1.0 : half_width)};
		if (dtype == 0)
		{
			run_typed<float>(a, profile, half_width, shape);
		}
		else
		{
			run_typed<double>(a, profile, half_width, shape);
		}
		return 0;
	}
	catch (...)
	{
		return 1;
	}
}

double ref_bessel_i0(double x)
{
	return bessel_i0(x);
}

double ref_wavelength(double voltage)
{
	return em::ctf_optics(voltage, 2.7, 0.1).get_wavelength();
}

// The float32 CTF of the reference at (k_x, k_y), for images of `extent`.
float ref_ctf_f32(double voltage, double cs, double q, double pixel_size, std::int64_t extent,
	double du, double dv, double angle, double phase, float k_x, float k_y)
{
	const contrast_transfer_function<float> f(em::ctf_optics(voltage, cs, q), pixel_size,
		static_cast<std::size_t>(extent), du, dv, angle, phase);
	return f(k_x, k_y);
}

// The float32 Kaiser-Bessel window of the reference at a distance.
float ref_kaiser_bessel_f32(double half_width, double shape, float distance)
{
	const kaiser_bessel_window<float> w(em::gridding_window::kaiser_bessel(half_width, shape));
	return w.evaluate(distance);
}

// fill_axis_phase of the reference, float32, interleaved (re, im).
void ref_axis_phase_f32(float *out, std::int64_t count, std::int64_t extent, double shift)
{
	fill_axis_phase(reinterpret_cast<std::complex<float>*>(out),
		static_cast<std::size_t>(count), static_cast<std::size_t>(extent), shift);
}

// choose_tile_extent of the reference.
std::int64_t ref_choose_tile_extent(std::int64_t extent, std::int64_t row_bytes,
	std::int64_t budget, std::int64_t task_floor, double half_width)
{
	return static_cast<std::int64_t>(choose_tile_extent(static_cast<std::size_t>(extent),
		static_cast<std::size_t>(row_bytes), static_cast<std::size_t>(budget),
		static_cast<std::size_t>(task_floor), half_width));
}

} // extern "C"
