// Fourier slice insertion on gfx950.
//
// Contract: include/rexlib/em/ops/reconstruction/insert_slices_operation.hpp:31-112
// (and the CTF / weighted variants). The per-sample arithmetic follows the
// reference's CPU row kernel in float32
// (src/backends/cpu/builders/em/reconstruction/row_insertion_kernel.inl:132-361):
//   - the row base k_y * Y is formed in double and rounded, the position is
//     base + k_x * X in float (row_insertion_kernel.inl:157-166);
//   - the shell is round-half-up |k|, exact in integers (row_insertion_kernel.inl:180-220);
//   - v = X(k) * x_phase * y_phase, then times r * (numerator weight), the
//     imaginary part negated for the -k copy (row_insertion_kernel.inl:222-279);
//   - the volume phase is (-1)^(x+y+z) for even N and a product of three table
//     entries for odd N (alternating_sign.inl:10-20, tabulated_phase.inl:12-37).
//
// Two schedules accumulate the same sums:
//   - scatter: one thread per stored slice element, float atomics into the
//     accumulators in HBM (every window, every size);
//   - brick: each workgroup owns a brick of the volume in LDS, gathers the
//     samples of every slice that reach it, and adds the brick to HBM once.
//     This is the GPU form of the reference's tile-owned schedule
//     (slice_insertion_functor.inl:105-156, 271-465).
#include "launch.hpp"

#include <cmath>
#include <mutex>
#include <unordered_map>
#include <vector>

namespace rexlib_hip
{

namespace
{

constexpr double pi_d = 3.14159265358979323846;
constexpr int max_taps = 8;

// ---------------------------------------------------------------------------
// Settle: one block per slice.
// ---------------------------------------------------------------------------

__global__ void settle_kernel(insertion_params p)
{
	const int n = p.n;
	const int h = p.h;
	for (int64_t s = blockIdx.x; s < p.slice_count; s += gridDim.x)
	{
		const int64_t kept = s / p.map.reduced_count;
		const int64_t red = s % p.map.reduced_count;
		const auto offset = [&](int slot)
		{
			return kept_offset(p.map, slot, kept) + reduced_offset(p.map, slot, red);
		};
		const float *t = p.transforms + offset(slot_transforms);
		const auto at = [&](int r, int c)
		{
			return static_cast<double>(t[r * p.transform_s0 + c * p.transform_s1]);
		};
		const double scale = static_cast<double>(p.big) / static_cast<double>(n);
		const double centre = static_cast<double>(n / 2);
		const int64_t radial_offset = offset(slot_radial);

		if (threadIdx.x == 0)
		{
			slice_entry e;
			bool finite = true;
			for (int a = 0; a < 3; ++a)
			{
				const double xs = at(0, a) * scale;
				const double ys = at(1, a) * scale;
				finite = finite && isfinite(xs) && isfinite(ys);
				e.x_step[a] = static_cast<float>(xs);
				e.y_step[a] = ys;
			}
			int last = -1;
			const float *radial = p.radial + radial_offset;
			for (int shell = 0; shell < h; ++shell)
			{
				if (radial[shell * p.radial_s0] != 0.0f)
				{
					last = shell;
				}
			}
			// A non-finite frame bounds no rows (src/backends/cpu/kernels/row_interval.cpp).
			e.last_shell = finite ? last : -1;
			e.ctf_radial = e.ctf_axial = e.ctf_diagonal = e.ctf_aberration = e.ctf_phase = 0.0f;
			e.pad = 0;
			e.extra_offset = p.extra != nullptr ? offset(slot_extra) : 0;
			if (p.variant == variant_ctf)
			{
				// contrast_transfer_function.inl:25-70: coefficients in double,
				// rounded to float.
				const float *c = p.extra + e.extra_offset;
				const double du = c[0];
				const double dv = c[p.extra_s0];
				const double angle = c[2 * p.extra_s0];
				const double phase = c[3 * p.extra_s0];
				const double half_difference = 0.5 * (du - dv);
				e.ctf_radial = static_cast<float>(p.ctf_defocus_scale * 0.5 * (du + dv));
				e.ctf_axial = static_cast<float>(p.ctf_defocus_scale * half_difference * cos(2.0 * angle));
				e.ctf_diagonal = static_cast<float>(
					2.0 * p.ctf_defocus_scale * half_difference * sin(2.0 * angle));
				e.ctf_aberration = static_cast<float>(p.ctf_aberration);
				e.ctf_phase = static_cast<float>(phase + p.ctf_phase0);
			}
			e.slice_offset = offset(slot_slices);
			e.radial_offset = radial_offset;
			e.numerator_offset = kept_offset(p.map, slot_numerator_out, kept);
			e.denominator_offset = kept_offset(p.map, slot_denominator_out, kept);
			e.kept = kept;
			p.entries[s] = e;
		}

		// Shift phases exp(2 pi i k (c + t) / n) (axis_phase.inl:14-36), in
		// double, rounded to float.
		float2 *table = p.tables + s * static_cast<int64_t>(h + n);
		const double x_scale = 2.0 * pi_d * (centre + at(0, 3)) / static_cast<double>(n);
		const double y_scale = 2.0 * pi_d * (centre + at(1, 3)) / static_cast<double>(n);
		for (int j = threadIdx.x; j < h + n; j += blockDim.x)
		{
			double angle;
			if (j < h)
			{
				angle = x_scale * static_cast<double>(j);
			}
			else
			{
				const int r = j - h;
				const int f = r <= n / 2 ? r : r - n;
				angle = y_scale * static_cast<double>(f);
			}
			double sn, cs;
			sincos(angle, &sn, &cs);
			table[j] = make_float2(static_cast<float>(cs), static_cast<float>(sn));
		}
	}
}

// Kaiser-Bessel table: entry i holds the window at i / count of the
// half-width, then two zeros (kaiser_bessel_window.inl:18-50).
__device__ double bessel_i0_series(double x)
{
	const double quarter_square = 0.25 * x * x;
	double sum = 1.0;
	double term = 1.0;
	for (double k = 1.0; k < 500.0; k += 1.0)
	{
		term *= quarter_square / (k * k);
		sum += term;
		if (term <= 2.220446049250313e-16 * sum)
		{
			break;
		}
	}
	return sum;
}

__global__ void kb_table_kernel(float *table, int count, double shape)
{
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= count + 3)
	{
		return;
	}
	if (i > count)
	{
		table[i] = 0.0f;
		return;
	}
	const double ratio = static_cast<double>(i) / static_cast<double>(count);
	table[i] = static_cast<float>(bessel_i0_series(shape * sqrt(1.0 - ratio * ratio)) /
		bessel_i0_series(shape));
}

// exp(-2 pi i f(j) C / N) for the stored index j of an odd N (tabulated_phase.inl:12-22).
__global__ void volume_phase_kernel(float2 *table, int big)
{
	const int j = blockIdx.x * blockDim.x + threadIdx.x;
	if (j >= big)
	{
		return;
	}
	const int half = big / 2;
	const int f = j <= half ? j : j - big;
	const double angle = 2.0 * pi_d * (-static_cast<double>(half)) * static_cast<double>(f) /
		static_cast<double>(big);
	double sn, cs;
	sincos(angle, &sn, &cs);
	table[j] = make_float2(static_cast<float>(cs), static_cast<float>(sn));
}

// Out-of-place insertions start from a copy of the accumulators.
__global__ void copy_accumulators_kernel(insertion_params p, const float2 *num_in, int64_t ni0, int64_t ni1,
	int64_t ni2, const float *den_in, int64_t di0, int64_t di1, int64_t di2)
{
	const int64_t volume = static_cast<int64_t>(p.big) * p.big * p.half;
	const int64_t total = volume * p.map.kept_count;
	for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < total;
		i += static_cast<int64_t>(gridDim.x) * blockDim.x)
	{
		const int64_t kept = i / volume;
		int64_t v = i % volume;
		const int64_t x = v % p.half;
		v /= p.half;
		const int64_t y = v % p.big;
		const int64_t z = v / p.big;
		const float2 a = num_in[kept_offset(p.map, slot_numerator, kept) + z * ni0 + y * ni1 + x * ni2];
		const float b = den_in[kept_offset(p.map, slot_denominator, kept) + z * di0 + y * di1 + x * di2];
		p.numerator[kept_offset(p.map, slot_numerator_out, kept) + z * p.num_s0 + y * p.num_s1 + x * p.num_s2] = a;
		p.denominator[kept_offset(p.map, slot_denominator_out, kept) + z * p.den_s0 + y * p.den_s1 + x * p.den_s2] = b;
	}
}

// ---------------------------------------------------------------------------
// Per-sample arithmetic shared by the schedules
// ---------------------------------------------------------------------------

// Round-half-up |k| of an integer frequency, exactly: the largest s with
// 4 s^2 - 4 s + 1 <= 4 |k|^2 (row_insertion_kernel.inl:180-220).
__device__ inline int shell_of(int kx, int ky)
{
	const int q4 = 4 * (kx * kx + ky * ky);
	int s = static_cast<int>(sqrtf(static_cast<float>(kx * kx + ky * ky)) + 0.5f);
	while (4 * s * s + 4 * s + 1 <= q4)
	{
		++s;
	}
	while (s > 0 && 4 * s * s - 4 * s + 1 > q4)
	{
		--s;
	}
	return s;
}

// The window over one axis: the first voxel and the tap weights.
template <int PROFILE>
struct window_taps;

// Box: round halves away from 0 (box_window.inl:16-34).
template <>
struct window_taps<REXLIB_HIP_BOX>
{
	static constexpr int taps = 1;
	__device__ static inline int place(const insertion_params &, float c, float *w)
	{
		w[0] = 1.0f;
		return static_cast<int>(c + (c < 0.0f ? -0.5f : 0.5f));
	}
};

// Triangle: floor and the fraction (triangle_window.inl:16-38).
template <>
struct window_taps<REXLIB_HIP_TRIANGLE>
{
	static constexpr int taps = 2;
	__device__ static inline int place(const insertion_params &, float c, float *w)
	{
		const int t = static_cast<int>(c);
		const int below = t - (c < static_cast<float>(t) ? 1 : 0);
		const float f = c - static_cast<float>(below);
		w[0] = 1.0f - f;
		w[1] = f;
		return below;
	}
};

// Kaiser-Bessel: table with linear interpolation (kaiser_bessel_window.inl:65-136).
template <>
struct window_taps<REXLIB_HIP_KAISER_BESSEL>
{
	static constexpr int taps = max_taps;
	__device__ static inline int place(const insertion_params &p, float c, float *w)
	{
		const float low = c - p.kb_half_width;
		const int t = static_cast<int>(low);
		const int first = t - (low < static_cast<float>(t) ? 1 : 0) + 1;
		for (int tap = 0; tap < max_taps; ++tap)
		{
			if (tap < p.taps)
			{
				const float d = fabsf(static_cast<float>(first + tap) - c);
				const float position = d * p.kb_scale;
				const int step = static_cast<int>(position);
				const int index = step < p.kb_count ? step : p.kb_count + 1;
				const float fraction = position - static_cast<float>(index);
				const float below = p.kb_table[index];
				w[tap] = below + fraction * (p.kb_table[index + 1] - below);
			}
			else
			{
				w[tap] = 0.0f;
			}
		}
		return first;
	}
};

// The sample of stored element (row, column) of slice s: its value v * r *
// (numerator weight) before the sign, its denominator weight, and false when
// it is not inserted.
template <int VARIANT>
__device__ inline bool prepare_sample(const insertion_params &p, const slice_entry &e, int64_t s, int row,
	int column, float2 &value, float &count)
{
	const int n = p.n;
	if ((n % 2 == 0) && (row == n / 2 || column == n / 2))
	{
		return false;
	}
	const int ky = row <= n / 2 ? row : row - n;
	const int kx = column;
	const int shell = shell_of(kx, ky);
	if (shell > e.last_shell)
	{
		return false;
	}
	const float r = p.radial[e.radial_offset + shell * p.radial_s0];
	if (r == 0.0f)
	{
		return false;
	}
	const float2 x = p.slices[e.slice_offset + row * p.slice_s0 + column * p.slice_s1];
	const float2 *table = p.tables + s * static_cast<int64_t>(p.h + n);
	const float2 shifted = cmul(x, table[column]);
	const float2 v = cmul(shifted, table[p.h + row]);
	float numerator_weight = 1.0f;
	float denominator_weight = 1.0f;
	if (VARIANT == variant_weighted)
	{
		denominator_weight = p.extra[e.extra_offset + row * p.extra_s0 + column * p.extra_s1];
	}
	else if (VARIANT == variant_ctf)
	{
		// contrast_transfer_function.inl:73-87 and ctf_weight.inl:72-103.
		const float fx = static_cast<float>(kx);
		const float fy = static_cast<float>(ky);
		const float xx = fx * fx;
		const float yy = fy * fy;
		const float norm = xx + yy;
		const float chi = e.ctf_radial * norm + e.ctf_axial * (xx - yy) + e.ctf_diagonal * (fx * fy) -
			e.ctf_aberration * (norm * norm) + e.ctf_phase;
		numerator_weight = sinf(chi);
		denominator_weight = numerator_weight * numerator_weight;
	}
	const float scale = r * numerator_weight;
	value = make_float2(v.x * scale, v.y * scale);
	count = r * denominator_weight;
	return true;
}

// Volume position of sample k_x of row k_y with the sign applied.
__device__ inline void sample_position(const slice_entry &e, int kx, int ky, float sign, float *c)
{
	const float fx = static_cast<float>(kx);
	for (int a = 0; a < 3; ++a)
	{
		const float base = static_cast<float>(static_cast<double>(ky) * (static_cast<double>(sign) * e.y_step[a]));
		const float step = sign * e.x_step[a];
		c[a] = base + fx * step;
	}
}

// Volume phase of the stored voxel (x, sy, sz), applied to a contribution.
template <bool ODD>
__device__ inline float2 volume_phase(const insertion_params &p, float2 value, int x, int sy, int sz)
{
	if (ODD)
	{
		return cmul(cmul(cmul(value, p.volume_phase[x]), p.volume_phase[sy]), p.volume_phase[sz]);
	}
	return ((x + sy + sz) & 1) ? make_float2(-value.x, -value.y) : value;
}

// ---------------------------------------------------------------------------
// Scatter schedule: a thread per stored slice element, atomics in HBM.
// ---------------------------------------------------------------------------

template <int PROFILE, int VARIANT, bool ODD>
__global__ __launch_bounds__(256) void scatter_kernel(insertion_params p)
{
	const int64_t elements = static_cast<int64_t>(p.n) * p.h;
	const int64_t idx = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
	if (idx >= elements)
	{
		return;
	}
	const int row = static_cast<int>(idx / p.h);
	const int column = static_cast<int>(idx % p.h);
	const int big = p.big;
	const int half_extent = big / 2;
	for (int64_t s = blockIdx.y; s < p.slice_count; s += gridDim.y)
	{
		const slice_entry &e = p.entries[s];
		if (e.last_shell < 0)
		{
			continue;
		}
		float2 value;
		float count;
		if (!prepare_sample<VARIANT>(p, e, s, row, column, value, count))
		{
			continue;
		}
		const int ky = row <= p.n / 2 ? row : row - p.n;
		float2 *num = p.numerator + e.numerator_offset;
		float *den = p.denominator + e.denominator_offset;
		for (int sign_index = 0; sign_index < 2; ++sign_index)
		{
			if (sign_index == 1 && column == 0)
			{
				break;
			}
			const float sign = sign_index == 0 ? 1.0f : -1.0f;
			const float2 signed_value = sign_index == 0 ? value : cconj(value);
			float c[3];
			sample_position(e, column, ky, sign, c);
			float wx[max_taps], wy[max_taps], wz[max_taps];
			const int x0 = window_taps<PROFILE>::place(p, c[0], wx);
			const int y0 = window_taps<PROFILE>::place(p, c[1], wy);
			const int z0 = window_taps<PROFILE>::place(p, c[2], wz);
			const int taps = PROFILE == REXLIB_HIP_KAISER_BESSEL ? p.taps : window_taps<PROFILE>::taps;
			for (int tz = 0; tz < taps; ++tz)
			{
				const int z = z0 + tz;
				if (z < -half_extent || z > half_extent)
				{
					continue;
				}
				const int sz = z < 0 ? z + big : z;
				for (int ty = 0; ty < taps; ++ty)
				{
					const int y = y0 + ty;
					if (y < -half_extent || y > half_extent)
					{
						continue;
					}
					const int sy = y < 0 ? y + big : y;
					const float wzy = wz[tz] * wy[ty];
					for (int tx = 0; tx < taps; ++tx)
					{
						const int x = x0 + tx;
						if (x < 0 || x > half_extent)
						{
							continue;
						}
						const float w = wzy * wx[tx];
						const float2 contribution = volume_phase<ODD>(p,
							make_float2(signed_value.x * w, signed_value.y * w), x, sy, sz);
						float2 *target = num + sz * p.num_s0 + sy * p.num_s1 + x * p.num_s2;
						atomicAdd(&target->x, contribution.x);
						atomicAdd(&target->y, contribution.y);
						atomicAdd(den + sz * p.den_s0 + sy * p.den_s1 + x * p.den_s2, count * w);
					}
				}
			}
		}
	}
}

template <int PROFILE, int VARIANT>
void launch_scatter(const insertion_params &p, hipStream_t stream)
{
	const int64_t elements = static_cast<int64_t>(p.n) * p.h;
	const dim3 block(256);
	const dim3 grid(static_cast<unsigned>((elements + 255) / 256),
		static_cast<unsigned>(std::min<int64_t>(p.slice_count, 65535)));
	if (p.big % 2 == 1)
	{
		scatter_kernel<PROFILE, VARIANT, true><<<grid, block, 0, stream>>>(p);
	}
	else
	{
		scatter_kernel<PROFILE, VARIANT, false><<<grid, block, 0, stream>>>(p);
	}
}

template <int VARIANT>
void dispatch_scatter(const insertion_params &p, hipStream_t stream)
{
	switch (p.profile)
	{
	case REXLIB_HIP_BOX:
		launch_scatter<REXLIB_HIP_BOX, VARIANT>(p, stream);
		break;
	case REXLIB_HIP_TRIANGLE:
		launch_scatter<REXLIB_HIP_TRIANGLE, VARIANT>(p, stream);
		break;
	default:
		launch_scatter<REXLIB_HIP_KAISER_BESSEL, VARIANT>(p, stream);
		break;
	}
}

struct kb_key
{
	int device;
	int count;
	double shape;
	bool operator==(const kb_key &o) const
	{
		return device == o.device && count == o.count && shape == o.shape;
	}
};

struct kb_hash
{
	std::size_t operator()(const kb_key &k) const
	{
		return std::hash<double>()(k.shape) ^ (static_cast<std::size_t>(k.count) << 8) ^
			static_cast<std::size_t>(k.device);
	}
};

} // anonymous namespace

// Host entry of every insertion: fills the workspace and launches.
void run_insertion(insertion_params p, int schedule, const float2 *num_in, const int64_t *num_in_strides,
	const float *den_in, const int64_t *den_in_strides, bool in_place, double shape, hipStream_t stream)
{
	const int64_t slices = p.slice_count;
	const int n = p.n;
	const int h = p.h;

	// Workspace: entries, phase tables, the window table, the volume phases.
	const std::size_t entry_bytes = align_up(sizeof(slice_entry) * std::max<int64_t>(slices, 1), 256);
	const std::size_t table_bytes = align_up(sizeof(float2) * (h + n) * std::max<int64_t>(slices, 1), 256);
	const std::size_t kb_bytes = align_up(sizeof(float) * (p.kb_count + 3), 256);
	const std::size_t phase_bytes = align_up(sizeof(float2) * p.big, 256);
	char *base = static_cast<char *>(
		workspace_area_bytes(area_insertion, entry_bytes + table_bytes + kb_bytes + phase_bytes, stream));
	p.entries = reinterpret_cast<slice_entry *>(base);
	p.tables = reinterpret_cast<float2 *>(base + entry_bytes);
	float *kb = reinterpret_cast<float *>(base + entry_bytes + table_bytes);
	float2 *vphase = reinterpret_cast<float2 *>(base + entry_bytes + table_bytes + kb_bytes);
	p.kb_table = kb;
	p.volume_phase = vphase;

	if (!in_place)
	{
		const int64_t total = static_cast<int64_t>(p.big) * p.big * p.half * p.map.kept_count;
		if (total > 0)
		{
			const unsigned blocks = static_cast<unsigned>(std::min<int64_t>((total + 255) / 256, 1 << 20));
			copy_accumulators_kernel<<<blocks, 256, 0, stream>>>(p, num_in, num_in_strides[0], num_in_strides[1],
				num_in_strides[2], den_in, den_in_strides[0], den_in_strides[1], den_in_strides[2]);
			REXLIB_HIP_CHECK(hipGetLastError());
		}
	}
	if (slices == 0)
	{
		return;
	}

	if (p.profile == REXLIB_HIP_KAISER_BESSEL)
	{
		kb_table_kernel<<<(p.kb_count + 3 + 255) / 256, 256, 0, stream>>>(kb, p.kb_count, shape);
		REXLIB_HIP_CHECK(hipGetLastError());
	}
	if (p.big % 2 == 1)
	{
		volume_phase_kernel<<<(p.big + 255) / 256, 256, 0, stream>>>(vphase, p.big);
		REXLIB_HIP_CHECK(hipGetLastError());
	}
	settle_kernel<<<static_cast<unsigned>(std::min<int64_t>(slices, 1 << 20)), 128, 0, stream>>>(p);
	REXLIB_HIP_CHECK(hipGetLastError());

	(void)schedule;
	switch (p.variant)
	{
	case variant_plain:
		dispatch_scatter<variant_plain>(p, stream);
		break;
	case variant_weighted:
		dispatch_scatter<variant_weighted>(p, stream);
		break;
	default:
		dispatch_scatter<variant_ctf>(p, stream);
		break;
	}
	REXLIB_HIP_CHECK(hipGetLastError());
}

} // namespace rexlib_hip
