// Per-sample arithmetic of the insertion and the wavefront helpers its
// kernels share (the scatter, gather and binned schedules and the sample
// producers).
#pragma once

#include "insertion.hpp"

namespace rexlib_hip
{

constexpr int max_taps = 8;

// ---------------------------------------------------------------------------
// Per-sample arithmetic shared by the schedules
// ---------------------------------------------------------------------------

// Round-half-up |k| of an integer frequency, exactly: the largest s with
// 4 s^2 - 4 s + 1 <= 4 |k|^2 (row_insertion_kernel.inl:180-220). The float
// estimate floor(sqrt(|k|^2) + 1/2) is within one of it for every |k|^2
// below 2^31, so one step each way, by selects, makes it exact.
__device__ inline int shell_of(int kx, int ky)
{
	const int q = kx * kx + ky * ky;
	const int q4 = 4 * q;
	int s = static_cast<int>(sqrtf(static_cast<float>(q)) + 0.5f);
	s += (4 * s * s + 4 * s + 1 <= q4) ? 1 : 0;
	s -= (s > 0 && 4 * s * s - 4 * s + 1 > q4) ? 1 : 0;
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
	__device__ static inline int first(const insertion_params &, float c)
	{
		return static_cast<int>(c + (c < 0.0f ? -0.5f : 0.5f));
	}
	__device__ static inline int place(const insertion_params &p, float c, float *w)
	{
		w[0] = 1.0f;
		return first(p, c);
	}
};

// Triangle: floor and the fraction (triangle_window.inl:16-38).
template <>
struct window_taps<REXLIB_HIP_TRIANGLE>
{
	static constexpr int taps = 2;
	__device__ static inline int first(const insertion_params &, float c)
	{
		const int t = static_cast<int>(c);
		return t - (c < static_cast<float>(t) ? 1 : 0);
	}
	__device__ static inline int place(const insertion_params &p, float c, float *w)
	{
		const int below = first(p, c);
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
	__device__ static inline int first(const insertion_params &p, float c)
	{
		const float low = c - p.kb_half_width;
		const int t = static_cast<int>(low);
		return t - (low < static_cast<float>(t) ? 1 : 0) + 1;
	}
	// Weights of the first `count` taps (at most p.taps are nonzero).
	template <int COUNT = max_taps>
	__device__ static inline int place(const insertion_params &p, float c, float *w)
	{
		const int first = window_taps::first(p, c);
		for (int tap = 0; tap < COUNT; ++tap)
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
// sin x for the CTF phase x (radians, |x| up to about 1e5). Cody-Waite
// reduction: x = k pi/2 + r with the product k pi/2 taken in double, so that
// r is exact to float precision; then the Taylor polynomials of sin (through
// r^9) and cos (through r^8) on |r| <= pi/4, whose truncation errors are
// below 2e-9 and 2.5e-8, selected and signed by the quadrant k mod 4.
// (The library sinf takes its multi-word reduction for phases of a few
// hundred radians, several times the cost of the whole sample.)
__host__ __device__ inline float phase_sin(float x)
{
	const float k = rintf(x * 0.636619772367581343f); // 2 / pi
	const float r = static_cast<float>(fma(-static_cast<double>(k), 1.5707963267948966192, static_cast<double>(x)));
	const float r2 = r * r;
	const float s = r + r * r2 * (-1.0f / 6.0f + r2 * (1.0f / 120.0f + r2 * (-1.0f / 5040.0f + r2 * (1.0f / 362880.0f))));
	const float c = 1.0f + r2 * (-0.5f + r2 * (1.0f / 24.0f + r2 * (-1.0f / 720.0f + r2 * (1.0f / 40320.0f))));
	const int q = static_cast<int>(k) & 3;
	const float v = (q & 1) ? c : s;
	return (q & 2) ? -v : v;
}

// The element's value comes from `load`, called only for inserted samples.
template <int VARIANT, typename Load>
__device__ inline bool prepare_value(const insertion_params &p, const slice_entry &e, int64_t s, int row,
	int column, Load load, float2 &value, float &count)
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
	const float2 x = load();
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
		numerator_weight = phase_sin(chi);
		denominator_weight = numerator_weight * numerator_weight;
	}
	const float scale = r * numerator_weight;
	value = make_float2(v.x * scale, v.y * scale);
	count = r * denominator_weight;
	return true;
}

template <int VARIANT>
__device__ inline bool prepare_sample(const insertion_params &p, const slice_entry &e, int64_t s, int row,
	int column, float2 &value, float &count)
{
	return prepare_value<VARIANT>(p, e, s, row, column,
		[&] { return p.slices[e.slice_offset + row * p.slice_s0 + column * p.slice_s1]; }, value, count);
}

// Volume position of sample k_x of row k_y with the sign applied.
__device__ inline void sample_position(const slice_entry &e, int kx, int ky, float sign, float *c)
{
#pragma clang fp contract(off)
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


// A prepared sample of a slice: v r w (the value with its radial and
// numerator weights) and r w_den, 12 bytes; zero when not inserted.
struct prepared_sample
{
	float re;
	float im;
	float den;
};

// Inclusive prefix sum over the 64 lanes by DPP row shifts (within rows of
// 16) and row broadcasts (across them); lanes without a source add 0.
__device__ __forceinline__ int wave_inclusive_scan(int v)
{
	v += __builtin_amdgcn_update_dpp(0, v, 0x111, 0xf, 0xf, true); // row_shr:1
	v += __builtin_amdgcn_update_dpp(0, v, 0x112, 0xf, 0xf, true); // row_shr:2
	v += __builtin_amdgcn_update_dpp(0, v, 0x114, 0xf, 0xf, true); // row_shr:4
	v += __builtin_amdgcn_update_dpp(0, v, 0x118, 0xf, 0xf, true); // row_shr:8
	v += __builtin_amdgcn_update_dpp(0, v, 0x142, 0xa, 0xf, true); // row_bcast:15
	v += __builtin_amdgcn_update_dpp(0, v, 0x143, 0xc, 0xf, true); // row_bcast:31
	return v;
}

// Exclusive starts and wave totals of two per-lane counts (each at least 0).
// Two 16-bit counts share an int while every count is below 1024 (64 lanes
// then sum below 2^16), else each takes a scan of its own. Called with every
// lane of the wavefront active.
__device__ __forceinline__ void scan_two(int a, int b, int &start_a, int &start_b, int &total_a, int &total_b)
{
	if (__ballot(a > 1023 || b > 1023) == 0)
	{
		const int v = a | (b << 16);
		const int inclusive = wave_inclusive_scan(v);
		const int exclusive = inclusive - v;
		const int total = __builtin_amdgcn_readlane(inclusive, 63);
		start_a = exclusive & 0xffff;
		start_b = exclusive >> 16;
		total_a = total & 0xffff;
		total_b = total >> 16;
	}
	else
	{
		const int inclusive_a = wave_inclusive_scan(a);
		const int inclusive_b = wave_inclusive_scan(b);
		start_a = inclusive_a - a;
		start_b = inclusive_b - b;
		total_a = __builtin_amdgcn_readlane(inclusive_a, 63);
		total_b = __builtin_amdgcn_readlane(inclusive_b, 63);
	}
}

// Prefix maximum over the 64 lanes (lane i gets the maximum of lanes 0..i),
// by DPP row shifts within rows of 16 and row broadcasts across them.
// Lanes without a source read 0 (bound_ctrl), which leaves the result
// unchanged wherever lane 0 holds a value of at least 0.
__device__ __forceinline__ int wave_prefix_max(int v)
{
	v = max(v, __builtin_amdgcn_update_dpp(0, v, 0x111, 0xf, 0xf, true)); // row_shr:1
	v = max(v, __builtin_amdgcn_update_dpp(0, v, 0x112, 0xf, 0xf, true)); // row_shr:2
	v = max(v, __builtin_amdgcn_update_dpp(0, v, 0x114, 0xf, 0xf, true)); // row_shr:4
	v = max(v, __builtin_amdgcn_update_dpp(0, v, 0x118, 0xf, 0xf, true)); // row_shr:8
	v = max(v, __builtin_amdgcn_update_dpp(0, v, 0x142, 0xa, 0xf, true)); // row_bcast:15
	v = max(v, __builtin_amdgcn_update_dpp(0, v, 0x143, 0xc, 0xf, true)); // row_bcast:31
	return v;
}

// A float held by lane `lane` (uniform), as a scalar.
__device__ inline float lane_value(float v, int lane)
{
	return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}

// Lane i + 1's value by a DPP wavefront shift (wave_shl:1); lane 63 gets 0
// (bound_ctrl). Called with every lane of the wavefront active.
__device__ __forceinline__ int next_lane(int v)
{
	return __builtin_amdgcn_update_dpp(0, v, 0x130, 0xf, 0xf, true);
}

// Lane i - 1's value (wave_shr:1); lane 0 gets 0.
__device__ __forceinline__ int previous_lane(int v)
{
	return __builtin_amdgcn_update_dpp(0, v, 0x138, 0xf, 0xf, true);
}

__device__ __forceinline__ float next_lane(float v)
{
	return __int_as_float(next_lane(__float_as_int(v)));
}

// Keeps the compiler from moving LDS accesses across it: the taps of other
// lanes may alias, which it cannot see.
__device__ inline void lds_order()
{
	__asm__ __volatile__("" ::: "memory");
}

// prepare_value without branches, for prepare_kernel and the fused column
// passes: every load is in bounds whatever the sample (the radial index is
// clamped to the last shell, the phase tables cover every row and column),
// and the sample is zeroed afterwards where prepare_value rejects it, so the
// loads of several elements can be in flight together. The arithmetic is
// prepare_value's, in the same order. `radial` is the slice's radial
// weights, `y_table` its row phases.
// Whether stored element (row, column) of a slice is inserted at all: not on
// a Nyquist row or column of an even n, within the last shell, and with a
// nonzero radial weight (row_insertion_kernel.inl:180-220). The same test as
// prepare_unbranched's, from the geometry alone.
__device__ inline bool sample_inserted(const insertion_params &p, const slice_entry &e, const float *radial, int row,
	int column)
{
	const int n = p.n;
	const int ky = row <= n / 2 ? row : row - n;
	const int shell = shell_of(column, ky);
	const bool nyquist = (n % 2 == 0) && (row == n / 2 || column == n / 2);
	const float r = radial[min(shell, max(e.last_shell, 0)) * p.radial_s0];
	return !nyquist && shell <= e.last_shell && r != 0.0f;
}

template <int VARIANT>
__device__ inline prepared_sample prepare_unbranched(const insertion_params &p, const slice_entry &e,
	const float *radial, const float2 *y_table, float2 x_phase, int row, int column, float2 x, bool &inserted_out)
{
	const int n = p.n;
	const int ky = row <= n / 2 ? row : row - n;
	const int shell = shell_of(column, ky);
	const bool nyquist = (n % 2 == 0) && (row == n / 2 || column == n / 2);
	const float r = radial[min(shell, max(e.last_shell, 0)) * p.radial_s0];
	const float2 y_phase = y_table[row];
	float numerator_weight = 1.0f;
	float denominator_weight = 1.0f;
	if (VARIANT == variant_weighted)
	{
		denominator_weight = p.extra[e.extra_offset + row * p.extra_s0 + column * p.extra_s1];
	}
	else if (VARIANT == variant_ctf)
	{
		const float fx = static_cast<float>(column);
		const float fy = static_cast<float>(ky);
		const float xx = fx * fx;
		const float yy = fy * fy;
		const float norm = xx + yy;
		const float chi = e.ctf_radial * norm + e.ctf_axial * (xx - yy) + e.ctf_diagonal * (fx * fy) -
			e.ctf_aberration * (norm * norm) + e.ctf_phase;
		numerator_weight = phase_sin(chi);
		denominator_weight = numerator_weight * numerator_weight;
	}
	const float2 v = cmul(cmul(x, x_phase), y_phase);
	const float scale = r * numerator_weight;
	const bool inserted = !nyquist && shell <= e.last_shell && r != 0.0f;
	inserted_out = inserted;
	return inserted ? prepared_sample{v.x * scale, v.y * scale, r * denominator_weight} :
		prepared_sample{0.0f, 0.0f, 0.0f};
}

template <int VARIANT>
__device__ inline prepared_sample prepare_unbranched(const insertion_params &p, const slice_entry &e,
	const float *radial, const float2 *y_table, float2 x_phase, int row, int column, float2 x)
{
	bool inserted;
	return prepare_unbranched<VARIANT>(p, e, radial, y_table, x_phase, row, column, x, inserted);
}

} // namespace rexlib_hip
