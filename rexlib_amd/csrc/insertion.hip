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
#include "fft_device.hpp"
#include "launch.hpp"
#include "sample.hpp"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdlib>
#include <tuple>
#include <map>
#include <mutex>
#include <unordered_map>
#include <vector>

namespace rexlib_hip
{

namespace
{

constexpr double pi_d = 3.14159265358979323846;

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
			// Two samples of one parity class differ by 2 a X + 2 b Y; when the
			// smallest eigenvalue of the Gram matrix of (X, Y) exceeds 3/4, that
			// is at least one voxel along some axis, so no two of them fall in
			// the same voxel cell.
			{
				double xx = 0.0, xy = 0.0, yy = 0.0;
				for (int a = 0; a < 3; ++a)
				{
					xx += static_cast<double>(e.x_step[a]) * e.x_step[a];
					xy += static_cast<double>(e.x_step[a]) * e.y_step[a];
					yy += e.y_step[a] * e.y_step[a];
				}
				const double trace = xx + yy;
				const double gap = sqrt(fmax(0.0, (xx - yy) * (xx - yy) + 4.0 * xy * xy));
				const double smallest = 0.5 * (trace - gap);
				e.lattice_free = finite && smallest > 0.76 ? 1 : 0;
			}
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
// Scatter schedule: a thread per stored slice element, atomics in HBM.
// ---------------------------------------------------------------------------

template <int PROFILE, int VARIANT, bool ODD>
__global__ __launch_bounds__(256) void scatter_kernel(insertion_params p)
{
	// A slice holds n (n/2 + 1) < 2^31 elements: 32-bit indices (a 64-bit
	// division per element would cost more than the sample itself).
	const int elements32 = p.n * p.h;
	const int64_t elements = elements32;
	const int idx = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
	if (idx >= elements32)
	{
		return;
	}
	const int row = static_cast<int>(static_cast<unsigned>(idx) / static_cast<unsigned>(p.h));
	const int column = idx - row * p.h;
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


// ---------------------------------------------------------------------------
// Brick schedule: one wavefront owns a brick of the volume in LDS.
//
// First, a prepare kernel computes every stored element of every slice once:
// v * r * (numerator weight) and r * (denominator weight), or zeros when the
// element is not inserted (row_insertion_kernel.inl:180-279).
//
// Bricks tile the first-tap cells of the volume (signed x in [0, N/2], y and
// z in [-N/2, N/2], each range extended downwards by taps - 1 so that every
// sample reaching the volume has its first tap in some brick). A sample copy
// belongs to the brick holding its first tap, so each copy is added by one
// brick; the brick's LDS holds its cells and the taps - 1 voxels above them.
//
// The wavefront tests 64 slices at a time for planes that pass near the
// brick; for each such slice it bounds the rows that can hold its samples and
// clips each row to them -- bound_rows and clip_row of
// src/backends/cpu/kernels/row_interval.cpp, restated per brick as the
// reference restates them per tile (slice_insertion_functor.inl:391-440).
// The rows of all hit slices are then packed densely over the lanes (64
// rows at a time, each with its base position, step and first sample in a
// small LDS table), and their samples over the lanes (64 at a time), so a
// lane carries a sample whatever slice it comes from.
//
// The samples are added by plain LDS read-modify-write, not LDS float
// atomics (which retire at about a third of a lane per clock per CU on
// gfx950, against eight for a read-modify-write). Per round every lane
// holding a sample claims its first-tap cell (a byte per cell: it writes its
// lane, then reads it back); the lanes that read their own lane add their
// taps one after another, the others keep their sample for the next round,
// whose new samples go to the free lanes only. No two lanes of one tap
// instruction then touch one voxel (a tap is a constant offset from the
// cell). LDS runs a wavefront's accesses in program order and lds_order
// keeps the compiler from moving them; the wavefront is the brick's only
// writer, so there is no barrier anywhere.
//
// The volume phase exp(-2 pi i g C / N) depends on the voxel alone, so the
// brick sums v * w and applies the phase once per voxel when it adds the
// brick to the accumulators (alternating_sign.inl:10-20,
// tabulated_phase.inl:12-37).
//
// Every plane passes through the origin, so bricks near it see every slice
// and bricks far out see few (about 13 / r of them). A task is a brick and a
// share of the slices: heavy bricks are split, and the heaviest tasks start
// first. Voxels that only one task can touch are added to the accumulators
// by plain read-add-write; the others (the taps - 1 layers at each face that
// the neighbouring bricks also reach, every voxel of a split brick, and bricks
// holding the y or z frequency +-N/2 of an even N, two signed frequencies of
// one stored element) add with atomics.
// ---------------------------------------------------------------------------


// Taps per axis an LDS brick is laid out for: the window's own (box 1,
// triangle 2), at most 4 for Kaiser-Bessel windows on this schedule.
template <int PROFILE>
struct brick_reach
{
	static constexpr int taps = PROFILE == REXLIB_HIP_BOX ? 1 : (PROFILE == REXLIB_HIP_TRIANGLE ? 2 : 4);
};

// LDS layout of a brick of first-tap cells EX x EY x EZ with the taps - 1
// voxels above them: odd pitches, so that lanes stepping along any axis fall
// in different banks.
template <int T, int EX, int EY, int EZ>
struct brick_layout
{
	static constexpr int lx = EX + T - 1;
	static constexpr int ly = EY + T - 1;
	static constexpr int lz = EZ + T - 1;
	static constexpr int pitch_y = lx % 2 == 0 ? lx + 1 : lx;
	static constexpr int pitch_z = (pitch_y * ly) % 2 == 0 ? pitch_y * ly + 1 : pitch_y * ly;
	static constexpr int voxels = pitch_z * lz;
	static constexpr int cells = EX * EY * EZ;
};

// A voxel of a brick in LDS: 12 bytes. (16-byte voxels, one 128-bit access
// per tap, measured slower: the larger brick costs wavefronts per CU.)
struct brick_voxel
{
	float re;
	float im;
	float den;
};

struct brick_box
{
	int lo[3];  // signed (x, y, z) of the first cell
	int ext[3]; // cells along (x, y, z)
	bool alias_free;
};

__device__ inline brick_box unpack_brick(int4 b)
{
	brick_box box;
	box.lo[0] = b.x;
	box.lo[1] = b.y;
	box.lo[2] = b.z;
	box.ext[0] = b.w & 0xff;
	box.ext[1] = (b.w >> 8) & 0xff;
	box.ext[2] = (b.w >> 16) & 0xff;
	box.alias_free = ((b.w >> 24) & 1) != 0;
	return box;
}

template <int VARIANT>
__global__ __launch_bounds__(256) void prepare_kernel(insertion_params p, slice_chunk chunk, prepared_sample *prepared)
{
	// A slice holds n (n/2 + 1) < 2^31 elements: 32-bit indices (a 64-bit
	// division per element would cost more than the sample itself).
	const int elements32 = p.n * p.h;
	const int64_t elements = elements32;
	const int idx = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
	if (idx >= elements32)
	{
		return;
	}
	const int row = static_cast<int>(static_cast<unsigned>(idx) / static_cast<unsigned>(p.h));
	const int column = idx - row * p.h;
	const int64_t slices = p.map.kept_count * chunk.count;
	for (int64_t j = blockIdx.y; j < slices; j += gridDim.y)
	{
		const int64_t kept = j / chunk.count;
		const int64_t s = kept * p.map.reduced_count + chunk.begin + j % chunk.count;
		const slice_entry &e = p.entries[s];
		// Branch-free, as in the fused column pass: every element is loaded
		// and the rejected ones zeroed afterwards.
		const float2 *table = p.tables + s * static_cast<int64_t>(p.h + p.n);
		const float2 x = p.slices[e.slice_offset + row * p.slice_s0 + column * p.slice_s1];
		prepared[j * elements + idx] = prepare_unbranched<VARIANT>(p, e, p.radial + e.radial_offset, table + p.h,
			table[column], row, column, x);
	}
}

// The column transforms of 256-point slices fused with the preparation.
// The rows of each slice are transformed already (in p.slices, at the
// slices' own offsets and strides); a block transforms 16 columns k_x of the
// chunk's slices by the four-step decomposition and writes the prepared
// samples of those columns directly, so the finished spectra never pass
// through HBM. The Nyquist column k_x = n/2 holds no inserted sample and the
// brick schedule never reads it, so it is skipped: the lines are (slice j of
// the chunk, k_x < n/2), numbered j n/2 + k_x, and a block's 16 columns never
// straddle two slices.
template <int VARIANT>
__global__ __launch_bounds__(256) void columns_prepare256_kernel(insertion_params p, slice_chunk chunk,
	const float2 *tw, prepared_sample *prepared)
{
	__shared__ float2 exchange[fft256_lines * fft256_line_stride];
	__shared__ float2 twl[256];
	__shared__ int64_t line_slice[fft256_lines];
	__shared__ int64_t line_src[fft256_lines];
	__shared__ int64_t line_dst[fft256_lines];
	__shared__ int line_column[fft256_lines];
	const int t = threadIdx.x;
	const int l = t & (fft256_lines - 1);
	const int j = t / fft256_lines; // n1 in the first half of the transform, k1 in the second
	twl[t] = tw[t];
	const int h = p.h;
	const int columns = p.n / 2;
	const int64_t elements = static_cast<int64_t>(p.n) * h;
	const int64_t total = p.map.kept_count * chunk.count * columns;
	for (int64_t first = xcd_block(blockIdx.x, gridDim.x) * static_cast<int64_t>(fft256_lines); first < total;
		first += static_cast<int64_t>(gridDim.x) * fft256_lines)
	{
		const int lines = static_cast<int>(total - first < fft256_lines ? total - first : fft256_lines);
		if (t < fft256_lines)
		{
			int64_t slice = 0;
			int64_t src = 0;
			int64_t dst = 0;
			int column = 0;
			if (t < lines)
			{
				const int64_t line = first + t;
				const int64_t jj = line / columns;
				column = static_cast<int>(line - jj * columns);
				const int64_t kept = jj / chunk.count;
				slice = kept * p.map.reduced_count + chunk.begin + jj % chunk.count;
				src = p.entries[slice].slice_offset + column * p.slice_s1;
				dst = jj * elements + column;
			}
			line_slice[t] = slice;
			line_src[t] = src;
			line_dst[t] = dst;
			line_column[t] = column;
		}
		__syncthreads();
		const bool live = l < lines;
		// Elements j + 16 k of the column, by pointer steps.
		const float2 *in = p.slices + line_src[l] + j * p.slice_s0;
		const int64_t in_step = 16 * p.slice_s0;
		float2 v[16];
#pragma unroll
		for (int k = 0; k < 16; ++k)
		{
			v[k] = live ? in[k * in_step] : make_float2(0.0f, 0.0f);
		}
		float2 *x = exchange + l * fft256_line_stride;
		fft256_four_step(v, x, twl, j, -1.0f);
		// The outputs go back to LDS (each thread to slots only it reads), so
		// the preparation loops over them without holding all 16.
		__syncthreads();
#pragma unroll
		for (int k2 = 0; k2 < 16; ++k2)
		{
			x[k2 * 17 + j] = v[k2];
		}
		if (live)
		{
			const int64_t slice = line_slice[l];
			const slice_entry &e = p.entries[slice];
			const int column = line_column[l];
			const float2 *table = p.tables + slice * static_cast<int64_t>(h + p.n);
			const float2 x_phase = table[column];
			const float *radial = p.radial + e.radial_offset;
			prepared_sample *out = prepared + line_dst[l] + j * h;
#pragma unroll 4
			for (int k2 = 0; k2 < 16; ++k2)
			{
				out[k2 * 16 * h] = prepare_unbranched<VARIANT>(p, e, radial, table + h, x_phase, j + 16 * k2, column,
					x[k2 * 17 + j]);
			}
		}
		__syncthreads();
	}
}

// The same fusion for the other lengths the FFT kernels take: a block
// stages `group` columns of the chunk's slices in LDS, transforms them with
// the Stockham stages (compile-time for N > 0, planned for N = 0), and
// prepares every element from LDS. The lines are (slice j of the chunk,
// column), numbered j columns + column, where columns is n/2 for even n
// (the Nyquist column, never inserted, is skipped) and n/2 + 1 for odd n.
template <int VARIANT, int N>
__global__ __launch_bounds__(256) void columns_prepare_kernel(insertion_params p, slice_chunk chunk, fft_plan plan,
	const float2 *tw, int group, prepared_sample *prepared)
{
	extern __shared__ float2 lds[];
	const int n = N > 0 ? N : plan.n;
	const int stride = n + 1;
	float2 *a = lds;
	float2 *b = lds + group * stride;
	float2 *twl = lds + 2 * group * stride;
	int64_t *line_slice = reinterpret_cast<int64_t *>(twl + n);
	int64_t *line_src = line_slice + group;
	int64_t *line_dst = line_src + group;
	int *line_column = reinterpret_cast<int *>(line_dst + group);
	for (int i = threadIdx.x; i < n; i += blockDim.x)
	{
		twl[i] = tw[i];
	}
	const int h = p.h;
	const int columns = n % 2 == 0 ? n / 2 : h;
	const int64_t elements = static_cast<int64_t>(n) * h;
	const int64_t total = p.map.kept_count * chunk.count * columns;
	for (int64_t first = xcd_block(blockIdx.x, gridDim.x) * static_cast<int64_t>(group); first < total;
		first += static_cast<int64_t>(gridDim.x) * group)
	{
		const int lines = static_cast<int>(total - first < group ? total - first : group);
		for (int q = threadIdx.x; q < lines; q += blockDim.x)
		{
			const int64_t line = first + q;
			const int64_t jj = line / columns;
			const int column = static_cast<int>(line - jj * columns);
			const int64_t kept = jj / chunk.count;
			const int64_t slice = kept * p.map.reduced_count + chunk.begin + jj % chunk.count;
			line_slice[q] = slice;
			line_src[q] = p.entries[slice].slice_offset + column * p.slice_s1;
			line_dst[q] = jj * elements + column;
			line_column[q] = column;
		}
		__syncthreads();
		for (int t = threadIdx.x; t < group * n; t += blockDim.x)
		{
			const int e = t / group;
			const int l = t - e * group;
			if (l < lines)
			{
				a[l * stride + e] = p.slices[line_src[l] + e * p.slice_s0];
			}
		}
		__syncthreads();
		const float2 *z = transform<N>(a, b, lines, stride, plan, twl, -1.0f);
		for (int t = threadIdx.x; t < group * n; t += blockDim.x)
		{
			const int e = t / group;
			const int l = t - e * group;
			if (l < lines)
			{
				const int64_t slice = line_slice[l];
				const slice_entry &entry = p.entries[slice];
				const float2 *table = p.tables + slice * static_cast<int64_t>(h + n);
				const int column = line_column[l];
				prepared[line_dst[l] + static_cast<int64_t>(e) * h] = prepare_unbranched<VARIANT>(p, entry,
					p.radial + entry.radial_offset, table + h, table[column], e, column, z[l * stride + e]);
			}
		}
		__syncthreads();
	}
}

// bound_rows of one slice and sign: the rows [r0, r1) whose samples k_x in
// [first, widest] can reach the box [lower, upper]; r0 >= r1 when none.
// Evaluated per lane, one slice per lane.
// `inv_y` and `inv_x` hold 1 / y_step_f and 1 / x_step per axis (0 where the
// step is 0); a product with them stands for the division, the box being
// widened far beyond its rounding.
//
// The reference bounds each axis separately over the whole k_x range
// (row_interval.cpp:34-136), which keeps rows whose line meets the box only
// at k_x the slice does not hold. Here the bound is refined once: the k_x
// (integers) that can reach the box from those rows, then the rows again
// for that k_x range. Each step keeps every row holding a sample in the
// box, and about half the rows go.
__device__ inline void bound_rows(const insertion_params &p, const float *x_step, const float *y_step_f,
	const float *inv_y, const float *inv_x, int last_shell, const float *lower, const float *upper, int sign_index,
	int &r0, int &r1)
{
	r0 = 0;
	r1 = 0;
	const int n = p.n;
	const int h = p.h;
	const bool even = n % 2 == 0;
	const int last_column = h - (even ? 2 : 1);
	const int highest_row = even ? n / 2 - 1 : n / 2;
	const int widest = min(last_column, last_shell);
	const int first_row = max(-highest_row, -widest);
	const int last_row = min(highest_row, widest);
	const float sign = sign_index == 0 ? 1.0f : -1.0f;
	const int first = sign_index;
	if (last_shell < 0 || first > widest || first_row > last_row)
	{
		return;
	}
	float low = static_cast<float>(first_row);
	float high = static_cast<float>(last_row);
	bool none = false;
	for (int a = 0; a < 3; ++a)
	{
		const float step = sign * x_step[a];
		const float y_step = sign * y_step_f[a];
		const float at_first = static_cast<float>(first) * step;
		const float at_last = static_cast<float>(widest) * step;
		const float lo = lower[a] - fmaxf(at_first, at_last);
		const float hi = upper[a] - fminf(at_first, at_last);
		if (y_step == 0.0f)
		{
			none = none || !(lo <= 0.0f && 0.0f <= hi);
		}
		else
		{
			const float inv = sign * inv_y[a];
			const float l = lo * inv;
			const float u = hi * inv;
			low = fmaxf(low, fminf(l, u));
			high = fminf(high, fmaxf(l, u));
		}
	}
	if (none || low > high)
	{
		return;
	}
	float k_low = static_cast<float>(first);
	float k_high = static_cast<float>(widest);
	for (int a = 0; a < 3; ++a)
	{
		const float step = sign * x_step[a];
		if (step != 0.0f)
		{
			const float y_step = sign * y_step_f[a];
			const float b0 = low * y_step;
			const float b1 = high * y_step;
			const float inv = sign * inv_x[a];
			const float l = (lower[a] - fmaxf(b0, b1)) * inv;
			const float u = (upper[a] - fminf(b0, b1)) * inv;
			k_low = fmaxf(k_low, fminf(l, u));
			k_high = fminf(k_high, fmaxf(l, u));
		}
	}
	k_low = ceilf(k_low);
	k_high = floorf(k_high);
	if (k_low > k_high)
	{
		return;
	}
	for (int a = 0; a < 3; ++a)
	{
		const float step = sign * x_step[a];
		const float y_step = sign * y_step_f[a];
		if (y_step != 0.0f)
		{
			const float at_first = k_low * step;
			const float at_last = k_high * step;
			const float inv = sign * inv_y[a];
			const float l = (lower[a] - fmaxf(at_first, at_last)) * inv;
			const float u = (upper[a] - fminf(at_first, at_last)) * inv;
			low = fmaxf(low, fminf(l, u));
			high = fminf(high, fmaxf(l, u));
		}
	}
	if (low > high)
	{
		return;
	}
	// The box is already widened past the window, so rounding inwards keeps
	// every sample that can reach it.
	r0 = max(static_cast<int>(ceilf(low)), first_row);
	r1 = min(static_cast<int>(floorf(high)) + 1, last_row + 1);
}

// Position of sample k_x along one axis from its row's base and step (the
// sign applied): the row base k_y * Y is formed in double and rounded to
// float once per row, then base + k_x * X in float, uncontracted
// (row_insertion_kernel.inl:157-166).
__device__ inline float row_position(float base, float step, int kx)
{
#pragma clang fp contract(off)
	const float product = static_cast<float>(kx) * step;
	return base + product;
}

// clip_row of one row k_y and sign: the samples k_x in [begin, end) (end <=
// begin when none) within the shell radius whose positions can fall in
// [lower, upper], a box widened by 0.01 voxel beyond every rounding here.
// `inv_step` holds 1 / step per axis (0 where the step is 0); a product with
// it stands for the division.
__device__ inline void clip_row(int ky, int first, int widest, float radius, const float *step, const float *y_step,
	const float *inv_step, const float *lower, const float *upper, int &begin, int &end)
{
	begin = 0;
	end = 0;
	const float kyf = static_cast<float>(ky);
	const float room = radius * radius - kyf * kyf;
	if (room < 0.0f)
	{
		return;
	}
	const int last = min(widest, static_cast<int>(sqrtf(room)));
	float lo_k = static_cast<float>(first);
	float hi_k = static_cast<float>(last);
	bool none = first > last;
#pragma unroll
	for (int a = 0; a < 3; ++a)
	{
		const float base = kyf * y_step[a];
		if (step[a] == 0.0f)
		{
			none = none || !(lower[a] <= base && base <= upper[a]);
		}
		else
		{
			const float l = (lower[a] - base) * inv_step[a];
			const float u = (upper[a] - base) * inv_step[a];
			lo_k = fmaxf(lo_k, fminf(l, u));
			hi_k = fminf(hi_k, fmaxf(l, u));
		}
	}
	if (!none && lo_k <= hi_k)
	{
		begin = max(static_cast<int>(ceilf(lo_k)), first);
		end = max(begin, min(static_cast<int>(floorf(hi_k)) + 1, last + 1));
	}
}

// A wavefront owns the brick of one task (a brick and a share of the
// slices): EX x EY x EZ first-tap cells and the voxels their taps reach, in
// LDS.
template <int PROFILE, bool ODD, int EX, int EY, int EZ>
__global__ __launch_bounds__(64) void brick_kernel(insertion_params p, slice_chunk chunk, const prepared_sample *prepared,
	const int4 *tasks, int task_count)
{
	constexpr int T = brick_reach<PROFILE>::taps;
	using L = brick_layout<T, EX, EY, EZ>;
	__shared__ brick_voxel acc[L::voxels];
	// Per first-tap cell: the lane that last claimed it.
	__shared__ unsigned char claim[L::cells];
	// Per lane: the last row (or slice) starting at or before it.
	__shared__ int marker[64];

	const int lane = threadIdx.x;
	const int big = p.big;
	const int n = p.n;
	const int h = p.h;
	const int64_t elements = static_cast<int64_t>(n) * h;
	const int64_t kept = blockIdx.y;
	const int taps = PROFILE == REXLIB_HIP_KAISER_BESSEL ? p.taps : T;
	const int last_column = h - (n % 2 == 0 ? 2 : 1);
	// A first tap f holds the positions [f + offset, f + 1 + offset).
	const float cell_offset = PROFILE == REXLIB_HIP_BOX ? -0.5f :
		(PROFILE == REXLIB_HIP_TRIANGLE ? 0.0f : p.kb_half_width - 1.0f);
	const uint64_t below = lane == 0 ? 0ull : (~0ull >> (64 - lane));

	for (int ti = blockIdx.x; ti < task_count; ti += gridDim.x)
	{
		const brick_box box = unpack_brick(tasks[2 * ti]);
		const int4 share = tasks[2 * ti + 1];
		const int64_t share_begin = chunk.count * share.x / share.y;
		const int64_t share_end = chunk.count * (share.x + 1) / share.y;
		for (int v = lane; v < L::voxels; v += 64)
		{
			acc[v] = brick_voxel{0.0f, 0.0f, 0.0f};
		}
		lds_order();
		// The positions whose first tap lies in the brick, widened so that
		// rounding never drops a sample (each sample is then tested exactly).
		float lower[3];
		float upper[3];
		float centre[3];
		float half_extent[3];
		float half_diagonal2 = 0.0f;
		for (int a = 0; a < 3; ++a)
		{
			lower[a] = static_cast<float>(box.lo[a]) + cell_offset - 0.01f;
			upper[a] = static_cast<float>(box.lo[a] + box.ext[a]) + cell_offset + 0.01f;
			centre[a] = 0.5f * (lower[a] + upper[a]);
			half_extent[a] = 0.5f * (upper[a] - lower[a]) + 0.01f;
			half_diagonal2 += half_extent[a] * half_extent[a];
		}
		const float half_diagonal = sqrtf(half_diagonal2);
		const float centre_norm = sqrtf(centre[0] * centre[0] + centre[1] * centre[1] + centre[2] * centre[2]);

		// The sample this lane holds, not yet added: value, cell (bits 0-23,
		// bit 24 set while held) and per axis the fraction past the first tap
		// (triangle) or the position (Kaiser-Bessel).
		bool pending = false;
		float held_re = 0.0f, held_im = 0.0f, held_den = 0.0f;
		int held_cell = 0;
		float held_w[3] = {0.0f, 0.0f, 0.0f};

		// One round: every holding lane claims its cell and the winners add
		// their taps; meanwhile (when `fill`) the lanes free after the round
		// load the next samples of the current rows, which they hold from the
		// next round on.
		int next = 0;
		int total = 0;
		int current = 0;
		int row_start = 0;
		int row_count = 0;
		// Lane k describes row k of the current 64 rows (registers, not LDS,
		// so that more wavefronts fit a CU): base position and the first k_x
		// less the row's first sample number (so sample s of the batch has
		// k_x = that + s); step; the prepared element of (row, k_x = 0), bit
		// 31 set for the -k copy.
		float4 row_base = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
		float4 row_step = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
		uint32_t row_index = 0;
		auto round = [&](bool fill)
		{
			// Claim: each holding lane writes its lane into the byte of its
			// first-tap cell and reads it back. A plain LDS read after the
			// store (lds_order keeps the compiler from reusing the stored
			// value); a volatile pointer would turn it into a flat load, which
			// is not ordered after the LDS store.
			const int lx = held_cell & 0xff;
			const int ly = (held_cell >> 8) & 0xff;
			const int lz = (held_cell >> 16) & 0xff;
			const int cell = lx + EX * (ly + EY * lz);
			if (pending)
			{
				claim[cell] = static_cast<unsigned char>(lane);
			}
			lds_order();
			const bool win = pending && claim[cell] == lane;
			lds_order();

			// The fill starts before the winners' taps: the lanes free after
			// this round (holding nothing, or winning now) take the next
			// samples of the current rows, and the loads of those samples are
			// in flight while the taps run. The samples join the claim of the
			// next round.
			int take = 0;
			int r = 0;
			bool mine = false;
			float4 b = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
			float4 st = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
			uint32_t ix = 0;
			int kx = 0;
			prepared_sample q{0.0f, 0.0f, 0.0f};
			if (fill)
			{
				const bool free = !pending || win;
				const uint64_t free_mask = __ballot(free);
				take = min(static_cast<int>(__popcll(free_mask)), total - next);
				if (take > 0)
				{
					// The row of sample next + lane: rows starting in [next, next + 64)
					// mark their slot, slot 0 starts with the row holding `next`.
					marker[lane] = lane == 0 ? current : 0;
					lds_order();
					if (row_count > 0 && row_start >= next && row_start < next + 64)
					{
						marker[row_start - next] = lane;
					}
					lds_order();
					r = wave_prefix_max(marker[lane]);
					lds_order();
					current = __builtin_amdgcn_readlane(r, take - 1);
					// Free lane k (in lane order) takes sample next + k itself: the
					// row of that sample, then the row's data, by shuffles.
					const int rank = static_cast<int>(__popcll(free_mask & below));
					mine = free && rank < take;
					const int row = __shfl(r, mine ? rank : 0, 64);
					b = make_float4(__shfl(row_base.x, row, 64), __shfl(row_base.y, row, 64),
						__shfl(row_base.z, row, 64), __shfl(row_base.w, row, 64));
					st = make_float4(__shfl(row_step.x, row, 64), __shfl(row_step.y, row, 64),
						__shfl(row_step.z, row, 64), 0.0f);
					ix = static_cast<uint32_t>(__shfl(static_cast<int>(row_index), row, 64));
					kx = __float_as_int(b.w) + next + rank;
					if (mine)
					{
						q = prepared[(ix & 0x7fffffffu) + static_cast<uint32_t>(kx)];
					}
				}
			}

			if (win)
			{
				const int base = lx + L::pitch_y * ly + L::pitch_z * lz;
				float wx[T], wy[T], wz[T];
				if constexpr (PROFILE == REXLIB_HIP_TRIANGLE)
				{
					wx[0] = 1.0f - held_w[0];
					wx[1] = held_w[0];
					wy[0] = 1.0f - held_w[1];
					wy[1] = held_w[1];
					wz[0] = 1.0f - held_w[2];
					wz[1] = held_w[2];
				}
				else if constexpr (PROFILE == REXLIB_HIP_BOX)
				{
					wx[0] = 1.0f;
					wy[0] = 1.0f;
					wz[0] = 1.0f;
				}
				else
				{
					window_taps<PROFILE>::template place<T>(p, held_w[0], wx);
					window_taps<PROFILE>::template place<T>(p, held_w[1], wy);
					window_taps<PROFILE>::template place<T>(p, held_w[2], wz);
				}
#pragma unroll
				for (int tz = 0; tz < T; ++tz)
				{
#pragma unroll
					for (int ty = 0; ty < T; ++ty)
					{
						const float wzy = wz[tz] * wy[ty];
#pragma unroll
						for (int tx = 0; tx < T; ++tx)
						{
							if (tx < taps && ty < taps && tz < taps)
							{
								const float wt = wzy * wx[tx];
								brick_voxel &a = acc[base + tx + L::pitch_y * ty + L::pitch_z * tz];
								brick_voxel v = a;
								v.re += held_re * wt;
								v.im += held_im * wt;
								v.den += held_den * wt;
								a = v;
							}
							lds_order();
						}
					}
				}
			}
			lds_order();
			pending = pending && !win;

			if (take > 0)
			{
				const float c[3] = {row_position(b.x, st.x, kx), row_position(b.y, st.y, kx),
					row_position(b.z, st.z, kx)};
				int local[3];
				float w[3];
#pragma unroll
				for (int a = 0; a < 3; ++a)
				{
					const int f = window_taps<PROFILE>::first(p, c[a]);
					local[a] = f - box.lo[a];
					w[a] = PROFILE == REXLIB_HIP_TRIANGLE ? c[a] - static_cast<float>(f) : c[a];
				}
				const bool own = static_cast<unsigned>(local[0]) < static_cast<unsigned>(box.ext[0]) &&
					static_cast<unsigned>(local[1]) < static_cast<unsigned>(box.ext[1]) &&
					static_cast<unsigned>(local[2]) < static_cast<unsigned>(box.ext[2]);
				const bool value = !(q.re == 0.0f && q.im == 0.0f && q.den == 0.0f);
				if (mine && own && value)
				{
					pending = true;
					held_re = q.re;
					held_im = (ix >> 31) != 0 ? -q.im : q.im;
					held_den = q.den;
					held_cell = local[0] | (local[1] << 8) | (local[2] << 16) | (1 << 24);
					held_w[0] = w[0];
					held_w[1] = w[1];
					held_w[2] = w[2];
				}
				next += take;
			}
		};

		for (int64_t window = share_begin; window < share_end; window += 64)
		{
			// Slices whose plane passes near the brick, one per lane.
			const int64_t j = window + lane;
			bool near = false;
			float x0 = 0.0f, x1 = 0.0f, x2 = 0.0f, y0 = 0.0f, y1 = 0.0f, y2 = 0.0f;
			int last_shell = -1;
			if (j < share_end)
			{
				const slice_entry &e = p.entries[kept * p.map.reduced_count + chunk.begin + j];
				last_shell = e.last_shell;
				if (e.last_shell >= 0)
				{
					x0 = e.x_step[0];
					x1 = e.x_step[1];
					x2 = e.x_step[2];
					y0 = static_cast<float>(e.y_step[0]);
					y1 = static_cast<float>(e.y_step[1]);
					y2 = static_cast<float>(e.y_step[2]);
					const float nx = x1 * y2 - x2 * y1;
					const float ny = x2 * y0 - x0 * y2;
					const float nz = x0 * y1 - x1 * y0;
					// The plane meets the box iff |n . c| is at most the box's
					// support sum_a |n_a| h_a along the normal n.
					const float distance = fabsf(nx * centre[0] + ny * centre[1] + nz * centre[2]);
					const float support = fabsf(nx) * half_extent[0] + fabsf(ny) * half_extent[1] +
						fabsf(nz) * half_extent[2];
					const float radius = static_cast<float>(e.last_shell) + 0.5f;
					const float spread =
						radius * sqrtf(x0 * x0 + x1 * x1 + x2 * x2 + y0 * y0 + y1 * y1 + y2 * y2);
					near = distance <= support * 1.001f + 1e-3f &&
						centre_norm <= spread * 1.001f + half_diagonal + 1e-3f;
				}
			}
			// bound_rows of both signs, one slice per lane, and the
			// reciprocal steps of the frame (shared by both signs).
			int r0_plus = 0, r1_plus = 0, r0_minus = 0, r1_minus = 0;
			float ix0 = 0.0f, ix1 = 0.0f, ix2 = 0.0f;
			if (near)
			{
				const float xs[3] = {x0, x1, x2};
				const float ys[3] = {y0, y1, y2};
				// The hardware reciprocal (within 1 ulp): the bounds it feeds
				// are widened by 0.01 voxel, far beyond its error.
				float iy[3];
				for (int a = 0; a < 3; ++a)
				{
					iy[a] = ys[a] == 0.0f ? 0.0f : __builtin_amdgcn_rcpf(ys[a]);
				}
				ix0 = x0 == 0.0f ? 0.0f : __builtin_amdgcn_rcpf(x0);
				ix1 = x1 == 0.0f ? 0.0f : __builtin_amdgcn_rcpf(x1);
				ix2 = x2 == 0.0f ? 0.0f : __builtin_amdgcn_rcpf(x2);
				const float ixs[3] = {ix0, ix1, ix2};
				bound_rows(p, xs, ys, iy, ixs, last_shell, lower, upper, 0, r0_plus, r1_plus);
				bound_rows(p, xs, ys, iy, ixs, last_shell, lower, upper, 1, r0_minus, r1_minus);
			}
			const int plus_rows = near ? max(r1_plus - r0_plus, 0) : 0;
			const int minus_rows = near ? max(r1_minus - r0_minus, 0) : 0;
			const int rows = plus_rows + minus_rows;
			const int rows_inclusive = wave_inclusive_scan(rows);
			const int rows_first = rows_inclusive - rows;
			const int total_rows = __builtin_amdgcn_readlane(rows_inclusive, 63);
			for (int rb = 0; rb < total_rows; rb += 64)
			{
				// The slice of row rb + lane: every slice with rows in
				// [rb, rb + 64) marks the slot of its first one (slot 0 for
				// the slice holding row rb); a prefix maximum hands each row
				// its slice's lane.
				marker[lane] = 0;
				lds_order();
				if (rows > 0 && rows_first < rb + 64 && rows_first + rows > rb)
				{
					marker[max(rows_first - rb, 0)] = lane;
				}
				lds_order();
				const int s = wave_prefix_max(marker[lane]);
				lds_order();
				const int r = rb + lane;
				// Every shuffle runs with all lanes active: a lane reads nothing
				// from a source lane that is switched off.
				const int idx = r - __shfl(rows_first, s, 64);
				const int np = __shfl(plus_rows, s, 64);
				const int first_plus = __shfl(r0_plus, s, 64);
				const int first_minus = __shfl(r0_minus, s, 64);
				const bool minus = idx >= np;
				const int ky = minus ? first_minus + (idx - np) : first_plus + idx;
				const float sign = minus ? -1.0f : 1.0f;
				const float frame_x[3] = {__shfl(x0, s, 64), __shfl(x1, s, 64), __shfl(x2, s, 64)};
				const float frame_y[3] = {__shfl(y0, s, 64), __shfl(y1, s, 64), __shfl(y2, s, 64)};
				const float frame_ix[3] = {__shfl(ix0, s, 64), __shfl(ix1, s, 64), __shfl(ix2, s, 64)};
				const int shell = __shfl(last_shell, s, 64);
				const int64_t js = window + s;
				float step[3];
				float y_step[3];
				float inv_step[3];
				float base[3];
				int begin = 0;
				int end = 0;
				if (r < total_rows)
				{
					const slice_entry &e = p.entries[kept * p.map.reduced_count + chunk.begin + js];
					for (int a = 0; a < 3; ++a)
					{
						step[a] = sign * frame_x[a];
						y_step[a] = sign * frame_y[a];
						inv_step[a] = sign * frame_ix[a];
						base[a] = static_cast<float>(static_cast<double>(ky) *
							(minus ? -e.y_step[a] : e.y_step[a]));
					}
					clip_row(ky, minus ? 1 : 0, min(last_column, shell), static_cast<float>(shell) + 0.5f, step, y_step,
						inv_step, lower, upper, begin, end);
				}
				else
				{
					for (int a = 0; a < 3; ++a)
					{
						step[a] = 0.0f;
						base[a] = 0.0f;
					}
				}
				row_count = end - begin;
				const int count_inclusive = wave_inclusive_scan(row_count);
				row_start = count_inclusive - row_count;
				total = __builtin_amdgcn_readlane(count_inclusive, 63);
				const int stored_row = ky < 0 ? ky + n : ky;
				row_base = make_float4(base[0], base[1], base[2], __int_as_float(begin - row_start));
				row_step = make_float4(step[0], step[1], step[2], 0.0f);
				row_index = r < total_rows ? (static_cast<uint32_t>((kept * chunk.count + js) * elements +
					static_cast<int64_t>(stored_row) * h) | (minus ? 0x80000000u : 0u)) : 0u;
				lds_order();
				next = 0;
				current = 0;
				while (next < total)
				{
					round(true);
				}
				lds_order();
			}
		}
		// Samples still held after the last rows.
		while (__ballot(pending) != 0)
		{
			round(false);
		}
		lds_order();

		// Add the brick to the accumulators of its kept batch element.
		// Voxels no other task touches (at least taps - 1 past the lower faces,
		// inside the cells, a brick of one task without aliasing) take a plain
		// read-add-write, the others atomics.
		const bool single = share.y == 1 && box.alias_free;
		const int he = big / 2;
		const int64_t num_base = kept_offset(p.map, slot_numerator_out, kept);
		const int64_t den_base = kept_offset(p.map, slot_denominator_out, kept);
		for (int v = lane; v < L::voxels; v += 64)
		{
			const int lz = v / L::pitch_z;
			const int rest = v - lz * L::pitch_z;
			const int ly = rest / L::pitch_y;
			const int lx = rest - ly * L::pitch_y;
			if (lx >= L::lx || ly >= L::ly)
			{
				continue;
			}
			const brick_voxel sum = acc[v];
			if (sum.re == 0.0f && sum.im == 0.0f && sum.den == 0.0f)
			{
				continue;
			}
			const int x = box.lo[0] + lx;
			const int y = box.lo[1] + ly;
			const int z = box.lo[2] + lz;
			if (x < 0 || x > he || y < -he || y > he || z < -he || z > he)
			{
				continue;
			}
			const int sy = y < 0 ? y + big : y;
			const int sz = z < 0 ? z + big : z;
			const float2 value = volume_phase<ODD>(p, make_float2(sum.re, sum.im), x, sy, sz);
			float2 *target = p.numerator + num_base + sz * p.num_s0 + sy * p.num_s1 + x * p.num_s2;
			float *dtarget = p.denominator + den_base + sz * p.den_s0 + sy * p.den_s1 + x * p.den_s2;
			const bool exclusive = single && lx >= T - 1 && ly >= T - 1 && lz >= T - 1 && lx < box.ext[0] &&
				ly < box.ext[1] && lz < box.ext[2];
			if (exclusive)
			{
				float2 t = *target;
				t.x += value.x;
				t.y += value.y;
				*target = t;
				*dtarget += sum.den;
			}
			else
			{
				atomicAdd(&target->x, value.x);
				atomicAdd(&target->y, value.y);
				atomicAdd(dtarget, sum.den);
			}
		}
		lds_order();
	}
}

// Brick shapes (first-tap cells along x, y, z): 16 x 8 x 6 (the default),
// 16 x 8 x 8, 16 x 16 x 8, 8 x 8 x 8, 8 x 8 x 6, 8 x 6 x 6 and 8 x 8 x 4, and
// the expected plane hits per task that splits busy bricks. Smaller bricks
// take less LDS (more wavefronts per CU) but hold more halo voxels per cell.
// One entry per volume extent N (0: every extent without an entry of its
// own); a call reads its extent's entry once, so concurrent callers each see
// a whole entry.
struct brick_tuning
{
	int ex;
	int ey;
	int ez;
	int share;
};

constexpr brick_tuning default_brick_tuning{16, 8, 6, 256};

std::mutex tuning_mutex;
std::map<int, brick_tuning> tuning_table;

bool valid_brick_shape(int ex, int ey, int ez)
{
	return (ex == 16 && ey == 8 && (ez == 6 || ez == 8)) || (ex == 16 && ey == 16 && ez == 8) ||
		(ex == 8 && ey == 8 && (ez == 8 || ez == 6 || ez == 4)) || (ex == 8 && ey == 6 && ez == 6);
}

// Table keys: the extent (every window), or the extent plus taps_key times
// the taps per axis of the brick layout (1, 2 or 4) for one window family.
constexpr int taps_key = 1 << 20;

// The tuning of an extent for a window of `taps` per axis (0: any window):
// the window's own entry, else the extent's, else the entry for every
// extent, else the built-in default.
brick_tuning tuning_for(int extent, int taps = 0)
{
	std::lock_guard<std::mutex> lock(tuning_mutex);
	auto found = taps > 0 ? tuning_table.find(extent + taps * taps_key) : tuning_table.end();
	if (found == tuning_table.end())
	{
		found = tuning_table.find(extent);
	}
	if (found == tuning_table.end())
	{
		found = tuning_table.find(0);
	}
	return found != tuning_table.end() ? found->second : default_brick_tuning;
}

template <int PROFILE, int EX, int EY, int EZ>
void launch_brick_shape(const insertion_params &p, slice_chunk chunk, const prepared_sample *prepared, const int4 *tasks,
	int task_count, hipStream_t stream)
{
	const dim3 grid(static_cast<unsigned>(task_count), static_cast<unsigned>(p.map.kept_count));
	if (p.big % 2 == 1)
	{
		brick_kernel<PROFILE, true, EX, EY, EZ><<<grid, 64, 0, stream>>>(p, chunk, prepared, tasks, task_count);
	}
	else
	{
		brick_kernel<PROFILE, false, EX, EY, EZ><<<grid, 64, 0, stream>>>(p, chunk, prepared, tasks, task_count);
	}
}

template <int PROFILE>
void launch_brick(const insertion_params &p, slice_chunk chunk, const prepared_sample *prepared, const int4 *tasks,
	int task_count, const brick_tuning &t, hipStream_t stream)
{
	if (t.ex == 16 && t.ey == 8 && t.ez == 8)
	{
		launch_brick_shape<PROFILE, 16, 8, 8>(p, chunk, prepared, tasks, task_count, stream);
	}
	else if (t.ex == 16 && t.ey == 16)
	{
		launch_brick_shape<PROFILE, 16, 16, 8>(p, chunk, prepared, tasks, task_count, stream);
	}
	else if (t.ex == 8 && t.ey == 8 && t.ez == 8)
	{
		launch_brick_shape<PROFILE, 8, 8, 8>(p, chunk, prepared, tasks, task_count, stream);
	}
	else if (t.ex == 8 && t.ey == 8 && t.ez == 6)
	{
		launch_brick_shape<PROFILE, 8, 8, 6>(p, chunk, prepared, tasks, task_count, stream);
	}
	else if (t.ex == 8 && t.ey == 8)
	{
		launch_brick_shape<PROFILE, 8, 8, 4>(p, chunk, prepared, tasks, task_count, stream);
	}
	else if (t.ex == 8)
	{
		launch_brick_shape<PROFILE, 8, 6, 6>(p, chunk, prepared, tasks, task_count, stream);
	}
	else
	{
		launch_brick_shape<PROFILE, 16, 8, 6>(p, chunk, prepared, tasks, task_count, stream);
	}
}

void dispatch_brick(const insertion_params &p, slice_chunk chunk, const prepared_sample *prepared, const int4 *tasks,
	int task_count, const brick_tuning &tuning, hipStream_t stream)
{
	switch (p.profile)
	{
	case REXLIB_HIP_BOX:
		launch_brick<REXLIB_HIP_BOX>(p, chunk, prepared, tasks, task_count, tuning, stream);
		break;
	case REXLIB_HIP_TRIANGLE:
		launch_brick<REXLIB_HIP_TRIANGLE>(p, chunk, prepared, tasks, task_count, tuning, stream);
		break;
	default:
		launch_brick<REXLIB_HIP_KAISER_BESSEL>(p, chunk, prepared, tasks, task_count, tuning, stream);
		break;
	}
}

void dispatch_prepare(const insertion_params &p, slice_chunk chunk, prepared_sample *prepared, hipStream_t stream)
{
	const int64_t elements = static_cast<int64_t>(p.n) * p.h;
	const dim3 grid(static_cast<unsigned>((elements + 255) / 256),
		static_cast<unsigned>(std::min<int64_t>(p.map.kept_count * chunk.count, 65535)));
	switch (p.variant)
	{
	case variant_plain:
		prepare_kernel<variant_plain><<<grid, 256, 0, stream>>>(p, chunk, prepared);
		break;
	case variant_weighted:
		prepare_kernel<variant_weighted><<<grid, 256, 0, stream>>>(p, chunk, prepared);
		break;
	default:
		prepare_kernel<variant_ctf><<<grid, 256, 0, stream>>>(p, chunk, prepared);
		break;
	}
}

template <int VARIANT>
void launch_columns_prepare(const insertion_params &p, slice_chunk chunk, prepared_sample *prepared,
	hipStream_t stream)
{
	const int n = p.n;
	const int columns = n % 2 == 0 ? n / 2 : p.h;
	const int64_t lines = p.map.kept_count * chunk.count * columns;
	const fft_plan plan = make_fft_plan(n);
	const float2 *tw = twiddles(n);
	const int group = lines_per_group(n);
	const std::size_t shared = sizeof(float2) * (2 * group * (n + 1) + n) + group * (3 * sizeof(int64_t) + sizeof(int));
	const unsigned grid = xcd_grid(static_cast<unsigned>(std::min<int64_t>((lines + group - 1) / group, 1 << 20)));
	switch (n)
	{
	case 64:
		columns_prepare_kernel<VARIANT, 64><<<grid, 256, shared, stream>>>(p, chunk, plan, tw, group, prepared);
		break;
	case 128:
		columns_prepare_kernel<VARIANT, 128><<<grid, 256, shared, stream>>>(p, chunk, plan, tw, group, prepared);
		break;
	case 512:
		columns_prepare_kernel<VARIANT, 512><<<grid, 256, shared, stream>>>(p, chunk, plan, tw, group, prepared);
		break;
	default:
		columns_prepare_kernel<VARIANT, 0><<<grid, 256, shared, stream>>>(p, chunk, plan, tw, group, prepared);
		break;
	}
}

void dispatch_columns_prepare(const insertion_params &p, slice_chunk chunk, prepared_sample *prepared,
	hipStream_t stream)
{
	if (p.n != 256)
	{
		switch (p.variant)
		{
		case variant_plain:
			launch_columns_prepare<variant_plain>(p, chunk, prepared, stream);
			break;
		case variant_weighted:
			launch_columns_prepare<variant_weighted>(p, chunk, prepared, stream);
			break;
		default:
			launch_columns_prepare<variant_ctf>(p, chunk, prepared, stream);
			break;
		}
		return;
	}
	const int64_t lines = p.map.kept_count * chunk.count * (p.n / 2);
	const unsigned grid = xcd_grid(static_cast<unsigned>(
		std::min<int64_t>((lines + fft256_lines - 1) / fft256_lines, 1 << 20)));
	const float2 *tw = twiddles(256);
	switch (p.variant)
	{
	case variant_plain:
		columns_prepare256_kernel<variant_plain><<<grid, 256, 0, stream>>>(p, chunk, tw, prepared);
		break;
	case variant_weighted:
		columns_prepare256_kernel<variant_weighted><<<grid, 256, 0, stream>>>(p, chunk, tw, prepared);
		break;
	default:
		columns_prepare256_kernel<variant_ctf><<<grid, 256, 0, stream>>>(p, chunk, tw, prepared);
		break;
	}
}

// The tasks of a brick grid as (packed box, (share, shares)) pairs,
// heaviest first, for slice counts up to `bucket`, on a device that runs
// `slots` brick wavefronts at once.
//
// A plane through the origin meets a brick at distance r with probability
// about min(1, 13 / r), and the samples it leaves there grow as the brick
// nears the origin (the plane's section of the brick widens), about as
// 1 / r down to the brick's half diagonal. The expected load of a brick is
// taken as bucket * 13 / max(r, half diagonal) plane hits, and a brick is
// split into shares of at most `target` of them -- or fewer, so that no task
// carries more than the average load of a slot (the total over `slots`):
// a task longer than that would outlast the rest of the launch.
std::vector<int4> brick_task_items(int big, int64_t bucket, int target, int ex, int ey, int ez, int taps, int slots)
{
	const int half = big / 2;
	// First-tap cells: x in [-(taps - 1), N/2]; y, z in [-(N/2) - (taps - 1),
	// N/2] (even N: the ends -N/2 and N/2 are one stored element; odd N: N
	// values from -(N/2)).
	const int reach = taps - 1;
	// Whether the voxels [first, first + count + taps - 1) hold -N/2 or N/2.
	const auto aliased = [&](int first, int count)
	{
		const int end = first + count + reach;
		return (first <= -half && -half < end) || (first <= half && half < end);
	};
	const auto starts = [](int low, int high, int edge)
	{
		std::vector<std::pair<int, int>> out;
		for (int s = low; s <= high; s += edge)
		{
			out.emplace_back(s, std::min(edge, high - s + 1));
		}
		return out;
	};
	struct item
	{
		int4 box;
		int4 share;
		double load;
	};
	struct brick
	{
		int4 box;
		double load;
	};
	std::vector<brick> bricks;
	double total = 0.0;
	const double half_diagonal = 0.5 * std::sqrt(static_cast<double>(ex * ex + ey * ey + ez * ez));
	for (const auto &z : starts(-half - reach, half, ez))
	{
		for (const auto &y : starts(-half - reach, half, ey))
		{
			for (const auto &x : starts(-reach, half, ex))
			{
				const bool alias = big % 2 == 0 && (aliased(z.first, z.second) || aliased(y.first, y.second));
				const int w = x.second | (y.second << 8) | (z.second << 16) | ((alias ? 0 : 1) << 24);
				const double cx = x.first + 0.5 * (x.second - 1);
				const double cy = y.first + 0.5 * (y.second - 1);
				const double cz = z.first + 0.5 * (z.second - 1);
				const double r = std::sqrt(cx * cx + cy * cy + cz * cz);
				const double load = static_cast<double>(bucket) * 13.0 / std::max(r, std::max(half_diagonal, 1.0));
				bricks.push_back({make_int4(x.first, y.first, z.first, w), load});
				total += load;
			}
		}
	}
	const double limit = std::min<double>(target, std::max(16.0, total / std::max(slots, 1)));
	std::vector<item> items;
	for (const brick &b : bricks)
	{
		const int shares = static_cast<int>(std::min<double>(64.0, std::max(1.0, std::ceil(b.load / limit))));
		for (int c = 0; c < shares; ++c)
		{
			items.push_back({b.box, make_int4(c, shares, 0, 0), b.load / shares});
		}
	}
	std::stable_sort(items.begin(), items.end(), [](const item &a, const item &b) { return a.load > b.load; });
	std::vector<int4> host;
	for (const auto &i : items)
	{
		host.push_back(i.box);
		host.push_back(i.share);
	}
	return host;
}

// The tasks of a volume of extent N receiving `slices` slices per
// accumulator, cached per device, extent, slice-count bucket, share target
// and brick shape.
struct task_list
{
	std::vector<int4> host;
	int4 *device = nullptr;
};

std::mutex task_mutex;
std::map<std::tuple<int, int, int64_t, int, int, int, int, int>, task_list> task_cache;

// The cached list for a call (made on first use); the caller holds task_mutex.
task_list &brick_task_list(int big, int64_t slices, const brick_tuning &tuning, int taps)
{
	int device = 0;
	REXLIB_HIP_CHECK(hipGetDevice(&device));
	int64_t bucket = 1;
	while (bucket < slices)
	{
		bucket *= 2;
	}
	const auto key = std::make_tuple(device, big, bucket, tuning.share, tuning.ex, tuning.ey, tuning.ez, taps);
	auto found = task_cache.find(key);
	if (found == task_cache.end())
	{
		// Brick wavefronts resident at once: as many per CU as its 160 KiB of
		// LDS holds (the brick's voxels, claim bytes and 64 markers, in
		// 512-byte granules).
		int units = 0;
		REXLIB_HIP_CHECK(hipDeviceGetAttribute(&units, hipDeviceAttributeMultiprocessorCount, device));
		const int lx = tuning.ex + taps - 1, ly = tuning.ey + taps - 1, lz = tuning.ez + taps - 1;
		const int pitch_y = lx % 2 == 0 ? lx + 1 : lx;
		const int pitch_z = (pitch_y * ly) % 2 == 0 ? pitch_y * ly + 1 : pitch_y * ly;
		const int lds = static_cast<int>(align_up(static_cast<std::size_t>(pitch_z * lz * 12 +
			tuning.ex * tuning.ey * tuning.ez + 64 * 4), 512));
		found = task_cache.emplace(key, task_list{}).first;
		found->second.host = brick_task_items(big, bucket, tuning.share, tuning.ex, tuning.ey, tuning.ez, taps,
			units * std::max(1, (160 * 1024) / lds));
	}
	return found->second;
}

// Workspace bytes the task list of a call takes (with its alignment).
std::size_t brick_task_bytes(int big, int64_t slices, const brick_tuning &tuning, int taps)
{
	std::lock_guard<std::mutex> lock(task_mutex);
	return align_up(sizeof(int4) * brick_task_list(big, slices, tuning, taps).host.size(), 256) + 256;
}

// The device copy of a call's tasks. With a fixed workspace buffer the list
// is copied into it from the host copy the cache keeps (nothing allocated);
// otherwise the device copy is made once and cached.
const int4 *brick_tasks(int big, int64_t slices, const brick_tuning &tuning, int taps, workspace_source &workspace,
	hipStream_t stream, int &count)
{
	std::lock_guard<std::mutex> lock(task_mutex);
	task_list &list = brick_task_list(big, slices, tuning, taps);
	count = static_cast<int>(list.host.size() / 2);
	const std::size_t bytes = sizeof(int4) * list.host.size();
	if (workspace.arena != nullptr)
	{
		int4 *at = static_cast<int4 *>(workspace.take(area_insertion, bytes, stream));
		REXLIB_HIP_CHECK(hipMemcpyAsync(at, list.host.data(), bytes, hipMemcpyHostToDevice, stream));
		return at;
	}
	if (list.device == nullptr)
	{
		REXLIB_HIP_CHECK(hipMalloc(&list.device, bytes));
		REXLIB_HIP_CHECK(hipMemcpy(list.device, list.host.data(), bytes, hipMemcpyHostToDevice));
	}
	return list.device;
}

} // anonymous namespace

void set_insertion_tuning(int extent, int taps, int extent_x, int extent_y, int extent_z, int share)
{
	if (extent < 0 || extent >= taps_key)
	{
		fail(REXLIB_HIP_INVALID_ARGUMENT, "insertion tuning: the extent must be 0 (every extent) or positive, "
			"below 2^20");
	}
	if (taps != 0 && taps != 1 && taps != 2 && taps != 4)
	{
		fail(REXLIB_HIP_INVALID_ARGUMENT, "insertion tuning: taps per axis are 0 (every window), 1 (nearest), "
			"2 (trilinear) or 4 (Kaiser-Bessel)");
	}
	if (!valid_brick_shape(extent_x, extent_y, extent_z))
	{
		fail(REXLIB_HIP_INVALID_ARGUMENT, "insertion tuning: brick extents (x, y, z) are (16, 8, 6), (16, 8, 8), "
			"(16, 16, 8), (8, 8, 8), (8, 8, 6), (8, 8, 4) or (8, 6, 6)");
	}
	if (share < 16 || share > (1 << 20))
	{
		fail(REXLIB_HIP_INVALID_ARGUMENT, "insertion tuning: plane hits per task lie in [16, 2^20]");
	}
	std::lock_guard<std::mutex> lock(tuning_mutex);
	tuning_table[extent + taps * taps_key] = brick_tuning{extent_x, extent_y, extent_z, share};
}

void get_insertion_tuning(int extent, int *extent_x, int *extent_y, int *extent_z, int *share)
{
	const brick_tuning t = tuning_for(extent);
	*extent_x = t.ex;
	*extent_y = t.ey;
	*extent_z = t.ez;
	*share = t.share;
}

void clear_insertion_tuning()
{
	std::lock_guard<std::mutex> lock(tuning_mutex);
	tuning_table.clear();
}

// The row transforms of images [first, last) of a pending image insertion.
void launch_rows(const pending_columns &columns, int64_t first, int64_t last, hipStream_t stream)
{
	const pending_rows &rows = *columns.rows;
	if (last <= first)
	{
		return;
	}
	line_set in = rows.in_rows;
	in.count = (last - first) * rows.m;
	line_set out = rows.out_rows;
	out.count = (last - first) * rows.n;
	const char *images = static_cast<const char *>(rows.images) +
		first * rows.in_rows.s_outer * static_cast<int64_t>(rows.element_bytes);
	const float *scales = rows.scales != nullptr ? rows.scales + first * rows.scale_stride : nullptr;
	launch_rfft_rows(rows.kind, images, in, rows.m, scales, rows.scale_stride,
		columns.spectra + first * rows.out_rows.s_outer, out, rows.n, stream);
}

// Host entry of every insertion: fills the workspace and launches.
void run_insertion(insertion_params p, int schedule, const float2 *num_in, const int64_t *num_in_strides,
	const float *den_in, const int64_t *den_in_strides, bool in_place, double shape, const pending_columns *columns,
	workspace_source &workspace, hipStream_t stream)
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
		workspace.take(area_insertion, entry_bytes + table_bytes + kb_bytes + phase_bytes, stream));
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
	// Prepared samples take 12 bytes per stored element; the slices of one
	// call go through in chunks that keep them within 4 GiB, and with a fixed
	// workspace buffer within what it has left beside the task list. The
	// brick schedule needs the window to span at most 4 voxels; wider
	// Kaiser-Bessel windows take the scatter schedule, and so do calls whose
	// workspace buffer cannot hold one slice's prepared samples.
	const int64_t per_slice = static_cast<int64_t>(p.n) * p.h * static_cast<int64_t>(sizeof(prepared_sample)) *
		p.map.kept_count;
	const int64_t limit = int64_t(4) << 30;
	// Taps per axis of the brick layout: 1, 2, or 4 for every Kaiser-Bessel window.
	const int layout_taps = p.profile == REXLIB_HIP_BOX ? 1 : (p.profile == REXLIB_HIP_TRIANGLE ? 2 : 4);
	const brick_tuning tuning = tuning_for(p.big, layout_taps);
	int64_t chunk_count = 0;
	if (schedule != schedule_scatter && p.taps <= 4 && p.map.kept_count <= 65535)
	{
		const auto room = [&](std::size_t reserved)
		{
			const std::size_t left = workspace.remaining();
			return left <= reserved + 512 ? int64_t(0) :
				std::min<int64_t>(limit, static_cast<int64_t>(std::min<std::size_t>(left - reserved - 512,
					std::size_t(1) << 40)));
		};
		chunk_count = std::min<int64_t>(p.map.reduced_count, room(0) / per_slice);
		if (chunk_count >= 1 && workspace.arena != nullptr)
		{
			// Tasks made for more slices stay valid for fewer (their shares
			// are fractions of a chunk).
			chunk_count = std::min<int64_t>(chunk_count,
				room(brick_task_bytes(p.big, chunk_count, tuning, layout_taps)) / per_slice);
		}
	}
	const bool brick = chunk_count >= 1;
	// Pending column transforms run inside the preparation of each chunk on
	// the brick schedule; on the scatter schedule they run here, in place.
	const bool fused_columns = columns != nullptr && brick;
	// Pending row transforms run per chunk (just before the chunk's column
	// pass, so a chunk's spectra are still in the caches) when every chunk is
	// a range of the images (one volume), else here for all images.
	const bool chunk_rows = fused_columns && columns->rows != nullptr && p.map.kept_count == 1;
	if (columns != nullptr && columns->rows != nullptr && !chunk_rows)
	{
		launch_rows(*columns, 0, columns->rows->in_rows.count / columns->rows->m, stream);
	}
	if (columns != nullptr && !fused_columns && columns->lines.count > 0)
	{
		launch_fft_lines(columns->spectra, columns->lines, columns->spectra, columns->lines, p.n, -1.0f, 1.0f,
			stream);
	}
	settle_kernel<<<static_cast<unsigned>(std::min<int64_t>(slices, 1 << 20)), 128, 0, stream>>>(p);
	REXLIB_HIP_CHECK(hipGetLastError());

	if (brick)
	{
		int count = 0;
		const int4 *tasks = brick_tasks(p.big, chunk_count, tuning, layout_taps, workspace, stream, count);
		prepared_sample *prepared = static_cast<prepared_sample *>(workspace.take(area_prepared,
			static_cast<std::size_t>(per_slice * chunk_count), stream));
		for (int64_t begin = 0; begin < p.map.reduced_count; begin += chunk_count)
		{
			const slice_chunk chunk{begin, std::min(chunk_count, p.map.reduced_count - begin)};
			if (chunk_rows)
			{
				launch_rows(*columns, chunk.begin, chunk.begin + chunk.count, stream);
			}
			const int prepare_token = timing_begin(REXLIB_HIP_KERNEL_PREPARE, stream);
			if (fused_columns)
			{
				dispatch_columns_prepare(p, chunk, prepared, stream);
			}
			else
			{
				dispatch_prepare(p, chunk, prepared, stream);
			}
			REXLIB_HIP_CHECK(hipGetLastError());
			timing_end(prepare_token, stream);
			const int brick_token = timing_begin(REXLIB_HIP_KERNEL_BRICK, stream);
			dispatch_brick(p, chunk, prepared, tasks, count, tuning, stream);
			REXLIB_HIP_CHECK(hipGetLastError());
			timing_end(brick_token, stream);
		}
	}
	else
	{
		const int scatter_token = timing_begin(REXLIB_HIP_KERNEL_SCATTER, stream);
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
		timing_end(scatter_token, stream);
	}
	REXLIB_HIP_CHECK(hipGetLastError());
}

} // namespace rexlib_hip
