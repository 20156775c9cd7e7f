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
// The brick is a box of signed frequencies (x in [0, N/2], y and z in
// [-N/2, N/2]). The wavefront tests 64 slices at a time for planes that pass
// near the brick; for each such slice it bounds the rows that reach the
// brick and clips each row to the samples that reach it -- bound_rows and
// clip_row of src/backends/cpu/kernels/row_interval.cpp, restated per brick
// as the reference restates them per tile (slice_insertion_functor.inl:391-440).
//
// A voxel takes 12 bytes of LDS (a 16 x 8 x 6 brick with its padding is
// 10 KB), which lets 16 wavefronts share a CU: the adds of a tap wait on LDS
// latency, and more wavefronts hide more of it.
//
// The samples are added by plain LDS read-modify-write, not LDS float
// atomics (which retire at about a third of a lane per clock per CU on
// gfx950, against eight for a read-modify-write). A lane owns one sample per
// instruction, and the samples are taken in four parity classes of
// (k_x mod 2, k_y mod 2): two samples of a class differ by 2 a X + 2 b Y,
// which is at least one voxel along some axis when the frame (X, Y) is close
// enough to orthogonal with steps of at least about 0.87 voxel (every rotation
// with N >= n), so no two lanes of an instruction touch one voxel. Slices
// whose frames fail that test take LDS atomics instead. The wavefront is the
// brick's only writer, so there is no barrier anywhere.
//
// Every plane passes through the origin, so bricks near it see every slice
// and bricks far out see few (about 13 / r of them). A task is a brick and a
// share of the slices: heavy bricks are split, and the heaviest tasks start
// first. A brick that one task owns is added to the accumulators by plain
// read-add-write; split bricks, and bricks holding the y or z frequency
// +-N/2 of an even N (two signed frequencies, one stored element), add with
// atomics.
// ---------------------------------------------------------------------------

constexpr int brick_edge = 16; // along x; the extents along y and z are parameters
// Slots of 64 samples whose loads a wavefront issues together.
constexpr int max_slots = 4;
// LDS pitches of a brick: odd, so that lanes stepping along any axis fall in
// different banks.
constexpr int pitch_y = brick_edge + 1;

__host__ __device__ constexpr int brick_pitch_z(int extent_y)
{
	return pitch_y * extent_y + 1;
}


// A voxel of a brick in LDS: 12 bytes.
struct brick_voxel
{
	float re;
	float im;
	float den;
};

struct brick_box
{
	int lo[3];  // signed (x, y, z) of the first voxel
	int ext[3]; // voxels along (x, y, z)
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

// Box and triangle windows on an even N: per sample, the first tap's brick
// index, its volume phase (-1)^(x+y+z) (a wrapped index keeps its parity)
// and the in-brick bits of the taps; each tap is then a constant offset and a
// constant sign flip. The sums are those of the general form below.
template <int PROFILE, bool ATOMIC, int EY, int EZ>
__device__ __forceinline__ void add_sample_even(const insertion_params &p, const brick_box &box, brick_voxel *acc,
	const float *c, float vr, float vi, float weight)
{
	constexpr int taps = window_taps<PROFILE>::taps;
	constexpr int pitch_z = brick_pitch_z(EY);
	float w[3][taps];
	bool in[3][taps];
	int g[3];
	int local[3];
#pragma unroll
	for (int a = 0; a < 3; ++a)
	{
		if constexpr (PROFILE == REXLIB_HIP_TRIANGLE)
		{
			// floor and fraction, as window_taps<TRIANGLE>::place.
			const float below = floorf(c[a]);
			g[a] = static_cast<int>(below);
			w[a][1] = c[a] - below;
			w[a][0] = 1.0f - w[a][1];
		}
		else
		{
			g[a] = window_taps<PROFILE>::place(p, c[a], w[a]);
		}
		local[a] = g[a] - box.lo[a];
#pragma unroll
		for (int t = 0; t < taps; ++t)
		{
			in[a][t] = static_cast<unsigned>(local[a] + t) < static_cast<unsigned>(box.ext[a]);
		}
	}
	const int base = local[0] + pitch_y * local[1] + pitch_z * local[2];
	const bool odd = ((g[0] + g[1] + g[2]) & 1) != 0;
	const float sr = odd ? -vr : vr;
	const float si = odd ? -vi : vi;
#pragma unroll
	for (int tz = 0; tz < taps; ++tz)
	{
#pragma unroll
		for (int ty = 0; ty < taps; ++ty)
		{
			const float wzy = w[2][tz] * w[1][ty];
			const bool in_zy = in[2][tz] && in[1][ty];
#pragma unroll
			for (int tx = 0; tx < taps; ++tx)
			{
				const float ww = wzy * w[0][tx];
				const bool flip = ((tx + ty + tz) & 1) != 0;
				const float re = (flip ? -sr : sr) * ww;
				const float im = (flip ? -si : si) * ww;
				const int idx = base + tx + pitch_y * ty + pitch_z * tz;
				if (in_zy && in[0][tx])
				{
					if (ATOMIC)
					{
						atomicAdd(&acc[idx].re, re);
						atomicAdd(&acc[idx].im, im);
						atomicAdd(&acc[idx].den, weight * ww);
					}
					else
					{
						brick_voxel v = acc[idx];
						v.re += re;
						v.im += im;
						v.den += weight * ww;
						acc[idx] = v;
					}
				}
				lds_order();
			}
		}
	}
}

// add_sample_even for a slot that holds rows of one k_y parity with every
// k_x, consecutive k_x of a row in consecutive lanes. Two samples of such a
// slot fall in one voxel cell only if they are neighbours in a row (see
// insert_slice_sign), so a lane whose cell is that of the lane before it
// hands its sample to that lane, which adds both in one pass over the taps:
// the read-modify-write of a tap then never has two lanes on one voxel.
// Every lane of the wavefront calls it (inactive ones with active false).
template <int PROFILE, int EY, int EZ>
__device__ __forceinline__ void add_sample_pair(const insertion_params &p, const brick_box &box, brick_voxel *acc,
	const float *c, bool active, float vr, float vi, float weight)
{
	static_assert(PROFILE != REXLIB_HIP_KAISER_BESSEL, "box and triangle windows only");
	constexpr int taps = window_taps<PROFILE>::taps;
	constexpr int pitch_z = brick_pitch_z(EY);
	const int lane = threadIdx.x & 63;
	float w[3][taps];
	bool in[3][taps];
	int g[3];
	int local[3];
#pragma unroll
	for (int a = 0; a < 3; ++a)
	{
		if constexpr (PROFILE == REXLIB_HIP_TRIANGLE)
		{
			const float below = floorf(c[a]);
			g[a] = static_cast<int>(below);
			w[a][1] = c[a] - below;
			w[a][0] = 1.0f - w[a][1];
		}
		else
		{
			g[a] = window_taps<PROFILE>::place(p, c[a], w[a]);
		}
		local[a] = g[a] - box.lo[a];
#pragma unroll
		for (int t = 0; t < taps; ++t)
		{
			in[a][t] = static_cast<unsigned>(local[a] + t) < static_cast<unsigned>(box.ext[a]);
		}
	}
	// The cell of an active sample, which lies within a few voxels of the
	// brick, in 10 bits per axis (never 0); inactive lanes hold keys no other
	// lane has. The lanes at the ends read 0 for their missing neighbour.
	const int key = active ? (((local[0] + 512) & 1023) | (((local[1] + 512) & 1023) << 10) |
		(((local[2] + 512) & 1023) << 20)) : ((1 << 30) | lane);
	const bool merge = next_lane(key) == key;
	const bool follower = previous_lane(key) == key;
	// The next lane's sample (zero unless this lane takes it) and weights.
	const float vr_next = next_lane(vr);
	const float vi_next = next_lane(vi);
	const float weight_next = next_lane(weight);
	float w_next[3][taps];
#pragma unroll
	for (int a = 0; a < 3; ++a)
	{
		if constexpr (PROFILE == REXLIB_HIP_TRIANGLE)
		{
			w_next[a][1] = next_lane(w[a][1]);
			w_next[a][0] = 1.0f - w_next[a][1];
		}
		else
		{
			w_next[a][0] = 1.0f;
		}
	}
	const float nr = merge ? vr_next : 0.0f;
	const float ni = merge ? vi_next : 0.0f;
	const float nd = merge ? weight_next : 0.0f;
	// A follower's sample is added by the lane before it.
#pragma unroll
	for (int t = 0; t < taps; ++t)
	{
		in[2][t] = in[2][t] && !follower;
	}
	const int base = local[0] + pitch_y * local[1] + pitch_z * local[2];
	const bool odd = ((g[0] + g[1] + g[2]) & 1) != 0;
	const float sr = odd ? -vr : vr;
	const float si = odd ? -vi : vi;
	const float tr = odd ? -nr : nr;
	const float ti = odd ? -ni : ni;
#pragma unroll
	for (int tz = 0; tz < taps; ++tz)
	{
#pragma unroll
		for (int ty = 0; ty < taps; ++ty)
		{
			const float wzy = w[2][tz] * w[1][ty];
			const float wzy_next = w_next[2][tz] * w_next[1][ty];
			const bool in_zy = in[2][tz] && in[1][ty];
#pragma unroll
			for (int tx = 0; tx < taps; ++tx)
			{
				const float ww = wzy * w[0][tx];
				const float ww_next = wzy_next * w_next[0][tx];
				const bool flip = ((tx + ty + tz) & 1) != 0;
				const int idx = base + tx + pitch_y * ty + pitch_z * tz;
				if (in_zy && in[0][tx])
				{
					// Two fused multiply-adds per component: the partner's term,
					// then this lane's.
					brick_voxel v = acc[idx];
					v.re = fmaf(flip ? -sr : sr, ww, fmaf(flip ? -tr : tr, ww_next, v.re));
					v.im = fmaf(flip ? -si : si, ww, fmaf(flip ? -ti : ti, ww_next, v.im));
					v.den = fmaf(weight, ww, fmaf(nd, ww_next, v.den));
					acc[idx] = v;
				}
				lds_order();
			}
		}
	}
}

template <int PROFILE, bool ODD, bool ATOMIC, int EY, int EZ>
__device__ inline void add_sample(const insertion_params &p, const brick_box &box, brick_voxel *acc, const float *c,
	float vr, float vi, float weight)
{
	if constexpr (!ODD && PROFILE != REXLIB_HIP_KAISER_BESSEL)
	{
		add_sample_even<PROFILE, ATOMIC, EY, EZ>(p, box, acc, c, vr, vi, weight);
		return;
	}
	constexpr int pitch_z = brick_pitch_z(EY);
	float wx[window_taps<PROFILE>::taps];
	float wy[window_taps<PROFILE>::taps];
	float wz[window_taps<PROFILE>::taps];
	const int gx0 = window_taps<PROFILE>::place(p, c[0], wx);
	const int gy0 = window_taps<PROFILE>::place(p, c[1], wy);
	const int gz0 = window_taps<PROFILE>::place(p, c[2], wz);
	const int taps = PROFILE == REXLIB_HIP_KAISER_BESSEL ? p.taps : window_taps<PROFILE>::taps;
	const int big = p.big;
	for (int tz = 0; tz < taps; ++tz)
	{
		const int lz = gz0 + tz - box.lo[2];
		const bool in_z = static_cast<unsigned>(lz) < static_cast<unsigned>(box.ext[2]);
		const int gz = gz0 + tz;
		const int sz = gz < 0 ? gz + big : gz;
		for (int ty = 0; ty < taps; ++ty)
		{
			const int ly = gy0 + ty - box.lo[1];
			const bool in_zy = in_z && static_cast<unsigned>(ly) < static_cast<unsigned>(box.ext[1]);
			const int gy = gy0 + ty;
			const int sy = gy < 0 ? gy + big : gy;
			const float wzy = wz[tz] * wy[ty];
			for (int tx = 0; tx < taps; ++tx)
			{
				const int lx = gx0 + tx - box.lo[0];
				const bool inside = in_zy && static_cast<unsigned>(lx) < static_cast<unsigned>(box.ext[0]);
				const int gx = gx0 + tx;
				const float w = wzy * wx[tx];
				const float2 contribution = volume_phase<ODD>(p, make_float2(vr * w, vi * w), gx, sy, sz);
				const int idx = lx + pitch_y * ly + pitch_z * lz;
				if (inside)
				{
					if (ATOMIC)
					{
						atomicAdd(&acc[idx].re, contribution.x);
						atomicAdd(&acc[idx].im, contribution.y);
						atomicAdd(&acc[idx].den, weight * w);
					}
					else
					{
						brick_voxel v = acc[idx];
						v.re += contribution.x;
						v.im += contribution.y;
						v.den += weight * w;
						acc[idx] = v;
					}
				}
				lds_order();
			}
		}
	}
}

// bound_rows of one slice and sign: the rows [r0, r1) whose samples k_x in
// [first, widest] can reach the box [lower, upper]; r0 >= r1 when none.
// Evaluated per lane, one slice per lane.
// `inv_y` holds 1 / y_step_f per axis (0 where the step is 0); a product with
// it stands for the division, the box being widened far beyond its rounding.
__device__ inline void bound_rows(const insertion_params &p, const float *x_step, const float *y_step_f,
	const float *inv_y, int last_shell, const float *lower, const float *upper, int sign_index, int &r0, int &r1)
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
	// The box is already widened past the window, so rounding inwards keeps
	// every sample that can reach it.
	r0 = max(static_cast<int>(ceilf(low)), first_row);
	r1 = min(static_cast<int>(floorf(high)) + 1, last_row + 1);
}

// Position of sample k_x of row k_y along one axis, the sign applied to the
// frame: the row base k_y * Y formed in double and rounded to float, plus
// k_x * X in float, uncontracted (row_insertion_kernel.inl:157-166). Lanes
// without a sample sit far outside every brick.
__device__ inline float row_position(bool active, double y_step, float x_step, int kx, int ky)
{
#pragma clang fp contract(off)
	const float base = static_cast<float>(static_cast<double>(ky) * y_step);
	const float at = base + static_cast<float>(kx) * x_step;
	return active ? at : -1.0e6f;
}

// One slice, one sign, into a brick, over the rows [r0, r1) that reach it;
// the whole wavefront takes part.
// `inv_x` holds 1 / x_step per axis (0 where the step is 0); `y_step_d` the
// row step in double, which the positions use (y_step_f, its float, only
// bounds the rows).
template <int PROFILE, bool ODD, bool ATOMIC, int EY, int EZ>
__device__ void insert_slice_sign(const insertion_params &p, const float *x_step, const float *y_step_f,
	const double *y_step_d, const float *inv_x, int last_shell, const prepared_sample *slice, const brick_box &box,
	const float *lower, const float *upper, int sign_index, int r0, int r1, brick_voxel *acc, int *row_table)
{
	const int lane = threadIdx.x & 63;
	const int n = p.n;
	const int h = p.h;
	const bool even = n % 2 == 0;
	const int last_column = h - (even ? 2 : 1);
	const float radius = static_cast<float>(last_shell) + 0.5f;
	const int widest = min(last_column, last_shell);
	const float sign = sign_index == 0 ? 1.0f : -1.0f;
	const int first = sign_index;
	float step[3];
	float y_step[3];
	double y_exact[3];
	float inv_step[3];
	for (int a = 0; a < 3; ++a)
	{
		step[a] = sign * x_step[a];
		y_step[a] = sign * y_step_f[a];
		y_exact[a] = sign_index == 0 ? y_step_d[a] : -y_step_d[a];
		inv_step[a] = sign * inv_x[a];
	}
	for (int rc = r0; rc < r1; rc += 64)
	{
		// clip_row: the samples of row k_y that can reach the box.
		const int ky = rc + lane;
		int begin = 0;
		int end = 0;
		if (ky < r1)
		{
			const float kyf = static_cast<float>(ky);
			const float room = radius * radius - kyf * kyf;
			if (room >= 0.0f)
			{
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
						// Reciprocal steps: the box is widened by 0.01 voxel,
						// far beyond this rounding.
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
		}
		// Box and triangle windows on an even N, lattice-free slices: the two
		// classes of one k_y parity share a slot whenever both fit. Samples of
		// one slot then differ by a X + 2 b Y; unless (a, b) = (+-1, 0) that
		// is at least sqrt(4 * 0.76) > sqrt(3) voxels (the bound behind
		// lattice_free), more than a voxel cell spans, so only neighbours in a
		// row can share a cell, and add_sample_pair merges those.
		if constexpr (!ODD && PROFILE != REXLIB_HIP_KAISER_BESSEL && !ATOMIC)
		{
			// Per row: every k_x of [begin, end), counted for its k_y parity.
			const int row_count = max(end - begin, 0);
			const int pair_count[2] = {(ky & 1) == 0 ? row_count : 0, (ky & 1) != 0 ? row_count : 0};
			int pair_start[2];
			int pair_total[2];
			scan_two(pair_count[0], pair_count[1], pair_start[0], pair_start[1], pair_total[0], pair_total[1]);
			if (pair_total[0] <= 64 && pair_total[1] <= 64)
			{
				prepared_sample q[2];
				int sample_x[2];
				int sample_y[2];
				bool valid[2];
#pragma unroll
				for (int c = 0; c < 2; ++c)
				{
					q[c] = prepared_sample{0.0f, 0.0f, 0.0f};
					sample_x[c] = 0;
					sample_y[c] = 0;
					valid[c] = false;
					if (pair_total[c] == 0)
					{
						continue;
					}
					row_table[lane] = 0;
					lds_order();
					if (pair_count[c] > 0)
					{
						atomicMax(&row_table[pair_start[c] & 63], lane);
					}
					lds_order();
					const int r = wave_prefix_max(row_table[lane]) & 63;
					lds_order();
					const int row_begin = __shfl(begin, r, 64);
					const int row_start = __shfl(pair_start[c], r, 64);
					const int sky = rc + r;
					const int kx = row_begin + (lane - row_start);
					const int row = sky < 0 ? sky + n : sky;
					sample_x[c] = kx;
					sample_y[c] = sky;
					valid[c] = lane < pair_total[c];
					q[c] = slice[valid[c] ? min(max(row * h + kx, 0), n * h - 1) : 0];
				}
#pragma unroll
				for (int c = 0; c < 2; ++c)
				{
					if (pair_total[c] == 0)
					{
						continue;
					}
					const bool active = valid[c] && !(q[c].re == 0.0f && q[c].im == 0.0f && q[c].den == 0.0f);
					float position[3];
					for (int a = 0; a < 3; ++a)
					{
						position[a] = row_position(active, y_exact[a], step[a], sample_x[c], sample_y[c]);
					}
					add_sample_pair<PROFILE, EY, EZ>(p, box, acc, position, active, q[c].re,
						sign_index == 0 ? q[c].im : -q[c].im, q[c].den);
				}
				continue;
			}
		}
		// Four parity classes; a class holds rows of one parity of k_y and the
		// k_x of one parity within them. Each class is packed densely over the
		// lanes (a scan over the rows), in slots of 64 samples.
		int class_begin[4];
		int class_start[4];
		int class_total[4];
		int slot_end[4]; // running count of slots through each class
		int count[4];
		for (int cls = 0; cls < 4; ++cls)
		{
			const int py = cls >> 1;
			const int px = cls & 1;
			class_begin[cls] = begin + ((begin - px) & 1); // first k_x >= begin of parity px
			count[cls] = 0;
			if (((ky - py) & 1) == 0 && class_begin[cls] < end)
			{
				count[cls] = (end - class_begin[cls] + 1) / 2;
			}
		}
		scan_two(count[0], count[1], class_start[0], class_start[1], class_total[0], class_total[1]);
		scan_two(count[2], count[3], class_start[2], class_start[3], class_total[2], class_total[3]);
		int slots = 0;
		for (int cls = 0; cls < 4; ++cls)
		{
			slots += (class_total[cls] + 63) >> 6;
			slot_end[cls] = slots;
		}
		// The common case, every class within one slot: class c is slot c,
		// and the row of each sample comes from the wave's row table (each
		// non-empty row records its lane at its first sample, then a prefix
		// maximum over the lanes) rather than from a search. Every load index
		// is checked against the slice before use.
		if (class_total[0] <= 64 && class_total[1] <= 64 && class_total[2] <= 64 && class_total[3] <= 64)
		{
			prepared_sample q[4];
			int sample_x[4];
			int sample_y[4];
			bool valid[4];
#pragma unroll
			for (int c = 0; c < 4; ++c)
			{
				q[c] = prepared_sample{0.0f, 0.0f, 0.0f};
				sample_x[c] = 0;
				sample_y[c] = 0;
				valid[c] = false;
				if (class_total[c] == 0)
				{
					continue;
				}
				// The first sample of a class always starts a row, so lane 0
				// holds a row index of at least 0 and the prefix maximum
				// gives every sample lane the row it falls in.
				row_table[lane] = 0;
				lds_order();
				if (count[c] > 0)
				{
					atomicMax(&row_table[class_start[c] & 63], lane);
				}
				lds_order();
				const int r = wave_prefix_max(row_table[lane]) & 63;
				lds_order();
				const int row_begin = __shfl(class_begin[c], r, 64);
				const int row_start = __shfl(class_start[c], r, 64);
				const int sky = rc + r;
				const int kx = row_begin + 2 * (lane - row_start);
				const int row = sky < 0 ? sky + n : sky;
				sample_x[c] = kx;
				sample_y[c] = sky;
				valid[c] = lane < class_total[c];
				// Clamped into the slice whatever the lane holds.
				q[c] = slice[valid[c] ? min(max(row * h + kx, 0), n * h - 1) : 0];
			}
#pragma unroll
			for (int c = 0; c < 4; ++c)
			{
				if (class_total[c] == 0)
				{
					continue;
				}
				// Lanes without a sample, and zero samples (not inserted), stay
				// in step with the other lanes and add nothing.
				const bool active = valid[c] && !(q[c].re == 0.0f && q[c].im == 0.0f && q[c].den == 0.0f);
				float position[3];
				for (int a = 0; a < 3; ++a)
				{
					position[a] = row_position(active, y_exact[a], step[a], sample_x[c], sample_y[c]);
				}
				add_sample<PROFILE, ODD, ATOMIC, EY, EZ>(p, box, acc, position, q[c].re,
					sign_index == 0 ? q[c].im : -q[c].im, q[c].den);
			}
			continue;
		}
		// Gather, then scatter: the samples of up to max_slots slots are
		// located and loaded first, so that their loads overlap, and added
		// afterwards one slot at a time (one class per instruction).
		for (int group = 0; group < slots; group += max_slots)
		{
			prepared_sample q[max_slots];
			int sample_x[max_slots];
			int sample_y[max_slots];
#pragma unroll
			for (int k = 0; k < max_slots; ++k)
			{
				q[k] = prepared_sample{0.0f, 0.0f, 0.0f};
				sample_x[k] = 0;
				sample_y[k] = 0;
				const int g = group + k;
				if (g >= slots)
				{
					continue;
				}
				const int cls = g < slot_end[0] ? 0 : (g < slot_end[1] ? 1 : (g < slot_end[2] ? 2 : 3));
				const int first_slot = cls == 0 ? 0 : slot_end[cls - 1];
				const int start = class_start[cls];
				const int i = ((g - first_slot) << 6) + lane;
				// The row holding sample i: the last lane whose start is at or below i.
				int lo_r = 0;
				int hi_r = 63;
				for (int step_r = 0; step_r < 6; ++step_r)
				{
					const int mid = (lo_r + hi_r + 1) >> 1;
					const int at = __shfl(start, mid, 64);
					if (at <= i)
					{
						lo_r = mid;
					}
					else
					{
						hi_r = mid - 1;
					}
				}
				const int row_begin = __shfl(class_begin[cls], lo_r, 64);
				const int row_start = __shfl(start, lo_r, 64);
				const int sky = rc + lo_r;
				const int kx = row_begin + 2 * (i - row_start);
				sample_x[k] = kx;
				sample_y[k] = sky;
				if (i < class_total[cls])
				{
					const int row = sky < 0 ? sky + n : sky;
					q[k] = slice[row * h + kx];
				}
			}
#pragma unroll
			for (int k = 0; k < max_slots; ++k)
			{
				if (group + k >= slots)
				{
					continue;
				}
				// Zero samples (not inserted) stay in step with the other lanes.
				const bool active = !(q[k].re == 0.0f && q[k].im == 0.0f && q[k].den == 0.0f);
				float c[3];
				for (int a = 0; a < 3; ++a)
				{
					c[a] = row_position(active, y_exact[a], step[a], sample_x[k], sample_y[k]);
				}
				add_sample<PROFILE, ODD, ATOMIC, EY, EZ>(p, box, acc, c, q[k].re, sign_index == 0 ? q[k].im : -q[k].im,
					q[k].den);
			}
		}
	}
}

template <int PROFILE, bool ODD, int EY, int EZ>
__global__ __launch_bounds__(64) void brick_kernel(insertion_params p, slice_chunk chunk, const prepared_sample *prepared,
	const int4 *tasks, int task_count, float reach)
{
	constexpr int pitch_z = brick_pitch_z(EY);
	constexpr int voxels_max = pitch_z * EZ;
	__shared__ brick_voxel acc[voxels_max];
	// Per sample slot: the row of each sample (see insert_slice_sign).
	__shared__ int row_table[64];

	const int lane = threadIdx.x;
	const int big = p.big;
	const int64_t elements = static_cast<int64_t>(p.n) * p.h;
	const int64_t kept = blockIdx.y;

	for (int ti = blockIdx.x; ti < task_count; ti += gridDim.x)
	{
		const brick_box box = unpack_brick(tasks[2 * ti]);
		const int4 share = tasks[2 * ti + 1];
		const int64_t share_begin = chunk.count * share.x / share.y;
		const int64_t share_end = chunk.count * (share.x + 1) / share.y;
		const int voxels = box.ext[0] * box.ext[1] * box.ext[2];
		for (int v = lane; v < voxels_max; v += 64)
		{
			acc[v] = brick_voxel{0.0f, 0.0f, 0.0f};
		}
		lds_order();
		// The box of positions that can reach the brick, widened beyond the
		// window so that rounding never drops a sample.
		float lower[3];
		float upper[3];
		float centre[3];
		float half_extent[3];
		float half_diagonal2 = 0.0f;
		for (int a = 0; a < 3; ++a)
		{
			lower[a] = static_cast<float>(box.lo[a]) - reach - 0.01f;
			upper[a] = static_cast<float>(box.lo[a] + box.ext[a] - 1) + reach + 0.01f;
			centre[a] = static_cast<float>(box.lo[a]) + 0.5f * static_cast<float>(box.ext[a] - 1);
			half_extent[a] = 0.5f * static_cast<float>(box.ext[a] - 1) + reach + 0.02f;
			half_diagonal2 += half_extent[a] * half_extent[a];
		}
		const float half_diagonal = sqrtf(half_diagonal2);
		const float centre_norm = sqrtf(centre[0] * centre[0] + centre[1] * centre[1] + centre[2] * centre[2]);

		for (int64_t window = share_begin; window < share_end; window += 64)
		{
			// Slices whose plane passes within reach of the brick, one per lane.
			const int64_t j = window + lane;
			bool near = false;
			// The frame of this lane's slice, kept for the insertions below.
			float x0 = 0.0f, x1 = 0.0f, x2 = 0.0f, y0 = 0.0f, y1 = 0.0f, y2 = 0.0f;
			int last_shell = -1;
			int lattice_free = 0;
			if (j < share_end)
			{
				const slice_entry &e = p.entries[kept * p.map.reduced_count + chunk.begin + j];
				last_shell = e.last_shell;
				lattice_free = e.lattice_free;
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
				bound_rows(p, xs, ys, iy, last_shell, lower, upper, 0, r0_plus, r1_plus);
				bound_rows(p, xs, ys, iy, last_shell, lower, upper, 1, r0_minus, r1_minus);
				near = r0_plus < r1_plus || r0_minus < r1_minus;
			}
			unsigned long long mask = __ballot(near);
			while (mask != 0)
			{
				const int bit = __ffsll(static_cast<long long>(mask)) - 1;
				mask &= mask - 1;
				const int64_t js = window + bit;
				const prepared_sample *slice = prepared + (kept * chunk.count + js) * elements;
				// The hit slice's frame, read from the lane that tested it.
				const float x_step[3] = {lane_value(x0, bit), lane_value(x1, bit), lane_value(x2, bit)};
				const float y_step[3] = {lane_value(y0, bit), lane_value(y1, bit), lane_value(y2, bit)};
				const float inv_x[3] = {lane_value(ix0, bit), lane_value(ix1, bit), lane_value(ix2, bit)};
				// The row step in double, from the slice's entry (uniform).
				const slice_entry &hit = p.entries[kept * p.map.reduced_count + chunk.begin + js];
				const double y_step_d[3] = {hit.y_step[0], hit.y_step[1], hit.y_step[2]};
				const int shell = __builtin_amdgcn_readlane(last_shell, bit);
				const bool exact = __builtin_amdgcn_readlane(lattice_free, bit) != 0;
				for (int sign_index = 0; sign_index < 2; ++sign_index)
				{
					const int r0 = __builtin_amdgcn_readlane(sign_index == 0 ? r0_plus : r0_minus, bit);
					const int r1 = __builtin_amdgcn_readlane(sign_index == 0 ? r1_plus : r1_minus, bit);
					if (r0 >= r1)
					{
						continue;
					}
					if (exact)
					{
						insert_slice_sign<PROFILE, ODD, false, EY, EZ>(p, x_step, y_step, y_step_d, inv_x, shell, slice, box,
							lower, upper, sign_index, r0, r1, acc, row_table);
					}
					else
					{
						insert_slice_sign<PROFILE, ODD, true, EY, EZ>(p, x_step, y_step, y_step_d, inv_x, shell, slice, box,
							lower, upper, sign_index, r0, r1, acc, row_table);
					}
				}
			}
		}
		lds_order();
		// Add the brick to the accumulators of its kept batch element.
		const bool exclusive = box.alias_free && share.y == 1;
		const int64_t num_base = kept_offset(p.map, slot_numerator_out, kept);
		const int64_t den_base = kept_offset(p.map, slot_denominator_out, kept);
		for (int v = lane; v < voxels; v += 64)
		{
			const int lx = v % box.ext[0];
			const int rest = v / box.ext[0];
			const int ly = rest % box.ext[1];
			const int lz = rest / box.ext[1];
			const int idx = lx + pitch_y * ly + pitch_z * lz;
			const brick_voxel sum = acc[idx];
			const float re = sum.re;
			const float im = sum.im;
			const float d = sum.den;
			if (re == 0.0f && im == 0.0f && d == 0.0f)
			{
				continue;
			}
			const int gx = box.lo[0] + lx;
			const int gy = box.lo[1] + ly;
			const int gz = box.lo[2] + lz;
			const int sy = gy < 0 ? gy + big : gy;
			const int sz = gz < 0 ? gz + big : gz;
			float2 *target = p.numerator + num_base + sz * p.num_s0 + sy * p.num_s1 + gx * p.num_s2;
			float *dtarget = p.denominator + den_base + sz * p.den_s0 + sy * p.den_s1 + gx * p.den_s2;
			if (exclusive)
			{
				float2 t = *target;
				t.x += re;
				t.y += im;
				*target = t;
				*dtarget += d;
			}
			else
			{
				atomicAdd(&target->x, re);
				atomicAdd(&target->y, im);
				atomicAdd(dtarget, d);
			}
		}
		lds_order();
	}
}

// Brick shapes: 16 x 8 x 6 (the default: 10 KB of LDS, 16 wavefronts per
// CU), 16 x 8 x 8, 16 x 8 x 16, 16 x 16 x 8 and 16 x 16 x 16 (x is 16), and
// the expected plane hits per task that splits busy bricks. One entry per
// volume extent N (0: every extent without an entry of its own); a call
// reads its extent's entry once, so concurrent callers each see a whole entry.
struct brick_tuning
{
	int ey;
	int ez;
	int share;
};

constexpr brick_tuning default_brick_tuning{8, 6, 256};

std::mutex tuning_mutex;
std::map<int, brick_tuning> tuning_table;

bool valid_brick_shape(int ey, int ez)
{
	return (ey == 8 && (ez == 6 || ez == 8 || ez == 16)) || (ey == 16 && (ez == 8 || ez == 16));
}

brick_tuning tuning_for(int extent)
{
	std::lock_guard<std::mutex> lock(tuning_mutex);
	auto found = tuning_table.find(extent);
	if (found == tuning_table.end())
	{
		found = tuning_table.find(0);
	}
	return found != tuning_table.end() ? found->second : default_brick_tuning;
}

template <int PROFILE, int EY, int EZ>
void launch_brick_shape(const insertion_params &p, slice_chunk chunk, const prepared_sample *prepared, const int4 *tasks,
	int task_count, float reach, hipStream_t stream)
{
	const dim3 grid(static_cast<unsigned>(task_count), static_cast<unsigned>(p.map.kept_count));
	if (p.big % 2 == 1)
	{
		brick_kernel<PROFILE, true, EY, EZ><<<grid, 64, 0, stream>>>(p, chunk, prepared, tasks, task_count, reach);
	}
	else
	{
		brick_kernel<PROFILE, false, EY, EZ><<<grid, 64, 0, stream>>>(p, chunk, prepared, tasks, task_count, reach);
	}
}

template <int PROFILE>
void launch_brick(const insertion_params &p, slice_chunk chunk, const prepared_sample *prepared, const int4 *tasks,
	int task_count, float reach, const brick_tuning &tuning, hipStream_t stream)
{
	const int ey = tuning.ey;
	const int ez = tuning.ez;
	if (ey == 8 && ez == 6)
	{
		launch_brick_shape<PROFILE, 8, 6>(p, chunk, prepared, tasks, task_count, reach, stream);
	}
	else if (ey == 8 && ez == 8)
	{
		launch_brick_shape<PROFILE, 8, 8>(p, chunk, prepared, tasks, task_count, reach, stream);
	}
	else if (ey == 8)
	{
		launch_brick_shape<PROFILE, 8, 16>(p, chunk, prepared, tasks, task_count, reach, stream);
	}
	else if (ez == 8)
	{
		launch_brick_shape<PROFILE, 16, 8>(p, chunk, prepared, tasks, task_count, reach, stream);
	}
	else
	{
		launch_brick_shape<PROFILE, 16, 16>(p, chunk, prepared, tasks, task_count, reach, stream);
	}
}

void dispatch_brick(const insertion_params &p, slice_chunk chunk, const prepared_sample *prepared, const int4 *tasks,
	int task_count, float reach, const brick_tuning &tuning, hipStream_t stream)
{
	switch (p.profile)
	{
	case REXLIB_HIP_BOX:
		launch_brick<REXLIB_HIP_BOX>(p, chunk, prepared, tasks, task_count, reach, tuning, stream);
		break;
	case REXLIB_HIP_TRIANGLE:
		launch_brick<REXLIB_HIP_TRIANGLE>(p, chunk, prepared, tasks, task_count, reach, tuning, stream);
		break;
	default:
		launch_brick<REXLIB_HIP_KAISER_BESSEL>(p, chunk, prepared, tasks, task_count, reach, tuning, stream);
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
// heaviest first, for slice counts up to `bucket`.
std::vector<int4> brick_task_items(int big, int64_t bucket, int target, int ey, int ez)
{
	const int half = big / 2;
	// Signed ranges: x in [0, N/2]; y, z in [-(N/2), N/2] (even N: N + 1
	// values whose two ends are one stored element; odd N: N values).
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
	std::vector<item> items;
	for (const auto &z : starts(-half, half, ez))
	{
		for (const auto &y : starts(-half, half, ey))
		{
			for (const auto &x : starts(0, half, brick_edge))
			{
				const bool alias = big % 2 == 0 &&
					(z.first == -half || z.first + z.second - 1 == half || y.first == -half ||
						y.first + y.second - 1 == half);
				const int w = x.second | (y.second << 8) | (z.second << 16) | ((alias ? 0 : 1) << 24);
				const double cx = x.first + 0.5 * (x.second - 1);
				const double cy = y.first + 0.5 * (y.second - 1);
				const double cz = z.first + 0.5 * (z.second - 1);
				const double r = std::sqrt(cx * cx + cy * cy + cz * cz);
				// A plane through the origin passes a brick at distance r with
				// probability about min(1, 13 / r).
				const double hits = static_cast<double>(bucket) * std::min(1.0, 13.0 / std::max(r, 1.0));
				const int shares = static_cast<int>(std::min<double>(64.0, std::max(1.0, std::ceil(hits / target))));
				for (int c = 0; c < shares; ++c)
				{
					items.push_back({make_int4(x.first, y.first, z.first, w), make_int4(c, shares, 0, 0),
						hits / shares});
				}
			}
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
std::map<std::tuple<int, int, int64_t, int, int, int>, task_list> task_cache;

// The cached list for a call (made on first use); the caller holds task_mutex.
task_list &brick_task_list(int big, int64_t slices, const brick_tuning &tuning)
{
	int device = 0;
	REXLIB_HIP_CHECK(hipGetDevice(&device));
	int64_t bucket = 1;
	while (bucket < slices)
	{
		bucket *= 2;
	}
	const auto key = std::make_tuple(device, big, bucket, tuning.share, tuning.ey, tuning.ez);
	auto found = task_cache.find(key);
	if (found == task_cache.end())
	{
		found = task_cache.emplace(key, task_list{}).first;
		found->second.host = brick_task_items(big, bucket, tuning.share, tuning.ey, tuning.ez);
	}
	return found->second;
}

// Workspace bytes the task list of a call takes (with its alignment).
std::size_t brick_task_bytes(int big, int64_t slices, const brick_tuning &tuning)
{
	std::lock_guard<std::mutex> lock(task_mutex);
	return align_up(sizeof(int4) * brick_task_list(big, slices, tuning).host.size(), 256) + 256;
}

// The device copy of a call's tasks. With a fixed workspace buffer the list
// is copied into it from the host copy the cache keeps (nothing allocated);
// otherwise the device copy is made once and cached.
const int4 *brick_tasks(int big, int64_t slices, const brick_tuning &tuning, workspace_source &workspace,
	hipStream_t stream, int &count)
{
	std::lock_guard<std::mutex> lock(task_mutex);
	task_list &list = brick_task_list(big, slices, tuning);
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

void set_insertion_tuning(int extent, int extent_y, int extent_z, int share)
{
	if (extent < 0)
	{
		fail(REXLIB_HIP_INVALID_ARGUMENT, "insertion tuning: the extent must be 0 (every extent) or positive");
	}
	if (!valid_brick_shape(extent_y, extent_z))
	{
		fail(REXLIB_HIP_INVALID_ARGUMENT, "insertion tuning: brick extents (y, z) are (8, 6), (8, 8), (8, 16), "
			"(16, 8) or (16, 16)");
	}
	if (share < 16 || share > (1 << 20))
	{
		fail(REXLIB_HIP_INVALID_ARGUMENT, "insertion tuning: plane hits per task lie in [16, 2^20]");
	}
	std::lock_guard<std::mutex> lock(tuning_mutex);
	tuning_table[extent] = brick_tuning{extent_y, extent_z, share};
}

void get_insertion_tuning(int extent, int *extent_y, int *extent_z, int *share)
{
	const brick_tuning t = tuning_for(extent);
	*extent_y = t.ey;
	*extent_z = t.ez;
	*share = t.share;
}

void clear_insertion_tuning()
{
	std::lock_guard<std::mutex> lock(tuning_mutex);
	tuning_table.clear();
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
	const brick_tuning tuning = tuning_for(p.big);
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
				room(brick_task_bytes(p.big, chunk_count, tuning)) / per_slice);
		}
	}
	const bool brick = chunk_count >= 1;
	// Pending column transforms run inside the preparation of each chunk on
	// the brick schedule; on the scatter schedule they run here, in place.
	const bool fused_columns = columns != nullptr && brick;
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
		const int4 *tasks = brick_tasks(p.big, chunk_count, tuning, workspace, stream, count);
		prepared_sample *prepared = static_cast<prepared_sample *>(workspace.take(area_prepared,
			static_cast<std::size_t>(per_slice * chunk_count), stream));
		const float reach = p.profile == REXLIB_HIP_BOX ? 0.5f : (p.profile == REXLIB_HIP_TRIANGLE ? 1.0f :
			p.kb_half_width);
		for (int64_t begin = 0; begin < p.map.reduced_count; begin += chunk_count)
		{
			const slice_chunk chunk{begin, std::min(chunk_count, p.map.reduced_count - begin)};
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
			dispatch_brick(p, chunk, prepared, tasks, count, reach, tuning, stream);
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
