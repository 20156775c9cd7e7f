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
#include "bins.hpp"
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
// Brick schedule: the sample producers. Each computes the prepared sample of
// every stored element of the chunk's slices -- v * r * (numerator weight)
// and r * (denominator weight) (row_insertion_kernel.inl:180-279) -- and
// writes both of its copies into their bricks' bins (bins.hpp); bins.hip
// counts the bins beforehand and adds them to the accumulators afterwards.
// ---------------------------------------------------------------------------

template <int VARIANT>
__global__ __launch_bounds__(256) void prepare_kernel(insertion_params p, slice_chunk chunk, bin_sink sink)
{
	// A slice holds n (n/2 + 1) < 2^31 elements: 32-bit indices (a 64-bit
	// division per element would cost more than the sample itself). Every
	// thread stays to the end: the bins are written by whole wavefronts.
	const int elements32 = p.n * p.h;
	const int idx = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
	const bool live = idx < elements32;
	const int clamped = live ? idx : 0;
	const int row = static_cast<int>(static_cast<unsigned>(clamped) / static_cast<unsigned>(p.h));
	const int column = clamped - row * p.h;
	const int64_t slices = p.map.kept_count * chunk.count;
	for (int64_t j = blockIdx.y; j < slices; j += gridDim.y)
	{
		const int64_t kept = j / chunk.count;
		const int64_t local = j - kept * chunk.count;
		const int64_t s = kept * p.map.reduced_count + chunk.begin + local;
		const slice_entry &e = p.entries[s];
		// Branch-free, as in the fused column pass: every element is loaded
		// and the rejected ones zeroed afterwards.
		const float2 *table = p.tables + s * static_cast<int64_t>(p.h + p.n);
		const float2 x = p.slices[e.slice_offset + row * p.slice_s0 + column * p.slice_s1];
		bool inserted = false;
		const prepared_sample v = prepare_unbranched<VARIANT>(p, e, p.radial + e.radial_offset, table + p.h,
			table[column], row, column, x, inserted);
		bin_put<false>(p, sink, nullptr, e, kept, static_cast<uint32_t>(local), row, column,
			live && e.last_shell >= 0 && inserted, v);
	}
}

// The column transforms of 256-point slices fused with the preparation.
// The rows of each slice are transformed already (in p.slices, at the
// slices' own offsets and strides); a block transforms 16 columns k_x of the
// chunk's slices by the four-step decomposition and writes the prepared
// samples of those columns into their bins directly, so the finished
// spectra never pass through HBM. The Nyquist column k_x = n/2 holds no
// inserted sample, so it is skipped: the lines are (slice j of the chunk,
// k_x < n/2), numbered j n/2 + k_x, and a block's 16 columns never straddle
// two slices.
template <int VARIANT>
__global__ __launch_bounds__(256) void columns_prepare256_kernel(insertion_params p, slice_chunk chunk,
	const float2 *tw, bin_sink sink)
{
	__shared__ float2 exchange[fft256_lines * fft256_line_stride];
	__shared__ float2 twl[256];
	__shared__ int64_t line_slice[fft256_lines];
	__shared__ int64_t line_src[fft256_lines];
	__shared__ int64_t line_kept[fft256_lines];
	__shared__ int line_local[fft256_lines];
	__shared__ int line_column[fft256_lines];
	const int t = threadIdx.x;
	const int l = t & (fft256_lines - 1);
	const int j = t / fft256_lines; // n1 in the first half of the transform, k1 in the second
	twl[t] = tw[t];
	const int h = p.h;
	const int columns = p.n / 2;
	const int64_t total = p.map.kept_count * chunk.count * columns;
	for (int64_t first = xcd_block(blockIdx.x, gridDim.x) * static_cast<int64_t>(fft256_lines); first < total;
		first += static_cast<int64_t>(gridDim.x) * fft256_lines)
	{
		const int lines = static_cast<int>(total - first < fft256_lines ? total - first : fft256_lines);
		if (t < fft256_lines)
		{
			int64_t slice = 0;
			int64_t src = 0;
			int64_t kept = 0;
			int local = 0;
			int column = 0;
			if (t < lines)
			{
				const int64_t line = first + t;
				const int64_t jj = line / columns;
				column = static_cast<int>(line - jj * columns);
				kept = jj / chunk.count;
				local = static_cast<int>(jj - kept * chunk.count);
				slice = kept * p.map.reduced_count + chunk.begin + local;
				src = p.entries[slice].slice_offset + column * p.slice_s1;
			}
			line_slice[t] = slice;
			line_src[t] = src;
			line_kept[t] = kept;
			line_local[t] = local;
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
		{
			// Every thread takes part (the bins are written by whole
			// wavefronts); lines past the block's last hold slice 0 and are
			// not inserted.
			const int64_t slice = line_slice[l];
			const slice_entry &e = p.entries[slice];
			const int column = line_column[l];
			const int64_t kept = line_kept[l];
			const uint32_t local = static_cast<uint32_t>(line_local[l]);
			const bool insertable = live && e.last_shell >= 0;
			const float2 *table = p.tables + slice * static_cast<int64_t>(h + p.n);
			const float2 x_phase = table[column];
			const float *radial = p.radial + e.radial_offset;
#pragma unroll 2
			for (int k2 = 0; k2 < 16; ++k2)
			{
				const int row = j + 16 * k2;
				bool inserted = false;
				const prepared_sample v = prepare_unbranched<VARIANT>(p, e, radial, table + h, x_phase, row, column,
					x[k2 * 17 + j], inserted);
				bin_put<false>(p, sink, nullptr, e, kept, local, row, column, insertable && inserted, v);
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
	const float2 *tw, int group, bin_sink sink)
{
	extern __shared__ float2 lds[];
	const int n = N > 0 ? N : plan.n;
	const int stride = n + 1;
	float2 *a = lds;
	float2 *b = lds + group * stride;
	float2 *twl = lds + 2 * group * stride;
	int64_t *line_slice = reinterpret_cast<int64_t *>(twl + n);
	int64_t *line_src = line_slice + group;
	int64_t *line_kept = line_src + group;
	int *line_column = reinterpret_cast<int *>(line_kept + group);
	int *line_local = line_column + group;
	for (int i = threadIdx.x; i < n; i += blockDim.x)
	{
		twl[i] = tw[i];
	}
	const int h = p.h;
	const int columns = n % 2 == 0 ? n / 2 : h;
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
			const int local = static_cast<int>(jj - kept * chunk.count);
			const int64_t slice = kept * p.map.reduced_count + chunk.begin + local;
			line_slice[q] = slice;
			line_src[q] = p.entries[slice].slice_offset + column * p.slice_s1;
			line_kept[q] = kept;
			line_column[q] = column;
			line_local[q] = local;
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
		// Every thread runs every round (the bins are written by whole
		// wavefronts); lanes past the block's elements take line 0's first
		// element and insert nothing.
		for (int round = 0; round < group * n; round += blockDim.x)
		{
			const int t = round + static_cast<int>(threadIdx.x);
			const int et = t / group;
			const int lt = t - et * group;
			const bool live = t < group * n && lt < lines;
			const int e = live ? et : 0;
			const int l = live ? lt : 0;
			const int64_t slice = line_slice[l];
			const slice_entry &entry = p.entries[slice];
			const float2 *table = p.tables + slice * static_cast<int64_t>(h + n);
			const int column = line_column[l];
			bool inserted = false;
			const prepared_sample v = prepare_unbranched<VARIANT>(p, entry, p.radial + entry.radial_offset, table + h,
				table[column], e, column, z[l * stride + e], inserted);
			bin_put<false>(p, sink, nullptr, entry, line_kept[l], static_cast<uint32_t>(line_local[l]), e, column,
				live && entry.last_shell >= 0 && inserted, v);
		}
		__syncthreads();
	}
}

void dispatch_prepare(const insertion_params &p, slice_chunk chunk, const bin_sink &sink, hipStream_t stream)
{
	const int64_t elements = static_cast<int64_t>(p.n) * p.h;
	const dim3 grid(static_cast<unsigned>((elements + 255) / 256),
		static_cast<unsigned>(std::min<int64_t>(p.map.kept_count * chunk.count, 65535)));
	switch (p.variant)
	{
	case variant_plain:
		prepare_kernel<variant_plain><<<grid, 256, 0, stream>>>(p, chunk, sink);
		break;
	case variant_weighted:
		prepare_kernel<variant_weighted><<<grid, 256, 0, stream>>>(p, chunk, sink);
		break;
	default:
		prepare_kernel<variant_ctf><<<grid, 256, 0, stream>>>(p, chunk, sink);
		break;
	}
}

template <int VARIANT>
void launch_columns_prepare(const insertion_params &p, slice_chunk chunk, const bin_sink &sink,
	hipStream_t stream)
{
	const int n = p.n;
	const int columns = n % 2 == 0 ? n / 2 : p.h;
	const int64_t lines = p.map.kept_count * chunk.count * columns;
	const fft_plan plan = make_fft_plan(n);
	const float2 *tw = twiddles(n);
	const int group = lines_per_group(n);
	const std::size_t shared = sizeof(float2) * (2 * group * (n + 1) + n) + group * (3 * sizeof(int64_t) + 2 * sizeof(int));
	const unsigned grid = xcd_grid(static_cast<unsigned>(std::min<int64_t>((lines + group - 1) / group, 1 << 20)));
	switch (n)
	{
	case 64:
		columns_prepare_kernel<VARIANT, 64><<<grid, 256, shared, stream>>>(p, chunk, plan, tw, group, sink);
		break;
	case 128:
		columns_prepare_kernel<VARIANT, 128><<<grid, 256, shared, stream>>>(p, chunk, plan, tw, group, sink);
		break;
	case 512:
		columns_prepare_kernel<VARIANT, 512><<<grid, 256, shared, stream>>>(p, chunk, plan, tw, group, sink);
		break;
	default:
		columns_prepare_kernel<VARIANT, 0><<<grid, 256, shared, stream>>>(p, chunk, plan, tw, group, sink);
		break;
	}
}

void dispatch_columns_prepare(const insertion_params &p, slice_chunk chunk, const bin_sink &sink,
	hipStream_t stream)
{
	if (p.n != 256)
	{
		switch (p.variant)
		{
		case variant_plain:
			launch_columns_prepare<variant_plain>(p, chunk, sink, stream);
			break;
		case variant_weighted:
			launch_columns_prepare<variant_weighted>(p, chunk, sink, stream);
			break;
		default:
			launch_columns_prepare<variant_ctf>(p, chunk, sink, stream);
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
		columns_prepare256_kernel<variant_plain><<<grid, 256, 0, stream>>>(p, chunk, tw, sink);
		break;
	case variant_weighted:
		columns_prepare256_kernel<variant_weighted><<<grid, 256, 0, stream>>>(p, chunk, tw, sink);
		break;
	default:
		columns_prepare256_kernel<variant_ctf><<<grid, 256, 0, stream>>>(p, chunk, tw, sink);
		break;
	}
}

} // anonymous namespace

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
	// The brick schedule needs the window to span at most 4 voxels; wider
	// Kaiser-Bessel windows take the scatter schedule, and so do calls whose
	// bins or record keys would not fit 32 bits.
	// Per-stream areas: about 6 GiB of records per chunk; a fixed buffer:
	// what it has left.
	const std::size_t budget = std::min<std::size_t>(workspace.remaining(), (std::size_t(6) << 30) + (std::size_t(1) << 28));
	const bin_layout layout = schedule != schedule_scatter && p.taps <= 4 ? plan_bins(p, budget) : bin_layout{};
	const bool brick = schedule != schedule_scatter && p.taps <= 4 && layout.usable;
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
		// Per bin a count, an end and a cursor; the tasks; the task count and
		// the next task; and the records of one chunk of slices.
		const std::size_t bin_bytes = align_up(sizeof(uint32_t) * static_cast<std::size_t>(layout.bins), 256);
		const std::size_t task_bytes = align_up(sizeof(int4) * static_cast<std::size_t>(layout.task_capacity), 256);
		char *area = static_cast<char *>(workspace.take(area_bins, 3 * bin_bytes + task_bytes + 256, stream));
		bin_state bins{};
		bins.counts = reinterpret_cast<uint32_t *>(area);
		bins.end = reinterpret_cast<uint32_t *>(area + bin_bytes);
		bins.tasks = reinterpret_cast<int4 *>(area + 3 * bin_bytes);
		bins.task_count = reinterpret_cast<uint32_t *>(area + 3 * bin_bytes + task_bytes);
		bins.task_next = bins.task_count + 1;
		bins.bins = layout.bins;
		bins.records_per_task = layout.records_per_task;
		bins.task_capacity = static_cast<uint32_t>(layout.task_capacity);
		bins.sink.g = layout.g;
		bins.sink.cursor = reinterpret_cast<uint32_t *>(area + 2 * bin_bytes);
		bins.sink.end = bins.end;
		bins.sink.records = static_cast<bin_record *>(workspace.take(area_prepared,
			sizeof(bin_record) * static_cast<std::size_t>(layout.record_capacity), stream));
		bins.sink.capacity = static_cast<uint32_t>(layout.record_capacity);
		for (int64_t begin = 0; begin < p.map.reduced_count; begin += layout.chunk_slices)
		{
			const slice_chunk chunk{begin, std::min(layout.chunk_slices, p.map.reduced_count - begin)};
			const int count_token = timing_begin(REXLIB_HIP_KERNEL_BIN_COUNT, stream);
			launch_bin_count(p, chunk, bins, stream);
			timing_end(count_token, stream);
			const int prepare_token = timing_begin(REXLIB_HIP_KERNEL_PREPARE, stream);
			if (fused_columns)
			{
				dispatch_columns_prepare(p, chunk, bins.sink, stream);
			}
			else
			{
				dispatch_prepare(p, chunk, bins.sink, stream);
			}
			REXLIB_HIP_CHECK(hipGetLastError());
			timing_end(prepare_token, stream);
			const int brick_token = timing_begin(REXLIB_HIP_KERNEL_BRICK, stream);
			launch_bin_brick(p, chunk, bins, stream);
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
