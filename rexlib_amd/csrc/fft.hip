// Batched FFTs on gfx950: Stockham autosort transforms in LDS.
//
// Every kernel stages a group of lines in LDS, runs the radix-4/2/3/5
// stages there (ping-pong between two LDS buffers, one butterfly per thread
// per step), and writes the lines back. Real transforms pack two real rows
// into one complex line and separate the two half spectra afterwards, so a
// real row costs half a complex transform.
//
// Conventions follow pocketfft as rexlib calls it
// (src/backends/cpu/kernels/pocketfft_fourier_transform.cpp:74-163): the
// forward transform is unscaled with exp(-2 pi i), the inverse uses
// exp(+2 pi i) and a scale of 1/n per axis
// (src/ops/fourier/fourier_normalization.cpp:56-71). A complex-to-real
// transform ignores the imaginary parts of the self-conjugate elements.
#include "fft.hpp"
#include "fft_device.hpp"

#include <hip/hip_bf16.h>
#include <hip/hip_fp16.h>

#include <cmath>
#include <cstdint>
#include <map>
#include <mutex>
#include <type_traits>
#include <vector>

namespace rexlib_hip
{

fft_plan make_fft_plan(int64_t n)
{
	fft_plan plan{};
	plan.n = static_cast<int32_t>(n);
	int64_t rest = n;
	const int radices[] = {4, 2, 3, 5};
	// Radix 4 first, then a single 2, then 3 and 5.
	while (rest % 4 == 0)
	{
		plan.radix[plan.stages++] = 4;
		rest /= 4;
	}
	for (const int r : radices)
	{
		while (rest % r == 0)
		{
			if (plan.stages == max_fft_stages)
			{
				fail(REXLIB_HIP_UNSUPPORTED, "fft: length has too many factors");
			}
			plan.radix[plan.stages++] = r;
			rest /= r;
		}
	}
	if (rest != 1)
	{
		fail(REXLIB_HIP_UNSUPPORTED,
			"fft: the gfx950 transforms take lengths whose prime factors are 2, 3, and 5, not " +
				std::to_string(n));
	}
	return plan;
}

const float2 *twiddles(int64_t n)
{
	static std::mutex mutex;
	static std::map<std::pair<int, int64_t>, float2 *> cache;
	int device = 0;
	REXLIB_HIP_CHECK(hipGetDevice(&device));
	std::lock_guard<std::mutex> lock(mutex);
	const auto key = std::make_pair(device, n);
	const auto found = cache.find(key);
	if (found != cache.end())
	{
		return found->second;
	}
	std::vector<float2> host(static_cast<std::size_t>(n));
	for (int64_t m = 0; m < n; ++m)
	{
		const double angle = -2.0 * 3.14159265358979323846 * static_cast<double>(m) / static_cast<double>(n);
		host[m] = make_float2(static_cast<float>(std::cos(angle)), static_cast<float>(std::sin(angle)));
	}
	float2 *device_table = nullptr;
	REXLIB_HIP_CHECK(hipMalloc(&device_table, sizeof(float2) * static_cast<std::size_t>(n)));
	REXLIB_HIP_CHECK(hipMemcpy(device_table, host.data(), sizeof(float2) * static_cast<std::size_t>(n),
		hipMemcpyHostToDevice));
	cache[key] = device_table;
	return device_table;
}

namespace
{

template <int KIND>
__device__ inline float load_real(const void *base, int64_t index);

template <>
__device__ inline float load_real<real_f32>(const void *base, int64_t index)
{
	return static_cast<const float *>(base)[index];
}

template <>
__device__ inline float load_real<real_bf16>(const void *base, int64_t index)
{
	return __bfloat162float(static_cast<const __hip_bfloat16 *>(base)[index]);
}

template <>
__device__ inline float load_real<real_f16>(const void *base, int64_t index)
{
	return __half2float(static_cast<const __half *>(base)[index]);
}

// Four consecutive elements from `index` (a multiple of 4, with the base
// aligned to four elements), as floats.
template <int KIND>
__device__ inline float4 load_real4(const void *base, int64_t index)
{
	if constexpr (KIND == real_f32)
	{
		return *reinterpret_cast<const float4 *>(static_cast<const float *>(base) + index);
	}
	else
	{
		const uint2 bits = *reinterpret_cast<const uint2 *>(static_cast<const unsigned short *>(base) + index);
		const unsigned short e[4] = {static_cast<unsigned short>(bits.x & 0xffff), static_cast<unsigned short>(bits.x >> 16),
			static_cast<unsigned short>(bits.y & 0xffff), static_cast<unsigned short>(bits.y >> 16)};
		float f[4];
#pragma unroll
		for (int i = 0; i < 4; ++i)
		{
			if constexpr (KIND == real_bf16)
			{
				f[i] = __uint_as_float(static_cast<unsigned>(e[i]) << 16);
			}
			else
			{
				f[i] = __half2float(__ushort_as_half(e[i]));
			}
		}
		return make_float4(f[0], f[1], f[2], f[3]);
	}
}

// Rows of real images, two at a time, to half spectra. Images of extent m
// are zero-padded to the transform length n with their origin m/2 moved to
// n/2. `rows` counts the padded rows, n per image, in line set `in` of images
// (middle = image row, outer = image); the output rows follow line set `out`.
template <int KIND, int N>
__global__ void rfft_rows_kernel(const void *in, line_set in_rows, int m, const float *scales,
	int64_t scale_stride, float2 *out, line_set out_rows, fft_plan plan, const float2 *tw, int pairs)
{
	extern __shared__ float2 lds[];
	const int n = N > 0 ? N : plan.n;
	const int h = n / 2 + 1;
	const int stride = n + 1;
	float2 *a = lds;
	float2 *b = lds + pairs * stride;
	float2 *twl = lds + 2 * pairs * stride;
	// Per padded row of the group: its first input element (-1: a zero row
	// of the padding or past the end) and its first output element.
	int64_t *row_in = reinterpret_cast<int64_t *>(twl + n);
	int64_t *row_out = row_in + 2 * pairs;
	float *row_scale = reinterpret_cast<float *>(row_out + 2 * pairs);
	for (int i = threadIdx.x; i < n; i += blockDim.x)
	{
		twl[i] = tw[i];
	}
	const int64_t total = out_rows.count; // padded rows of all images
	const int offset = n / 2 - m / 2;
	for (int64_t first = blockIdx.x * static_cast<int64_t>(2 * pairs); first < total;
		first += static_cast<int64_t>(gridDim.x) * 2 * pairs)
	{
		for (int q = threadIdx.x; q < 2 * pairs; q += blockDim.x)
		{
			const int64_t row = first + q;
			int64_t src = -1;
			int64_t dst = -1;
			float scale = 1.0f;
			if (row < total)
			{
				const int64_t image = row / n;
				const int r = static_cast<int>(row - image * n) - offset;
				if (r >= 0 && r < m)
				{
					src = line_offset(in_rows, image * m + r);
					if (scales != nullptr)
					{
						scale = scales[image * scale_stride];
					}
				}
				dst = line_offset(out_rows, row);
			}
			row_in[q] = src;
			row_out[q] = dst;
			row_scale[q] = scale;
		}
		__syncthreads();
		for (int t = threadIdx.x; t < pairs * n; t += blockDim.x)
		{
			const int p = t / n;
			const int i = t - p * n;
			const int c = i - offset;
			float value[2] = {0.0f, 0.0f};
			for (int half = 0; half < 2; ++half)
			{
				const int64_t src = row_in[2 * p + half];
				if (src >= 0 && c >= 0 && c < m)
				{
					value[half] = load_real<KIND>(in, src + c * in_rows.s_elem) * row_scale[2 * p + half];
				}
			}
			a[p * stride + i] = make_float2(value[0], value[1]);
		}
		__syncthreads();
		const float2 *z = transform<N>(a, b, pairs, stride, plan, twl, -1.0f);
		for (int t = threadIdx.x; t < pairs * h; t += blockDim.x)
		{
			const int p = t / h;
			const int k = t - p * h;
			const float2 zk = z[p * stride + k];
			const float2 zm = cconj(z[p * stride + (n - k) % n]);
			const int64_t dst0 = row_out[2 * p];
			const int64_t dst1 = row_out[2 * p + 1];
			if (dst0 >= 0)
			{
				const float2 sum = cadd(zk, zm);
				out[dst0 + k * out_rows.s_elem] = make_float2(0.5f * sum.x, 0.5f * sum.y);
			}
			if (dst1 >= 0)
			{
				// (zk - zm) / (2 i)
				const float2 diff = csub(zk, zm);
				out[dst1 + k * out_rows.s_elem] = make_float2(0.5f * diff.y, -0.5f * diff.x);
			}
		}
		__syncthreads();
	}
}

// Complex lines from `src` to `dst` (may be the same memory), `group` lines
// per step; consecutive lines of a group are loaded by consecutive threads.
template <int N>
__global__ void fft_lines_kernel(const float2 *src, line_set src_lines, float2 *dst, line_set dst_lines,
	fft_plan plan, const float2 *tw, float sign, float scale, int group)
{
	extern __shared__ float2 lds[];
	const int n = N > 0 ? N : plan.n;
	const int stride = n + 1;
	float2 *a = lds;
	float2 *b = lds + group * stride;
	float2 *twl = lds + 2 * group * stride;
	int64_t *line_src = reinterpret_cast<int64_t *>(twl + n);
	int64_t *line_dst = line_src + group;
	for (int i = threadIdx.x; i < n; i += blockDim.x)
	{
		twl[i] = tw[i];
	}
	const int64_t total = src_lines.count;
	for (int64_t first = blockIdx.x * static_cast<int64_t>(group); first < total;
		first += static_cast<int64_t>(gridDim.x) * group)
	{
		const int lines = static_cast<int>(total - first < group ? total - first : group);
		for (int q = threadIdx.x; q < lines; q += blockDim.x)
		{
			line_src[q] = line_offset(src_lines, first + q);
			line_dst[q] = line_offset(dst_lines, first + q);
		}
		__syncthreads();
		for (int t = threadIdx.x; t < group * n; t += blockDim.x)
		{
			const int e = t / group;
			const int l = t - e * group;
			if (l < lines)
			{
				a[l * stride + e] = src[line_src[l] + e * src_lines.s_elem];
			}
		}
		__syncthreads();
		const float2 *z = transform<N>(a, b, lines, stride, plan, twl, sign);
		for (int t = threadIdx.x; t < group * n; t += blockDim.x)
		{
			const int e = t / group;
			const int l = t - e * group;
			if (l < lines)
			{
				const float2 v = z[l * stride + e];
				dst[line_dst[l] + e * dst_lines.s_elem] = make_float2(v.x * scale, v.y * scale);
			}
		}
		__syncthreads();
	}
}

__global__ __launch_bounds__(256) void fft256_lines_kernel(const float2 *src, line_set src_lines, float2 *dst,
	line_set dst_lines, const float2 *tw, float sign, float scale)
{
	// Per line 16 x 17 values (odd pitches against bank conflicts).
	__shared__ float2 exchange[fft256_lines * fft256_line_stride];
	__shared__ float2 twl[256];
	__shared__ int64_t line_src[fft256_lines];
	__shared__ int64_t line_dst[fft256_lines];
	const int t = threadIdx.x;
	const int l = t & (fft256_lines - 1);
	const int j = t / fft256_lines; // n1 in the first half, k1 in the second
	twl[t] = tw[t];
	const int64_t total = src_lines.count;
	for (int64_t first = xcd_block(blockIdx.x, gridDim.x) * static_cast<int64_t>(fft256_lines); first < total;
		first += static_cast<int64_t>(gridDim.x) * fft256_lines)
	{
		const int lines = static_cast<int>(total - first < fft256_lines ? total - first : fft256_lines);
		if (t < fft256_lines)
		{
			line_src[t] = t < lines ? line_offset(src_lines, first + t) : 0;
			line_dst[t] = t < lines ? line_offset(dst_lines, first + t) : 0;
		}
		__syncthreads();
		float2 v[16];
		const bool live = l < lines;
		const float2 *in = src + line_src[l];
#pragma unroll
		for (int k = 0; k < 16; ++k)
		{
			v[k] = live ? in[(j + 16 * k) * src_lines.s_elem] : make_float2(0.0f, 0.0f);
		}
		float2 *x = exchange + l * fft256_line_stride;
		fft256_four_step(v, x, twl, j, sign);
		// The outputs go back to LDS (each thread to slots only it reads),
		// so the stores loop over them without holding all 16.
		__syncthreads();
#pragma unroll
		for (int k2 = 0; k2 < 16; ++k2)
		{
			x[k2 * 17 + j] = v[k2];
		}
		if (live)
		{
			float2 *out = dst + line_dst[l] + j * dst_lines.s_elem;
			const int64_t step = 16 * dst_lines.s_elem;
#pragma unroll 4
			for (int k2 = 0; k2 < 16; ++k2)
			{
				const float2 y = x[k2 * 17 + j];
				out[k2 * step] = make_float2(y.x * scale, y.y * scale);
			}
		}
		__syncthreads();
	}
}

// Real 256-point rows, two to a complex line, by the four-step
// decomposition: a block takes 16 lines (32 rows). Thread (line l, j) loads
// elements j + 16 k of its two rows (16 consecutive elements per line and
// load across the lanes), transforms them with fft256_four_step, and stores
// Z[j + 16 k2] to LDS. The separation of the two half spectra
// (X_a = (Z_k + conj Z_{n-k}) / 2, X_b = (Z_k - conj Z_{n-k}) / 2i) then
// reads pairs from LDS, and writes each output row's 129 values, 16
// consecutive per lane group. Rows of the padding, and columns outside the
// image, are zeros, as in rfft_rows_kernel.
template <int KIND, bool VEC>
__global__ __launch_bounds__(256) void rfft256_rows_kernel(const void *in, line_set in_rows, int m,
	const float *scales, int64_t scale_stride, float2 *out, line_set out_rows, const float2 *tw)
{
	constexpr int n = 256;
	constexpr int h = n / 2 + 1;
	__shared__ float2 exchange[fft256_lines * fft256_line_stride];
	__shared__ float2 twl[256];
	__shared__ int64_t row_in[2 * fft256_lines];
	__shared__ int64_t row_out[2 * fft256_lines];
	__shared__ float row_scale[2 * fft256_lines];
	const int t = threadIdx.x;
	const int j = t & 15;
	const int l = t >> 4;
	twl[t] = tw[t];
	const int64_t total = out_rows.count; // padded rows of all images
	const int offset = n / 2 - m / 2;
	for (int64_t first = xcd_block(blockIdx.x, gridDim.x) * static_cast<int64_t>(2 * fft256_lines); first < total;
		first += static_cast<int64_t>(gridDim.x) * 2 * fft256_lines)
	{
		if (t < 2 * fft256_lines)
		{
			const int64_t row = first + t;
			int64_t src = -1;
			int64_t dst = -1;
			float scale = 1.0f;
			if (row < total)
			{
				const int64_t image = row / n;
				const int r = static_cast<int>(row - image * n) - offset;
				if (r >= 0 && r < m)
				{
					src = line_offset(in_rows, image * m + r);
					if (scales != nullptr)
					{
						scale = scales[image * scale_stride];
					}
				}
				dst = line_offset(out_rows, row);
			}
			row_in[t] = src;
			row_out[t] = dst;
			row_scale[t] = scale;
		}
		__syncthreads();
		float2 *x = exchange + l * fft256_line_stride;
		float2 v[16];
		if constexpr (VEC)
		{
			// Rows whose every run of 4 elements is aligned to 4 elements:
			// each thread loads 4 consecutive elements of a row pair at a
			// time (a wave reads whole rows), and the elements reach their
			// threads through the exchange, element i of line l at
			// (i / 16) 17 + i % 16.
#pragma unroll
			for (int step = 0; step < 4; ++step)
			{
				const int q = step * 256 + t;  // float4 index within the block's 16 lines
				const int ll = q >> 6;         // line
				const int i = (q & 63) * 4;    // first element
				const int c = i - offset;
				const int64_t sa = row_in[2 * ll];
				const int64_t sb = row_in[2 * ll + 1];
				float4 a = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
				float4 b = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
				if (c >= 0 && c < m)
				{
					if (sa >= 0)
					{
						const float g = row_scale[2 * ll];
						a = load_real4<KIND>(in, sa + c);
						a = make_float4(a.x * g, a.y * g, a.z * g, a.w * g);
					}
					if (sb >= 0)
					{
						const float g = row_scale[2 * ll + 1];
						b = load_real4<KIND>(in, sb + c);
						b = make_float4(b.x * g, b.y * g, b.z * g, b.w * g);
					}
				}
				float2 *xl = exchange + ll * fft256_line_stride + (i >> 4) * 17 + (i & 15);
				xl[0] = make_float2(a.x, b.x);
				xl[1] = make_float2(a.y, b.y);
				xl[2] = make_float2(a.z, b.z);
				xl[3] = make_float2(a.w, b.w);
			}
			__syncthreads();
#pragma unroll
			for (int k = 0; k < 16; ++k)
			{
				v[k] = x[k * 17 + j];
			}
			__syncthreads();
		}
		else
		{
			const int64_t src_a = row_in[2 * l];
			const int64_t src_b = row_in[2 * l + 1];
			const float scale_a = row_scale[2 * l];
			const float scale_b = row_scale[2 * l + 1];
#pragma unroll
			for (int k = 0; k < 16; ++k)
			{
				const int c = j + 16 * k - offset;
				const bool inside = c >= 0 && c < m;
				const float a =
					src_a >= 0 && inside ? load_real<KIND>(in, src_a + c * in_rows.s_elem) * scale_a : 0.0f;
				const float b =
					src_b >= 0 && inside ? load_real<KIND>(in, src_b + c * in_rows.s_elem) * scale_b : 0.0f;
				v[k] = make_float2(a, b);
			}
		}
		fft256_four_step(v, x, twl, j, -1.0f);
		__syncthreads();
#pragma unroll
		for (int k2 = 0; k2 < 16; ++k2)
		{
			x[k2 * 17 + j] = v[k2]; // Z[j + 16 k2]
		}
		__syncthreads();
		const int64_t dst_a = row_out[2 * l];
		const int64_t dst_b = row_out[2 * l + 1];
#pragma unroll
		for (int q = 0; q < 9; ++q)
		{
			const int k = j + 16 * q;
			if (k < h)
			{
				const int km = (n - k) & (n - 1);
				const float2 zk = x[q * 17 + j];
				const float2 zm = cconj(x[(km >> 4) * 17 + (km & 15)]);
				if (dst_a >= 0)
				{
					const float2 sum = cadd(zk, zm);
					out[dst_a + k * out_rows.s_elem] = make_float2(0.5f * sum.x, 0.5f * sum.y);
				}
				if (dst_b >= 0)
				{
					// (zk - zm) / (2 i)
					const float2 diff = csub(zk, zm);
					out[dst_b + k * out_rows.s_elem] = make_float2(0.5f * diff.y, -0.5f * diff.x);
				}
			}
		}
		__syncthreads();
	}
}

// Half-spectrum rows, two at a time, to real rows, scaled.
template <int N>
__global__ void irfft_rows_kernel(const float2 *src, line_set src_rows, float *dst, line_set dst_rows,
	fft_plan plan, const float2 *tw, float scale, int pairs)
{
	extern __shared__ float2 lds[];
	const int n = N > 0 ? N : plan.n;
	const int h = n / 2 + 1;
	const int stride = n + 1;
	float2 *a = lds;
	float2 *b = lds + pairs * stride;
	float2 *twl = lds + 2 * pairs * stride;
	int64_t *row_src = reinterpret_cast<int64_t *>(twl + n);
	int64_t *row_dst = row_src + 2 * pairs;
	for (int i = threadIdx.x; i < n; i += blockDim.x)
	{
		twl[i] = tw[i];
	}
	const int64_t total = src_rows.count;
	for (int64_t first = blockIdx.x * static_cast<int64_t>(2 * pairs); first < total;
		first += static_cast<int64_t>(gridDim.x) * 2 * pairs)
	{
		for (int q = threadIdx.x; q < 2 * pairs; q += blockDim.x)
		{
			const int64_t row = first + q;
			row_src[q] = row < total ? line_offset(src_rows, row) : -1;
			row_dst[q] = row < total ? line_offset(dst_rows, row) : -1;
		}
		__syncthreads();
		for (int t = threadIdx.x; t < pairs * n; t += blockDim.x)
		{
			const int p = t / n;
			const int k = t - p * n;
			const bool mirrored = k >= h;
			const int kk = mirrored ? n - k : k;
			const bool self_conjugate = kk == 0 || (n % 2 == 0 && kk == n / 2);
			float2 va = make_float2(0.0f, 0.0f);
			float2 vb = make_float2(0.0f, 0.0f);
			const int64_t src0 = row_src[2 * p];
			const int64_t src1 = row_src[2 * p + 1];
			if (src0 >= 0)
			{
				va = src[src0 + kk * src_rows.s_elem];
			}
			if (src1 >= 0)
			{
				vb = src[src1 + kk * src_rows.s_elem];
			}
			if (self_conjugate)
			{
				va.y = 0.0f;
				vb.y = 0.0f;
			}
			if (mirrored)
			{
				va = cconj(va);
				vb = cconj(vb);
			}
			// Z = A + i B
			a[p * stride + k] = make_float2(va.x - vb.y, va.y + vb.x);
		}
		__syncthreads();
		const float2 *z = transform<N>(a, b, pairs, stride, plan, twl, 1.0f);
		for (int t = threadIdx.x; t < pairs * n; t += blockDim.x)
		{
			const int p = t / n;
			const int i = t - p * n;
			const float2 v = z[p * stride + i];
			const int64_t dst0 = row_dst[2 * p];
			const int64_t dst1 = row_dst[2 * p + 1];
			if (dst0 >= 0)
			{
				dst[dst0 + i * dst_rows.s_elem] = v.x * scale;
			}
			if (dst1 >= 0)
			{
				dst[dst1 + i * dst_rows.s_elem] = v.y * scale;
			}
		}
		__syncthreads();
	}
}

unsigned grid_for(int64_t work, int64_t per_block)
{
	const int64_t blocks = (work + per_block - 1) / per_block;
	return static_cast<unsigned>(std::max<int64_t>(1, std::min<int64_t>(blocks, 1 << 16)));
}

} // anonymous namespace

int lines_per_group(int n)
{
	// About 2048 complex elements (16 KiB) per LDS buffer: small blocks keep
	// more wavefronts per CU (measured faster than 4096 or 8192 at n = 256).
	return std::max(1, std::min(8, 2048 / std::max(n, 1)));
}

// Calls f.template operator()<N>() with N the length when it is a power of
// two from 64 to 512, else 0 (the planned transform).
template <typename F>
void with_fixed_length(int64_t n, F &&f)
{
	switch (n)
	{
	case 64: f(std::integral_constant<int, 64>()); break;
	case 128: f(std::integral_constant<int, 128>()); break;
	case 256: f(std::integral_constant<int, 256>()); break;
	case 512: f(std::integral_constant<int, 512>()); break;
	default: f(std::integral_constant<int, 0>()); break;
	}
}

void launch_rfft_rows(int kind, const void *in, const line_set &in_rows, int m, const float *scales,
	int64_t scale_stride, float2 *out, const line_set &out_rows, int64_t n, hipStream_t stream)
{
	const fft_plan plan = make_fft_plan(n);
	const float2 *tw = twiddles(n);
	if (n == 256)
	{
		const unsigned blocks = xcd_grid(grid_for(out_rows.count, 2 * fft256_lines));
		// Rows take loads of 4 elements when every run of 4 elements is
		// aligned and contiguous.
		const std::uintptr_t element = kind == real_f32 ? 4 : 2;
		const bool aligned = in_rows.s_elem == 1 && m % 4 == 0 && (n / 2 - m / 2) % 4 == 0 &&
			in_rows.s_outer % 4 == 0 && in_rows.s_middle % 4 == 0 && in_rows.s_inner % 4 == 0 &&
			reinterpret_cast<std::uintptr_t>(in) % (4 * element) == 0;
		auto launch = [&](auto kind_constant, auto vec_constant)
		{
			rfft256_rows_kernel<decltype(kind_constant)::value, decltype(vec_constant)::value>
				<<<blocks, 256, 0, stream>>>(in, in_rows, m, scales, scale_stride, out, out_rows, tw);
		};
		using vec = std::true_type;
		using scalar = std::false_type;
		switch (kind)
		{
		case real_f32:
			aligned ? launch(std::integral_constant<int, real_f32>(), vec()) :
				launch(std::integral_constant<int, real_f32>(), scalar());
			break;
		case real_bf16:
			aligned ? launch(std::integral_constant<int, real_bf16>(), vec()) :
				launch(std::integral_constant<int, real_bf16>(), scalar());
			break;
		default:
			aligned ? launch(std::integral_constant<int, real_f16>(), vec()) :
				launch(std::integral_constant<int, real_f16>(), scalar());
			break;
		}
		REXLIB_HIP_CHECK(hipGetLastError());
		return;
	}
	const int pairs = std::max(1, lines_per_group(static_cast<int>(n)) / 2);
	// Two buffers, the twiddles, then per row two offsets and a scale.
	const std::size_t shared = sizeof(float2) * (2 * pairs * (n + 1) + n) + 2 * pairs * (2 * sizeof(int64_t) + 4);
	const unsigned grid = grid_for(out_rows.count, 2 * pairs);
	with_fixed_length(n, [&](auto length)
		{
			constexpr int N = decltype(length)::value;
			switch (kind)
			{
			case real_f32:
				rfft_rows_kernel<real_f32, N><<<grid, 256, shared, stream>>>(in, in_rows, m, scales, scale_stride,
					out, out_rows, plan, tw, pairs);
				break;
			case real_bf16:
				rfft_rows_kernel<real_bf16, N><<<grid, 256, shared, stream>>>(in, in_rows, m, scales, scale_stride,
					out, out_rows, plan, tw, pairs);
				break;
			default:
				rfft_rows_kernel<real_f16, N><<<grid, 256, shared, stream>>>(in, in_rows, m, scales, scale_stride,
					out, out_rows, plan, tw, pairs);
				break;
			}
		});
	REXLIB_HIP_CHECK(hipGetLastError());
}

void launch_fft_lines(const float2 *src, const line_set &src_lines, float2 *dst, const line_set &dst_lines,
	int64_t n, float sign, float scale, hipStream_t stream)
{
	const fft_plan plan = make_fft_plan(n);
	const float2 *tw = twiddles(n);
	const int group = lines_per_group(static_cast<int>(n));
	const std::size_t shared = sizeof(float2) * (2 * group * (n + 1) + n) + 2 * group * sizeof(int64_t);
	const unsigned grid = grid_for(src_lines.count, group);
	if (n == 256)
	{
		const unsigned blocks = xcd_grid(grid_for(src_lines.count, fft256_lines));
		fft256_lines_kernel<<<blocks, 256, 0, stream>>>(src, src_lines, dst, dst_lines, tw, sign, scale);
		REXLIB_HIP_CHECK(hipGetLastError());
		return;
	}
	with_fixed_length(n, [&](auto length)
		{
			constexpr int N = decltype(length)::value;
			fft_lines_kernel<N><<<grid, 256, shared, stream>>>(src, src_lines, dst, dst_lines, plan, tw, sign, scale,
				group);
		});
	REXLIB_HIP_CHECK(hipGetLastError());
}

void launch_irfft_rows(const float2 *src, const line_set &src_rows, float *dst, const line_set &dst_rows,
	int64_t n, float scale, hipStream_t stream)
{
	const fft_plan plan = make_fft_plan(n);
	const float2 *tw = twiddles(n);
	const int pairs = std::max(1, lines_per_group(static_cast<int>(n)) / 2);
	const std::size_t shared = sizeof(float2) * (2 * pairs * (n + 1) + n) + 2 * pairs * 2 * sizeof(int64_t);
	const unsigned grid = grid_for(src_rows.count, 2 * pairs);
	with_fixed_length(n, [&](auto length)
		{
			constexpr int N = decltype(length)::value;
			irfft_rows_kernel<N><<<grid, 256, shared, stream>>>(src, src_rows, dst, dst_rows, plan, tw, scale, pairs);
		});
	REXLIB_HIP_CHECK(hipGetLastError());
}

} // namespace rexlib_hip
