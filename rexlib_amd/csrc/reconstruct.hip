// The steps after the insertion: weight normalisation, gridding correction,
// and per-shell sums in float64.
//
// The reference has no function for these; its only reconstruction is a test
// recipe that divides on the host where the denominator exceeds 1e-3
// (tests/integration/src/em/functional/test_cpu_reconstruction.cpp:457-464).
#include "common.hpp"

#include <cmath>

namespace rexlib_hip
{

namespace
{

__global__ void normalize_kernel(const float2 *num, const float *den, float2 *out, int64_t count, float threshold)
{
	for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < count;
		i += static_cast<int64_t>(gridDim.x) * blockDim.x)
	{
		const float d = den[i];
		const float2 v = num[i];
		out[i] = d > threshold ? make_float2(v.x / d, v.y / d) : make_float2(0.0f, 0.0f);
	}
}

// The 1-D transform of the window at real-space offset u = (x - C) / N,
// normalised to 1 at u = 0, in double.
__device__ double window_transform(int profile, double half_width, double shape, double u)
{
	const double pi = 3.14159265358979323846;
	if (profile == REXLIB_HIP_BOX || profile == REXLIB_HIP_TRIANGLE)
	{
		const double a = pi * u;
		const double s = fabs(a) < 1e-12 ? 1.0 : sin(a) / a;
		return profile == REXLIB_HIP_BOX ? s : s * s;
	}
	// Kaiser-Bessel: sinh(sqrt(b^2 - (2 pi w u)^2)) / sqrt(...), over its value at 0.
	const auto kb = [](double z2)
	{
		if (z2 > 1e-12)
		{
			const double r = sqrt(z2);
			return sinh(r) / r;
		}
		if (z2 < -1e-12)
		{
			const double r = sqrt(-z2);
			return sin(r) / r;
		}
		return 1.0;
	};
	const double t = 2.0 * pi * half_width * u;
	return kb(shape * shape - t * t) / kb(shape * shape);
}

// out (crop^3 per volume) = the central crop^3 of in (big^3 per volume),
// divided by the window's transform along each axis. The crop starts at
// big/2 - crop/2, so that the volume's centre big/2 stays at crop/2 (the
// central n^3 of a reconstruction inserted with 2x padding).
__global__ void gridding_kernel(const float *in, float *out, int big, int crop, int64_t volumes, int profile,
	double half_width, double shape)
{
	extern __shared__ float table[];
	const int offset = big / 2 - crop / 2;
	for (int i = threadIdx.x; i < crop; i += blockDim.x)
	{
		const double u = static_cast<double>(i + offset - big / 2) / static_cast<double>(big);
		table[i] = static_cast<float>(1.0 / window_transform(profile, half_width, shape, u));
	}
	__syncthreads();
	const int64_t volume = static_cast<int64_t>(crop) * crop * crop;
	const int64_t big_volume = static_cast<int64_t>(big) * big * big;
	const int64_t count = volume * volumes;
	for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < count;
		i += static_cast<int64_t>(gridDim.x) * blockDim.x)
	{
		const int64_t b = i / volume;
		int64_t v = i - b * volume;
		const int x = static_cast<int>(v % crop);
		v /= crop;
		const int y = static_cast<int>(v % crop);
		const int z = static_cast<int>(v / crop);
		const int64_t source = b * big_volume + ((static_cast<int64_t>(z + offset) * big + (y + offset)) * big) +
			(x + offset);
		out[i] = in[source] * table[x] * table[y] * table[z];
	}
}

// Per-shell sums with block-private float64 partials in LDS.
__global__ void shell_sums_kernel(const float2 *a, const float2 *b, double *sums, int big)
{
	extern __shared__ double partial[]; // 4 * shells
	const int half = big / 2 + 1;
	const int shells = half;
	for (int i = threadIdx.x; i < 4 * shells; i += blockDim.x)
	{
		partial[i] = 0.0;
	}
	__syncthreads();
	const int64_t count = static_cast<int64_t>(big) * big * half;
	for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < count;
		i += static_cast<int64_t>(gridDim.x) * blockDim.x)
	{
		int64_t v = i;
		const int x = static_cast<int>(v % half);
		v /= half;
		const int sy = static_cast<int>(v % big);
		const int sz = static_cast<int>(v / big);
		const int y = sy <= big / 2 ? sy : sy - big;
		const int z = sz <= big / 2 ? sz : sz - big;
		const int64_t q4 = 4ll * (static_cast<int64_t>(x) * x + static_cast<int64_t>(y) * y +
			static_cast<int64_t>(z) * z);
		int64_t s = static_cast<int64_t>(sqrt(static_cast<double>(q4) * 0.25) + 0.5);
		while (4 * s * s + 4 * s + 1 <= q4)
		{
			++s;
		}
		while (s > 0 && 4 * s * s - 4 * s + 1 > q4)
		{
			--s;
		}
		if (s >= shells)
		{
			continue;
		}
		const float2 va = a[i];
		const float2 vb = b[i];
		const double ar = va.x, ai = va.y, br = vb.x, bi = vb.y;
		atomicAdd(&partial[4 * s + 0], ar * br + ai * bi);
		atomicAdd(&partial[4 * s + 1], ar * ar + ai * ai);
		atomicAdd(&partial[4 * s + 2], br * br + bi * bi);
		atomicAdd(&partial[4 * s + 3], 1.0);
	}
	__syncthreads();
	for (int i = threadIdx.x; i < 4 * shells; i += blockDim.x)
	{
		if (partial[i] != 0.0)
		{
			atomicAdd(&sums[(i % 4) * shells + i / 4], partial[i]);
		}
	}
}

unsigned blocks_for(int64_t count)
{
	return static_cast<unsigned>(std::max<int64_t>(1, std::min<int64_t>((count + 255) / 256, 8192)));
}

} // anonymous namespace

void launch_normalize(const float2 *num, const float *den, float2 *out, int64_t count, double threshold,
	hipStream_t stream)
{
	normalize_kernel<<<blocks_for(count), 256, 0, stream>>>(num, den, out, count, static_cast<float>(threshold));
	REXLIB_HIP_CHECK(hipGetLastError());
}

void launch_gridding(const float *in, float *out, int big, int crop, int64_t volumes, int profile,
	double half_width, double shape, hipStream_t stream)
{
	const int64_t count = static_cast<int64_t>(crop) * crop * crop * volumes;
	gridding_kernel<<<blocks_for(count), 256, sizeof(float) * crop, stream>>>(in, out, big, crop, volumes, profile,
		half_width, shape);
	REXLIB_HIP_CHECK(hipGetLastError());
}

void launch_shell_sums(const float2 *a, const float2 *b, double *sums, int big, hipStream_t stream)
{
	const int64_t count = static_cast<int64_t>(big) * big * (big / 2 + 1);
	const int shells = big / 2 + 1;
	shell_sums_kernel<<<std::min<unsigned>(blocks_for(count), 1024), 256, sizeof(double) * 4 * shells, stream>>>(
		a, b, sums, big);
	REXLIB_HIP_CHECK(hipGetLastError());
}

} // namespace rexlib_hip
