// Register-level DFT pieces shared by the FFT kernels and the fused column
// pass of the image insertion.
#pragma once

#include "common.hpp"

namespace rexlib_hip
{

__device__ inline float2 times_i(float2 z, float sign)
{
	// (sign * i) * z
	return make_float2(-sign * z.y, sign * z.x);
}

// Radix-R DFT in registers with exp(sign 2 pi i q r / R).
template <int R>
__device__ inline void dft(float2 *v, float sign);

template <>
__device__ inline void dft<2>(float2 *v, float)
{
	const float2 a = v[0];
	const float2 b = v[1];
	v[0] = cadd(a, b);
	v[1] = csub(a, b);
}

template <>
__device__ inline void dft<4>(float2 *v, float sign)
{
	const float2 s02 = cadd(v[0], v[2]);
	const float2 d02 = csub(v[0], v[2]);
	const float2 s13 = cadd(v[1], v[3]);
	const float2 d13 = times_i(csub(v[1], v[3]), sign);
	v[0] = cadd(s02, s13);
	v[2] = csub(s02, s13);
	v[1] = cadd(d02, d13);
	v[3] = csub(d02, d13);
}

template <>
__device__ inline void dft<3>(float2 *v, float sign)
{
	const float s3 = 0.86602540378443864676f;
	const float2 t = cadd(v[1], v[2]);
	const float2 d = times_i(csub(v[1], v[2]), sign * s3);
	const float2 m = make_float2(v[0].x - 0.5f * t.x, v[0].y - 0.5f * t.y);
	v[0] = cadd(v[0], t);
	v[1] = cadd(m, d);
	v[2] = csub(m, d);
}

template <>
__device__ inline void dft<5>(float2 *v, float sign)
{
	const float c1 = 0.30901699437494742410f;  // cos(2 pi / 5)
	const float c2 = -0.80901699437494742410f; // cos(4 pi / 5)
	const float s1 = 0.95105651629515357212f;  // sin(2 pi / 5)
	const float s2 = 0.58778525229247312917f;  // sin(4 pi / 5)
	const float2 t1 = cadd(v[1], v[4]);
	const float2 t2 = cadd(v[2], v[3]);
	const float2 t3 = csub(v[1], v[4]);
	const float2 t4 = csub(v[2], v[3]);
	const float2 a1 = make_float2(v[0].x + c1 * t1.x + c2 * t2.x, v[0].y + c1 * t1.y + c2 * t2.y);
	const float2 a2 = make_float2(v[0].x + c2 * t1.x + c1 * t2.x, v[0].y + c2 * t1.y + c1 * t2.y);
	const float2 b1 = times_i(make_float2(s1 * t3.x + s2 * t4.x, s1 * t3.y + s2 * t4.y), sign);
	const float2 b2 = times_i(make_float2(s2 * t3.x - s1 * t4.x, s2 * t3.y - s1 * t4.y), sign);
	v[0] = cadd(v[0], cadd(t1, t2));
	v[1] = cadd(a1, b1);
	v[4] = csub(a1, b1);
	v[2] = cadd(a2, b2);
	v[3] = csub(a2, b2);
}

// 256-point lines by the four-step decomposition 256 = 16 x 16:
//   X[k1 + 16 k2] = sum_n1 W16^(n1 k2) W256^(n1 k1) sum_n2 x[n1 + 16 n2] W16^(n2 k1).
// Thread (line l, n1) holds x[n1 + 16 n2], n2 = 0..15, and computes the inner
// 16-point DFT in registers (radix 4 x 4), multiplies by W256^(n1 k1), and
// hands the values over through LDS to thread (l, k1), which computes the
// outer 16-point DFT. One LDS exchange per line instead of four.
constexpr int fft256_lines = 16; // lines per block of 256 threads
// LDS float2 values per line of the exchange: 16 x 17 values (odd pitch within a
// line) plus one, so that lines start 546 dwords apart, 2 banks on from one
// another, and the 16 lines of a wave's accesses spread over all 32 banks.
constexpr int fft256_line_stride = 16 * 17 + 1;

// exp(sign 2 pi i m / 16) for m = 0..9 (the products n1 k1 of a 4 x 4 split).
__device__ inline float2 unit16(int m, float sign)
{
	constexpr float c1 = 0.923879532511286756128f; // cos(pi / 8)
	constexpr float s1 = 0.382683432365089771728f; // sin(pi / 8)
	constexpr float c2 = 0.707106781186547524401f; // cos(pi / 4)
	float c = 1.0f;
	float s = 0.0f;
	switch (m)
	{
	case 1: c = c1; s = s1; break;
	case 2: c = c2; s = c2; break;
	case 3: c = s1; s = c1; break;
	case 4: c = 0.0f; s = 1.0f; break;
	case 6: c = -c2; s = c2; break;
	case 9: c = -c1; s = -s1; break;
	default: break;
	}
	return make_float2(c, sign * s);
}

// In-register 16-point DFT: v[k] = sum_n v[n] exp(sign 2 pi i n k / 16).
__device__ inline void dft16(float2 *v, float sign)
{
	float2 a[16];
#pragma unroll
	for (int n1 = 0; n1 < 4; ++n1)
	{
		float2 q[4] = {v[n1], v[n1 + 4], v[n1 + 8], v[n1 + 12]};
		dft<4>(q, sign);
#pragma unroll
		for (int k1 = 0; k1 < 4; ++k1)
		{
			a[4 * n1 + k1] = (n1 * k1 == 0) ? q[k1] : cmul(q[k1], unit16(n1 * k1, sign));
		}
	}
#pragma unroll
	for (int k1 = 0; k1 < 4; ++k1)
	{
		float2 q[4] = {a[k1], a[4 + k1], a[8 + k1], a[12 + k1]};
		dft<4>(q, sign);
#pragma unroll
		for (int k2 = 0; k2 < 4; ++k2)
		{
			v[k1 + 4 * k2] = q[k2];
		}
	}
}

// The four-step transform of one 256-point line in a block of 256 threads,
// 16 lines per block: thread (l, j) enters with v[k] = x_l[j + 16 k] and
// leaves with v[k2] = X_l[j + 16 k2]. `line` is line l's 16 x 17 slice of
// the block's exchange buffer; `twl` holds exp(-2 pi i m / 256), m < 256.
// Every thread of the block must call it (it holds a barrier).
__device__ inline void fft256_four_step(float2 *v, float2 *line, const float2 *twl, int j, float sign)
{
	dft16(v, sign);
#pragma unroll
	for (int k1 = 0; k1 < 16; ++k1)
	{
		float2 w = twl[(j * k1) & 255];
		w.y *= -sign; // the table holds exp(-2 pi i m / 256)
		line[k1 * 17 + j] = k1 == 0 ? v[k1] : cmul(v[k1], w);
	}
	__syncthreads();
#pragma unroll
	for (int n1 = 0; n1 < 16; ++n1)
	{
		v[n1] = line[j * 17 + n1];
	}
	dft16(v, sign);
}

// The chip's eight XCDs take workgroups in turn (workgroup b runs on XCD
// b mod 8), each behind its own L2. With a grid that is a multiple of 8,
// workgroup b takes logical block (b mod 8) (G / 8) + b / 8: each XCD then
// works through one contiguous range of logical blocks, so neighbouring
// blocks, whose lines share cache lines, meet in one L2. Other grids keep
// b.
constexpr unsigned xcd_count = 8;

__device__ inline unsigned xcd_block(unsigned b, unsigned grid)
{
	return grid % xcd_count == 0 ? (b % xcd_count) * (grid / xcd_count) + b / xcd_count : b;
}

__host__ inline unsigned xcd_grid(unsigned blocks)
{
	return (blocks + xcd_count - 1) / xcd_count * xcd_count;
}

} // namespace rexlib_hip
