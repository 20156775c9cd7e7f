// Register-level DFT pieces shared by the FFT kernels and the fused column
// pass of the image insertion.
#pragma once

#include "common.hpp"
#include "fft.hpp"

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

// Stockham stages over `lines` lines of an LDS buffer, ping-pong between a
// and b; the planned form for any length of radices 4, 2, 3, 5, the fixed
// form for the powers of two with compile-time indices.
template <int R>
__device__ inline void stage(const float2 *a, float2 *b, int lines, int n, int line_stride, int ns,
	const float2 *tw, float sign)
{
	const int m = n / R;
	const int twiddle_step = n / (ns * R);
	for (int t = threadIdx.x; t < lines * m; t += blockDim.x)
	{
		const int line = t / m;
		const int j = t - line * m;
		const int k = j % ns;
		const float2 *src = a + line * line_stride;
		float2 v[R];
#pragma unroll
		for (int r = 0; r < R; ++r)
		{
			v[r] = src[j + r * m];
		}
#pragma unroll
		for (int r = 1; r < R; ++r)
		{
			float2 w = tw[k * r * twiddle_step];
			w.y *= -sign; // the table holds exp(-2 pi i m / n)
			v[r] = cmul(v[r], w);
		}
		dft<R>(v, sign);
		const int d = (j / ns) * ns * R + k;
		float2 *dst = b + line * line_stride;
#pragma unroll
		for (int r = 0; r < R; ++r)
		{
			dst[d + r * ns] = v[r];
		}
	}
}

// Transform `lines` lines of LDS buffer a (line stride line_stride); returns
// the buffer holding the result. sign -1: forward, +1: inverse (unscaled).
__device__ inline float2 *block_fft(float2 *a, float2 *b, int lines, int line_stride, const fft_plan &plan,
	const float2 *tw, float sign)
{
	int ns = 1;
	const int n = plan.n;
	for (int st = 0; st < plan.stages; ++st)
	{
		const int r = plan.radix[st];
		switch (r)
		{
		case 4: stage<4>(a, b, lines, n, line_stride, ns, tw, sign); break;
		case 2: stage<2>(a, b, lines, n, line_stride, ns, tw, sign); break;
		case 3: stage<3>(a, b, lines, n, line_stride, ns, tw, sign); break;
		default: stage<5>(a, b, lines, n, line_stride, ns, tw, sign); break;
		}
		__syncthreads();
		float2 *swap = a;
		a = b;
		b = swap;
		ns *= r;
	}
	return a;
}

// The same transform for a power-of-two length N known at compile time:
// radix 4 while it divides, then one radix 2 (the order make_fft_plan gives),
// with every index computed by shifts and masks.
template <int N, int R, int NS>
__device__ inline void stage_fixed(const float2 *a, float2 *b, int lines, int line_stride, const float2 *tw,
	float sign)
{
	constexpr int m = N / R;
	constexpr int twiddle_step = N / (NS * R);
	for (int t = threadIdx.x; t < lines * m; t += blockDim.x)
	{
		const int line = t / m;
		const int j = t % m;
		const int k = j % NS;
		const float2 *src = a + line * line_stride;
		float2 v[R];
#pragma unroll
		for (int r = 0; r < R; ++r)
		{
			v[r] = src[j + r * m];
		}
#pragma unroll
		for (int r = 1; r < R; ++r)
		{
			float2 w = tw[k * r * twiddle_step];
			w.y *= -sign; // the table holds exp(-2 pi i m / n)
			v[r] = cmul(v[r], w);
		}
		dft<R>(v, sign);
		const int d = (j / NS) * NS * R + k;
		float2 *dst = b + line * line_stride;
#pragma unroll
		for (int r = 0; r < R; ++r)
		{
			dst[d + r * NS] = v[r];
		}
	}
}

template <int N, int NS = 1>
__device__ inline float2 *block_fft_fixed(float2 *a, float2 *b, int lines, int line_stride, const float2 *tw,
	float sign)
{
	if constexpr (NS >= N)
	{
		return a;
	}
	else
	{
		constexpr int R = (N / NS) % 4 == 0 ? 4 : 2;
		stage_fixed<N, R, NS>(a, b, lines, line_stride, tw, sign);
		__syncthreads();
		return block_fft_fixed<N, NS * R>(b, a, lines, line_stride, tw, sign);
	}
}

// The fixed form for N > 0, the planned one otherwise.
template <int N>
__device__ inline float2 *transform(float2 *a, float2 *b, int lines, int line_stride, const fft_plan &plan,
	const float2 *tw, float sign)
{
	if constexpr (N > 0)
	{
		return block_fft_fixed<N>(a, b, lines, line_stride, tw, sign);
	}
	else
	{
		return block_fft(a, b, lines, line_stride, plan, tw, sign);
	}
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
