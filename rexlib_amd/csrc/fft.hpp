// Batched FFTs of the path: line descriptors and the block transform.
#pragma once

#include "common.hpp"

namespace rexlib_hip
{

constexpr int max_fft_stages = 16;

// A factorisation of the length into radices 4, 2, 3, 5.
struct fft_plan
{
	int32_t n;
	int32_t stages;
	int32_t radix[max_fft_stages];
};

// Lines of a batched transform: the line index l decomposes into
// (outer, middle, inner) = (l / (cm * ci), (l / ci) % cm, l % ci), and its
// first element and element stride are, in elements,
//   outer * s_outer + middle * s_middle + inner * s_inner, s_elem.
struct line_set
{
	int64_t count;
	int64_t middle_count;
	int64_t inner_count;
	int64_t s_outer;
	int64_t s_middle;
	int64_t s_inner;
	int64_t s_elem;
};

__host__ __device__ inline int64_t line_offset(const line_set &l, int64_t line)
{
	const int64_t inner = line % l.inner_count;
	const int64_t rest = line / l.inner_count;
	const int64_t middle = rest % l.middle_count;
	const int64_t outer = rest / l.middle_count;
	return outer * l.s_outer + middle * l.s_middle + inner * l.s_inner;
}

fft_plan make_fft_plan(int64_t n);

// Device twiddles exp(-2 pi i m / n), m < n, computed in double, cached per
// device and length.
const float2 *twiddles(int64_t n);

// Lines per block of the LDS transforms of length n.
int lines_per_group(int n);

// Stage callers: the element type of real lines.
enum real_kind : int
{
	real_f32 = 0,
	real_bf16 = 1,
	real_f16 = 2
};

} // namespace rexlib_hip
