// Elementwise programs the reconstruction's quotient is written with:
// fill (ops::fill_operation, src/functional/creation.cpp:301-349),
// divide (include/rexlib/ops/arithmetic/divide_operation.hpp),
// greater (include/rexlib/ops/compare/greater_operation.hpp) and
// where (include/rexlib/ops/selection/where_operation.hpp), over strided
// operands that broadcast against the result as numpy's rules do (an axis of
// extent 1, or a missing leading axis, repeats).
//
// A thread takes one result element per grid-stride step; the element's
// offset in every operand comes from its multi-index, computed once.
#include "launch.hpp"

#include <cmath>
#include <cstring>
#include <string>

namespace rexlib_hip
{

namespace
{

constexpr int max_rank = REXLIB_HIP_MAX_RANK;

// Strides of up to four operands aligned to the result's axes (0 where an
// operand broadcasts).
struct walk
{
	int32_t rank;
	int64_t count;
	int64_t extents[max_rank];
	int64_t strides[4][max_rank];
};

__device__ inline void offsets_of(const walk &w, int64_t i, int operands, int64_t *out)
{
	if (w.rank == 1)
	{
		// Dense operands (and scalars) collapse to one axis: no division.
		for (int k = 0; k < operands; ++k)
		{
			out[k] = i * w.strides[k][0];
		}
		return;
	}
	for (int k = 0; k < operands; ++k)
	{
		out[k] = 0;
	}
	for (int a = w.rank - 1; a >= 0; --a)
	{
		const int64_t e = w.extents[a];
		const int64_t q = i / e;
		const int64_t r = i - q * e;
		i = q;
		for (int k = 0; k < operands; ++k)
		{
			out[k] += r * w.strides[k][a];
		}
	}
}

template <typename T>
__global__ void fill_kernel(T *out, walk w, T value)
{
	for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < w.count;
		i += static_cast<int64_t>(gridDim.x) * blockDim.x)
	{
		int64_t o[1];
		offsets_of(w, i, 1, o);
		out[o[0]] = value;
	}
}

__device__ inline float divide_of(float a, float b)
{
	return a / b;
}

// (a + bi) / (c + di) with the larger of |c|, |d| scaled out first (Smith's
// method), so that neither product overflows for finite operands.
__device__ inline float2 divide_of(float2 a, float2 b)
{
	if (fabsf(b.x) >= fabsf(b.y))
	{
		const float r = b.y / b.x;
		const float d = b.x + b.y * r;
		return make_float2((a.x + a.y * r) / d, (a.y - a.x * r) / d);
	}
	const float r = b.x / b.y;
	const float d = b.x * r + b.y;
	return make_float2((a.x * r + a.y) / d, (a.y * r - a.x) / d);
}

template <typename T>
__global__ void divide_kernel(T *out, const T *a, const T *b, walk w)
{
	for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < w.count;
		i += static_cast<int64_t>(gridDim.x) * blockDim.x)
	{
		int64_t o[3];
		offsets_of(w, i, 3, o);
		out[o[0]] = divide_of(a[o[1]], b[o[2]]);
	}
}

__global__ void greater_kernel(uint8_t *out, const float *a, const float *b, walk w)
{
	for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < w.count;
		i += static_cast<int64_t>(gridDim.x) * blockDim.x)
	{
		int64_t o[3];
		offsets_of(w, i, 3, o);
		out[o[0]] = a[o[1]] > b[o[2]] ? 1 : 0;
	}
}

template <typename T>
__global__ void where_kernel(T *out, const uint8_t *condition, const T *x, const T *y, walk w)
{
	for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < w.count;
		i += static_cast<int64_t>(gridDim.x) * blockDim.x)
	{
		int64_t o[4];
		offsets_of(w, i, 4, o);
		out[o[0]] = condition[o[1]] != 0 ? x[o[2]] : y[o[3]];
	}
}

unsigned blocks_for(int64_t count)
{
	return static_cast<unsigned>(std::max<int64_t>(1, std::min<int64_t>((count + 255) / 256, 8192)));
}

// The walk of a result and its operands: each operand's axes align to the
// result's last ones, and each of its extents equals the result's or is 1.
walk make_walk(const char *what, const rexlib_hip_array &result, const rexlib_hip_array *const *operands, int count)
{
	walk w{};
	w.rank = result.rank;
	w.count = 1;
	for (int a = 0; a < result.rank; ++a)
	{
		w.extents[a] = result.extents[a];
		w.count *= result.extents[a];
		w.strides[0][a] = result.strides[a];
	}
	for (int k = 0; k < count; ++k)
	{
		const rexlib_hip_array &x = *operands[k];
		if (x.rank > result.rank)
		{
			fail(REXLIB_HIP_INVALID_ARGUMENT, std::string(what) + ": an operand has more axes than the result");
		}
		const int lead = result.rank - x.rank;
		for (int a = 0; a < result.rank; ++a)
		{
			if (a < lead)
			{
				w.strides[k + 1][a] = 0;
				continue;
			}
			const int64_t e = x.extents[a - lead];
			if (e != result.extents[a] && e != 1)
			{
				fail(REXLIB_HIP_INVALID_ARGUMENT, std::string(what) + ": operand shapes do not broadcast to the result");
			}
			w.strides[k + 1][a] = e == 1 ? 0 : x.strides[a - lead];
		}
	}
	// Merge axis a into axis a + 1 wherever every operand steps across the
	// pair as across one axis (a broadcast operand steps 0 across both), and
	// drop axes of extent 1: dense operands of one shape become one axis.
	int rank = 0;
	walk c = w;
	for (int a = 0; a < w.rank; ++a)
	{
		if (w.extents[a] == 1 && w.rank > 1)
		{
			continue;
		}
		bool merge = rank > 0;
		for (int k = 0; k <= count && merge; ++k)
		{
			merge = c.strides[k][rank - 1] == w.strides[k][a] * w.extents[a];
		}
		if (merge)
		{
			c.extents[rank - 1] *= w.extents[a];
			for (int k = 0; k <= count; ++k)
			{
				c.strides[k][rank - 1] = w.strides[k][a];
			}
			continue;
		}
		c.extents[rank] = w.extents[a];
		for (int k = 0; k <= count; ++k)
		{
			c.strides[k][rank] = w.strides[k][a];
		}
		++rank;
	}
	if (rank == 0)
	{
		// Every axis of extent 1: one element.
		c.extents[0] = 1;
		for (int k = 0; k <= count; ++k)
		{
			c.strides[k][0] = 0;
		}
		rank = 1;
	}
	c.rank = rank;
	return c;
}

void check_operand(const rexlib_hip_array *a, const char *what)
{
	if (a == nullptr)
	{
		fail(REXLIB_HIP_INVALID_ARGUMENT, std::string(what) + ": missing operand");
	}
	if (a->rank < 0 || a->rank > max_rank)
	{
		fail(REXLIB_HIP_INVALID_ARGUMENT, std::string(what) + ": rank beyond " + std::to_string(max_rank));
	}
	bool empty = false;
	for (int i = 0; i < a->rank; ++i)
	{
		if (a->extents[i] < 0)
		{
			fail(REXLIB_HIP_INVALID_ARGUMENT, std::string(what) + ": negative extent");
		}
		empty = empty || a->extents[i] == 0;
	}
	// An operand with no elements may have no storage.
	if (a->data == nullptr && !empty)
	{
		fail(REXLIB_HIP_INVALID_ARGUMENT, std::string(what) + ": missing operand");
	}
}

} // anonymous namespace

} // namespace rexlib_hip

using namespace rexlib_hip;

extern "C" {

int rexlib_hip_fill(const rexlib_hip_array *out, double real, double imaginary, void *stream)
{
	return guarded([&]
	{
		check_operand(out, "fill");
		const walk w = make_walk("fill", *out, nullptr, 0);
		if (w.count == 0)
		{
			return;
		}
		const hipStream_t s = static_cast<hipStream_t>(stream);
		const unsigned blocks = blocks_for(w.count);
		switch (out->dtype)
		{
		case REXLIB_HIP_FLOAT32:
			fill_kernel<float><<<blocks, 256, 0, s>>>(static_cast<float *>(out->data), w, static_cast<float>(real));
			break;
		case REXLIB_HIP_FLOAT64:
			fill_kernel<double><<<blocks, 256, 0, s>>>(static_cast<double *>(out->data), w, real);
			break;
		case REXLIB_HIP_COMPLEX64:
			fill_kernel<float2><<<blocks, 256, 0, s>>>(static_cast<float2 *>(out->data), w,
				make_float2(static_cast<float>(real), static_cast<float>(imaginary)));
			break;
		case REXLIB_HIP_COMPLEX128:
			fill_kernel<double2><<<blocks, 256, 0, s>>>(static_cast<double2 *>(out->data), w,
				make_double2(real, imaginary));
			break;
		case REXLIB_HIP_FLOAT16:
		case REXLIB_HIP_BFLOAT16:
		{
			uint16_t bits = 0;
			if (out->dtype == REXLIB_HIP_FLOAT16)
			{
				const _Float16 h = static_cast<_Float16>(real);
				std::memcpy(&bits, &h, sizeof(bits));
			}
			else
			{
				// Round to nearest even from float32.
				const float f = static_cast<float>(real);
				uint32_t u = 0;
				std::memcpy(&u, &f, sizeof(u));
				bits = std::isnan(f) ? static_cast<uint16_t>(0x7fc0) :
					static_cast<uint16_t>((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
			}
			fill_kernel<uint16_t><<<blocks, 256, 0, s>>>(static_cast<uint16_t *>(out->data), w, bits);
			break;
		}
		case REXLIB_HIP_BOOL:
			fill_kernel<uint8_t><<<blocks, 256, 0, s>>>(static_cast<uint8_t *>(out->data), w,
				static_cast<uint8_t>(real != 0.0 ? 1 : 0));
			break;
		default:
			fail(REXLIB_HIP_UNSUPPORTED, "fill: no gfx950 program for this element type");
		}
		REXLIB_HIP_CHECK(hipGetLastError());
	});
}

int rexlib_hip_divide(const rexlib_hip_array *dividend, const rexlib_hip_array *divisor,
	const rexlib_hip_array *result, void *stream)
{
	return guarded([&]
	{
		check_operand(dividend, "divide dividend");
		check_operand(divisor, "divide divisor");
		check_operand(result, "divide result");
		if (dividend->dtype != divisor->dtype || dividend->dtype != result->dtype)
		{
			fail(REXLIB_HIP_INVALID_ARGUMENT, "divide: the operands and the result share one element type");
		}
		const rexlib_hip_array *operands[2] = {dividend, divisor};
		const walk w = make_walk("divide", *result, operands, 2);
		if (w.count == 0)
		{
			return;
		}
		const hipStream_t s = static_cast<hipStream_t>(stream);
		switch (result->dtype)
		{
		case REXLIB_HIP_FLOAT32:
			divide_kernel<float><<<blocks_for(w.count), 256, 0, s>>>(static_cast<float *>(result->data),
				static_cast<const float *>(dividend->data), static_cast<const float *>(divisor->data), w);
			break;
		case REXLIB_HIP_COMPLEX64:
			divide_kernel<float2><<<blocks_for(w.count), 256, 0, s>>>(static_cast<float2 *>(result->data),
				static_cast<const float2 *>(dividend->data), static_cast<const float2 *>(divisor->data), w);
			break;
		default:
			fail(REXLIB_HIP_UNSUPPORTED, "divide: the gfx950 program takes float32 and complex64");
		}
		REXLIB_HIP_CHECK(hipGetLastError());
	});
}

int rexlib_hip_greater(const rexlib_hip_array *a, const rexlib_hip_array *b, const rexlib_hip_array *result,
	void *stream)
{
	return guarded([&]
	{
		check_operand(a, "greater a");
		check_operand(b, "greater b");
		check_operand(result, "greater result");
		if (result->dtype != REXLIB_HIP_BOOL)
		{
			fail(REXLIB_HIP_INVALID_ARGUMENT, "greater: the result is boolean");
		}
		if (a->dtype != b->dtype)
		{
			fail(REXLIB_HIP_INVALID_ARGUMENT, "greater: the operands share one element type");
		}
		if (a->dtype != REXLIB_HIP_FLOAT32)
		{
			fail(REXLIB_HIP_UNSUPPORTED, "greater: the gfx950 program takes float32");
		}
		const rexlib_hip_array *operands[2] = {a, b};
		const walk w = make_walk("greater", *result, operands, 2);
		if (w.count == 0)
		{
			return;
		}
		greater_kernel<<<blocks_for(w.count), 256, 0, static_cast<hipStream_t>(stream)>>>(
			static_cast<uint8_t *>(result->data), static_cast<const float *>(a->data),
			static_cast<const float *>(b->data), w);
		REXLIB_HIP_CHECK(hipGetLastError());
	});
}

int rexlib_hip_where(const rexlib_hip_array *condition, const rexlib_hip_array *x, const rexlib_hip_array *y,
	const rexlib_hip_array *result, void *stream)
{
	return guarded([&]
	{
		check_operand(condition, "where condition");
		check_operand(x, "where x");
		check_operand(y, "where y");
		check_operand(result, "where result");
		if (condition->dtype != REXLIB_HIP_BOOL)
		{
			fail(REXLIB_HIP_INVALID_ARGUMENT, "where: the condition is boolean");
		}
		if (x->dtype != y->dtype || x->dtype != result->dtype)
		{
			fail(REXLIB_HIP_INVALID_ARGUMENT, "where: x, y and the result share one element type");
		}
		const rexlib_hip_array *operands[3] = {condition, x, y};
		const walk w = make_walk("where", *result, operands, 3);
		if (w.count == 0)
		{
			return;
		}
		const hipStream_t s = static_cast<hipStream_t>(stream);
		const uint8_t *c = static_cast<const uint8_t *>(condition->data);
		switch (result->dtype)
		{
		case REXLIB_HIP_FLOAT32:
			where_kernel<float><<<blocks_for(w.count), 256, 0, s>>>(static_cast<float *>(result->data), c,
				static_cast<const float *>(x->data), static_cast<const float *>(y->data), w);
			break;
		case REXLIB_HIP_COMPLEX64:
			where_kernel<float2><<<blocks_for(w.count), 256, 0, s>>>(static_cast<float2 *>(result->data), c,
				static_cast<const float2 *>(x->data), static_cast<const float2 *>(y->data), w);
			break;
		default:
			fail(REXLIB_HIP_UNSUPPORTED, "where: the gfx950 program takes float32 and complex64");
		}
		REXLIB_HIP_CHECK(hipGetLastError());
	});
}

} // extern "C"
