// Background standardization of images
// (include/rexlib/em/ops/preprocessing/standardize_background_operation.hpp:15-76;
// CPU kernels src/backends/cpu/builders/em/preprocessing/).
//
// One workgroup per image (the value's part along the image axes; every
// other axis is a batch axis). The workgroup reads its image three times:
//   1. count, sum and sum of squares, in double, of the finite background
//      elements -- those whose squared distance (in elements) from the
//      centre, index n / 2 along each image axis, exceeds radius^2
//      (background_geometry.inl:46-55) -- giving a mean and 1 / deviation
//      (population), or (0, 1) without any (background_moments_kernel.inl:117-135);
//   2. the same over the finite background elements whose value,
//      standardized with the first statistics in double, lies within the
//      bounds (background_moments_kernel.inl:206-213);
//   3. result = (x - mean) * scale in the element type, with mean and scale
//      rounded to it; an element that is not finite or lies outside the
//      bounds takes a standard normal draw instead
//      (standardization_map_kernel.inl:60-131).
// The draw of counter c is the Box-Muller pair of the Philox4x32-10 block
// (c / 2, 0, 0, 0) keyed by the image's key, its cosine branch for even c
// and its sine branch for odd c (src/backends/cpu/kernels/standard_normal.inl:20-52).
// The counter of an element is its row-major position among the elements
// sharing its key: the image axes and the batch axes the key broadcasts
// along, in the value's axis order.
//
// Two kernels run these steps. The resident kernel takes float32 images
// that are the value's last two axes, stored densely with a multiple of
// four columns, 16-byte aligned, of at most 65,536 elements: a workgroup
// loads its image once, 16 bytes a lane, keeps it in registers through both
// statistics and writes the result, so the image crosses HBM once each way.
// The general kernel takes every other layout and rank and reads its image
// three times.
#include "launch.hpp"

#include <algorithm>
#include <cmath>
#include <limits>
#include <string>

namespace rexlib_hip
{

namespace
{

constexpr int max_rank = REXLIB_HIP_MAX_RANK;
constexpr int threads = 256;

struct standardize_plan
{
	int32_t image_rank;
	int32_t batch_rank;
	int64_t image_count;
	int64_t image_elements;
	// Image axes, in the value's order.
	int64_t image_extent[max_rank];
	int64_t image_value_stride[max_rank];
	int64_t image_result_stride[max_rank];
	int64_t image_counter_stride[max_rank];
	int64_t centre[max_rank];
	// Batch axes, in the value's order.
	int64_t batch_extent[max_rank];
	int64_t batch_value_stride[max_rank];
	int64_t batch_result_stride[max_rank];
	int64_t batch_key_stride[max_rank];
	int64_t batch_counter_stride[max_rank];
	double squared_radius;
	double lower_bound;
	double upper_bound;
};

// Philox4x32-10 (Salmon, Moraes, Dror and Shaw, "Parallel random numbers: as
// easy as 1, 2, 3", SC 2011): ten rounds of two 32 x 32 -> 64 bit products,
// the key bumped by the Weyl constants between rounds.
__device__ inline uint4 philox4x32_10(uint4 c, uint2 k)
{
	constexpr uint32_t m0 = 0xD2511F53u;
	constexpr uint32_t m1 = 0xCD9E8D57u;
	constexpr uint32_t w0 = 0x9E3779B9u;
	constexpr uint32_t w1 = 0xBB67AE85u;
	for (int round = 0; round < 10; ++round)
	{
		if (round > 0)
		{
			k.x += w0;
			k.y += w1;
		}
		const uint32_t hi0 = __umulhi(m0, c.x);
		const uint32_t lo0 = m0 * c.x;
		const uint32_t hi1 = __umulhi(m1, c.z);
		const uint32_t lo1 = m1 * c.z;
		c = make_uint4(hi1 ^ c.y ^ k.x, lo1, hi0 ^ c.w ^ k.y, lo0);
	}
	return c;
}

// 64 random bits to a uniform value in (0, 1]: the top 53 bits plus one,
// times 2^-53 (standard_normal.inl:20-25).
__device__ inline double uniform_open_closed(uint64_t bits)
{
	return static_cast<double>((bits >> 11) + 1) * (1.0 / 9007199254740992.0);
}

// The standard normal value of counter c for a key.
__device__ inline double standard_normal(uint64_t key, uint64_t counter)
{
	const uint64_t block = counter / 2;
	const uint4 r = philox4x32_10(
		make_uint4(static_cast<uint32_t>(block), static_cast<uint32_t>(block >> 32), 0u, 0u),
		make_uint2(static_cast<uint32_t>(key), static_cast<uint32_t>(key >> 32)));
	const uint64_t first = (static_cast<uint64_t>(r.y) << 32) | r.x;
	const uint64_t second = (static_cast<uint64_t>(r.w) << 32) | r.z;
	const double radius = sqrt(-2.0 * log(uniform_open_closed(first)));
	const double angle = 6.283185307179586476925 * uniform_open_closed(second);
	return (counter & 1) == 0 ? radius * cos(angle) : radius * sin(angle);
}

__device__ inline bool finite_double(double x)
{
	return fabs(x) <= 1.7976931348623157e308;
}

// Sum of (count, sum, sum of squares) over a workgroup of `block` lanes.
template <int block = threads>
__device__ inline void block_sum(double &count, double &sum, double &squares, double *scratch)
{
	for (int offset = 32; offset > 0; offset /= 2)
	{
		count += __shfl_down(count, offset, 64);
		sum += __shfl_down(sum, offset, 64);
		squares += __shfl_down(squares, offset, 64);
	}
	const int lane = threadIdx.x & 63;
	const int wave = threadIdx.x >> 6;
	if (lane == 0)
	{
		scratch[3 * wave] = count;
		scratch[3 * wave + 1] = sum;
		scratch[3 * wave + 2] = squares;
	}
	__syncthreads();
	// Lane w of every wavefront takes wave w's partial sums; a butterfly
	// over those lanes and a broadcast of lane 0 give every lane the total
	// (the partials never sit in registers all at once).
	constexpr int waves = block / 64;
	count = lane < waves ? scratch[3 * lane] : 0.0;
	sum = lane < waves ? scratch[3 * lane + 1] : 0.0;
	squares = lane < waves ? scratch[3 * lane + 2] : 0.0;
	for (int offset = waves / 2; offset > 0; offset /= 2)
	{
		count += __shfl_xor(count, offset, 64);
		sum += __shfl_xor(sum, offset, 64);
		squares += __shfl_xor(squares, offset, 64);
	}
	count = __shfl(count, 0, 64);
	sum = __shfl(sum, 0, 64);
	squares = __shfl(squares, 0, 64);
	__syncthreads();
}

template <typename T>
__global__ __launch_bounds__(threads) void standardize_kernel(const T *value, const uint64_t *key, T *result,
	standardize_plan p)
{
	__shared__ double scratch[3 * (threads / 64)];
	for (int64_t image = blockIdx.x; image < p.image_count; image += gridDim.x)
	{
		int64_t value_base = 0, result_base = 0, key_offset = 0, counter_base = 0;
		int64_t rest = image;
		for (int a = p.batch_rank - 1; a >= 0; --a)
		{
			const int64_t i = rest % p.batch_extent[a];
			rest /= p.batch_extent[a];
			value_base += i * p.batch_value_stride[a];
			result_base += i * p.batch_result_stride[a];
			key_offset += i * p.batch_key_stride[a];
			counter_base += i * p.batch_counter_stride[a];
		}
		const uint64_t image_key = key[key_offset];

		double mean = 0.0;
		double scale = 1.0;
		for (int pass = 0; pass < 2; ++pass)
		{
			const double lower = pass == 0 ? -INFINITY : p.lower_bound;
			const double upper = pass == 0 ? INFINITY : p.upper_bound;
			double count = 0.0, sum = 0.0, squares = 0.0;
			for (int64_t e = threadIdx.x; e < p.image_elements; e += threads)
			{
				int64_t offset = 0, squared_distance = 0;
				int64_t r = e;
				for (int a = p.image_rank - 1; a >= 0; --a)
				{
					const int64_t i = r % p.image_extent[a];
					r /= p.image_extent[a];
					offset += i * p.image_value_stride[a];
					const int64_t d = i - p.centre[a];
					squared_distance += d * d;
				}
				const double x = static_cast<double>(value[value_base + offset]);
				const double standardized = (x - mean) * scale;
				const bool valid = static_cast<double>(squared_distance) > p.squared_radius && finite_double(x) &&
					standardized >= lower && standardized <= upper;
				count += valid ? 1.0 : 0.0;
				sum += valid ? x : 0.0;
				squares += valid ? x * x : 0.0;
			}
			block_sum(count, sum, squares, scratch);
			if (count == 0.0)
			{
				mean = 0.0;
				scale = 1.0;
			}
			else
			{
				mean = sum / count;
				const double variance = squares / count - mean * mean;
				scale = variance > 0.0 ? 1.0 / sqrt(variance) : 1.0;
			}
		}

		const T mean_t = static_cast<T>(mean);
		const T scale_t = static_cast<T>(scale);
		const T lower_t = static_cast<T>(p.lower_bound);
		const T upper_t = static_cast<T>(p.upper_bound);
		for (int64_t e = threadIdx.x; e < p.image_elements; e += threads)
		{
			int64_t in = 0, out = 0, counter = counter_base;
			int64_t r = e;
			for (int a = p.image_rank - 1; a >= 0; --a)
			{
				const int64_t i = r % p.image_extent[a];
				r /= p.image_extent[a];
				in += i * p.image_value_stride[a];
				out += i * p.image_result_stride[a];
				counter += i * p.image_counter_stride[a];
			}
			T standardized;
			{
#pragma clang fp contract(off)
				standardized = (value[value_base + in] - mean_t) * scale_t;
			}
			const bool kept = standardized >= lower_t && standardized <= upper_t &&
				fabs(static_cast<double>(standardized)) <= static_cast<double>(std::numeric_limits<T>::max());
			result[result_base + out] = kept ? standardized :
				static_cast<T>(standard_normal(image_key, static_cast<uint64_t>(counter)));
		}
	}
}

// The mean and 1 / deviation (population) of (count, sum, sum of squares),
// or (0, 1) without any element (background_moments_kernel.inl:117-135).
__device__ inline void moments(double count, double sum, double squares, double &mean, double &scale)
{
	if (count == 0.0)
	{
		mean = 0.0;
		scale = 1.0;
		return;
	}
	mean = sum / count;
	const double variance = squares / count - mean * mean;
	scale = variance > 0.0 ? 1.0 / sqrt(variance) : 1.0;
}

// standard_normal out of line for the resident kernel: inlined, its
// polynomial constants are hoisted out of the image loop into registers
// that the held image needs.
__device__ __noinline__ float draw_float(uint64_t key, uint64_t counter)
{
	return static_cast<float>(standard_normal(key, counter));
}

// An empty register constraint: the compiler takes `v` as rewritten here, so
// it converts the held floats to double afresh in the second statistics
// instead of keeping the first statistics' doubles (twice the registers)
// alive across the workgroup sum.
__device__ inline void rewritten(float4 &v)
{
	asm volatile("" : "+v"(v.x), "+v"(v.y), "+v"(v.z), "+v"(v.w));
}

// Geometry of the resident kernel's images: `rows` x `cols` float32
// elements stored densely (row-major, no gaps), `cols` a multiple of 4,
// rows * cols <= 4 * block * quads. An element is background when its
// squared distance from (rows / 2, cols / 2) exceeds `threshold` (the
// squared radius rounded down, clamped to INT32_MAX: for an integer d2,
// d2 > r^2 exactly when d2 > floor(r^2)).
struct resident_images
{
	int32_t rows;
	int32_t cols;
	int32_t threshold;
};

// One workgroup of `block` lanes per image; lane t holds the 16-byte
// groups t, t + block, ... (`quads` of them): the first quads - staged in
// registers, the last `staged` in LDS (each lane reads back only what it
// wrote, so no barrier orders them). Per element the three steps
// are those of standardize_kernel; the background and finiteness of each
// held element is kept as one bit across the two statistics.
template <int block, int quads, int staged>
__global__ __launch_bounds__(block) void standardize_resident_kernel(const float *value, const uint64_t *key,
	float *result, standardize_plan p, resident_images g)
{
	static_assert(4 * quads <= 64, "the background bits of a lane fit 64 bits");
	static_assert(staged < quads, "some groups stay in registers");
	constexpr int kept_quads = quads - staged;
	__shared__ double scratch[3 * (block / 64)];
	__shared__ float4 stage[staged > 0 ? staged * block : 1];
	const int groups_per_row = g.cols / 4;
	const int groups = g.rows * groups_per_row;
	const int centre_row = g.rows / 2;
	const int centre_col = g.cols / 2;
	for (int64_t image = blockIdx.x; image < p.image_count; image += gridDim.x)
	{
		int64_t value_base = 0, result_base = 0, key_offset = 0, counter_base = 0;
		int64_t rest = image;
		for (int a = p.batch_rank - 1; a >= 0; --a)
		{
			const int64_t i = rest % p.batch_extent[a];
			rest /= p.batch_extent[a];
			value_base += i * p.batch_value_stride[a];
			result_base += i * p.batch_result_stride[a];
			key_offset += i * p.batch_key_stride[a];
			counter_base += i * p.batch_counter_stride[a];
		}

		// The lane index passes through an empty register constraint once
		// per image, so the compiler recomputes the per-group geometry in
		// each image rather than hoisting it out of the image loop into
		// registers the held image needs.
		int lane = threadIdx.x;
		asm volatile("" : "+v"(lane));
		// Groups past the image's end load its last group again and count
		// nowhere (no branch, so the held image stays in registers).
		// Every group's load is issued before any is used.
		float4 loaded[quads];
		const float4 *in = reinterpret_cast<const float4 *>(value + value_base);
#pragma unroll
		for (int k = 0; k < quads; ++k)
		{
			loaded[k] = in[static_cast<unsigned>(min(lane + k * block, groups - 1))];
		}
		float4 x[kept_quads];
#pragma unroll
		for (int k = 0; k < quads; ++k)
		{
			if (k < kept_quads)
			{
				x[k < kept_quads ? k : 0] = loaded[k];
			}
			else
			{
				stage[(k - kept_quads) * block + lane] = loaded[k];
			}
		}
		uint64_t background = 0;
		double count = 0.0, sum = 0.0, squares = 0.0;
#pragma unroll
		for (int k = 0; k < quads; ++k)
		{
			const int q = lane + k * block;
			const int row = q / groups_per_row;
			const int col = 4 * (q - row * groups_per_row);
			const int dr = row - centre_row;
			const float lanes[4] = {loaded[k].x, loaded[k].y, loaded[k].z, loaded[k].w};
#pragma unroll
			for (int c = 0; c < 4; ++c)
			{
				const int dc = col + c - centre_col;
				const bool valid = (q < groups) & (dr * dr + dc * dc > g.threshold) &
					(fabsf(lanes[c]) <= 3.402823466e38f);
				background |= static_cast<uint64_t>(valid) << (4 * k + c);
				const double v = static_cast<double>(lanes[c]);
				count += valid ? 1.0 : 0.0;
				sum += valid ? v : 0.0;
				squares += valid ? v * v : 0.0;
			}
		}
		block_sum<block>(count, sum, squares, scratch);
		double mean, scale;
		moments(count, sum, squares, mean, scale);

		count = 0.0;
		sum = 0.0;
		squares = 0.0;
#pragma unroll
		for (int k = 0; k < quads; ++k)
		{
			float4 held = k < kept_quads ? x[k < kept_quads ? k : 0] : stage[(k - kept_quads) * block + lane];
			rewritten(held);
			const float lanes[4] = {held.x, held.y, held.z, held.w};
#pragma unroll
			for (int c = 0; c < 4; ++c)
			{
				const double v = static_cast<double>(lanes[c]);
				const double standardized = (v - mean) * scale;
				const bool valid = (((background >> (4 * k + c)) & 1u) != 0) & (standardized >= p.lower_bound) &
					(standardized <= p.upper_bound);
				count += valid ? 1.0 : 0.0;
				sum += valid ? v : 0.0;
				squares += valid ? v * v : 0.0;
			}
		}
		block_sum<block>(count, sum, squares, scratch);
		moments(count, sum, squares, mean, scale);

		const float mean_t = static_cast<float>(mean);
		const float scale_t = static_cast<float>(scale);
		const float lower_t = static_cast<float>(p.lower_bound);
		const float upper_t = static_cast<float>(p.upper_bound);
		// Standardized values go out as held; the elements to be replaced
		// (few: outliers and non-finite values) are marked and then
		// overwritten one by one by the same lane, after the held image is
		// dead.
		uint64_t drawn = 0;
		float4 *out = reinterpret_cast<float4 *>(result + result_base);
		asm volatile("" : "+v"(lane));
#pragma unroll
		for (int k = 0; k < quads; ++k)
		{
			const int q = lane + k * block;
			const float4 held = k < kept_quads ? x[k < kept_quads ? k : 0] : stage[(k - kept_quads) * block + lane];
			float lanes[4] = {held.x, held.y, held.z, held.w};
#pragma unroll
			for (int c = 0; c < 4; ++c)
			{
				float standardized;
				{
#pragma clang fp contract(off)
					standardized = (lanes[c] - mean_t) * scale_t;
				}
				const bool kept = (standardized >= lower_t) & (standardized <= upper_t) &
					(fabsf(standardized) <= 3.402823466e38f);
				drawn |= static_cast<uint64_t>(!kept & (q < groups)) << (4 * k + c);
				lanes[c] = standardized;
			}
			if (q < groups)
			{
				out[static_cast<unsigned>(q)] = make_float4(lanes[0], lanes[1], lanes[2], lanes[3]);
			}
		}
		if (drawn != 0)
		{
			const uint64_t image_key = key[key_offset];
			do
			{
				const int bit = __builtin_ctzll(drawn);
				drawn &= drawn - 1;
				const int e = 4 * (lane + (bit / 4) * block) + bit % 4;
				result[result_base + e] = draw_float(image_key, static_cast<uint64_t>(counter_base + e));
			} while (drawn != 0);
		}
	}
}

template <int block, int quads, int staged = 0>
void launch_resident(unsigned grid, hipStream_t s, const float *value, const uint64_t *key, float *result,
	const standardize_plan &p, const resident_images &g)
{
	standardize_resident_kernel<block, quads, staged><<<grid, block, 0, s>>>(value, key, result, p, g);
}

// Launches the resident kernel when the operands fit it; false otherwise.
bool try_resident(hipStream_t s, const rexlib_hip_array *value, const rexlib_hip_array *result, const uint64_t *key,
	uint32_t image_axes, const standardize_plan &p, unsigned grid)
{
	const int rank = value->rank;
	if (value->dtype != REXLIB_HIP_FLOAT32 || p.image_rank != 2 || rank < 2 || image_axes != (3u << (rank - 2)))
	{
		return false;
	}
	const int64_t rows = p.image_extent[0];
	const int64_t cols = p.image_extent[1];
	if (cols % 4 != 0 || rows * cols > 65536 || p.image_value_stride[1] != 1 || p.image_result_stride[1] != 1 ||
		p.image_counter_stride[0] != cols || p.image_counter_stride[1] != 1 || p.image_value_stride[0] != cols ||
		p.image_result_stride[0] != cols || reinterpret_cast<uintptr_t>(value->data) % 16 != 0 ||
		reinterpret_cast<uintptr_t>(result->data) % 16 != 0)
	{
		return false;
	}
	for (int k = 0; k < p.batch_rank; ++k)
	{
		if (p.batch_value_stride[k] % 4 != 0 || p.batch_result_stride[k] % 4 != 0)
		{
			return false;
		}
	}
	resident_images g{};
	g.rows = static_cast<int32_t>(rows);
	g.cols = static_cast<int32_t>(cols);
	g.threshold = static_cast<int32_t>(std::min(std::floor(p.squared_radius), static_cast<double>(INT32_MAX)));
	const float *v = static_cast<const float *>(value->data);
	float *r = static_cast<float *>(result->data);
	const int64_t elements = rows * cols;
	if (elements <= 256 * 4 * 1)
	{
		launch_resident<256, 1>(grid, s, v, key, r, p, g);
	}
	else if (elements <= 256 * 4 * 2)
	{
		launch_resident<256, 2>(grid, s, v, key, r, p, g);
	}
	else if (elements <= 256 * 4 * 4)
	{
		launch_resident<256, 4>(grid, s, v, key, r, p, g);
	}
	else if (elements <= 1024 * 4 * 4)
	{
		launch_resident<1024, 4>(grid, s, v, key, r, p, g);
	}
	else if (elements <= 1024 * 4 * 8)
	{
		launch_resident<1024, 8, 0>(grid, s, v, key, r, p, g);
	}
	else
	{
		launch_resident<1024, 16, 9>(grid, s, v, key, r, p, g);
	}
	return true;
}

} // anonymous namespace

} // namespace rexlib_hip

using namespace rexlib_hip;

extern "C" {

int rexlib_hip_standardize_background(const rexlib_hip_array *value, const rexlib_hip_array *key,
	uint32_t image_axes, double radius, double lower_bound, double upper_bound, const rexlib_hip_array *result,
	void *stream)
{
	return guarded([&]
	{
		check_array(value, "standardize_background value", 0);
		check_array(result, "standardize_background result", 0);
		check_array(key, "standardize_background key", 0);
		if (value->dtype != REXLIB_HIP_FLOAT32 && value->dtype != REXLIB_HIP_FLOAT64)
		{
			fail(REXLIB_HIP_UNSUPPORTED, std::string("standardize_background: the gfx950 program takes float32 or "
				"float64 values, not ") + dtype_name(value->dtype));
		}
		if (result->dtype != value->dtype || result->rank != value->rank)
		{
			fail(REXLIB_HIP_INVALID_ARGUMENT, "standardize_background: the result must have the value's type and rank");
		}
		if (key->dtype != REXLIB_HIP_UINT64)
		{
			fail(REXLIB_HIP_INVALID_ARGUMENT, "standardize_background: the key must be uint64");
		}
		if (!(radius >= 0.0) || !(lower_bound <= 0.0 && 0.0 <= upper_bound))
		{
			fail(REXLIB_HIP_INVALID_ARGUMENT, "standardize_background: the radius must be at least 0 and the bounds "
				"must lie on either side of 0");
		}
		const int rank = value->rank;
		if (image_axes == 0 || (rank < 32 && (image_axes >> rank) != 0))
		{
			fail(REXLIB_HIP_INVALID_ARGUMENT, "standardize_background: the image axes must name axes of the value");
		}
		standardize_plan p{};
		p.squared_radius = radius * radius;
		p.lower_bound = lower_bound;
		p.upper_bound = upper_bound;
		p.image_count = 1;
		p.image_elements = 1;
		for (int a = 0; a < rank; ++a)
		{
			if (result->extents[a] != value->extents[a])
			{
				fail(REXLIB_HIP_INVALID_ARGUMENT, "standardize_background: the result must have the value's extents");
			}
			if ((image_axes >> a) & 1u)
			{
				const int k = p.image_rank++;
				p.image_extent[k] = value->extents[a];
				p.image_value_stride[k] = value->strides[a];
				p.image_result_stride[k] = result->strides[a];
				p.centre[k] = value->extents[a] / 2;
				p.image_elements *= value->extents[a];
			}
			else
			{
				const int k = p.batch_rank++;
				p.batch_extent[k] = value->extents[a];
				p.batch_value_stride[k] = value->strides[a];
				p.batch_result_stride[k] = result->strides[a];
				p.image_count *= value->extents[a];
			}
		}
		// The key broadcasts against the batch axes from the right.
		if (key->rank > p.batch_rank)
		{
			fail(REXLIB_HIP_INVALID_ARGUMENT, "standardize_background: the key has more axes than the batch");
		}
		bool shared[max_rank];
		for (int k = 0; k < p.batch_rank; ++k)
		{
			const int ka = k - (p.batch_rank - key->rank);
			const int64_t e = ka >= 0 ? key->extents[ka] : 1;
			if (e != 1 && e != p.batch_extent[k])
			{
				fail(REXLIB_HIP_INVALID_ARGUMENT, "standardize_background: the key does not broadcast against the "
					"batch axes");
			}
			shared[k] = e == 1;
			p.batch_key_stride[k] = e == 1 ? 0 : key->strides[ka];
		}
		// Counter strides: row-major over the image axes and the batch axes
		// the key is shared along, in the value's axis order.
		int64_t step = 1;
		int image_k = p.image_rank;
		int batch_k = p.batch_rank;
		for (int a = rank - 1; a >= 0; --a)
		{
			if ((image_axes >> a) & 1u)
			{
				--image_k;
				p.image_counter_stride[image_k] = step;
				step *= value->extents[a];
			}
			else
			{
				--batch_k;
				p.batch_counter_stride[batch_k] = shared[batch_k] ? step : 0;
				step *= shared[batch_k] ? value->extents[a] : 1;
			}
		}
		if (p.image_count == 0 || p.image_elements == 0)
		{
			return;
		}
		const hipStream_t s = static_cast<hipStream_t>(stream);
		const unsigned grid = static_cast<unsigned>(std::min<int64_t>(p.image_count, 1 << 20));
		const uint64_t *k = static_cast<const uint64_t *>(key->data);
		if (try_resident(s, value, result, k, image_axes, p, grid))
		{
		}
		else if (value->dtype == REXLIB_HIP_FLOAT32)
		{
			standardize_kernel<float><<<grid, threads, 0, s>>>(static_cast<const float *>(value->data), k,
				static_cast<float *>(result->data), p);
		}
		else
		{
			standardize_kernel<double><<<grid, threads, 0, s>>>(static_cast<const double *>(value->data), k,
				static_cast<double *>(result->data), p);
		}
		REXLIB_HIP_CHECK(hipGetLastError());
	});
}

} // extern "C"
