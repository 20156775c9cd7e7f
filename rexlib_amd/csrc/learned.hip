// Density compensation of the accumulators: the iterative estimate and a
// small learned predictor of it.
//
// The reference has no density compensation (its reconstruction divides the
// numerator by the denominator, tests/integration/src/em/functional/test_cpu_reconstruction.cpp:457-464).
// Here the compensation is a weight W per voxel of the half spectrum that
// flattens the gridded sampling density:
//   - iterative (the reference estimate and the fallback): starting from
//     W = 1, each pass convolves W * den with K3 = t x t x t, t = (1/6, 2/3,
//     1/6) -- the autocorrelation of the trilinear window at integer lags --
//     and divides: W <- W / max(K3 * (W den), eps);
//   - learned: a multilayer perceptron 10 -> 64 -> 64 -> 64 -> 1 (ReLU)
//     maps per-voxel features to log W. Its weights are float16 and its
//     products run on v_mfma_f32_16x16x16_f16 with float32 accumulation, in
//     the forward pass and in the backward pass that trains it.
//
// Features of a voxel g: |g| / (N/2); log(1 + den(g)); for each of the 8
// octants (sx, sy, sz) in {-1, 1}^3, log(1 + the mean of den over the 7
// neighbours g + (a sx, b sy, c sz), (a, b, c) in {0, 1}^3 \ 0) -- an 8-bin
// histogram of where the sampling around g lies.
//
// The half spectrum is Hermitian: the full-spectrum element (x, y, z) with x
// beyond the stored half is the element (N - x, -y, -z), and a real field
// such as den takes the same value there.
#include "common.hpp"

#include <hip/hip_fp16.h>

#include <cmath>

namespace rexlib_hip
{

namespace
{

// The stored index of full-spectrum frequency (x, y, z), each taken modulo
// N, for x in [-1, N) and y, z in [-N, 2N) (the neighbours of a stored
// voxel): frequencies with x >= N/2 + 1 are read at their Hermitian mirror
// (-x, -y, -z).
__device__ inline int64_t neighbour_index(int x, int y, int z, int big, int half)
{
	x = x < 0 ? x + big : x;
	y = y < 0 ? y + big : (y >= big ? y - big : y);
	z = z < 0 ? z + big : (z >= big ? z - big : z);
	if (x >= half)
	{
		x = big - x;
		y = y == 0 ? 0 : big - y;
		z = z == 0 ? 0 : big - z;
	}
	return (static_cast<int64_t>(z) * big + y) * half + x;
}

// Threads over a (ceil(N half / blockDim), N) grid: (x, y) of plane z.
__device__ inline bool voxel_of_thread(int big, int half, int &x, int &y, int &z)
{
	const int xy = blockIdx.x * blockDim.x + threadIdx.x;
	if (xy >= big * half)
	{
		return false;
	}
	y = xy / half;
	x = xy - y * half;
	z = blockIdx.y;
	return true;
}

__global__ __launch_bounds__(256) void compensation_pass_kernel(const float *den, const float *w_in, float *w_out,
	int big, float floor_value)
{
	const int half = big / 2 + 1;
	const float t[3] = {1.0f / 6.0f, 2.0f / 3.0f, 1.0f / 6.0f};
	int x, y, z;
	if (!voxel_of_thread(big, half, x, y, z))
	{
		return;
	}
	const int64_t i = (static_cast<int64_t>(z) * big + y) * half + x;
	float c = 0.0f;
	for (int dz = -1; dz <= 1; ++dz)
	{
		for (int dy = -1; dy <= 1; ++dy)
		{
			for (int dx = -1; dx <= 1; ++dx)
			{
				const int64_t j = neighbour_index(x + dx, y + dy, z + dz, big, half);
				c += t[dz + 1] * t[dy + 1] * t[dx + 1] * w_in[j] * den[j];
			}
		}
	}
	const float w = w_in[i];
	w_out[i] = den[i] > 0.0f ? w / fmaxf(c, floor_value) : 0.0f;
}

__global__ void fill_kernel(float *out, int64_t count, float value)
{
	for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < count;
		i += static_cast<int64_t>(gridDim.x) * blockDim.x)
	{
		out[i] = value;
	}
}

// Angular coverage of the slice normals: for radius bin r (of the 8
// spanning |g| in [0, N/2]) and direction u on a 33 x 64 (theta, phi) grid
// (theta = k pi / 32 from pole to pole, phi = k 2 pi / 64: the axes and the
// coordinate planes of the lattice fall on grid directions, so a voxel's
// nearest direction is never a tie between two),
// the fractions of the slices whose plane passes within one voxel of the
// point at that radius and direction, widened by half the grid's spacing
// (|n . u| < 1 / radius + 0.05, so that no plane slips between directions), that fall in
// each of 8 bins of the normal's angle around u, taken modulo pi (n and -n
// are one plane), measured from the local frame e1 = d u / d theta,
// e2 = d u / d phi. All zeros where no plane passes.
constexpr int coverage_radii = 8;
constexpr int coverage_theta = 33;
constexpr int coverage_phi = 64;
constexpr int coverage_bins = 8;
constexpr int coverage_tile = 256;

__device__ inline float3 coverage_direction(int it, int ip, float3 &e1, float3 &e2)
{
	const float theta = static_cast<float>(it) * (3.14159265358979f / (coverage_theta - 1));
	const float phi = static_cast<float>(ip) * (6.28318530717959f / coverage_phi);
	float st, ct, sp, cp;
	sincosf(theta, &st, &ct);
	sincosf(phi, &sp, &cp);
	e1 = make_float3(ct * cp, ct * sp, -st);
	e2 = make_float3(-sp, cp, 0.0f);
	return make_float3(st * cp, st * sp, ct);
}

// One thread per table entry and one slice range per blockIdx.y; the
// workgroup stages the unit normals of 256 slices at a time in LDS. The
// counts of the range are added to `counts` (the table's own memory, as
// uint32, zeroed first) with integer atomics, so the result does not depend
// on the order of the ranges.
__global__ __launch_bounds__(256) void coverage_count_kernel(const float *transforms, int64_t slices, int64_t ts0,
	int64_t ts1, int64_t ts2, int big, int64_t per_block, unsigned *counts)
{
	__shared__ float3 normals[coverage_tile];
	const int entry = blockIdx.x * blockDim.x + threadIdx.x;
	constexpr int entries = coverage_radii * coverage_theta * coverage_phi;
	const bool live = entry < entries;
	const int r = entry / (coverage_theta * coverage_phi);
	const int rest = entry - r * coverage_theta * coverage_phi;
	const int it = rest / coverage_phi;
	const int ip = rest - it * coverage_phi;
	float3 e1, e2;
	const float3 u = coverage_direction(it, ip, e1, e2);
	const float radius = (static_cast<float>(r) + 0.5f) / coverage_radii * (0.5f * static_cast<float>(big));
	const float band = fminf(1.0f, 1.0f / radius + 0.05f);
	unsigned count[coverage_bins];
	for (int b = 0; b < coverage_bins; ++b)
	{
		count[b] = 0;
	}
	const int64_t begin = blockIdx.y * per_block;
	const int64_t end = min<int64_t>(slices, begin + per_block);
	for (int64_t first = begin; first < end; first += coverage_tile)
	{
		const int64_t j = first + threadIdx.x;
		float3 n = make_float3(0.0f, 0.0f, 0.0f);
		if (j < end)
		{
			const float *t = transforms + j * ts0;
			const float x0 = t[0], x1 = t[ts2], x2 = t[2 * ts2];
			const float y0 = t[ts1], y1 = t[ts1 + ts2], y2 = t[ts1 + 2 * ts2];
			n = make_float3(x1 * y2 - x2 * y1, x2 * y0 - x0 * y2, x0 * y1 - x1 * y0);
			const float length = sqrtf(n.x * n.x + n.y * n.y + n.z * n.z);
			const float inv = length > 0.0f ? 1.0f / length : 0.0f;
			n = make_float3(n.x * inv, n.y * inv, n.z * inv);
		}
		__syncthreads();
		normals[threadIdx.x] = n;
		__syncthreads();
		const int tile = static_cast<int>(min<int64_t>(coverage_tile, end - first));
		for (int k = 0; k < tile; ++k)
		{
			const float3 m = normals[k];
			const float along = m.x * u.x + m.y * u.y + m.z * u.z;
			const float a1 = m.x * e1.x + m.y * e1.y + m.z * e1.z;
			const float a2 = m.x * e2.x + m.y * e2.y + m.z * e2.z;
			if (fabsf(along) < band && (a1 != 0.0f || a2 != 0.0f))
			{
				// The angle modulo pi, in [0, pi).
				float angle = atan2f(a2, a1);
				angle = angle < 0.0f ? angle + 3.14159265358979f : angle;
				angle = angle >= 3.14159265358979f ? angle - 3.14159265358979f : angle;
				const int bin = min(coverage_bins - 1, static_cast<int>(angle * (coverage_bins / 3.14159265358979f)));
				for (int b = 0; b < coverage_bins; ++b)
				{
					count[b] += bin == b ? 1u : 0u;
				}
			}
		}
	}
	if (live)
	{
		for (int b = 0; b < coverage_bins; ++b)
		{
			if (count[b] != 0)
			{
				atomicAdd(counts + static_cast<int64_t>(entry) * coverage_bins + b, count[b]);
			}
		}
	}
}

// In place: the uint32 counts of each entry become the fractions of their
// total (zeros when no plane passes).
__global__ __launch_bounds__(256) void coverage_normalize_kernel(float *table)
{
	const int entry = blockIdx.x * blockDim.x + threadIdx.x;
	if (entry >= coverage_radii * coverage_theta * coverage_phi)
	{
		return;
	}
	const unsigned *counts = reinterpret_cast<const unsigned *>(table) + static_cast<int64_t>(entry) * coverage_bins;
	unsigned c[coverage_bins];
	unsigned total = 0;
	for (int b = 0; b < coverage_bins; ++b)
	{
		c[b] = counts[b];
		total += c[b];
	}
	const float scale = total > 0 ? 1.0f / static_cast<float>(total) : 0.0f;
	for (int b = 0; b < coverage_bins; ++b)
	{
		table[static_cast<int64_t>(entry) * coverage_bins + b] = static_cast<float>(c[b]) * scale;
	}
}

// Features of every voxel, as 16 float16 per voxel (the last 6 zero):
// the shell radius |g| / (N/2), log(1 + den * density_scale) (the local
// sample count against the mean of the positive weights), and the 8
// coverage fractions of the table entry nearest to the voxel's radius and
// direction. The 32 bytes are stored as two 16-byte words.
__global__ __launch_bounds__(256) void features_kernel(const float *den, const float *table, float density_scale,
	__half *features, int big)
{
	const int half = big / 2 + 1;
	int x, sy, sz;
	if (!voxel_of_thread(big, half, x, sy, sz))
	{
		return;
	}
	const int64_t i = (static_cast<int64_t>(sz) * big + sy) * half + x;
	// Signed frequencies as numpy's fftfreq: index N/2 is -N/2.
	const int y = sy < big / 2 ? sy : sy - big;
	const int z = sz < big / 2 ? sz : sz - big;
	// log(1 + v) by the hardware logarithm, corrected for the rounding of
	// 1 + v: with u = fl(1 + v), log(u) v / (u - 1) keeps the relative error
	// near one ulp for small v as well (u = 1 gives v itself).
	auto log1 = [](float v)
	{
		const float u = 1.0f + v;
		return u == 1.0f ? v : __logf(u) * (v / (u - 1.0f));
	};
	const float length = sqrtf(static_cast<float>(x * x + y * y + z * z));
	float f[16];
	f[0] = length / (0.5f * static_cast<float>(big));
	f[1] = log1(fmaxf(den[i], 0.0f) * density_scale);
	const int r = min(coverage_radii - 1, static_cast<int>(f[0] * coverage_radii));
	const float theta = length > 0.0f ? acosf(fminf(1.0f, fmaxf(-1.0f, static_cast<float>(z) / length))) : 0.0f;
	float phi = atan2f(static_cast<float>(y), static_cast<float>(x));
	phi = phi < 0.0f ? phi + 6.28318530717959f : phi;
	// The nearest grid direction (phi wraps around).
	const int it = min(coverage_theta - 1, static_cast<int>(theta * ((coverage_theta - 1) / 3.14159265358979f) + 0.5f));
	const int ip = static_cast<int>(phi * (coverage_phi / 6.28318530717959f) + 0.5f) % coverage_phi;
	const float *entry = table + (static_cast<int64_t>(r * coverage_theta + it) * coverage_phi + ip) * coverage_bins;
	for (int b = 0; b < coverage_bins; ++b)
	{
		f[2 + b] = entry[b];
	}
	for (int k = 10; k < 16; ++k)
	{
		f[k] = 0.0f;
	}
	unsigned packed[8];
	for (int k = 0; k < 8; ++k)
	{
		packed[k] = static_cast<unsigned>(__half_as_ushort(__float2half(f[2 * k]))) |
			(static_cast<unsigned>(__half_as_ushort(__float2half(f[2 * k + 1]))) << 16);
	}
	uint4 *out = reinterpret_cast<uint4 *>(features + i * 16);
	out[0] = make_uint4(packed[0], packed[1], packed[2], packed[3]);
	out[1] = make_uint4(packed[4], packed[5], packed[6], packed[7]);
}

// ---------------------------------------------------------------------------
// The perceptron on MFMA
// ---------------------------------------------------------------------------

// Each wavefront works on its own tile; LDS operations of one wavefront
// retire in order, so only the compiler must be kept from moving LDS
// accesses across these points (no workgroup barrier: the wavefronts of a
// workgroup run different numbers of tiles).
__device__ inline void wave_order()
{
	__asm__ __volatile__("" ::: "memory");
}

typedef _Float16 half4_t __attribute__((ext_vector_type(4)));
typedef _Float16 half8_t __attribute__((ext_vector_type(8)));
typedef float float4_t __attribute__((ext_vector_type(4)));

constexpr int hidden = 64;
constexpr int inputs = 16; // 10 features, padded

// Parameter layout (float16 copy and float32 master share it), row-major
// W[k][n]: W1 16x64, W2 64x64, W3 64x64, W4 64x16 (column 0 used), then the
// biases b1, b2, b3 (64 each) and b4 (16, element 0 used).
constexpr int off_w1 = 0;
constexpr int off_w2 = off_w1 + inputs * hidden;
constexpr int off_w3 = off_w2 + hidden * hidden;
constexpr int off_w4 = off_w3 + hidden * hidden;
constexpr int off_b1 = off_w4 + hidden * 16;
constexpr int off_b2 = off_b1 + hidden;
constexpr int off_b3 = off_b2 + hidden;
constexpr int off_b4 = off_b3 + hidden;
constexpr int parameter_count = off_b4 + 16;

// D (16x16, float) += A (16x16, fp16) B (16x16, fp16) on one wavefront. Lane
// l holds A[l % 16][4 (l / 16) + j], B[4 (l / 16) + j][l % 16] and
// D[4 (l / 16) + j][l % 16], j = 0..3.
__device__ inline float4_t mfma16(half4_t a, half4_t b, float4_t c)
{
	return __builtin_amdgcn_mfma_f32_16x16x16f16(a, b, c, 0, 0, 0);
}

// D (16x16, float) += A (16x32, fp16) B (32x16, fp16), gfx950's double-depth
// form: twice the products of mfma16 per instruction. Lane l holds eight
// values of A's row l % 16 and eight of B's column l % 16, the same eight
// depth positions in both; D is laid out as for mfma16.
__device__ inline float4_t mfma32(half8_t a, half8_t b, float4_t c)
{
	return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
}

// Two half4 fragments as one half8.
__device__ inline half8_t join(half4_t lo, half4_t hi)
{
	return half8_t{lo[0], lo[1], lo[2], lo[3], hi[0], hi[1], hi[2], hi[3]};
}

// Load the B operand tile (rows k0.., columns n0..) of a row-major K x N
// float16 matrix in LDS with row length `cols`.
__device__ inline half4_t load_b(const _Float16 *w, int cols, int k0, int n0, int lane)
{
	half4_t b;
	const int k = k0 + 4 * (lane >> 4);
	const int n = n0 + (lane & 15);
	for (int j = 0; j < 4; ++j)
	{
		b[j] = w[(k + j) * cols + n];
	}
	return b;
}

// Load the A operand tile (rows m0.., columns k0..) of a row-major M x K
// float16 matrix in LDS with row length `cols`.
__device__ inline half4_t load_a(const _Float16 *m, int cols, int m0, int k0, int lane)
{
	half4_t a;
	const int row = m0 + (lane & 15);
	const int k = k0 + 4 * (lane >> 4);
	for (int j = 0; j < 4; ++j)
	{
		a[j] = m[row * cols + k + j];
	}
	return a;
}

// A operand of the transpose of a row-major K x M matrix: A[m][k] = T[k][m].
__device__ inline half4_t load_a_transposed(const _Float16 *t, int cols, int m0, int k0, int lane)
{
	half4_t a;
	const int m = m0 + (lane & 15);
	const int k = k0 + 4 * (lane >> 4);
	for (int j = 0; j < 4; ++j)
	{
		a[j] = t[(k + j) * cols + m];
	}
	return a;
}

// B operand of the transpose of a row-major N x K matrix: B[k][n] = T[n][k].
__device__ inline half4_t load_b_transposed(const _Float16 *t, int cols, int k0, int n0, int lane)
{
	half4_t b;
	const int k = k0 + 4 * (lane >> 4);
	const int n = n0 + (lane & 15);
	for (int j = 0; j < 4; ++j)
	{
		b[j] = t[n * cols + k + j];
	}
	return b;
}

// One dense layer on a 16-row tile: out (16 x 64) = act(in (16 x K) W (K x 64) + b).
template <int K, bool RELU>
__device__ inline void dense_tile(const _Float16 *in, const _Float16 *w, const float *b, _Float16 *out, int lane)
{
	for (int nt = 0; nt < hidden / 16; ++nt)
	{
		float4_t acc = {0.0f, 0.0f, 0.0f, 0.0f};
		for (int kt = 0; kt < K / 16; ++kt)
		{
			acc = mfma16(load_a(in, K, 0, kt * 16, lane), load_b(w, hidden, kt * 16, nt * 16, lane), acc);
		}
		const int col = nt * 16 + (lane & 15);
		for (int j = 0; j < 4; ++j)
		{
			const int row = 4 * (lane >> 4) + j;
			float v = acc[j] + b[col];
			if (RELU)
			{
				v = fmaxf(v, 0.0f);
			}
			out[row * hidden + col] = static_cast<_Float16>(v);
		}
	}
}

// Stage the float16 parameters and float32 biases in LDS.
struct mlp_lds
{
	_Float16 w[off_b1];
	float b[parameter_count - off_b1];
};

__device__ inline void stage_parameters(mlp_lds &s, const _Float16 *params_h, const float *params_f)
{
	for (int i = threadIdx.x; i < off_b1; i += blockDim.x)
	{
		s.w[i] = params_h[i];
	}
	for (int i = threadIdx.x; i < parameter_count - off_b1; i += blockDim.x)
	{
		s.b[i] = params_f[off_b1 + i];
	}
	__syncthreads();
}

// Per-wave activations: the input tile and the three hidden layers.
struct tile_lds
{
	_Float16 x[16 * inputs];
	_Float16 h1[16 * hidden];
	_Float16 h2[16 * hidden];
	_Float16 h3[16 * hidden];
	float g[16 * hidden]; // gradient of one layer's output
	_Float16 gh[16 * hidden]; // gradients as float16 MFMA operands, two in turn
	_Float16 gh2[16 * hidden];
};

constexpr int mlp_waves = 4;

// out[i] = MLP(features[i]) for i < count, in transposed form: each layer
// computes Y^T = W^T X^T on a tile of 16 samples, the 16 columns of the B
// operand. The weights are the A operands and stay in registers for the
// whole kernel. An MFMA's output fragment (lane l: rows 4 (l / 16) + j,
// column l % 16) is laid out as a B operand, so after bias, ReLU and the
// float16 rounding the activations never leave registers.
// - The first layer (depth 16) takes mfma16: lane l of A holds
//   W^T[m0 + l % 16][4 (l / 16) + j] = W[4 (l / 16) + j][m0 + l % 16].
// - The depth-64 layers take mfma32 over two chunks of 32 hidden units. A
//   sum is the same in any order of its terms, so chunk t gives lane group
//   g = l / 16 the units 32 t + 4 g + j (j < 4) and 32 t + 16 + 4 g + j - 4
//   (j >= 4): its B operand is the two output fragments of m-tiles 2t and
//   2t + 1 side by side, and its A operand holds the weights of the same
//   units.
__global__ __launch_bounds__(256) void mlp_forward_kernel(const __half *features, const _Float16 *params_h,
	const float *params_f, float *out, int64_t count)
{
	// Biases per (layer, m-tile, lane group): the 4 rows a lane's output holds.
	__shared__ float4_t bias[3][4][4];
	if (threadIdx.x < 48)
	{
		const int layer = threadIdx.x / 16;
		const int mt = (threadIdx.x / 4) % 4;
		const int g = threadIdx.x % 4;
		const float *b = params_f + off_b1 + layer * hidden + mt * 16 + 4 * g;
		bias[layer][mt][g] = float4_t{b[0], b[1], b[2], b[3]};
	}
	__syncthreads();
	const int lane = threadIdx.x & 63;
	const int g = lane >> 4;
	const int c = lane & 15;
	half4_t w1[4];
	half8_t w2[4][2];
	half8_t w3[4][2];
	half8_t w4[2];
#pragma unroll
	for (int j = 0; j < 4; ++j)
	{
#pragma unroll
		for (int mt = 0; mt < 4; ++mt)
		{
			w1[mt][j] = params_h[off_w1 + (4 * g + j) * hidden + mt * 16 + c];
		}
	}
#pragma unroll
	for (int j = 0; j < 8; ++j)
	{
#pragma unroll
		for (int t = 0; t < 2; ++t)
		{
			// Hidden unit of depth position (g, j) in chunk t.
			const int unit = 32 * t + (j < 4 ? 4 * g + j : 16 + 4 * g + j - 4);
#pragma unroll
			for (int mt = 0; mt < 4; ++mt)
			{
				w2[mt][t][j] = params_h[off_w2 + unit * hidden + mt * 16 + c];
				w3[mt][t][j] = params_h[off_w3 + unit * hidden + mt * 16 + c];
			}
			w4[t][j] = params_h[off_w4 + unit * 16 + c];
		}
	}
	const float b4 = params_f[off_b4];
	const int64_t tiles = (count + 15) / 16;
	const int64_t waves = static_cast<int64_t>(gridDim.x) * (blockDim.x / 64);
	for (int64_t tile = static_cast<int64_t>(blockIdx.x) * (blockDim.x / 64) + (threadIdx.x >> 6); tile < tiles;
		tile += waves)
	{
		const int64_t row = tile * 16 + c;
		// B fragment of X^T: B[4 g + j][c] = X[row][4 g + j], 8 bytes per lane.
		half4_t x = {0, 0, 0, 0};
		if (row < count)
		{
			x = *reinterpret_cast<const half4_t *>(features + row * inputs + 4 * g);
		}
		half4_t h1[4];
		half4_t h2[4];
#pragma unroll
		for (int mt = 0; mt < 4; ++mt)
		{
			const float4_t acc = mfma16(w1[mt], x, float4_t{0.0f, 0.0f, 0.0f, 0.0f});
			const float4_t b = bias[0][mt][g];
#pragma unroll
			for (int j = 0; j < 4; ++j)
			{
				h1[mt][j] = static_cast<_Float16>(fmaxf(acc[j] + b[j], 0.0f));
			}
		}
#pragma unroll
		for (int mt = 0; mt < 4; ++mt)
		{
			float4_t acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
			for (int t = 0; t < 2; ++t)
			{
				acc = mfma32(w2[mt][t], join(h1[2 * t], h1[2 * t + 1]), acc);
			}
			const float4_t b = bias[1][mt][g];
#pragma unroll
			for (int j = 0; j < 4; ++j)
			{
				h2[mt][j] = static_cast<_Float16>(fmaxf(acc[j] + b[j], 0.0f));
			}
		}
#pragma unroll
		for (int mt = 0; mt < 4; ++mt)
		{
			float4_t acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
			for (int t = 0; t < 2; ++t)
			{
				acc = mfma32(w3[mt][t], join(h2[2 * t], h2[2 * t + 1]), acc);
			}
			const float4_t b = bias[2][mt][g];
#pragma unroll
			for (int j = 0; j < 4; ++j)
			{
				h1[mt][j] = static_cast<_Float16>(fmaxf(acc[j] + b[j], 0.0f)); // layer 3, reusing h1
			}
		}
		float4_t acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
		for (int t = 0; t < 2; ++t)
		{
			acc = mfma32(w4[t], join(h1[2 * t], h1[2 * t + 1]), acc);
		}
		// Output row 0 (the one used column of W4) sits in lanes 0..15.
		if (g == 0 && row < count)
		{
			out[row] = acc[0] + b4;
		}
	}
}

// Add the gradient of a weight matrix: dW (K x 64) += act_in^T (K x 16) dY (16 x 64), for one tile.
template <int K>
__device__ inline void accumulate_weight_gradient(const _Float16 *act_in, int in_cols, const _Float16 *dy,
	float *grad, float scale, int lane)
{
	for (int mt = 0; mt < K / 16; ++mt)
	{
		for (int nt = 0; nt < hidden / 16; ++nt)
		{
			float4_t acc = {0.0f, 0.0f, 0.0f, 0.0f};
			acc = mfma16(load_a_transposed(act_in, in_cols, mt * 16, 0, lane), load_b(dy, hidden, 0, nt * 16, lane),
				acc);
			const int col = nt * 16 + (lane & 15);
			for (int j = 0; j < 4; ++j)
			{
				const int row = mt * 16 + 4 * (lane >> 4) + j;
				atomicAdd(&grad[row * hidden + col], acc[j] * scale);
			}
		}
	}
}

// dX (16 x K) = dY (16 x 64) W^T (64 x K) masked by act_x > 0; written to t.g and t.gh.
template <int K>
__device__ inline void backward_input(const _Float16 *dy, const _Float16 *w, const _Float16 *act_x, float *g,
	_Float16 *gh, int lane)
{
	for (int nt = 0; nt < K / 16; ++nt)
	{
		float4_t acc = {0.0f, 0.0f, 0.0f, 0.0f};
		for (int kt = 0; kt < hidden / 16; ++kt)
		{
			acc = mfma16(load_a(dy, hidden, 0, kt * 16, lane), load_b_transposed(w, hidden, kt * 16, nt * 16, lane),
				acc);
		}
		const int col = nt * 16 + (lane & 15);
		for (int j = 0; j < 4; ++j)
		{
			const int row = 4 * (lane >> 4) + j;
			const float v = static_cast<float>(act_x[row * K + col]) > 0.0f ? acc[j] : 0.0f;
			g[row * hidden + col] = v;
			gh[row * hidden + col] = static_cast<_Float16>(v);
		}
	}
}

// One pass of forward and backward over `count` samples: the mean squared
// error of MLP(features) against `target`, scaled by `loss_scale` for the
// float16 gradients; grads (float32, parameter layout) and loss[0] accumulate.
__global__ __launch_bounds__(64 * mlp_waves) void mlp_backward_kernel(const __half *features, const float *target,
	const _Float16 *params_h, const float *params_f, float *grads, float *loss, int64_t count, float loss_scale)
{
	__shared__ mlp_lds s;
	__shared__ tile_lds tiles[mlp_waves];
	stage_parameters(s, params_h, params_f);
	const int lane = threadIdx.x & 63;
	const int wave = threadIdx.x >> 6;
	tile_lds &t = tiles[wave];
	const int64_t tiles_total = (count + 15) / 16;
	const float inverse = 1.0f / static_cast<float>(count);
	for (int64_t tile = blockIdx.x * static_cast<int64_t>(mlp_waves) + wave; tile < tiles_total;
		tile += static_cast<int64_t>(gridDim.x) * mlp_waves)
	{
		const int64_t row0 = tile * 16;
		for (int e = lane; e < 16 * inputs; e += 64)
		{
			const int64_t r = row0 + e / inputs;
			t.x[e] = r < count ? static_cast<_Float16>(__half2float(features[r * inputs + e % inputs])) : _Float16(0);
		}
		wave_order();
		dense_tile<inputs, true>(t.x, s.w + off_w1, s.b, t.h1, lane);
		wave_order();
		dense_tile<hidden, true>(t.h1, s.w + off_w2, s.b + (off_b2 - off_b1), t.h2, lane);
		wave_order();
		dense_tile<hidden, true>(t.h2, s.w + off_w3, s.b + (off_b3 - off_b1), t.h3, lane);
		wave_order();
		float4_t acc = {0.0f, 0.0f, 0.0f, 0.0f};
		for (int kt = 0; kt < hidden / 16; ++kt)
		{
			acc = mfma16(load_a(t.h3, hidden, 0, kt * 16, lane), load_b(s.w + off_w4, 16, kt * 16, 0, lane), acc);
		}
		// dOut = 2 (out - target) / count, scaled; rows beyond count carry 0.
		// Lanes with column 0 hold the outputs; share them through t.g.
		if ((lane & 15) == 0)
		{
			for (int j = 0; j < 4; ++j)
			{
				const int row = 4 * (lane >> 4) + j;
				const int64_t r = row0 + row;
				float d = 0.0f;
				if (r < count)
				{
					const float out = acc[j] + s.b[off_b4 - off_b1];
					const float diff = out - target[r];
					atomicAdd(loss, diff * diff * inverse);
					d = 2.0f * diff * inverse * loss_scale;
				}
				t.g[row] = d;
			}
		}
		wave_order();
		// Layer 4: dW4[c][0] += sum_r h3[r][c] dOut[r]; db4 += sum_r dOut[r];
		// dH3[r][c] = dOut[r] W4[c][0] (h3 > 0).
		{
			const int c = lane;
			float sum = 0.0f;
			for (int r = 0; r < 16; ++r)
			{
				sum += static_cast<float>(t.h3[r * hidden + c]) * t.g[r];
			}
			atomicAdd(&grads[off_w4 + c * 16], sum);
			if (lane == 0)
			{
				float db = 0.0f;
				for (int r = 0; r < 16; ++r)
				{
					db += t.g[r];
				}
				atomicAdd(&grads[off_b4], db);
			}
			const float w4 = static_cast<float>(s.w[off_w4 + c * 16]);
			float dout[16];
			for (int r = 0; r < 16; ++r)
			{
				dout[r] = t.g[r];
			}
			wave_order();
			for (int r = 0; r < 16; ++r)
			{
				const float v = static_cast<float>(t.h3[r * hidden + c]) > 0.0f ? dout[r] * w4 : 0.0f;
				t.g[r * hidden + c] = v;
				t.gh[r * hidden + c] = static_cast<_Float16>(v);
			}
		}
		wave_order();
		// Layer 3: dW3 += h2^T dH3; db3 += column sums; dH2 = dH3 W3^T (h2 > 0).
		accumulate_weight_gradient<hidden>(t.h2, hidden, t.gh, grads + off_w3, 1.0f, lane);
		{
			float db = 0.0f;
			for (int r = 0; r < 16; ++r)
			{
				db += t.g[r * hidden + lane];
			}
			atomicAdd(&grads[off_b3 + lane], db);
		}
		wave_order();
		backward_input<hidden>(t.gh, s.w + off_w3, t.h2, t.g, t.gh2, lane);
		wave_order();
		// Layer 2.
		accumulate_weight_gradient<hidden>(t.h1, hidden, t.gh2, grads + off_w2, 1.0f, lane);
		{
			float db = 0.0f;
			for (int r = 0; r < 16; ++r)
			{
				db += t.g[r * hidden + lane];
			}
			atomicAdd(&grads[off_b2 + lane], db);
		}
		wave_order();
		backward_input<hidden>(t.gh2, s.w + off_w2, t.h1, t.g, t.gh, lane);
		wave_order();
		// Layer 1: dW1 += x^T dH1; db1 += column sums.
		accumulate_weight_gradient<inputs>(t.x, inputs, t.gh, grads + off_w1, 1.0f, lane);
		{
			float db = 0.0f;
			for (int r = 0; r < 16; ++r)
			{
				db += t.g[r * hidden + lane];
			}
			atomicAdd(&grads[off_b1 + lane], db);
		}
		wave_order();
	}
}

// Adam on the float32 master parameters; refreshes the float16 copy. The
// gradients arrive multiplied by `loss_scale` (except the output layer's,
// which arrive multiplied too) and are cleared for the next step.
__global__ void adam_kernel(float *params, _Float16 *params_h, float *grads, float *m, float *v, int count, float lr,
	float beta1, float beta2, float eps, float correction1, float correction2, float loss_scale)
{
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= count)
	{
		return;
	}
	const float g = grads[i] / loss_scale;
	grads[i] = 0.0f;
	const float mi = beta1 * m[i] + (1.0f - beta1) * g;
	const float vi = beta2 * v[i] + (1.0f - beta2) * g * g;
	m[i] = mi;
	v[i] = vi;
	const float update = lr * (mi / correction1) / (sqrtf(vi / correction2) + eps);
	const float p = params[i] - update;
	params[i] = p;
	params_h[i] = static_cast<_Float16>(p);
}

// The (ceil(N (N/2 + 1) / 256), N) grid of voxel_of_thread.
dim3 volume_grid(int big)
{
	const int64_t plane = static_cast<int64_t>(big) * (big / 2 + 1);
	return dim3(static_cast<unsigned>((plane + 255) / 256), static_cast<unsigned>(big));
}

unsigned blocks_of(int64_t count, int per)
{
	return static_cast<unsigned>(std::max<int64_t>(1, std::min<int64_t>((count + per - 1) / per, 16384)));
}

} // anonymous namespace

int mlp_parameter_count()
{
	return parameter_count;
}

void launch_compensation(const float *den, float *w, float *scratch, int big, int passes, float floor_value,
	hipStream_t stream)
{
	const int64_t count = static_cast<int64_t>(big) * big * (big / 2 + 1);
	// Ping-pong between w and scratch, starting so that the last pass
	// writes w.
	float *a = passes % 2 == 0 ? w : scratch;
	float *b = passes % 2 == 0 ? scratch : w;
	fill_kernel<<<blocks_of(count, 256), 256, 0, stream>>>(a, count, 1.0f);
	for (int p = 0; p < passes; ++p)
	{
		compensation_pass_kernel<<<volume_grid(big), 256, 0, stream>>>(den, a, b, big, floor_value);
		float *t = a;
		a = b;
		b = t;
	}
	if (passes == 0)
	{
		fill_kernel<<<blocks_of(count, 256), 256, 0, stream>>>(w, count, 1.0f);
	}
	REXLIB_HIP_CHECK(hipGetLastError());
}

void launch_coverage(const float *transforms, int64_t slices, int64_t ts0, int64_t ts1, int64_t ts2, int big,
	float *table, hipStream_t stream)
{
	constexpr int entries = coverage_radii * coverage_theta * coverage_phi;
	constexpr int64_t table_size = static_cast<int64_t>(entries) * coverage_bins;
	// Up to 64 slice ranges of whole tiles: 66 x 64 workgroups for large stacks.
	const int64_t tiles = std::max<int64_t>(1, (slices + coverage_tile - 1) / coverage_tile);
	const int64_t per_block = ((tiles + 63) / 64) * coverage_tile;
	const unsigned ranges = static_cast<unsigned>(std::max<int64_t>(1, (slices + per_block - 1) / per_block));
	fill_kernel<<<(table_size + 255) / 256, 256, 0, stream>>>(table, table_size, 0.0f);
	REXLIB_HIP_CHECK(hipGetLastError());
	coverage_count_kernel<<<dim3((entries + 255) / 256, ranges), 256, 0, stream>>>(transforms, slices, ts0, ts1, ts2,
		big, per_block, reinterpret_cast<unsigned *>(table));
	REXLIB_HIP_CHECK(hipGetLastError());
	coverage_normalize_kernel<<<(entries + 255) / 256, 256, 0, stream>>>(table);
	REXLIB_HIP_CHECK(hipGetLastError());
}

void launch_features(const float *den, const float *table, float density_scale, void *features, int big,
	hipStream_t stream)
{
	features_kernel<<<volume_grid(big), 256, 0, stream>>>(den, table, density_scale, static_cast<__half *>(features),
		big);
	REXLIB_HIP_CHECK(hipGetLastError());
}

void launch_mlp_forward(const void *features, const void *params_h, const float *params_f, float *out, int64_t count,
	hipStream_t stream)
{
	mlp_forward_kernel<<<blocks_of((count + 15) / 16, 4 * 32), 256, 0, stream>>>(
		static_cast<const __half *>(features), static_cast<const _Float16 *>(params_h), params_f, out, count);
	REXLIB_HIP_CHECK(hipGetLastError());
}

void launch_mlp_backward(const void *features, const float *target, const void *params_h, const float *params_f,
	float *grads, float *loss, int64_t count, float loss_scale, hipStream_t stream)
{
	mlp_backward_kernel<<<blocks_of((count + 15) / 16, mlp_waves), 64 * mlp_waves, 0, stream>>>(
		static_cast<const __half *>(features), target, static_cast<const _Float16 *>(params_h), params_f, grads, loss,
		count, loss_scale);
	REXLIB_HIP_CHECK(hipGetLastError());
}

void launch_adam(float *params, void *params_h, float *grads, float *m, float *v, float lr, float beta1, float beta2,
	float eps, int step, float loss_scale, hipStream_t stream)
{
	const float c1 = 1.0f - std::pow(beta1, static_cast<float>(step));
	const float c2 = 1.0f - std::pow(beta2, static_cast<float>(step));
	adam_kernel<<<(parameter_count + 255) / 256, 256, 0, stream>>>(params, static_cast<_Float16 *>(params_h), grads, m,
		v, parameter_count, lr, beta1, beta2, eps, c1, c2, loss_scale);
	REXLIB_HIP_CHECK(hipGetLastError());
}

} // namespace rexlib_hip
