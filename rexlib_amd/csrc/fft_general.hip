// General transforms: complex-to-complex, real-to-complex and
// complex-to-real DFTs over any set of axes of strided operands, of any
// length, with rexlib's three normalisations -- what pocketfft does for the
// reference's CPU backend (src/backends/cpu/kernels/pocketfft_fourier_transform.cpp:74-163,
// src/ops/fourier/fourier_normalization.cpp:56-71).
//
// A complex transform between dense operands over axes of 2-3-5 lengths
// runs its axis passes from the input straight into the output. Otherwise
// the operand is gathered into a contiguous complex buffer; each axis is
// then a batch of lines (outer, line, inner) transformed in place by the
// Stockham kernels (fft.hip) when its length factors into 2, 3 and 5, and by
// Bluestein's chirp-z algorithm otherwise: x_k w_k, zero-padded to a power
// of two m >= 2n - 1, convolved with conj(w) by two length-m transforms, then
// times w_k, with w_k = exp(-+ i pi k^2 / n) computed in double from k^2
// mod 2n. A real-to-complex transform transforms the real input as complex
// and keeps the first n/2 + 1 coefficients of its last axis. A
// complex-to-real transform runs the inverse over every axis but the last,
// extends each last-axis line of n/2 + 1 coefficients to n by Hermitian
// symmetry (the imaginary parts of the self-conjugate coefficients are
// dropped, as pocketfft's c2r does), transforms it, and keeps the real part.
#include "launch.hpp"

#include <algorithm>
#include <cmath>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <type_traits>
#include <vector>

namespace rexlib_hip
{

namespace
{

constexpr int max_rank = REXLIB_HIP_MAX_RANK;

struct strided
{
	int32_t rank;
	int64_t extents[max_rank];
	int64_t strides[max_rank];
};

__device__ inline int64_t strided_offset(const strided &s, int64_t i)
{
	int64_t o = 0;
	for (int a = s.rank - 1; a >= 0; --a)
	{
		const int64_t q = i / s.extents[a];
		o += (i - q * s.extents[a]) * s.strides[a];
		i = q;
	}
	return o;
}

int64_t element_count(const strided &s)
{
	int64_t c = 1;
	for (int a = 0; a < s.rank; ++a)
	{
		c *= s.extents[a];
	}
	return c;
}

// Contiguous complex buffer <- strided real or complex operand.
template <typename T>
__global__ void gather_kernel(const T *in, strided s, int64_t count, float2 *out)
{
	for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < count;
		i += static_cast<int64_t>(gridDim.x) * blockDim.x)
	{
		const T v = in[strided_offset(s, i)];
		if constexpr (std::is_same<T, float>::value)
		{
			out[i] = make_float2(v, 0.0f);
		}
		else
		{
			out[i] = v;
		}
	}
}

// Strided complex operand <- contiguous buffer times `scale` (a sub-box of
// it: `box` extents, read with the buffer's own contiguous strides `from`),
// or its real parts for a real operand.
template <typename T>
__global__ void scatter_kernel_box(const float2 *in, strided box, strided from, T *out, strided to, int64_t count,
	float scale)
{
	for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < count;
		i += static_cast<int64_t>(gridDim.x) * blockDim.x)
	{
		int64_t rest = i;
		int64_t src = 0;
		int64_t dst = 0;
		for (int a = box.rank - 1; a >= 0; --a)
		{
			const int64_t q = rest / box.extents[a];
			const int64_t r = rest - q * box.extents[a];
			rest = q;
			src += r * from.strides[a];
			dst += r * to.strides[a];
		}
		const float2 v = cscale(in[src], scale);
		if constexpr (std::is_same<T, float>::value)
		{
			out[dst] = v.x;
		}
		else
		{
			out[dst] = v;
		}
	}
}

// exp(sign i pi k^2 / n), from k^2 mod 2n in integers and the angle in double.
__device__ inline float2 chirp(int64_t k, int64_t n, float sign)
{
	const int64_t q = (k % (2 * n)) * (k % (2 * n)) % (2 * n);
	double s, c;
	sincospi(static_cast<double>(q) / static_cast<double>(n), &s, &c);
	return make_float2(static_cast<float>(c), static_cast<float>(sign * s));
}

// Bluestein step 1: a[l][k] = x[l][k] w[k] for k < n, 0 up to m, with
// w[k] = exp(sign i pi k^2 / n) from the cached table and m = 2^m_shift.
__global__ void chirp_in_kernel(const float2 *x, line_set lines, int64_t n, int m_shift, const float2 *w, float2 *a)
{
	const int64_t m = int64_t(1) << m_shift;
	const int64_t total = lines.count * m;
	for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < total;
		i += static_cast<int64_t>(gridDim.x) * blockDim.x)
	{
		const int64_t l = i >> m_shift;
		const int64_t k = i & (m - 1);
		float2 v = make_float2(0.0f, 0.0f);
		if (k < n)
		{
			v = cmul(x[line_offset(lines, l) + k * lines.s_elem], w[k]);
		}
		a[i] = v;
	}
}

// Bluestein step 3: a[l][k] *= B[k] (the transform of the conjugate chirp).
__global__ void chirp_multiply_kernel(float2 *a, const float2 *b, int64_t m, int64_t total)
{
	for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < total;
		i += static_cast<int64_t>(gridDim.x) * blockDim.x)
	{
		a[i] = cmul(a[i], b[i % m]);
	}
}

// Bluestein step 5: x[l][k] = scale exp(sign i pi k^2 / n) a[l][k], k < n.
__global__ void chirp_out_kernel(const float2 *a, int64_t n, int64_t m, const float2 *w, float scale, float2 *x,
	line_set lines)
{
	const int64_t total = lines.count * n;
	for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < total;
		i += static_cast<int64_t>(gridDim.x) * blockDim.x)
	{
		const int64_t l = i / n;
		const int64_t k = i - l * n;
		x[line_offset(lines, l) + k * lines.s_elem] = cscale(cmul(a[l * m + k], w[k]), scale);
	}
}

// The conjugate chirp, wrapped to length m: b[k] = b[m - k] = exp(-sign i pi k^2 / n) for k < n.
__global__ void chirp_table_kernel(int64_t n, int64_t m, float sign, float2 *b)
{
	for (int64_t k = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; k < m;
		k += static_cast<int64_t>(gridDim.x) * blockDim.x)
	{
		float2 v = make_float2(0.0f, 0.0f);
		if (k < n)
		{
			v = chirp(k, n, -sign);
		}
		else if (m - k < n)
		{
			v = chirp(m - k, n, -sign);
		}
		b[k] = v;
	}
}

// Hermitian extension along the halved axis: each line of n/2 + 1
// coefficients (lines `half`) becomes n (lines `full`), with the imaginary
// parts of the self-conjugate coefficients (0, and n/2 for even n) dropped.
__global__ void hermitian_kernel(const float2 *in, line_set half, int64_t n, float2 *out, line_set full)
{
	const int64_t h = n / 2 + 1;
	const int64_t total = full.count * n;
	for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < total;
		i += static_cast<int64_t>(gridDim.x) * blockDim.x)
	{
		const int64_t l = i / n;
		const int64_t k = i - l * n;
		const float2 *line = in + line_offset(half, l);
		float2 v;
		if (k < h)
		{
			v = line[k * half.s_elem];
			if (k == 0 || 2 * k == n)
			{
				v.y = 0.0f;
			}
		}
		else
		{
			v = cconj(line[(n - k) * half.s_elem]);
		}
		out[line_offset(full, l) + k * full.s_elem] = v;
	}
}

unsigned blocks_for(int64_t count)
{
	return static_cast<unsigned>(std::max<int64_t>(1, std::min<int64_t>((count + 255) / 256, 16384)));
}

bool smooth(int64_t n)
{
	for (const int r : {2, 3, 5})
	{
		while (n % r == 0)
		{
			n /= r;
		}
	}
	return n == 1;
}

// The Stockham kernels hold two copies of a line and its twiddles in LDS:
// lengths up to 4096 (96 KiB).
constexpr int64_t longest_direct = 4096;

int64_t bluestein_length(int64_t n)
{
	int64_t m = 1;
	while (m < 2 * n - 1)
	{
		m *= 2;
	}
	return m;
}

// The chirp w[k] = exp(sign i pi k^2 / n), k < n.
__global__ void chirp_values_kernel(int64_t n, float sign, float2 *w)
{
	for (int64_t k = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; k < n;
		k += static_cast<int64_t>(gridDim.x) * blockDim.x)
	{
		w[k] = chirp(k, n, sign);
	}
}

// B = FFT_m(conjugate chirp) in the first m elements, then the chirp w
// itself (n elements), per device, length and sign; built once.
const float2 *chirp_spectrum(int64_t n, int64_t m, float sign, hipStream_t stream)
{
	static std::mutex mutex;
	static std::map<std::tuple<int, int64_t, int>, float2 *> cache;
	int device = 0;
	REXLIB_HIP_CHECK(hipGetDevice(&device));
	std::lock_guard<std::mutex> lock(mutex);
	const auto key = std::make_tuple(device, n, sign < 0 ? -1 : 1);
	const auto found = cache.find(key);
	if (found != cache.end())
	{
		return found->second;
	}
	float2 *b = nullptr;
	REXLIB_HIP_CHECK(hipMalloc(&b, sizeof(float2) * static_cast<std::size_t>(m + n)));
	chirp_table_kernel<<<blocks_for(m), 256, 0, stream>>>(n, m, sign, b);
	chirp_values_kernel<<<blocks_for(n), 256, 0, stream>>>(n, sign, b + m);
	const line_set one{1, 1, 1, 0, 0, 0, 1};
	launch_fft_lines(b, one, b, one, m, -1.0f, 1.0f, stream);
	REXLIB_HIP_CHECK(hipStreamSynchronize(stream));
	cache[key] = b;
	return b;
}

// Lines along `axis` of a contiguous buffer of `extents`.
line_set axis_lines(const std::vector<int64_t> &extents, int axis)
{
	int64_t outer = 1;
	int64_t inner = 1;
	for (int a = 0; a < axis; ++a)
	{
		outer *= extents[a];
	}
	for (std::size_t a = axis + 1; a < extents.size(); ++a)
	{
		inner *= extents[a];
	}
	const int64_t n = extents[axis];
	return line_set{outer * inner, 1, inner, n * inner, 0, 1, inner};
}

// In-place transform of the lines along `axis` of the contiguous buffer,
// sign -1 forward, +1 inverse, unscaled.
void transform_axis(float2 *buffer, const std::vector<int64_t> &extents, int axis, float sign, float2 *scratch,
	hipStream_t stream)
{
	const int64_t n = extents[axis];
	const line_set lines = axis_lines(extents, axis);
	if (n <= 1 || lines.count == 0)
	{
		return;
	}
	if (smooth(n) && n <= longest_direct)
	{
		launch_fft_lines(buffer, lines, buffer, lines, n, sign, 1.0f, stream);
		return;
	}
	const int64_t m = bluestein_length(n);
	if (m > longest_direct)
	{
		fail(REXLIB_HIP_UNSUPPORTED, "fft: lengths whose transform needs more than " +
			std::to_string(longest_direct) + " points are not supported on gfx950 (length " + std::to_string(n) + ")");
	}
	const float2 *b = chirp_spectrum(n, m, sign, stream);
	const int64_t total = lines.count * m;
	int m_shift = 0;
	while ((int64_t(1) << m_shift) < m)
	{
		++m_shift;
	}
	chirp_in_kernel<<<blocks_for(total), 256, 0, stream>>>(buffer, lines, n, m_shift, b + m, scratch);
	const line_set padded{lines.count, 1, 1, m, 0, 0, 1};
	launch_fft_lines(scratch, padded, scratch, padded, m, -1.0f, 1.0f, stream);
	chirp_multiply_kernel<<<blocks_for(total), 256, 0, stream>>>(scratch, b, m, total);
	launch_fft_lines(scratch, padded, scratch, padded, m, 1.0f, 1.0f / static_cast<float>(m), stream);
	chirp_out_kernel<<<blocks_for(lines.count * n), 256, 0, stream>>>(scratch, n, m, b + m, 1.0f, buffer, lines);
	REXLIB_HIP_CHECK(hipGetLastError());
}

// Bytes of Bluestein scratch a transform along these axes needs.
int64_t bluestein_elements(const std::vector<int64_t> &extents, const std::vector<int> &axes)
{
	int64_t most = 0;
	int64_t count = 1;
	for (const int64_t e : extents)
	{
		count *= e;
	}
	for (const int a : axes)
	{
		const int64_t n = extents[a];
		if (n > 1 && !(smooth(n) && n <= longest_direct))
		{
			most = std::max(most, count / n * bluestein_length(n));
		}
	}
	return most;
}

// Normalised axes: each in [0, rank), distinct, in the given order.
std::vector<int> normalise_axes(const int32_t *axes, int32_t count, int rank, const char *what)
{
	if (axes == nullptr || count < 1 || count > rank)
	{
		fail(REXLIB_HIP_INVALID_ARGUMENT, std::string(what) + ": between 1 and rank axes are required");
	}
	std::vector<int> out;
	for (int i = 0; i < count; ++i)
	{
		int a = axes[i] < 0 ? axes[i] + rank : axes[i];
		if (a < 0 || a >= rank || std::find(out.begin(), out.end(), a) != out.end())
		{
			fail(REXLIB_HIP_INVALID_ARGUMENT, std::string(what) + ": axes must be distinct and within the rank");
		}
		out.push_back(a);
	}
	return out;
}

// The factor of a transform of `count` points: backward 1 forward and 1/count
// inverse; forward the reverse; ortho 1/sqrt(count) both ways.
double normalisation_factor(int normalization, bool inverse, double count)
{
	switch (normalization)
	{
	case 0:
		return inverse ? 1.0 / count : 1.0;
	case 1:
		return 1.0 / std::sqrt(count);
	case 2:
		return inverse ? 1.0 : 1.0 / count;
	default:
		fail(REXLIB_HIP_INVALID_ARGUMENT, "fft: normalisation is 0 (backward), 1 (ortho) or 2 (forward)");
	}
}

strided strided_of(const rexlib_hip_array &a)
{
	strided s{};
	s.rank = a.rank;
	for (int i = 0; i < a.rank; ++i)
	{
		s.extents[i] = a.extents[i];
		s.strides[i] = a.strides[i];
	}
	return s;
}

strided contiguous_of(const std::vector<int64_t> &extents)
{
	strided s{};
	s.rank = static_cast<int32_t>(extents.size());
	int64_t stride = 1;
	for (int a = s.rank - 1; a >= 0; --a)
	{
		s.extents[a] = extents[a];
		s.strides[a] = stride;
		stride *= extents[a];
	}
	return s;
}

void check_fft_array(const rexlib_hip_array *a, const char *what)
{
	if (a == nullptr || a->rank < 1 || a->rank > max_rank)
	{
		fail(REXLIB_HIP_INVALID_ARGUMENT, std::string(what) + ": an operand of rank 1 to 8 is required");
	}
	for (int i = 0; i < a->rank; ++i)
	{
		if (a->extents[i] < 0)
		{
			fail(REXLIB_HIP_INVALID_ARGUMENT, std::string(what) + ": negative extent");
		}
	}
}

// The whole general transform. kind: 0 c2c, 1 r2c, 2 c2r.
// True when the array's elements are stored row-major with no gaps (axes of
// extent 1 may carry any stride).
bool dense(const rexlib_hip_array &a)
{
	int64_t step = 1;
	for (int d = a.rank - 1; d >= 0; --d)
	{
		if (a.extents[d] != 1 && a.strides[d] != step)
		{
			return false;
		}
		step *= a.extents[d];
	}
	return true;
}

// A complex transform between dense operands (the same buffer, or buffers
// that do not overlap) over axes whose lengths factor into 2, 3 and 5: the
// first axis pass reads the input and writes the output, the later passes
// run in place on the output, the last one applying the normalisation. No
// gather, scatter or workspace. Returns false when the operands do not fit.
bool direct_complex_transform(const rexlib_hip_array *in, const rexlib_hip_array *out, const std::vector<int> &axes,
	const std::vector<int64_t> &full, float sign, float scale, int64_t count, hipStream_t stream)
{
	if (!dense(*in) || !dense(*out))
	{
		return false;
	}
	const char *a = static_cast<const char *>(in->data);
	const char *b = static_cast<const char *>(out->data);
	const int64_t bytes = count * static_cast<int64_t>(sizeof(float2));
	if (a != b && a < b + bytes && b < a + bytes)
	{
		return false;
	}
	std::vector<int> live;
	for (const int axis : axes)
	{
		const int64_t n = full[axis];
		if (n > 1)
		{
			if (!smooth(n) || n > longest_direct)
			{
				return false;
			}
			live.push_back(axis);
		}
	}
	if (live.empty())
	{
		return false;
	}
	const float2 *src = static_cast<const float2 *>(in->data);
	float2 *dst = static_cast<float2 *>(out->data);
	for (std::size_t i = 0; i < live.size(); ++i)
	{
		const line_set lines = axis_lines(full, live[i]);
		launch_fft_lines(i == 0 ? src : dst, lines, dst, lines, full[live[i]], sign,
			i + 1 == live.size() ? scale : 1.0f, stream);
	}
	return true;
}

void general_transform(int kind, const rexlib_hip_array *in, const rexlib_hip_array *out, const int32_t *axes_in,
	int32_t axis_count, bool inverse, int normalization, hipStream_t stream)
{
	const char *what = kind == 0 ? "fft" : (kind == 1 ? "rfft" : "irfft");
	check_fft_array(in, what);
	check_fft_array(out, what);
	if (in->rank != out->rank)
	{
		fail(REXLIB_HIP_INVALID_ARGUMENT, std::string(what) + ": input and output ranks differ");
	}
	const int rank = in->rank;
	const std::vector<int> axes = normalise_axes(axes_in, axis_count, rank, what);
	const int last = axes.back();
	const int expect_in = kind == 1 ? REXLIB_HIP_FLOAT32 : REXLIB_HIP_COMPLEX64;
	const int expect_out = kind == 2 ? REXLIB_HIP_FLOAT32 : REXLIB_HIP_COMPLEX64;
	if (in->dtype != expect_in || out->dtype != expect_out)
	{
		fail(in->dtype == REXLIB_HIP_FLOAT64 || in->dtype == REXLIB_HIP_COMPLEX128 ? REXLIB_HIP_UNSUPPORTED :
			REXLIB_HIP_INVALID_ARGUMENT, std::string(what) + ": float32 / complex64 operands of the transform's kind");
	}
	// The full (complex) extents of the transform.
	std::vector<int64_t> full(in->extents, in->extents + rank);
	if (kind == 2)
	{
		full = std::vector<int64_t>(out->extents, out->extents + rank);
	}
	for (int a = 0; a < rank; ++a)
	{
		const int64_t expect = (kind == 1 && a == last) ? full[a] / 2 + 1 : full[a];
		const int64_t have = kind == 2 ? in->extents[a] : out->extents[a];
		const int64_t want = kind == 2 ? ((a == last) ? full[a] / 2 + 1 : full[a]) : expect;
		if (have != want)
		{
			fail(REXLIB_HIP_INVALID_ARGUMENT, std::string(what) + ": the output extents do not fit the input's");
		}
	}
	int64_t count = 1;
	for (const int64_t e : full)
	{
		count *= e;
	}
	if (count == 0)
	{
		return;
	}
	double points = 1.0;
	for (const int a : axes)
	{
		points *= static_cast<double>(full[a]);
	}
	const float scale = static_cast<float>(normalisation_factor(normalization, inverse, points));
	const float sign = inverse ? 1.0f : -1.0f;
	if (kind == 0 && direct_complex_transform(in, out, axes, full, sign, scale, count, stream))
	{
		REXLIB_HIP_CHECK(hipGetLastError());
		return;
	}

	// Workspace: the contiguous buffer, the c2r half lines, and Bluestein scratch.
	std::vector<int64_t> work_extents = full;
	if (kind == 2)
	{
		work_extents[last] = full[last] / 2 + 1;
	}
	int64_t work = 1;
	for (const int64_t e : work_extents)
	{
		work *= e;
	}
	const int64_t buffer_elements = std::max(count, work);
	const int64_t blue = std::max(bluestein_elements(full, axes), bluestein_elements(work_extents, axes));
	float2 *buffer = static_cast<float2 *>(workspace_area_bytes(area_fft,
		sizeof(float2) * static_cast<std::size_t>(buffer_elements + (kind == 2 ? count : 0) + blue), stream));
	float2 *extended = buffer + buffer_elements;
	float2 *scratch = extended + (kind == 2 ? count : 0);

	// Gather the input.
	const strided in_view = strided_of(*in);
	const int64_t in_count = element_count(in_view);
	if (kind == 1)
	{
		gather_kernel<float><<<blocks_for(in_count), 256, 0, stream>>>(static_cast<const float *>(in->data), in_view,
			in_count, buffer);
	}
	else
	{
		gather_kernel<float2><<<blocks_for(in_count), 256, 0, stream>>>(static_cast<const float2 *>(in->data), in_view,
			in_count, buffer);
	}

	float2 *result = buffer;
	std::vector<int64_t> result_extents = kind == 2 ? work_extents : full;
	if (kind == 2)
	{
		// Inverse over every axis but the halved one, on the half lines; then
		// the Hermitian extension of the halved axis and its transform.
		for (const int a : axes)
		{
			if (a != last)
			{
				transform_axis(buffer, work_extents, a, sign, scratch, stream);
			}
		}
		const int64_t n = full[last];
		const line_set half_lines = axis_lines(work_extents, last);
		const line_set full_lines = axis_lines(full, last);
		hermitian_kernel<<<blocks_for(full_lines.count * n), 256, 0, stream>>>(buffer, half_lines, n, extended,
			full_lines);
		transform_axis(extended, full, last, sign, scratch, stream);
		result = extended;
		result_extents = full;
	}
	else
	{
		for (const int a : axes)
		{
			transform_axis(buffer, full, a, sign, scratch, stream);
		}
	}

	// Scatter the output (r2c: the first n/2 + 1 coefficients of the last axis).
	const strided out_view = strided_of(*out);
	const strided from = contiguous_of(result_extents);
	const int64_t out_count = element_count(out_view);
	if (kind == 2)
	{
		scatter_kernel_box<float><<<blocks_for(out_count), 256, 0, stream>>>(result, out_view, from,
			static_cast<float *>(out->data), out_view, out_count, scale);
	}
	else
	{
		scatter_kernel_box<float2><<<blocks_for(out_count), 256, 0, stream>>>(result, out_view, from,
			static_cast<float2 *>(out->data), out_view, out_count, scale);
	}
	REXLIB_HIP_CHECK(hipGetLastError());
}

} // anonymous namespace

} // namespace rexlib_hip

using namespace rexlib_hip;

extern "C" {

int rexlib_hip_fft(const rexlib_hip_array *in, const rexlib_hip_array *out, const int32_t *axes, int32_t axis_count,
	int32_t inverse, int32_t normalization, void *stream)
{
	return guarded([&]
	{
		general_transform(0, in, out, axes, axis_count, inverse != 0, normalization, static_cast<hipStream_t>(stream));
	});
}

int rexlib_hip_rfft(const rexlib_hip_array *in, const rexlib_hip_array *out, const int32_t *axes, int32_t axis_count,
	int32_t normalization, void *stream)
{
	return guarded([&]
	{
		general_transform(1, in, out, axes, axis_count, false, normalization, static_cast<hipStream_t>(stream));
	});
}

int rexlib_hip_irfft(const rexlib_hip_array *in, const rexlib_hip_array *out, const int32_t *axes,
	int32_t axis_count, int32_t normalization, void *stream)
{
	return guarded([&]
	{
		general_transform(2, in, out, axes, axis_count, true, normalization, static_cast<hipStream_t>(stream));
	});
}

} // extern "C"
