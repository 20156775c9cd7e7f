// Binned insertion on gfx950: the counting pass, the bin plan, and the brick
// kernel that adds each bin's samples to its brick in LDS.
//
// Per chunk of slices, on one stream:
//   1. bin_count_kernel: for every inserted sample copy, its brick, from the
//      geometry alone (frame, shell, radial weight); runs of equal bricks
//      along a wavefront add their length to the bin's count with one atomic;
//   2. bin_plan_kernel: one block turns the counts into bin ends, cursors
//      and tasks (a bin's records cut into shares of at most
//      records_per_task), and resets the task counter;
//   3. a producer (prepare_kernel or a fused column pass, insertion.hip)
//      writes every sample copy into its bin (bins.hpp, bin_put);
//   4. bin_brick_kernel: persistent wavefronts take tasks from the counter;
//      each sums its share of a bin into the brick and its halo in LDS and
//      adds the brick to the accumulators.
// The tile-owned gather of the reference
// (src/backends/cpu/builders/em/reconstruction/slice_insertion_functor.inl:105-465)
// visits every slice per tile and bounds its rows; here the samples are
// sorted to their bricks once, so a brick reads only its own samples.
#include "bins.hpp"
#include "launch.hpp"

#include <algorithm>
#include <map>
#include <mutex>
#include <tuple>

namespace rexlib_hip
{

namespace
{

// One thread per stored element (row-major, so a wavefront's lanes run
// along k_x and mostly share bricks), blocks along y over the chunk's slices.
__global__ __launch_bounds__(256) void bin_count_kernel(insertion_params p, slice_chunk chunk, bin_grid g,
	uint32_t *counts)
{
	const int elements = p.n * p.h;
	const int idx = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
	const bool live = idx < elements;
	const int clamped = live ? idx : 0;
	const int row = static_cast<int>(static_cast<unsigned>(clamped) / static_cast<unsigned>(p.h));
	const int column = clamped - row * p.h;
	bin_sink sink{};
	sink.g = g;
	const int64_t slices = p.map.kept_count * chunk.count;
	for (int64_t j = blockIdx.y; j < slices; j += gridDim.y)
	{
		const int64_t kept = j / chunk.count;
		const int64_t local = j - kept * chunk.count;
		const int64_t s = kept * p.map.reduced_count + chunk.begin + local;
		const slice_entry &e = p.entries[s];
		const bool inserted = live && e.last_shell >= 0 && sample_inserted(p, e, p.radial + e.radial_offset, row, column);
		bin_put<true>(p, sink, counts, e, kept, static_cast<uint32_t>(local), row, column, inserted,
			prepared_sample{0.0f, 0.0f, 0.0f});
	}
}

// Exclusive sums over the 1024 threads of a block of two per-thread values.
__device__ inline void block_exclusive_scan2(uint32_t &a, uint32_t &b, uint32_t &total_a, uint32_t &total_b)
{
	__shared__ uint32_t wave_a[16];
	__shared__ uint32_t wave_b[16];
	const int lane = threadIdx.x & 63;
	const int wave = threadIdx.x >> 6;
	const uint32_t ia = static_cast<uint32_t>(wave_inclusive_scan(static_cast<int>(a)));
	const uint32_t ib = static_cast<uint32_t>(wave_inclusive_scan(static_cast<int>(b)));
	if (lane == 63)
	{
		wave_a[wave] = ia;
		wave_b[wave] = ib;
	}
	__syncthreads();
	uint32_t before_a = 0, before_b = 0;
	total_a = 0;
	total_b = 0;
	for (int w = 0; w < 16; ++w)
	{
		if (w < wave)
		{
			before_a += wave_a[w];
			before_b += wave_b[w];
		}
		total_a += wave_a[w];
		total_b += wave_b[w];
	}
	a = before_a + ia - a;
	b = before_b + ib - b;
}

// One block of 1024 threads. Each thread takes a contiguous range of bins;
// bins with no record get no task.
__global__ __launch_bounds__(1024) void bin_plan_kernel(const uint32_t *counts, int64_t bins, uint32_t per_task,
	uint32_t *end, uint32_t *cursor, int4 *tasks, uint32_t task_capacity, uint32_t *task_count, uint32_t *task_next)
{
	const int64_t per_thread = (bins + 1023) / 1024;
	const int64_t b0 = min(bins, static_cast<int64_t>(threadIdx.x) * per_thread);
	const int64_t b1 = min(bins, b0 + per_thread);
	uint32_t records = 0;
	uint32_t shares = 0;
	for (int64_t b = b0; b < b1; ++b)
	{
		const uint32_t c = counts[b];
		records += c;
		shares += (c + per_task - 1) / per_task;
	}
	uint32_t record_base = records;
	uint32_t task_base = shares;
	uint32_t total_records = 0, total_tasks = 0;
	block_exclusive_scan2(record_base, task_base, total_records, total_tasks);
	for (int64_t b = b0; b < b1; ++b)
	{
		const uint32_t c = counts[b];
		end[b] = record_base + c;
		cursor[b] = record_base;
		const uint32_t k = (c + per_task - 1) / per_task;
		for (uint32_t i = 0; i < k; ++i)
		{
			if (task_base + i < task_capacity)
			{
				const uint32_t begin = record_base + static_cast<uint32_t>(static_cast<uint64_t>(c) * i / k);
				const uint32_t stop = record_base + static_cast<uint32_t>(static_cast<uint64_t>(c) * (i + 1) / k);
				tasks[task_base + i] = make_int4(static_cast<int>(b), static_cast<int>(begin), static_cast<int>(stop),
					static_cast<int>(k));
			}
		}
		record_base += c;
		task_base += k;
	}
	if (threadIdx.x == 0)
	{
		*task_count = min(total_tasks, task_capacity);
		*task_next = 0;
	}
}

__global__ void fill_u32_kernel(uint32_t *a, int64_t count, uint32_t value)
{
	for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < count;
		i += static_cast<int64_t>(gridDim.x) * blockDim.x)
	{
		a[i] = value;
	}
}

// A voxel of a brick in LDS: 12 bytes.
struct brick_voxel
{
	float re;
	float im;
	float den;
};

template <int PROFILE>
struct profile_reach
{
	// Taps per axis the LDS brick is sized for: the window's own, at most 4
	// for Kaiser-Bessel windows on this schedule.
	static constexpr int taps = PROFILE == REXLIB_HIP_BOX ? 1 : (PROFILE == REXLIB_HIP_TRIANGLE ? 2 : 4);
};

// A wavefront owns a brick of 2^SX x 2^SY x 2^SZ first-tap cells and the
// taps - 1 voxels beyond it along each axis, all in LDS, and takes one task
// (a share of one bin) at a time.
//
// Lane l streams records begin + l, begin + l + 64, ... of the share. Per
// round every lane holding a record claims its cell (writes its lane into
// the cell's claim word, then reads it back); the lanes that read their own
// lane add their sample's taps, the others keep their record for the next
// round. So no two lanes of one round share a cell. A tap is a constant
// offset from the cell, so no two lanes of one tap instruction touch one
// voxel either, and the taps of a round are added one after another: LDS
// runs a wavefront's reads and writes in program order, and lds_order keeps
// the compiler from moving a tap's reads above the previous tap's writes.
// The wavefront is the brick's only writer, so no barrier is needed.
//
// The volume phase exp(-2 pi i g C / N) is a function of the voxel alone, so
// the brick sums v * w and applies the phase once per voxel when it adds the
// brick to the accumulators (alternating_sign.inl, tabulated_phase.inl).
// Voxels no other task can touch (inside the brick, at least taps - 1 past
// its lower faces, not a y or z frequency of +-N/2 on an even N, in a bin of
// one share) are added by plain read-add-write, the others with atomics.
template <int PROFILE, int SX, int SY, int SZ>
__global__ __launch_bounds__(64) void bin_brick_kernel(insertion_params p, slice_chunk chunk, bin_grid g,
	const bin_record *records, uint32_t capacity, const int4 *tasks, const uint32_t *task_count, uint32_t *task_next)
{
	constexpr int T = profile_reach<PROFILE>::taps;
	constexpr int EX = 1 << SX;
	constexpr int EY = 1 << SY;
	constexpr int EZ = 1 << SZ;
	constexpr int LX = EX + T - 1;
	constexpr int LY = EY + T - 1;
	constexpr int LZ = EZ + T - 1;
	constexpr int voxels = LX * LY * LZ;
	__shared__ brick_voxel acc[voxels];
	__shared__ int claim[EX * EY * EZ];
	volatile int *claim_read = claim;

	const int lane = threadIdx.x;
	const int taps = PROFILE == REXLIB_HIP_KAISER_BESSEL ? p.taps : T;
	const int reach = taps - 1;
	const int he = g.half_extent;
	const int big = p.big;
	const bool odd = (big & 1) != 0;
	const uint32_t column_mask = (1u << g.column_bits) - 1u;
	const uint32_t row_mask = (1u << g.row_bits) - 1u;
	const int sign_shift = g.column_bits + g.row_bits;
	const uint32_t total = *task_count;

	for (;;)
	{
		uint32_t t = 0;
		if (lane == 0)
		{
			t = atomicAdd(task_next, 1u);
		}
		t = static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(static_cast<int>(t)));
		if (t >= total)
		{
			break;
		}
		const int4 task = tasks[t];
		const int64_t bin = static_cast<uint32_t>(task.x);
		// Shares never pass the records the buffer holds.
		const uint32_t begin = min(static_cast<uint32_t>(task.y), capacity);
		const uint32_t end = min(static_cast<uint32_t>(task.z), capacity);
		const bool single = task.w == 1;
		const int64_t kept = bin / g.bricks;
		const int64_t brick = bin - kept * g.bricks;
		const int bx = static_cast<int>(brick % g.count[0]);
		const int64_t rest = brick / g.count[0];
		const int by = static_cast<int>(rest % g.count[1]);
		const int bz = static_cast<int>(rest / g.count[1]);
		// Signed coordinates of LDS voxel (0, 0, 0).
		const int ox = (bx << SX) - reach;
		const int oy = (by << SY) - he - reach;
		const int oz = (bz << SZ) - he - reach;
		const slice_entry *entries = p.entries + kept * p.map.reduced_count + chunk.begin;

		for (int v = lane; v < voxels; v += 64)
		{
			acc[v] = brick_voxel{0.0f, 0.0f, 0.0f};
		}
		lds_order();

		uint32_t at = begin + lane;
		bool have = at < end;
		bin_record record = have ? records[at] : bin_record{0.0f, 0.0f, 0.0f, 0u};
		uint32_t ahead = at + 64;
		bool have_ahead = ahead < end;
		bin_record next = have_ahead ? records[ahead] : bin_record{0.0f, 0.0f, 0.0f, 0u};

		while (__ballot(have) != 0)
		{
			const uint32_t key = record.key;
			const int column = static_cast<int>(key & column_mask);
			const int row = static_cast<int>((key >> g.column_bits) & row_mask);
			const bool negative = ((key >> sign_shift) & 1u) != 0;
			const uint32_t slice = key >> (sign_shift + 1);
			// Keys come from this chunk's producers; the test only keeps a
			// damaged record from reading outside the chunk's entries.
			const bool usable = have && slice < static_cast<uint32_t>(chunk.count);
			const slice_entry &e = entries[usable ? slice : 0];
			const int ky = row <= p.n / 2 ? row : row - p.n;
			float c[3];
			bin_position(e, column, ky, negative, c);
			float wx[T], wy[T], wz[T];
			int fx, fy, fz;
			if constexpr (PROFILE == REXLIB_HIP_KAISER_BESSEL)
			{
				fx = window_taps<PROFILE>::template place<T>(p, c[0], wx);
				fy = window_taps<PROFILE>::template place<T>(p, c[1], wy);
				fz = window_taps<PROFILE>::template place<T>(p, c[2], wz);
			}
			else
			{
				fx = window_taps<PROFILE>::place(p, c[0], wx);
				fy = window_taps<PROFILE>::place(p, c[1], wy);
				fz = window_taps<PROFILE>::place(p, c[2], wz);
			}
			// The cell in the brick; the producers placed the record by the
			// same arithmetic, so the clamps change nothing but keep every
			// LDS access inside the brick.
			const int lx = min(max(fx - ox, 0), EX - 1);
			const int ly = min(max(fy - oy, 0), EY - 1);
			const int lz = min(max(fz - oz, 0), EZ - 1);
			const int cell = lx + EX * (ly + EY * lz);
			if (usable)
			{
				claim[cell] = lane;
			}
			lds_order();
			const bool win = usable && claim_read[cell] == lane;
			lds_order();
			if (win)
			{
				const int base = lx + LX * (ly + LY * lz);
				const float re = record.re;
				const float im = record.im;
				const float den = record.den;
#pragma unroll
				for (int tz = 0; tz < T; ++tz)
				{
#pragma unroll
					for (int ty = 0; ty < T; ++ty)
					{
						const float wzy = wz[tz] * wy[ty];
#pragma unroll
						for (int tx = 0; tx < T; ++tx)
						{
							if (PROFILE == REXLIB_HIP_KAISER_BESSEL && (tx >= taps || ty >= taps || tz >= taps))
							{
								continue;
							}
							const float w = wzy * wx[tx];
							brick_voxel &a = acc[base + tx + LX * (ty + LY * tz)];
							brick_voxel s = a;
							s.re += re * w;
							s.im += im * w;
							s.den += den * w;
							a = s;
							lds_order();
						}
					}
				}
			}
			lds_order();
			// Lanes that added their sample move on; a record that cannot be
			// used is dropped.
			if (win || (have && !usable))
			{
				record = next;
				have = have_ahead;
				ahead += 64;
				have_ahead = ahead < end;
				if (have_ahead)
				{
					next = records[ahead];
				}
			}
		}
		lds_order();

		// Add the brick to the accumulators of its kept batch element.
		const int64_t num_base = kept_offset(p.map, slot_numerator_out, kept);
		const int64_t den_base = kept_offset(p.map, slot_denominator_out, kept);
		for (int v = lane; v < voxels; v += 64)
		{
			const int lx = v % LX;
			const int r = v / LX;
			const int ly = r % LY;
			const int lz = r / LY;
			const int x = ox + lx;
			const int y = oy + ly;
			const int z = oz + lz;
			if (x < 0 || x > he || y < -he || y > he || z < -he || z > he)
			{
				continue;
			}
			const brick_voxel sum = acc[v];
			if (sum.re == 0.0f && sum.im == 0.0f && sum.den == 0.0f)
			{
				continue;
			}
			const int sy = y < 0 ? y + big : y;
			const int sz = z < 0 ? z + big : z;
			float2 value = make_float2(sum.re, sum.im);
			if (odd)
			{
				value = cmul(cmul(cmul(value, p.volume_phase[x]), p.volume_phase[sy]), p.volume_phase[sz]);
			}
			else if ((x + sy + sz) & 1)
			{
				value = make_float2(-value.x, -value.y);
			}
			const bool alias = !odd && (y == he || y == -he || z == he || z == -he);
			// Lower neighbours' taps reach the first `reach` layers; the layers
			// from the interior extent on are the upper neighbours' cells.
			const bool owned = single && !alias && lx >= reach && ly >= reach && lz >= reach && lx < EX && ly < EY &&
				lz < EZ;
			float2 *target = p.numerator + num_base + sz * p.num_s0 + sy * p.num_s1 + x * p.num_s2;
			float *dtarget = p.denominator + den_base + sz * p.den_s0 + sy * p.den_s1 + x * p.den_s2;
			if (owned)
			{
				float2 t2 = *target;
				t2.x += value.x;
				t2.y += value.y;
				*target = t2;
				*dtarget += sum.den;
			}
			else
			{
				atomicAdd(&target->x, value.x);
				atomicAdd(&target->y, value.y);
				atomicAdd(dtarget, sum.den);
			}
		}
		lds_order();
	}
}

// 256-byte aligned bytes of one uint32 per bin.
std::size_t bin_array_bytes(int64_t bins)
{
	return (sizeof(uint32_t) * static_cast<std::size_t>(bins) + 255) / 256 * 256;
}

int bit_width(int64_t v)
{
	int bits = 0;
	while (v > 0)
	{
		++bits;
		v >>= 1;
	}
	return bits;
}

// Stored elements of an n-point slice that any radial weight can insert:
// not on a Nyquist row or column of an even n, shell (round-half-up |k|) at
// most n/2, i.e. 4 |k|^2 < (n + 1)^2.
int64_t insertable_elements(int n)
{
	const int h = n / 2 + 1;
	const int64_t limit = static_cast<int64_t>(2 * (h - 1) + 1) * (2 * (h - 1) + 1);
	int64_t count = 0;
	for (int row = 0; row < n; ++row)
	{
		if (n % 2 == 0 && row == n / 2)
		{
			continue;
		}
		const int64_t ky = row <= n / 2 ? row : row - n;
		for (int kx = 0; kx < h; ++kx)
		{
			if (n % 2 == 0 && kx == n / 2)
			{
				break;
			}
			if (4 * (static_cast<int64_t>(kx) * kx + ky * ky) >= limit)
			{
				break;
			}
			++count;
		}
	}
	return count;
}

struct tuning_entry
{
	int shift[3];
	int records_per_task;
};

std::mutex tuning_mutex;
std::map<int, tuning_entry> tuning_table;

bool shape_supported(int sx, int sy, int sz)
{
	return (sx == 3 && sy == 3 && sz == 3) || (sx == 4 && sy == 3 && sz == 3) || (sx == 3 && sy == 3 && sz == 2) ||
		(sx == 4 && sy == 3 && sz == 2);
}

template <int PROFILE, int SX, int SY, int SZ>
unsigned persistent_blocks()
{
	static std::mutex mutex;
	static std::map<int, unsigned> cache;
	int device = 0;
	REXLIB_HIP_CHECK(hipGetDevice(&device));
	std::lock_guard<std::mutex> lock(mutex);
	const auto found = cache.find(device);
	if (found != cache.end())
	{
		return found->second;
	}
	int per_cu = 0;
	REXLIB_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu,
		reinterpret_cast<const void *>(bin_brick_kernel<PROFILE, SX, SY, SZ>), 64, 0));
	hipDeviceProp_t prop;
	REXLIB_HIP_CHECK(hipGetDeviceProperties(&prop, device));
	const unsigned blocks = static_cast<unsigned>(std::max(1, per_cu) * std::max(1, prop.multiProcessorCount));
	cache[device] = blocks;
	return blocks;
}

template <int PROFILE, int SX, int SY, int SZ>
void launch_bin_brick_shape(const insertion_params &p, slice_chunk chunk, const bin_state &s, hipStream_t stream)
{
	const unsigned blocks = persistent_blocks<PROFILE, SX, SY, SZ>();
	bin_brick_kernel<PROFILE, SX, SY, SZ><<<blocks, 64, 0, stream>>>(p, chunk, s.sink.g, s.sink.records,
		s.sink.capacity, s.tasks, s.task_count, s.task_next);
}

template <int PROFILE>
void launch_bin_brick_profile(const insertion_params &p, slice_chunk chunk, const bin_state &s, hipStream_t stream)
{
	const int *sh = s.sink.g.shift;
	if (sh[0] == 4 && sh[2] == 3)
	{
		launch_bin_brick_shape<PROFILE, 4, 3, 3>(p, chunk, s, stream);
	}
	else if (sh[0] == 4)
	{
		launch_bin_brick_shape<PROFILE, 4, 3, 2>(p, chunk, s, stream);
	}
	else if (sh[2] == 2)
	{
		launch_bin_brick_shape<PROFILE, 3, 3, 2>(p, chunk, s, stream);
	}
	else
	{
		launch_bin_brick_shape<PROFILE, 3, 3, 3>(p, chunk, s, stream);
	}
}

} // anonymous namespace

void set_insertion_tuning(int extent, int shift_x, int shift_y, int shift_z, int records_per_task)
{
	if (extent < 0)
	{
		fail(REXLIB_HIP_INVALID_ARGUMENT, "insertion tuning: the extent must be 0 (every extent) or positive");
	}
	if (!shape_supported(shift_x, shift_y, shift_z))
	{
		fail(REXLIB_HIP_INVALID_ARGUMENT, "insertion tuning: brick shapes (log2 x, y, z) are (3, 3, 3), (4, 3, 3), "
			"(3, 3, 2) and (4, 3, 2)");
	}
	if (records_per_task < 256 || records_per_task > (1 << 24))
	{
		fail(REXLIB_HIP_INVALID_ARGUMENT, "insertion tuning: records per task lie in [256, 2^24]");
	}
	std::lock_guard<std::mutex> lock(tuning_mutex);
	tuning_table[extent] = tuning_entry{{shift_x, shift_y, shift_z}, records_per_task};
}

void get_insertion_tuning(int extent, int *shift_x, int *shift_y, int *shift_z, int *records_per_task)
{
	std::lock_guard<std::mutex> lock(tuning_mutex);
	auto found = tuning_table.find(extent);
	if (found == tuning_table.end())
	{
		found = tuning_table.find(0);
	}
	const tuning_entry t = found != tuning_table.end() ? found->second : tuning_entry{{3, 3, 3}, 8192};
	*shift_x = t.shift[0];
	*shift_y = t.shift[1];
	*shift_z = t.shift[2];
	*records_per_task = t.records_per_task;
}

void clear_insertion_tuning()
{
	std::lock_guard<std::mutex> lock(tuning_mutex);
	tuning_table.clear();
}

bin_layout plan_bins(const insertion_params &p, std::size_t budget)
{
	bin_layout l{};
	int shift[3];
	get_insertion_tuning(p.big, &shift[0], &shift[1], &shift[2], &l.records_per_task);
	bin_grid &g = l.g;
	const int he = p.big / 2;
	g.taps = p.taps;
	g.half_extent = he;
	const int reach = p.taps - 1;
	const int64_t cells[3] = {static_cast<int64_t>(he) + reach + 1, 2 * static_cast<int64_t>(he) + reach + 1,
		2 * static_cast<int64_t>(he) + reach + 1};
	g.bricks = 1;
	for (int a = 0; a < 3; ++a)
	{
		g.shift[a] = shift[a];
		g.count[a] = static_cast<int32_t>(((cells[a] - 1) >> shift[a]) + 1);
		g.bricks *= g.count[a];
	}
	g.column_bits = std::max(1, bit_width(p.h - 1));
	g.row_bits = std::max(1, bit_width(p.n - 1));
	const int slice_bits = 31 - g.column_bits - g.row_bits;
	l.bins = p.map.kept_count * g.bricks;
	l.per_slice = 2 * insertable_elements(p.n) * p.map.kept_count;
	l.usable = false;
	if (slice_bits < 1 || l.bins >= (int64_t(1) << 32) || l.per_slice == 0)
	{
		return l;
	}
	// The bin arrays (count, end, cursor), the two counters and one task per
	// bin come first; each record then costs 16 bytes and 16 / per_task of a
	// task. The slices of a chunk fit the rest and the key's slice field.
	const int64_t fixed = 3 * static_cast<int64_t>(bin_array_bytes(l.bins)) + 1024 +
		static_cast<int64_t>(sizeof(int4)) * (l.bins + 1);
	if (budget <= static_cast<std::size_t>(fixed))
	{
		return l;
	}
	const double per_record = static_cast<double>(sizeof(bin_record)) +
		static_cast<double>(sizeof(int4)) / static_cast<double>(l.records_per_task);
	const int64_t records = static_cast<int64_t>(static_cast<double>(static_cast<int64_t>(budget) - fixed) /
		per_record);
	l.chunk_slices = std::min<int64_t>(int64_t(1) << slice_bits, records / l.per_slice);
	l.chunk_slices = std::min<int64_t>(l.chunk_slices, std::max<int64_t>(1, p.map.reduced_count));
	if (l.chunk_slices < 1)
	{
		return l;
	}
	l.record_capacity = l.per_slice * l.chunk_slices;
	l.task_capacity = l.bins + l.record_capacity / l.records_per_task + 1;
	l.usable = l.record_capacity < (int64_t(1) << 32) && l.task_capacity < (int64_t(1) << 31);
	return l;
}

void launch_bin_count(const insertion_params &p, slice_chunk chunk, const bin_state &s, hipStream_t stream)
{
	const int64_t bins = s.bins;
	fill_u32_kernel<<<static_cast<unsigned>(std::min<int64_t>((bins + 255) / 256, 4096)), 256, 0, stream>>>(
		s.counts, bins, 0u);
	const int64_t elements = static_cast<int64_t>(p.n) * p.h;
	const dim3 grid(static_cast<unsigned>((elements + 255) / 256),
		static_cast<unsigned>(std::min<int64_t>(p.map.kept_count * chunk.count, 65535)));
	bin_count_kernel<<<grid, 256, 0, stream>>>(p, chunk, s.sink.g, s.counts);
	bin_plan_kernel<<<1, 1024, 0, stream>>>(s.counts, bins, static_cast<uint32_t>(s.records_per_task), s.end,
		s.sink.cursor, s.tasks, s.task_capacity, s.task_count, s.task_next);
	REXLIB_HIP_CHECK(hipGetLastError());
}

void launch_bin_brick(const insertion_params &p, slice_chunk chunk, const bin_state &s, hipStream_t stream)
{
	switch (p.profile)
	{
	case REXLIB_HIP_BOX:
		launch_bin_brick_profile<REXLIB_HIP_BOX>(p, chunk, s, stream);
		break;
	case REXLIB_HIP_TRIANGLE:
		launch_bin_brick_profile<REXLIB_HIP_TRIANGLE>(p, chunk, s, stream);
		break;
	default:
		launch_bin_brick_profile<REXLIB_HIP_KAISER_BESSEL>(p, chunk, s, stream);
		break;
	}
	REXLIB_HIP_CHECK(hipGetLastError());
}

} // namespace rexlib_hip
