// Binned insertion: the sample records and the device code that the sample
// producers and the brick kernel share.
//
// The volume is cut into bricks of first-tap cells: a sample belongs to the
// brick holding the first voxel of its window (the floor of its position
// for the triangle window), and its taps reach at most taps - 1 voxels past
// that brick, into a halo the brick also holds in LDS. A counting pass sizes
// one bin of records per (accumulator, brick); the producers then write
// every sample once, into its bin, and a wavefront owns one bin's share at a
// time. This is the reference's tile-owned gather
// (src/backends/cpu/builders/em/reconstruction/slice_insertion_functor.inl:105-156,
// 271-465), with the per-tile row bounding replaced by sorting the samples to
// their tiles once.
#pragma once

#include "bin_layout.hpp"
#include "sample.hpp"

namespace rexlib_hip
{

// Volume position of sample k_x of row k_y, or of its -k copy: the row base
// k_y Y in double rounded to float, then base + k_x X as one fused
// multiply-add, so that the producers and the brick kernel compute the same
// bits (row_insertion_kernel.inl:157-166). The -k copy is the exact negation.
__device__ inline void bin_position(const slice_entry &e, int kx, int ky, bool negative, float *c)
{
	const float fx = static_cast<float>(kx);
	for (int a = 0; a < 3; ++a)
	{
		const float base = static_cast<float>(static_cast<double>(ky) * e.y_step[a]);
		const float position = __builtin_fmaf(fx, e.x_step[a], base);
		c[a] = negative ? -position : position;
	}
}

// The first voxel of the window along one axis (window_taps<>::place).
__device__ inline int first_tap(const insertion_params &p, float c)
{
	switch (p.profile)
	{
	case REXLIB_HIP_BOX:
		return window_taps<REXLIB_HIP_BOX>::first(p, c);
	case REXLIB_HIP_TRIANGLE:
		return window_taps<REXLIB_HIP_TRIANGLE>::first(p, c);
	default:
		return window_taps<REXLIB_HIP_KAISER_BESSEL>::first(p, c);
	}
}

// The brick of a sample at c, or false when none of its taps lands on a
// stored voxel (x in [0, N/2], y and z in [-(N/2), N/2]).
__device__ inline bool bin_locate(const insertion_params &p, const bin_grid &g, const float *c, uint32_t &brick)
{
	const int reach = g.taps - 1;
	const int he = g.half_extent;
	int u[3];
	bool inside = true;
	for (int a = 0; a < 3; ++a)
	{
		const int first = first_tap(p, c[a]);
		const int low = a == 0 ? 0 : -he;
		inside = inside && first <= he && first >= low - reach;
		u[a] = first - low + reach;
	}
	if (!inside)
	{
		return false;
	}
	brick = static_cast<uint32_t>(((u[2] >> g.shift[2]) * g.count[1] + (u[1] >> g.shift[1])) * g.count[0] +
		(u[0] >> g.shift[0]));
	return true;
}

// One record per lane (valid or not) into its bin, or (COUNT) one count.
// Called by every lane of the wavefront. Lanes of equal bins in a run take
// consecutive records: the first lane of each run reserves the run with one
// atomic on the bin's cursor and hands the base to the others. A record that
// would pass its bin's end (the counting pass saw fewer) is dropped rather
// than written into another bin.
template <bool COUNT>
__device__ inline void bin_emit(const bin_sink &sink, uint32_t *counts, bool valid, uint32_t bin,
	const bin_record &record)
{
	if (__ballot(valid) == 0)
	{
		return;
	}
	const int lane = static_cast<int>(__lane_id());
	const uint32_t key = valid ? bin : 0xffffffffu;
	const uint32_t before = static_cast<uint32_t>(previous_lane(static_cast<int>(key)));
	const bool head = valid && (lane == 0 || before != key);
	const unsigned long long bounds = __ballot(head || !valid);
	const unsigned long long above = lane == 63 ? 0ull : bounds & (~0ull << (lane + 1));
	const int next = above != 0ull ? __ffsll(static_cast<long long>(above)) - 1 : 64;
	const uint32_t length = static_cast<uint32_t>(next - lane);
	if (COUNT)
	{
		if (head)
		{
			atomicAdd(&counts[bin], length);
		}
		return;
	}
	const int leader = wave_prefix_max(head ? lane : 0);
	uint32_t base = 0;
	uint32_t room = 0;
	if (head)
	{
		base = atomicAdd(&sink.cursor[bin], length);
		const uint32_t end = min(sink.end[bin], sink.capacity);
		room = end > base ? end - base : 0u;
	}
	base = static_cast<uint32_t>(__shfl(static_cast<int>(base), leader));
	room = static_cast<uint32_t>(__shfl(static_cast<int>(room), leader));
	const uint32_t rank = static_cast<uint32_t>(lane - leader);
	if (valid && rank < room)
	{
		sink.records[base + rank] = record;
	}
}

// Both copies of stored element (row, column) of slice `slice` (of the
// chunk) into the bins of accumulator `kept`: the +k copy, and the -k copy
// unless column is 0 (row_insertion_kernel.inl:222-279). The copy whose
// position has x >= 0 goes first, so that the second round, whose taps reach
// x >= 0 only near the x = 0 plane, is mostly skipped. `inserted` as
// sample_inserted; `v` is the prepared sample (unused when counting).
template <bool COUNT>
__device__ inline void bin_put(const insertion_params &p, const bin_sink &sink, uint32_t *counts,
	const slice_entry &e, int64_t kept, uint32_t slice, int row, int column, bool inserted, prepared_sample v)
{
	const int ky = row <= p.n / 2 ? row : row - p.n;
	float c[3];
	bin_position(e, column, ky, false, c);
	const bool first_negative = column != 0 && c[0] < 0.0f;
	const uint64_t kept_base = static_cast<uint64_t>(kept) * static_cast<uint64_t>(sink.g.bricks);
	for (int round = 0; round < 2; ++round)
	{
		const bool negative = round == 0 ? first_negative : !first_negative;
		const float sc[3] = {negative ? -c[0] : c[0], negative ? -c[1] : c[1], negative ? -c[2] : c[2]};
		uint32_t brick = 0;
		const bool valid = inserted && (round == 0 || column != 0) && bin_locate(p, sink.g, sc, brick);
		const bin_record record{v.re, negative ? -v.im : v.im, v.den, bin_key(sink.g, slice, negative, row, column)};
		bin_emit<COUNT>(sink, counts, valid, static_cast<uint32_t>(kept_base + brick), record);
	}
}

} // namespace rexlib_hip
