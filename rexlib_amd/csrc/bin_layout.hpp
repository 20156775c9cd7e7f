// The layout of the binned insertion's sample bins, shared by the host
// launchers and the kernels (bins.hpp, bins.hip, insertion.hip).
#pragma once

#include "insertion.hpp"

namespace rexlib_hip
{

// A sample in its bin: the value with its radial and numerator weights
// (conjugated for the -k copy), the denominator weight, and its key:
// (slice of the chunk, sign, row, column) as bit fields, column lowest.
struct alignas(16) bin_record
{
	float re;
	float im;
	float den;
	uint32_t key;
};

// The brick grid over first-tap coordinates. Along x the first taps of
// inserted samples run over [-(taps - 1), N/2], along y and z over
// [-(N/2) - (taps - 1), N/2]; u = first - low + (taps - 1) counts them from 0,
// and brick b along an axis holds u in [b << shift, (b + 1) << shift).
struct bin_grid
{
	int32_t shift[3];     // log2 of the interior extents along (x, y, z)
	int32_t count[3];     // bricks along (x, y, z)
	int32_t taps;         // window taps per axis (at most 4)
	int32_t half_extent;  // N / 2 (integer division): voxels x in [0, it], y and z in [-it, it]
	int32_t column_bits;  // key field widths
	int32_t row_bits;
	int64_t bricks;       // count[0] * count[1] * count[2]
};

// Where the producers write: per bin a cursor (next free record) and an end.
struct bin_sink
{
	bin_grid g;
	uint32_t *cursor;
	const uint32_t *end;
	bin_record *records;
	uint32_t capacity; // records the buffer holds
};

__host__ __device__ inline uint32_t bin_key(const bin_grid &g, uint32_t slice, bool negative, int row, int column)
{
	return (slice << (g.row_bits + g.column_bits + 1)) | (static_cast<uint32_t>(negative) << (g.row_bits + g.column_bits)) |
		(static_cast<uint32_t>(row) << g.column_bits) | static_cast<uint32_t>(column);
}

// What plan_bins decides for one insertion call: the brick grid, the
// records per task, and the chunking of the slices so that one chunk's
// records (at most both copies of every insertable element) and its key
// fields fit.
struct bin_layout
{
	bin_grid g;
	int records_per_task;
	int64_t bins;            // accumulators x bricks
	int64_t per_slice;       // records one slice can produce, over all accumulators
	int64_t chunk_slices;
	int64_t record_capacity; // per chunk
	int64_t task_capacity;
	bool usable;             // false: the scatter schedule takes the call
};

// Device buffers of one call's bins.
struct bin_state
{
	bin_sink sink;        // grid, cursors, ends, records
	uint32_t *counts;
	uint32_t *end;
	int4 *tasks;          // (bin, first record, end record, shares of the bin)
	uint32_t *task_count;
	uint32_t *task_next;
	int64_t bins;
	int records_per_task;
	uint32_t task_capacity;
};

} // namespace rexlib_hip
