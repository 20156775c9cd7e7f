// Operand checks and the batch plan of the insertions, on the host.
//
// Shape and type rules follow include/rexlib/em/ops/reconstruction/slice_insertion_rules.hpp:41-183
// (cores (N, N, N/2+1), (n, n/2+1), (3, 4), (4), (n/2+1); one element type,
// complex for the numerator and the slices). The batch plan follows
// src/backends/cpu/plans/slice_insertion_layout_plan.cpp:81-185: the batch
// prefixes broadcast together, right-aligned; an axis the accumulator lacks
// or holds with an extent of 1, while the frame has more, is reduced (summed
// into the accumulator); the others are kept.
#pragma once

#include "common.hpp"

#include <algorithm>
#include <sstream>
#include <vector>

namespace rexlib_hip
{

constexpr int max_batch_rank = REXLIB_HIP_MAX_RANK;

// Operand slots of an insertion, in a fixed order independent of the variant.
enum slot : int
{
	slot_numerator = 0,
	slot_denominator = 1,
	slot_slices = 2,
	slot_transforms = 3,
	slot_radial = 4,
	slot_extra = 5, // weights or CTF parameters
	slot_numerator_out = 6,
	slot_denominator_out = 7,
	slot_count = 8
};

// Batch addressing handed to the device: the linear index of a slice is
// kept * reduced_count + reduced, and each index decomposes over its axes
// (last axis fastest) into element offsets per operand slot.
struct batch_map
{
	int32_t kept_rank;
	int32_t reduced_rank;
	int64_t kept_count;
	int64_t reduced_count;
	int64_t kept_extents[max_batch_rank];
	int64_t reduced_extents[max_batch_rank];
	int64_t kept_strides[slot_count][max_batch_rank];
	int64_t reduced_strides[slot_count][max_batch_rank];
};

inline const char *dtype_name(int dtype)
{
	switch (dtype)
	{
	case REXLIB_HIP_FLOAT32: return "float32";
	case REXLIB_HIP_FLOAT64: return "float64";
	case REXLIB_HIP_COMPLEX64: return "complex64";
	case REXLIB_HIP_COMPLEX128: return "complex128";
	case REXLIB_HIP_FLOAT16: return "float16";
	case REXLIB_HIP_BFLOAT16: return "bfloat16";
	case REXLIB_HIP_BOOL: return "bool";
	case REXLIB_HIP_UINT64: return "uint64";
	default: return "unknown";
	}
}

inline std::string shape_string(const rexlib_hip_array &a)
{
	std::ostringstream oss;
	oss << "(";
	for (int i = 0; i < a.rank; ++i)
	{
		oss << (i ? ", " : "") << a.extents[i];
	}
	oss << ")";
	return oss.str();
}

inline void check_array(const rexlib_hip_array *a, const char *name, int core_rank)
{
	if (a == nullptr)
	{
		fail(REXLIB_HIP_INVALID_ARGUMENT, std::string(name) + ": missing operand");
	}
	if (a->rank < core_rank || a->rank > REXLIB_HIP_MAX_RANK)
	{
		std::ostringstream oss;
		oss << name << ": rank " << a->rank << " does not hold the " << core_rank << " core axes";
		fail(REXLIB_HIP_INVALID_ARGUMENT, oss.str());
	}
	for (int i = 0; i < a->rank; ++i)
	{
		if (a->extents[i] < 0)
		{
			fail(REXLIB_HIP_INVALID_ARGUMENT, std::string(name) + ": negative extent");
		}
	}
}

inline int64_t core_extent(const rexlib_hip_array &a, int from_end)
{
	return a.extents[a.rank - from_end];
}

inline int64_t core_stride(const rexlib_hip_array &a, int from_end)
{
	return a.strides[a.rank - from_end];
}

inline void expect_core(const rexlib_hip_array &a, const char *name, std::initializer_list<int64_t> core)
{
	int k = static_cast<int>(core.size());
	int i = 0;
	for (const auto e : core)
	{
		if (a.extents[a.rank - k + i] != e)
		{
			std::ostringstream oss;
			oss << name << ": shape " << shape_string(a) << " does not end with the expected core (";
			int j = 0;
			for (const auto f : core)
			{
				oss << (j++ ? ", " : "") << f;
			}
			oss << ")";
			fail(REXLIB_HIP_INVALID_ARGUMENT, oss.str());
		}
		++i;
	}
}

// Element types: every operand float32, complex64 for the numerator and the
// slices. A consistent float64 request is valid for the operation but has no
// program here (unsupported, so a dispatcher can fall through to the CPU).
inline void check_types(const rexlib_hip_array *const *ops, const bool *complex_slot, int count,
	const char *const *names)
{
	bool all_double = true;
	bool all_single = true;
	for (int i = 0; i < count; ++i)
	{
		if (ops[i] == nullptr)
		{
			continue;
		}
		const int want_single = complex_slot[i] ? REXLIB_HIP_COMPLEX64 : REXLIB_HIP_FLOAT32;
		const int want_double = complex_slot[i] ? REXLIB_HIP_COMPLEX128 : REXLIB_HIP_FLOAT64;
		all_single = all_single && ops[i]->dtype == want_single;
		all_double = all_double && ops[i]->dtype == want_double;
	}
	if (all_single)
	{
		return;
	}
	if (all_double)
	{
		fail(REXLIB_HIP_UNSUPPORTED, "insertion: float64 operands have no gfx950 program");
	}
	std::ostringstream oss;
	oss << "insertion: operand element types must be float32 (complex64 for the numerator and "
		"the slices), but they are";
	for (int i = 0; i < count; ++i)
	{
		if (ops[i] != nullptr)
		{
			oss << " " << names[i] << "=" << dtype_name(ops[i]->dtype);
		}
	}
	fail(REXLIB_HIP_INVALID_ARGUMENT, oss.str());
}

// Broadcast the batch prefixes and split kept from reduced axes.
inline batch_map make_batch_map(const rexlib_hip_array *const ops[slot_count], const int core_ranks[slot_count],
	const int accumulator_slot_of[slot_count])
{
	// The frame: the broadcast of every input prefix (slots 0..5).
	int frame_rank = 0;
	for (int s = 0; s <= slot_extra; ++s)
	{
		if (ops[s] != nullptr)
		{
			frame_rank = std::max(frame_rank, ops[s]->rank - core_ranks[s]);
		}
	}
	std::vector<int64_t> frame(static_cast<std::size_t>(frame_rank), 1);
	for (int s = 0; s <= slot_extra; ++s)
	{
		if (ops[s] == nullptr)
		{
			continue;
		}
		const int prefix = ops[s]->rank - core_ranks[s];
		for (int i = 0; i < prefix; ++i)
		{
			const int ax = frame_rank - prefix + i;
			const int64_t e = ops[s]->extents[i];
			if (frame[ax] == 1)
			{
				frame[ax] = e;
			}
			else if (e != 1 && e != frame[ax])
			{
				fail(REXLIB_HIP_INVALID_ARGUMENT,
					"insertion: the batch axes of the operands do not broadcast together");
			}
		}
	}

	// Strides of each slot over the frame (0 where broadcast).
	std::vector<std::vector<int64_t>> strides(slot_count, std::vector<int64_t>(frame_rank, 0));
	for (int s = 0; s < slot_count; ++s)
	{
		if (ops[s] == nullptr)
		{
			continue;
		}
		const int prefix = ops[s]->rank - core_ranks[s];
		for (int i = 0; i < prefix; ++i)
		{
			const int ax = frame_rank - prefix + i;
			strides[s][ax] = ops[s]->extents[i] == 1 ? 0 : ops[s]->strides[i];
		}
	}

	// The accumulator decides: its axes of extent 1 (or missing) under a
	// frame axis of more than 1 are reduced.
	const rexlib_hip_array *acc = ops[slot_numerator];
	const int acc_prefix = acc->rank - core_ranks[slot_numerator];
	batch_map map{};
	for (int ax = 0; ax < frame_rank; ++ax)
	{
		const int i = ax - (frame_rank - acc_prefix);
		const int64_t acc_extent = i >= 0 ? acc->extents[i] : 1;
		const bool reduced = acc_extent == 1 && frame[ax] > 1;
		if (reduced)
		{
			const int r = map.reduced_rank++;
			map.reduced_extents[r] = frame[ax];
			for (int s = 0; s < slot_count; ++s)
			{
				map.reduced_strides[s][r] = strides[s][ax];
			}
		}
		else
		{
			const int k = map.kept_rank++;
			map.kept_extents[k] = frame[ax];
			for (int s = 0; s < slot_count; ++s)
			{
				map.kept_strides[s][k] = strides[s][ax];
			}
		}
	}
	map.kept_count = 1;
	for (int k = 0; k < map.kept_rank; ++k)
	{
		map.kept_count *= map.kept_extents[k];
	}
	map.reduced_count = 1;
	for (int r = 0; r < map.reduced_rank; ++r)
	{
		map.reduced_count *= map.reduced_extents[r];
	}
	(void)accumulator_slot_of;
	return map;
}

__host__ __device__ inline int64_t kept_offset(const batch_map &m, int slot, int64_t kept)
{
	int64_t offset = 0;
	for (int k = m.kept_rank - 1; k >= 0; --k)
	{
		const int64_t e = m.kept_extents[k];
		offset += (kept % e) * m.kept_strides[slot][k];
		kept /= e;
	}
	return offset;
}

__host__ __device__ inline int64_t reduced_offset(const batch_map &m, int slot, int64_t reduced)
{
	int64_t offset = 0;
	for (int r = m.reduced_rank - 1; r >= 0; --r)
	{
		const int64_t e = m.reduced_extents[r];
		offset += (reduced % e) * m.reduced_strides[slot][r];
		reduced /= e;
	}
	return offset;
}

} // namespace rexlib_hip
