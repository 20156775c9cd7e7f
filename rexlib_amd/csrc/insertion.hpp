// Device-side types of the insertion, shared by its kernels.
#pragma once

#include "layout.hpp"

namespace rexlib_hip
{

enum variant_kind : int
{
	variant_plain = 0,
	variant_weighted = 1,
	variant_ctf = 2
};

// What settle_slices computes for each slice
// (src/backends/cpu/builders/em/reconstruction/slice_insertion_functor.inl:158-269):
// the frame, the last shell with a nonzero radial weight, the CTF
// coefficients, and the element offsets of the slice's operands.
struct alignas(16) slice_entry
{
	float x_step[3];   // (N / n) * first row of A, rounded to float
	int32_t last_shell; // -1: nothing to insert (no weight, or a non-finite frame)
	double y_step[3];  // (N / n) * second row of A, in double for the row base
	float ctf_radial;  // pi L D s^2 = radial |k|^2 + axial (kx^2 - ky^2) + diagonal kx ky
	float ctf_axial;
	float ctf_diagonal;
	float ctf_aberration; // (pi / 2) Cs L^3 s^4 = aberration |k|^4
	float ctf_phase;      // phase shift + asin(Q)
	int32_t lattice_free; // 1: samples two steps apart on both axes never share a voxel cell
	int64_t slice_offset;  // elements of the complex slices
	int64_t extra_offset;  // elements of the weights (weighted) or CTF parameters
	int64_t radial_offset;
	int64_t numerator_offset;   // of the output numerator
	int64_t denominator_offset; // of the output denominator
	int64_t kept;
};

// The range of slices one launch covers, per kept accumulator: reduced
// indices [begin, begin + count).
struct slice_chunk
{
	int64_t begin;
	int64_t count;
};

// Everything an insertion kernel reads, by value.
struct insertion_params
{
	batch_map map;
	int32_t n;          // slice extent
	int32_t h;          // n / 2 + 1
	int32_t big;        // volume extent N
	int32_t half;       // N / 2 + 1
	int32_t variant;
	int32_t profile;
	int32_t taps;       // window taps per axis
	int32_t kb_count;   // Kaiser-Bessel table steps over the half-width
	float kb_half_width;
	float kb_scale;     // steps per voxel
	int64_t slice_count; // kept * reduced
	// Operand base addresses and core strides (elements).
	const float2 *slices;
	int64_t slice_s0, slice_s1;
	const float *transforms;
	int64_t transform_s0, transform_s1;
	const float *radial;
	int64_t radial_s0;
	const float *extra;
	int64_t extra_s0, extra_s1;
	float2 *numerator;
	int64_t num_s0, num_s1, num_s2;
	float *denominator;
	int64_t den_s0, den_s1, den_s2;
	// CTF constants, computed in double on the host.
	double ctf_defocus_scale; // pi L / (n p)^2
	double ctf_aberration;    // (pi / 2) Cs L^3 / (n p)^4
	double ctf_phase0;        // asin(Q)
	// Workspace.
	slice_entry *entries;
	float2 *tables;           // per slice: h x phases then n y phases
	const float *kb_table;    // kb_count + 3 entries
	const float2 *volume_phase; // N entries: exp(-2 pi i f(j) C / N) (odd N)
};

} // namespace rexlib_hip
