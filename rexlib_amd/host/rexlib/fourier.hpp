// Real Fourier transforms of rexlib 0.1.0, as the insertion path calls them:
//   fourier_normalization      include/rexlib/ops/fourier/fourier_normalization.hpp
//   rfft_operation             include/rexlib/ops/fourier/rfft_operation.hpp:20-72
//   irfft_operation            include/rexlib/ops/fourier/irfft_operation.hpp:22-75
//   rfft2 / irfft2 / irfft3    include/rexlib/functional/fourier.hpp:483-760
// Axes are given from the end (-2, -1) or as absolute indices.
#pragma once

#include "array.hpp"

namespace rexlib
{
namespace ops
{

enum class fourier_normalization
{
	backward,
	ortho,
	forward
};

using axis_list = std::vector<std::ptrdiff_t>;

class REXLIB_API rfft_operation final : public basic_operation<rfft_operation>
{
public:
	explicit rfft_operation(axis_list axes, fourier_normalization normalization = fourier_normalization::backward)
		: m_axes(std::move(axes)), m_normalization(normalization)
	{
	}
	const axis_list &get_axes() const noexcept { return m_axes; }
	fourier_normalization get_normalization() const noexcept { return m_normalization; }
	const char *get_name() const noexcept override { return "rfft"; }

private:
	axis_list m_axes;
	fourier_normalization m_normalization;
};

class REXLIB_API irfft_operation final : public basic_operation<irfft_operation>
{
public:
	irfft_operation(axis_list axes, std::size_t extent,
		fourier_normalization normalization = fourier_normalization::backward)
		: m_axes(std::move(axes)), m_extent(extent), m_normalization(normalization)
	{
	}
	const axis_list &get_axes() const noexcept { return m_axes; }
	std::size_t get_extent() const noexcept { return m_extent; }
	fourier_normalization get_normalization() const noexcept { return m_normalization; }
	const char *get_name() const noexcept override { return "irfft"; }

private:
	axis_list m_axes;
	std::size_t m_extent;
	fourier_normalization m_normalization;
};

} // namespace ops

REXLIB_API array rfft2(const_array_ref x, ops::fourier_normalization normalization, const execution_context &context,
	array *out = nullptr);
REXLIB_API array rfft2(const_array_ref x, const execution_context &context, array *out = nullptr);
REXLIB_API array irfft2(const_array_ref x, std::size_t extent, const execution_context &context,
	array *out = nullptr);
REXLIB_API array irfft3(const_array_ref x, std::size_t extent, ops::fourier_normalization normalization,
	const execution_context &context, array *out = nullptr);
REXLIB_API array irfft3(const_array_ref x, std::size_t extent, const execution_context &context,
	array *out = nullptr);

} // namespace rexlib
