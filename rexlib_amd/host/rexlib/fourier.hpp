// Fourier transforms of rexlib 0.1.0:
//   fourier_normalization      include/rexlib/ops/fourier/fourier_normalization.hpp
//   fft_operation              include/rexlib/ops/fourier/fft_operation.hpp
//   ifft_operation             include/rexlib/ops/fourier/ifft_operation.hpp
//   rfft_operation             include/rexlib/ops/fourier/rfft_operation.hpp:20-72
//   irfft_operation            include/rexlib/ops/fourier/irfft_operation.hpp:22-75
//   fftn / ifftn / rfftn / irfftn, rfft2 / irfft2 / irfft3
//                              include/rexlib/functional/fourier.hpp
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

// Complex transforms over `axes`, the remaining axes a batch.
class REXLIB_API fft_operation final : public basic_operation<fft_operation>
{
public:
	explicit fft_operation(axis_list axes, fourier_normalization normalization = fourier_normalization::backward)
		: m_axes(std::move(axes)), m_normalization(normalization)
	{
	}
	const axis_list &get_axes() const noexcept { return m_axes; }
	fourier_normalization get_normalization() const noexcept { return m_normalization; }
	const char *get_name() const noexcept override { return "fft"; }

private:
	axis_list m_axes;
	fourier_normalization m_normalization;
};

class REXLIB_API ifft_operation final : public basic_operation<ifft_operation>
{
public:
	explicit ifft_operation(axis_list axes, fourier_normalization normalization = fourier_normalization::backward)
		: m_axes(std::move(axes)), m_normalization(normalization)
	{
	}
	const axis_list &get_axes() const noexcept { return m_axes; }
	fourier_normalization get_normalization() const noexcept { return m_normalization; }
	const char *get_name() const noexcept override { return "ifft"; }

private:
	axis_list m_axes;
	fourier_normalization m_normalization;
};

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

// Transforms over any axes (every axis when `axes` is empty); for rfftn
// and irfftn the last axis listed is the halved one, and irfftn's `extent`
// is the output's extent along it.
REXLIB_API array fftn(const_array_ref x, ops::axis_list axes, const execution_context &context,
	ops::fourier_normalization normalization = ops::fourier_normalization::backward, array *out = nullptr);
REXLIB_API array ifftn(const_array_ref x, ops::axis_list axes, const execution_context &context,
	ops::fourier_normalization normalization = ops::fourier_normalization::backward, array *out = nullptr);
REXLIB_API array rfftn(const_array_ref x, ops::axis_list axes, const execution_context &context,
	ops::fourier_normalization normalization = ops::fourier_normalization::backward, array *out = nullptr);
REXLIB_API array irfftn(const_array_ref x, ops::axis_list axes, std::size_t extent, const execution_context &context,
	ops::fourier_normalization normalization = ops::fourier_normalization::backward, array *out = nullptr);

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
