// Fourier transforms of rexlib 0.1.0:
//   fourier_normalization      include/rexlib/ops/fourier/fourier_normalization.hpp
//   fft_operation              include/rexlib/ops/fourier/fft_operation.hpp
//   ifft_operation             include/rexlib/ops/fourier/ifft_operation.hpp
//   rfft_operation             include/rexlib/ops/fourier/rfft_operation.hpp:20-72
//   irfft_operation            include/rexlib/ops/fourier/irfft_operation.hpp:22-75
//   fftn / ifftn / rfftn / irfftn, rfft2 / irfft2 / irfft3
//                              include/rexlib/functional/fourier.hpp
// The operations hold absolute axes; the functions also take them from the end (-2, -1).
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

// Absolute axes (ops/parameters/axis_list.hpp).
using axis_list = std::vector<std::size_t>;

// The axes a transform runs over (ops/parameters/axis_selection.hpp).
class axis_selection
{
public:
	explicit axis_selection(axis_list axes) noexcept : m_axes(std::move(axes)) {}
	span<const std::size_t> get_axes() const noexcept { return span<const std::size_t>(m_axes.data(), m_axes.size()); }

private:
	axis_list m_axes;
};

// The axes of an inverse real transform and the restored extent along the
// last of them (ops/parameters/restored_signal_axes.hpp).
class restored_signal_axes
{
public:
	restored_signal_axes(axis_list axes, std::size_t extent) noexcept : m_axes(std::move(axes)), m_extent(extent) {}
	span<const std::size_t> get_axes() const noexcept { return span<const std::size_t>(m_axes.data(), m_axes.size()); }
	std::size_t get_extent() const noexcept { return m_extent; }

private:
	axis_list m_axes;
	std::size_t m_extent;
};

// A transform whose parameters are an axis selection (fft, ifft, rfft) or
// restored signal axes (irfft), read through get_parameters()
// (core/operation/parametric_operation.hpp:45).
template <typename Derived, typename Parameters>
class fourier_operation : public basic_operation<Derived>
{
public:
	using parameters_type = Parameters;
	template <typename... Args>
	explicit fourier_operation(fourier_normalization normalization, Args &&...args)
		: m_parameters(std::forward<Args>(args)...), m_normalization(normalization)
	{
	}
	const parameters_type &get_parameters() const noexcept { return m_parameters; }
	fourier_normalization get_normalization() const noexcept { return m_normalization; }

private:
	parameters_type m_parameters;
	fourier_normalization m_normalization;
};

// Complex transforms over `axes`, the remaining axes a batch.
class REXLIB_API fft_operation final : public fourier_operation<fft_operation, axis_selection>
{
public:
	explicit fft_operation(axis_list axes, fourier_normalization normalization = fourier_normalization::backward)
		: fourier_operation(normalization, std::move(axes))
	{
	}
	const char *get_name() const noexcept override { return "fft"; }
};

class REXLIB_API ifft_operation final : public fourier_operation<ifft_operation, axis_selection>
{
public:
	explicit ifft_operation(axis_list axes, fourier_normalization normalization = fourier_normalization::backward)
		: fourier_operation(normalization, std::move(axes))
	{
	}
	const char *get_name() const noexcept override { return "ifft"; }
};

class REXLIB_API rfft_operation final : public fourier_operation<rfft_operation, axis_selection>
{
public:
	explicit rfft_operation(axis_list axes, fourier_normalization normalization = fourier_normalization::backward)
		: fourier_operation(normalization, std::move(axes))
	{
	}
	const char *get_name() const noexcept override { return "rfft"; }
};

class REXLIB_API irfft_operation final : public fourier_operation<irfft_operation, restored_signal_axes>
{
public:
	irfft_operation(axis_list axes, std::size_t extent,
		fourier_normalization normalization = fourier_normalization::backward)
		: fourier_operation(normalization, std::move(axes), extent)
	{
	}
	const char *get_name() const noexcept override { return "irfft"; }
};

} // namespace ops

// Axes counted from the end when negative (-1 the last).
using signed_axes = std::vector<std::ptrdiff_t>;

// Transforms over any axes (every axis when `axes` is empty); for rfftn
// and irfftn the last axis listed is the halved one, and irfftn's `extent`
// is the output's extent along it.
REXLIB_API array fftn(const_array_ref x, signed_axes axes, const execution_context &context,
	ops::fourier_normalization normalization = ops::fourier_normalization::backward, array *out = nullptr);
REXLIB_API array ifftn(const_array_ref x, signed_axes axes, const execution_context &context,
	ops::fourier_normalization normalization = ops::fourier_normalization::backward, array *out = nullptr);
REXLIB_API array rfftn(const_array_ref x, signed_axes axes, const execution_context &context,
	ops::fourier_normalization normalization = ops::fourier_normalization::backward, array *out = nullptr);
REXLIB_API array irfftn(const_array_ref x, signed_axes axes, std::size_t extent, const execution_context &context,
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
