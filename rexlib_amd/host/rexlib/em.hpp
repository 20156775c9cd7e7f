// rexlib::em reconstruction API: parameters, operations and functions, with
// the names, signatures and argument meaning of rexlib 0.1.0:
//   gridding_profile           include/rexlib/em/ops/parameters/gridding_profile.hpp:14-44
//   gridding_window            include/rexlib/em/ops/parameters/gridding_window.hpp:13-90
//   ctf_optics                 include/rexlib/em/ops/parameters/ctf_optics.hpp; wavelength
//                              src/em/ops/parameters/ctf_optics.cpp:81-85
//   insert_*_operation         include/rexlib/em/ops/reconstruction/insert_slices_operation.hpp:16-141,
//                              insert_weighted_slices_operation.hpp, insert_ctf_slices_operation.hpp:17-116
//   insert_* functions         include/rexlib/em/functional/reconstruction.hpp:47-149,
//                              src/em/functional/reconstruction.cpp:24-137
//   standardize_background_operation
//                              include/rexlib/em/ops/preprocessing/standardize_background_operation.hpp:15-118
//   standardize_background_operation
//                              include/rexlib/em/ops/preprocessing/standardize_background_operation.hpp:15-118
//
// The functions check parameters as the reference's operation constructors
// do (std::invalid_argument) and leave operand checks to the program, which
// reports them as std::invalid_argument too. Outputs default to new arrays
// of the accumulators' layout; passing the accumulators themselves as
// numerator_out / denominator_out accumulates in place.
#pragma once

#include "fourier.hpp"

#include <utility>

namespace rexlib
{
namespace em
{

enum class gridding_profile
{
	box,
	triangle,
	kaiser_bessel
};

class REXLIB_API gridding_window
{
public:
	static gridding_window nearest() noexcept;
	static gridding_window linear() noexcept;
	// Throws std::invalid_argument unless half_width > 0 and shape >= 0 are finite.
	static gridding_window kaiser_bessel(double half_width, double shape);
	gridding_profile get_profile() const noexcept { return m_profile; }
	double get_half_width() const noexcept { return m_half_width; }
	double get_shape() const noexcept { return m_shape; }

private:
	gridding_window(gridding_profile profile, double half_width, double shape) noexcept
		: m_profile(profile), m_half_width(half_width), m_shape(shape)
	{
	}
	gridding_profile m_profile;
	double m_half_width;
	double m_shape;
};

class REXLIB_API ctf_optics
{
public:
	// voltage in kV, spherical aberration in mm, amplitude contrast in [0, 1);
	// throws std::invalid_argument otherwise.
	ctf_optics(double voltage, double spherical_aberration, double amplitude_contrast);
	double get_voltage() const noexcept { return m_voltage; }
	double get_spherical_aberration() const noexcept { return m_spherical_aberration; }
	double get_amplitude_contrast() const noexcept { return m_amplitude_contrast; }
	// Relativistic electron wavelength in angstrom.
	double get_wavelength() const noexcept;

private:
	double m_voltage;
	double m_spherical_aberration;
	double m_amplitude_contrast;
};

class REXLIB_API insert_slices_operation final : public basic_operation<insert_slices_operation>
{
public:
	explicit insert_slices_operation(const gridding_window &window) noexcept : m_window(window) {}
	const gridding_window &get_window() const noexcept { return m_window; }
	const char *get_name() const noexcept override { return "insert_slices"; }

private:
	gridding_window m_window;
};

class REXLIB_API insert_weighted_slices_operation final : public basic_operation<insert_weighted_slices_operation>
{
public:
	explicit insert_weighted_slices_operation(const gridding_window &window) noexcept : m_window(window) {}
	const gridding_window &get_window() const noexcept { return m_window; }
	const char *get_name() const noexcept override { return "insert_weighted_slices"; }

private:
	gridding_window m_window;
};

class REXLIB_API insert_ctf_slices_operation final : public basic_operation<insert_ctf_slices_operation>
{
public:
	// Throws std::invalid_argument unless pixel_size is finite and positive.
	insert_ctf_slices_operation(const gridding_window &window, const ctf_optics &optics, double pixel_size);
	const gridding_window &get_window() const noexcept { return m_window; }
	const ctf_optics &get_optics() const noexcept { return m_optics; }
	double get_pixel_size() const noexcept { return m_pixel_size; }
	const char *get_name() const noexcept override { return "insert_ctf_slices"; }

private:
	gridding_window m_window;
	ctf_optics m_optics;
	double m_pixel_size;
};

// Standardizes the images along `axes` against their background (a
// result from a value and a uint64 key); throws std::invalid_argument when
// an axis is repeated, the radius is negative or NaN, or the bounds do not
// lie on either side of 0.
class REXLIB_API standardize_background_operation final : public basic_operation<standardize_background_operation>
{
public:
	using parameters_type = ops::axis_selection;
	standardize_background_operation(ops::axis_list axes, double radius, double lower_bound, double upper_bound);
	const ops::axis_selection &get_parameters() const noexcept { return m_axes; }
	double get_radius() const noexcept { return m_radius; }
	double get_lower_bound() const noexcept { return m_lower_bound; }
	double get_upper_bound() const noexcept { return m_upper_bound; }
	const char *get_name() const noexcept override { return "standardize_background"; }

private:
	ops::axis_selection m_axes;
	double m_radius;
	double m_lower_bound;
	double m_upper_bound;
};

REXLIB_API array standardize_background(const_array_ref value, const_array_ref key, ops::axis_list axes, double radius,
	double lower_bound, double upper_bound, const execution_context &context, array *out = nullptr);

REXLIB_API std::pair<array, array> insert_slices(const_array_ref numerator, const_array_ref denominator,
	const_array_ref slices, const_array_ref transforms, const_array_ref radial_weights,
	const gridding_window &window, const execution_context &context, array *numerator_out = nullptr,
	array *denominator_out = nullptr);

REXLIB_API std::pair<array, array> insert_weighted_slices(const_array_ref numerator, const_array_ref denominator,
	const_array_ref slices, const_array_ref weights, const_array_ref transforms, const_array_ref radial_weights,
	const gridding_window &window, const execution_context &context, array *numerator_out = nullptr,
	array *denominator_out = nullptr);

REXLIB_API std::pair<array, array> insert_ctf_slices(const_array_ref numerator, const_array_ref denominator,
	const_array_ref slices, const_array_ref transforms, const_array_ref ctf, const_array_ref radial_weights,
	const gridding_window &window, const ctf_optics &optics, double pixel_size, const execution_context &context,
	array *numerator_out = nullptr, array *denominator_out = nullptr);

} // namespace em
} // namespace rexlib
