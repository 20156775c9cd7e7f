/*
 * TEST INFRASTRUCTURE — the oracle. Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load this library, and only as a checker.
 *
 * A double-precision, brute-force restatement of rexlib's slice-insertion
 * contract, written from the contract text and from the literal oracle in
 * the reference's tests:
 *   - contract: include/rexlib/em/ops/reconstruction/insert_slices_operation.hpp:31-112,
 *     insert_ctf_slices_operation.hpp:33-65, insert_weighted_slices_operation.hpp:32-44;
 *   - fixture oracle: tests/unitary/src/backends/cpu/builders/em/reconstruction/
 *     fixtures/slice_insertion_cases.hpp:300-322 (window weights),
 *     327-358 (reached voxels), 363-385 (reference_ctf), 396-563 (reference()).
 * Every frequency of the full spectrum of each slice is visited once, with
 * both signs, and every voxel its window reaches is updated, in double.
 *
 * Batches are expanded by the caller: K accumulators, each receiving B
 * slices, with one slice, transform, radial, and extra row per (k, b).
 */
#include <math.h>
#include <stdint.h>
#include <stddef.h>
#include <stdlib.h>

#define ORACLE_PI 3.14159265358979323846

enum { PROFILE_BOX = 0, PROFILE_TRIANGLE = 1, PROFILE_KAISER_BESSEL = 2 };
enum { VARIANT_PLAIN = 0, VARIANT_WEIGHTED = 1, VARIANT_CTF = 2 };

/* Power series of I0, summed until a term falls below one epsilon of the
 * sum (src/backends/cpu/kernels/bessel_i0.cpp:12-26). */
double oracle_bessel_i0(double x)
{
	const double quarter_square = 0.25 * x * x;
	double sum = 1.0, term = 1.0;
	for (double k = 1.0;; k += 1.0) {
		term *= quarter_square / (k * k);
		sum += term;
		if (term <= 2.220446049250313e-16 * sum)
			return sum;
	}
}

/* lambda = 12.2643247 / sqrt(V (1 + 0.978466e-6 V)), V in volts
 * (src/em/ops/parameters/ctf_optics.cpp:81-85). */
double oracle_wavelength(double voltage_kv)
{
	const double volts = voltage_kv * 1e3;
	return 12.2643247 / sqrt(volts * (1.0 + 0.978466e-6 * volts));
}

/* The window weight of one axis (fixture slice_insertion_cases.hpp:300-322). */
double oracle_window_weight(int profile, double half_width, double shape, double distance)
{
	switch (profile) {
	case PROFILE_BOX:
		return 1.0;
	case PROFILE_TRIANGLE:
		return 1.0 - fabs(distance);
	default: {
		const double ratio = distance / half_width;
		if (!(fabs(ratio) < 1.0))
			return 0.0;
		return oracle_bessel_i0(shape * sqrt(1.0 - ratio * ratio)) / oracle_bessel_i0(shape);
	}
	}
}

/* The voxels a sample at `c` reaches along one axis, as [first, last]
 * (fixture slice_insertion_cases.hpp:327-358). Box rounds halves away from 0
 * like std::round. */
static void reached(int profile, double half_width, double c, int64_t *first, int64_t *last)
{
	switch (profile) {
	case PROFILE_BOX:
		*first = *last = (int64_t)round(c);
		return;
	case PROFILE_TRIANGLE:
		*first = (int64_t)floor(c);
		*last = *first + 1;
		return;
	default:
		*first = (int64_t)floor(c - half_width) + 1;
		*last = (int64_t)ceil(c + half_width) - 1;
		return;
	}
}

/* The CTF of a frequency from the contract formula (insert_ctf_slices_operation.hpp:43-58;
 * fixture reference_ctf at slice_insertion_cases.hpp:363-385). */
double oracle_ctf(const double *p, int64_t extent, double pixel_size,
	double voltage_kv, double cs_mm, double amplitude_contrast, double k_x, double k_y)
{
	const double wavelength = oracle_wavelength(voltage_kv);
	const double aberration = cs_mm * 1e7;
	const double side = (double)extent * pixel_size;
	const double s2 = (k_x * k_x + k_y * k_y) / (side * side);
	const double direction = atan2(k_y, k_x);
	const double defocus = 0.5 * (p[0] + p[1]) + 0.5 * (p[0] - p[1]) * cos(2.0 * (direction - p[2]));
	return sin(ORACLE_PI * wavelength * defocus * s2
		- 0.5 * ORACLE_PI * aberration * wavelength * wavelength * wavelength * s2 * s2
		+ p[3] + asin(amplitude_contrast));
}

/*
 * Insert K*B slices into K accumulators, in double.
 *   num: K x N x N x H complex (interleaved re, im); den: K x N x N x H.
 *   slices: K*B x n x h complex; transforms: K*B x 12 (row-major 3x4);
 *   radial: K*B x h; extra: K*B x 4 (ctf) or K*B x n x h (weighted) or NULL.
 * The accumulators are updated in place.
 */
void oracle_insert_f64(int variant, int profile, double half_width, double shape,
	int64_t n, int64_t big, int64_t K, int64_t B,
	double *num, double *den,
	const double *slices, const double *transforms, const double *radial, const double *extra,
	double voltage_kv, double cs_mm, double amplitude_contrast, double pixel_size)
{
	const int64_t h = n / 2 + 1;
	const int64_t half = big / 2 + 1;
	const double scale = (double)big / (double)n;
	const double centre = (double)(n / 2);
	const double volume_centre = (double)(big / 2);
	const int64_t volume_count = big * big * half;

	for (int64_t k = 0; k < K; ++k)
	for (int64_t b = 0; b < B; ++b) {
		const int64_t s = k * B + b;
		const double *slice = slices + s * n * h * 2;
		const double *t = transforms + s * 12;
		const double *r_row = radial + s * h;
		double *nk = num + k * volume_count * 2;
		double *dk = den + k * volume_count;

		for (int64_t row = 0; row < n; ++row)
		for (int64_t column = 0; column < h; ++column) {
			if (n % 2 == 0 && (row == n / 2 || column == n / 2))
				continue;
			const double k_y = (double)(row <= n / 2 ? row : row - n);
			const double k_x = (double)column;
			const int64_t shell = (int64_t)floor(sqrt(k_x * k_x + k_y * k_y) + 0.5);
			if (shell >= h)
				continue;
			const double r = r_row[shell];
			if (r == 0.0)
				continue;

			double alpha = 1.0, beta = 1.0;
			if (variant == VARIANT_WEIGHTED) {
				beta = extra[s * n * h + row * h + column];
			} else if (variant == VARIANT_CTF) {
				alpha = oracle_ctf(extra + s * 4, n, pixel_size, voltage_kv, cs_mm,
					amplitude_contrast, k_x, k_y);
				beta = alpha * alpha;
			}

			const double xr = slice[(row * h + column) * 2];
			const double xi = slice[(row * h + column) * 2 + 1];
			const double shift = 2.0 * ORACLE_PI *
				(k_x * (centre + t[3]) + k_y * (centre + t[7])) / (double)n;
			const double vr = xr * cos(shift) - xi * sin(shift);
			const double vi = xr * sin(shift) + xi * cos(shift);

			for (int sign = 1; sign >= -1; sign -= 2) {
				if (sign < 0 && column == 0)
					continue;
				const double value_r = vr;
				const double value_i = sign > 0 ? vi : -vi;
				double big_k[3];
				for (int j = 0; j < 3; ++j)
					big_k[j] = sign * scale * (k_x * t[j] + k_y * t[4 + j]);

				int64_t z0, z1, y0, y1, x0, x1;
				reached(profile, half_width, big_k[2], &z0, &z1);
				reached(profile, half_width, big_k[1], &y0, &y1);
				reached(profile, half_width, big_k[0], &x0, &x1);
				for (int64_t gz = z0; gz <= z1; ++gz)
				for (int64_t gy = y0; gy <= y1; ++gy)
				for (int64_t gx = x0; gx <= x1; ++gx) {
					if (gx < 0 || gx > big / 2 || llabs(gy) > big / 2 || llabs(gz) > big / 2)
						continue;
					const double kappa =
						oracle_window_weight(profile, half_width, shape, gx - big_k[0]) *
						oracle_window_weight(profile, half_width, shape, gy - big_k[1]) *
						oracle_window_weight(profile, half_width, shape, gz - big_k[2]);
					const double angle = -2.0 * ORACLE_PI * (double)(gx + gy + gz) *
						volume_centre / (double)big;
					const double pr = cos(angle), pi_ = sin(angle);
					const int64_t sy = gy < 0 ? gy + big : gy;
					const int64_t sz = gz < 0 ? gz + big : gz;
					const int64_t index = (sz * big + sy) * half + gx;
					const double w = kappa * r * alpha;
					nk[index * 2] += w * (value_r * pr - value_i * pi_);
					nk[index * 2 + 1] += w * (value_r * pi_ + value_i * pr);
					dk[index] += kappa * r * beta;
				}
			}
		}
	}
}
