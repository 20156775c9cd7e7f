// Test driver of the plugin boundary: a C++ caller of the rexlib API that
// loads librexlib-hip.so the way rexlib's loader does and calls
// rexlib::em::insert_*_slices / rfft2 / irfft3 unchanged. The Python tests
// run it as a child process and check what it prints and writes.
//
//   plugin_driver info <plugin>            registration, compatibility and suitability (no GPU needed)
//   plugin_driver run <plugin> <dir>       a reconstruction on hip:0 from the raw files in <dir>
//   plugin_driver errors <plugin>          error behaviour on hip:0
//   plugin_driver bench <plugin> <images> <launch> <steps>
//                                          throughput of rfft2 + insert_ctf_slices per launch of
//                                          256^2 images into 256^3, then irfft3, through the plugin
//
// Every line printed is a JSON object.
#include "rexlib/em.hpp"
#include "rexlib/fourier.hpp"

#include <dlfcn.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <random>
#include <fstream>
#include <iostream>
#include <map>
#include <sstream>

using namespace rexlib;

namespace
{

// A queue of no backend: every HIP builder must decline it.
class foreign_queue final : public command_queue
{
public:
	command_token submit(command) override { throw std::logic_error("foreign_queue runs nothing"); }
};

// Device memory of no backend.
class foreign_resource final : public memory_resource
{
public:
	memory_resource_kind get_kind() const noexcept override { return memory_resource_kind::device_local; }
	std::shared_ptr<memory_allocator> create_allocator() const override { return nullptr; }
};

std::string quote(const std::string &s)
{
	std::string out = "\"";
	for (char c : s)
	{
		if (c == '"' || c == '\\')
		{
			out += '\\';
		}
		out += (c == '\n' ? ' ' : c);
	}
	return out + "\"";
}

template <typename F>
std::string outcome(F &&f)
{
	try
	{
		f();
		return "ok";
	}
	catch (const unsupported_operation_error &)
	{
		return "unsupported_operation_error";
	}
	catch (const plugin_load_error &)
	{
		return "plugin_load_error";
	}
	catch (const std::invalid_argument &)
	{
		return "invalid_argument";
	}
	catch (const std::exception &e)
	{
		return std::string("exception: ") + e.what();
	}
}

int info(const std::string &path)
{
	// The raw entry point, as the loader reads it before anything else.
	void *handle = dlopen(path.c_str(), RTLD_NOW | RTLD_LOCAL);
	if (!handle)
	{
		std::cout << "{\"error\": " << quote(dlerror()) << "}" << std::endl;
		return 1;
	}
	using compatibility_entry = const plugin_compatibility *(*)();
	auto entry = reinterpret_cast<compatibility_entry>(dlsym(handle, "rexlib_get_plugin_compatibility"));
	const plugin_compatibility *r = entry ? entry() : nullptr;
	if (!r)
	{
		std::cout << "{\"error\": \"no compatibility record\"}" << std::endl;
		return 1;
	}
	std::cout << "{\"compatibility\": [" << r->size << ", " << r->rexlib_version_major << ", "
			  << r->rexlib_version_minor << ", " << r->rexlib_version_patch << ", " << r->abi_family << ", "
			  << r->standard_library << ", " << r->standard_library_abi << ", " << r->runtime << "]}" << std::endl;

	// Compatibility rules, on records derived from the plugin's.
	plugin_compatibility newer = *r;
	newer.rexlib_version_minor += 1;
	plugin_compatibility short_record = *r;
	short_record.size = 16;
	plugin_compatibility other_abi = *r;
	other_abi.standard_library_abi ^= 1;
	plugin_compatibility patch = *r;
	patch.rexlib_version_patch += 3;
	const plugin_compatibility &library = get_library_compatibility();
	std::cout << "{\"check_same\": " << quote(outcome([&] { check_plugin_compatibility(*r, library); }))
			  << ", \"check_patch\": " << quote(outcome([&] { check_plugin_compatibility(patch, library); }))
			  << ", \"check_minor\": " << quote(outcome([&] { check_plugin_compatibility(newer, library); }))
			  << ", \"check_size\": " << quote(outcome([&] { check_plugin_compatibility(short_record, library); }))
			  << ", \"check_abi\": " << quote(outcome([&] { check_plugin_compatibility(other_abi, library); }))
			  << "}" << std::endl;

	service_catalog catalog;
	const plugin &p = load_plugin(path, catalog);
	std::vector<std::string> backends;
	catalog.get_service<device_service>()->enumerate_backends(backends);
	std::vector<device_index> devices;
	catalog.get_service<device_service>()->enumerate_devices(devices);
	std::cout << "{\"name\": " << quote(p.get_name()) << ", \"version\": " << quote(p.get_version().to_string())
			  << ", \"backends\": [";
	for (std::size_t i = 0; i < backends.size(); ++i)
	{
		std::cout << (i ? ", " : "") << quote(backends[i]);
	}
	std::cout << "], \"devices\": " << devices.size() << ", \"operations\": [";
	const auto operations = catalog.get_service<program_service>()->enumerate_operations();
	for (std::size_t i = 0; i < operations.size(); ++i)
	{
		std::cout << (i ? ", " : "") << quote(operations[i]);
	}
	std::cout << "]}" << std::endl;

	// No builder takes a queue of another backend: the service reports
	// unsupported_operation_error rather than a builder throwing.
	foreign_queue queue;
	foreign_resource resource;
	const auto layout3 = strided_layout::make_contiguous({16, 16, 9});
	const std::vector<operand_descriptor> outs = {operand_descriptor(layout3, numerical_type::complex_float32, &resource),
		operand_descriptor(layout3, numerical_type::float32, &resource)};
	const std::vector<operand_descriptor> ins = {outs[0], outs[1],
		operand_descriptor(strided_layout::make_contiguous({4, 16, 9}), numerical_type::complex_float32, &resource),
		operand_descriptor(strided_layout::make_contiguous({4, 3, 4}), numerical_type::float32, &resource),
		operand_descriptor(strided_layout::make_contiguous({4, 4}), numerical_type::float32, &resource),
		operand_descriptor(strided_layout::make_contiguous({9}), numerical_type::float32, &resource)};
	const em::insert_ctf_slices_operation op(em::gridding_window::linear(), em::ctf_optics(300, 2.7, 0.1), 1.5);
	std::cout << "{\"foreign_queue\": " << quote(outcome([&] {
		catalog.get_service<program_service>()->build(op, {outs.data(), outs.size()}, {ins.data(), ins.size()}, queue);
	})) << ", \"bad_pixel_size\": " << quote(outcome([&] {
		em::insert_ctf_slices_operation(em::gridding_window::linear(), em::ctf_optics(300, 2.7, 0.1), -1.0);
	})) << ", \"bad_optics\": " << quote(outcome([&] { em::ctf_optics(300, 2.7, 1.0); }))
			  << ", \"bad_window\": " << quote(outcome([&] { em::gridding_window::kaiser_bessel(0.0, 2.0); }))
			  << ", \"wavelength_300kV\": " << em::ctf_optics(300, 2.7, 0.1).get_wavelength() << "}" << std::endl;
	return 0;
}

// Raw little-endian files of the run: name -> bytes.
std::vector<char> read_file(const std::string &path)
{
	std::ifstream f(path, std::ios::binary);
	if (!f)
	{
		throw std::runtime_error("cannot read " + path);
	}
	return std::vector<char>(std::istreambuf_iterator<char>(f), std::istreambuf_iterator<char>());
}

void write_file(const std::string &path, const array &host)
{
	std::ofstream f(path, std::ios::binary);
	const std::size_t bytes = host.get_layout().compute_element_count() * get_size(host.get_data_type());
	f.write(static_cast<const char *>(host.get_buffer()->get_host_ptr()), static_cast<std::streamsize>(bytes));
}

array host_from_file(const std::string &path, std::vector<std::size_t> extents, numerical_type type)
{
	array a = make_host_array(extents, type);
	const std::vector<char> bytes = read_file(path);
	const std::size_t need = a.get_layout().compute_element_count() * get_size(type);
	if (bytes.size() != need)
	{
		throw std::runtime_error(path + ": " + std::to_string(bytes.size()) + " bytes, expected " +
			std::to_string(need));
	}
	std::memcpy(a.get_buffer()->get_host_ptr(), bytes.data(), need);
	return a;
}

// meta.txt: key value lines (count n big window half_width shape voltage cs q pixel split).
std::map<std::string, double> read_meta(const std::string &dir)
{
	std::map<std::string, double> meta;
	std::ifstream f(dir + "/meta.txt");
	std::string key;
	double value;
	while (f >> key >> value)
	{
		meta[key] = value;
	}
	return meta;
}

int run(const std::string &path, const std::string &dir)
{
	service_catalog catalog;
	load_plugin(path, catalog);
	const execution_context ctx = make_execution_context(catalog, device_index("hip", 0));
	auto meta = read_meta(dir);
	const std::size_t count = static_cast<std::size_t>(meta.at("count"));
	const std::size_t n = static_cast<std::size_t>(meta.at("n"));
	const std::size_t big = static_cast<std::size_t>(meta.at("big"));
	const std::size_t split = static_cast<std::size_t>(meta.at("split"));
	const std::size_t h = n / 2 + 1;
	const int profile = static_cast<int>(meta.at("window"));
	const em::gridding_window window = profile == 0 ? em::gridding_window::nearest()
		: profile == 1                             ? em::gridding_window::linear()
												   : em::gridding_window::kaiser_bessel(meta.at("half_width"), meta.at("shape"));
	const em::ctf_optics optics(meta.at("voltage"), meta.at("cs"), meta.at("q"));
	const double pixel = meta.at("pixel");

	// Images -> spectra on the device (rfft2), then insertion in two calls
	// that accumulate in place, as production callers do.
	const array images = to_device(host_from_file(dir + "/images.bin", {count, n, n}, numerical_type::float32), ctx);
	const array transforms =
		to_device(host_from_file(dir + "/transforms.bin", {count, 3, 4}, numerical_type::float32), ctx);
	const array ctf = to_device(host_from_file(dir + "/ctf.bin", {count, 4}, numerical_type::float32), ctx);
	const array radial = to_device(host_from_file(dir + "/radial.bin", {h}, numerical_type::float32), ctx);
	const array spectra = rfft2(images, ctx);
	array num = zeros({big, big, big / 2 + 1}, numerical_type::complex_float32, ctx);
	array den = zeros({big, big, big / 2 + 1}, numerical_type::float32, ctx);

	// The two halves of the batch are views of the same device arrays.
	auto part = [&](const array &a, std::size_t begin, std::size_t end) {
		const strided_layout &l = a.get_layout();
		std::vector<std::size_t> e;
		std::vector<std::ptrdiff_t> s;
		l.get_extents(e);
		l.get_strides(s);
		e[0] = end - begin;
		const std::ptrdiff_t offset = l.get_offset() + static_cast<std::ptrdiff_t>(begin) * s[0];
		return a.view(array_descriptor(strided_layout(e, s, offset), a.get_data_type()));
	};
	const std::size_t halves[3] = {0, split, count};
	for (int k = 0; k < 2; ++k)
	{
		if (halves[k] == halves[k + 1])
		{
			continue;
		}
		em::insert_ctf_slices(num, den, part(spectra, halves[k], halves[k + 1]),
			part(transforms, halves[k], halves[k + 1]), part(ctf, halves[k], halves[k + 1]), radial, window, optics,
			pixel, ctx, &num, &den);
	}
	const array volume = irfft3(num, big, ctx);
	// The quotient of the reference's recipe
	// (tests/integration/src/em/functional/test_cpu_reconstruction.cpp:457-464)
	// on the device: where(den > 1e-3, num / den, 0). divide takes one type,
	// so the weights enter as complex numbers (made on the host here).
	const array den_host = to_host(den, ctx);
	array den_c_host = make_host_array({big, big, big / 2 + 1}, numerical_type::complex_float32);
	{
		const float *d = static_cast<const float *>(den_host.get_buffer()->get_host_ptr());
		float *c = static_cast<float *>(den_c_host.get_buffer()->get_host_ptr());
		for (std::size_t i = 0; i < den_host.get_layout().compute_element_count(); ++i)
		{
			c[2 * i] = d[i] > 0.0f ? d[i] : 1.0f;
			c[2 * i + 1] = 0.0f;
		}
	}
	const array den_c = to_device(den_c_host, ctx);
	const array threshold = full({1}, 1e-3, numerical_type::float32, ctx);
	const array zero = full({1}, 0.0, numerical_type::complex_float32, ctx);
	const array quotient = where(greater(den, threshold, ctx), divide(num, den_c, ctx), zero, ctx);
	write_file(dir + "/quotient_out.bin", to_host(quotient, ctx));

	// Background standardization of the images, a key per image.
	array keys_host = make_host_array({count}, numerical_type::uint64);
	{
		std::uint64_t *k = static_cast<std::uint64_t *>(keys_host.get_buffer()->get_host_ptr());
		for (std::size_t i = 0; i < count; ++i)
		{
			k[i] = 1000 + i;
		}
	}
	const array standardized = em::standardize_background(images, to_device(keys_host, ctx), {1, 2},
		static_cast<double>(n) / 2 - 2, -3.0, 3.0, ctx);
	write_file(dir + "/standardized_out.bin", to_host(standardized, ctx));
	write_file(dir + "/spectra_out.bin", to_host(spectra, ctx));
	write_file(dir + "/num_out.bin", to_host(num, ctx));
	write_file(dir + "/den_out.bin", to_host(den, ctx));
	write_file(dir + "/volume_out.bin", to_host(volume, ctx));
	std::cout << "{\"run\": \"ok\", \"device\": "
			  << quote(catalog.get_service<device_service>()->create_device_session(device_index("hip", 0))
						   ->get_properties()
						   .get_name())
			  << "}" << std::endl;
	return 0;
}

int errors(const std::string &path)
{
	service_catalog catalog;
	load_plugin(path, catalog);
	const execution_context ctx = make_execution_context(catalog, device_index("hip", 0));
	const auto window = em::gridding_window::linear();
	array num = zeros({16, 16, 9}, numerical_type::complex_float32, ctx);
	array den = zeros({16, 16, 9}, numerical_type::float32, ctx);
	const array slices = zeros({4, 16, 9}, numerical_type::complex_float32, ctx);
	const array transforms = zeros({4, 3, 4}, numerical_type::float32, ctx);
	const array radial = zeros({9}, numerical_type::float32, ctx);
	const array bad_radial = zeros({7}, numerical_type::float32, ctx);
	const array bad_transforms = zeros({4, 2, 4}, numerical_type::float32, ctx);
	const array num64 = zeros({16, 16, 9}, numerical_type::complex_float64, ctx);
	const array den64 = zeros({16, 16, 9}, numerical_type::float64, ctx);
	const array slices64 = zeros({4, 16, 9}, numerical_type::complex_float64, ctx);
	const array transforms64 = zeros({4, 3, 4}, numerical_type::float64, ctx);
	const array radial64 = zeros({9}, numerical_type::float64, ctx);
	const array host_slices = make_host_array({4, 16, 9}, numerical_type::complex_float32);
	std::cout << "{\"valid\": " << quote(outcome([&] {
		em::insert_slices(num, den, slices, transforms, radial, window, ctx, &num, &den);
		num.wait();
	})) << ", \"radial_extent\": " << quote(outcome([&] {
		em::insert_slices(num, den, slices, transforms, bad_radial, window, ctx, &num, &den);
	})) << ", \"transform_shape\": " << quote(outcome([&] {
		em::insert_slices(num, den, slices, bad_transforms, radial, window, ctx, &num, &den);
	})) << ", \"float64\": " << quote(outcome([&] {
		em::insert_slices(num64, den64, slices64, transforms64, radial64, window, ctx);
	})) << ", \"host_operand\": " << quote(outcome([&] {
		em::insert_slices(num, den, host_slices, transforms, radial, window, ctx, &num, &den);
	})) << ", \"new_outputs\": " << quote(outcome([&] {
		auto out = em::insert_slices(num, den, slices, transforms, radial, window, ctx);
		if (out.first.get_buffer() == num.get_buffer() || out.second.is_empty())
		{
			throw std::logic_error("outputs were not allocated");
		}
		out.first.wait();
	})) << "}" << std::endl;
	return 0;
}

// The benchmark's call sequence through the plugin: `images` 256^2 images in
// launches of `launch` (the first `launch` random images reused by every
// launch, with their own random rotations and CTF), each launch an rfft2 into
// a spectra array and an insert_ctf_slices accumulating in place, then one
// irfft3 of the numerator; one untimed step, then `steps` timed ones. The
// accumulators are not cleared between steps (the work is the same).
int bench(const std::string &path, std::size_t images, std::size_t launch, int steps)
{
	service_catalog catalog;
	load_plugin(path, catalog);
	const execution_context ctx = make_execution_context(catalog, device_index("hip", 0));
	const std::size_t n = 256, big = 256, h = n / 2 + 1;
	launch = std::min(launch, images);
	std::mt19937_64 rng(2024);
	std::normal_distribution<float> normal(0.0f, 1.0f);
	std::uniform_real_distribution<double> uniform(0.0, 1.0);
	array images_host = make_host_array({launch, n, n}, numerical_type::float32);
	{
		float *p = static_cast<float *>(images_host.get_buffer()->get_host_ptr());
		for (std::size_t i = 0; i < launch * n * n; ++i)
		{
			p[i] = normal(rng);
		}
	}
	array transforms_host = make_host_array({images, 3, 4}, numerical_type::float32);
	array ctf_host = make_host_array({images, 4}, numerical_type::float32);
	{
		float *t = static_cast<float *>(transforms_host.get_buffer()->get_host_ptr());
		float *c = static_cast<float *>(ctf_host.get_buffer()->get_host_ptr());
		for (std::size_t i = 0; i < images; ++i)
		{
			// A uniform rotation from a random unit quaternion.
			double q[4];
			double norm = 0.0;
			for (double &x : q)
			{
				x = std::sqrt(-2.0 * std::log(1.0 - uniform(rng))) * std::cos(2.0 * M_PI * uniform(rng));
				norm += x * x;
			}
			norm = std::sqrt(norm);
			const double a = q[0] / norm, b = q[1] / norm, cq = q[2] / norm, d = q[3] / norm;
			const double r[9] = {a * a + b * b - cq * cq - d * d, 2 * (b * cq - a * d), 2 * (b * d + a * cq),
				2 * (b * cq + a * d), a * a - b * b + cq * cq - d * d, 2 * (cq * d - a * b),
				2 * (b * d - a * cq), 2 * (cq * d + a * b), a * a - b * b - cq * cq + d * d};
			for (int row = 0; row < 3; ++row)
			{
				for (int col = 0; col < 3; ++col)
				{
					t[i * 12 + row * 4 + col] = static_cast<float>(r[row * 3 + col]);
				}
				t[i * 12 + row * 4 + 3] = 0.0f;
			}
			c[i * 4 + 0] = static_cast<float>(10000.0 + 20000.0 * uniform(rng));
			c[i * 4 + 1] = static_cast<float>(10000.0 + 20000.0 * uniform(rng));
			c[i * 4 + 2] = static_cast<float>(M_PI * uniform(rng));
			c[i * 4 + 3] = 0.0f;
		}
	}
	const array image_stack = to_device(images_host, ctx);
	const array transforms = to_device(transforms_host, ctx);
	const array ctf = to_device(ctf_host, ctx);
	const array radial = full({h}, 1.0, numerical_type::float32, ctx);
	array num = zeros({big, big, big / 2 + 1}, numerical_type::complex_float32, ctx);
	array den = zeros({big, big, big / 2 + 1}, numerical_type::float32, ctx);
	array spectra = zeros({launch, n, h}, numerical_type::complex_float32, ctx);
	const em::gridding_window window = em::gridding_window::linear();
	const em::ctf_optics optics(300.0, 2.7, 0.1);
	auto part = [&](const array &a, std::size_t begin, std::size_t end) {
		const strided_layout &l = a.get_layout();
		std::vector<std::size_t> e;
		std::vector<std::ptrdiff_t> st;
		l.get_extents(e);
		l.get_strides(st);
		e[0] = end - begin;
		const std::ptrdiff_t offset = l.get_offset() + static_cast<std::ptrdiff_t>(begin) * st[0];
		return a.view(array_descriptor(strided_layout(e, st, offset), a.get_data_type()));
	};
	array volume;
	auto step = [&]() {
		for (std::size_t s = 0; s < images; s += launch)
		{
			const std::size_t e = std::min(s + launch, images);
			array out = part(spectra, 0, e - s);
			rfft2(part(image_stack, 0, e - s), ctx, &out);
			em::insert_ctf_slices(num, den, out, part(transforms, s, e), part(ctf, s, e), radial, window, optics,
				1.5, ctx, &num, &den);
		}
		volume = irfft3(num, big, ctx);
		volume.wait();
		den.wait();
	};
	step();
	const auto t0 = std::chrono::steady_clock::now();
	for (int k = 0; k < steps; ++k)
	{
		step();
	}
	const double seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
	std::cout << "{\"images\": " << images << ", \"launch\": " << launch << ", \"steps\": " << steps
			  << ", \"ms_per_step\": " << 1e3 * seconds / steps << ", \"images_per_s\": " << images * steps / seconds
			  << "}" << std::endl;
	return 0;
}

} // namespace

int main(int argc, char **argv)
{
	if (argc < 3)
	{
		std::cerr << "usage: plugin_driver info|run|errors|bench <plugin> [dir | images launch steps]" << std::endl;
		return 2;
	}
	const std::string mode = argv[1];
	try
	{
		if (mode == "info")
		{
			return info(argv[2]);
		}
		if (mode == "run" && argc >= 4)
		{
			return run(argv[2], argv[3]);
		}
		if (mode == "errors")
		{
			return errors(argv[2]);
		}
		if (mode == "bench" && argc >= 6)
		{
			return bench(argv[2], std::stoul(argv[3]), std::stoul(argv[4]), std::stoi(argv[5]));
		}
	}
	catch (const std::exception &e)
	{
		std::cout << "{\"error\": " << quote(e.what()) << "}" << std::endl;
		return 1;
	}
	std::cerr << "unknown mode " << mode << std::endl;
	return 2;
}
