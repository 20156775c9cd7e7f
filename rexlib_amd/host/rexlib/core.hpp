// The part of rexlib's core API that a device plugin and its callers use,
// restated with the same names, signatures and meaning as rexlib 0.1.0, so
// that a plugin and the em reconstruction callers written against it build
// here, where rexlib itself cannot be built (its submodules are absent).
//
// Interfaces mirrored (reference file:line):
//   plugin                     include/rexlib/core/plugin.hpp:23-54
//   plugin_compatibility       include/rexlib/core/plugin_compatibility.hpp:23-62
//   version                    include/rexlib/core/version.hpp
//   backend_priority           include/rexlib/core/backend_priority.hpp
//   numerical_type             include/rexlib/core/numerical/numerical_type.hpp:18-46
//   memory_resource_kind       include/rexlib/core/hardware/memory_resource_kind.hpp
//   memory_resource_affinity   include/rexlib/core/hardware/memory_resource_affinity.hpp
//   strided_layout             include/rexlib/core/layout/strided_layout.hpp:48-205 (extents,
//                              strides and offset in elements)
//   operand_descriptor         include/rexlib/core/dispatch/operand_descriptor.hpp
//   buffer, memory_resource, memory_allocator, program, command,
//   command_token, command_timeline, command_queue, device, device_backend
//                              include/rexlib/core/hardware/*.hpp
//   device_service             include/rexlib/core/hardware/device_service.hpp
//   program_builder            include/rexlib/core/dispatch/program_builder.hpp:47-99
//   program_service            include/rexlib/core/dispatch/program_service.hpp:51-80
//   service_catalog            include/rexlib/core/service_catalog.hpp
//   operation, operation_id    include/rexlib/core/operation/operation.hpp, operation_id.hpp:15-50
//
// Left out, because no caller of the insertion path needs them: operation
// declaration macros and shape/type rules (each operation here is a plain
// class), program caches, scratch requirements beyond the accessor, the
// MSVC branches of the ABI record.
#pragma once

#include <complex>
#include <cstddef>
#include <cstdint>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <typeindex>
#include <typeinfo>
#include <vector>

#define REXLIB_API __attribute__((visibility("default")))
#define REXLIB_EXPORT __attribute__((visibility("default")))

// C++ ABI of the code being compiled (include/rexlib/core/platform/cpp_abi.hpp).
#define REXLIB_CPP_ABI_ITANIUM 1
#define REXLIB_CPP_LIBSTDCXX 1
#define REXLIB_CPP_LIBCXX 2
#define REXLIB_CPP_NO_MSVC_RUNTIME 0
#define REXLIB_CPP_ABI_FAMILY REXLIB_CPP_ABI_ITANIUM
#if defined(_LIBCPP_VERSION)
#define REXLIB_CPP_STANDARD_LIBRARY REXLIB_CPP_LIBCXX
#define REXLIB_CPP_STANDARD_LIBRARY_ABI _LIBCPP_ABI_VERSION
#else
#define REXLIB_CPP_STANDARD_LIBRARY REXLIB_CPP_LIBSTDCXX
#define REXLIB_CPP_STANDARD_LIBRARY_ABI _GLIBCXX_USE_CXX11_ABI
#endif
#define REXLIB_CPP_RUNTIME REXLIB_CPP_NO_MSVC_RUNTIME

// The rexlib release these headers state.
#define REXLIB_VERSION_MAJOR 0
#define REXLIB_VERSION_MINOR 1
#define REXLIB_VERSION_PATCH 0

namespace rexlib
{

// ---------------------------------------------------------------------------
// Basic values
// ---------------------------------------------------------------------------

// A contiguous view: pointer and count.
template <typename T>
class span
{
public:
	span() noexcept = default;
	span(T *data, std::size_t size) noexcept : m_data(data), m_size(size) {}
	template <typename U>
	span(const std::vector<U> &v) noexcept : m_data(v.data()), m_size(v.size()) {}
	template <typename U>
	span(std::vector<U> &v) noexcept : m_data(v.data()), m_size(v.size()) {}
	T *data() const noexcept { return m_data; }
	std::size_t size() const noexcept { return m_size; }
	bool empty() const noexcept { return m_size == 0; }
	T &operator[](std::size_t i) const noexcept { return m_data[i]; }
	T *begin() const noexcept { return m_data; }
	T *end() const noexcept { return m_data + m_size; }

private:
	T *m_data = nullptr;
	std::size_t m_size = 0;
};

class version
{
public:
	version() = default;
	constexpr version(std::uint32_t major, std::uint32_t minor, std::uint32_t patch) noexcept
		: m_major(major), m_minor(minor), m_patch(patch)
	{
	}
	constexpr std::uint32_t get_major() const noexcept { return m_major; }
	constexpr std::uint32_t get_minor() const noexcept { return m_minor; }
	constexpr std::uint32_t get_patch() const noexcept { return m_patch; }
	friend bool operator==(const version &a, const version &b) noexcept
	{
		return a.m_major == b.m_major && a.m_minor == b.m_minor && a.m_patch == b.m_patch;
	}
	friend bool operator!=(const version &a, const version &b) noexcept { return !(a == b); }
	std::string to_string() const
	{
		return std::to_string(m_major) + "." + std::to_string(m_minor) + "." + std::to_string(m_patch);
	}

private:
	std::uint32_t m_major = 0;
	std::uint32_t m_minor = 0;
	std::uint32_t m_patch = 0;
};

enum class backend_priority
{
	unsupported = -2147483647 - 1,
	fallback = -1024,
	normal = 0,
	optimal = 1024,
};

REXLIB_API const char *to_string(backend_priority priority) noexcept;

enum class numerical_type
{
	unknown = -1,
	boolean,
	char8,
	int8,
	uint8,
	int16,
	uint16,
	int32,
	uint32,
	int64,
	uint64,
	float16,
	float32,
	float64,
	complex_float16,
	complex_float32,
	complex_float64,
	count
};

REXLIB_API std::size_t get_size(numerical_type type) noexcept;
REXLIB_API const char *to_string(numerical_type type) noexcept;

// A scalar of any numerical type, held by value
// (core/numerical/scalar_value.hpp:34-77): constructed from a value of a
// supported type, read back with get_if<T>() for the type it holds.
template <typename T>
struct scalar_type_of;
template <> struct scalar_type_of<bool> { static constexpr numerical_type value = numerical_type::boolean; };
template <> struct scalar_type_of<std::int8_t> { static constexpr numerical_type value = numerical_type::int8; };
template <> struct scalar_type_of<std::uint8_t> { static constexpr numerical_type value = numerical_type::uint8; };
template <> struct scalar_type_of<std::int16_t> { static constexpr numerical_type value = numerical_type::int16; };
template <> struct scalar_type_of<std::uint16_t> { static constexpr numerical_type value = numerical_type::uint16; };
template <> struct scalar_type_of<std::int32_t> { static constexpr numerical_type value = numerical_type::int32; };
template <> struct scalar_type_of<std::uint32_t> { static constexpr numerical_type value = numerical_type::uint32; };
template <> struct scalar_type_of<std::int64_t> { static constexpr numerical_type value = numerical_type::int64; };
template <> struct scalar_type_of<std::uint64_t> { static constexpr numerical_type value = numerical_type::uint64; };
template <> struct scalar_type_of<float> { static constexpr numerical_type value = numerical_type::float32; };
template <> struct scalar_type_of<double> { static constexpr numerical_type value = numerical_type::float64; };
template <> struct scalar_type_of<std::complex<float>> {
	static constexpr numerical_type value = numerical_type::complex_float32; };
template <> struct scalar_type_of<std::complex<double>> {
	static constexpr numerical_type value = numerical_type::complex_float64; };

class scalar_value
{
public:
	template <typename T>
	scalar_value(const T &value) noexcept : m_data_type(scalar_type_of<T>::value)
	{
		static_assert(sizeof(T) <= sizeof(m_storage), "scalar too large");
		std::memcpy(m_storage, &value, sizeof(T));
	}
	template <typename T>
	const T *get_if() const noexcept
	{
		return m_data_type == scalar_type_of<T>::value ? reinterpret_cast<const T *>(m_storage) : nullptr;
	}
	numerical_type get_data_type() const noexcept { return m_data_type; }

private:
	alignas(std::complex<double>) unsigned char m_storage[sizeof(std::complex<double>)];
	numerical_type m_data_type;
};

enum class memory_resource_kind
{
	unified,
	managed,
	device_local,
	device_mapped,
	host_staging,
	host,
};

REXLIB_API int score_host_affinity(memory_resource_kind kind) noexcept;
REXLIB_API int score_device_affinity(memory_resource_kind kind) noexcept;
REXLIB_API bool is_host_accessible(memory_resource_kind kind) noexcept;
REXLIB_API bool is_device_accessible(memory_resource_kind kind) noexcept;

enum class memory_resource_affinity
{
	host,
	device,
	count
};

// ---------------------------------------------------------------------------
// Exceptions (include/rexlib/core/exceptions/)
// ---------------------------------------------------------------------------

class REXLIB_API unsupported_operation_error : public std::runtime_error
{
public:
	using std::runtime_error::runtime_error;
};

class REXLIB_API plugin_load_error : public std::runtime_error
{
public:
	using std::runtime_error::runtime_error;
};

// ---------------------------------------------------------------------------
// Layouts and operands
// ---------------------------------------------------------------------------

// Extents, strides and offset in elements. A stride of 0 broadcasts an axis.
class REXLIB_API strided_layout
{
public:
	strided_layout() = default;
	strided_layout(std::vector<std::size_t> extents, std::vector<std::ptrdiff_t> strides, std::ptrdiff_t offset = 0);
	// Row-major contiguous layout of the given extents.
	static strided_layout make_contiguous(std::vector<std::size_t> extents);

	// Accessors as strided_layout.hpp:48-88 declares them (extents and
	// strides through out-parameters).
	std::size_t get_rank() const noexcept { return m_extents.size(); }
	void get_extents(std::vector<std::size_t> &out) const { out = m_extents; }
	void get_strides(std::vector<std::ptrdiff_t> &out) const { out = m_strides; }
	std::ptrdiff_t get_offset() const noexcept { return m_offset; }
	std::size_t compute_element_count() const noexcept;

	friend bool operator==(const strided_layout &a, const strided_layout &b) noexcept
	{
		return a.m_extents == b.m_extents && a.m_strides == b.m_strides && a.m_offset == b.m_offset;
	}

private:
	std::vector<std::size_t> m_extents;
	std::vector<std::ptrdiff_t> m_strides;
	std::ptrdiff_t m_offset = 0;
};

class array_descriptor
{
public:
	array_descriptor() = default;
	array_descriptor(strided_layout layout, numerical_type type) : m_layout(std::move(layout)), m_type(type) {}
	const strided_layout &get_layout() const noexcept { return m_layout; }
	numerical_type get_data_type() const noexcept { return m_type; }
	friend bool operator==(const array_descriptor &a, const array_descriptor &b) noexcept
	{
		return a.m_layout == b.m_layout && a.m_type == b.m_type;
	}

private:
	strided_layout m_layout;
	numerical_type m_type = numerical_type::unknown;
};

class memory_resource;

class REXLIB_API operand_descriptor
{
public:
	operand_descriptor() noexcept = default;
	explicit operand_descriptor(const array_descriptor &descriptor, const memory_resource *resource = nullptr) noexcept
		: m_descriptor(descriptor), m_memory_region(resource)
	{
	}
	operand_descriptor(const strided_layout &layout, numerical_type data_type,
		const memory_resource *resource = nullptr) noexcept
		: m_descriptor(layout, data_type), m_memory_region(resource)
	{
	}
	const strided_layout &get_layout() const noexcept { return m_descriptor.get_layout(); }
	numerical_type get_data_type() const noexcept { return m_descriptor.get_data_type(); }
	const array_descriptor &get_array_descriptor() const noexcept { return m_descriptor; }
	const memory_resource *get_memory_resource() const noexcept { return m_memory_region; }
	void set_memory_resource(const memory_resource *resource) noexcept { m_memory_region = resource; }

private:
	array_descriptor m_descriptor;
	const memory_resource *m_memory_region = nullptr;
};

// ---------------------------------------------------------------------------
// Hardware abstraction
// ---------------------------------------------------------------------------

class memory_allocator;
class command_queue;
class command_token;

class REXLIB_API buffer
{
public:
	buffer() noexcept;
	buffer(const buffer &) = delete;
	buffer &operator=(const buffer &) = delete;
	virtual ~buffer();
	// Null when the memory cannot be read by the host.
	virtual void *get_host_ptr() noexcept = 0;
	virtual const void *get_host_ptr() const noexcept = 0;
	virtual std::size_t get_size() const noexcept = 0;
	virtual const memory_resource &get_memory_resource() const noexcept = 0;
};

class REXLIB_API memory_resource
{
public:
	memory_resource() noexcept;
	memory_resource(const memory_resource &) = delete;
	memory_resource &operator=(const memory_resource &) = delete;
	virtual ~memory_resource();
	virtual memory_resource_kind get_kind() const noexcept = 0;
	virtual std::shared_ptr<memory_allocator> create_allocator() const = 0;
};

// Plain host memory (operator new), as the core's own resource.
REXLIB_API const memory_resource &get_host_memory_resource() noexcept;

class REXLIB_API command_timeline
{
public:
	command_timeline() noexcept;
	command_timeline(const command_timeline &) = delete;
	command_timeline &operator=(const command_timeline &) = delete;
	virtual ~command_timeline();
	virtual void wait(std::size_t id) const = 0;
	virtual bool is_complete(std::size_t id) const = 0;
};

// A point on a timeline; empty tokens are complete.
class REXLIB_API command_token
{
public:
	command_token() noexcept;
	command_token(std::shared_ptr<const command_timeline> timeline, std::size_t id) noexcept;
	const std::shared_ptr<const command_timeline> &get_timeline() const noexcept { return m_timeline; }
	std::size_t get_id() const noexcept { return m_id; }
	bool is_empty() const noexcept { return !m_timeline; }
	bool is_complete() const;
	void wait() const;

private:
	std::shared_ptr<const command_timeline> m_timeline;
	std::size_t m_id = 0;
};

class REXLIB_API memory_allocator
{
public:
	memory_allocator() noexcept;
	memory_allocator(const memory_allocator &) = delete;
	memory_allocator &operator=(const memory_allocator &) = delete;
	virtual ~memory_allocator();
	virtual const memory_resource &get_memory_resource() const noexcept = 0;
	virtual std::size_t get_max_alignment() const noexcept = 0;
	// `pending` receives a token the buffer's first use must wait for (empty
	// when it is ready at once).
	virtual std::shared_ptr<buffer> allocate(std::size_t size, std::size_t alignment, command_queue *queue_hint,
		command_token &pending) = 0;
};

struct program_scratch_requirement
{
	std::size_t size;
	std::size_t alignment;
	memory_resource_affinity affinity;
};

class REXLIB_API program
{
public:
	program() noexcept;
	program(const program &) = delete;
	program &operator=(const program &) = delete;
	virtual ~program();
	virtual span<const program_scratch_requirement> get_scratch_requirements() const;
};

class REXLIB_API command
{
public:
	command() noexcept;
	explicit command(std::shared_ptr<const program> program) noexcept;
	command &bind_outputs(std::vector<std::shared_ptr<buffer>> outputs) noexcept;
	command &bind_inputs(std::vector<std::shared_ptr<const buffer>> inputs) noexcept;
	command &bind_scratch(std::vector<std::shared_ptr<buffer>> scratch) noexcept;
	command &bind_dependencies(std::vector<command_token> dependencies) noexcept;
	const std::shared_ptr<const program> &get_program() const noexcept { return m_program; }
	span<const std::shared_ptr<buffer>> get_outputs() const noexcept { return {m_outputs.data(), m_outputs.size()}; }
	span<const std::shared_ptr<const buffer>> get_inputs() const noexcept { return {m_inputs.data(), m_inputs.size()}; }
	span<const std::shared_ptr<buffer>> get_scratch() const noexcept { return {m_scratch.data(), m_scratch.size()}; }
	span<const command_token> get_dependencies() const noexcept
	{
		return {m_dependencies.data(), m_dependencies.size()};
	}

private:
	std::shared_ptr<const program> m_program;
	std::vector<std::shared_ptr<buffer>> m_outputs;
	std::vector<std::shared_ptr<const buffer>> m_inputs;
	std::vector<std::shared_ptr<buffer>> m_scratch;
	std::vector<command_token> m_dependencies;
};

class REXLIB_API command_queue
{
public:
	command_queue() noexcept;
	command_queue(const command_queue &) = delete;
	command_queue &operator=(const command_queue &) = delete;
	virtual ~command_queue();
	virtual command_token submit(command cmd) = 0;
};

class REXLIB_API device
{
public:
	device() noexcept;
	device(const device &) = delete;
	device &operator=(const device &) = delete;
	virtual ~device();
	virtual const memory_resource &get_memory_resource(memory_resource_affinity affinity) const = 0;
	virtual std::shared_ptr<command_queue> create_command_queue() const = 0;
};

// What a backend reports of a device (hardware/device_properties.hpp:26-153):
// set by the backend, read by the caller; out-of-line accessors.
enum class device_type
{
	unknown = -1,
	cpu,
	gpu,
	integrated_gpu,
};

class REXLIB_API device_properties
{
public:
	device_properties();
	~device_properties();
	device_properties(const device_properties &other);
	device_properties &operator=(const device_properties &other);

	void set_type(device_type type);
	device_type get_type() const noexcept;
	void set_name(const std::string &name);
	const std::string &get_name() const noexcept;
	void set_physical_location(const std::string &location);
	const std::string &get_physical_location() const noexcept;
	void set_total_memory_bytes(std::size_t bytes);
	std::size_t get_total_memory_bytes() const noexcept;
	void set_optimal_data_alignment(std::size_t alignment);
	std::size_t get_optimal_data_alignment() const noexcept;

private:
	struct implementation;
	std::unique_ptr<implementation> m_implementation;
};

class REXLIB_API device_backend
{
public:
	device_backend() noexcept;
	device_backend(const device_backend &) = delete;
	device_backend &operator=(const device_backend &) = delete;
	virtual ~device_backend();
	virtual std::string get_name() const = 0;
	virtual version get_version() const = 0;
	virtual void enumerate_devices(std::vector<std::size_t> &ids) const = 0;
	virtual bool get_device_properties(std::size_t id, device_properties &desc) const = 0;
	virtual std::shared_ptr<device> create_device(std::size_t id) const = 0;
};

class device_index
{
public:
	device_index() = default;
	device_index(std::string backend_name, std::size_t device_id)
		: m_backend(std::move(backend_name)), m_id(device_id)
	{
	}
	const std::string &get_backend_name() const noexcept { return m_backend; }
	std::size_t get_device_id() const noexcept { return m_id; }

private:
	std::string m_backend;
	std::size_t m_id = 0;
};

// A device opened for work: the device, its default queue and one allocator
// per affinity.
class REXLIB_API device_session
{
public:
	device_session(device_index index, std::shared_ptr<device> dev, device_properties properties);
	const device_index &get_index() const noexcept { return m_index; }
	const std::shared_ptr<device> &get_device() const noexcept { return m_device; }
	const device_properties &get_properties() const noexcept { return m_properties; }
	const std::shared_ptr<command_queue> &get_default_queue() const noexcept { return m_queue; }
	const std::shared_ptr<memory_allocator> &get_allocator(memory_resource_affinity affinity) const noexcept
	{
		return m_allocators[static_cast<std::size_t>(affinity)];
	}

private:
	device_index m_index;
	std::shared_ptr<device> m_device;
	device_properties m_properties;
	std::shared_ptr<command_queue> m_queue;
	std::shared_ptr<memory_allocator> m_allocators[static_cast<std::size_t>(memory_resource_affinity::count)];
};

// ---------------------------------------------------------------------------
// Operations and programs
// ---------------------------------------------------------------------------

class REXLIB_API operation_id
{
public:
	template <typename T>
	static operation_id of() noexcept
	{
		return operation_id(std::type_index(typeid(T)));
	}
	const char *get_name() const noexcept { return m_index.name(); }
	friend bool operator==(const operation_id &a, const operation_id &b) noexcept { return a.m_index == b.m_index; }
	friend bool operator!=(const operation_id &a, const operation_id &b) noexcept { return a.m_index != b.m_index; }
	std::size_t hash() const noexcept { return m_index.hash_code(); }

private:
	explicit operation_id(std::type_index index) noexcept : m_index(index) {}
	std::type_index m_index;
};

class REXLIB_API operation
{
public:
	operation() noexcept;
	virtual ~operation();
	virtual operation_id get_id() const noexcept = 0;
	virtual const char *get_name() const noexcept = 0;
};

// An operation with an identity of its own type.
template <typename Derived>
class basic_operation : public operation
{
public:
	operation_id get_id() const noexcept override { return operation_id::of<Derived>(); }
};

template <typename T>
const T &operation_cast(const operation &op)
{
	if (op.get_id() != operation_id::of<T>())
	{
		throw std::invalid_argument(std::string("operation_cast: the operation is ") + op.get_name());
	}
	return static_cast<const T &>(op);
}

class program_cache;

class REXLIB_API program_builder
{
public:
	program_builder() noexcept;
	program_builder(const program_builder &) = delete;
	program_builder &operator=(const program_builder &) = delete;
	virtual ~program_builder();
	virtual operation_id get_operation_id() const noexcept = 0;
	// unsupported (never an exception) for launch configurations the builder cannot take.
	virtual backend_priority get_suitability(const operation &operation,
		span<const operand_descriptor> output_descriptors, span<const operand_descriptor> input_descriptors,
		command_queue &queue) const = 0;
	virtual std::shared_ptr<program> build(const operation &operation,
		span<const operand_descriptor> output_descriptors, span<const operand_descriptor> input_descriptors,
		command_queue &queue, program_cache *cache) const = 0;
};

// ---------------------------------------------------------------------------
// Services and plugins
// ---------------------------------------------------------------------------

class REXLIB_API service
{
public:
	virtual ~service();
};

class REXLIB_API device_service final : public service
{
public:
	device_service();
	~device_service() override;
	void enumerate_backends(std::vector<std::string> &names) const;
	device_backend *get_backend(const std::string &name) const;
	bool register_backend(std::unique_ptr<device_backend> backend);
	void enumerate_devices(std::vector<device_index> &indices) const;
	bool get_device_properties(const device_index &index, device_properties &properties) const;
	std::shared_ptr<device_session> create_device_session(const device_index &index) const;

private:
	std::vector<std::unique_ptr<device_backend>> m_backends;
};

class REXLIB_API program_service : public service
{
public:
	program_service() noexcept;
	~program_service() override;
	bool register_builder(std::unique_ptr<program_builder> builder);
	// The program of the most suitable builder; unsupported_operation_error
	// when every builder declines (src/core/dispatch/program_service.cpp:84-89).
	std::shared_ptr<program> build(const operation &operation, span<const operand_descriptor> output_descriptors,
		span<const operand_descriptor> input_descriptors, command_queue &queue, program_cache *cache = nullptr) const;
	// The operations that have at least one builder.
	std::vector<std::string> enumerate_operations() const;

private:
	std::vector<std::unique_ptr<program_builder>> m_builders;
};

class REXLIB_API service_catalog
{
public:
	service_catalog();
	~service_catalog();
	service_catalog(const service_catalog &) = delete;
	service_catalog &operator=(const service_catalog &) = delete;
	template <typename T>
	std::shared_ptr<T> get_service()
	{
		return std::static_pointer_cast<T>(get_service(std::type_index(typeid(T))));
	}

private:
	std::shared_ptr<service> get_service(std::type_index key);
	std::shared_ptr<device_service> m_devices;
	std::shared_ptr<program_service> m_programs;
};

class REXLIB_API plugin
{
public:
	plugin() = default;
	virtual ~plugin() = default;
	virtual const std::string &get_name() const noexcept = 0;
	virtual version get_version() const noexcept = 0;
	virtual void register_at(service_catalog &catalog) const = 0;
};

struct plugin_compatibility
{
	std::uint32_t size;
	std::uint32_t rexlib_version_major;
	std::uint32_t rexlib_version_minor;
	std::uint32_t rexlib_version_patch;
	std::uint32_t abi_family;
	std::uint32_t standard_library;
	std::uint32_t standard_library_abi;
	std::uint32_t runtime;
};

// The record of the library itself.
REXLIB_API const plugin_compatibility &get_library_compatibility() noexcept;

// Throws plugin_load_error when a plugin built as `plugin_record` cannot be
// loaded by a library built as `library_record` (src/core/plugin_compatibility_check.cpp).
REXLIB_API void check_plugin_compatibility(const plugin_compatibility &plugin_record,
	const plugin_compatibility &library_record);

// Loads the plugin in the shared object at `path` and registers it at
// `catalog`; the object is never unloaded. Returns the plugin.
REXLIB_API const plugin &load_plugin(const std::string &path, service_catalog &catalog);

} // namespace rexlib

#define REXLIB_PLUGIN_COMPATIBILITY                                                                                  \
	{                                                                                                                \
		static_cast<std::uint32_t>(sizeof(::rexlib::plugin_compatibility)), REXLIB_VERSION_MAJOR,                   \
			REXLIB_VERSION_MINOR, REXLIB_VERSION_PATCH, REXLIB_CPP_ABI_FAMILY, REXLIB_CPP_STANDARD_LIBRARY,          \
			REXLIB_CPP_STANDARD_LIBRARY_ABI, REXLIB_CPP_RUNTIME                                                     \
	}

// The two C entry points through which the loader finds a plugin
// (include/rexlib/core/plugin_definition.hpp:26-50): the compatibility
// record, read before anything else, and the plugin, constructed on the
// first call (null when its constructor throws).
#define REXLIB_DEFINE_PLUGIN(type)                                                                                  \
	extern "C" REXLIB_EXPORT const ::rexlib::plugin_compatibility *rexlib_get_plugin_compatibility()                 \
	{                                                                                                                \
		static const ::rexlib::plugin_compatibility compatibility = REXLIB_PLUGIN_COMPATIBILITY;                     \
		return &compatibility;                                                                                       \
	}                                                                                                                \
	extern "C" REXLIB_EXPORT const ::rexlib::plugin *rexlib_get_plugin() noexcept                                    \
	{                                                                                                                \
		try                                                                                                          \
		{                                                                                                            \
			static const type instance{};                                                                            \
			return &instance;                                                                                        \
		}                                                                                                            \
		catch (...)                                                                                                  \
		{                                                                                                            \
			return nullptr;                                                                                          \
		}                                                                                                            \
	}                                                                                                                \
	static_assert(std::is_base_of<::rexlib::plugin, type>::value,                                                    \
		"REXLIB_DEFINE_PLUGIN takes a class derived from rexlib::plugin")
