// The core side of the mirrored rexlib API: services, plugin loading, arrays,
// the eager dispatcher and the functions of the insertion path. It holds no
// computation: every operation runs as a program of a registered plugin.
#include "rexlib/em.hpp"
#include "rexlib/fourier.hpp"

#include <cxxabi.h>
#include <dlfcn.h>

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <sstream>

namespace rexlib
{

// ---------------------------------------------------------------------------
// Values
// ---------------------------------------------------------------------------

const char *to_string(backend_priority priority) noexcept
{
	switch (priority)
	{
	case backend_priority::unsupported:
		return "unsupported";
	case backend_priority::fallback:
		return "fallback";
	case backend_priority::normal:
		return "normal";
	case backend_priority::optimal:
		return "optimal";
	}
	return "unknown";
}

std::size_t get_size(numerical_type type) noexcept
{
	switch (type)
	{
	case numerical_type::boolean:
	case numerical_type::char8:
	case numerical_type::int8:
	case numerical_type::uint8:
		return 1;
	case numerical_type::int16:
	case numerical_type::uint16:
	case numerical_type::float16:
		return 2;
	case numerical_type::int32:
	case numerical_type::uint32:
	case numerical_type::float32:
	case numerical_type::complex_float16:
		return 4;
	case numerical_type::int64:
	case numerical_type::uint64:
	case numerical_type::float64:
	case numerical_type::complex_float32:
		return 8;
	case numerical_type::complex_float64:
		return 16;
	default:
		return 0;
	}
}

const char *to_string(numerical_type type) noexcept
{
	static const char *names[] = {"boolean", "char8", "int8", "uint8", "int16", "uint16", "int32", "uint32",
		"int64", "uint64", "float16", "float32", "float64", "complex_float16", "complex_float32",
		"complex_float64"};
	const int i = static_cast<int>(type);
	return i >= 0 && i < static_cast<int>(numerical_type::count) ? names[i] : "unknown";
}

// Scores of src/core/hardware/memory_resource_kind.cpp.
int score_host_affinity(memory_resource_kind kind) noexcept
{
	switch (kind)
	{
	case memory_resource_kind::host:
	case memory_resource_kind::host_staging:
	case memory_resource_kind::device_mapped:
	case memory_resource_kind::unified:
		return 2;
	case memory_resource_kind::managed:
		return 1;
	default:
		return 0;
	}
}

int score_device_affinity(memory_resource_kind kind) noexcept
{
	switch (kind)
	{
	case memory_resource_kind::device_local:
		return 4;
	case memory_resource_kind::unified:
		return 3;
	case memory_resource_kind::managed:
		return 2;
	case memory_resource_kind::device_mapped:
		return 1;
	default:
		return 0;
	}
}

bool is_host_accessible(memory_resource_kind kind) noexcept
{
	return score_host_affinity(kind) > 0;
}

bool is_device_accessible(memory_resource_kind kind) noexcept
{
	return score_device_affinity(kind) > 0;
}

// ---------------------------------------------------------------------------
// Layouts
// ---------------------------------------------------------------------------

strided_layout::strided_layout(std::vector<std::size_t> extents, std::vector<std::ptrdiff_t> strides,
	std::ptrdiff_t offset)
	: m_extents(std::move(extents)), m_strides(std::move(strides)), m_offset(offset)
{
	if (m_extents.size() != m_strides.size())
	{
		throw std::invalid_argument("strided_layout: extents and strides differ in rank");
	}
}

strided_layout strided_layout::make_contiguous(std::vector<std::size_t> extents)
{
	std::vector<std::ptrdiff_t> strides(extents.size());
	std::ptrdiff_t step = 1;
	for (std::size_t i = extents.size(); i-- > 0;)
	{
		strides[i] = step;
		step *= static_cast<std::ptrdiff_t>(extents[i]);
	}
	return strided_layout(std::move(extents), std::move(strides), 0);
}

std::size_t strided_layout::compute_element_count() const noexcept
{
	std::size_t count = 1;
	for (std::size_t e : m_extents)
	{
		count *= e;
	}
	return count;
}

namespace
{

// Elements spanned from the offset to the last addressed element, plus one.
std::size_t span_elements(const strided_layout &layout)
{
	if (layout.compute_element_count() == 0)
	{
		return 0;
	}
	std::vector<std::size_t> extents;
	std::vector<std::ptrdiff_t> strides;
	layout.get_extents(extents);
	layout.get_strides(strides);
	std::ptrdiff_t last = 0;
	for (std::size_t i = 0; i < extents.size(); ++i)
	{
		last += static_cast<std::ptrdiff_t>(extents[i] - 1) * std::max<std::ptrdiff_t>(strides[i], 0);
	}
	return static_cast<std::size_t>(last + 1);
}

} // namespace

// ---------------------------------------------------------------------------
// Device properties
// ---------------------------------------------------------------------------

struct device_properties::implementation
{
	device_type type = device_type::unknown;
	std::string name;
	std::string location;
	std::size_t total_memory = 0;
	std::size_t alignment = 256;
};

device_properties::device_properties() : m_implementation(new implementation) {}
device_properties::~device_properties() = default;
device_properties::device_properties(const device_properties &other)
	: m_implementation(new implementation(*other.m_implementation))
{
}
device_properties &device_properties::operator=(const device_properties &other)
{
	*m_implementation = *other.m_implementation;
	return *this;
}
void device_properties::set_type(device_type type) { m_implementation->type = type; }
device_type device_properties::get_type() const noexcept { return m_implementation->type; }
void device_properties::set_name(const std::string &name) { m_implementation->name = name; }
const std::string &device_properties::get_name() const noexcept { return m_implementation->name; }
void device_properties::set_physical_location(const std::string &location) { m_implementation->location = location; }
const std::string &device_properties::get_physical_location() const noexcept { return m_implementation->location; }
void device_properties::set_total_memory_bytes(std::size_t bytes) { m_implementation->total_memory = bytes; }
std::size_t device_properties::get_total_memory_bytes() const noexcept { return m_implementation->total_memory; }
void device_properties::set_optimal_data_alignment(std::size_t alignment)
{
	if (alignment == 0 || (alignment & (alignment - 1)) != 0)
	{
		throw std::invalid_argument("device_properties: the alignment must be a power of two");
	}
	m_implementation->alignment = alignment;
}
std::size_t device_properties::get_optimal_data_alignment() const noexcept { return m_implementation->alignment; }

// ---------------------------------------------------------------------------
// Hardware abstraction: out-of-line constructors and destructors
// ---------------------------------------------------------------------------

buffer::buffer() noexcept = default;
buffer::~buffer() = default;
memory_resource::memory_resource() noexcept = default;
memory_resource::~memory_resource() = default;
memory_allocator::memory_allocator() noexcept = default;
memory_allocator::~memory_allocator() = default;
command_timeline::command_timeline() noexcept = default;
command_timeline::~command_timeline() = default;
program::program() noexcept = default;
program::~program() = default;
command_queue::command_queue() noexcept = default;
command_queue::~command_queue() = default;
device::device() noexcept = default;
device::~device() = default;
device_backend::device_backend() noexcept = default;
device_backend::~device_backend() = default;
operation::operation() noexcept = default;
operation::~operation() = default;
program_builder::program_builder() noexcept = default;
program_builder::~program_builder() = default;
service::~service() = default;
dispatcher::dispatcher() noexcept = default;
dispatcher::~dispatcher() = default;

span<const program_scratch_requirement> program::get_scratch_requirements() const
{
	return {};
}

command_token::command_token() noexcept = default;

command_token::command_token(std::shared_ptr<const command_timeline> timeline, std::size_t id) noexcept
	: m_timeline(std::move(timeline)), m_id(id)
{
}

bool command_token::is_complete() const
{
	return !m_timeline || m_timeline->is_complete(m_id);
}

void command_token::wait() const
{
	if (m_timeline)
	{
		m_timeline->wait(m_id);
	}
}

command::command() noexcept = default;

command::command(std::shared_ptr<const program> program) noexcept : m_program(std::move(program)) {}

command &command::bind_outputs(std::vector<std::shared_ptr<buffer>> outputs) noexcept
{
	m_outputs = std::move(outputs);
	return *this;
}

command &command::bind_inputs(std::vector<std::shared_ptr<const buffer>> inputs) noexcept
{
	m_inputs = std::move(inputs);
	return *this;
}

command &command::bind_scratch(std::vector<std::shared_ptr<buffer>> scratch) noexcept
{
	m_scratch = std::move(scratch);
	return *this;
}

command &command::bind_dependencies(std::vector<command_token> dependencies) noexcept
{
	m_dependencies = std::move(dependencies);
	return *this;
}

// Host memory of the core: plain heap blocks, ready at once.
namespace
{

class host_resource final : public memory_resource
{
public:
	memory_resource_kind get_kind() const noexcept override { return memory_resource_kind::host; }
	std::shared_ptr<memory_allocator> create_allocator() const override;
};

class host_buffer final : public buffer
{
public:
	host_buffer(std::size_t size, std::size_t alignment, const memory_resource &resource)
		: m_size(size), m_resource(resource)
	{
		const std::size_t a = std::max<std::size_t>(alignment, alignof(std::max_align_t));
		m_data = std::aligned_alloc(a, std::max<std::size_t>((size + a - 1) / a * a, a));
		if (!m_data)
		{
			throw std::bad_alloc();
		}
		std::memset(m_data, 0, size);
	}
	~host_buffer() override { std::free(m_data); }
	void *get_host_ptr() noexcept override { return m_data; }
	const void *get_host_ptr() const noexcept override { return m_data; }
	std::size_t get_size() const noexcept override { return m_size; }
	const memory_resource &get_memory_resource() const noexcept override { return m_resource; }

private:
	void *m_data = nullptr;
	std::size_t m_size;
	const memory_resource &m_resource;
};

class host_allocator final : public memory_allocator
{
public:
	explicit host_allocator(const memory_resource &resource) : m_resource(resource) {}
	const memory_resource &get_memory_resource() const noexcept override { return m_resource; }
	std::size_t get_max_alignment() const noexcept override { return 4096; }
	std::shared_ptr<buffer> allocate(std::size_t size, std::size_t alignment, command_queue *, command_token &pending)
		override
	{
		pending = command_token();
		return std::make_shared<host_buffer>(size, alignment, m_resource);
	}

private:
	const memory_resource &m_resource;
};

std::shared_ptr<memory_allocator> host_resource::create_allocator() const
{
	return std::make_shared<host_allocator>(*this);
}

} // namespace

const memory_resource &get_host_memory_resource() noexcept
{
	static const host_resource resource;
	return resource;
}

device_session::device_session(device_index index, std::shared_ptr<device> dev, device_properties properties)
	: m_index(std::move(index)), m_device(std::move(dev)), m_properties(std::move(properties))
{
	m_queue = m_device->create_command_queue();
	m_allocators[static_cast<std::size_t>(memory_resource_affinity::host)] =
		m_device->get_memory_resource(memory_resource_affinity::host).create_allocator();
	m_allocators[static_cast<std::size_t>(memory_resource_affinity::device)] =
		m_device->get_memory_resource(memory_resource_affinity::device).create_allocator();
}

// ---------------------------------------------------------------------------
// Services
// ---------------------------------------------------------------------------

device_service::device_service() = default;
device_service::~device_service() = default;

void device_service::enumerate_backends(std::vector<std::string> &names) const
{
	names.clear();
	for (const auto &b : m_backends)
	{
		names.push_back(b->get_name());
	}
}

device_backend *device_service::get_backend(const std::string &name) const
{
	for (const auto &b : m_backends)
	{
		if (b->get_name() == name)
		{
			return b.get();
		}
	}
	return nullptr;
}

bool device_service::register_backend(std::unique_ptr<device_backend> backend)
{
	if (!backend || get_backend(backend->get_name()))
	{
		return false;
	}
	m_backends.push_back(std::move(backend));
	return true;
}

void device_service::enumerate_devices(std::vector<device_index> &indices) const
{
	indices.clear();
	for (const auto &b : m_backends)
	{
		std::vector<std::size_t> ids;
		b->enumerate_devices(ids);
		for (std::size_t id : ids)
		{
			indices.emplace_back(b->get_name(), id);
		}
	}
}

bool device_service::get_device_properties(const device_index &index, device_properties &properties) const
{
	const device_backend *b = get_backend(index.get_backend_name());
	return b && b->get_device_properties(index.get_device_id(), properties);
}

std::shared_ptr<device_session> device_service::create_device_session(const device_index &index) const
{
	const device_backend *b = get_backend(index.get_backend_name());
	if (!b)
	{
		throw std::invalid_argument("create_device_session: no backend named " + index.get_backend_name());
	}
	device_properties properties;
	if (!b->get_device_properties(index.get_device_id(), properties))
	{
		throw std::invalid_argument("create_device_session: " + index.get_backend_name() + " has no device " +
			std::to_string(index.get_device_id()));
	}
	return std::make_shared<device_session>(index, b->create_device(index.get_device_id()), properties);
}

program_service::program_service() noexcept = default;
program_service::~program_service() = default;

bool program_service::register_builder(std::unique_ptr<program_builder> builder)
{
	if (!builder)
	{
		return false;
	}
	m_builders.push_back(std::move(builder));
	return true;
}

std::shared_ptr<program> program_service::build(const operation &operation,
	span<const operand_descriptor> output_descriptors, span<const operand_descriptor> input_descriptors,
	command_queue &queue, program_cache *cache) const
{
	const program_builder *best = nullptr;
	backend_priority best_priority = backend_priority::unsupported;
	for (const auto &b : m_builders)
	{
		if (b->get_operation_id() != operation.get_id())
		{
			continue;
		}
		const backend_priority p = b->get_suitability(operation, output_descriptors, input_descriptors, queue);
		if (p != backend_priority::unsupported && (!best || static_cast<int>(p) > static_cast<int>(best_priority)))
		{
			best = b.get();
			best_priority = p;
		}
	}
	if (!best)
	{
		throw unsupported_operation_error(std::string("Could not find a suitable operation program builder for the "
													  "requested operation (") +
			operation.get_name() + ")");
	}
	return best->build(operation, output_descriptors, input_descriptors, queue, cache);
}

std::vector<std::string> program_service::enumerate_operations() const
{
	std::vector<std::string> names;
	for (const auto &b : m_builders)
	{
		int status = 0;
		char *readable = abi::__cxa_demangle(b->get_operation_id().get_name(), nullptr, nullptr, &status);
		const std::string name = status == 0 && readable ? readable : b->get_operation_id().get_name();
		std::free(readable);
		if (std::find(names.begin(), names.end(), name) == names.end())
		{
			names.push_back(name);
		}
	}
	return names;
}

service_catalog::service_catalog()
	: m_devices(std::make_shared<device_service>()), m_programs(std::make_shared<program_service>())
{
}

service_catalog::~service_catalog() = default;

std::shared_ptr<service> service_catalog::get_service(std::type_index key)
{
	if (key == std::type_index(typeid(device_service)))
	{
		return m_devices;
	}
	if (key == std::type_index(typeid(program_service)))
	{
		return m_programs;
	}
	return nullptr;
}

// ---------------------------------------------------------------------------
// Plugins
// ---------------------------------------------------------------------------

const plugin_compatibility &get_library_compatibility() noexcept
{
	static const plugin_compatibility record = REXLIB_PLUGIN_COMPATIBILITY;
	return record;
}

void check_plugin_compatibility(const plugin_compatibility &plugin_record, const plugin_compatibility &library_record)
{
	const std::size_t first_size = offsetof(plugin_compatibility, runtime) + sizeof(std::uint32_t);
	if (plugin_record.size < first_size)
	{
		throw plugin_load_error("The compatibility record of the plugin holds " + std::to_string(plugin_record.size) +
			" bytes, fewer than the " + std::to_string(first_size) + " every plugin states.");
	}
	// Before 1.0 every minor release may break the interface.
	const bool same_major = plugin_record.rexlib_version_major == library_record.rexlib_version_major;
	const bool version_ok = same_major &&
		(library_record.rexlib_version_major == 0
				? plugin_record.rexlib_version_minor == library_record.rexlib_version_minor
				: plugin_record.rexlib_version_minor <= library_record.rexlib_version_minor);
	if (!version_ok)
	{
		std::ostringstream message;
		message << "The plugin was built against rexlib " << plugin_record.rexlib_version_major << "."
				<< plugin_record.rexlib_version_minor << "." << plugin_record.rexlib_version_patch << ", but rexlib "
				<< library_record.rexlib_version_major << "." << library_record.rexlib_version_minor << "."
				<< library_record.rexlib_version_patch << " is loaded.";
		throw plugin_load_error(message.str());
	}
	if (plugin_record.abi_family != library_record.abi_family ||
		plugin_record.standard_library != library_record.standard_library ||
		plugin_record.standard_library_abi != library_record.standard_library_abi ||
		plugin_record.runtime != library_record.runtime)
	{
		std::ostringstream message;
		message << "The plugin was built for C++ ABI " << plugin_record.abi_family << " with standard library "
				<< plugin_record.standard_library << " (ABI " << plugin_record.standard_library_abi
				<< "), which differs from the library's (" << library_record.abi_family << ", "
				<< library_record.standard_library << ", " << library_record.standard_library_abi << ").";
		throw plugin_load_error(message.str());
	}
}

const plugin &load_plugin(const std::string &path, service_catalog &catalog)
{
	// Never closed: the plugin's objects live as long as the process.
	void *handle = dlopen(path.c_str(), RTLD_NOW | RTLD_LOCAL);
	if (!handle)
	{
		const char *why = dlerror();
		throw plugin_load_error("Could not load " + path + ": " + (why ? why : "unknown error"));
	}
	using compatibility_entry = const plugin_compatibility *(*)();
	using plugin_entry = const plugin *(*)();
	auto get_compatibility = reinterpret_cast<compatibility_entry>(dlsym(handle, "rexlib_get_plugin_compatibility"));
	auto get_plugin = reinterpret_cast<plugin_entry>(dlsym(handle, "rexlib_get_plugin"));
	if (!get_compatibility || !get_plugin)
	{
		throw plugin_load_error(path + " does not define a rexlib plugin (REXLIB_DEFINE_PLUGIN)");
	}
	const plugin_compatibility *record = get_compatibility();
	if (!record)
	{
		throw plugin_load_error(path + " returned no compatibility record");
	}
	check_plugin_compatibility(*record, get_library_compatibility());
	const plugin *instance = get_plugin();
	if (!instance)
	{
		throw plugin_load_error("The plugin in " + path + " could not be constructed");
	}
	instance->register_at(catalog);
	return *instance;
}

// ---------------------------------------------------------------------------
// Arrays and contexts
// ---------------------------------------------------------------------------

array::array(std::shared_ptr<buffer> storage, array_descriptor descriptor)
	: m_buffer(std::move(storage)), m_descriptor(std::move(descriptor))
{
}

namespace
{
std::mutex hazard_mutex;
}

std::vector<command_token> array::get_write_hazards() const
{
	std::lock_guard<std::mutex> lock(hazard_mutex);
	std::vector<command_token> tokens = m_hazards->reads;
	tokens.push_back(m_hazards->last_write);
	return tokens;
}

command_token array::get_read_hazard() const
{
	std::lock_guard<std::mutex> lock(hazard_mutex);
	return m_hazards->last_write;
}

void array::record_write(const command_token &token) const
{
	std::lock_guard<std::mutex> lock(hazard_mutex);
	m_hazards->last_write = token;
	m_hazards->reads.clear();
}

void array::record_read(const command_token &token) const
{
	std::lock_guard<std::mutex> lock(hazard_mutex);
	auto &reads = m_hazards->reads;
	reads.erase(std::remove_if(reads.begin(), reads.end(), [](const command_token &t) { return t.is_complete(); }),
		reads.end());
	reads.push_back(token);
}

void array::wait() const
{
	for (const command_token &t : get_write_hazards())
	{
		t.wait();
	}
}

device_context::device_context(std::shared_ptr<const device_session> session) : m_session(std::move(session))
{
	m_active_queue = m_session->get_default_queue();
	for (std::size_t a = 0; a < static_cast<std::size_t>(memory_resource_affinity::count); ++a)
	{
		m_allocators[a] = m_session->get_allocator(static_cast<memory_resource_affinity>(a));
	}
}

device_context device_context::on_queue(std::shared_ptr<command_queue> queue) const
{
	device_context c = *this;
	c.m_active_queue = std::move(queue);
	return c;
}

execution_context::execution_context(device_context device_context, std::shared_ptr<dispatcher> dispatcher)
	: m_device_context(std::move(device_context)), m_dispatcher(std::move(dispatcher))
{
}

execution_context execution_context::with_device_context(device_context context) const
{
	return execution_context(std::move(context), m_dispatcher);
}

namespace
{

std::size_t bit_ceil(std::size_t v)
{
	std::size_t p = 1;
	while (p < v)
	{
		p <<= 1;
	}
	return p;
}

class eager_dispatcher final : public dispatcher
{
public:
	explicit eager_dispatcher(std::shared_ptr<const program_service> programs) : m_programs(std::move(programs)) {}

	void dispatch(const operation &operation, span<array> outputs, span<const array_descriptor> planned,
		span<const const_array_ref> inputs, const device_context &device) override
	{
		command_queue &queue = *device.get_active_queue();
		std::vector<command_token> dependencies;
		// Missing outputs come from the device allocator, aligned to
		// min(allocator maximum, device optimum, bit_ceil(size)).
		for (std::size_t i = 0; i < outputs.size(); ++i)
		{
			if (!outputs[i].is_empty())
			{
				continue;
			}
			if (i >= planned.size())
			{
				throw std::invalid_argument(std::string(operation.get_name()) + ": output " + std::to_string(i) +
					" has no planned layout");
			}
			const array_descriptor &d = planned[i];
			const std::size_t size = span_elements(d.get_layout()) * get_size(d.get_data_type());
			const auto &allocator = device.get_allocator(memory_resource_affinity::device);
			const std::size_t alignment = std::min({allocator->get_max_alignment(),
				device.get_device_session()->get_properties().get_optimal_data_alignment(),
				bit_ceil(std::max<std::size_t>(size, 1))});
			command_token pending;
			outputs[i] = array(allocator->allocate(std::max<std::size_t>(size, 1), alignment, &queue, pending), d);
			dependencies.push_back(pending);
		}
		std::vector<operand_descriptor> out_d, in_d;
		std::vector<std::shared_ptr<buffer>> out_b;
		std::vector<std::shared_ptr<const buffer>> in_b;
		for (const array &a : outputs)
		{
			out_d.push_back(a.get_operand_descriptor());
			out_b.push_back(a.get_buffer());
			for (const command_token &t : a.get_write_hazards())
			{
				dependencies.push_back(t);
			}
		}
		for (const const_array_ref &r : inputs)
		{
			in_d.push_back(r->get_operand_descriptor());
			in_b.push_back(r->get_buffer());
			dependencies.push_back(r->get_read_hazard());
		}
		dependencies.erase(std::remove_if(dependencies.begin(), dependencies.end(),
							   [](const command_token &t) { return t.is_empty(); }),
			dependencies.end());
		std::shared_ptr<program> p = m_programs->build(operation, {out_d.data(), out_d.size()},
			{in_d.data(), in_d.size()}, queue, nullptr);
		command cmd(p);
		cmd.bind_outputs(std::move(out_b)).bind_inputs(std::move(in_b)).bind_dependencies(std::move(dependencies));
		const command_token token = queue.submit(std::move(cmd));
		for (const const_array_ref &r : inputs)
		{
			r->record_read(token);
		}
		for (const array &a : outputs)
		{
			a.record_write(token);
		}
	}

private:
	std::shared_ptr<const program_service> m_programs;
};

void check_context(const execution_context &context)
{
	if (!context.get_dispatcher() || !context.get_device_context().get_device_session())
	{
		throw std::invalid_argument("execution_context: no device session or dispatcher");
	}
}

} // namespace

std::shared_ptr<dispatcher> make_eager_dispatcher(std::shared_ptr<const program_service> programs)
{
	return std::make_shared<eager_dispatcher>(std::move(programs));
}

execution_context make_execution_context(service_catalog &catalog, const device_index &index)
{
	auto session = catalog.get_service<device_service>()->create_device_session(index);
	return execution_context(device_context(session), make_eager_dispatcher(catalog.get_service<program_service>()));
}

array make_host_array(std::vector<std::size_t> extents, numerical_type type)
{
	strided_layout layout = strided_layout::make_contiguous(std::move(extents));
	const std::size_t size = std::max<std::size_t>(layout.compute_element_count() * get_size(type), 1);
	command_token pending;
	auto allocator = get_host_memory_resource().create_allocator();
	return array(allocator->allocate(size, 64, nullptr, pending), array_descriptor(std::move(layout), type));
}

array zeros(std::vector<std::size_t> extents, numerical_type type, const execution_context &context)
{
	check_context(context);
	array out;
	const array_descriptor planned(strided_layout::make_contiguous(std::move(extents)), type);
	context.get_dispatcher()->dispatch(ops::fill_operation(scalar_value(0.0)), {&out, 1}, {&planned, 1}, {},
		context.get_device_context());
	return out;
}

array to_device(const_array_ref x, const execution_context &context, array *out)
{
	check_context(context);
	array local;
	array &target = out ? *out : local;
	const array_descriptor planned(strided_layout::make_contiguous(x->get_extents()), x->get_data_type());
	const const_array_ref input[] = {x};
	context.get_dispatcher()->dispatch(ops::copy_operation(), {&target, 1}, {&planned, 1}, {input, 1},
		context.get_device_context());
	return target;
}

array to_host(const_array_ref x, const execution_context &context)
{
	check_context(context);
	array host = make_host_array(x->get_extents(), x->get_data_type());
	const const_array_ref input[] = {x};
	context.get_dispatcher()->dispatch(ops::copy_operation(), {&host, 1}, {}, {input, 1},
		context.get_device_context());
	host.wait();
	return host;
}

// ---------------------------------------------------------------------------
// Elementwise functions
// ---------------------------------------------------------------------------

array full(std::vector<std::size_t> extents, double value, numerical_type type, const execution_context &context)
{
	check_context(context);
	array out;
	const array_descriptor planned(strided_layout::make_contiguous(std::move(extents)), type);
	context.get_dispatcher()->dispatch(ops::fill_operation(scalar_value(value)), {&out, 1}, {&planned, 1}, {},
		context.get_device_context());
	return out;
}

namespace
{

// Extents of operands broadcast together, numpy's rules (axes aligned from
// the last; an extent of 1 repeats).
std::vector<std::size_t> broadcast_extents(std::initializer_list<const_array_ref> operands, const char *what)
{
	std::size_t rank = 0;
	for (const auto &x : operands)
	{
		rank = std::max(rank, x->get_extents().size());
	}
	std::vector<std::size_t> out(rank, 1);
	for (const auto &x : operands)
	{
		const auto &e = x->get_extents();
		const std::size_t lead = rank - e.size();
		for (std::size_t a = 0; a < e.size(); ++a)
		{
			std::size_t &o = out[lead + a];
			if (e[a] == o || e[a] == 1)
			{
				continue;
			}
			if (o != 1)
			{
				throw std::invalid_argument(std::string(what) + ": operand extents do not broadcast");
			}
			o = e[a];
		}
	}
	return out;
}

array run_elementwise(const operation &op, std::initializer_list<const_array_ref> inputs, numerical_type out_type,
	const execution_context &context, array *out, const char *what)
{
	check_context(context);
	array local;
	array &target = out ? *out : local;
	const array_descriptor planned(strided_layout::make_contiguous(broadcast_extents(inputs, what)), out_type);
	const std::vector<const_array_ref> in(inputs);
	context.get_dispatcher()->dispatch(op, {&target, 1}, {&planned, 1}, {in.data(), in.size()},
		context.get_device_context());
	return target;
}

} // namespace

array divide(const_array_ref x, const_array_ref y, const execution_context &context, array *out)
{
	if (x->get_data_type() != y->get_data_type() || x->get_data_type() == numerical_type::boolean)
	{
		throw std::invalid_argument("divide: both operands take one non-boolean type");
	}
	return run_elementwise(ops::divide_operation(), {x, y}, x->get_data_type(), context, out, "divide");
}

array greater(const_array_ref x, const_array_ref y, const execution_context &context, array *out)
{
	if (x->get_data_type() != y->get_data_type())
	{
		throw std::invalid_argument("greater: both operands take one type");
	}
	return run_elementwise(ops::greater_operation(), {x, y}, numerical_type::boolean, context, out, "greater");
}

array where(const_array_ref condition, const_array_ref x, const_array_ref y, const execution_context &context,
	array *out)
{
	if (condition->get_data_type() != numerical_type::boolean || x->get_data_type() != y->get_data_type())
	{
		throw std::invalid_argument("where: a boolean condition and two operands of one type");
	}
	return run_elementwise(ops::where_operation(), {condition, x, y}, x->get_data_type(), context, out, "where");
}

// ---------------------------------------------------------------------------
// Fourier transforms
// ---------------------------------------------------------------------------

namespace
{

array run_transform(const operation &op, const_array_ref x, std::vector<std::size_t> out_extents,
	numerical_type out_type, const execution_context &context, array *out)
{
	check_context(context);
	array local;
	array &target = out ? *out : local;
	const array_descriptor planned(strided_layout::make_contiguous(std::move(out_extents)), out_type);
	const const_array_ref input[] = {x};
	context.get_dispatcher()->dispatch(op, {&target, 1}, {&planned, 1}, {input, 1}, context.get_device_context());
	return target;
}

numerical_type complex_of(numerical_type t)
{
	switch (t)
	{
	case numerical_type::float16:
		return numerical_type::complex_float16;
	case numerical_type::float32:
		return numerical_type::complex_float32;
	case numerical_type::float64:
		return numerical_type::complex_float64;
	default:
		throw std::invalid_argument(std::string("rfft: a real floating-point signal is required, not ") +
			to_string(t));
	}
}

numerical_type real_of(numerical_type t)
{
	switch (t)
	{
	case numerical_type::complex_float16:
		return numerical_type::float16;
	case numerical_type::complex_float32:
		return numerical_type::float32;
	case numerical_type::complex_float64:
		return numerical_type::float64;
	default:
		throw std::invalid_argument(std::string("irfft: a complex spectrum is required, not ") + to_string(t));
	}
}

void require_rank(const_array_ref x, std::size_t rank, const char *what)
{
	if (x->get_extents().size() < rank)
	{
		throw std::invalid_argument(std::string(what) + ": the operand has fewer than " + std::to_string(rank) +
			" axes");
	}
}

} // namespace

namespace
{

// Axes as absolute indices, every axis when empty; repeated axes refused.
ops::axis_list absolute_axes(signed_axes given, std::size_t rank, const char *what)
{
	if (given.empty())
	{
		for (std::size_t a = 0; a < rank; ++a)
		{
			given.push_back(static_cast<std::ptrdiff_t>(a));
		}
	}
	ops::axis_list axes;
	for (auto a : given)
	{
		if (a < 0)
		{
			a += static_cast<std::ptrdiff_t>(rank);
		}
		if (a < 0 || a >= static_cast<std::ptrdiff_t>(rank))
		{
			throw std::invalid_argument(std::string(what) + ": an axis lies outside the operand");
		}
		axes.push_back(static_cast<std::size_t>(a));
	}
	for (std::size_t i = 0; i < axes.size(); ++i)
	{
		for (std::size_t j = i + 1; j < axes.size(); ++j)
		{
			if (axes[i] == axes[j])
			{
				throw std::invalid_argument(std::string(what) + ": an axis is repeated");
			}
		}
	}
	return axes;
}

numerical_type complex_result(numerical_type t, const char *what)
{
	switch (t)
	{
	case numerical_type::float32:
	case numerical_type::complex_float32:
		return numerical_type::complex_float32;
	case numerical_type::float64:
	case numerical_type::complex_float64:
		return numerical_type::complex_float64;
	default:
		throw std::invalid_argument(std::string(what) + ": a floating-point operand is required, not " +
			to_string(t));
	}
}

} // namespace

array fftn(const_array_ref x, signed_axes given, const execution_context &context,
	ops::fourier_normalization normalization, array *out)
{
	const ops::axis_list axes = absolute_axes(std::move(given), x->get_extents().size(), "fftn");
	return run_transform(ops::fft_operation(axes, normalization), x, x->get_extents(),
		complex_result(x->get_data_type(), "fftn"), context, out);
}

array ifftn(const_array_ref x, signed_axes given, const execution_context &context,
	ops::fourier_normalization normalization, array *out)
{
	const ops::axis_list axes = absolute_axes(std::move(given), x->get_extents().size(), "ifftn");
	return run_transform(ops::ifft_operation(axes, normalization), x, x->get_extents(),
		complex_result(x->get_data_type(), "ifftn"), context, out);
}

array rfftn(const_array_ref x, signed_axes given, const execution_context &context,
	ops::fourier_normalization normalization, array *out)
{
	const ops::axis_list axes = absolute_axes(std::move(given), x->get_extents().size(), "rfftn");
	std::vector<std::size_t> e = x->get_extents();
	e[static_cast<std::size_t>(axes.back())] = e[static_cast<std::size_t>(axes.back())] / 2 + 1;
	return run_transform(ops::rfft_operation(axes, normalization), x, e, complex_of(x->get_data_type()), context,
		out);
}

array irfftn(const_array_ref x, signed_axes given, std::size_t extent, const execution_context &context,
	ops::fourier_normalization normalization, array *out)
{
	const ops::axis_list axes = absolute_axes(std::move(given), x->get_extents().size(), "irfftn");
	std::vector<std::size_t> e = x->get_extents();
	const std::size_t last = static_cast<std::size_t>(axes.back());
	if (e[last] != extent / 2 + 1)
	{
		throw std::invalid_argument("irfftn: the halved axis must hold extent / 2 + 1 coefficients");
	}
	e[last] = extent;
	return run_transform(ops::irfft_operation(axes, extent, normalization), x, e, real_of(x->get_data_type()),
		context, out);
}

array rfft2(const_array_ref x, ops::fourier_normalization normalization, const execution_context &context, array *out)
{
	require_rank(x, 2, "rfft2");
	std::vector<std::size_t> e = x->get_extents();
	e.back() = e.back() / 2 + 1;
	return run_transform(ops::rfft_operation(absolute_axes({-2, -1}, e.size(), "rfft2"), normalization), x, e,
		complex_of(x->get_data_type()), context, out);
}

array rfft2(const_array_ref x, const execution_context &context, array *out)
{
	return rfft2(x, ops::fourier_normalization::backward, context, out);
}

array irfft2(const_array_ref x, std::size_t extent, const execution_context &context, array *out)
{
	require_rank(x, 2, "irfft2");
	if (x->get_extents().back() != extent / 2 + 1)
	{
		throw std::invalid_argument("irfft2: the last axis must hold extent / 2 + 1 coefficients");
	}
	std::vector<std::size_t> e = x->get_extents();
	e.back() = extent;
	const ops::axis_list axes = absolute_axes({-2, -1}, e.size(), "irfft2");
	return run_transform(ops::irfft_operation(axes, extent), x, e, real_of(x->get_data_type()), context, out);
}

array irfft3(const_array_ref x, std::size_t extent, ops::fourier_normalization normalization,
	const execution_context &context, array *out)
{
	require_rank(x, 3, "irfft3");
	if (x->get_extents().back() != extent / 2 + 1)
	{
		throw std::invalid_argument("irfft3: the last axis must hold extent / 2 + 1 coefficients");
	}
	std::vector<std::size_t> e = x->get_extents();
	e.back() = extent;
	const ops::axis_list axes = absolute_axes({-3, -2, -1}, e.size(), "irfft3");
	return run_transform(ops::irfft_operation(axes, extent, normalization), x, e, real_of(x->get_data_type()), context,
		out);
}

array irfft3(const_array_ref x, std::size_t extent, const execution_context &context, array *out)
{
	return irfft3(x, extent, ops::fourier_normalization::backward, context, out);
}

// ---------------------------------------------------------------------------
// em: parameters and insertion
// ---------------------------------------------------------------------------

namespace em
{

gridding_window gridding_window::nearest() noexcept
{
	return gridding_window(gridding_profile::box, 0.5, 0.0);
}

gridding_window gridding_window::linear() noexcept
{
	return gridding_window(gridding_profile::triangle, 1.0, 0.0);
}

gridding_window gridding_window::kaiser_bessel(double half_width, double shape)
{
	if (!(std::isfinite(half_width) && half_width > 0.0))
	{
		std::ostringstream oss;
		oss << "gridding_window: the half-width of a Kaiser-Bessel window must be finite and positive, but it is "
			<< half_width << ".";
		throw std::invalid_argument(oss.str());
	}
	if (!(std::isfinite(shape) && shape >= 0.0))
	{
		std::ostringstream oss;
		oss << "gridding_window: the shape of a Kaiser-Bessel window must be finite and 0 or more, but it is "
			<< shape << ".";
		throw std::invalid_argument(oss.str());
	}
	return gridding_window(gridding_profile::kaiser_bessel, half_width, shape);
}

ctf_optics::ctf_optics(double voltage, double spherical_aberration, double amplitude_contrast)
	: m_voltage(voltage), m_spherical_aberration(spherical_aberration), m_amplitude_contrast(amplitude_contrast)
{
	if (!(std::isfinite(voltage) && voltage > 0.0))
	{
		throw std::invalid_argument("ctf_optics: the voltage must be finite and positive, but it is " +
			std::to_string(voltage) + ".");
	}
	if (!std::isfinite(spherical_aberration))
	{
		throw std::invalid_argument("ctf_optics: the spherical aberration must be finite, but it is " +
			std::to_string(spherical_aberration) + ".");
	}
	if (!(amplitude_contrast >= 0.0 && amplitude_contrast < 1.0))
	{
		throw std::invalid_argument("ctf_optics: the amplitude contrast must lie in [0, 1), but it is " +
			std::to_string(amplitude_contrast) + ".");
	}
}

double ctf_optics::get_wavelength() const noexcept
{
	// lambda = h / sqrt(2 m e V (1 + e V / (2 m c^2))) in angstrom, V in volts.
	const double volts = m_voltage * 1e3;
	return 12.2643247 / std::sqrt(volts * (1.0 + 0.978466e-6 * volts));
}

standardize_background_operation::standardize_background_operation(ops::axis_list axes, double radius,
	double lower_bound, double upper_bound)
	: m_axes(axes), m_radius(radius), m_lower_bound(lower_bound), m_upper_bound(upper_bound)
{
	for (std::size_t i = 0; i < axes.size(); ++i)
	{
		for (std::size_t j = i + 1; j < axes.size(); ++j)
		{
			if (axes[i] == axes[j])
			{
				throw std::invalid_argument("standardize_background: an axis is repeated");
			}
		}
	}
	if (!(radius >= 0.0))
	{
		throw std::invalid_argument("standardize_background: the radius must be at least 0");
	}
	if (!(lower_bound <= 0.0 && 0.0 <= upper_bound))
	{
		throw std::invalid_argument("standardize_background: the bounds must lie on either side of 0");
	}
}

array standardize_background(const_array_ref value, const_array_ref key, ops::axis_list axes, double radius,
	double lower_bound, double upper_bound, const execution_context &context, array *out)
{
	check_context(context);
	const standardize_background_operation op(std::move(axes), radius, lower_bound, upper_bound);
	array local;
	array &target = out ? *out : local;
	const array_descriptor planned(strided_layout::make_contiguous(value->get_extents()), value->get_data_type());
	const const_array_ref inputs[2] = {value, key};
	context.get_dispatcher()->dispatch(op, {&target, 1}, {&planned, 1}, {inputs, 2}, context.get_device_context());
	return target;
}

insert_ctf_slices_operation::insert_ctf_slices_operation(const gridding_window &window, const ctf_optics &optics,
	double pixel_size)
	: m_window(window), m_optics(optics), m_pixel_size(pixel_size)
{
	if (!(std::isfinite(pixel_size) && pixel_size > 0.0))
	{
		std::ostringstream oss;
		oss << "insert_ctf_slices: the pixel size must be finite and positive, but it is " << pixel_size << ".";
		throw std::invalid_argument(oss.str());
	}
}

namespace
{

std::pair<array, array> run_insertion(const operation &op, const_array_ref numerator, const_array_ref denominator,
	std::vector<const_array_ref> inputs, const execution_context &context, array *numerator_out,
	array *denominator_out)
{
	check_context(context);
	array outputs[2];
	if (numerator_out)
	{
		outputs[0] = *numerator_out;
	}
	if (denominator_out)
	{
		outputs[1] = *denominator_out;
	}
	const array_descriptor planned[2] = {
		array_descriptor(strided_layout::make_contiguous(numerator->get_extents()), numerator->get_data_type()),
		array_descriptor(strided_layout::make_contiguous(denominator->get_extents()), denominator->get_data_type())};
	context.get_dispatcher()->dispatch(op, {outputs, 2}, {planned, 2}, {inputs.data(), inputs.size()},
		context.get_device_context());
	if (numerator_out)
	{
		*numerator_out = outputs[0];
	}
	if (denominator_out)
	{
		*denominator_out = outputs[1];
	}
	return {outputs[0], outputs[1]};
}

} // namespace

std::pair<array, array> insert_slices(const_array_ref numerator, const_array_ref denominator, const_array_ref slices,
	const_array_ref transforms, const_array_ref radial_weights, const gridding_window &window,
	const execution_context &context, array *numerator_out, array *denominator_out)
{
	return run_insertion(insert_slices_operation(window), numerator, denominator,
		{numerator, denominator, slices, transforms, radial_weights}, context, numerator_out, denominator_out);
}

std::pair<array, array> insert_weighted_slices(const_array_ref numerator, const_array_ref denominator,
	const_array_ref slices, const_array_ref weights, const_array_ref transforms, const_array_ref radial_weights,
	const gridding_window &window, const execution_context &context, array *numerator_out, array *denominator_out)
{
	return run_insertion(insert_weighted_slices_operation(window), numerator, denominator,
		{numerator, denominator, slices, weights, transforms, radial_weights}, context, numerator_out,
		denominator_out);
}

std::pair<array, array> insert_ctf_slices(const_array_ref numerator, const_array_ref denominator,
	const_array_ref slices, const_array_ref transforms, const_array_ref ctf, const_array_ref radial_weights,
	const gridding_window &window, const ctf_optics &optics, double pixel_size, const execution_context &context,
	array *numerator_out, array *denominator_out)
{
	return run_insertion(insert_ctf_slices_operation(window, optics, pixel_size), numerator, denominator,
		{numerator, denominator, slices, transforms, ctf, radial_weights}, context, numerator_out, denominator_out);
}

} // namespace em
} // namespace rexlib
