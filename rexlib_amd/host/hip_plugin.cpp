// librexlib-hip.so: the MI355X backend as a rexlib plugin.
//
// Defined with REXLIB_DEFINE_PLUGIN, so the loader finds it through the two C
// entry points rexlib_get_plugin_compatibility and rexlib_get_plugin
// (include/rexlib/core/plugin_definition.hpp:26-50). register_at adds
//   - a device backend "hip" (include/rexlib/core/hardware/device_backend.hpp:33-113):
//     one device per HIP device, zero devices (not an error) without a GPU,
//     as plugins/cuda/src/hardware/cuda_device_backend.cpp:37-42 does;
//   - per device, a device-local memory resource (hipMalloc) and a host
//     staging resource (pinned, hipHostMalloc), and command queues, each a
//     HIP stream of default priority with its own event timeline;
//   - program builders for insert_slices, insert_weighted_slices,
//     insert_ctf_slices, fft, ifft, rfft and irfft (any axes; the square
//     images and cubic grids of the path take dedicated kernels), fill, copy,
//     divide, greater, where and standardize_background, which run the kernels of
//     librexlib_hip.so through its C ABI (include/rexlib_hip.h).
//
// Queue semantics follow plugins/cuda/src/hardware/cuda_command_queue.cpp:102-215:
// a command waits for its dependencies (an event wait for another HIP queue,
// a host wait for a foreign timeline, nothing for its own in-order stream),
// its buffers stay alive until its event completes, and a failed launch
// synchronises the stream before the error propagates. A builder returns
// unsupported, never throws, for a queue, type or memory it cannot take, so
// that another backend can take the operation.

// rexlib's own headers, by the paths rexlib installs (built here against the
// restatements under rexlib_amd/host/rexlib; tests/test_plugin_boundary.py
// checks this file against rexlib's headers themselves).
#include <rexlib/core/dispatch/operand_descriptor.hpp>
#include <rexlib/core/dispatch/program_builder.hpp>
#include <rexlib/core/dispatch/program_service.hpp>
#include <rexlib/core/exceptions/unsupported_operation_error.hpp>
#include <rexlib/core/hardware/buffer.hpp>
#include <rexlib/core/hardware/command.hpp>
#include <rexlib/core/hardware/command_queue.hpp>
#include <rexlib/core/hardware/command_timeline.hpp>
#include <rexlib/core/hardware/command_token.hpp>
#include <rexlib/core/hardware/device.hpp>
#include <rexlib/core/hardware/device_backend.hpp>
#include <rexlib/core/hardware/device_properties.hpp>
#include <rexlib/core/hardware/device_service.hpp>
#include <rexlib/core/hardware/memory_allocator.hpp>
#include <rexlib/core/hardware/memory_resource.hpp>
#include <rexlib/core/hardware/program.hpp>
#include <rexlib/core/layout/strided_layout.hpp>
#include <rexlib/core/operation/operation_cast.hpp>
#include <rexlib/core/plugin.hpp>
#include <rexlib/core/plugin_definition.hpp>
#include <rexlib/core/service_catalog.hpp>
#include <rexlib/em/ops/reconstruction/insert_ctf_slices_operation.hpp>
#include <rexlib/em/ops/reconstruction/insert_slices_operation.hpp>
#include <rexlib/em/ops/reconstruction/insert_weighted_slices_operation.hpp>
#include <rexlib/em/ops/preprocessing/standardize_background_operation.hpp>
#include <rexlib/ops/arithmetic/divide_operation.hpp>
#include <rexlib/ops/assignment/copy_operation.hpp>
#include <rexlib/ops/assignment/fill_operation.hpp>
#include <rexlib/ops/compare/greater_operation.hpp>
#include <rexlib/ops/fourier/fft_operation.hpp>
#include <rexlib/ops/fourier/ifft_operation.hpp>
#include <rexlib/ops/fourier/irfft_operation.hpp>
#include <rexlib/ops/fourier/rfft_operation.hpp>
#include <rexlib/ops/selection/where_operation.hpp>

#include "rexlib_hip.h"

#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <deque>
#include <mutex>
#include <string>

namespace rexlib_hip_plugin
{
namespace
{

using namespace rexlib;

void hip_check(hipError_t e, const char *what)
{
	if (e != hipSuccess)
	{
		throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
	}
}

// Status of the C ABI as the exception the reference throws.
void abi_check(int status)
{
	switch (status)
	{
	case REXLIB_HIP_OK:
		return;
	case REXLIB_HIP_INVALID_ARGUMENT:
		throw std::invalid_argument(rexlib_hip_last_error());
	case REXLIB_HIP_UNSUPPORTED:
		throw unsupported_operation_error(rexlib_hip_last_error());
	default:
		throw std::runtime_error(rexlib_hip_last_error());
	}
}

// ---------------------------------------------------------------------------
// Memory
// ---------------------------------------------------------------------------

class hip_memory_resource final : public memory_resource
{
public:
	hip_memory_resource(int device, memory_resource_kind kind) : m_device(device), m_kind(kind) {}
	memory_resource_kind get_kind() const noexcept override { return m_kind; }
	std::shared_ptr<memory_allocator> create_allocator() const override;
	int get_device() const noexcept { return m_device; }

private:
	int m_device;
	memory_resource_kind m_kind;
};

class hip_buffer final : public buffer
{
public:
	hip_buffer(const hip_memory_resource &resource, std::size_t size) : m_resource(resource), m_size(size)
	{
		hip_check(hipSetDevice(resource.get_device()), "hipSetDevice");
		if (resource.get_kind() == memory_resource_kind::device_local)
		{
			hip_check(hipMalloc(&m_data, size), "hipMalloc");
		}
		else
		{
			hip_check(hipHostMalloc(&m_data, size, hipHostMallocDefault), "hipHostMalloc");
		}
	}
	~hip_buffer() override
	{
		(void)hipSetDevice(m_resource.get_device());
		if (m_resource.get_kind() == memory_resource_kind::device_local)
		{
			(void)hipFree(m_data);
		}
		else
		{
			(void)hipHostFree(m_data);
		}
	}
	void *get_host_ptr() noexcept override { return host_visible() ? m_data : nullptr; }
	const void *get_host_ptr() const noexcept override { return host_visible() ? m_data : nullptr; }
	std::size_t get_size() const noexcept override { return m_size; }
	const memory_resource &get_memory_resource() const noexcept override { return m_resource; }
	void *get_address() const noexcept { return m_data; }

private:
	bool host_visible() const noexcept { return m_resource.get_kind() != memory_resource_kind::device_local; }
	const hip_memory_resource &m_resource;
	std::size_t m_size;
	void *m_data = nullptr;
};

class hip_allocator final : public memory_allocator
{
public:
	explicit hip_allocator(const hip_memory_resource &resource) : m_resource(resource) {}
	const memory_resource &get_memory_resource() const noexcept override { return m_resource; }
	// hipMalloc and hipHostMalloc return blocks aligned to at least 256 bytes.
	std::size_t get_max_alignment() const noexcept override { return 256; }
	std::shared_ptr<buffer> allocate(std::size_t size, std::size_t alignment, command_queue *, command_token &pending)
		override
	{
		if (alignment > get_max_alignment())
		{
			throw std::invalid_argument("hip_allocator: alignment above 256 bytes");
		}
		pending = command_token();
		return std::make_shared<hip_buffer>(m_resource, std::max<std::size_t>(size, 1));
	}

private:
	const hip_memory_resource &m_resource;
};

std::shared_ptr<memory_allocator> hip_memory_resource::create_allocator() const
{
	return std::make_shared<hip_allocator>(*this);
}

// The address of an operand: a HIP buffer's own, or the host pointer of a
// host buffer of the core.
void *operand_address(const buffer &b)
{
	if (const auto *h = dynamic_cast<const hip_buffer *>(&b))
	{
		return h->get_address();
	}
	return const_cast<void *>(b.get_host_ptr());
}

// ---------------------------------------------------------------------------
// Queues
// ---------------------------------------------------------------------------

// Events of one stream, numbered from 0; completed events at the front are
// released, and ids below the first kept one are complete.
class hip_timeline final : public command_timeline
{
public:
	explicit hip_timeline(int device) : m_device(device) {}
	~hip_timeline() override
	{
		for (hipEvent_t e : m_events)
		{
			(void)hipEventDestroy(e);
		}
	}

	std::size_t record(hipStream_t stream)
	{
		hipEvent_t e = nullptr;
		hip_check(hipEventCreateWithFlags(&e, hipEventDisableTiming), "hipEventCreate");
		hip_check(hipEventRecord(e, stream), "hipEventRecord");
		std::lock_guard<std::mutex> lock(m_mutex);
		m_events.push_back(e);
		const std::size_t id = m_first + m_events.size() - 1;
		while (m_events.size() > 1 && hipEventQuery(m_events.front()) == hipSuccess)
		{
			(void)hipEventDestroy(m_events.front());
			m_events.pop_front();
			++m_first;
		}
		return id;
	}

	// The event stays locked while it is used: record() releases completed
	// events from another thread.
	void wait(std::size_t id) const override
	{
		std::lock_guard<std::mutex> lock(m_mutex);
		if (hipEvent_t e = event(id))
		{
			hip_check(hipEventSynchronize(e), "hipEventSynchronize");
		}
	}

	bool is_complete(std::size_t id) const override
	{
		std::lock_guard<std::mutex> lock(m_mutex);
		hipEvent_t e = event(id);
		return !e || hipEventQuery(e) == hipSuccess;
	}

	// Makes `stream` wait for event `id` of this timeline.
	void enqueue_wait(std::size_t id, hipStream_t stream) const
	{
		std::lock_guard<std::mutex> lock(m_mutex);
		if (hipEvent_t e = event(id))
		{
			hip_check(hipStreamWaitEvent(stream, e, 0), "hipStreamWaitEvent");
		}
	}

private:
	// Caller holds m_mutex.
	hipEvent_t event(std::size_t id) const
	{
		if (id < m_first || id >= m_first + m_events.size())
		{
			return nullptr;
		}
		return m_events[id - m_first];
	}
	int m_device;
	mutable std::mutex m_mutex;
	std::deque<hipEvent_t> m_events;
	std::size_t m_first = 0;
};

// Programs of this plugin enqueue their kernels on a stream.
class hip_program : public program
{
public:
	virtual void enqueue(span<const std::shared_ptr<buffer>> outputs,
		span<const std::shared_ptr<const buffer>> inputs, hipStream_t stream) const = 0;
};

class hip_command_queue final : public command_queue
{
public:
	explicit hip_command_queue(int device) : m_device(device), m_timeline(std::make_shared<hip_timeline>(device))
	{
		hip_check(hipSetDevice(device), "hipSetDevice");
		hip_check(hipStreamCreateWithFlags(&m_stream, hipStreamNonBlocking), "hipStreamCreate");
	}
	~hip_command_queue() override
	{
		(void)hipSetDevice(m_device);
		(void)hipStreamSynchronize(m_stream);
		m_in_flight.clear();
		(void)hipStreamDestroy(m_stream);
	}
	hipStream_t get_stream() const noexcept { return m_stream; }
	int get_device() const noexcept { return m_device; }

	command_token submit(command cmd) override
	{
		std::lock_guard<std::mutex> lock(m_mutex);
		hip_check(hipSetDevice(m_device), "hipSetDevice");
		for (const command_token &dep : cmd.get_dependencies())
		{
			if (dep.is_empty() || dep.get_timeline() == m_timeline)
			{
				continue;
			}
			if (const auto *other = dynamic_cast<const hip_timeline *>(dep.get_timeline().get()))
			{
				other->enqueue_wait(dep.get_id(), m_stream);
			}
			else
			{
				dep.wait();
			}
		}
		const auto *p = dynamic_cast<const hip_program *>(cmd.get_program().get());
		if (!p)
		{
			throw std::invalid_argument("hip_command_queue: the program was not built for a HIP queue");
		}
		try
		{
			p->enqueue(cmd.get_outputs(), cmd.get_inputs(), m_stream);
		}
		catch (...)
		{
			(void)hipStreamSynchronize(m_stream);
			throw;
		}
		const std::size_t id = m_timeline->record(m_stream);
		m_in_flight.emplace_back(id, std::move(cmd));
		while (!m_in_flight.empty() && m_timeline->is_complete(m_in_flight.front().first))
		{
			m_in_flight.pop_front();
		}
		return command_token(m_timeline, id);
	}

private:
	int m_device;
	hipStream_t m_stream = nullptr;
	std::shared_ptr<hip_timeline> m_timeline;
	std::mutex m_mutex;
	std::deque<std::pair<std::size_t, command>> m_in_flight;
};

// ---------------------------------------------------------------------------
// Devices
// ---------------------------------------------------------------------------

class hip_device final : public device
{
public:
	explicit hip_device(int id)
		: m_id(id), m_local(id, memory_resource_kind::device_local), m_staging(id, memory_resource_kind::host_staging)
	{
	}
	const memory_resource &get_memory_resource(memory_resource_affinity affinity) const override
	{
		return affinity == memory_resource_affinity::device ? static_cast<const memory_resource &>(m_local)
															  : static_cast<const memory_resource &>(m_staging);
	}
	std::shared_ptr<command_queue> create_command_queue() const override
	{
		return std::make_shared<hip_command_queue>(m_id);
	}

private:
	int m_id;
	hip_memory_resource m_local;
	hip_memory_resource m_staging;
};

class hip_device_backend final : public device_backend
{
public:
	std::string get_name() const override { return "hip"; }
	version get_version() const override { return version(0, 1, 0); }
	void enumerate_devices(std::vector<std::size_t> &ids) const override
	{
		ids.clear();
		int count = rexlib_hip_device_count();
		for (int i = 0; i < count; ++i)
		{
			ids.push_back(static_cast<std::size_t>(i));
		}
	}
	bool get_device_properties(std::size_t id, device_properties &desc) const override
	{
		if (static_cast<int>(id) >= rexlib_hip_device_count())
		{
			return false;
		}
		hipDeviceProp_t p{};
		if (hipGetDeviceProperties(&p, static_cast<int>(id)) != hipSuccess)
		{
			return false;
		}
		desc.set_type(device_type::gpu);
		desc.set_name(std::string(p.name) + " (" + p.gcnArchName + ")");
		desc.set_total_memory_bytes(p.totalGlobalMem);
		desc.set_optimal_data_alignment(256);
		return true;
	}
	std::shared_ptr<device> create_device(std::size_t id) const override
	{
		if (static_cast<int>(id) >= rexlib_hip_device_count())
		{
			throw std::invalid_argument("hip: no device " + std::to_string(id));
		}
		return std::make_shared<hip_device>(static_cast<int>(id));
	}
};

// ---------------------------------------------------------------------------
// Operand checks and conversion to the C ABI
// ---------------------------------------------------------------------------

bool on_hip_queue(command_queue &queue)
{
	return dynamic_cast<hip_command_queue *>(&queue) != nullptr;
}

// An operand that the queue's device addresses: memory of this plugin's own
// resources on that device (device-local, or the pinned host staging),
// never another backend's or another device's.
bool device_accessible(const operand_descriptor &d, const command_queue &queue)
{
	const auto *q = dynamic_cast<const hip_command_queue *>(&queue);
	const auto *r = dynamic_cast<const hip_memory_resource *>(d.get_memory_resource());
	return q != nullptr && r != nullptr && r->get_device() == q->get_device() && is_device_accessible(r->get_kind());
}

bool all_device_accessible(span<const operand_descriptor> ds, const command_queue &queue)
{
	return std::all_of(ds.begin(), ds.end(), [&](const operand_descriptor &d) { return device_accessible(d, queue); });
}

bool types_are(span<const operand_descriptor> ds, std::initializer_list<numerical_type> types)
{
	if (ds.size() != types.size())
	{
		return false;
	}
	std::size_t i = 0;
	for (numerical_type t : types)
	{
		if (ds[i++].get_data_type() != t)
		{
			return false;
		}
	}
	return true;
}

int abi_type(numerical_type t)
{
	switch (t)
	{
	case numerical_type::float32:
		return REXLIB_HIP_FLOAT32;
	case numerical_type::float64:
		return REXLIB_HIP_FLOAT64;
	case numerical_type::complex_float32:
		return REXLIB_HIP_COMPLEX64;
	case numerical_type::complex_float64:
		return REXLIB_HIP_COMPLEX128;
	case numerical_type::float16:
		return REXLIB_HIP_FLOAT16;
	case numerical_type::boolean:
		return REXLIB_HIP_BOOL;
	case numerical_type::uint64:
		return REXLIB_HIP_UINT64;
	default:
		throw unsupported_operation_error(std::string("hip: no element type ") + to_string(t));
	}
}

rexlib_hip_array abi_array(const operand_descriptor &d, const buffer &b)
{
	const strided_layout &layout = d.get_layout();
	if (layout.get_rank() > REXLIB_HIP_MAX_RANK)
	{
		throw std::invalid_argument("hip: operands hold at most 8 axes");
	}
	rexlib_hip_array a{};
	a.data = static_cast<char *>(operand_address(b)) +
		layout.get_offset() * static_cast<std::ptrdiff_t>(get_size(d.get_data_type()));
	a.dtype = abi_type(d.get_data_type());
	a.rank = static_cast<int32_t>(layout.get_rank());
	std::vector<std::size_t> extents;
	std::vector<std::ptrdiff_t> strides;
	layout.get_extents(extents);
	layout.get_strides(strides);
	for (std::size_t i = 0; i < layout.get_rank(); ++i)
	{
		a.extents[i] = static_cast<int64_t>(extents[i]);
		a.strides[i] = static_cast<int64_t>(strides[i]);
	}
	return a;
}

// Extents of an operand, and whether its layout is row-major contiguous.
std::vector<std::size_t> extents_of(const operand_descriptor &d)
{
	std::vector<std::size_t> e;
	d.get_layout().get_extents(e);
	return e;
}

bool contiguous(const strided_layout &layout)
{
	std::vector<std::size_t> extents;
	std::vector<std::ptrdiff_t> strides;
	layout.get_extents(extents);
	layout.get_strides(strides);
	std::ptrdiff_t step = 1;
	for (std::size_t i = extents.size(); i-- > 0;)
	{
		if (extents[i] != 1 && strides[i] != step)
		{
			return false;
		}
		step *= static_cast<std::ptrdiff_t>(extents[i]);
	}
	return true;
}

rexlib_hip_window abi_window(const em::gridding_window &w)
{
	rexlib_hip_window out{};
	out.profile = w.get_profile() == em::gridding_profile::box
		? REXLIB_HIP_BOX
		: (w.get_profile() == em::gridding_profile::triangle ? REXLIB_HIP_TRIANGLE : REXLIB_HIP_KAISER_BESSEL);
	out.half_width = w.get_half_width();
	out.shape = w.get_shape();
	return out;
}

// ---------------------------------------------------------------------------
// Programs and builders
// ---------------------------------------------------------------------------

// A program that remembers the operand descriptors it was built for.
class bound_program : public hip_program
{
public:
	bound_program(span<const operand_descriptor> outputs, span<const operand_descriptor> inputs)
		: m_outputs(outputs.begin(), outputs.end()), m_inputs(inputs.begin(), inputs.end())
	{
	}

protected:
	std::vector<rexlib_hip_array> arrays(span<const std::shared_ptr<buffer>> outputs,
		span<const std::shared_ptr<const buffer>> inputs, std::vector<rexlib_hip_array> &out) const
	{
		if (outputs.size() != m_outputs.size() || inputs.size() != m_inputs.size())
		{
			throw std::invalid_argument("hip: the command binds other operands than the program was built for");
		}
		std::vector<rexlib_hip_array> in;
		out.clear();
		for (std::size_t i = 0; i < inputs.size(); ++i)
		{
			in.push_back(abi_array(m_inputs[i], *inputs[i]));
		}
		for (std::size_t i = 0; i < outputs.size(); ++i)
		{
			out.push_back(abi_array(m_outputs[i], *outputs[i]));
		}
		return in;
	}
	std::vector<operand_descriptor> m_outputs;
	std::vector<operand_descriptor> m_inputs;
};

// What an insertion program keeps of its operation (operations are not
// copyable): the window, and for the CTF form the optics and pixel size.
struct insertion_parameters
{
	rexlib_hip_window window;
	rexlib_hip_ctf_optics optics;
	double pixel_size;
};

template <typename Operation>
class insertion_program final : public bound_program
{
public:
	insertion_program(const insertion_parameters &parameters, span<const operand_descriptor> outputs,
		span<const operand_descriptor> inputs)
		: bound_program(outputs, inputs), m_parameters(parameters)
	{
	}
	void enqueue(span<const std::shared_ptr<buffer>> outputs, span<const std::shared_ptr<const buffer>> inputs,
		hipStream_t stream) const override;

private:
	insertion_parameters m_parameters;
};

template <>
void insertion_program<em::insert_slices_operation>::enqueue(span<const std::shared_ptr<buffer>> outputs,
	span<const std::shared_ptr<const buffer>> inputs, hipStream_t stream) const
{
	std::vector<rexlib_hip_array> out;
	const auto in = arrays(outputs, inputs, out);
	abi_check(rexlib_hip_insert_slices(&in[0], &in[1], &in[2], &in[3], &in[4], &m_parameters.window, &out[0],
		&out[1], stream));
}

template <>
void insertion_program<em::insert_weighted_slices_operation>::enqueue(span<const std::shared_ptr<buffer>> outputs,
	span<const std::shared_ptr<const buffer>> inputs, hipStream_t stream) const
{
	std::vector<rexlib_hip_array> out;
	const auto in = arrays(outputs, inputs, out);
	abi_check(rexlib_hip_insert_weighted_slices(&in[0], &in[1], &in[2], &in[3], &in[4], &in[5], &m_parameters.window,
		&out[0], &out[1], stream));
}

template <>
void insertion_program<em::insert_ctf_slices_operation>::enqueue(span<const std::shared_ptr<buffer>> outputs,
	span<const std::shared_ptr<const buffer>> inputs, hipStream_t stream) const
{
	std::vector<rexlib_hip_array> out;
	const auto in = arrays(outputs, inputs, out);
	abi_check(rexlib_hip_insert_ctf_slices(&in[0], &in[1], &in[2], &in[3], &in[4], &in[5], &m_parameters.window,
		&m_parameters.optics, m_parameters.pixel_size, &out[0], &out[1], stream));
}

template <typename Operation>
insertion_parameters parameters_of(const Operation &op)
{
	return insertion_parameters{abi_window(op.get_window()), rexlib_hip_ctf_optics{}, 0.0};
}

template <>
insertion_parameters parameters_of(const em::insert_ctf_slices_operation &op)
{
	const em::ctf_optics &o = op.get_optics();
	return insertion_parameters{abi_window(op.get_window()),
		rexlib_hip_ctf_optics{o.get_voltage(), o.get_spherical_aberration(), o.get_amplitude_contrast()},
		op.get_pixel_size()};
}

// Input types of each insertion (accumulators, slices, then real operands).
template <typename Operation>
std::initializer_list<numerical_type> insertion_inputs();

template <>
std::initializer_list<numerical_type> insertion_inputs<em::insert_slices_operation>()
{
	static const std::initializer_list<numerical_type> t = {numerical_type::complex_float32, numerical_type::float32,
		numerical_type::complex_float32, numerical_type::float32, numerical_type::float32};
	return t;
}

template <>
std::initializer_list<numerical_type> insertion_inputs<em::insert_weighted_slices_operation>()
{
	static const std::initializer_list<numerical_type> t = {numerical_type::complex_float32, numerical_type::float32,
		numerical_type::complex_float32, numerical_type::float32, numerical_type::float32, numerical_type::float32};
	return t;
}

template <>
std::initializer_list<numerical_type> insertion_inputs<em::insert_ctf_slices_operation>()
{
	static const std::initializer_list<numerical_type> t = {numerical_type::complex_float32, numerical_type::float32,
		numerical_type::complex_float32, numerical_type::float32, numerical_type::float32, numerical_type::float32};
	return t;
}

// The insertion programs take float32 / complex64 operands in device memory
// on a HIP queue (float64 stays with the reference's CPU backend).
template <typename Operation>
class insertion_builder final : public program_builder
{
public:
	operation_id get_operation_id() const noexcept override { return operation_id::of<Operation>(); }
	backend_priority get_suitability(const operation &, span<const operand_descriptor> outputs,
		span<const operand_descriptor> inputs, command_queue &queue) const override
	{
		if (!on_hip_queue(queue) || !all_device_accessible(outputs, queue) || !all_device_accessible(inputs, queue) ||
			!types_are(outputs, {numerical_type::complex_float32, numerical_type::float32}) ||
			!types_are(inputs, insertion_inputs<Operation>()))
		{
			return backend_priority::unsupported;
		}
		return backend_priority::optimal;
	}
	std::shared_ptr<program> build(const operation &op, span<const operand_descriptor> outputs,
		span<const operand_descriptor> inputs, command_queue &, program_cache *) const override
	{
		return std::make_shared<insertion_program<Operation>>(parameters_of(operation_cast<Operation>(op)), outputs,
			inputs);
	}
};

// Lengths the fast square-image kernels take: prime factors 2, 3 and 5.
bool smooth_length(std::size_t n)
{
	if (n == 0)
	{
		return false;
	}
	for (const std::size_t r : {2u, 3u, 5u})
	{
		while (n % r == 0)
		{
			n /= r;
		}
	}
	return n == 1;
}

// Whether the (absolute) axes of an operation on `rank` axes are its last
// `count`, in order.
bool last_axes(span<const std::size_t> axes, std::size_t rank, std::size_t count)
{
	if (axes.size() != count || rank < count)
	{
		return false;
	}
	for (std::size_t i = 0; i < count; ++i)
	{
		if (axes[i] != rank - count + i)
		{
			return false;
		}
	}
	return true;
}

// Any other transform over any axes: the general programs of the C ABI.
enum class transform_kind
{
	fft,
	ifft,
	rfft,
	irfft
};

int32_t abi_normalization(ops::fourier_normalization n)
{
	switch (n)
	{
	case ops::fourier_normalization::ortho:
		return 1;
	case ops::fourier_normalization::forward:
		return 2;
	default:
		return 0;
	}
}

class general_transform_program final : public bound_program
{
public:
	general_transform_program(span<const operand_descriptor> outputs, span<const operand_descriptor> inputs,
		transform_kind kind, span<const std::size_t> axes, ops::fourier_normalization normalization)
		: bound_program(outputs, inputs), m_kind(kind), m_normalization(abi_normalization(normalization))
	{
		for (std::size_t i = 0; i < axes.size(); ++i)
		{
			m_axes.push_back(static_cast<int32_t>(axes[i]));
		}
	}
	void enqueue(span<const std::shared_ptr<buffer>> outputs, span<const std::shared_ptr<const buffer>> inputs,
		hipStream_t stream) const override
	{
		std::vector<rexlib_hip_array> out;
		const auto in = arrays(outputs, inputs, out);
		const int32_t count = static_cast<int32_t>(m_axes.size());
		switch (m_kind)
		{
		case transform_kind::fft:
		case transform_kind::ifft:
			abi_check(rexlib_hip_fft(&in[0], &out[0], m_axes.data(), count, m_kind == transform_kind::ifft ? 1 : 0,
				m_normalization, stream));
			break;
		case transform_kind::rfft:
			abi_check(rexlib_hip_rfft(&in[0], &out[0], m_axes.data(), count, m_normalization, stream));
			break;
		case transform_kind::irfft:
			abi_check(rexlib_hip_irfft(&in[0], &out[0], m_axes.data(), count, m_normalization, stream));
			break;
		}
	}

private:
	transform_kind m_kind;
	int32_t m_normalization;
	std::vector<int32_t> m_axes;
};

bool general_transform_suitable(span<const operand_descriptor> outputs, span<const operand_descriptor> inputs,
	command_queue &queue, numerical_type in_type, numerical_type out_type, span<const std::size_t> axes)
{
	return on_hip_queue(queue) && outputs.size() == 1 && inputs.size() == 1 && all_device_accessible(outputs, queue) &&
		all_device_accessible(inputs, queue) && inputs[0].get_data_type() == in_type &&
		outputs[0].get_data_type() == out_type && !axes.empty() &&
		inputs[0].get_layout().get_rank() <= REXLIB_HIP_MAX_RANK &&
		inputs[0].get_layout().get_rank() == outputs[0].get_layout().get_rank();
}

template <typename Operation, transform_kind Kind>
class complex_transform_builder final : public program_builder
{
public:
	operation_id get_operation_id() const noexcept override { return operation_id::of<Operation>(); }
	backend_priority get_suitability(const operation &op, span<const operand_descriptor> outputs,
		span<const operand_descriptor> inputs, command_queue &queue) const override
	{
		const auto &f = operation_cast<Operation>(op);
		return general_transform_suitable(outputs, inputs, queue, numerical_type::complex_float32,
				   numerical_type::complex_float32, f.get_parameters().get_axes())
			? backend_priority::optimal
			: backend_priority::unsupported;
	}
	std::shared_ptr<program> build(const operation &op, span<const operand_descriptor> outputs,
		span<const operand_descriptor> inputs, command_queue &, program_cache *) const override
	{
		const auto &f = operation_cast<Operation>(op);
		return std::make_shared<general_transform_program>(outputs, inputs, Kind, f.get_parameters().get_axes(),
			f.get_normalization());
	}
};

class rfft_program final : public bound_program
{
public:
	using bound_program::bound_program;
	void enqueue(span<const std::shared_ptr<buffer>> outputs, span<const std::shared_ptr<const buffer>> inputs,
		hipStream_t stream) const override
	{
		std::vector<rexlib_hip_array> out;
		const auto in = arrays(outputs, inputs, out);
		abi_check(rexlib_hip_rfft2(&in[0], &out[0], stream));
	}
};

class rfft_builder final : public program_builder
{
public:
	operation_id get_operation_id() const noexcept override { return operation_id::of<ops::rfft_operation>(); }
	backend_priority get_suitability(const operation &op, span<const operand_descriptor> outputs,
		span<const operand_descriptor> inputs, command_queue &queue) const override
	{
		const auto &f = operation_cast<ops::rfft_operation>(op);
		if (fast(f, outputs, inputs, queue) ||
			general_transform_suitable(outputs, inputs, queue, numerical_type::float32, numerical_type::complex_float32,
				f.get_parameters().get_axes()))
		{
			return backend_priority::optimal;
		}
		return backend_priority::unsupported;
	}
	std::shared_ptr<program> build(const operation &op, span<const operand_descriptor> outputs,
		span<const operand_descriptor> inputs, command_queue &queue, program_cache *) const override
	{
		const auto &f = operation_cast<ops::rfft_operation>(op);
		if (fast(f, outputs, inputs, queue))
		{
			return std::make_shared<rfft_program>(outputs, inputs);
		}
		return std::make_shared<general_transform_program>(outputs, inputs, transform_kind::rfft,
			f.get_parameters().get_axes(),
			f.get_normalization());
	}

private:
	// rfft2 of square images: the four-step and packed-row kernels.
	static bool fast(const ops::rfft_operation &f, span<const operand_descriptor> outputs,
		span<const operand_descriptor> inputs, command_queue &queue)
	{
		if (!on_hip_queue(queue) || outputs.size() != 1 || inputs.size() != 1 || !all_device_accessible(outputs, queue) ||
			!all_device_accessible(inputs, queue) || f.get_normalization() != ops::fourier_normalization::backward ||
			!last_axes(f.get_parameters().get_axes(), inputs[0].get_layout().get_rank(), 2) ||
			outputs[0].get_data_type() != numerical_type::complex_float32 ||
			(inputs[0].get_data_type() != numerical_type::float32 &&
				inputs[0].get_data_type() != numerical_type::float16))
		{
			return false;
		}
		const std::vector<std::size_t> e = extents_of(inputs[0]);
		return e[e.size() - 1] == e[e.size() - 2] && smooth_length(e.back());
	}
};

class irfft_program final : public bound_program
{
public:
	irfft_program(span<const operand_descriptor> outputs, span<const operand_descriptor> inputs, int dims)
		: bound_program(outputs, inputs), m_dims(dims)
	{
	}
	void enqueue(span<const std::shared_ptr<buffer>> outputs, span<const std::shared_ptr<const buffer>> inputs,
		hipStream_t stream) const override
	{
		std::vector<rexlib_hip_array> out;
		const auto in = arrays(outputs, inputs, out);
		abi_check(m_dims == 3 ? rexlib_hip_irfft3(&in[0], &out[0], stream)
							  : rexlib_hip_irfft2(&in[0], &out[0], stream));
	}

private:
	int m_dims;
};

class irfft_builder final : public program_builder
{
public:
	operation_id get_operation_id() const noexcept override { return operation_id::of<ops::irfft_operation>(); }
	backend_priority get_suitability(const operation &op, span<const operand_descriptor> outputs,
		span<const operand_descriptor> inputs, command_queue &queue) const override
	{
		const auto &f = operation_cast<ops::irfft_operation>(op);
		if (fast(f, outputs, inputs, queue) ||
			general_transform_suitable(outputs, inputs, queue, numerical_type::complex_float32, numerical_type::float32,
				f.get_parameters().get_axes()))
		{
			return backend_priority::optimal;
		}
		return backend_priority::unsupported;
	}
	std::shared_ptr<program> build(const operation &op, span<const operand_descriptor> outputs,
		span<const operand_descriptor> inputs, command_queue &queue, program_cache *) const override
	{
		const auto &f = operation_cast<ops::irfft_operation>(op);
		if (fast(f, outputs, inputs, queue))
		{
			return std::make_shared<irfft_program>(outputs, inputs, static_cast<int>(f.get_parameters().get_axes().size()));
		}
		return std::make_shared<general_transform_program>(outputs, inputs, transform_kind::irfft,
			f.get_parameters().get_axes(), f.get_normalization());
	}

private:
	// irfft2 / irfft3 of square or cubic grids of a smooth extent.
	static bool fast(const ops::irfft_operation &f, span<const operand_descriptor> outputs,
		span<const operand_descriptor> inputs, command_queue &queue)
	{
		if (!on_hip_queue(queue) || outputs.size() != 1 || inputs.size() != 1 || !all_device_accessible(outputs, queue) ||
			!all_device_accessible(inputs, queue) || f.get_normalization() != ops::fourier_normalization::backward ||
			outputs[0].get_data_type() != numerical_type::float32 ||
			inputs[0].get_data_type() != numerical_type::complex_float32)
		{
			return false;
		}
		const std::size_t rank = inputs[0].get_layout().get_rank();
		const span<const std::size_t> axes = f.get_parameters().get_axes();
		const std::size_t dims = last_axes(axes, rank, 2) ? 2 : (last_axes(axes, rank, 3) ? 3 : 0);
		if (dims == 0)
		{
			return false;
		}
		const std::vector<std::size_t> e = extents_of(outputs[0]);
		for (std::size_t a = rank - dims; a < rank; ++a)
		{
			if (e[a] != e.back())
			{
				return false;
			}
		}
		return smooth_length(e.back());
	}
};

// Fill of an operand in device memory with the operation's value, by a fill
// kernel (not hipMemsetAsync, so that a captured graph holds kernel nodes
// only), any strides.
// The fill value as real and imaginary parts; false for a scalar type the
// fill program does not take.
bool fill_value_parts(const scalar_value &v, double &re, double &im)
{
	re = 0.0;
	im = 0.0;
	if (const float *x = v.get_if<float>())
	{
		re = *x;
	}
	else if (const double *x = v.get_if<double>())
	{
		re = *x;
	}
	else if (const std::complex<float> *x = v.get_if<std::complex<float>>())
	{
		re = x->real();
		im = x->imag();
	}
	else if (const std::complex<double> *x = v.get_if<std::complex<double>>())
	{
		re = x->real();
		im = x->imag();
	}
	else if (const bool *x = v.get_if<bool>())
	{
		re = *x ? 1.0 : 0.0;
	}
	else if (const std::int32_t *x = v.get_if<std::int32_t>())
	{
		re = static_cast<double>(*x);
	}
	else if (const std::int64_t *x = v.get_if<std::int64_t>())
	{
		re = static_cast<double>(*x);
	}
	else
	{
		return false;
	}
	return true;
}

class fill_program final : public bound_program
{
public:
	fill_program(span<const operand_descriptor> outputs, span<const operand_descriptor> inputs, double re, double im)
		: bound_program(outputs, inputs), m_re(re), m_im(im)
	{
	}
	void enqueue(span<const std::shared_ptr<buffer>> outputs, span<const std::shared_ptr<const buffer>> inputs,
		hipStream_t stream) const override
	{
		std::vector<rexlib_hip_array> out;
		arrays(outputs, inputs, out);
		abi_check(rexlib_hip_fill(&out[0], m_re, m_im, stream));
	}

private:
	double m_re;
	double m_im;
};

bool fill_type(numerical_type t)
{
	return t == numerical_type::float32 || t == numerical_type::float64 || t == numerical_type::complex_float32 ||
		t == numerical_type::complex_float64 || t == numerical_type::float16 || t == numerical_type::boolean;
}

class fill_builder final : public program_builder
{
public:
	operation_id get_operation_id() const noexcept override { return operation_id::of<ops::fill_operation>(); }
	backend_priority get_suitability(const operation &op, span<const operand_descriptor> outputs,
		span<const operand_descriptor> inputs, command_queue &queue) const override
	{
		double re = 0.0, im = 0.0;
		if (!on_hip_queue(queue) || outputs.size() != 1 || !inputs.empty() || !all_device_accessible(outputs, queue) ||
			!fill_type(outputs[0].get_data_type()) || outputs[0].get_layout().get_rank() > REXLIB_HIP_MAX_RANK ||
			!fill_value_parts(operation_cast<ops::fill_operation>(op).get_fill_value(), re, im))
		{
			return backend_priority::unsupported;
		}
		return backend_priority::normal;
	}
	std::shared_ptr<program> build(const operation &op, span<const operand_descriptor> outputs,
		span<const operand_descriptor> inputs, command_queue &, program_cache *) const override
	{
		double re = 0.0, im = 0.0;
		if (!fill_value_parts(operation_cast<ops::fill_operation>(op).get_fill_value(), re, im))
		{
			throw unsupported_operation_error("hip: fill values of this scalar type are not supported");
		}
		return std::make_shared<fill_program>(outputs, inputs, re, im);
	}
};

// divide, greater and where through the elementwise programs of the C ABI:
// float32 and complex64 operands (float32 for greater), boolean conditions
// and comparisons, broadcasting, any strides.
enum class elementwise_kind
{
	divide,
	greater,
	where
};

class elementwise_program final : public bound_program
{
public:
	elementwise_program(span<const operand_descriptor> outputs, span<const operand_descriptor> inputs,
		elementwise_kind kind)
		: bound_program(outputs, inputs), m_kind(kind)
	{
	}
	void enqueue(span<const std::shared_ptr<buffer>> outputs, span<const std::shared_ptr<const buffer>> inputs,
		hipStream_t stream) const override
	{
		std::vector<rexlib_hip_array> out;
		const auto in = arrays(outputs, inputs, out);
		switch (m_kind)
		{
		case elementwise_kind::divide:
			abi_check(rexlib_hip_divide(&in[0], &in[1], &out[0], stream));
			break;
		case elementwise_kind::greater:
			abi_check(rexlib_hip_greater(&in[0], &in[1], &out[0], stream));
			break;
		case elementwise_kind::where:
			abi_check(rexlib_hip_where(&in[0], &in[1], &in[2], &out[0], stream));
			break;
		}
	}

private:
	elementwise_kind m_kind;
};

template <typename Operation, elementwise_kind Kind>
class elementwise_builder final : public program_builder
{
public:
	operation_id get_operation_id() const noexcept override { return operation_id::of<Operation>(); }
	backend_priority get_suitability(const operation &, span<const operand_descriptor> outputs,
		span<const operand_descriptor> inputs, command_queue &queue) const override
	{
		const std::size_t arity = Kind == elementwise_kind::where ? 3 : 2;
		if (!on_hip_queue(queue) || outputs.size() != 1 || inputs.size() != arity || !all_device_accessible(outputs, queue) ||
			!all_device_accessible(inputs, queue))
		{
			return backend_priority::unsupported;
		}
		for (const auto &d : inputs)
		{
			if (d.get_layout().get_rank() > REXLIB_HIP_MAX_RANK)
			{
				return backend_priority::unsupported;
			}
		}
		const numerical_type value = inputs[Kind == elementwise_kind::where ? 1 : 0].get_data_type();
		const bool real = value == numerical_type::float32;
		const bool ok = Kind == elementwise_kind::greater
			? real && inputs[1].get_data_type() == value && outputs[0].get_data_type() == numerical_type::boolean
			: (real || value == numerical_type::complex_float32) && outputs[0].get_data_type() == value &&
				inputs[arity - 1].get_data_type() == value &&
				(Kind != elementwise_kind::where || inputs[0].get_data_type() == numerical_type::boolean);
		return ok ? backend_priority::optimal : backend_priority::unsupported;
	}
	std::shared_ptr<program> build(const operation &, span<const operand_descriptor> outputs,
		span<const operand_descriptor> inputs, command_queue &, program_cache *) const override
	{
		return std::make_shared<elementwise_program>(outputs, inputs, Kind);
	}
};

// Copies between contiguous operands of one type and extents, host or device.
class copy_program final : public bound_program
{
public:
	using bound_program::bound_program;
	void enqueue(span<const std::shared_ptr<buffer>> outputs, span<const std::shared_ptr<const buffer>> inputs,
		hipStream_t stream) const override
	{
		const std::size_t element = get_size(m_outputs[0].get_data_type());
		char *dst = static_cast<char *>(operand_address(*outputs[0])) +
			m_outputs[0].get_layout().get_offset() * static_cast<std::ptrdiff_t>(element);
		const char *src = static_cast<const char *>(operand_address(*inputs[0])) +
			m_inputs[0].get_layout().get_offset() * static_cast<std::ptrdiff_t>(element);
		hip_check(hipMemcpyAsync(dst, src, m_outputs[0].get_layout().compute_element_count() * element, hipMemcpyDefault,
					  stream),
			"hipMemcpyAsync");
	}
};

class copy_builder final : public program_builder
{
public:
	operation_id get_operation_id() const noexcept override { return operation_id::of<ops::copy_operation>(); }
	backend_priority get_suitability(const operation &, span<const operand_descriptor> outputs,
		span<const operand_descriptor> inputs, command_queue &queue) const override
	{
		if (!on_hip_queue(queue) || outputs.size() != 1 || inputs.size() != 1 ||
			outputs[0].get_memory_resource() == nullptr || inputs[0].get_memory_resource() == nullptr ||
			(!device_accessible(outputs[0], queue) && !device_accessible(inputs[0], queue)) ||
			outputs[0].get_data_type() != inputs[0].get_data_type() ||
			extents_of(outputs[0]) != extents_of(inputs[0]) || !contiguous(outputs[0].get_layout()) ||
			!contiguous(inputs[0].get_layout()))
		{
			return backend_priority::unsupported;
		}
		return backend_priority::normal;
	}
	std::shared_ptr<program> build(const operation &, span<const operand_descriptor> outputs,
		span<const operand_descriptor> inputs, command_queue &, program_cache *) const override
	{
		return std::make_shared<copy_program>(outputs, inputs);
	}
};

// Background standardization over any axes of float32 / float64 images.
class standardize_program final : public bound_program
{
public:
	standardize_program(span<const operand_descriptor> outputs, span<const operand_descriptor> inputs, uint32_t axes,
		double radius, double lower_bound, double upper_bound)
		: bound_program(outputs, inputs), m_axes(axes), m_radius(radius), m_lower_bound(lower_bound),
		  m_upper_bound(upper_bound)
	{
	}
	void enqueue(span<const std::shared_ptr<buffer>> outputs, span<const std::shared_ptr<const buffer>> inputs,
		hipStream_t stream) const override
	{
		std::vector<rexlib_hip_array> out;
		const auto in = arrays(outputs, inputs, out);
		abi_check(rexlib_hip_standardize_background(&in[0], &in[1], m_axes, m_radius, m_lower_bound, m_upper_bound,
			&out[0], stream));
	}

private:
	uint32_t m_axes;
	double m_radius;
	double m_lower_bound;
	double m_upper_bound;
};

class standardize_builder final : public program_builder
{
public:
	operation_id get_operation_id() const noexcept override
	{
		return operation_id::of<em::standardize_background_operation>();
	}
	backend_priority get_suitability(const operation &op, span<const operand_descriptor> outputs,
		span<const operand_descriptor> inputs, command_queue &queue) const override
	{
		const auto &s = operation_cast<em::standardize_background_operation>(op);
		if (!on_hip_queue(queue) || outputs.size() != 1 || inputs.size() != 2 || !all_device_accessible(outputs, queue) ||
			!all_device_accessible(inputs, queue))
		{
			return backend_priority::unsupported;
		}
		const numerical_type t = inputs[0].get_data_type();
		const std::size_t rank = inputs[0].get_layout().get_rank();
		if ((t != numerical_type::float32 && t != numerical_type::float64) || outputs[0].get_data_type() != t ||
			inputs[1].get_data_type() != numerical_type::uint64 || rank > REXLIB_HIP_MAX_RANK ||
			axis_mask(s.get_parameters().get_axes(), rank) == 0)
		{
			return backend_priority::unsupported;
		}
		return backend_priority::optimal;
	}
	std::shared_ptr<program> build(const operation &op, span<const operand_descriptor> outputs,
		span<const operand_descriptor> inputs, command_queue &, program_cache *) const override
	{
		const auto &s = operation_cast<em::standardize_background_operation>(op);
		return std::make_shared<standardize_program>(outputs, inputs,
			axis_mask(s.get_parameters().get_axes(), inputs[0].get_layout().get_rank()), s.get_radius(),
			s.get_lower_bound(), s.get_upper_bound());
	}

private:
	// The axes as a bit set (0 for an axis beyond the rank).
	static uint32_t axis_mask(span<const std::size_t> axes, std::size_t rank)
	{
		uint32_t mask = 0;
		for (std::size_t i = 0; i < axes.size(); ++i)
		{
			if (axes[i] >= rank)
			{
				return 0;
			}
			mask |= 1u << axes[i];
		}
		return mask;
	}
};

// ---------------------------------------------------------------------------
// The plugin
// ---------------------------------------------------------------------------

class hip_plugin final : public plugin
{
public:
	const std::string &get_name() const noexcept override
	{
		static const std::string name("rexlib-hip");
		return name;
	}
	version get_version() const noexcept override { return version(0, 1, 0); }
	void register_at(service_catalog &catalog) const override
	{
		catalog.get_service<device_service>()->register_backend(std::make_unique<hip_device_backend>());
		auto programs = catalog.get_service<program_service>();
		programs->register_builder(std::make_unique<insertion_builder<em::insert_slices_operation>>());
		programs->register_builder(std::make_unique<insertion_builder<em::insert_weighted_slices_operation>>());
		programs->register_builder(std::make_unique<insertion_builder<em::insert_ctf_slices_operation>>());
		programs->register_builder(std::make_unique<rfft_builder>());
		programs->register_builder(
			std::make_unique<complex_transform_builder<ops::fft_operation, transform_kind::fft>>());
		programs->register_builder(
			std::make_unique<complex_transform_builder<ops::ifft_operation, transform_kind::ifft>>());
		programs->register_builder(std::make_unique<irfft_builder>());
		programs->register_builder(std::make_unique<fill_builder>());
		programs->register_builder(std::make_unique<copy_builder>());
		programs->register_builder(
			std::make_unique<elementwise_builder<ops::divide_operation, elementwise_kind::divide>>());
		programs->register_builder(
			std::make_unique<elementwise_builder<ops::greater_operation, elementwise_kind::greater>>());
		programs->register_builder(
			std::make_unique<elementwise_builder<ops::where_operation, elementwise_kind::where>>());
		programs->register_builder(std::make_unique<standardize_builder>());
	}
};

} // namespace
} // namespace rexlib_hip_plugin

REXLIB_DEFINE_PLUGIN(rexlib_hip_plugin::hip_plugin);
