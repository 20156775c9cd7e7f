// Arrays, execution contexts and the eager dispatcher, with the names and
// meaning of rexlib 0.1.0:
//   array / const_array_ref    include/rexlib/core/ndarray/array.hpp, const_array_ref.hpp
//   device_context             include/rexlib/core/hardware/device_context.hpp
//   execution_context          include/rexlib/core/dispatch/execution_context.hpp
//   dispatcher                 include/rexlib/core/dispatch/dispatcher.hpp
//   eager dispatch             src/core/dispatch/eager_dispatcher.cpp:126-176 (allocation of
//                              missing outputs), 273-313 and 433-448 (hazards and tokens)
//
// An array is a buffer, a layout and a type, plus the hazards of the work
// queued on it: the token of its last write and the tokens of the reads
// since. The dispatcher makes every command wait for the hazards of its
// operands (reads wait for the last write; writes wait for the last write
// and every read since), which is what rexlib's access_hazard_tracker
// provides. Host access (to_host, wait) waits for them.
#pragma once

#include "core.hpp"

namespace rexlib
{

class REXLIB_API array
{
public:
	array() = default;
	array(std::shared_ptr<buffer> storage, array_descriptor descriptor);

	bool is_empty() const noexcept { return !m_buffer; }
	const std::shared_ptr<buffer> &get_buffer() const noexcept { return m_buffer; }
	const array_descriptor &get_descriptor() const noexcept { return m_descriptor; }
	const strided_layout &get_layout() const noexcept { return m_descriptor.get_layout(); }
	numerical_type get_data_type() const noexcept { return m_descriptor.get_data_type(); }
	std::vector<std::size_t> get_extents() const
	{
		std::vector<std::size_t> e;
		get_layout().get_extents(e);
		return e;
	}
	const memory_resource *get_memory_resource() const noexcept
	{
		return m_buffer ? &m_buffer->get_memory_resource() : nullptr;
	}
	operand_descriptor get_operand_descriptor() const
	{
		return operand_descriptor(m_descriptor, get_memory_resource());
	}
	// Another layout over the same buffer, sharing the hazards (a view).
	array view(array_descriptor descriptor) const
	{
		array v = *this;
		v.m_descriptor = std::move(descriptor);
		return v;
	}

	// Hazards (shared by the copies of an array).
	std::vector<command_token> get_write_hazards() const;
	command_token get_read_hazard() const;
	void record_write(const command_token &token) const;
	void record_read(const command_token &token) const;
	// Waits for all the work queued on the array.
	void wait() const;

private:
	struct hazards
	{
		command_token last_write;
		std::vector<command_token> reads;
	};
	std::shared_ptr<buffer> m_buffer;
	array_descriptor m_descriptor;
	std::shared_ptr<hazards> m_hazards = std::make_shared<hazards>();
};

class const_array_ref
{
public:
	const_array_ref(const array &a) noexcept : m_array(&a) {}
	const array &get() const noexcept { return *m_array; }
	const array *operator->() const noexcept { return m_array; }

private:
	const array *m_array;
};

class REXLIB_API device_context
{
public:
	device_context() noexcept = default;
	explicit device_context(std::shared_ptr<const device_session> session);
	const std::shared_ptr<const device_session> &get_device_session() const noexcept { return m_session; }
	const std::shared_ptr<command_queue> &get_active_queue() const noexcept { return m_active_queue; }
	const std::shared_ptr<memory_allocator> &get_allocator(memory_resource_affinity affinity) const noexcept
	{
		return m_allocators[static_cast<std::size_t>(affinity)];
	}
	device_context on_queue(std::shared_ptr<command_queue> queue) const;

private:
	std::shared_ptr<const device_session> m_session;
	std::shared_ptr<command_queue> m_active_queue;
	std::shared_ptr<memory_allocator> m_allocators[static_cast<std::size_t>(memory_resource_affinity::count)];
};

class REXLIB_API dispatcher
{
public:
	dispatcher() noexcept;
	dispatcher(const dispatcher &) = delete;
	dispatcher &operator=(const dispatcher &) = delete;
	virtual ~dispatcher();
	// Outputs that are empty arrays are allocated with the layout and type
	// in `planned_outputs` (same order) from the device's allocator.
	virtual void dispatch(const operation &operation, span<array> output_operands,
		span<const array_descriptor> planned_outputs, span<const const_array_ref> input_operands,
		const device_context &device) = 0;
};

REXLIB_API std::shared_ptr<dispatcher> make_eager_dispatcher(std::shared_ptr<const program_service> programs);

class REXLIB_API execution_context
{
public:
	execution_context() noexcept = default;
	execution_context(device_context device_context, std::shared_ptr<dispatcher> dispatcher);
	const device_context &get_device_context() const noexcept { return m_device_context; }
	const std::shared_ptr<dispatcher> &get_dispatcher() const noexcept { return m_dispatcher; }
	execution_context with_device_context(device_context context) const;

private:
	device_context m_device_context;
	std::shared_ptr<dispatcher> m_dispatcher;
};

// A context on `index` of the catalog's devices with the eager dispatcher.
REXLIB_API execution_context make_execution_context(service_catalog &catalog, const device_index &index);

namespace ops
{

// Sets every element to a value (src/functional/creation.cpp:301-349 dispatches it for zeros).
class REXLIB_API fill_operation final : public basic_operation<fill_operation>
{
public:
	explicit fill_operation(const scalar_value &fill_value) noexcept : m_fill_value(fill_value) {}
	const scalar_value &get_fill_value() const noexcept { return m_fill_value; }
	const char *get_name() const noexcept override { return "fill"; }

private:
	scalar_value m_fill_value;
};

// Copies one operand into another of the same extents and type (src/functional/transfer.cpp:60-78).
class REXLIB_API copy_operation final : public basic_operation<copy_operation>
{
public:
	const char *get_name() const noexcept override { return "copy"; }
};

// result = dividend / divisor, elementwise with broadcasting, one element
// type (include/rexlib/ops/arithmetic/divide_operation.hpp).
class REXLIB_API divide_operation final : public basic_operation<divide_operation>
{
public:
	const char *get_name() const noexcept override { return "divide"; }
};

// result = x > y, boolean (include/rexlib/ops/compare/greater_operation.hpp).
class REXLIB_API greater_operation final : public basic_operation<greater_operation>
{
public:
	const char *get_name() const noexcept override { return "greater"; }
};

// result = condition ? x : y (include/rexlib/ops/selection/where_operation.hpp).
class REXLIB_API where_operation final : public basic_operation<where_operation>
{
public:
	const char *get_name() const noexcept override { return "where"; }
};

} // namespace ops

// Host arrays, in the core's host memory.
REXLIB_API array make_host_array(std::vector<std::size_t> extents, numerical_type type);
REXLIB_API array zeros(std::vector<std::size_t> extents, numerical_type type, const execution_context &context);
REXLIB_API array to_device(const_array_ref x, const execution_context &context, array *out = nullptr);
// Copies to a new host array and waits for the copy.
REXLIB_API array to_host(const_array_ref x, const execution_context &context);

// Elementwise functions (include/rexlib/functional/arithmetic.hpp:125-130,
// compare.hpp:101-106, selection.hpp:27-33, creation.hpp): operands broadcast
// against each other; the result takes the broadcast extents.
REXLIB_API array full(std::vector<std::size_t> extents, double value, numerical_type type,
	const execution_context &context);
REXLIB_API array divide(const_array_ref x, const_array_ref y, const execution_context &context, array *out = nullptr);
REXLIB_API array greater(const_array_ref x, const_array_ref y, const execution_context &context, array *out = nullptr);
REXLIB_API array where(const_array_ref condition, const_array_ref x, const_array_ref y,
	const execution_context &context, array *out = nullptr);

} // namespace rexlib
