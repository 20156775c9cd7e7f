// The insertion entry point that allocates nothing: its workspace is a
// fixed device array of this library, for callers that cannot allocate
// (code under graph capture, serving loops that forbid hipMalloc).
#include "launch.hpp"

#include <map>
#include <mutex>

// 128 MiB per device, held from the library's load until the process ends.
// Host code never uses this name as an address: kernels and copies get the
// address that hipGetSymbolAddress returns for the current device.
static constexpr std::size_t scratch_bytes = std::size_t(128) << 20;
static __device__ __align__(256) unsigned char scratch[scratch_bytes];

namespace rexlib_hip
{

namespace
{

std::mutex scratch_mutex;
std::map<int, char *> scratch_addresses;

// The scratch array's address on the current device, looked up once per
// device with that device current.
char *device_scratch()
{
	int device = 0;
	REXLIB_HIP_CHECK(hipGetDevice(&device));
	std::lock_guard<std::mutex> lock(scratch_mutex);
	const auto found = scratch_addresses.find(device);
	if (found != scratch_addresses.end())
	{
		return found->second;
	}
	void *address = nullptr;
	REXLIB_HIP_CHECK(hipGetSymbolAddress(&address, HIP_SYMBOL(scratch)));
	if (address == nullptr)
	{
		fail(REXLIB_HIP_RUNTIME_ERROR, "insert_ctf_slices_no_alloc: no device address for the scratch array");
	}
	scratch_addresses[device] = static_cast<char *>(address);
	return static_cast<char *>(address);
}

} // anonymous namespace

} // namespace rexlib_hip

using namespace rexlib_hip;

extern "C" {

int rexlib_hip_insert_ctf_slices_no_alloc(const rexlib_hip_array *numerator, const rexlib_hip_array *denominator,
	const rexlib_hip_array *slices, const rexlib_hip_array *transforms, const rexlib_hip_array *ctf,
	const rexlib_hip_array *radial_weights, const rexlib_hip_window *window, const rexlib_hip_ctf_optics *optics,
	double pixel_size, const rexlib_hip_array *numerator_out, const rexlib_hip_array *denominator_out, void *stream)
{
	return guarded([&]
	{
		workspace_source workspace;
		workspace.arena = device_scratch();
		workspace.arena_bytes = scratch_bytes;
		insert_ctf_slices_within(numerator, denominator, slices, transforms, ctf, radial_weights, window, optics,
			pixel_size, numerator_out, denominator_out, workspace, stream);
	});
}

} // extern "C"
