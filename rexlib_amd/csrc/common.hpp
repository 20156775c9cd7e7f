// Shared host and device helpers of the gfx950 backend.
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/rexlib_hip.h"

#include <cstdint>
#include <stdexcept>
#include <string>

namespace rexlib_hip
{

// Errors carry the C ABI status they map to.
struct error : std::runtime_error
{
	int status;
	error(int s, const std::string &message) : std::runtime_error(message), status(s) {}
};

[[noreturn]] inline void fail(int status, const std::string &message)
{
	throw error(status, message);
}

inline void check_hip(hipError_t e, const char *what)
{
	if (e != hipSuccess)
	{
		fail(REXLIB_HIP_RUNTIME_ERROR, std::string(what) + ": " + hipGetErrorString(e));
	}
}

#define REXLIB_HIP_CHECK(expr) ::rexlib_hip::check_hip((expr), #expr)

void set_last_error(const std::string &message);

// Runs `body`, turning exceptions into a status and the thread's last error.
template <typename F>
int guarded(F &&body)
{
	try
	{
		body();
		set_last_error("");
		return REXLIB_HIP_OK;
	}
	catch (const error &e)
	{
		set_last_error(e.what());
		return e.status;
	}
	catch (const std::exception &e)
	{
		set_last_error(e.what());
		return REXLIB_HIP_RUNTIME_ERROR;
	}
	catch (...)
	{
		set_last_error("unknown error");
		return REXLIB_HIP_RUNTIME_ERROR;
	}
}

inline std::size_t align_up(std::size_t value, std::size_t alignment)
{
	return (value + alignment - 1) / alignment * alignment;
}

// ------------------------------------------------------------------------
// Device-side helpers
// ------------------------------------------------------------------------

struct cfloat
{
	float x, y;
};

__host__ __device__ inline float2 cmul(float2 a, float2 b)
{
	return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

__host__ __device__ inline float2 cadd(float2 a, float2 b)
{
	return make_float2(a.x + b.x, a.y + b.y);
}

__host__ __device__ inline float2 csub(float2 a, float2 b)
{
	return make_float2(a.x - b.x, a.y - b.y);
}

__host__ __device__ inline float2 cconj(float2 a)
{
	return make_float2(a.x, -a.y);
}

__host__ __device__ inline float2 cscale(float2 a, float s)
{
	return make_float2(a.x * s, a.y * s);
}

} // namespace rexlib_hip
