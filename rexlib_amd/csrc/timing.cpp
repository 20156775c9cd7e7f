// Event timing of single kernels, for measurement: when switched on, the
// launchers bracket the kernels they name with a pair of HIP events recorded
// on the launch stream, and the totals are read back on request. Off by
// default; recording is skipped while the stream is being captured into a
// graph.
#include "launch.hpp"

#include <atomic>
#include <mutex>
#include <vector>

namespace rexlib_hip
{

namespace
{

struct timed_launch
{
	int kernel;
	hipEvent_t start;
	hipEvent_t end;
};

std::mutex timing_mutex;
std::atomic<bool> timing_on{false};
std::vector<timed_launch> timing_records;

void release_records()
{
	for (const timed_launch &r : timing_records)
	{
		(void)hipEventDestroy(r.start);
		(void)hipEventDestroy(r.end);
	}
	timing_records.clear();
}

} // namespace

void set_kernel_timing(bool enable)
{
	std::lock_guard<std::mutex> lock(timing_mutex);
	release_records();
	timing_on.store(enable);
}

int timing_begin(int kernel, hipStream_t stream)
{
	if (!timing_on.load())
	{
		return -1;
	}
	hipStreamCaptureStatus status = hipStreamCaptureStatusNone;
	REXLIB_HIP_CHECK(hipStreamIsCapturing(stream, &status));
	if (status != hipStreamCaptureStatusNone)
	{
		return -1;
	}
	timed_launch r{kernel, nullptr, nullptr};
	REXLIB_HIP_CHECK(hipEventCreate(&r.start));
	REXLIB_HIP_CHECK(hipEventCreate(&r.end));
	REXLIB_HIP_CHECK(hipEventRecord(r.start, stream));
	std::lock_guard<std::mutex> lock(timing_mutex);
	timing_records.push_back(r);
	return static_cast<int>(timing_records.size()) - 1;
}

void timing_end(int token, hipStream_t stream)
{
	if (token < 0)
	{
		return;
	}
	std::lock_guard<std::mutex> lock(timing_mutex);
	if (token < static_cast<int>(timing_records.size()))
	{
		REXLIB_HIP_CHECK(hipEventRecord(timing_records[token].end, stream));
	}
}

void kernel_timing(int kernel, double *total_ms, int64_t *launches)
{
	std::lock_guard<std::mutex> lock(timing_mutex);
	double total = 0.0;
	int64_t count = 0;
	for (const timed_launch &r : timing_records)
	{
		if (r.kernel != kernel)
		{
			continue;
		}
		REXLIB_HIP_CHECK(hipEventSynchronize(r.end));
		float ms = 0.0f;
		REXLIB_HIP_CHECK(hipEventElapsedTime(&ms, r.start, r.end));
		total += ms;
		++count;
	}
	*total_ms = total;
	*launches = count;
}

} // namespace rexlib_hip
