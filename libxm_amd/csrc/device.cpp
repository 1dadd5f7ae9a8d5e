/*
 * device.cpp - the device context: one non-blocking HIP stream at the
 * default priority for all library work, a second one (same priority) for
 * copies that stage host-resident blocks into HBM, reusable device scratch
 * buffers (descriptor arrays, operand panels), pinned staging for
 * uploads and pinned, device-mapped buffers for file-resident blocks.
 */
#include "internal.h"

#include <cstdlib>

static std::once_flag init_once;
static hipStream_t stream, side;
static int device_id;

static void
init_device(void)
{
	int n = 0;
	if (hipGetDeviceCount(&n) != hipSuccess || n == 0)
		XM_FATAL("libxm", "no HIP device available");
	HIP_OK("libxm", hipGetDevice(&device_id));
	HIP_OK("libxm", hipStreamCreateWithFlags(&stream,
	    hipStreamNonBlocking));
	HIP_OK("libxm", hipStreamCreateWithFlags(&side,
	    hipStreamNonBlocking));
}

hipStream_t
xm_stream(void)
{
	std::call_once(init_once, init_device);
	return stream;
}

hipStream_t
xm_side_stream(void)
{
	std::call_once(init_once, init_device);
	return side;
}

struct scratch_slot {
	void *dev = nullptr;
	size_t bytes = 0;
	void *host = nullptr;	/* pinned staging for upload() */
	size_t host_bytes = 0;
};

static const int NSLOTS = 32;
static scratch_slot slots[NSLOTS];
static std::mutex scratch_mu;

void *
scratch_buffer(int slot, size_t bytes)
{
	std::lock_guard<std::mutex> g(scratch_mu);
	scratch_slot &s = slots[slot];
	if (bytes == 0)
		bytes = 1;
	if (s.bytes < bytes) {
		HIP_OK("libxm", hipStreamSynchronize(xm_stream()));
		if (s.dev)
			HIP_OK("libxm", hipFree(s.dev));
		size_t nb = std::max(bytes, s.bytes + s.bytes / 2);
		nb = (nb + 4095) & ~(size_t)4095;
		HIP_OK("libxm", hipMalloc(&s.dev, nb));
		s.bytes = nb;
	}
	return s.dev;
}

void *
upload(int slot, const void *host, size_t bytes)
{
	void *d = scratch_buffer(slot, bytes);
	if (bytes == 0)
		return d;
	scratch_slot &s = slots[slot];
	{
		std::lock_guard<std::mutex> g(scratch_mu);
		if (s.host_bytes < bytes) {
			HIP_OK("libxm", hipStreamSynchronize(xm_stream()));
			if (s.host)
				HIP_OK("libxm", hipHostFree(s.host));
			size_t nb = std::max(bytes, s.host_bytes * 2);
			HIP_OK("libxm", hipHostMalloc(&s.host, nb,
			    hipHostMallocDefault));
			s.host_bytes = nb;
		}
	}
	/* the staging copy may still be read by an earlier async copy */
	HIP_OK("libxm", hipStreamSynchronize(xm_stream()));
	memcpy(s.host, host, bytes);
	HIP_OK("libxm", hipMemcpyAsync(d, s.host, bytes, hipMemcpyHostToDevice,
	    xm_stream()));
	return d;
}

static struct {
	void *p = nullptr;
	size_t bytes = 0;
} pinned[NSLOTS];

void *
pinned_buffer(int slot, size_t bytes)
{
	std::lock_guard<std::mutex> g(scratch_mu);
	auto &s = pinned[slot];
	if (s.bytes < std::max<size_t>(bytes, 1)) {
		if (s.p)
			HIP_OK("libxm", hipHostFree(s.p));
		size_t nb = std::max(bytes, s.bytes + s.bytes / 2);
		nb = (nb + 4095) & ~(size_t)4095;
		HIP_OK("libxm", hipHostMalloc(&s.p, nb, hipHostMallocMapped));
		s.bytes = nb;
	}
	return s.p;
}

double
elapsed_ms(hipEvent_t a, hipEvent_t b)
{
	float ms = 0;
	HIP_OK("libxm", hipEventElapsedTime(&ms, a, b));
	return ms;
}

/*
 * Stage timestamps of graph-launched plans (events recorded inside a graph
 * cannot be timed): pinned host slots the stamp kernel writes the device
 * wall clock into, and that clock's rate.  Calls are synchronous, so every
 * plan's graph can share the slots.
 */
uint64_t *
stamp_slots(void)
{
	static uint64_t *slots = nullptr;
	if (slots == nullptr) {
		HIP_OK("libxm", hipHostMalloc((void **)&slots,
		    STAMP_SLOTS * sizeof(uint64_t), hipHostMallocMapped));
		memset(slots, 0, STAMP_SLOTS * sizeof(uint64_t));
	}
	return slots;
}

double
stamp_ms(uint64_t a, uint64_t b)
{
	static int khz = 0;
	if (khz == 0) {
		int dev = 0;
		HIP_OK("libxm", hipGetDevice(&dev));
		HIP_OK("libxm", hipDeviceGetAttribute(&khz,
		    hipDeviceAttributeWallClockRate, dev));
		if (khz <= 0)
			XM_FATAL("libxm", "device wall clock rate unavailable");
	}
	return b >= a ? (double)(b - a) / khz : 0.0;
}

XM_API int
xm_amd_synchronize(void)
{
	HIP_OK("xm_amd_synchronize", hipStreamSynchronize(xm_stream()));
	HIP_OK("xm_amd_synchronize", hipStreamSynchronize(xm_side_stream()));
	return 0;
}
