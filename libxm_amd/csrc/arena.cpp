/*
 * arena.cpp - block storage in HBM (replaces libxm's disk pagefile
 * allocator, libxm src/alloc.c:147-347).
 *
 * Storage is a list of slabs obtained from hipMalloc, carved first-fit
 * into 256-byte aligned allocations.  Slabs grow geometrically (64 MiB up
 * to 8 GiB) so a tensor of many blocks costs few hipMalloc calls and the
 * arena can fill the 288 GB of an MI355X.  data_ptr is the device address
 * of the allocation.  When HBM is exhausted, or the arena reaches the
 * limit given by XM_ARENA_BYTES, new slabs come from pinned host memory
 * that the GPU reads directly over the host link, so a tensor larger than
 * HBM still works (slowly).  All entry points take the allocator's lock.
 */
#include "internal.h"

#include <algorithm>
#include <atomic>
#include <cstdio>
#include <cstdlib>

static const size_t ALIGN = 256;
static const size_t SLAB_MIN = 64ull << 20;
static const size_t SLAB_MAX = 8ull << 30;
static std::atomic<uint64_t> next_alloc_id{1};

static size_t
round_up(size_t x, size_t a)
{
	return (x + a - 1) / a * a;
}

XM_API xm_allocator_t *
xm_allocator_create(const char *path)
{
	xm_allocator_t *a = new (std::nothrow) xm_allocator_t;
	if (a == nullptr) {
		perror("xm_allocator_create");
		return nullptr;
	}
	a->has_path = path != nullptr;
	if (path)
		a->path = path;
	a->device_bytes = a->host_bytes = 0;
	a->limit_bytes = SIZE_MAX;
	a->id = next_alloc_id++;
	if (const char *lim = getenv("XM_ARENA_BYTES"))
		a->limit_bytes = strtoull(lim, nullptr, 10);
	return a;
}

XM_API const char *
xm_allocator_get_path(xm_allocator_t *a)
{
	return a->has_path ? a->path.c_str() : nullptr;
}

static bool
new_slab(xm_allocator_t *a, size_t need)
{
	size_t grow = a->slabs.empty() ? SLAB_MIN :
	    std::min(SLAB_MAX, a->slabs.back().bytes * 2);
	size_t bytes = round_up(std::max(need, grow), 2ull << 20);
	arena_slab s;
	s.bytes = bytes;
	s.host = false;
	s.base = nullptr;
	if (a->device_bytes + bytes <= a->limit_bytes) {
		if (hipMalloc((void **)&s.base, bytes) != hipSuccess) {
			(void)hipGetLastError();
			s.base = nullptr;
			/* fall back to the exact request if a big slab failed */
			if (bytes > round_up(need, 2ull << 20)) {
				bytes = round_up(need, 2ull << 20);
				if (hipMalloc((void **)&s.base, bytes) !=
				    hipSuccess) {
					(void)hipGetLastError();
					s.base = nullptr;
				}
				s.bytes = bytes;
			}
		}
	}
	if (s.base == nullptr) {
		bytes = round_up(need, 2ull << 20);
		if (hipHostMalloc((void **)&s.base, bytes,
		    hipHostMallocMapped) != hipSuccess) {
			(void)hipGetLastError();
			return false;
		}
		s.bytes = bytes;
		s.host = true;
		a->host_bytes += bytes;
	} else {
		a->device_bytes += s.bytes;
	}
	s.free_list.push_back({0, s.bytes});
	a->slabs.push_back(std::move(s));
	return true;
}

static uint64_t
carve(arena_slab &s, size_t bytes)
{
	for (size_t i = 0; i < s.free_list.size(); i++) {
		auto &f = s.free_list[i];
		if (f.second >= bytes) {
			size_t off = f.first;
			f.first += bytes;
			f.second -= bytes;
			if (f.second == 0)
				s.free_list.erase(s.free_list.begin() + i);
			return (uint64_t)(uintptr_t)(s.base + off);
		}
	}
	return XM_NULL_PTR;
}

uint64_t
arena_host_slab(xm_allocator_t *a, uint64_t ptr)
{
	if (a == nullptr || ptr == XM_NULL_PTR)
		return 0;
	std::lock_guard<std::mutex> g(a->mu);
	if (a->host_bytes == 0)
		return 0;
	for (const arena_slab &s : a->slabs)
		if (s.host && ptr >= (uint64_t)(uintptr_t)s.base &&
		    ptr < (uint64_t)(uintptr_t)s.base + s.bytes)
			return (uint64_t)(uintptr_t)s.base;
	return 0;
}

XM_API uint64_t
xm_allocator_allocate(xm_allocator_t *a, size_t size_bytes)
{
	size_t bytes = round_up(size_bytes ? size_bytes : 1, ALIGN);
	std::lock_guard<std::mutex> g(a->mu);
	uint64_t p = XM_NULL_PTR;
	(void)xm_stream();	/* the device comes up with the first block */
	for (auto &s : a->slabs)
		if ((p = carve(s, bytes)) != XM_NULL_PTR)
			break;
	if (p == XM_NULL_PTR) {
		if (!new_slab(a, bytes))
			return XM_NULL_PTR;
		p = carve(a->slabs.back(), bytes);
	}
	auto it = std::lower_bound(a->live.begin(), a->live.end(),
	    std::make_pair(p, (size_t)0));
	a->live.insert(it, {p, bytes});
	return p;
}

static arena_slab *
slab_of(xm_allocator_t *a, uint64_t p)
{
	for (auto &s : a->slabs)
		if (p >= (uint64_t)(uintptr_t)s.base &&
		    p < (uint64_t)(uintptr_t)s.base + s.bytes)
			return &s;
	return nullptr;
}

XM_API void
xm_allocator_deallocate(xm_allocator_t *a, uint64_t p)
{
	if (p == XM_NULL_PTR)
		return;
	bump_data_epoch();
	std::lock_guard<std::mutex> g(a->mu);
	auto it = std::lower_bound(a->live.begin(), a->live.end(),
	    std::make_pair(p, (size_t)0));
	if (it == a->live.end() || it->first != p)
		XM_FATAL("xm_allocator_deallocate", "unknown data pointer");
	size_t bytes = it->second;
	a->live.erase(it);
	arena_slab *s = slab_of(a, p);
	size_t off = p - (uint64_t)(uintptr_t)s->base;
	auto &fl = s->free_list;
	auto pos = std::lower_bound(fl.begin(), fl.end(),
	    std::make_pair(off, (size_t)0));
	pos = fl.insert(pos, {off, bytes});
	/* coalesce with the neighbours */
	if (pos + 1 != fl.end() && pos->first + pos->second == (pos + 1)->first) {
		pos->second += (pos + 1)->second;
		fl.erase(pos + 1);
	}
	if (pos != fl.begin() && (pos - 1)->first + (pos - 1)->second ==
	    pos->first) {
		(pos - 1)->second += pos->second;
		fl.erase(pos);
	}
}

static void
check_range(const char *fn, xm_allocator_t *a, uint64_t p, size_t bytes)
{
	if (p == XM_NULL_PTR)
		XM_FATAL(fn, "data pointer is NULL");
	std::lock_guard<std::mutex> g(a->mu);
	auto it = std::upper_bound(a->live.begin(), a->live.end(),
	    std::make_pair(p, SIZE_MAX));
	if (it == a->live.begin())
		XM_FATAL(fn, "unknown data pointer");
	--it;
	if (p + bytes > it->first + it->second)
		XM_FATAL(fn, "access beyond the allocation");
}

XM_API void
xm_allocator_read(xm_allocator_t *a, uint64_t p, void *mem, size_t bytes)
{
	check_range("xm_allocator_read", a, p, bytes);
	HIP_OK("xm_allocator_read", hipMemcpyAsync(mem,
	    (const void *)(uintptr_t)p, bytes, hipMemcpyDeviceToHost,
	    xm_stream()));
	HIP_OK("xm_allocator_read", hipStreamSynchronize(xm_stream()));
}

XM_API void
xm_allocator_write(xm_allocator_t *a, uint64_t p, const void *mem,
    size_t bytes)
{
	check_range("xm_allocator_write", a, p, bytes);
	bump_data_epoch();
	HIP_OK("xm_allocator_write", hipMemcpyAsync((void *)(uintptr_t)p, mem,
	    bytes, hipMemcpyHostToDevice, xm_stream()));
	HIP_OK("xm_allocator_write", hipStreamSynchronize(xm_stream()));
}

void bwd_cache_drop_alloc(xm_allocator_t *a);

XM_API void
xm_allocator_destroy(xm_allocator_t *a)
{
	if (a == nullptr)
		return;
	bwd_cache_drop_alloc(a);
	if (!a->slabs.empty())
		HIP_OK("xm_allocator_destroy",
		    hipStreamSynchronize(xm_stream()));
	for (auto &s : a->slabs) {
		if (s.host)
			HIP_OK("xm_allocator_destroy", hipHostFree(s.base));
		else
			HIP_OK("xm_allocator_destroy", hipFree(s.base));
	}
	delete a;
}

XM_API void
xm_amd_allocator_usage(xm_allocator_t *a, size_t *dev, size_t *host)
{
	std::lock_guard<std::mutex> g(a->mu);
	if (dev)
		*dev = a->device_bytes;
	if (host)
		*host = a->host_bytes;
}
