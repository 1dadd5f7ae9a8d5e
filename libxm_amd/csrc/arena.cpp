/*
 * arena.cpp - block storage (replaces libxm's pagefile allocator, libxm
 * src/alloc.c:147-347), in three tiers.
 *
 * HBM: slabs obtained from hipMalloc, carved first-fit into 256-byte
 * aligned allocations.  Slabs grow geometrically (64 MiB up to 8 GiB) so
 * a tensor of many blocks costs few hipMalloc calls and the arena can fill
 * the 288 GB of an MI355X.  data_ptr is the device address of the
 * allocation.
 * Pinned host: when HBM is exhausted, or the arena reaches the limit
 * given by XM_ARENA_BYTES, new slabs come from pinned host memory that the
 * GPU reads directly over the host link, up to XM_HOST_BYTES (default: no
 * limit without a path, half the physical memory with one).
 * File: allocators created with a path (libxm's pagefile allocators) put
 * what does not fit the first two tiers in that file, created at
 * xm_allocator_create and removed by xm_allocator_destroy.  data_ptr is
 * then a file pointer (filestore.h) that no kernel dereferences: the
 * operations read such blocks into pinned buffers and write them back.
 * All entry points take the allocator's lock.
 */
#include "internal.h"

#include <algorithm>
#include <atomic>
#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <unistd.h>

static const size_t ALIGN = 256;
static const size_t SLAB_MIN = 64ull << 20;
static const size_t SLAB_MAX = 8ull << 30;
static std::atomic<uint64_t> next_alloc_id{1};

/* Allocators with a backing file, by the slot their file pointers carry. */
static std::mutex freg_mu;
static std::vector<xm_allocator_t *> freg;

static size_t
round_up(size_t x, size_t a)
{
	return (x + a - 1) / a * a;
}

static size_t
env_bytes(const char *name, size_t dflt)
{
	const char *e = getenv(name);
	return e ? (size_t)strtoull(e, nullptr, 10) : dflt;
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
	a->device_bytes = a->host_bytes = a->file_bytes = 0;
	a->file = nullptr;
	a->file_slot = 0;
	a->id = next_alloc_id++;
	a->limit_bytes = env_bytes("XM_ARENA_BYTES", SIZE_MAX);
	size_t half = SIZE_MAX;
	if (path) {
		long pages = sysconf(_SC_PHYS_PAGES), psz = sysconf(_SC_PAGESIZE);
		if (pages > 0 && psz > 0)
			half = (size_t)pages * (size_t)psz / 2;
	}
	a->host_limit = env_bytes("XM_HOST_BYTES", half);
	if (path) {
		file_store *f = new file_store;
		if (!f->create(path)) {
			delete f;
			delete a;
			return nullptr;
		}
		std::lock_guard<std::mutex> g(freg_mu);
		size_t slot = 0;
		while (slot < freg.size() && freg[slot] != nullptr)
			slot++;
		if (slot > FILE_SLOT_MAX) {
			fprintf(stderr, "libxm: xm_allocator_create: more than %u "
			    "allocators with a path\n", FILE_SLOT_MAX + 1);
			f->destroy();
			delete f;
			delete a;
			return nullptr;
		}
		if (slot == freg.size())
			freg.push_back(nullptr);
		freg[slot] = a;
		a->file = f;
		a->file_slot = (uint32_t)slot;
	}
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
		if (a->host_bytes + bytes > a->host_limit)
			return false;
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
	/* a file counts as one slab */
	if (file_ptr(ptr))
		return ptr & ~FILE_OFF_MASK;
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
		if (new_slab(a, bytes)) {
			p = carve(a->slabs.back(), bytes);
		} else if (a->file) {
			const uint64_t off = a->file->allocate(bytes);
			if (off == UINT64_MAX)
				return XM_NULL_PTR;
			p = file_make_ptr(a->file_slot, off);
			a->file_bytes += bytes;
		} else {
			return XM_NULL_PTR;
		}
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
	if (file_ptr(p)) {
		a->file->release(file_off(p), bytes);
		a->file_bytes -= bytes;
		return;
	}
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
	if (file_ptr(p)) {
		arena_file_read("xm_allocator_read", p, mem, bytes);
		return;
	}
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
	if (file_ptr(p)) {
		arena_file_write("xm_allocator_write", p, mem, bytes);
		return;
	}
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
	if (a->file) {
		{
			std::lock_guard<std::mutex> g(freg_mu);
			freg[a->file_slot] = nullptr;
		}
		a->file->destroy();
		delete a->file;
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

/* Bytes of live allocations in the allocator's file.  No libxm
 * counterpart. */
XM_API size_t
xm_amd_allocator_file_bytes(xm_allocator_t *a)
{
	std::lock_guard<std::mutex> g(a->mu);
	return a->file_bytes;
}

static file_store *
store_of(const char *fn, uint64_t p)
{
	std::lock_guard<std::mutex> g(freg_mu);
	const uint32_t slot = file_slot(p);
	if (slot >= freg.size() || freg[slot] == nullptr)
		XM_FATAL(fn, "file pointer of no open allocator");
	return freg[slot]->file;
}

void
arena_file_read(const char *fn, uint64_t p, void *buf, size_t bytes)
{
	file_store *f = store_of(fn, p);
	errno = 0;
	if (!f->read(file_off(p), buf, bytes))
		xm_fatal("%s: pread: %s", fn, errno ? strerror(errno) :
		    "end of file");
}

void
arena_file_write(const char *fn, uint64_t p, const void *buf, size_t bytes)
{
	file_store *f = store_of(fn, p);
	errno = 0;
	if (!f->write(file_off(p), buf, bytes))
		xm_fatal("%s: pwrite: %s", fn, errno ? strerror(errno) :
		    "no bytes written");
}

bool
tensor_in_file(const xm_tensor_t *t)
{
	if (t == nullptr)
		return false;
	for (const xm_block &b : t->blocks)
		if (b.type == XM_BLOCK_TYPE_CANONICAL && file_ptr(b.data))
			return true;
	return false;
}

/* ---- staging of file ranges for element kernels -------------------- */

size_t
file_batch_bytes(void)
{
	return env_bytes("XM_STAGE_BYTES", 1ull << 30);
}

void
file_stage::load(const char *fn)
{
	rd.finish();
	wr.finish();
	buf = (char *)pinned_buffer(slot, rd.bytes + wr.bytes);
	for (const auto &x : rd.r)
		arena_file_read(fn, x.p, buf + x.off, x.n);
	for (const auto &x : wr.r)
		if (x.load)
			arena_file_read(fn, x.p, buf + rd.bytes + x.off, x.n);
}

uint64_t
file_stage::src(uint64_t p) const
{
	if (!file_ptr(p))
		return p;
	uint64_t o = rd.offset(p);
	if (o == UINT64_MAX)
		XM_FATAL("libxm", "file range not staged");
	return (uint64_t)(uintptr_t)(buf + o);
}

uint64_t
file_stage::dst(uint64_t p) const
{
	if (!file_ptr(p))
		return p;
	uint64_t o = wr.offset(p);
	if (o == UINT64_MAX)
		XM_FATAL("libxm", "file range not staged");
	return (uint64_t)(uintptr_t)(buf + rd.bytes + o);
}

void
file_stage::store(const char *fn)
{
	for (const auto &x : wr.r)
		arena_file_write(fn, x.p, buf + rd.bytes + x.off, x.n);
}
