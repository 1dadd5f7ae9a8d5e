/*
 * filestore.cpp - backing file of path allocators and the layout of file
 * ranges in staging buffers (declarations and pointer format in
 * filestore.h).
 */
#include "filestore.h"

#include <algorithm>
#include <cerrno>
#include <cstdio>
#include <fcntl.h>
#include <sys/stat.h>
#include <unistd.h>

/* Growth step of the file when no free range fits: the current length,
 * at least 64 MiB and at most 4 GiB beyond what the request needs. */
static const uint64_t GROW_MIN = 64ull << 20;
static const uint64_t GROW_MAX = 4ull << 30;
/* Largest single pread/pwrite. */
static const size_t IO_MAX = 1ull << 30;

static uint64_t
round_up(uint64_t x, uint64_t a)
{
	return (x + a - 1) / a * a;
}

bool
file_store::create(const char *p)
{
	/* libxm src/alloc.c:163-176 */
	fd = open(p, O_CREAT | O_RDWR, S_IRUSR | S_IWUSR);
	if (fd == -1) {
		const int e = errno;
		perror("open");
		errno = e;
		return false;
	}
	if (ftruncate(fd, 0)) {
		const int e = errno;
		perror("ftruncate");
		if (close(fd))
			perror("close");
		fd = -1;
		errno = e;
		return false;
	}
	path = p;
	length = used = 0;
	free_list.clear();
	return true;
}

/* Insert [off, off + n) into the sorted free list, merging neighbours. */
static void
insert_free(std::vector<std::pair<uint64_t, uint64_t>> &fl, uint64_t off,
    uint64_t n)
{
	auto pos = std::lower_bound(fl.begin(), fl.end(),
	    std::make_pair(off, (uint64_t)0));
	pos = fl.insert(pos, {off, n});
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

uint64_t
file_store::allocate(uint64_t bytes)
{
	const uint64_t n = round_up(bytes ? bytes : 1, FILE_ALIGN);
	for (size_t i = 0; i < free_list.size(); i++) {
		auto &f = free_list[i];
		if (f.second >= n) {
			const uint64_t off = f.first;
			f.first += n;
			f.second -= n;
			if (f.second == 0)
				free_list.erase(free_list.begin() + i);
			used += n;
			return off;
		}
	}
	/* grow: a free range at the end of the file counts towards n */
	uint64_t tail = 0;
	if (!free_list.empty() && free_list.back().first +
	    free_list.back().second == length)
		tail = free_list.back().second;
	const uint64_t grow = std::min(GROW_MAX, std::max(GROW_MIN, length));
	const uint64_t nlen = round_up(length + (n - tail) + grow, FILE_ALIGN);
	if (nlen > FILE_OFF_MASK) {
		fprintf(stderr, "libxm: %s: the file would exceed 2^48 bytes\n",
		    path.c_str());
		return UINT64_MAX;
	}
	if (ftruncate(fd, (off_t)nlen)) {
		perror("ftruncate");
		return UINT64_MAX;
	}
	insert_free(free_list, length, nlen - length);
	length = nlen;
	return allocate(n);
}

void
file_store::release(uint64_t off, uint64_t bytes)
{
	const uint64_t n = round_up(bytes ? bytes : 1, FILE_ALIGN);
	used -= n;
	insert_free(free_list, off, n);
}

bool
file_store::read(uint64_t off, void *buf, size_t n) const
{
	char *p = (char *)buf;
	while (n > 0) {
		const ssize_t got = pread(fd, p, std::min(n, IO_MAX), (off_t)off);
		if (got < 0 && errno == EINTR)
			continue;
		if (got <= 0)
			return false;
		p += got;
		off += (uint64_t)got;
		n -= (size_t)got;
	}
	return true;
}

bool
file_store::write(uint64_t off, const void *buf, size_t n) const
{
	const char *p = (const char *)buf;
	while (n > 0) {
		const ssize_t put = pwrite(fd, p, std::min(n, IO_MAX),
		    (off_t)off);
		if (put < 0 && errno == EINTR)
			continue;
		if (put <= 0)
			return false;
		p += put;
		off += (uint64_t)put;
		n -= (size_t)put;
	}
	return true;
}

void
file_store::destroy()
{
	/* libxm src/alloc.c:331-336 */
	if (fd == -1)
		return;
	if (close(fd))
		perror("close");
	if (unlink(path.c_str()))
		perror("unlink");
	fd = -1;
}

void
stage_layout::add(uint64_t p, uint64_t n, bool ld, bool st)
{
	if (n)
		r.push_back({p, n, 0, ld, st});
}

void
stage_layout::finish()
{
	if (done)
		return;
	done = true;
	std::sort(r.begin(), r.end(), [](const range &x, const range &y) {
		return x.p < y.p;
	});
	std::vector<range> m;
	for (const range &x : r) {
		if (!m.empty() && file_slot(m.back().p) == file_slot(x.p) &&
		    x.p <= m.back().p + m.back().n) {
			range &y = m.back();
			y.n = std::max(y.p + y.n, x.p + x.n) - y.p;
			y.load |= x.load;
			y.store |= x.store;
		} else {
			m.push_back(x);
		}
	}
	bytes = 0;
	for (range &x : m) {
		x.off = bytes;
		bytes += round_up(x.n, 256);
	}
	r.swap(m);
}

uint64_t
stage_layout::offset(uint64_t p) const
{
	auto it = std::upper_bound(r.begin(), r.end(), p,
	    [](uint64_t v, const range &x) { return v < x.p; });
	if (it == r.begin())
		return UINT64_MAX;
	--it;
	if (p >= it->p + it->n)
		return UINT64_MAX;
	return it->off + (p - it->p);
}
