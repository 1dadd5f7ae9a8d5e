/*
 * Host checks of the file tier (libxm_amd/csrc/filestore.cpp): the file
 * pointer format, the backing file's allocator against a brute-force page
 * map over random allocate/release sequences, pread/pwrite round trips,
 * growth, creation failures, removal on destroy, and the staging layout's
 * merging and address mapping.  Prints "OK" on success.
 * Usage: filestore_check <scratch directory>
 */
#include "filestore.h"

#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <random>
#include <set>
#include <string>
#include <sys/stat.h>
#include <vector>

static int failures = 0;

#define CHECK(cond)                                                       \
	do {                                                              \
		if (!(cond)) {                                            \
			fprintf(stderr, "FAIL %s:%d: %s\n", __FILE__,     \
			    __LINE__, #cond);                             \
			failures++;                                       \
		}                                                         \
	} while (0)

static void
check_pointers()
{
	CHECK(!file_ptr(~0ull));			/* XM_NULL_PTR */
	CHECK(!file_ptr(0x7fffffffffffull));		/* top of user space */
	CHECK(!file_ptr(0x1000));
	for (uint32_t slot : {0u, 1u, 77u, (unsigned)FILE_SLOT_MAX})
		for (uint64_t off : {0ull, 4096ull, FILE_OFF_MASK}) {
			uint64_t p = file_make_ptr(slot, off);
			CHECK(file_ptr(p));
			CHECK(file_slot(p) == slot);
			CHECK(file_off(p) == off);
			CHECK(p != ~0ull);
		}
}

static bool
exists(const std::string &p)
{
	struct stat st;
	return stat(p.c_str(), &st) == 0;
}

static uint64_t
file_size(const std::string &p)
{
	struct stat st;
	return stat(p.c_str(), &st) == 0 ? (uint64_t)st.st_size : 0;
}

/* Random allocate/release against a page map: ranges never overlap, stay
 * inside the file, are page aligned; `used` and the free list account for
 * every page; data written to each live range reads back intact. */
static void
check_allocator(const std::string &dir)
{
	const std::string path = dir + "/pagefile";
	file_store f;
	CHECK(f.create(path.c_str()));
	CHECK(exists(path) && file_size(path) == 0);
	std::mt19937_64 rng(7);
	std::map<uint64_t, uint64_t> live;	/* off -> requested bytes */
	std::map<uint64_t, uint32_t> stamp;	/* off -> pattern seed */
	auto rounded = [](uint64_t n) {
		return (n ? n + FILE_ALIGN - 1 : FILE_ALIGN) / FILE_ALIGN *
		    FILE_ALIGN;
	};
	uint64_t max_len = 0;
	for (int step = 0; step < 3000; step++) {
		bool alloc = live.empty() || rng() % 3 != 0;
		if (alloc) {
			/* sizes from 0 bytes to 256 KiB */
			uint64_t n = rng() % 4 == 0 ? rng() % 8 :
			    rng() % (256u << 10);
			uint64_t off = f.allocate(n);
			CHECK(off != UINT64_MAX);
			CHECK(off % FILE_ALIGN == 0);
			CHECK(off + rounded(n) <= f.length);
			/* no overlap with any live range */
			auto it = live.lower_bound(off);
			if (it != live.end())
				CHECK(off + rounded(n) <= it->first);
			if (it != live.begin()) {
				--it;
				CHECK(it->first + rounded(it->second) <= off);
			}
			live[off] = n;
			if (n >= 8 && rng() % 8 == 0) {
				uint32_t seed = (uint32_t)rng();
				std::vector<uint32_t> buf(n / 4);
				for (size_t i = 0; i < buf.size(); i++)
					buf[i] = seed ^ (uint32_t)(i * 2654435761u);
				CHECK(f.write(off, buf.data(), buf.size() * 4));
				stamp[off] = seed;
			}
		} else {
			auto it = live.begin();
			std::advance(it, rng() % live.size());
			f.release(it->first, it->second);
			stamp.erase(it->first);
			live.erase(it);
		}
		max_len = std::max(max_len, f.length);
		if (step % 100 == 0) {
			uint64_t used = 0;
			for (auto &l : live)
				used += rounded(l.second);
			CHECK(f.used == used);
			uint64_t fr = 0;
			for (size_t i = 0; i < f.free_list.size(); i++) {
				fr += f.free_list[i].second;
				if (i)	/* sorted and coalesced */
					CHECK(f.free_list[i - 1].first +
					    f.free_list[i - 1].second <
					    f.free_list[i].first);
			}
			CHECK(used + fr == f.length);
			CHECK(file_size(path) == f.length);
		}
	}
	/* the stamped ranges still hold their data */
	for (auto &s : stamp) {
		uint64_t n = live[s.first];
		std::vector<uint32_t> buf(n / 4);
		CHECK(f.read(s.first, buf.data(), buf.size() * 4));
		bool same = true;
		for (size_t i = 0; i < buf.size(); i++)
			same &= buf[i] == (s.second ^ (uint32_t)(i * 2654435761u));
		CHECK(same);
	}
	/* releasing everything coalesces the file into one free range, and
	 * a request that fits is served from its start */
	for (auto &l : live)
		f.release(l.first, l.second);
	CHECK(f.used == 0);
	CHECK(f.free_list.size() == 1 && f.free_list[0].first == 0 &&
	    f.free_list[0].second == f.length);
	const uint64_t len = f.length;
	CHECK(f.allocate(1000) == 0);
	CHECK(f.length == len);
	/* a request beyond the file grows it; the tail free range counts */
	uint64_t big = f.allocate(len);
	CHECK(big != UINT64_MAX && big == FILE_ALIGN);
	CHECK(f.length > len && file_size(path) == f.length);
	/* reads past the end fail, reads of never-written pages give zeros */
	char z[64];
	memset(z, 1, sizeof z);
	CHECK(f.read(big + 4096, z, sizeof z));
	bool zero = true;
	for (char c : z)
		zero &= c == 0;
	CHECK(zero);
	CHECK(!f.read(f.length - 8, z, 64));
	f.destroy();
	CHECK(!exists(path));
	f.destroy();		/* a second destroy does nothing */
	(void)max_len;
}

static void
check_create_failure(const std::string &dir)
{
	file_store f;
	std::string bad = dir + "/no/such/dir/pagefile";
	CHECK(!f.create(bad.c_str()));
	CHECK(errno == ENOENT);
	CHECK(f.fd == -1);
	/* an existing file is truncated, as libxm's pagefile */
	std::string p = dir + "/existing";
	FILE *fp = fopen(p.c_str(), "w");
	fputs("old contents", fp);
	fclose(fp);
	CHECK(f.create(p.c_str()));
	CHECK(file_size(p) == 0);
	f.destroy();
	CHECK(!exists(p));
}

static void
check_layout()
{
	const uint64_t A = file_make_ptr(3, 0), B = file_make_ptr(4, 0);
	stage_layout l;
	l.add(A + 1000, 500, true, false);	/* [1000, 1500) */
	l.add(A + 1400, 300, false, true);	/* overlaps: [1000, 1700) */
	l.add(A + 1700, 100, true, false);	/* touches: [1000, 1800) */
	l.add(A + 5000, 10, true, false);	/* separate */
	l.add(B + 1000, 64, false, true);	/* other file, same offset */
	l.add(A + 9000, 0, true, true);		/* empty: ignored */
	l.finish();
	CHECK(l.r.size() == 3);
	CHECK(l.r[0].p == A + 1000 && l.r[0].n == 800);
	CHECK(l.r[0].load && l.r[0].store);
	CHECK(l.r[1].p == A + 5000 && l.r[1].n == 10);
	CHECK(l.r[1].load && !l.r[1].store);
	CHECK(l.r[2].p == B + 1000 && !l.r[2].load && l.r[2].store);
	for (auto &x : l.r)
		CHECK(x.off % 256 == 0);
	CHECK(l.r[0].off == 0 && l.r[1].off == 1024 && l.r[2].off == 1280);
	CHECK(l.bytes == 1536);
	CHECK(l.offset(A + 1000) == 0);
	CHECK(l.offset(A + 1799) == 799);
	CHECK(l.offset(A + 1800) == UINT64_MAX);
	CHECK(l.offset(A + 999) == UINT64_MAX);
	CHECK(l.offset(A + 5009) == 1024 + 9);
	CHECK(l.offset(B + 1000) == 1280);
	CHECK(l.offset(B + 999) == UINT64_MAX);
	l.finish();		/* idempotent */
	CHECK(l.r.size() == 3 && l.bytes == 1536);

	/* random requests: every requested byte maps into its merged range at
	 * the same distance from the range start, ranges are disjoint */
	std::mt19937_64 rng(11);
	for (int t = 0; t < 200; t++) {
		stage_layout m;
		std::vector<std::pair<uint64_t, uint64_t>> req;
		int n = 1 + (int)(rng() % 40);
		for (int i = 0; i < n; i++) {
			uint64_t p = file_make_ptr((uint32_t)(rng() % 3),
			    rng() % 100000);
			uint64_t len = 1 + rng() % 5000;
			req.push_back({p, len});
			m.add(p, len, rng() % 2, rng() % 2);
		}
		m.finish();
		for (size_t i = 1; i < m.r.size(); i++) {
			CHECK(m.r[i - 1].off + m.r[i - 1].n <= m.r[i].off);
			if (file_slot(m.r[i - 1].p) == file_slot(m.r[i].p))
				CHECK(m.r[i - 1].p + m.r[i - 1].n < m.r[i].p);
		}
		for (auto &q : req) {
			uint64_t o0 = m.offset(q.first);
			uint64_t o1 = m.offset(q.first + q.second - 1);
			CHECK(o0 != UINT64_MAX && o1 == o0 + q.second - 1);
			CHECK(o1 < m.bytes);
		}
	}
}

int
main(int argc, char **argv)
{
	if (argc < 2) {
		fprintf(stderr, "usage: %s <scratch directory>\n", argv[0]);
		return 2;
	}
	check_pointers();
	check_allocator(argv[1]);
	check_create_failure(argv[1]);
	check_layout();
	if (failures) {
		printf("%d failures\n", failures);
		return 1;
	}
	printf("OK\n");
	return 0;
}
