// Device runtime of libchemtensor_gpu: HIP error handling, the library stream, a caching
// device allocator (the DMRG sweep allocates and frees tensors of recurring sizes at every
// bond, which must not turn into hipMalloc/hipFree calls) and host <-> device transfers of
// block-sparse tensors in their packed (serialize-order) layout.

#include <algorithm>
#include <atomic>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <cstring>

#include "ct_internal.h"
#include "staging.h"

namespace ctg {

void hip_check(hipError_t e, const char* what, const char* file, int line)
{
	if (e == hipSuccess) return;
	fprintf(stderr, "chemtensor_gpu: HIP error '%s' (%d) in %s at %s:%d\n", hipGetErrorString(e), (int)e, what, file, line);
	throw HipError{ e, what };
}

hipStream_t default_stream()
{
	static hipStream_t s = nullptr;
	static std::once_flag once;
	std::call_once(once, [] { CTG_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking)); });
	return s;
}

//________________________________________________________________________________________________________________________
//
// caching allocator: freed blocks are kept per rounded size and handed out again

namespace {

// Freed blocks go to a cache of the freeing thread and are handed out again to that thread: its
// device work is ordered on the one stream it uses, so a block never returns to use while an
// earlier launch may still read it. A thread that has synchronised its stream (the serving pool
// after each task, the worker after each batch) passes its cached blocks to a shared cache, from
// which any thread may take them. The live map (block -> size) is sharded by address.
struct LiveShard
{
	std::mutex mu;
	std::unordered_map<void*, size_t> m;
};

struct ThreadCache
{
	std::mutex mu;     // taken by the owner, and by dev_trim
	std::multimap<size_t, void*> blocks;
};

// the shared cache: blocks by rounded size, sharded by size so that threads taking blocks of
// different sizes do not wait for each other
struct SharedShard
{
	std::mutex mu;
	std::unordered_map<size_t, std::vector<void*>> by_size;
};

struct Pool
{
	static constexpr int NSHARD = 64;
	LiveShard live[NSHARD];
	SharedShard shared[NSHARD];
	std::mutex mu;                         // the list of thread caches
	std::vector<ThreadCache*> caches;
	std::atomic<size_t> bytes_live{ 0 }, bytes_cached{ 0 };
	// Fibonacci hashing: block addresses and rounded sizes are multiples of large powers of two,
	// so their low bits alone would send most of them to one shard
	static int mix(uint64_t x) { return (int)((x * 0x9E3779B97F4A7C15ull) >> 58); }
	LiveShard& shard(void* p) { return live[mix((uint64_t)(uintptr_t)p)]; }
	SharedShard& shared_shard(size_t sz) { return shared[mix((uint64_t)sz)]; }
};

Pool& pool()
{
	static Pool* p = new Pool();
	return *p;
}

thread_local ThreadCache* tl_cache = nullptr;

ThreadCache& thread_cache()
{
	if (!tl_cache) {
		tl_cache = new ThreadCache();
		Pool& p = pool();
		std::lock_guard<std::mutex> g(p.mu);
		p.caches.push_back(tl_cache);
	}
	return *tl_cache;
}

size_t round_size(size_t n)
{
	if (n < 512) return 512;
	// 1/8 binade granularity
	size_t b = 1;
	while ((b << 3) < n) b <<= 1;
	return (n + b - 1) / b * b;
}

// a shared block of rounded size sz .. 5/4 sz (sizes in the cache are rounded sizes), or null
void* take_shared(Pool& p, size_t sz, size_t& got)
{
	for (size_t c = sz; c <= sz + sz / 4; c = round_size(c + 1)) {
		SharedShard& sh = p.shared_shard(c);
		std::lock_guard<std::mutex> g(sh.mu);
		auto it = sh.by_size.find(c);
		if (it != sh.by_size.end() && !it->second.empty()) {
			void* ptr = it->second.back();
			it->second.pop_back();
			got = c;
			return ptr;
		}
	}
	return nullptr;
}

// a block of at least sz bytes and at most 5/4 sz from a cache, or null
void* take_fit(std::multimap<size_t, void*>& m, size_t sz, size_t& got)
{
	auto it = m.lower_bound(sz);
	if (it == m.end() || it->first > sz + sz / 4) return nullptr;
	void* ptr = it->second;
	got = it->first;
	m.erase(it);
	return ptr;
}

void register_live(void* ptr, size_t sz)
{
	Pool& p = pool();
	LiveShard& sh = p.shard(ptr);
	std::lock_guard<std::mutex> g(sh.mu);
	sh.m[ptr] = sz;
	p.bytes_live += sz;
}

thread_local int tl_capture_depth = 0;

size_t cache_cap()
{
	static const size_t cap = [] {
		size_t fr = 0, tot = 0;
		if (hipMemGetInfo(&fr, &tot) != hipSuccess || tot == 0) {
			(void)hipGetLastError();
			return (size_t)16 << 30;
		}
		return tot / 8;
	}();
	return cap;
}

}  // namespace

void* dev_alloc(size_t nbytes)
{
	Pool& p = pool();
	const size_t sz = round_size(nbytes);
	size_t got = 0;
	void* ptr = nullptr;
	{
		ThreadCache& tc = thread_cache();
		std::lock_guard<std::mutex> g(tc.mu);
		ptr = take_fit(tc.blocks, sz, got);
	}
	if (!ptr) ptr = take_shared(p, sz, got);
	if (ptr) {
		p.bytes_cached -= got;
		register_live(ptr, got);
		return ptr;
	}
	// a long-running caller (the serving loop) sees ever new sizes: the caches are released once
	// they hold more than an eighth of the device memory
	if (p.bytes_cached > cache_cap() && tl_capture_depth == 0) dev_trim();
	hipError_t e = hipMalloc(&ptr, sz);
	if (e != hipSuccess) {
		// release the caches and retry once
		dev_trim();
		(void)hipGetLastError();
		CTG_HIP(hipMalloc(&ptr, sz));
	}
	register_live(ptr, sz);
	return ptr;
}

void dev_free(void* ptr)
{
	if (!ptr) return;
	Pool& p = pool();
	size_t sz = 0;
	{
		LiveShard& sh = p.shard(ptr);
		std::lock_guard<std::mutex> g(sh.mu);
		auto it = sh.m.find(ptr);
		if (it == sh.m.end()) return;
		sz = it->second;
		sh.m.erase(it);
	}
	p.bytes_live -= sz;
	ThreadCache& tc = thread_cache();
	std::lock_guard<std::mutex> g(tc.mu);
	tc.blocks.emplace(sz, ptr);
	p.bytes_cached += sz;
}

void dev_cache_to_shared()
{
	if (!tl_cache) return;
	std::multimap<size_t, void*> mine;
	{
		std::lock_guard<std::mutex> g(tl_cache->mu);
		mine.swap(tl_cache->blocks);
	}
	Pool& p = pool();
	// equal sizes are adjacent in the multimap: one lock per run of one size
	for (auto it = mine.begin(); it != mine.end();) {
		const size_t sz = it->first;
		SharedShard& sh = p.shared_shard(sz);
		std::lock_guard<std::mutex> g(sh.mu);
		std::vector<void*>& v = sh.by_size[sz];
		for (; it != mine.end() && it->first == sz; ++it) v.push_back(it->second);
	}
}

void dev_trim()
{
	Pool& p = pool();
	std::vector<void*> blocks;
	{
		// take every cached block out first: whatever work used them was enqueued before they
		// were freed, so the synchronisation below covers it
		size_t removed = 0;
		std::lock_guard<std::mutex> g(p.mu);
		for (SharedShard& sh : p.shared) {
			std::lock_guard<std::mutex> g2(sh.mu);
			for (auto& kv : sh.by_size) {
				for (void* b : kv.second) { blocks.push_back(b); removed += kv.first; }
				kv.second.clear();
			}
		}
		p.bytes_cached -= removed;
		removed = 0;
		for (ThreadCache* tc : p.caches) {
			std::lock_guard<std::mutex> g2(tc->mu);
			for (auto& b : tc->blocks) { blocks.push_back(b.second); removed += b.first; }
			tc->blocks.clear();
			p.bytes_cached -= removed;
			removed = 0;
		}
		p.bytes_cached -= removed;
	}
	if (blocks.empty()) return;
	(void)hipDeviceSynchronize();
	for (void* b : blocks) (void)hipFree(b);
}

//________________________________________________________________________________________________________________________
//
// host-to-device copies of host-side tables without host synchronisation: the data is copied into
// a pinned staging block owned by the calling thread and the DMA is enqueued on the stream; a block
// is reused only after the event recorded behind its copy has completed, and a thread holds at most
// STAGE_MAX_BLOCKS of them (staging.h). A thread's blocks are freed by release_thread_resources.

namespace {

// HIP operations of the staging pool; with quiet set (thread teardown) errors are ignored
struct HipStageOps
{
	typedef hipEvent_t Event;
	bool quiet = false;
	void* host_alloc(size_t n) { void* p = nullptr; CTG_HIP(hipHostMalloc(&p, n, hipHostMallocDefault)); return p; }
	void host_free(void* p) { const hipError_t e = hipHostFree(p); if (!quiet) CTG_HIP(e); }
	Event event_create() { hipEvent_t e; CTG_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming)); return e; }
	void event_destroy(Event e) { (void)hipEventDestroy(e); }
	bool event_done(Event e) { return hipEventQuery(e) == hipSuccess; }
	void event_wait(Event e) { const hipError_t r = hipEventSynchronize(e); if (!quiet) CTG_HIP(r); }
};

typedef StagePool<HipStageOps> Staging;
thread_local Staging* tl_staging = nullptr;

Staging& staging()
{
	if (!tl_staging) tl_staging = new Staging();
	return *tl_staging;
}

}  // namespace

void copy_h2d_async(void* dst, const void* src, size_t nbytes, hipStream_t s)
{
	const H2DPart part = { src, nbytes, 0 };
	copy_h2d_parts(dst, &part, 1, nbytes, s);
}

void copy_h2d_parts(void* dst, const H2DPart* parts, int nparts, size_t total, hipStream_t s)
{
	if (total == 0) return;
	Staging& S = staging();
	HipStageOps ops;
	Staging::Block b = S.acquire(ops, total);
	for (int i = 0; i < nparts; i++)
		if (parts[i].nbytes) memcpy(b.p + parts[i].dst_off, parts[i].src, parts[i].nbytes);
	CTG_HIP(hipMemcpyAsync(dst, b.p, total, hipMemcpyHostToDevice, s));
	CTG_HIP(hipEventRecord(b.ev, s));
	S.submit(b);
}

CaptureScope::CaptureScope() { tl_capture_depth++; }
CaptureScope::~CaptureScope() { tl_capture_depth--; }

void release_thread_resources()
{
	if (tl_staging) {
		HipStageOps ops;
		ops.quiet = true;
		tl_staging->release(ops);
		delete tl_staging;
		tl_staging = nullptr;
	}
	// device blocks cached for this thread go to the shared cache (its stream has completed the
	// work that used them); the thread's cache is unregistered
	if (tl_cache) {
		dev_cache_to_shared();
		Pool& p = pool();
		{
			std::lock_guard<std::mutex> g(p.mu);
			p.caches.erase(std::remove(p.caches.begin(), p.caches.end(), tl_cache), p.caches.end());
		}
		delete tl_cache;
		tl_cache = nullptr;
	}
}

size_t dev_bytes_in_use()
{
	return pool().bytes_live;
}

//________________________________________________________________________________________________________________________
//
// device tensors

DevTensor::DevTensor(Meta m) : meta(std::move(m))
{
	d = (double*)dev_alloc((size_t)(meta.nelem > 0 ? meta.nelem : 1) * sizeof(double));
}

DevTensor DevTensor::view(const DevTensor& o, Meta m)
{
	DevTensor v;
	v.meta = std::move(m);
	v.d = o.d;
	v.owner = false;
	return v;
}

DevTensor::DevTensor(DevTensor&& o) noexcept : meta(std::move(o.meta)), d(o.d), owner(o.owner) { o.d = nullptr; }

DevTensor& DevTensor::operator=(DevTensor&& o) noexcept
{
	if (this != &o) {
		if (owner) dev_free(d);
		meta = std::move(o.meta);
		d = o.d;
		owner = o.owner;
		o.d = nullptr;
	}
	return *this;
}

DevTensor::~DevTensor() { if (owner) dev_free(d); }

void DevTensor::zero(hipStream_t s) { launch_fill(d, 0.0, meta.nelem, s); }

DevTensor upload(const block_sparse_tensor* t, hipStream_t s)
{
	if (t->dtype != CT_DOUBLE_REAL) throw HipError{ hipErrorInvalidValue, "only CT_DOUBLE_REAL tensors are supported" };
	DevTensor dt(meta_of(t));
	std::vector<double> host((size_t)dt.meta.nelem);
	block_sparse_tensor_serialize_entries(t, host.data());
	copy_h2d_async(dt.d, host.data(), host.size() * sizeof(double), s);
	return dt;
}

void download(const DevTensor& t, block_sparse_tensor* out, hipStream_t s)
{
	const Meta& m = t.meta;
	std::vector<const qnumber*> qn(m.ndim);
	std::vector<tensor_axis_direction> ad(m.ndim);
	for (int i = 0; i < m.ndim; i++) { qn[i] = m.qnums_logical[i].data(); ad[i] = (tensor_axis_direction)m.axis_dir[i]; }
	allocate_block_sparse_tensor(CT_DOUBLE_REAL, m.ndim, m.dim_logical.data(), ad.data(), qn.data(), out);
	std::vector<double> host((size_t)m.nelem);
	CTG_HIP(hipMemcpyAsync(host.data(), t.d, host.size() * sizeof(double), hipMemcpyDeviceToHost, s));
	CTG_HIP(hipStreamSynchronize(s));
	block_sparse_tensor_deserialize_entries(out, host.data());
}

HostPacked download_packed(const DevTensor& t, hipStream_t s)
{
	HostPacked h;
	h.meta = t.meta;
	h.v.resize((size_t)t.meta.nelem);
	if (!h.v.empty()) CTG_HIP(hipMemcpyAsync(h.v.data(), t.d, h.v.size() * sizeof(double), hipMemcpyDeviceToHost, s));
	CTG_HIP(hipStreamSynchronize(s));
	return h;
}

DevTensor upload_packed(const HostPacked& h, hipStream_t s)
{
	DevTensor t(h.meta);
	copy_h2d_async(t.d, h.v.data(), h.v.size() * sizeof(double), s);
	return t;
}

}  // namespace ctg
