// Device runtime of libchemtensor_gpu: HIP error handling, the library stream, a caching
// device allocator (the DMRG sweep allocates and frees tensors of recurring sizes at every
// bond, which must not turn into hipMalloc/hipFree calls) and host <-> device transfers of
// block-sparse tensors in their packed (serialize-order) layout.

#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <cstring>

#include "ct_internal.h"

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

// Freed blocks are cached per host thread and handed out again only to the thread that freed
// them. A thread's device work is ordered on the stream it uses, and callers that work on several
// threads synchronise their stream at the end of each task, so a block never returns to use while
// another stream may still read it.
struct Pool
{
	std::mutex mu;
	std::map<std::thread::id, std::multimap<size_t, void*>> free_blocks;
	std::map<void*, size_t> live;
	size_t bytes_live = 0, bytes_cached = 0;
};

Pool& pool()
{
	static Pool* p = new Pool();
	return *p;
}

size_t round_size(size_t n)
{
	if (n < 512) return 512;
	// 1/8 binade granularity
	size_t b = 1;
	while ((b << 3) < n) b <<= 1;
	return (n + b - 1) / b * b;
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
	const size_t cap = cache_cap();
	bool over = false;
	{
		std::lock_guard<std::mutex> g(p.mu);
		auto& fb = p.free_blocks[std::this_thread::get_id()];
		auto it = fb.lower_bound(sz);
		if (it != fb.end() && it->first <= sz + sz / 4) {
			void* ptr = it->second;
			p.live[ptr] = it->first;
			p.bytes_live += it->first;
			p.bytes_cached -= it->first;
			fb.erase(it);
			return ptr;
		}
		over = p.bytes_cached > cap;
	}
	// a long-running caller (the serving loop) sees ever new sizes: the cache is released once it
	// holds more than an eighth of the device memory
	if (over && tl_capture_depth == 0) dev_trim();
	void* ptr = nullptr;
	hipError_t e = hipMalloc(&ptr, sz);
	if (e != hipSuccess) {
		// release the cache and retry once
		dev_trim();
		(void)hipGetLastError();
		CTG_HIP(hipMalloc(&ptr, sz));
	}
	std::lock_guard<std::mutex> g(p.mu);
	p.live[ptr] = sz;
	p.bytes_live += sz;
	return ptr;
}

void dev_free(void* ptr)
{
	if (!ptr) return;
	Pool& p = pool();
	std::lock_guard<std::mutex> g(p.mu);
	auto it = p.live.find(ptr);
	if (it == p.live.end()) return;
	p.free_blocks[std::this_thread::get_id()].emplace(it->second, ptr);
	p.bytes_cached += it->second;
	p.bytes_live -= it->second;
	p.live.erase(it);
}

void dev_trim()
{
	Pool& p = pool();
	std::lock_guard<std::mutex> g(p.mu);
	bool any = false;
	for (auto& kv : p.free_blocks) any = any || !kv.second.empty();
	if (any) (void)hipDeviceSynchronize();
	for (auto& kv : p.free_blocks) {
		for (auto& b : kv.second) (void)hipFree(b.second);
		kv.second.clear();
	}
	p.bytes_cached = 0;
}

//________________________________________________________________________________________________________________________
//
// host-to-device copies of host-side tables without host synchronisation: the data is copied into
// a pinned staging block owned by the calling thread and the DMA is enqueued on the stream; a block
// is reused only after the event recorded behind its copy has completed. Staging blocks live until
// the process ends.

namespace {

struct StageBlock { char* p; size_t cap; hipEvent_t ev; };

struct Staging
{
	std::vector<StageBlock> idle;
	std::deque<StageBlock> busy;
};

// at most this many staging blocks (pinned buffer + event) per thread: a thread with more copies in
// flight waits for its oldest one
constexpr size_t STAGE_MAX_BLOCKS = 32;

thread_local Staging* tl_staging = nullptr;

Staging& staging()
{
	if (!tl_staging) tl_staging = new Staging();
	return *tl_staging;
}

size_t stage_cap(size_t nbytes)
{
	size_t cap = 64 * 1024;
	while (cap < nbytes) cap *= 2;
	return cap;
}

}  // namespace

void copy_h2d_async(void* dst, const void* src, size_t nbytes, hipStream_t s)
{
	if (nbytes == 0) return;
	Staging& S = staging();
	while (!S.busy.empty() && hipEventQuery(S.busy.front().ev) == hipSuccess) {
		S.idle.push_back(S.busy.front());
		S.busy.pop_front();
	}
	int best = -1;
	for (int i = 0; i < (int)S.idle.size(); i++)
		if (S.idle[i].cap >= nbytes && (best < 0 || S.idle[i].cap < S.idle[best].cap)) best = i;
	StageBlock b;
	if (best >= 0) {
		b = S.idle[best];
		S.idle.erase(S.idle.begin() + best);
	}
	else if (S.idle.size() + S.busy.size() < STAGE_MAX_BLOCKS) {
		const size_t cap = stage_cap(nbytes);
		void* p = nullptr;
		CTG_HIP(hipHostMalloc(&p, cap, hipHostMallocDefault));
		b.p = (char*)p;
		b.cap = cap;
		CTG_HIP(hipEventCreateWithFlags(&b.ev, hipEventDisableTiming));
	}
	else {
		// all blocks in use or too small: wait for the oldest copy if none is idle, then grow the
		// largest idle block's buffer (its copy has completed)
		if (S.idle.empty()) {
			CTG_HIP(hipEventSynchronize(S.busy.front().ev));
			S.idle.push_back(S.busy.front());
			S.busy.pop_front();
		}
		int big = 0;
		for (int i = 1; i < (int)S.idle.size(); i++) if (S.idle[i].cap > S.idle[big].cap) big = i;
		b = S.idle[big];
		S.idle.erase(S.idle.begin() + big);
		if (b.cap < nbytes) {
			CTG_HIP(hipHostFree(b.p));
			const size_t cap = stage_cap(nbytes);
			void* p = nullptr;
			CTG_HIP(hipHostMalloc(&p, cap, hipHostMallocDefault));
			b.p = (char*)p;
			b.cap = cap;
		}
	}
	memcpy(b.p, src, nbytes);
	CTG_HIP(hipMemcpyAsync(dst, b.p, nbytes, hipMemcpyHostToDevice, s));
	CTG_HIP(hipEventRecord(b.ev, s));
	S.busy.push_back(b);
}

CaptureScope::CaptureScope() { tl_capture_depth++; }
CaptureScope::~CaptureScope() { tl_capture_depth--; }

void release_thread_resources()
{
	if (tl_staging) {
		Staging& S = *tl_staging;
		for (StageBlock& b : S.busy) {
			(void)hipEventSynchronize(b.ev);
			S.idle.push_back(b);
		}
		S.busy.clear();
		for (StageBlock& b : S.idle) {
			(void)hipHostFree(b.p);
			(void)hipEventDestroy(b.ev);
		}
		delete tl_staging;
		tl_staging = nullptr;
	}
	// device blocks cached for this thread (its stream has completed the work that used them)
	std::multimap<size_t, void*> mine;
	{
		Pool& p = pool();
		std::lock_guard<std::mutex> g(p.mu);
		auto it = p.free_blocks.find(std::this_thread::get_id());
		if (it != p.free_blocks.end()) {
			mine.swap(it->second);
			p.free_blocks.erase(it);
			for (auto& b : mine) p.bytes_cached -= b.first;
		}
	}
	for (auto& b : mine) (void)hipFree(b.second);
}

size_t dev_bytes_in_use()
{
	Pool& p = pool();
	std::lock_guard<std::mutex> g(p.mu);
	return p.bytes_live;
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
