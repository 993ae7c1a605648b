// Host-only check of the pinned staging pool's policy (chemtensor_amd/csrc/staging.h) with simulated
// buffers, events and a simulated stream: prints the largest number of blocks, events and pinned
// bytes a thread held, and whether a block was ever handed out while its copy was in flight.
//
// usage: staging_check <scenario: stalled | fast | growing | mixed> <requests> <seed>

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <random>
#include <string>
#include <vector>

#include "staging.h"

namespace {

struct FakeOps
{
	typedef int Event;
	// stream model: a copy recorded with sequence number q completes once done_upto >= q
	std::vector<uint64_t> rec;     // per event: sequence number of its last record (0: never recorded)
	std::vector<bool> alive;
	uint64_t issued = 0, done_upto = 0;
	std::map<void*, size_t> live;  // pinned buffers
	size_t bytes = 0, max_bytes = 0, max_buffers = 0, max_events = 0, events = 0;

	void* host_alloc(size_t n)
	{
		void* p = malloc(n);
		live[p] = n;
		bytes += n;
		if (bytes > max_bytes) max_bytes = bytes;
		if (live.size() > max_buffers) max_buffers = live.size();
		return p;
	}
	void host_free(void* p)
	{
		auto it = live.find(p);
		if (it == live.end()) { fprintf(stderr, "free of an unknown buffer\n"); exit(3); }
		bytes -= it->second;
		live.erase(it);
		free(p);
	}
	Event event_create()
	{
		rec.push_back(0);
		alive.push_back(true);
		events++;
		if (events > max_events) max_events = events;
		return (int)rec.size() - 1;
	}
	void event_destroy(Event e) { alive[(size_t)e] = false; events--; }
	bool event_done(Event e) { return rec[(size_t)e] <= done_upto; }
	void event_wait(Event e) { if (rec[(size_t)e] > done_upto) done_upto = rec[(size_t)e]; }
	void record(Event e) { rec[(size_t)e] = ++issued; }
};

}  // namespace

int main(int argc, char** argv)
{
	if (argc != 4) { fprintf(stderr, "usage: %s scenario requests seed\n", argv[0]); return 2; }
	const std::string sc = argv[1];
	const int nreq = atoi(argv[2]);
	std::mt19937_64 rng((uint64_t)atoll(argv[3]));
	FakeOps ops;
	ctg::StagePool<FakeOps> pool;
	size_t max_blocks = 0, max_req = 0;
	int reused_in_flight = 0, too_small = 0;
	for (int r = 0; r < nreq; r++) {
		size_t n;
		if (sc == "growing") n = (size_t)1024 << std::min(r / 8, 12);
		else if (sc == "mixed") n = (rng() % 4 == 0) ? (size_t)(rng() % (16u << 20)) + 1 : (size_t)(rng() % 4096) + 1;
		else n = (size_t)(rng() % (4u << 20)) + 1;
		if (n > max_req) max_req = n;
		auto b = pool.acquire(ops, n);
		if (!ops.event_done(b.ev)) reused_in_flight++;
		if (b.cap < n) too_small++;
		memset(b.p, 1, n);   // the copy into the block (faults if the buffer is smaller than claimed)
		ops.record(b.ev);
		pool.submit(b);
		if (pool.blocks() > max_blocks) max_blocks = pool.blocks();
		// the simulated GPU: "stalled" completes nothing on its own, "fast" completes every copy at
		// once, "growing" and "mixed" complete a random prefix of the stream now and then
		if (sc == "fast") ops.done_upto = ops.issued;
		else if (sc != "stalled" && rng() % 3 == 0) ops.done_upto += rng() % (ops.issued - ops.done_upto + 1);
	}
	pool.release(ops);
	printf("max_blocks %zu\nmax_buffers %zu\nmax_events %zu\nmax_bytes %zu\nmax_request %zu\ncap_of_max_request %zu\n",
		max_blocks, ops.max_buffers, ops.max_events, ops.max_bytes, max_req, ctg::stage_cap(max_req));
	printf("reused_in_flight %d\ntoo_small %d\nlive_buffers_after_release %zu\nlive_events_after_release %zu\n",
		reused_in_flight, too_small, ops.live.size(), ops.events);
	return 0;
}
