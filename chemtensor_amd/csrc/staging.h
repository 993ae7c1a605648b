// Per-thread pool of pinned host staging blocks behind copy_h2d_async: the policy, written against
// an abstract set of operations so that a host-only build can check its bounds.
//
// A block is a pinned buffer plus the event recorded behind the copy that last read it. A block is
// handed out again only once its event has completed (queried, or waited for), and a thread never
// holds more than STAGE_MAX_BLOCKS blocks: with all of them in flight it waits for its oldest copy.
// A block too small for a request is regrown only while idle. So per thread, at any time:
//   blocks (pinned buffers, events) <= STAGE_MAX_BLOCKS,
//   pinned bytes <= STAGE_MAX_BLOCKS * stage_cap(largest request so far).

#pragma once

#include <cstddef>
#include <deque>
#include <vector>

namespace ctg {

constexpr size_t STAGE_MAX_BLOCKS = 32;

inline size_t stage_cap(size_t nbytes)
{
	size_t cap = 64 * 1024;
	while (cap < nbytes) cap *= 2;
	return cap;
}

// Ops: Event type; void* host_alloc(size_t); void host_free(void*); Event event_create();
// void event_destroy(Event); bool event_done(Event); void event_wait(Event)
template <class Ops>
struct StagePool
{
	struct Block { char* p; size_t cap; typename Ops::Event ev; };
	std::vector<Block> idle;
	std::deque<Block> busy;

	size_t blocks() const { return idle.size() + busy.size(); }

	// a block of at least nbytes whose previous copy has completed
	Block acquire(Ops& ops, size_t nbytes)
	{
		while (!busy.empty() && ops.event_done(busy.front().ev)) {
			idle.push_back(busy.front());
			busy.pop_front();
		}
		int best = -1;
		for (int i = 0; i < (int)idle.size(); i++)
			if (idle[i].cap >= nbytes && (best < 0 || idle[i].cap < idle[best].cap)) best = i;
		Block b;
		if (best >= 0) {
			b = idle[best];
			idle.erase(idle.begin() + best);
		}
		else if (blocks() < STAGE_MAX_BLOCKS) {
			b.cap = stage_cap(nbytes);
			b.p = (char*)ops.host_alloc(b.cap);
			b.ev = ops.event_create();
		}
		else {
			// all blocks in use or too small: wait for the oldest copy if none is idle, then regrow
			// the largest idle block (its copy has completed)
			if (idle.empty()) {
				ops.event_wait(busy.front().ev);
				idle.push_back(busy.front());
				busy.pop_front();
			}
			int big = 0;
			for (int i = 1; i < (int)idle.size(); i++) if (idle[i].cap > idle[big].cap) big = i;
			b = idle[big];
			idle.erase(idle.begin() + big);
			if (b.cap < nbytes) {
				ops.host_free(b.p);
				b.cap = stage_cap(nbytes);
				b.p = (char*)ops.host_alloc(b.cap);
			}
		}
		return b;
	}

	// the block's copy has been enqueued and its event recorded behind it
	void submit(const Block& b) { busy.push_back(b); }

	// wait for every copy, then free every buffer and event
	void release(Ops& ops)
	{
		for (Block& b : busy) {
			ops.event_wait(b.ev);
			idle.push_back(b);
		}
		busy.clear();
		for (Block& b : idle) {
			ops.host_free(b.p);
			ops.event_destroy(b.ev);
		}
		idle.clear();
	}
};

}  // namespace ctg
