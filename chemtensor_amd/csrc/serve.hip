// Serving loop for independent DMRG ground-state solves with continuous batching.
//
// Each admitted request cycles on its own: [start] -> ready -> batched Lanczos -> split -> ready ...
// A worker thread owns the batched Lanczos (one HIP stream): whenever requests are ready it plans
// one batch over all of them (one grouped launch per stage for all their matvecs of a Lanczos step,
// one recurrence per request, grouped by Lanczos iteration count), solves it and hands every
// request to a pool of host threads (each with its own stream) for its split: device reshape and
// download, host block SVD with truncation, upload, environment update, next two-site tensor. A
// request rejoins the ready set when its split is done, so slow splits of large requests do not
// hold back small ones, and a request that arrives while others run is admitted (in arrival
// order, while the HBM budget and the slots allow) and starts on the pool at once. Requests that
// finish free their budget. The grouped launches use one fixed launch form per stage, every output
// tile is summed in a fixed order (by one wavefront, or by four with a fixed LDS reduction), and
// each request's Lanczos reductions cover only its own segment, so a request's results do not
// depend on the other requests in the batch.

#include <algorithm>
#include <string>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <deque>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <thread>

#include "ct_internal.h"

namespace ctg {

namespace {

struct Ticket
{
	int64_t id = 0;
	ct_serve_request req = {};
	double t_submit = 0, t_admit = -1, t_done = -1;
	int status = 1;
	std::vector<double> en;
	size_t reserved = 0;
	std::unique_ptr<DmrgState> state;
};

size_t estimate_bytes(const ct_serve_request& r)
{
	// dense upper bounds: MPS sites L chi^2 d, environments 2 L chi^2 M, merged MPO pairs L M^2 d^4,
	// Lanczos vectors (maxiter + 3) chi^2 d^2, grouped intermediates 2 M chi^2 d^2
	const double d = (r.model == 0) ? 2 : 4, M = (r.model == 0) ? 5 : 6;
	const double chi = (double)r.max_vdim, L = r.nsites;
	const double el = L * chi * chi * d + 2 * L * chi * chi * M + L * M * M * d * d * d * d
		+ (r.maxiter_lanczos + 3) * chi * chi * d * d + 2 * M * chi * chi * d * d;
	return (size_t)(8.0 * el);
}

// Persistent host threads, each with its own HIP stream, serving a task queue. Every task ends with
// its stream synchronised, so buffers it freed are reused (by its thread) only in stream order.
struct HostPool
{
	std::vector<std::thread> th;
	std::mutex mu;
	std::condition_variable cv;
	std::deque<std::function<void(hipStream_t)>> q;
	bool stop = false;

	void start(int n)
	{
		for (int i = 0; i < n; i++)
			th.emplace_back([this] {
				hipStream_t s = nullptr;
				if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) s = nullptr;
				host_svd_serial_on_this_thread(true);   // the pool's threads already split in parallel
				for (;;) {
					std::function<void(hipStream_t)> f;
					{
						std::unique_lock<std::mutex> g(mu);
						cv.wait(g, [&] { return stop || !q.empty(); });
						if (stop) break;
						f = std::move(q.front());
						q.pop_front();
					}
					f(s);
					if (s) (void)hipStreamSynchronize(s);
					dev_cache_to_shared();
				}
				release_thread_resources();
				if (s) (void)hipStreamDestroy(s);
			});
	}

	void submit(std::function<void(hipStream_t)> f)
	{
		{
			std::lock_guard<std::mutex> g(mu);
			q.push_back(std::move(f));
		}
		cv.notify_one();
	}

	// f(0 .. n-1) on the pool; returns when all have finished
	void run(int n, const std::function<void(int, hipStream_t)>& f)
	{
		if (n <= 0) return;
		std::mutex m;
		std::condition_variable done;
		int left = n;
		for (int i = 0; i < n; i++)
			submit([&, i](hipStream_t s) {
				f(i, s);
				if (s) (void)hipStreamSynchronize(s);
				std::lock_guard<std::mutex> g(m);
				if (--left == 0) done.notify_all();
			});
		std::unique_lock<std::mutex> g(m);
		done.wait(g, [&] { return left == 0; });
	}

	void shutdown()
	{
		{
			std::lock_guard<std::mutex> g(mu);
			stop = true;
			q.clear();
		}
		cv.notify_all();
		for (auto& t : th) t.join();
		th.clear();
	}
};

}  // namespace

struct Server
{
	HostPool hosts;
	// threads of the worker's batch plans only: a plan never queues behind the splits in 'hosts'
	HostPool planners;
	std::mutex mu;
	std::condition_variable cv_work, cv_done;
	std::deque<std::shared_ptr<Ticket>> pending;
	std::map<int64_t, std::shared_ptr<Ticket>> tickets;
	std::deque<std::shared_ptr<Ticket>> ready;     // started requests waiting for their next Lanczos solve
	int64_t running = 0;                            // admitted requests not yet finished
	int64_t next_id = 1;
	bool stop = false;
	double budget = 0;
	size_t used = 0;
	int max_active = 0;
	int host_threads = 8;
	int64_t ticks = 0, lanczos_batches = 0, completed = 0, max_batch = 0;
	double busy_s = 0;
	// seconds spent in each part of the ticks: start of requests, batch plans, batched Lanczos,
	// device part of the splits before / after the host SVDs, host SVDs
	double t_start = 0, t_plan = 0, t_lanczos = 0, t_split_dev = 0, t_split_host = 0, t_splits = 0;
	double t_plan_part[4] = { 0, 0, 0, 0 };       // the batch plans' parts (BatchPlan::host_s)
	hipStream_t st = nullptr;
	std::chrono::steady_clock::time_point t0;
	std::thread worker;

	double now() const { return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(); }

	void finish(Ticket& t, int status)
	{
		if (status < 0)
			fprintf(stderr, "chemtensor_gpu: serve request %lld (model %d, L %d, chi %lld) failed with status %d\n",
				(long long)t.id, t.req.model, t.req.nsites, (long long)t.req.max_vdim, status);
		t.state.reset();
		std::lock_guard<std::mutex> g(mu);
		t.status = status;
		t.t_done = now();
		used -= t.reserved;
		t.reserved = 0;
		completed++;
		running--;
		cv_done.notify_all();
		cv_work.notify_one();   // budget and a slot freed
	}

	void make_ready(const std::shared_ptr<Ticket>& t)
	{
		{
			std::lock_guard<std::mutex> g(mu);
			ready.push_back(t);
		}
		cv_work.notify_one();
	}

	int start_request(Ticket& t, hipStream_t s)
	{
		const ct_serve_request& r = t.req;
		mpo H = {};
		mps psi = {};
		if (r.model == 0) ct_construct_heisenberg_xxz_1d_mpo(r.nsites, r.param[0], r.param[1], r.param[2], &H);
		else ct_construct_fermi_hubbard_1d_mpo(r.nsites, r.param[0], r.param[1], r.param[2], &H);
		rng_state rng;
		seed_rng_state(r.seed, &rng);
		construct_random_mps(CT_DOUBLE_REAL, r.nsites, H.d, H.qsite, r.sector, r.max_vdim, &rng, &psi);
		t.state.reset(new DmrgState());
		t.state->st = s;
		t.state->plan_ahead = true;   // each split task plans the next batch tables of its request
		// the split's factorisation in the server: on the host unless CT_SERVE_SPLIT=device. The device
		// factorisation's long one-workgroup kernels slow the batched Lanczos of the other requests
		// more than they relieve the host (300 requests at 20/s: 3.7 against 8.2 solves/s)
		{
			const char* e = getenv("CT_SERVE_SPLIT");
			t.state->split_choice = (e && std::string(e) == "device") ? DmrgState::SPLIT_DEVICE_EIG : DmrgState::SPLIT_HOST;
		}
		int ret = CT_ERR_HIP;
		try {
			ret = t.state->init(&H, &psi, r.num_sweeps, r.maxiter_lanczos, r.tol_split, r.max_vdim, r.tol_energy);
		}
		catch (const HipError&) {
			ret = CT_ERR_HIP;
		}
		delete_mps(&psi);
		delete_mpo(&H);
		return ret;
	}

	// one batched Lanczos over the active requests with this iteration count, then their splits
	void step_group(std::vector<std::shared_ptr<Ticket>>& grp, int maxiter)
	{
		std::vector<BatchInput> in(grp.size());
		for (size_t i = 0; i < grp.size(); i++) {
			DmrgState& S = *grp[i]->state;
			in[i].theta = S.theta.meta;
			in[i].w = DevTensor::view(S.h2[S.bond], S.h2[S.bond].meta);
			in[i].l = DevTensor::view(S.lb[S.bond], S.lb[S.bond].meta);
			in[i].r = DevTensor::view(S.rb[S.bond + 1], S.rb[S.bond + 1].meta);
			in[i].pre = S.ahead.get();
		}
		double c0 = now();
		const ParallelFor par = [this](int n, const std::function<void(int)>& f) { planners.run(n, [&f](int i, hipStream_t) { f(i); }); };
		std::unique_ptr<BatchPlan> b(build_batch_plan(in, st, false, &par));
		double c1 = now();
		t_plan += c1 - c0;
		for (int i = 0; i < 4; i++) t_plan_part[i] += b->host_s[i];
		const int64_t ntot = b->off.back();
		double* v0 = (double*)dev_alloc((size_t)std::max<int64_t>(ntot, 1) * sizeof(double));
		double* u = (double*)dev_alloc((size_t)std::max<int64_t>(ntot, 1) * sizeof(double));
		for (size_t i = 0; i < grp.size(); i++) {
			const DmrgState& S = *grp[i]->state;
			if (S.theta.meta.nelem) launch_copy(S.theta.d, v0 + b->off[i], S.theta.meta.nelem, st);
		}
		std::vector<double> lam(grp.size());
		std::vector<int> status(grp.size());
		lanczos_ground_state_batch(*b, v0, maxiter, lam.data(), status.data(), u, false);
		lanczos_batches++;
		c0 = now();
		t_lanczos += c0 - c1;
		max_batch = std::max<int64_t>(max_batch, (int64_t)grp.size());
		// every request's ground state into its own buffer, then its split as an asynchronous host task
		std::vector<std::shared_ptr<DevTensor>> aopt(grp.size());
		for (size_t i = 0; i < grp.size(); i++) {
			if (status[i] != 0) continue;
			const DmrgState& S = *grp[i]->state;
			aopt[i] = std::make_shared<DevTensor>(S.theta.meta);
			if (S.theta.meta.nelem) launch_copy(u + b->off[i], aopt[i]->d, S.theta.meta.nelem, st);
		}
		CTG_HIP(hipStreamSynchronize(st));
		dev_free(v0);
		dev_free(u);
		b.reset();
		dev_cache_to_shared();   // the worker's stream has completed the batch
		for (size_t i = 0; i < grp.size(); i++) {
			std::shared_ptr<Ticket> t = grp[i];
			if (status[i] != 0) { finish(*t, status[i]); continue; }
			const double en = lam[i];
			std::shared_ptr<DevTensor> a = aopt[i];
			hosts.submit([this, t, a, en](hipStream_t s) {
				DmrgState& S = *t->state;
				int ret;
				double td = 0, th = 0;
				try {
					S.st = s;
					const double a0 = now();
					S.advance_begin(*a);
					const double a1 = now();
					S.advance_host();
					const double a2 = now();
					ret = S.advance_end(en);
					td = (a1 - a0) + (now() - a2);
					th = a2 - a1;
				}
				catch (const HipError&) {
					ret = CT_ERR_HIP;
				}
				{
					std::lock_guard<std::mutex> g(mu);
					t_split_dev += td;
					t_split_host += th;
				}
				if (ret != 0) finish(*t, ret);
				else if (S.done) { t->en = S.en_sweeps; finish(*t, 0); }
				else make_ready(t);
			});
		}
		t_splits += now() - c0;
	}

	void loop()
	{
		for (;;) {
			std::vector<std::shared_ptr<Ticket>> batch;
			{
				std::unique_lock<std::mutex> g(mu);
				auto admissible = [&] {
					if (pending.empty() || running >= max_active) return false;
					const size_t need = estimate_bytes(pending.front()->req);
					return used + need <= budget || running == 0;
				};
				cv_work.wait(g, [&] { return stop || !ready.empty() || admissible(); });
				if (stop) return;
				// admission in arrival order while the budget and the slots allow; the start of a
				// request (MPO, seeded MPS, upload, normalisation, environments) is a host task
				while (admissible()) {
					std::shared_ptr<Ticket> t = pending.front();
					pending.pop_front();
					t->reserved = estimate_bytes(t->req);
					used += t->reserved;
					running++;
					t->t_admit = now();
					hosts.submit([this, t](hipStream_t s) {
						const double c = now();
						const int ret = start_request(*t, s);
						{
							std::lock_guard<std::mutex> g2(mu);
							t_start += now() - c;
						}
						if (ret != 0) finish(*t, ret);
						else make_ready(t);
					});
				}
				batch.assign(ready.begin(), ready.end());
				ready.clear();
			}
			if (batch.empty()) continue;
			const double tick0 = now();
			std::map<int, std::vector<std::shared_ptr<Ticket>>> groups;
			for (auto& t : batch) groups[t->req.maxiter_lanczos].push_back(t);
			for (auto& kv : groups) {
				try {
					step_group(kv.second, kv.first);
				}
				catch (const HipError&) {
					for (auto& t : kv.second) if (t->status == 1) finish(*t, CT_ERR_HIP);
				}
			}
			ticks++;
			busy_s += now() - tick0;
		}
	}
};

}  // namespace ctg

using namespace ctg;

struct ct_serve
{
	std::unique_ptr<Server> s;
};

extern "C" {

int ct_serve_create(const double hbm_budget_bytes, const int max_active, ct_serve** server)
{
	if (!server || hbm_budget_bytes <= 0 || max_active < 1) return CT_ERR_ARG;
	try {
		std::unique_ptr<ct_serve> h(new ct_serve());
		h->s.reset(new Server());
		Server& S = *h->s;
		S.budget = hbm_budget_bytes;
		S.max_active = max_active;
		const unsigned hc = std::thread::hardware_concurrency();
		S.host_threads = (int)std::max(1u, std::min(16u, hc ? hc : 8u));
		S.t0 = std::chrono::steady_clock::now();
		CTG_HIP(hipStreamCreateWithFlags(&S.st, hipStreamNonBlocking));
		S.hosts.start(S.host_threads);
		S.planners.start(std::min(4, S.host_threads));
		S.worker = std::thread([&S] {
			S.loop();
			(void)hipStreamSynchronize(S.st);
			release_thread_resources();
		});
		*server = h.release();
		return 0;
	}
	catch (const HipError&) {
		return CT_ERR_HIP;
	}
}

int ct_serve_submit(ct_serve* server, const struct ct_serve_request* req, int64_t* ticket)
{
	if (!server || !req || !ticket) return CT_ERR_ARG;
	if (req->model < 0 || req->model > 1 || req->nsites < 2 || req->max_vdim < 1 || req->num_sweeps < 1 || req->maxiter_lanczos < 1) return CT_ERR_ARG;
	Server& S = *server->s;
	auto t = std::make_shared<Ticket>();
	t->req = *req;
	{
		std::lock_guard<std::mutex> g(S.mu);
		t->id = S.next_id++;
		t->t_submit = S.now();
		S.pending.push_back(t);
		S.tickets[t->id] = t;
	}
	S.cv_work.notify_one();
	*ticket = t->id;
	return 0;
}

static void fill_result(const Ticket& t, struct ct_serve_result* res)
{
	res->status = t.status;
	res->sweeps = (int)t.en.size();
	res->energy = t.en.empty() ? std::nan("") : t.en.back();
	res->energy_change = t.en.size() >= 2 ? std::fabs(t.en[t.en.size() - 1] - t.en[t.en.size() - 2]) : std::nan("");
	res->t_submit = t.t_submit;
	res->t_admit = t.t_admit;
	res->t_done = t.t_done;
}

int ct_serve_poll(ct_serve* server, const int64_t ticket, const int wait, struct ct_serve_result* res)
{
	if (!server || !res) return CT_ERR_ARG;
	Server& S = *server->s;
	std::unique_lock<std::mutex> g(S.mu);
	auto it = S.tickets.find(ticket);
	if (it == S.tickets.end()) return CT_ERR_ARG;
	const std::shared_ptr<Ticket> t = it->second;
	if (wait) S.cv_done.wait(g, [&] { return t->status != 1 || S.stop; });
	fill_result(*t, res);
	return t->status == 1 ? 1 : 0;
}

int ct_serve_energies(ct_serve* server, const int64_t ticket, double* en_sweeps, const int capacity)
{
	if (!server) return CT_ERR_ARG;
	Server& S = *server->s;
	std::lock_guard<std::mutex> g(S.mu);
	auto it = S.tickets.find(ticket);
	if (it == S.tickets.end()) return CT_ERR_ARG;
	const Ticket& t = *it->second;
	const int n = (int)t.en.size();
	for (int i = 0; i < std::min(n, capacity); i++) en_sweeps[i] = t.en[i];
	return n;
}

int ct_serve_stats(ct_serve* server, double* stats14)
{
	if (!server || !stats14) return CT_ERR_ARG;
	double* stats13 = stats14;
	double* stats8 = stats13;
	Server& S = *server->s;
	std::lock_guard<std::mutex> g(S.mu);
	stats8[0] = (double)S.ticks;
	stats8[1] = (double)S.lanczos_batches;
	stats8[2] = (double)S.completed;
	stats8[3] = (double)S.pending.size();
	stats8[4] = (double)S.used;
	stats8[5] = (double)S.max_batch;
	stats8[6] = S.busy_s;
	stats8[7] = S.now();
	stats13[8] = S.t_start;
	stats13[9] = S.t_plan;
	stats13[10] = S.t_lanczos;
	stats13[11] = S.t_split_dev;
	stats13[12] = S.t_split_host;
	stats13[13] = S.t_splits;
	return 0;
}

int ct_serve_destroy(ct_serve* server)
{
	if (!server) return CT_ERR_ARG;
	Server& S = *server->s;
	{
		std::lock_guard<std::mutex> g(S.mu);
		S.stop = true;
	}
	S.cv_work.notify_all();
	S.cv_done.notify_all();
	if (S.worker.joinable()) S.worker.join();
	// CT_SERVE_PROFILE=1: the worker's time by part
	const char* prof = getenv("CT_SERVE_PROFILE");
	if (prof && prof[0] == '1')
		fprintf(stderr, "chemtensor_gpu: serve worker: ticks %lld busy %.3f s: batch plans %.3f s (operands %.3f, tables %.3f, items+uploads %.3f, "
			"reduction blocks %.3f), batched Lanczos %.3f s, split hand-off %.3f s; pool: starts %.3f s, splits device %.3f s host %.3f s\n",
			(long long)S.ticks, S.busy_s, S.t_plan, S.t_plan_part[0], S.t_plan_part[1], S.t_plan_part[2], S.t_plan_part[3], S.t_lanczos,
			S.t_splits, S.t_start, S.t_split_dev, S.t_split_host);
	S.planners.shutdown();
	S.hosts.shutdown();
	for (auto& kv : S.tickets) kv.second->state.reset();
	(void)hipStreamSynchronize(S.st);
	(void)hipStreamDestroy(S.st);
	delete server;
	return 0;
}

}  // extern "C"
