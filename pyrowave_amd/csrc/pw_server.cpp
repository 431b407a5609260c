// pw_server.cpp -- continuous-batching serving loop over the codec.
//
// No counterpart in the reference (its callers drive one encoder per stream).
// One worker thread owns K streams ("lanes").  Whenever a lane is free it
// takes the frames that are ready at that moment -- the oldest queued frame of
// every session whose previous frame has finished, oldest submission first, up
// to max_batch -- builds the launch tables for exactly those frames into the
// lane's reusable storage and runs encode (+ result download + decode of the
// encoder's device result) as one launch per stage.  Frames that arrive while
// a batch runs join the next batch on another lane, so batches overlap on the
// device.  Sessions keep their encoder/decoder state resident between frames;
// admission compares a session's device footprint with the remaining budget.

#include "pw_host.hpp"

#include <chrono>
#include <algorithm>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <thread>
#include <unordered_map>

using namespace pw;

namespace
{
uint64_t now_ns()
{
	return uint64_t(std::chrono::duration_cast<std::chrono::nanoseconds>(
	                    std::chrono::steady_clock::now().time_since_epoch())
	                    .count());
}

struct Request
{
	uint64_t ticket = 0;
	uint32_t session = 0;
	pyrowave_hip_frame frame = {};
	size_t budget = 0;
	bool decode = false;
	pyrowave_hip_decode_target target = {};
	uint64_t submit_ns = 0, start_ns = 0;
};

struct Session
{
	uint32_t id = 0;
	pyrowave_encoder enc = nullptr;
	pyrowave_decoder dec = nullptr;
	size_t footprint = 0;
	std::deque<Request> queue;
	bool busy = false; // a frame of this session is on a lane
};

// Pinned host memory one frame's result is downloaded into: the encoder's
// rate-control words (error bits), the packetize info words, the stream
// (sequence header + blocks) and the packet table.
struct ResultBuffer
{
	HostBuffer mem;
	size_t stream_cap = 0; // stream bytes downloaded
	size_t table_off = 0;
	uint32_t max_packets = 0;
	const uint32_t *rc() const { return static_cast<const uint32_t *>(mem.ptr); }
	const uint32_t *info() const { return rc() + kRcCount; }
	const uint8_t *stream() const { return static_cast<const uint8_t *>(mem.ptr) + kStreamOff; }
	const uint32_t *table() const
	{
		return reinterpret_cast<const uint32_t *>(static_cast<const uint8_t *>(mem.ptr) + table_off);
	}
	static constexpr size_t kStreamOff = 64;
};

struct Lane
{
	hipStream_t stream = nullptr;
	hipEvent_t done = nullptr;
	bool busy = false;
	bool failed = false; // a table build or launch of the batch failed
	std::vector<Request> batch;
	std::vector<Session *> sessions;
	std::vector<std::unique_ptr<ResultBuffer>> results;
	PlanStorage enc_store, dec_store;
	Plan enc_plan, dec_plan;
};

struct Finished
{
	std::unique_ptr<ResultBuffer> buf;
	uint32_t bytes = 0, packets = 0;
};

// Bytes of the download that covers a result at this budget (the encoder
// never writes more than the budget's words).
uint32_t download_words(size_t budget)
{
	return uint32_t(target_words_of(budget & ~size_t(3))) + 64u;
}
} // namespace

struct pyrowave_hip_server_opaque
{
	pyrowave_device device = nullptr;
	Device *dev = nullptr;
	size_t budget = 0, used = 0;
	int precision = PYROWAVE_HIP_PRECISION_FP16;
	size_t boundary = 0;
	uint32_t max_batch = 16;

	std::mutex m;
	std::condition_variable work_cv, done_cv;
	std::map<uint32_t, std::unique_ptr<Session>> sessions;
	uint32_t next_session = 1;
	uint64_t next_ticket = 1;
	size_t queued = 0;   // requests in session queues
	size_t running = 0;  // requests on lanes
	std::deque<pyrowave_hip_server_completion> completions;
	std::unordered_map<uint64_t, Finished> results;
	bool stop = false;

	std::vector<std::unique_ptr<Lane>> lanes;
	std::thread worker;

	// Result buffers not in use (shared by the worker and take_bitstream,
	// under `m`).
	std::vector<std::unique_ptr<ResultBuffer>> pool;
	std::unique_ptr<ResultBuffer> acquire(size_t bytes);

	void run();
	void launch(Lane &L);
	void retire(Lane &L);
};

std::unique_ptr<ResultBuffer> pyrowave_hip_server_opaque::acquire(size_t bytes)
{
	{
		std::lock_guard<std::mutex> g(m);
		size_t best = pool.size();
		for (size_t i = 0; i < pool.size(); i++)
			if (pool[i]->mem.size >= bytes && (best == pool.size() || pool[i]->mem.size < pool[best]->mem.size))
				best = i;
		if (best != pool.size())
		{
			std::unique_ptr<ResultBuffer> b = std::move(pool[best]);
			pool.erase(pool.begin() + long(best));
			return b;
		}
	}
	std::unique_ptr<ResultBuffer> b(new ResultBuffer);
	if (!b->mem.alloc(bytes + bytes / 4 + 4096))
		return nullptr;
	return b;
}

// Worker: retire finished lanes, fill free lanes, sleep briefly while the
// device works.  Only this thread touches lanes; sessions are shared with the
// API under `m`.
void pyrowave_hip_server_opaque::run()
{
	(void)hipSetDevice(dev->hip_device);
	for (;;)
	{
		bool any_busy = false;
		for (auto &lane : lanes)
		{
			if (lane->busy && hipEventQuery(lane->done) != hipErrorNotReady)
				retire(*lane);
			any_busy |= lane->busy;
		}
		for (auto &lane : lanes)
		{
			if (lane->busy)
				continue;
			{
				std::lock_guard<std::mutex> g(m);
				// The oldest queued frame of every idle session, oldest first.
				std::vector<Session *> ready;
				for (auto &kv : sessions)
					if (!kv.second->busy && !kv.second->queue.empty())
						ready.push_back(kv.second.get());
				if (ready.empty())
					break;
				std::sort(ready.begin(), ready.end(), [](const Session *a, const Session *b) {
					return a->queue.front().ticket < b->queue.front().ticket;
				});
				if (ready.size() > max_batch)
					ready.resize(max_batch);
				lane->batch.clear();
				lane->sessions = ready;
				for (Session *s : ready)
				{
					lane->batch.push_back(s->queue.front());
					s->queue.pop_front();
					s->busy = true;
				}
				queued -= ready.size();
				running += ready.size();
			}
			launch(*lane);
			any_busy = true;
		}
		std::unique_lock<std::mutex> lk(m);
		if (stop && queued == 0 && running == 0)
			break;
		if (any_busy)
			work_cv.wait_for(lk, std::chrono::microseconds(20));
		else if (queued == 0 && !stop)
			work_cv.wait(lk, [&] { return stop || queued != 0; });
	}
}

void pyrowave_hip_server_opaque::launch(Lane &L)
{
	const uint64_t t0 = now_ns();
	const size_t n = L.batch.size();
	std::vector<Slot *> enc_slots(n), dec_slots;
	std::vector<size_t> dec_index;
	for (size_t i = 0; i < n; i++)
	{
		L.batch[i].start_ns = t0;
		enc_slots[i] = &L.sessions[i]->enc->slot;
		if (L.batch[i].decode)
		{
			dec_slots.push_back(&L.sessions[i]->dec->slot);
			dec_index.push_back(i);
		}
	}
	std::lock_guard<std::recursive_mutex> g(dev->lock);
	bool ok = L.enc_plan.build(enc_slots, true, &L.enc_store, L.stream);
	if (ok)
	{
		for (size_t i = 0; i < n; i++)
			fill_encode_call(L.enc_plan.call(int(i)), *enc_slots[i], L.batch[i].frame, L.batch[i].budget);
		run_encode(L.enc_plan, L.stream);
		// Packetize on the device, then one download per part straight into
		// the frame's result buffer: no host walk or copy of the blocks.
		std::vector<PacketizeJob> jobs(n);
		L.results.clear();
		L.results.resize(n);
		for (size_t i = 0; i < n && ok; i++)
		{
			Slot &S = *enc_slots[i];
			const Geometry &G = S.geom;
			S.have_result = true;
			S.result_on_host = false;
			PacketizeJob &J = jobs[i];
			J.meta = S.fs.meta;
			J.stream = S.fs.bitstream - 2;
			J.table = S.dtable;
			J.info = S.dpinfo;
			J.n32 = uint32_t(G.count32);
			J.h0 = uint32_t(G.width - 1) | (uint32_t(G.height - 1) << 14) | ((S.sequence & 7u) << 28) | (1u << 31);
			J.h1 = uint32_t(G.chroma == PYROWAVE_CHROMA_SUBSAMPLING_444) << 26;
			J.boundary = uint32_t(std::min<size_t>(boundary, UINT32_MAX));
			// Stream bytes are at most 8 + the budget's words; two adjacent
			// packets together exceed the boundary, so the count is bounded.
			const size_t stream_cap = std::min(size_t(download_words(L.batch[i].budget)) * 4 + 8,
			                                   (size_t(S.fs.capacity_words) + 2) * 4);
			const uint64_t by_size = 2 * uint64_t(stream_cap) / std::max<uint64_t>(J.boundary, 1) + 3;
			const uint32_t maxpk = uint32_t(std::min<uint64_t>(by_size, uint64_t(G.count32) + 1));
			const size_t table_off = ResultBuffer::kStreamOff + ((stream_cap + 63) & ~size_t(63));
			std::unique_ptr<ResultBuffer> rb = acquire(table_off + size_t(maxpk) * 8);
			if (!rb)
			{
				ok = false;
				break;
			}
			rb->stream_cap = stream_cap;
			rb->table_off = table_off;
			rb->max_packets = maxpk;
			L.results[i] = std::move(rb);
		}
		if (ok)
		{
			launch_packetize(jobs.data(), int(n), L.stream);
			for (size_t i = 0; i < n; i++)
			{
				Slot &S = *enc_slots[i];
				ResultBuffer &rb = *L.results[i];
				uint8_t *h = static_cast<uint8_t *>(rb.mem.ptr);
				(void)hipEventRecord(S.result_ready, L.stream);
				(void)hipMemcpyAsync(h, S.fs.rc, kRcCount * 4, hipMemcpyDeviceToHost, L.stream);
				(void)hipMemcpyAsync(h + kRcCount * 4, S.dpinfo, 8, hipMemcpyDeviceToHost, L.stream);
				(void)hipMemcpyAsync(h + ResultBuffer::kStreamOff, S.fs.bitstream - 2, rb.stream_cap,
				                     hipMemcpyDeviceToHost, L.stream);
				(void)hipMemcpyAsync(h + rb.table_off, S.dtable, size_t(rb.max_packets) * 8, hipMemcpyDeviceToHost,
				                     L.stream);
			}
		}
	}
	if (ok && !dec_slots.empty())
	{
		ok = L.dec_plan.build(dec_slots, false, &L.dec_store, L.stream);
		if (ok)
		{
			for (size_t k = 0; k < dec_slots.size(); k++)
			{
				const size_t i = dec_index[k];
				fill_decode_call(L.dec_plan.call(int(k)), *dec_slots[k], enc_slots[i], L.batch[i].target);
			}
			run_decode(L.dec_plan, L.stream);
		}
	}
	ok = hipEventRecord(L.done, L.stream) == hipSuccess && ok;
	if (!ok)
		(void)hipStreamSynchronize(L.stream);
	L.failed = !ok;
	L.busy = true;
}

void pyrowave_hip_server_opaque::retire(Lane &L)
{
	const bool device_ok = hipEventSynchronize(L.done) == hipSuccess && !L.failed;
	const uint64_t t1 = now_ns();
	std::vector<pyrowave_hip_server_completion> done(L.batch.size());
	std::vector<Finished> out(L.batch.size());
	for (size_t i = 0; i < L.batch.size(); i++)
	{
		const Request &R = L.batch[i];
		pyrowave_hip_server_completion &c = done[i];
		c.ticket = R.ticket;
		c.session = R.session;
		c.submit_ns = R.submit_ns;
		c.start_ns = R.start_ns;
		c.done_ns = t1;
		c.status = device_ok && i < L.results.size() && L.results[i] ? PYROWAVE_SUCCESS : PYROWAVE_ERROR_GENERIC;
		if (c.status != PYROWAVE_SUCCESS)
			continue;
		ResultBuffer &rb = *L.results[i];
		const uint32_t bytes = rb.info()[0], packets = rb.info()[1];
		// Device error bits, or a result the download did not cover.
		if (rb.rc()[kRcError] != 0 || bytes > rb.stream_cap || packets == 0 ||
		    packets > rb.max_packets)
		{
			c.status = PYROWAVE_ERROR_GENERIC;
			continue;
		}
		c.bytes = bytes;
		c.packets = packets;
		out[i].bytes = bytes;
		out[i].packets = packets;
		out[i].buf = std::move(L.results[i]);
	}
	{
		std::lock_guard<std::recursive_mutex> g(dev->lock);
		dev->resolve_timings(false);
	}
	{
		std::lock_guard<std::mutex> g(m);
		for (size_t i = 0; i < L.batch.size(); i++)
		{
			if (done[i].status == PYROWAVE_SUCCESS)
				results.emplace(done[i].ticket, std::move(out[i]));
			else if (i < L.results.size() && L.results[i])
				pool.push_back(std::move(L.results[i]));
			completions.push_back(done[i]);
			L.sessions[i]->busy = false;
		}
		running -= L.batch.size();
	}
	L.busy = false;
	L.batch.clear();
	L.sessions.clear();
	L.results.clear();
	done_cv.notify_all();
}

extern "C" {

PYROWAVE_PUBLIC_API pyrowave_result pyrowave_hip_server_create(const pyrowave_hip_server_create_info *info,
                                                               pyrowave_hip_server *server)
{
	if (!info || !server || !info->device)
		return PYROWAVE_ERROR_INVALID_ARGUMENT;
	if (!valid_precision(info->precision))
		return PYROWAVE_ERROR_INVALID_ARGUMENT;
	std::unique_ptr<pyrowave_hip_server_opaque> S(new pyrowave_hip_server_opaque);
	S->device = info->device;
	S->dev = &info->device->dev;
	S->budget = info->hbm_budget_bytes;
	S->precision = info->precision;
	S->boundary = info->packet_boundary ? info->packet_boundary : (size_t(1) << 40);
	S->max_batch = info->max_batch_frames ? info->max_batch_frames : 16u;
	const uint32_t nlanes = info->streams ? std::min(info->streams, 16u) : 4u;
	(void)hipSetDevice(S->dev->hip_device);
	for (uint32_t i = 0; i < nlanes; i++)
	{
		std::unique_ptr<Lane> L(new Lane);
		if (hipStreamCreateWithFlags(&L->stream, hipStreamNonBlocking) != hipSuccess ||
		    hipEventCreateWithFlags(&L->done, hipEventDisableTiming) != hipSuccess)
			return PYROWAVE_ERROR_GENERIC;
		S->lanes.push_back(std::move(L));
	}
	S->dev->live_objects++;
	pyrowave_hip_server_opaque *raw = S.release();
	raw->worker = std::thread([raw] { raw->run(); });
	*server = raw;
	return PYROWAVE_SUCCESS;
}

PYROWAVE_PUBLIC_API void pyrowave_hip_server_destroy(pyrowave_hip_server server)
{
	if (!server)
		return;
	{
		std::lock_guard<std::mutex> g(server->m);
		server->stop = true;
	}
	server->work_cv.notify_all();
	if (server->worker.joinable())
		server->worker.join();
	for (auto &kv : server->sessions)
	{
		pyrowave_encoder_destroy(kv.second->enc);
		pyrowave_decoder_destroy(kv.second->dec);
	}
	server->sessions.clear();
	(void)hipSetDevice(server->dev->hip_device);
	for (auto &L : server->lanes)
	{
		(void)hipStreamSynchronize(L->stream);
		L->enc_store.dev.release();
		L->dec_store.dev.release();
		(void)hipEventDestroy(L->done);
		(void)hipStreamDestroy(L->stream);
	}
	server->lanes.clear();
	server->dev->live_objects--;
	delete server;
}

PYROWAVE_PUBLIC_API pyrowave_result pyrowave_hip_server_session_footprint(pyrowave_hip_server server, int width,
                                                                          int height,
                                                                          pyrowave_chroma_subsampling chroma,
                                                                          size_t *bytes)
{
	if (!server || !bytes)
		return PYROWAVE_ERROR_INVALID_ARGUMENT;
	if (chroma == PYROWAVE_CHROMA_SUBSAMPLING_420 && ((width & 1) || (height & 1)))
		return PYROWAVE_ERROR_INVALID_ARGUMENT;
	Geometry G;
	if (!G.init(width, height, int(chroma), server->precision))
		return PYROWAVE_ERROR_INVALID_ARGUMENT;
	*bytes = slot_layout(G, true).total + slot_layout(G, false).total;
	return PYROWAVE_SUCCESS;
}

PYROWAVE_PUBLIC_API pyrowave_result pyrowave_hip_server_open_session(pyrowave_hip_server server, int width, int height,
                                                                     pyrowave_chroma_subsampling chroma,
                                                                     uint32_t *session)
{
	if (!server || !session)
		return PYROWAVE_ERROR_INVALID_ARGUMENT;
	size_t need = 0;
	pyrowave_result r = pyrowave_hip_server_session_footprint(server, width, height, chroma, &need);
	if (r != PYROWAVE_SUCCESS)
		return r;
	{
		// Reserve the footprint first so concurrent opens cannot overcommit.
		std::lock_guard<std::mutex> g(server->m);
		if (server->used + need > server->budget)
			return PYROWAVE_ERROR_OUT_OF_DEVICE_MEMORY;
		server->used += need;
	}
	std::unique_ptr<Session> s(new Session);
	s->footprint = need;
	pyrowave_hip_encoder_create_info ei = {server->device, width, height, chroma,
	                                       pyrowave_hip_precision(server->precision), 0};
	pyrowave_hip_decoder_create_info di = {server->device, width, height, chroma,
	                                       pyrowave_hip_precision(server->precision)};
	r = pyrowave_hip_encoder_create(&ei, &s->enc);
	if (r == PYROWAVE_SUCCESS)
		r = pyrowave_hip_decoder_create(&di, &s->dec);
	std::lock_guard<std::mutex> g(server->m);
	if (r != PYROWAVE_SUCCESS)
	{
		server->used -= need;
		if (s->enc)
			pyrowave_encoder_destroy(s->enc);
		return r;
	}
	s->id = server->next_session++;
	*session = s->id;
	server->sessions.emplace(s->id, std::move(s));
	return PYROWAVE_SUCCESS;
}

PYROWAVE_PUBLIC_API pyrowave_result pyrowave_hip_server_close_session(pyrowave_hip_server server, uint32_t session)
{
	if (!server)
		return PYROWAVE_ERROR_INVALID_ARGUMENT;
	std::unique_ptr<Session> s;
	{
		std::unique_lock<std::mutex> lk(server->m);
		auto it = server->sessions.find(session);
		if (it == server->sessions.end())
			return PYROWAVE_ERROR_INVALID_ARGUMENT;
		Session *p = it->second.get();
		server->done_cv.wait(lk, [&] { return p->queue.empty() && !p->busy; });
		it = server->sessions.find(session);
		s = std::move(it->second);
		server->sessions.erase(it);
		server->used -= s->footprint;
	}
	pyrowave_encoder_destroy(s->enc);
	pyrowave_decoder_destroy(s->dec);
	return PYROWAVE_SUCCESS;
}

PYROWAVE_PUBLIC_API pyrowave_result pyrowave_hip_server_submit(pyrowave_hip_server server, uint32_t session,
                                                               const pyrowave_hip_frame *frame,
                                                               size_t maximum_bitstream_size,
                                                               const pyrowave_hip_decode_target *target,
                                                               uint64_t *ticket)
{
	if (!server || !frame || !ticket)
		return PYROWAVE_ERROR_INVALID_ARGUMENT;
	const size_t words = maximum_bitstream_size & ~size_t(3);
	if (words == 0 || words > UINT32_MAX)
		return PYROWAVE_ERROR_INVALID_ARGUMENT;
	{
		std::lock_guard<std::mutex> g(server->m);
		auto it = server->sessions.find(session);
		if (it == server->sessions.end() || server->stop)
			return PYROWAVE_ERROR_INVALID_ARGUMENT;
		Session &s = *it->second;
		const Geometry &G = s.enc->slot.geom;
		if (!frame_fits(G, *frame) || (target && !target_fits(G, *target)))
			return PYROWAVE_ERROR_INVALID_ARGUMENT;
		Request R;
		R.ticket = server->next_ticket++;
		R.session = session;
		R.frame = *frame;
		R.budget = maximum_bitstream_size;
		R.decode = target != nullptr;
		if (target)
			R.target = *target;
		R.submit_ns = now_ns();
		s.queue.push_back(R);
		server->queued++;
		*ticket = R.ticket;
	}
	server->work_cv.notify_one();
	return PYROWAVE_SUCCESS;
}

PYROWAVE_PUBLIC_API pyrowave_result pyrowave_hip_server_poll(pyrowave_hip_server server,
                                                             pyrowave_hip_server_completion *completions, uint32_t max,
                                                             uint32_t *count, uint32_t timeout_us)
{
	if (!server || !count || (max && !completions))
		return PYROWAVE_ERROR_INVALID_ARGUMENT;
	std::unique_lock<std::mutex> lk(server->m);
	if (server->completions.empty() && timeout_us)
		server->done_cv.wait_for(lk, std::chrono::microseconds(timeout_us),
		                         [&] { return !server->completions.empty(); });
	uint32_t n = 0;
	while (n < max && !server->completions.empty())
	{
		completions[n++] = server->completions.front();
		server->completions.pop_front();
	}
	*count = n;
	return PYROWAVE_SUCCESS;
}

PYROWAVE_PUBLIC_API pyrowave_result pyrowave_hip_server_take_bitstream(pyrowave_hip_server server, uint64_t ticket,
                                                                       void *data, size_t size,
                                                                       pyrowave_packet *packets, uint32_t max_packets)
{
	if (!server || !data)
		return PYROWAVE_ERROR_INVALID_ARGUMENT;
	std::lock_guard<std::mutex> g(server->m);
	auto it = server->results.find(ticket);
	if (it == server->results.end())
		return PYROWAVE_ERROR_INVALID_ARGUMENT;
	Finished &f = it->second;
	if (size < f.bytes || (packets && max_packets < f.packets))
		return PYROWAVE_ERROR_INVALID_ARGUMENT;
	memcpy(data, f.buf->stream(), f.bytes);
	if (packets)
	{
		const uint32_t *t = f.buf->table();
		for (uint32_t i = 0; i < f.packets; i++)
			packets[i] = {t[2 * i], t[2 * i + 1]};
	}
	server->pool.push_back(std::move(f.buf));
	server->results.erase(it);
	return PYROWAVE_SUCCESS;
}

} // extern "C"
