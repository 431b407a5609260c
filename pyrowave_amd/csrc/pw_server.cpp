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

struct Lane
{
	hipStream_t stream = nullptr;
	hipEvent_t done = nullptr;
	bool busy = false;
	bool failed = false; // a table build or launch of the batch failed
	std::vector<Request> batch;
	std::vector<Session *> sessions;
	PlanStorage enc_store, dec_store;
	Plan enc_plan, dec_plan;
};

struct Finished
{
	std::vector<uint8_t> bytes;
	std::vector<pyrowave_packet> packets;
};

bool frame_fits(const Geometry &G, const pyrowave_hip_frame &f)
{
	for (int c = 0; c < 3; c++)
	{
		const pyrowave_hip_plane &p = f.planes[c];
		size_t step = p.pixel_stride_in_bytes ? p.pixel_stride_in_bytes : 1;
		if (!p.data || p.row_stride_in_bytes < (size_t(G.plane_w(c)) - 1) * step + 1)
			return false;
	}
	return true;
}

bool target_fits(const Geometry &G, const pyrowave_hip_decode_target &t)
{
	for (int c = 0; c < 3; c++)
		if (!t.planes[c].data || t.planes[c].row_stride_in_bytes < size_t(G.plane_w(c)))
			return false;
	return true;
}

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

	void run();
	void launch(Lane &L);
	void retire(Lane &L);
};

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
		for (size_t i = 0; i < n; i++)
		{
			Slot &S = *enc_slots[i];
			(void)hipEventRecord(S.result_ready, L.stream);
			S.have_result = true;
			S.result_on_host = false;
			ok = ok && queue_result_download(S, L.stream, download_words(L.batch[i].budget));
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
	{
		std::lock_guard<std::recursive_mutex> g(dev->lock);
		for (size_t i = 0; i < L.batch.size(); i++)
		{
			const Request &R = L.batch[i];
			pyrowave_encoder enc = L.sessions[i]->enc;
			enc->slot.result_on_host = device_ok;
			pyrowave_hip_server_completion &c = done[i];
			c.ticket = R.ticket;
			c.session = R.session;
			c.submit_ns = R.submit_ns;
			c.start_ns = R.start_ns;
			c.done_ns = t1;
			c.status = device_ok ? PYROWAVE_SUCCESS : PYROWAVE_ERROR_GENERIC;
			size_t np = 0;
			if (c.status == PYROWAVE_SUCCESS)
				c.status = pyrowave_encoder_compute_num_packets(enc, boundary, &np);
			if (c.status == PYROWAVE_SUCCESS)
			{
				Finished &f = out[i];
				f.packets.resize(np);
				f.bytes.resize(size_t(download_words(R.budget)) * 4 + 8);
				c.status = pyrowave_encoder_packetize(enc, f.packets.data(), boundary, &np, f.bytes.data(),
				                                      f.bytes.size());
				size_t total = 0;
				for (size_t k = 0; k < np; k++)
					total += f.packets[k].size;
				f.bytes.resize(total);
				f.packets.resize(np);
				c.bytes = total;
				c.packets = uint32_t(np);
			}
		}
		dev->resolve_timings(false);
	}
	{
		std::lock_guard<std::mutex> g(m);
		for (size_t i = 0; i < L.batch.size(); i++)
		{
			if (done[i].status == PYROWAVE_SUCCESS)
				results.emplace(done[i].ticket, std::move(out[i]));
			completions.push_back(done[i]);
			L.sessions[i]->busy = false;
		}
		running -= L.batch.size();
	}
	L.busy = false;
	L.batch.clear();
	L.sessions.clear();
	done_cv.notify_all();
}

extern "C" {

PYROWAVE_PUBLIC_API pyrowave_result pyrowave_hip_server_create(const pyrowave_hip_server_create_info *info,
                                                               pyrowave_hip_server *server)
{
	if (!info || !server || !info->device)
		return PYROWAVE_ERROR_INVALID_ARGUMENT;
	if (info->precision != PYROWAVE_HIP_PRECISION_FP16 && info->precision != PYROWAVE_HIP_PRECISION_FP32)
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
	const Finished &f = it->second;
	if (size < f.bytes.size() || (packets && max_packets < f.packets.size()))
		return PYROWAVE_ERROR_INVALID_ARGUMENT;
	if (!f.bytes.empty())
		memcpy(data, f.bytes.data(), f.bytes.size());
	if (packets && !f.packets.empty())
		memcpy(packets, f.packets.data(), f.packets.size() * sizeof(pyrowave_packet));
	server->results.erase(it);
	return PYROWAVE_SUCCESS;
}

} // extern "C"
