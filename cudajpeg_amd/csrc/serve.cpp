// Serving loop of the MI355X baseline-JPEG decoder (C ABI in
// include/cudajpeg_amd.h).  No counterpart in the reference, whose host
// program decodes one file per openJPG call (src/main.cu:18-50).
//
// Threads: callers submit / poll / release; a batcher thread forms batches
// (flush on max_batch waiting requests or on the oldest request's deadline,
// admission by decoded-frame bytes), stages them with the parallel host
// planner and launches them on the stream of a free batch slot; a completion
// thread waits for the slots in launch order and publishes their frames.  A
// slot is reused once every frame it holds has been released, so a frame's
// device memory is stable from completion to release.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/cudajpeg_amd.h"
#include "engine_internal.h"

namespace {

int64_t mono_ns() {
  return std::chrono::duration_cast<std::chrono::nanoseconds>(std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

struct Request {
  uint64_t id = 0;
  std::vector<uint8_t> data;
  uint64_t frame_bytes = 0;  // W * H * channels, from the frame header
  int width = 0, height = 0, channels = 0;
  int64_t t_arrival = 0, t_launch = 0, t_ready = 0;
  int error = 0;
  int slot = -1;  // -1: never admitted
  int batch_size = 0;
  bool done = false;
  void* frame = nullptr;
};

struct Slot {
  cjd_batch* b = nullptr;
  hipStream_t stream = nullptr;
  std::vector<Request*> reqs;
  int held = 0;       // completed requests of this slot not yet released
  bool busy = false;  // staged / in flight / holding frames
  int64_t redone_seen = 0;  // cjd_batch_fallbacks of b after its last finish
  int64_t protocol_seen = 0;  // cjd_batch_protocol_redos of b after its last finish
};

}  // namespace

struct cjd_server {
  cjd_context* ctx = nullptr;
  cjd_server_options opt{};
  std::mutex mu;
  std::condition_variable cv_work;      // batcher: requests, a freed slot or budget, stop
  std::condition_variable cv_inflight;  // completer: a launched slot, stop
  std::condition_variable cv_done;      // pollers: completions
  std::deque<Request*> pending;
  std::deque<int> inflight;  // launched slots, in launch order
  std::deque<Request*> completed;
  std::unordered_map<uint64_t, Request*> live;  // submitted and not released
  std::vector<Slot> slots;
  uint64_t next_id = 1;
  uint64_t admitted_bytes = 0;
  int64_t n_submitted = 0, n_completed = 0, n_redone = 0;
  int64_t n_batches = 0, t_stage = 0, t_launch = 0, t_finish = 0, t_slot_wait = 0, t_finish_max = 0,
          n_protocol = 0;
  int64_t t_device = 0, t_device_max = 0, t_turnaround = 0;
  bool stop = false;
  std::thread batcher, completer;

  void complete_locked(Request* r) {
    ++n_completed;
    r->done = true;
    if (!r->t_ready) r->t_ready = mono_ns();
    completed.push_back(r);
    cv_done.notify_all();
  }

  void batcher_loop() {
    (void)hipSetDevice(ctx->device);
    const int max_batch = opt.max_batch;
    const int64_t max_wait = (int64_t)opt.max_wait_us * 1000;
    std::unique_lock<std::mutex> lk(mu);
    for (;;) {
      cv_work.wait(lk, [&] { return stop || !pending.empty(); });
      if (stop) return;
      // flush on a full batch or on the oldest request's deadline
      const int64_t deadline = pending.front()->t_arrival + max_wait;
      while (!stop && (int)pending.size() < max_batch) {
        const int64_t now = mono_ns();
        if (now >= deadline) break;
        cv_work.wait_for(lk, std::chrono::nanoseconds(deadline - now));
      }
      if (stop) return;
      int si = -1;
      for (size_t k = 0; k < slots.size(); ++k)
        if (!slots[k].busy) {
          si = (int)k;
          break;
        }
      if (si < 0) {  // every slot in flight or holding frames
        const int64_t t0 = mono_ns();
        cv_work.wait(lk);
        t_slot_wait += mono_ns() - t0;
        continue;
      }
      // admission by decoded-frame bytes, in arrival order
      std::vector<Request*> take;
      while (!pending.empty() && (int)take.size() < max_batch) {
        Request* r = pending.front();
        if (admitted_bytes + r->frame_bytes > opt.frame_budget) {
          if (admitted_bytes == 0 && take.empty()) {  // can never fit
            pending.pop_front();
            r->error = CJD_OOM_ERROR;
            complete_locked(r);
            continue;
          }
          break;
        }
        admitted_bytes += r->frame_bytes;
        r->slot = si;
        take.push_back(r);
        pending.pop_front();
      }
      if (take.empty()) {  // wait for frames to be released
        if (!pending.empty()) cv_work.wait(lk);
        continue;
      }
      Slot& sl = slots[(size_t)si];
      sl.busy = true;
      sl.reqs = take;
      lk.unlock();
      const int n = (int)take.size();
      std::vector<const uint8_t*> files((size_t)n);
      std::vector<size_t> sizes((size_t)n);
      std::vector<int> errs((size_t)n, 0);
      for (int i = 0; i < n; ++i) {
        files[(size_t)i] = take[(size_t)i]->data.data();
        sizes[(size_t)i] = take[(size_t)i]->data.size();
      }
      const int64_t ts = mono_ns();
      int e = cjd_batch_stage_async(sl.b, n, files.data(), sizes.data(), errs.data(), sl.stream);
      const int64_t tl = mono_ns();
      if (!e) e = cjd_batch_decode(sl.b, sl.stream);
      const int64_t te = mono_ns();
      for (int i = 0; i < n; ++i) {
        Request* r = take[(size_t)i];
        r->t_launch = tl;
        r->batch_size = n;
        r->error = e ? e : errs[(size_t)i];
      }
      lk.lock();
      ++n_batches;
      t_stage += tl - ts;
      t_launch += te - tl;
      inflight.push_back(si);
      cv_inflight.notify_one();
    }
  }

  // Finishes the launched slots in launch order.
  void completer_loop() {
    (void)hipSetDevice(ctx->device);
    std::unique_lock<std::mutex> lk(mu);
    for (;;) {
      cv_inflight.wait(lk, [&] { return stop || !inflight.empty(); });
      if (inflight.empty()) return;  // stopping and drained
      const int si = inflight.front();
      inflight.pop_front();
      Slot& sl = slots[(size_t)si];
      lk.unlock();
      const int n = (int)sl.reqs.size();
      std::vector<int> errs((size_t)n, 0);
      const int64_t tf = mono_ns();
      const int e = cjd_batch_finish(sl.b, sl.stream, errs.data());
      const int64_t tr = mono_ns();
      const int64_t redone = cjd_batch_fallbacks(sl.b);  // read by the thread that owns the batch now
      const int64_t protocol = cjd_batch_protocol_redos(sl.b);
      float ms_ent = 0.f, ms_pix = 0.f, ms_all = 0.f;
      if (!e) (void)cjd_batch_timings(sl.b, &ms_ent, &ms_pix, &ms_all);
      const int64_t dev_ns = (int64_t)((double)ms_all * 1e6);
      const int64_t t_launched = n ? sl.reqs[0]->t_launch : tr;
      for (int i = 0; i < n; ++i) {
        Request* r = sl.reqs[(size_t)i];
        r->t_ready = tr;
        if (!r->error) r->error = e ? e : errs[(size_t)i];
        if (!r->error) {
          size_t bytes = 0;
          if (cjd_batch_device_ptr(sl.b, i, CJD_OUT_PIXELS, &r->frame, &bytes)) r->frame = nullptr;
        }
      }
      lk.lock();
      t_finish += tr - tf;
      t_finish_max = std::max(t_finish_max, tr - tf);
      n_protocol += protocol - sl.protocol_seen;
      sl.protocol_seen = protocol;
      t_device += dev_ns;
      t_device_max = std::max(t_device_max, dev_ns);
      t_turnaround += tr - t_launched;
      n_redone += redone - sl.redone_seen;
      sl.redone_seen = redone;
      sl.held = n;
      for (Request* r : sl.reqs) complete_locked(r);
    }
  }
};

extern "C" int64_t cjd_server_now_ns(void) { return mono_ns(); }

extern "C" int cjd_server_create(cjd_context* ctx, const cjd_server_options* opt, cjd_server** out) {
  if (!ctx || !out) return CJD_PROGRAMMER_ERROR;
  if (hipSetDevice(ctx->device) != hipSuccess) return CJD_DEVICE_MEM_ERROR;
  auto* s = new (std::nothrow) cjd_server;
  if (!s) return CJD_OOM_ERROR;
  s->ctx = ctx;
  if (opt) s->opt = *opt;
  if (s->opt.max_batch <= 0) s->opt.max_batch = 64;
  if (s->opt.max_wait_us <= 0) s->opt.max_wait_us = 2000;
  if (s->opt.frame_budget == 0) s->opt.frame_budget = 64ull << 30;
  if (s->opt.slots <= 0) s->opt.slots = 4;
  s->slots.resize((size_t)s->opt.slots);
  for (Slot& sl : s->slots) {
    int e = cjd_batch_create(ctx, &sl.b);
    if (!e) e = cjd_batch_set_options(sl.b, s->opt.idct_mode, CJD_ENTROPY_TILED, s->opt.tile_bits, 0);
    if (!e) e = cjd_batch_set_extensions(sl.b, s->opt.extensions);
    if (!e && hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking) != hipSuccess) e = CJD_DEVICE_MEM_ERROR;
    if (e) {
      for (Slot& t : s->slots) {
        if (t.stream) (void)hipStreamDestroy(t.stream);
        if (t.b) cjd_batch_destroy(t.b);
      }
      delete s;
      return e;
    }
  }
  s->batcher = std::thread([s] { s->batcher_loop(); });
  s->completer = std::thread([s] { s->completer_loop(); });
  *out = s;
  return 0;
}

extern "C" int cjd_server_submit(cjd_server* s, const uint8_t* data, size_t size, uint64_t* id) {
  if (!s || !data || !id) return CJD_PROGRAMMER_ERROR;
  cjd_frame_info f{};
  const int e = cjd_parse_ext(data, size, s->opt.extensions, &f);
  if (e) return e;
  auto* r = new (std::nothrow) Request;
  if (!r) return CJD_OOM_ERROR;
  r->data.assign(data, data + size);
  r->width = f.width;
  r->height = f.height;
  r->channels = f.num_channels == 3 ? 3 : 1;
  r->frame_bytes = (uint64_t)f.width * f.height * r->channels;
  r->t_arrival = mono_ns();
  std::lock_guard<std::mutex> lk(s->mu);
  r->id = s->next_id++;
  ++s->n_submitted;
  s->live[r->id] = r;
  s->pending.push_back(r);
  s->cv_work.notify_one();
  *id = r->id;
  return 0;
}

extern "C" int cjd_server_poll(cjd_server* s, cjd_completion* out, int max, int wait_us) {
  if (!s || !out || max <= 0) return -CJD_PROGRAMMER_ERROR;
  std::unique_lock<std::mutex> lk(s->mu);
  if (s->completed.empty() && wait_us > 0)
    s->cv_done.wait_for(lk, std::chrono::microseconds(wait_us), [&] { return !s->completed.empty(); });
  int n = 0;
  while (n < max && !s->completed.empty()) {
    Request* r = s->completed.front();
    s->completed.pop_front();
    cjd_completion& c = out[n++];
    c.id = r->id;
    c.error = r->error;
    c.batch_size = r->batch_size;
    c.frame = r->error ? nullptr : r->frame;
    c.frame_bytes = r->error ? 0 : r->frame_bytes;
    c.width = r->width;
    c.height = r->height;
    c.channels = r->channels;
    c.t_arrival_ns = r->t_arrival;
    c.t_launch_ns = r->t_launch;
    c.t_ready_ns = r->t_ready;
  }
  return n;
}

extern "C" int cjd_server_stats(cjd_server* s, cjd_server_counters* out) {
  if (!s || !out) return CJD_PROGRAMMER_ERROR;
  std::lock_guard<std::mutex> lk(s->mu);
  out->submitted = s->n_submitted;
  out->completed = s->n_completed;
  out->pending = (int64_t)s->pending.size();
  out->inflight = (int64_t)s->inflight.size();
  out->redone = s->n_redone;
  out->batches = s->n_batches;
  out->stage_ns = s->t_stage;
  out->launch_ns = s->t_launch;
  out->finish_ns = s->t_finish;
  out->slot_wait_ns = s->t_slot_wait;
  out->device_ns = s->t_device;
  out->device_max_ns = s->t_device_max;
  out->turnaround_ns = s->t_turnaround;
  out->finish_max_ns = s->t_finish_max;
  out->protocol_redos = s->n_protocol;
  return 0;
}

extern "C" int cjd_server_release(cjd_server* s, uint64_t id) {
  if (!s) return CJD_PROGRAMMER_ERROR;
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->live.find(id);
  if (it == s->live.end() || !it->second->done) return CJD_PROGRAMMER_ERROR;
  Request* r = it->second;
  s->live.erase(it);
  if (r->slot >= 0) {
    s->admitted_bytes -= r->frame_bytes;
    Slot& sl = s->slots[(size_t)r->slot];
    if (--sl.held == 0) {
      sl.busy = false;
      sl.reqs.clear();
    }
    s->cv_work.notify_one();
  }
  delete r;
  return 0;
}

extern "C" void cjd_server_destroy(cjd_server* s) {
  if (!s) return;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
  }
  s->cv_work.notify_all();
  if (s->batcher.joinable()) s->batcher.join();
  s->cv_inflight.notify_all();
  if (s->completer.joinable()) s->completer.join();
  for (Slot& sl : s->slots) {
    if (sl.stream) (void)hipStreamDestroy(sl.stream);
    if (sl.b) cjd_batch_destroy(sl.b);
  }
  for (auto& kv : s->live) delete kv.second;
  delete s;
}

extern "C" int cjd_server_replay(cjd_server* s, int n, const uint8_t* const* files, const size_t* sizes,
                                 const int64_t* arrival_ns, int64_t* latency_ns, int* errors, int64_t* wall_ns) {
  if (!s || n < 0 || (n && (!files || !sizes || !arrival_ns || !latency_ns || !errors))) return CJD_PROGRAMMER_ERROR;
  std::mutex mu;
  std::unordered_map<uint64_t, int> index;
  std::atomic<int> rejected{0};
  const int64_t start = mono_ns();
  std::thread submitter([&] {
    for (int k = 0; k < n; ++k) {
      const int64_t target = start + arrival_ns[k];
      for (int64_t now = mono_ns(); now < target; now = mono_ns())
        if (target - now > 200000) std::this_thread::sleep_for(std::chrono::nanoseconds(target - now - 100000));
      uint64_t id = 0;
      // not under the index lock: a submit parses and copies the file, and
      // the completion loop below must not wait for it to release frames
      const int e = cjd_server_submit(s, files[k], sizes[k], &id);
      if (e) {
        errors[k] = e;
        latency_ns[k] = 0;
        rejected.fetch_add(1);
      } else {
        std::lock_guard<std::mutex> lk(mu);
        index[id] = k;
      }
    }
  });
  int done = 0;
  int64_t last = start;
  std::vector<cjd_completion> buf(256);
  while (done + rejected.load() < n) {
    const int m = cjd_server_poll(s, buf.data(), (int)buf.size(), 1000);
    for (int i = 0; i < m; ++i) {
      const cjd_completion& c = buf[(size_t)i];
      // the id is indexed right after its submit returns; a completion can
      // be polled just before that
      int k = -1;
      while (k < 0) {
        {
          std::lock_guard<std::mutex> lk(mu);
          auto it = index.find(c.id);
          if (it != index.end()) {
            k = it->second;
            index.erase(it);
          }
        }
        if (k < 0) std::this_thread::yield();
      }
      latency_ns[k] = c.t_ready_ns - c.t_arrival_ns;
      errors[k] = c.error;
      last = std::max(last, c.t_ready_ns);
      (void)cjd_server_release(s, c.id);
      ++done;
    }
  }
  submitter.join();
  if (wall_ns) *wall_ns = last - start;
  return 0;
}
