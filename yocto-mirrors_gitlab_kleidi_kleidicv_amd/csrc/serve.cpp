// Serving engine: continuous batching of single-image requests on one GPU.
//
// Requests wait in arrival order for HBM budget (FIFO: the head blocks later
// requests), take their outputs from a device ring, and are launched on
// stream slots:
//   - large requests (>= large_px output pixels) fill the GPU alone, so they
//     go one at a time through their single-image entry point on slot 0, in
//     arrival order; when more of them wait than slot 0 queues (the device
//     is busy and the host sets the pace), up to large_batch share one batch
//     launch;
//   - small requests are packed into one-shot batch launches
//     (kleidicv_hip_batch_launch) on the other slots, so they never wait
//     behind a large one; slot 0 takes them too when no large request waits,
//     and a lone small request runs alone (cheaper than a batch).
// Each slot keeps up to `depth` launches queued on its stream, so the device
// does not wait for the host between them, and slots overlap on the device.
// A slot's work lists are uploaded on its own stream, so a queued batch can
// reuse the slot's workspace: stream order keeps the upload behind the
// kernels that read the previous lists.
//
// The same scheduler runs against a host stand-in (standin_polls > 0) in
// which nothing touches a device, so the policy is tested without a GPU.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <deque>
#include <new>
#include <unordered_map>
#include <vector>

#include "kleidicv_hip/kleidicv.h"
#include "prepare.h"

namespace {

using kcv::from_hip;

constexpr size_t kRingAlign = 256;
constexpr size_t kStandinBytesPerOp = 1024;

// Byte ranges of one device allocation handed out as a ring.  Blocks are
// taken at the head in order and released in any order; the oldest blocks'
// space returns once they and every block before them are released.
class Ring {
 public:
  void init(size_t size) { size_ = size; }
  size_t size() const { return size_; }

  // Offset of a new block (its sequence number in *seq), or false while the
  // ring is full.
  bool take(size_t nbytes, size_t *off, uint64_t *seq) {
    const size_t n = (nbytes + kRingAlign - 1) / kRingAlign * kRingAlign;
    if (n > size_) return false;
    if (blocks_.empty()) head_ = 0;
    const size_t tail = blocks_.empty() ? 0 : blocks_.front().off;
    size_t o;
    if (blocks_.empty() || head_ > tail) {  // free: [head, size) and [0, tail)
      if (size_ - head_ >= n) {
        o = head_;
      } else if (tail > n) {  // wrap; the end of the ring idles
        o = 0;
      } else {
        return false;
      }
    } else if (tail - head_ > n) {  // wrapped; free: [head, tail)
      o = head_;
    } else {
      return false;
    }
    blocks_.push_back(Block{o, n, false});
    head_ = o + n;
    *off = o;
    *seq = first_seq_ + blocks_.size() - 1;
    return true;
  }

  void release(uint64_t seq) {
    if (seq < first_seq_ || seq - first_seq_ >= blocks_.size()) return;
    blocks_[seq - first_seq_].released = true;
    while (!blocks_.empty() && blocks_.front().released) {
      blocks_.pop_front();
      ++first_seq_;
    }
  }

 private:
  struct Block {
    size_t off, n;
    bool released;
  };
  size_t size_ = 0, head_ = 0;
  uint64_t first_seq_ = 0;
  std::deque<Block> blocks_;
};

struct Request {
  kleidicv_hip_op_desc desc;
  kleidicv_hip_request_info info;
  size_t out_px = 0;
  uint64_t ring_seq = 0;
  bool in_ring = false;   // output block held in the ring
  bool budgeted = false;  // out_bytes counted in out_in_use
  bool taken = false;     // reported by take_completed
};

// A launch in flight: its completion event (or stand-in poll count) and
// the requests it carries.
struct Launch {
  hipEvent_t event = nullptr;
  unsigned polls_left = 0;
  std::vector<uint64_t> ids;
};

size_t out_channels(const kleidicv_hip_op_desc &d) {
  if (d.op == KLEIDICV_HIP_OP_YUV_SEMIPLANAR_TO_RGB)
    return (d.color_format & KLEIDICV_COLOR_CONVERSION_FLAG_ALPHA) ? 4 : 3;
  return d.channels;
}

// Output width and height: the filters and NV12 keep the source size.
void out_size(const kleidicv_hip_op_desc &d, size_t *w, size_t *h) {
  switch (d.op) {
    case KLEIDICV_HIP_OP_RESIZE_LINEAR:
    case KLEIDICV_HIP_OP_REMAP_S16POINT5:
    case KLEIDICV_HIP_OP_REMAP_F32:
    case KLEIDICV_HIP_OP_WARP_PERSPECTIVE:
      *w = d.dst_width;
      *h = d.dst_height;
      break;
    default:
      *w = d.src_width;
      *h = d.src_height;
      break;
  }
}

// One request through its single-image device entry point.
kleidicv_error_t launch_alone(const kleidicv_hip_op_desc &d, hipStream_t s) {
  void *st = static_cast<void *>(s);
  switch (d.op) {
    case KLEIDICV_HIP_OP_GAUSSIAN_BLUR:
      return kleidicv_hip_gaussian_blur_u8(d.src, d.src_stride, d.dst, d.dst_stride,
                                           d.src_width, d.src_height, d.channels,
                                           d.kernel_size, d.kernel_size, d.sigma, d.sigma,
                                           d.border_type, st);
    case KLEIDICV_HIP_OP_SEPARABLE_FILTER_2D:
      return kleidicv_hip_separable_filter_2d_u8(
          d.src, d.src_stride, d.dst, d.dst_stride, d.src_width, d.src_height, d.channels,
          d.kernel_x, d.kernel_size, d.kernel_y, d.kernel_size, d.border_type, st);
    case KLEIDICV_HIP_OP_MEDIAN_BLUR:
      return kleidicv_hip_median_blur_u8(d.src, d.src_stride, d.dst, d.dst_stride,
                                         d.src_width, d.src_height, d.channels, d.kernel_size,
                                         d.kernel_size, d.border_type, st);
    case KLEIDICV_HIP_OP_RESIZE_LINEAR:
      return kleidicv_hip_resize_linear_u8(d.src, d.src_stride, d.src_width, d.src_height,
                                           d.dst, d.dst_stride, d.dst_width, d.dst_height,
                                           d.channels, st);
    case KLEIDICV_HIP_OP_REMAP_S16POINT5:
      return kleidicv_hip_remap_s16point5_u8(
          d.src, d.src_stride, d.src_width, d.src_height, d.dst, d.dst_stride, d.dst_width,
          d.dst_height, d.channels, static_cast<const int16_t *>(d.aux0), d.aux0_stride,
          static_cast<const uint16_t *>(d.aux1), d.aux1_stride, d.border_type,
          d.border_value, st);
    case KLEIDICV_HIP_OP_REMAP_F32:
      return kleidicv_hip_remap_f32_u8(
          d.src, d.src_stride, d.src_width, d.src_height, d.dst, d.dst_stride, d.dst_width,
          d.dst_height, d.channels, static_cast<const float *>(d.aux0), d.aux0_stride,
          static_cast<const float *>(d.aux1), d.aux1_stride, d.interpolation, d.border_type,
          d.border_value, st);
    case KLEIDICV_HIP_OP_WARP_PERSPECTIVE:
      return kleidicv_hip_warp_perspective_u8(d.src, d.src_stride, d.src_width, d.src_height,
                                              d.dst, d.dst_stride, d.dst_width, d.dst_height,
                                              d.transformation, d.channels, d.interpolation,
                                              d.border_type, d.border_value, st);
    case KLEIDICV_HIP_OP_YUV_SEMIPLANAR_TO_RGB:
      return kleidicv_hip_yuv_semiplanar_to_rgb_u8(
          d.src, d.src_stride, static_cast<const uint8_t *>(d.aux0), d.aux0_stride, d.dst,
          d.dst_stride, d.src_width, d.src_height, d.color_format, st);
    default:
      return KLEIDICV_ERROR_NOT_IMPLEMENTED;
  }
}

double now_seconds() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

bool valid_config(const kleidicv_hip_server_config &c) {
  return c.slots >= 1 && c.max_batch_ops >= 1 && c.depth >= 1 && c.large_batch >= 1 &&
         c.release_done <= 1 && c.plain <= 1 && (c.plain || c.workspace_reserve > 0) &&
         c.workspace_reserve * c.slots < c.hbm_budget;
}

}  // namespace

struct kleidicv_hip_server {
  kleidicv_hip_server_config cfg{};
  bool standin = false;
  uint8_t *ring_base = nullptr;
  Ring ring;
  size_t out_budget = 0, out_in_use = 0, peak = 0;
  uint64_t next_id = 0, launches = 0;
  std::unordered_map<uint64_t, Request> reqs;
  std::deque<uint64_t> queue, ready, ready_large;
  std::vector<std::deque<Launch>> inflight;
  std::vector<hipStream_t> streams;
  std::vector<void *> ws;  // one work-list workspace per slot
  std::vector<hipEvent_t> free_events;
  std::deque<uint64_t> completed;
  std::vector<kleidicv_hip_op_desc> scratch_descs;
  std::vector<uint64_t> scratch_ids;

  size_t workspace_bytes() const { return cfg.workspace_reserve * cfg.slots; }
  size_t in_use() const { return out_in_use + workspace_bytes(); }
  // requests not yet completed: every record except the completed ones
  // waiting to be taken and the taken ones whose output is still held
  size_t pending() const { return reqs.size() - completed.size() - n_taken_kept; }
  size_t n_taken_kept = 0;

  // True when admission can never progress again: nothing is in flight or
  // ready and the head of the queue waits for ring space that only
  // kleidicv_hip_server_release() can return.
  bool stalled() const {
    if (queue.empty() || !ready.empty() || !ready_large.empty()) return false;
    for (const auto &q : inflight)
      if (!q.empty()) return false;
    return true;
  }

  ~kleidicv_hip_server() {
    if (!standin) {
      for (hipStream_t s : streams)
        if (s) (void)hipStreamSynchronize(s);
      for (auto &q : inflight)
        for (Launch &l : q)
          if (l.event) (void)hipEventDestroy(l.event);
      for (hipEvent_t e : free_events) (void)hipEventDestroy(e);
      for (void *w : ws)
        if (w) (void)hipFree(w);
      for (hipStream_t s : streams)
        if (s) (void)hipStreamDestroy(s);
    }
  }

  kleidicv_error_t init(const kleidicv_hip_server_config &c, void *ring_ptr, size_t ring_bytes) {
    cfg = c;
    standin = c.standin_polls > 0;
    out_budget = c.hbm_budget - c.workspace_reserve * c.slots;
    ring_base = static_cast<uint8_t *>(ring_ptr);
    ring.init(ring_ptr || standin ? ring_bytes : 0);
    inflight.resize(c.slots);
    streams.assign(c.slots, nullptr);
    ws.assign(c.slots, nullptr);
    if (standin) return KLEIDICV_OK;
    if (!kcv::runtime().ok) return KLEIDICV_ERROR_NOT_IMPLEMENTED;
    for (unsigned i = 0; i < c.slots; ++i) {
      KCV_TRY(from_hip(hipStreamCreateWithFlags(&streams[i], hipStreamNonBlocking)));
      if (c.workspace_reserve)
        KCV_TRY(from_hip(kcv::device_alloc(&ws[i], c.workspace_reserve, nullptr, false)));
    }
    peak = workspace_bytes();
    return KLEIDICV_OK;
  }

  // -- submission -----------------------------------------------------------
  kleidicv_error_t check_fits(const kleidicv_hip_op_desc &d, size_t *bytes, size_t *px) const {
    size_t w, h;
    out_size(d, &w, &h);
    *px = w * h;
    if (d.dst) {
      *bytes = d.dst_stride * h;
    } else {
      if (ring.size() == 0) return KLEIDICV_ERROR_NULL_POINTER;
      *bytes = w * out_channels(d) * h;
      if (*bytes > ring.size()) return KLEIDICV_ERROR_RANGE;
    }
    if (*bytes > out_budget) return KLEIDICV_ERROR_RANGE;
    return KLEIDICV_OK;
  }

  kleidicv_error_t submit(const kleidicv_hip_op_desc *ops, size_t count, const double *arrival,
                          uint64_t *first_id) {
    std::vector<size_t> bytes(count), px(count);
    for (size_t i = 0; i < count; ++i) KCV_TRY(check_fits(ops[i], &bytes[i], &px[i]));
    if (first_id) *first_id = next_id;
    for (size_t i = 0; i < count; ++i) {
      Request r;
      r.desc = ops[i];
      std::memset(&r.info, 0, sizeof(r.info));
      r.info.id = next_id;
      r.info.state = KLEIDICV_HIP_REQUEST_QUEUED;
      r.info.arrival = arrival ? arrival[i] : 0.0;
      r.info.out_bytes = bytes[i];
      r.out_px = px[i];
      reqs.emplace(next_id, r);
      queue.push_back(next_id++);
    }
    return KLEIDICV_OK;
  }

  // -- admission --------------------------------------------------------------
  void admit(double now) {
    const size_t cap = cfg.plain ? SIZE_MAX
                                 : size_t{cfg.depth} * cfg.max_batch_ops * cfg.slots;
    while (!queue.empty() && ready.size() + ready_large.size() < cap) {
      Request &r = reqs.at(queue.front());
      if (out_in_use + r.info.out_bytes > out_budget) return;  // FIFO: the head waits
      if (!r.desc.dst) {
        size_t w, h, off;
        out_size(r.desc, &w, &h);
        if (!ring.take(r.info.out_bytes, &off, &r.ring_seq)) return;  // ring full for now
        r.in_ring = true;
        // uintptr arithmetic: the stand-in's ring base may be 0
        r.desc.dst = reinterpret_cast<uint8_t *>(reinterpret_cast<uintptr_t>(ring_base) + off);
        r.desc.dst_stride = w * out_channels(r.desc);
      }
      r.info.dst = r.desc.dst;
      r.info.dst_stride = r.desc.dst_stride;
      out_in_use += r.info.out_bytes;
      r.budgeted = true;
      peak = std::max(peak, in_use());
      r.info.state = KLEIDICV_HIP_REQUEST_ADMITTED;
      r.info.admitted = now;
      queue.pop_front();
      if (cfg.plain || r.out_px >= cfg.large_px)
        ready_large.push_back(r.info.id);
      else
        ready.push_back(r.info.id);
    }
  }

  // -- completion -------------------------------------------------------------
  void finish(Request &r, double now, int state, kleidicv_error_t err) {
    r.info.state = state;
    r.info.error = err;
    r.info.done = now;
    if (r.budgeted) {
      out_in_use -= r.info.out_bytes;
      r.budgeted = false;
    }
    if (r.in_ring && (cfg.release_done || state == KLEIDICV_HIP_REQUEST_FAILED)) {
      ring.release(r.ring_seq);
      r.in_ring = false;
    }
    completed.push_back(r.info.id);
  }

  kleidicv_error_t query(Launch &l, bool *done) {
    if (standin) {
      *done = --l.polls_left == 0;
      return KLEIDICV_OK;
    }
    const hipError_t e = hipEventQuery(l.event);
    if (e == hipErrorNotReady) {
      *done = false;
      return KLEIDICV_OK;
    }
    *done = true;
    return from_hip(e);
  }

  kleidicv_error_t retire(double now) {
    for (auto &q : inflight) {
      while (!q.empty()) {
        bool done = false;
        KCV_TRY(query(q.front(), &done));
        if (!done) break;
        Launch l = std::move(q.front());
        q.pop_front();
        if (l.event) free_events.push_back(l.event);
        for (uint64_t id : l.ids) finish(reqs.at(id), now, KLEIDICV_HIP_REQUEST_DONE, KLEIDICV_OK);
      }
    }
    return KLEIDICV_OK;
  }

  // -- launching --------------------------------------------------------------
  kleidicv_error_t record(unsigned slot, std::vector<uint64_t> ids, double now, bool alone) {
    Launch l;
    if (standin) {
      l.polls_left = cfg.standin_polls;
    } else {
      if (free_events.empty()) {
        hipEvent_t e;
        KCV_TRY(from_hip(hipEventCreateWithFlags(&e, hipEventDisableTiming)));
        free_events.push_back(e);
      }
      l.event = free_events.back();
      free_events.pop_back();
      const kleidicv_error_t e = from_hip(hipEventRecord(l.event, streams[slot]));
      if (e != KLEIDICV_OK) {
        free_events.push_back(l.event);
        return e;
      }
    }
    ++launches;
    for (uint64_t id : ids) {
      Request &r = reqs.at(id);
      r.info.state = KLEIDICV_HIP_REQUEST_LAUNCHED;
      r.info.slot = slot;
      r.info.alone = alone ? 1 : 0;
      r.info.launch = launches;
      r.info.launched = now;
    }
    l.ids = std::move(ids);
    inflight[slot].push_back(std::move(l));
    return KLEIDICV_OK;
  }

  kleidicv_error_t run_alone(unsigned slot, uint64_t id, double now) {
    Request &r = reqs.at(id);
    if (!standin) {
      const kleidicv_error_t e = launch_alone(r.desc, streams[slot]);
      if (e != KLEIDICV_OK) {
        finish(r, now, KLEIDICV_HIP_REQUEST_FAILED, e);
        return KLEIDICV_OK;
      }
    }
    return record(slot, {id}, now, true);
  }

  // One batch launch on `slot` of up to `limit` requests from `q`.  When
  // their work lists exceed the slot's workspace the batch is halved (the
  // rest goes back to the front of the queue); a request that fails
  // validation completes as failed and the rest is launched without it.
  kleidicv_error_t run_batch(unsigned slot, std::deque<uint64_t> &q, size_t limit, double now) {
    std::vector<uint64_t> &ids = scratch_ids;
    ids.clear();
    while (!q.empty() && ids.size() < limit) {
      ids.push_back(q.front());
      q.pop_front();
    }
    for (;;) {
      if (ids.empty()) return KLEIDICV_OK;
      size_t need = 0, failed = SIZE_MAX;
      kleidicv_error_t e = KLEIDICV_OK;
      if (standin) {
        need = kStandinBytesPerOp * ids.size();
        if (need > cfg.workspace_reserve) e = KLEIDICV_ERROR_RANGE;
      } else {
        std::vector<kleidicv_hip_op_desc> &d = scratch_descs;
        d.clear();
        for (uint64_t id : ids) d.push_back(reqs.at(id).desc);
        e = kleidicv_hip_batch_launch(d.data(), d.size(), &cfg.tuning, ws[slot],
                                      cfg.workspace_reserve, streams[slot], &need, &failed);
      }
      if (e == KLEIDICV_OK) break;
      if (e == KLEIDICV_ERROR_RANGE && need > cfg.workspace_reserve) {
        if (ids.size() == 1) {  // its work lists alone exceed the reserve
          finish(reqs.at(ids[0]), now, KLEIDICV_HIP_REQUEST_FAILED, KLEIDICV_ERROR_RANGE);
          return KLEIDICV_OK;
        }
        const size_t keep = ids.size() / 2;
        for (size_t k = ids.size(); k-- > keep;) q.push_front(ids[k]);
        ids.resize(keep);
        continue;
      }
      if (failed < ids.size()) {
        finish(reqs.at(ids[failed]), now, KLEIDICV_HIP_REQUEST_FAILED, e);
        ids.erase(ids.begin() + static_cast<std::ptrdiff_t>(failed));
        continue;
      }
      return e;  // a device-level failure: the caller sees it
    }
    return record(slot, ids, now, false);
  }

  // Launches one request or batch on `slot`; *did says whether it did.
  kleidicv_error_t launch_one(unsigned slot, double now, bool *did) {
    *did = true;
    if (slot == 0 && !ready_large.empty()) {
      if (cfg.plain || cfg.large_batch == 1 || ready_large.size() <= cfg.depth) {
        const uint64_t id = ready_large.front();
        ready_large.pop_front();
        return run_alone(0, id, now);
      }
      return run_batch(0, ready_large, cfg.large_batch, now);
    }
    if (ready.empty()) {
      *did = false;
      return KLEIDICV_OK;
    }
    if (ready.size() == 1) {
      const uint64_t id = ready.front();
      ready.pop_front();
      return run_alone(slot, id, now);
    }
    return run_batch(slot, ready, cfg.max_batch_ops, now);
  }

  kleidicv_error_t launch(double now) {
    for (unsigned i = 0; i < cfg.slots; ++i) {
      bool did = true;
      while (did && (cfg.plain || inflight[i].size() < cfg.depth))
        KCV_TRY(launch_one(i, now, &did));
    }
    return KLEIDICV_OK;
  }

  kleidicv_error_t pump(double now) {
    KCV_TRY(retire(now));
    admit(now);
    return launch(now);
  }

  size_t take(kleidicv_hip_request_info *out, size_t capacity) {
    size_t n = 0;
    while (n < capacity && !completed.empty()) {
      auto it = reqs.find(completed.front());
      completed.pop_front();
      out[n++] = it->second.info;
      if (it->second.in_ring) {
        it->second.taken = true;  // kept until released
        ++n_taken_kept;
      } else {
        reqs.erase(it);
      }
    }
    return n;
  }
};

extern "C" {

kleidicv_error_t kleidicv_hip_run_ops(const kleidicv_hip_op_desc *ops, size_t count,
                                      void *stream, size_t *failed_index) {
  if (ops == nullptr && count > 0) return KLEIDICV_ERROR_NULL_POINTER;
  for (size_t i = 0; i < count; ++i) {
    const kleidicv_error_t e = launch_alone(ops[i], static_cast<hipStream_t>(stream));
    if (e != KLEIDICV_OK) {
      if (failed_index) *failed_index = i;
      return e;
    }
  }
  return KLEIDICV_OK;
}

double kleidicv_hip_server_clock(void) { return now_seconds(); }

kleidicv_error_t kleidicv_hip_server_create(const kleidicv_hip_server_config *config,
                                            void *ring, size_t ring_bytes,
                                            kleidicv_hip_server **server) {
  if (config == nullptr || server == nullptr) return KLEIDICV_ERROR_NULL_POINTER;
  *server = nullptr;
  if (!valid_config(*config)) return KLEIDICV_ERROR_RANGE;
  if (config->standin_polls == 0 && config->tuning.use_graph)
    return KLEIDICV_ERROR_NOT_IMPLEMENTED;
  auto *s = new (std::nothrow) kleidicv_hip_server;
  if (s == nullptr) return KLEIDICV_ERROR_ALLOCATION;
  const kleidicv_error_t e = s->init(*config, ring, ring_bytes);
  if (e != KLEIDICV_OK) {
    delete s;
    return e;
  }
  *server = s;
  return KLEIDICV_OK;
}

kleidicv_error_t kleidicv_hip_server_submit(kleidicv_hip_server *server,
                                            const kleidicv_hip_op_desc *ops, size_t count,
                                            const double *arrival, uint64_t *first_id) {
  if (server == nullptr || (ops == nullptr && count > 0)) return KLEIDICV_ERROR_NULL_POINTER;
  return server->submit(ops, count, arrival, first_id);
}

kleidicv_error_t kleidicv_hip_server_pump(kleidicv_hip_server *server, double now,
                                          size_t *pending) {
  if (server == nullptr) return KLEIDICV_ERROR_NULL_POINTER;
  const kleidicv_error_t e = server->pump(now);
  if (pending) *pending = server->pending();
  return e;
}

size_t kleidicv_hip_server_take_completed(kleidicv_hip_server *server,
                                          kleidicv_hip_request_info *out, size_t capacity) {
  if (server == nullptr || out == nullptr) return 0;
  return server->take(out, capacity);
}

kleidicv_error_t kleidicv_hip_server_info(const kleidicv_hip_server *server, uint64_t id,
                                          kleidicv_hip_request_info *info) {
  if (server == nullptr || info == nullptr) return KLEIDICV_ERROR_NULL_POINTER;
  auto it = server->reqs.find(id);
  if (it == server->reqs.end()) return KLEIDICV_ERROR_RANGE;
  *info = it->second.info;
  return KLEIDICV_OK;
}

kleidicv_error_t kleidicv_hip_server_release(kleidicv_hip_server *server, uint64_t id) {
  if (server == nullptr) return KLEIDICV_ERROR_NULL_POINTER;
  auto it = server->reqs.find(id);
  if (it == server->reqs.end()) return KLEIDICV_ERROR_RANGE;
  Request &r = it->second;
  if (r.info.state != KLEIDICV_HIP_REQUEST_DONE && r.info.state != KLEIDICV_HIP_REQUEST_FAILED)
    return KLEIDICV_ERROR_RANGE;  // still in flight
  if (r.in_ring) {
    server->ring.release(r.ring_seq);
    r.in_ring = false;
  }
  if (r.taken) {  // forget it once it has also been taken
    --server->n_taken_kept;
    server->reqs.erase(it);
  }
  return KLEIDICV_OK;
}

kleidicv_error_t kleidicv_hip_server_stats(const kleidicv_hip_server *server,
                                           kleidicv_hip_server_counters *stats) {
  if (server == nullptr || stats == nullptr) return KLEIDICV_ERROR_NULL_POINTER;
  stats->launches = server->launches;
  stats->pending = server->pending();
  stats->out_in_use = server->out_in_use;
  stats->workspace_bytes = server->workspace_bytes();
  stats->peak = server->peak;
  stats->stalled = server->stalled() ? 1 : 0;
  return KLEIDICV_OK;
}

kleidicv_error_t kleidicv_hip_server_replay(kleidicv_hip_server *server,
                                            const kleidicv_hip_op_desc *ops, size_t count,
                                            const double *offsets,
                                            kleidicv_hip_request_info *results, double *t0) {
  if (server == nullptr || ((ops == nullptr || offsets == nullptr) && count > 0))
    return KLEIDICV_ERROR_NULL_POINTER;
  for (size_t i = 1; i < count; ++i)
    if (offsets[i] < offsets[i - 1]) return KLEIDICV_ERROR_RANGE;
  // every request must be able to fit before anything runs
  for (size_t i = 0; i < count; ++i) {
    size_t b, p;
    KCV_TRY(server->check_fits(ops[i], &b, &p));
  }
  const double start = now_seconds();
  if (t0) *t0 = start;
  uint64_t first = 0;
  size_t next = 0, got = 0;
  std::vector<kleidicv_hip_request_info> done(256);
  while (next < count || server->pending() > 0) {
    const double now = now_seconds();
    size_t n = 0;
    while (next + n < count && offsets[next + n] <= now - start) ++n;
    if (n > 0) {
      std::vector<double> arr(n);
      for (size_t k = 0; k < n; ++k) arr[k] = start + offsets[next + k];
      uint64_t id0 = 0;
      KCV_TRY(server->submit(ops + next, n, arr.data(), &id0));
      if (next == 0) first = id0;
      next += n;
    }
    KCV_TRY(server->pump(now));
    for (size_t k; (k = server->take(done.data(), done.size())) > 0;) {
      for (size_t j = 0; j < k; ++j) {
        if (done[j].id < first || done[j].id - first >= count) continue;
        if (results) results[done[j].id - first] = done[j];
        ++got;
      }
    }
    // outputs held for release() can fill the ring for good
    if (next == count && server->stalled()) return KLEIDICV_ERROR_ALLOCATION;
  }
  return got == count ? KLEIDICV_OK : KLEIDICV_ERROR_RANGE;
}

void kleidicv_hip_server_destroy(kleidicv_hip_server *server) { delete server; }

}  // extern "C"
