// Batch engine host side: validates op descriptors with the single-image
// preparation code, builds per-(stage, family, variant) work lists in device
// memory and runs them as one persistent kernel launch per segment, in stage
// order on the caller's stream; it also tunes the scheduling parameters.
//
// Graph replay (use_graph = 1) is switched off: replaying the batch as an
// explicit multi-branch hipGraph ended once in a host segmentation fault
// inside hipGraphLaunch, with no defect found in this code; until the cause
// is known the option is rejected with KLEIDICV_ERROR_NOT_IMPLEMENTED.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <deque>
#include <mutex>
#include <new>
#include <vector>

#include "batch.h"
#include "kleidicv_hip/kleidicv.h"
#include "prepare.h"

namespace {

using kcv::Action;
using kcv::Family;
using kcv::from_hip;

constexpr size_t kAlign = 256;

size_t align_up(size_t v) { return (v + kAlign - 1) / kAlign * kAlign; }
int64_t div_up(int64_t a, int64_t b) { return (a + b - 1) / b; }

// One validated op, ready to be turned into jobs.
struct Prepared {
  Family family;
  unsigned stage;
  kcv::Job job;        // variant and args filled; geometry set at build time
  int resize_kind;     // for kFamResize
  size_t table_bytes;  // resize coordinate tables (0 when none)
};

struct SegmentPlan {
  Family family;
  int variant;
  unsigned stage;  // tables: ignored (run before everything that needs them)
  kcv::Segment seg;
  int grid;
};

kleidicv_hip_batch_tuning default_tuning() {
  kleidicv_hip_batch_tuning t;
  t.blocks_per_cu = 8;
  t.resize_band_rows = 0;
  t.xcd_order = 1;
  t.use_graph = 0;
  return t;
}

bool valid_tuning(const kleidicv_hip_batch_tuning &t) {
  return t.blocks_per_cu >= 1 && t.blocks_per_cu <= 16 && t.resize_band_rows <= 256 &&
         t.xcd_order <= 1 && t.use_graph <= 1;
}

// Graph replay is switched off (see the top of this file).
kleidicv_error_t check_supported(const kleidicv_hip_batch_tuning &t) {
  return t.use_graph ? KLEIDICV_ERROR_NOT_IMPLEMENTED : KLEIDICV_OK;
}

int gather_variant(int channels, int interpolation, int border) {
  return channels * 8 + interpolation * 4 + border;
}

// Validation and argument preparation for one descriptor.
kleidicv_error_t prepare_op(const kleidicv_hip_op_desc &d, Prepared *p, bool *skip) {
  *skip = false;
  std::memset(p, 0, sizeof(*p));
  p->stage = d.stage;
  Action act = Action::kNone;
  kcv::Job &j = p->job;
  switch (d.op) {
    case KLEIDICV_HIP_OP_GAUSSIAN_BLUR: {
      KCV_TRY(kcv::prepare_gaussian(d.src, d.src_stride, d.dst, d.dst_stride, d.src_width,
                                    d.src_height, d.channels, d.kernel_size, d.kernel_size,
                                    d.sigma, d.sigma, d.border_type, &j.u.sep, &act));
      if (act == Action::kCopy) {
        p->family = kcv::kFamCopy;
        kcv::CopyArgs c{};
        c.src = d.src;
        c.src_stride = d.src_stride;
        c.dst = d.dst;
        c.dst_stride = d.dst_stride;
        c.row_bytes = static_cast<int64_t>(d.src_width * d.channels);
        c.height = static_cast<int>(d.src_height);
        j.u.copy = c;
      } else {
        p->family = kcv::kFamSep;
        j.variant = (j.u.sep.mode * 32 + j.u.sep.ksize) * 8 + j.u.sep.channels;
      }
      break;
    }
    case KLEIDICV_HIP_OP_SEPARABLE_FILTER_2D:
      KCV_TRY(kcv::prepare_separable(d.src, d.src_stride, d.dst, d.dst_stride, d.src_width,
                                     d.src_height, d.channels, d.kernel_x, d.kernel_size,
                                     d.kernel_y, d.kernel_size, d.border_type, &j.u.sep,
                                     &act));
      p->family = kcv::kFamSep;
      j.variant = (j.u.sep.mode * 32 + j.u.sep.ksize) * 8 + j.u.sep.channels;
      break;
    case KLEIDICV_HIP_OP_MEDIAN_BLUR:
      KCV_TRY(kcv::prepare_median(d.src, d.src_stride, d.dst, d.dst_stride, d.src_width,
                                  d.src_height, d.channels, d.kernel_size, d.kernel_size,
                                  d.border_type, &j.u.median, &act));
      // batched stages hold the sorting-network sizes only
      if (act == Action::kLaunch && j.u.median.ksize > 7) return KLEIDICV_ERROR_NOT_IMPLEMENTED;
      p->family = kcv::kFamMedian;
      j.variant = j.u.median.ksize * 8 + j.u.median.channels;
      break;
    case KLEIDICV_HIP_OP_RESIZE_LINEAR: {
      kcv::ResizeArgs a;
      KCV_TRY(kcv::prepare_resize(d.src, d.src_stride, d.src_width, d.src_height, d.dst,
                                  d.dst_stride, d.dst_width, d.dst_height, d.channels, &a,
                                  &act));
      if (act != Action::kLaunch) break;
      if (a.dst_width == 0 || a.dst_height == 0) {
        act = Action::kNone;
        break;
      }
      p->family = kcv::kFamResize;
      switch (kcv::rsz::pick(a)) {
        case kcv::rsz::Path::kQuarter: p->resize_kind = kcv::kRsQuarter; break;
        case kcv::rsz::Path::kUp2: p->resize_kind = kcv::kRsExact2; break;
        case kcv::rsz::Path::kUp4: p->resize_kind = kcv::kRsExact4; break;
        case kcv::rsz::Path::kUp2x4: p->resize_kind = kcv::kRsUp2x4; break;
        default: p->resize_kind = kcv::kRsGeneric; break;
      }
      j.u.resize.a = a;
      j.u.resize.kind = p->resize_kind;
      // the quarter kernels are specialised per channel count
      j.variant = p->resize_kind == kcv::kRsQuarter ? kcv::kRsQuarter + 8 * a.channels
                  : p->resize_kind == kcv::kRsGeneric && kcv::rsz::generic_reads_direct(a)
                      ? kcv::kRsDirect
                      : p->resize_kind;
      p->table_bytes = kcv::resize_workspace_bytes(a);
      break;
    }
    case KLEIDICV_HIP_OP_REMAP_S16POINT5:
      KCV_TRY(kcv::prepare_remap_s16point5(
          d.src, d.src_stride, d.src_width, d.src_height, d.dst, d.dst_stride, d.dst_width,
          d.dst_height, d.channels, static_cast<const int16_t *>(d.aux0), d.aux0_stride,
          static_cast<const uint16_t *>(d.aux1), d.aux1_stride, d.border_type, d.border_value,
          &j.u.remap_s16, &act));
      p->family = kcv::kFamRemapS16;
      j.variant = j.u.remap_s16.channels * 8 + j.u.remap_s16.border;
      break;
    case KLEIDICV_HIP_OP_REMAP_F32:
      KCV_TRY(kcv::prepare_remap_f32(d.src, d.src_stride, d.src_width, d.src_height, d.dst,
                                     d.dst_stride, d.dst_width, d.dst_height, d.channels,
                                     static_cast<const float *>(d.aux0), d.aux0_stride,
                                     static_cast<const float *>(d.aux1), d.aux1_stride,
                                     d.interpolation, d.border_type, d.border_value,
                                     &j.u.remap_f32, &act));
      p->family = kcv::kFamRemapF32;
      j.variant = gather_variant(j.u.remap_f32.channels, j.u.remap_f32.interpolation,
                                 j.u.remap_f32.border);
      break;
    case KLEIDICV_HIP_OP_WARP_PERSPECTIVE:
      KCV_TRY(kcv::prepare_warp(d.src, d.src_stride, d.src_width, d.src_height, d.dst,
                                d.dst_stride, d.dst_width, d.dst_height, d.transformation,
                                d.channels, d.interpolation, d.border_type, d.border_value,
                                &j.u.warp, &act));
      p->family = kcv::kFamWarp;
      j.variant = gather_variant(j.u.warp.channels, j.u.warp.interpolation, j.u.warp.border);
      break;
    case KLEIDICV_HIP_OP_YUV_SEMIPLANAR_TO_RGB:
      KCV_TRY(kcv::prepare_yuv_sp(d.src, d.src_stride, static_cast<const uint8_t *>(d.aux0),
                                  d.aux0_stride, d.dst, d.dst_stride, d.src_width,
                                  d.src_height, d.color_format, &j.u.yuv, &act));
      p->family = kcv::kFamYuv;
      j.variant = j.u.yuv.alpha;
      break;
    default:
      return KLEIDICV_ERROR_NOT_IMPLEMENTED;
  }
  if (act == Action::kNone) {
    *skip = true;
    return KLEIDICV_OK;
  }
  // The batch runs asynchronously on device memory only.
  const void *ptrs[3] = {d.src, d.dst, d.aux0};
  const int nptr = (d.op == KLEIDICV_HIP_OP_YUV_SEMIPLANAR_TO_RGB ||
                    d.op == KLEIDICV_HIP_OP_REMAP_S16POINT5 ||
                    d.op == KLEIDICV_HIP_OP_REMAP_F32)
                       ? 3
                       : 2;
  for (int i = 0; i < nptr; ++i)
    if (!kcv::is_device_pointer(ptrs[i])) return KLEIDICV_ERROR_NOT_IMPLEMENTED;
  if ((d.op == KLEIDICV_HIP_OP_REMAP_S16POINT5 || d.op == KLEIDICV_HIP_OP_REMAP_F32) &&
      !kcv::is_device_pointer(d.aux1))
    return KLEIDICV_ERROR_NOT_IMPLEMENTED;
  return KLEIDICV_OK;
}

// Work items of a job under `t`; sets the job's tiles_x and band.
int64_t job_items(const Prepared &p, const kleidicv_hip_batch_tuning &t, kcv::Job *j) {
  using namespace kcv;
  switch (p.family) {
    case kFamSep: {
      const SepArgs &a = j->u.sep;
      j->tiles_x = static_cast<int>(div_up(int64_t{a.width} * a.channels, kFilterTileW));
      return j->tiles_x * div_up(a.height, kFilterTileH);
    }
    case kFamMedian: {
      const MedianArgs &a = j->u.median;
      j->tiles_x = static_cast<int>(div_up(int64_t{a.width} * a.channels, kFilterTileW));
      return j->tiles_x * div_up(a.height, kFilterTileH);
    }
    case kFamCopy: {
      const CopyArgs &a = j->u.copy;
      j->band = kFilterTileH;
      j->tiles_x = static_cast<int>(div_up(a.row_bytes, kCopyTileE));
      return j->tiles_x * div_up(a.height, j->band);
    }
    case kFamResize: {
      const ResizeArgs &a = j->u.resize.a;
      const int64_t n = int64_t{a.dst_width} * a.channels;
      if (p.resize_kind == kRsQuarter) {
        j->band = t.resize_band_rows ? static_cast<int>(t.resize_band_rows) : 2;
        j->tiles_x = static_cast<int>(
            div_up(n, int64_t{rsz::quarter_thread_bytes(a.channels)} * rsz::kThreads));
      } else if (p.resize_kind == kRsUp2x4) {
        // one item per row group (two output rows) of 1024 source pixels
        j->band = 1;
        j->tiles_x = static_cast<int>(div_up(a.src_width, kUp2x4TilePx));
        return j->tiles_x * (int64_t{a.src_height} + 1);
      } else {
        // staged: no taller than the band whose source window fits the LDS
        // stage (taller bands fall back to per-byte global reads); direct:
        // one group of rows in flight per thread
        j->band = rsz::generic_reads_direct(a)
                      ? rsz::direct_band_rows(rsz::kThreads / rsz::direct_lanes(a))
                      : rsz::band_rows(a);
        if (t.resize_band_rows) j->band = std::min(j->band, static_cast<int>(t.resize_band_rows));
        j->tiles_x = static_cast<int>(div_up(n, kResizeTileE));
      }
      return j->tiles_x * div_up(a.dst_height, j->band);
    }
    case kFamRemapS16: {
      const RemapS16Args &a = j->u.remap_s16;
      j->tiles_x = static_cast<int>(div_up(a.dst_width, gather_tile_w(a.channels)));
      return j->tiles_x * div_up(a.dst_height, gather_tile_h(a.channels));
    }
    case kFamRemapF32: {
      const RemapF32Args &a = j->u.remap_f32;
      j->tiles_x = static_cast<int>(div_up(a.dst_width, gather_tile_w(a.channels)));
      return j->tiles_x * div_up(a.dst_height, gather_tile_h(a.channels));
    }
    case kFamWarp: {
      const WarpArgs &a = j->u.warp;
      j->tiles_x = static_cast<int>(div_up(a.dst_width, gather_tile_w(a.channels)));
      return j->tiles_x * div_up(a.dst_height, gather_tile_h(a.channels));
    }
    case kFamYuv: {
      const YuvArgs &a = j->u.yuv;
      j->band = 4;  // row pairs per item
      j->tiles_x = static_cast<int>(div_up(a.width, kYuvTileW));
      return j->tiles_x * div_up(div_up(a.height, 2), j->band);
    }
    default:
      return 0;
  }
}

// Host-side plan of a batch placed at device address `base`: the work-list
// bytes to upload there (jobs, then first-item offsets), the total device
// bytes including the resize tables that the tables kernels fill, and the
// segments to launch.  Layout does not depend on `base`.
struct Plan {
  std::vector<unsigned char> blob;
  size_t bytes = 0;
  std::vector<SegmentPlan> segs;
  size_t items = 0;
  int blocks_per_cu = 1;
};

kleidicv_error_t make_plan(const std::vector<Prepared> &prep,
                           const kleidicv_hip_batch_tuning &t, char *base, Plan *out) {
  const int cus = kcv::runtime().compute_units > 0 ? kcv::runtime().compute_units : 256;
  *out = Plan{};
  out->blocks_per_cu = static_cast<int>(t.blocks_per_cu);

  // Jobs in (stage, family, variant) order; resize ops add a tables job.
  struct Entry {
    Family family;
    unsigned stage;
    kcv::Job job;
    int64_t items;
    size_t prep_index;
  };
  std::vector<Entry> entries;
  entries.reserve(prep.size() * 2);
  size_t table_total = 0;
  std::vector<size_t> table_offset(prep.size(), 0);
  for (size_t i = 0; i < prep.size(); ++i) {
    const Prepared &p = prep[i];
    Entry e{p.family, p.stage, p.job, 0, i};
    e.items = job_items(p, t, &e.job);
    entries.push_back(e);
    if (p.table_bytes) {
      table_offset[i] = table_total;
      table_total += align_up(p.table_bytes);
      Entry te{kcv::kFamTables, 0, kcv::Job{}, 0, i};
      const kcv::ResizeArgs &a = p.job.u.resize.a;
      te.job.u.tables.a = a;
      te.job.u.tables.kind = p.resize_kind;
      te.job.variant = p.resize_kind;
      if (p.resize_kind == kcv::kRsGeneric) te.job.u.tables.plan = kcv::rsz::make_plan(a);
      const int64_t n = int64_t{a.dst_width} * a.channels;
      te.job.tiles_x = static_cast<int>(
          div_up(std::max<int64_t>(n, a.dst_height), kcv::kTablesPerItem));
      te.items = te.job.tiles_x;
      entries.push_back(te);
    }
  }
  std::stable_sort(entries.begin(), entries.end(), [](const Entry &a, const Entry &b) {
    const bool ta = a.family == kcv::kFamTables, tb = b.family == kcv::kFamTables;
    if (ta != tb) return ta;  // tables first
    if (a.stage != b.stage) return a.stage < b.stage;
    if (a.family != b.family) return a.family < b.family;
    return a.job.variant < b.job.variant;
  });

  const size_t jobs_bytes = align_up(entries.size() * sizeof(kcv::Job));
  const size_t first_bytes = align_up(entries.size() * sizeof(uint32_t));
  out->bytes = jobs_bytes + first_bytes + table_total;
  out->blob.assign(jobs_bytes + first_bytes, 0);
  auto *hjobs = reinterpret_cast<kcv::Job *>(out->blob.data());
  auto *hfirst = reinterpret_cast<uint32_t *>(out->blob.data() + jobs_bytes);
  // Device addresses are formed as integers: `base` is null when only the
  // size is wanted.
  const uintptr_t b0 = reinterpret_cast<uintptr_t>(base);
  auto *djobs = reinterpret_cast<kcv::Job *>(b0);
  auto *dfirst = reinterpret_cast<uint32_t *>(b0 + jobs_bytes);
  const uintptr_t dtables = b0 + jobs_bytes + first_bytes;

  // Point resize and tables jobs at their tables.
  for (Entry &e : entries) {
    const Prepared &p = prep[e.prep_index];
    if (p.table_bytes == 0) continue;
    const kcv::ResizeArgs &a = p.job.u.resize.a;
    const uintptr_t c0 = dtables + table_offset[e.prep_index];
    auto *cols = reinterpret_cast<int2 *>(c0);
    auto *rows = reinterpret_cast<int2 *>(c0 + sizeof(int2) * int64_t{a.dst_width} * a.channels);
    if (e.family == kcv::kFamTables) {
      e.job.u.tables.cols = cols;
      e.job.u.tables.rows = rows;
    } else {
      e.job.u.resize.cols = cols;
      e.job.u.resize.rows = rows;
    }
  }

  size_t begin = 0;
  while (begin < entries.size()) {
    size_t end = begin;
    int64_t count = 0;
    while (end < entries.size() && entries[end].family == entries[begin].family &&
           entries[end].job.variant == entries[begin].job.variant &&
           (entries[end].stage == entries[begin].stage ||
            entries[begin].family == kcv::kFamTables)) {
      std::memcpy(&hjobs[end], &entries[end].job, sizeof(kcv::Job));
      hfirst[end] = static_cast<uint32_t>(count);
      count += entries[end].items;
      if (count >= (int64_t{1} << 31)) return KLEIDICV_ERROR_RANGE;
      ++end;
    }
    if (count > 0) {
      SegmentPlan sp{};
      sp.family = entries[begin].family;
      sp.variant = entries[begin].job.variant;
      sp.stage = entries[begin].stage;
      sp.seg.jobs = reinterpret_cast<const kcv::Job *>(reinterpret_cast<uintptr_t>(djobs) +
                                                       begin * sizeof(kcv::Job));
      sp.seg.first = reinterpret_cast<const uint32_t *>(reinterpret_cast<uintptr_t>(dfirst) +
                                                        begin * sizeof(uint32_t));
      sp.seg.njobs = static_cast<int>(end - begin);
      sp.seg.nitems = static_cast<uint32_t>(count);
      sp.seg.xcd_order = static_cast<int>(t.xcd_order);
      // Items are split statically over the grid, so a grid larger than the
      // kernel can keep resident would run its last workgroups as a second,
      // mostly empty wave: cap it at the kernel's own occupancy.
      int64_t per_cu = t.blocks_per_cu;
      const int resident = kcv::segment_blocks_per_cu(sp.family, sp.variant);
      if (resident > 0) per_cu = std::min<int64_t>(per_cu, resident);
      int64_t grid = std::min<int64_t>(count, int64_t{cus} * per_cu);
      if (t.xcd_order) grid = div_up(grid, 8) * 8;
      sp.grid = static_cast<int>(grid);
      out->segs.push_back(sp);
      out->items += static_cast<size_t>(count);
    }
    begin = end;
  }
  return KLEIDICV_OK;
}

// Copies a plan's work lists to `base`.  The copy runs on the library's own
// stream and is complete on return, so any caller stream may run the plan.
kleidicv_error_t upload(const Plan &plan, void *base) {
  if (plan.blob.empty()) return KLEIDICV_OK;
  hipStream_t up = kcv::runtime().stream;
  KCV_TRY(from_hip(hipMemcpyAsync(base, plan.blob.data(), plan.blob.size(),
                                  hipMemcpyHostToDevice, up)));
  return from_hip(hipStreamSynchronize(up));
}

// Pinned staging for one-shot launches: work lists are copied into a ring of
// pinned host memory and from there to the device on the caller's stream,
// so a launch never waits for the device.  A region of the ring is reused
// only after the copy that read it has completed (an event per copy).
class StagingRing {
 public:
  kleidicv_error_t copy(void *dst, const void *src, size_t len, hipStream_t s) {
    std::lock_guard<std::mutex> lock(mutex_);
    if (host_ == nullptr) {
      if (hipHostMalloc(reinterpret_cast<void **>(&host_), kCapacity, hipHostMallocDefault) !=
          hipSuccess) {
        (void)hipGetLastError();
        host_ = nullptr;
        return KLEIDICV_ERROR_ALLOCATION;
      }
    }
    if (len > kCapacity) return KLEIDICV_ERROR_RANGE;
    size_t off = head_;
    if (off + len > kCapacity) off = 0;
    // Wait for earlier copies still reading [off, off + len).
    for (auto it = live_.begin(); it != live_.end();) {
      if (it->off < off + len && off < it->off + it->len) {
        KCV_TRY(from_hip(hipEventSynchronize(it->ev)));
        free_events_.push_back(it->ev);
        it = live_.erase(it);
      } else {
        ++it;
      }
    }
    std::memcpy(host_ + off, src, len);
    KCV_TRY(from_hip(hipMemcpyAsync(dst, host_ + off, len, hipMemcpyHostToDevice, s)));
    hipEvent_t ev;
    if (free_events_.empty()) {
      KCV_TRY(from_hip(hipEventCreateWithFlags(&ev, hipEventDisableTiming)));
    } else {
      ev = free_events_.back();
      free_events_.pop_back();
    }
    KCV_TRY(from_hip(hipEventRecord(ev, s)));
    live_.push_back(Span{off, len, ev});
    head_ = off + len;
    return KLEIDICV_OK;
  }

 private:
  static constexpr size_t kCapacity = size_t{16} << 20;
  struct Span {
    size_t off, len;
    hipEvent_t ev;
  };
  std::mutex mutex_;
  char *host_ = nullptr;
  size_t head_ = 0;
  std::deque<Span> live_;
  std::vector<hipEvent_t> free_events_;
};

StagingRing &staging() {
  static StagingRing *ring = new StagingRing;  // lives until the process ends
  return *ring;
}

// Side streams for running the segments of one stage side by side: they
// are independent (an op only reads outputs of earlier stages), so while one
// segment's last workgroups drain, the next segment's fill the idle CUs.
// The caller's stream forks to the side streams through an event and joins
// them through one event each.  Streams and events are created once per
// device and kept until the process ends; an event may be re-recorded as
// soon as the waits on its previous recording have been enqueued.
class Fork {
 public:
  static constexpr int kSide = 3;

  kleidicv_error_t run(const SegmentPlan *segs, size_t n, hipStream_t s) {
    std::lock_guard<std::mutex> lock(mutex_);
    if (!ready_) {
      for (int k = 0; k < kSide; ++k) {
        KCV_TRY(from_hip(hipStreamCreateWithFlags(&side_[k], hipStreamNonBlocking)));
        KCV_TRY(from_hip(hipEventCreateWithFlags(&join_[k], hipEventDisableTiming)));
      }
      KCV_TRY(from_hip(hipEventCreateWithFlags(&fork_, hipEventDisableTiming)));
      ready_ = true;
    }
    const int used = static_cast<int>(std::min<size_t>(n - 1, kSide));
    KCV_TRY(from_hip(hipEventRecord(fork_, s)));
    for (int k = 0; k < used; ++k) KCV_TRY(from_hip(hipStreamWaitEvent(side_[k], fork_, 0)));
    for (size_t i = 0; i < n; ++i) {
      const SegmentPlan &sp = segs[i];
      const hipStream_t st = i % (used + 1) == 0 ? s : side_[i % (used + 1) - 1];
      KCV_TRY(from_hip(kcv::launch_segment(sp.family, sp.variant, sp.seg, sp.grid, st)));
    }
    for (int k = 0; k < used; ++k) {
      KCV_TRY(from_hip(hipEventRecord(join_[k], side_[k])));
      KCV_TRY(from_hip(hipStreamWaitEvent(s, join_[k], 0)));
    }
    return KLEIDICV_OK;
  }

 private:
  std::mutex mutex_;
  bool ready_ = false;
  hipStream_t side_[kSide] = {};
  hipEvent_t join_[kSide] = {};
  hipEvent_t fork_ = nullptr;
};

// Side streams per (device, caller stream): independent callers never wait
// on each other's side work.
Fork &fork_for(hipStream_t s) {
  static std::mutex mutex;
  static std::vector<std::pair<std::pair<int, hipStream_t>, Fork *>> forks;  // live until exit
  int dev = 0;
  (void)hipGetDevice(&dev);
  std::lock_guard<std::mutex> lock(mutex);
  for (auto &f : forks)
    if (f.first.first == dev && f.first.second == s) return *f.second;
  forks.push_back({{dev, s}, new Fork});
  return *forks.back().second;
}

bool same_group(const SegmentPlan &a, const SegmentPlan &b) {
  const bool ta = a.family == kcv::kFamTables, tb = b.family == kcv::kFamTables;
  return ta == tb && (ta || a.stage == b.stage);
}

// Segments in plan order.  The segments of one stage (or the tables) run
// side by side when they hold enough work to pay for the fork and join: at
// least four waves of workgroups over the whole GPU.
kleidicv_error_t enqueue(const Plan &plan, hipStream_t s) {
  const size_t n = plan.segs.size();
  const int64_t full = int64_t{4} * std::max(kcv::runtime().compute_units, 1) *
                       std::max<int64_t>(plan.blocks_per_cu, 1);
  for (size_t i = 0; i < n;) {
    size_t j = i + 1;
    int64_t items = plan.segs[i].seg.nitems;
    while (j < n && same_group(plan.segs[i], plan.segs[j])) items += plan.segs[j++].seg.nitems;
    if (j - i == 1 || items < full) {
      for (size_t k = i; k < j; ++k) {
        const SegmentPlan &sp = plan.segs[k];
        KCV_TRY(from_hip(kcv::launch_segment(sp.family, sp.variant, sp.seg, sp.grid, s)));
      }
    } else {
      KCV_TRY(fork_for(s).run(&plan.segs[i], j - i, s));
    }
    i = j;
  }
  return KLEIDICV_OK;
}

kleidicv_error_t prepare_all(const kleidicv_hip_op_desc *ops, size_t count,
                             std::vector<Prepared> *prep, size_t *failed_index) {
  prep->clear();
  prep->reserve(count);
  for (size_t i = 0; i < count; ++i) {
    Prepared p;
    bool skip = false;
    const kleidicv_error_t e = prepare_op(ops[i], &p, &skip);
    if (e != KLEIDICV_OK) {
      if (failed_index) *failed_index = i;
      return e;
    }
    if (!skip) prep->push_back(p);
  }
  return KLEIDICV_OK;
}

}  // namespace

struct kleidicv_hip_batch {
  std::vector<Prepared> prep;
  kleidicv_hip_batch_tuning tuning{};
  void *dev = nullptr;
  Plan plan;

  void release() {
    if (dev) (void)hipFree(dev);
    dev = nullptr;
    plan = Plan{};
  }

  // Builds the plan for `t` into a new allocation and swaps it in only once
  // it is complete on the device: on any failure the batch keeps its
  // previous plan, tuning and work lists, and still runs as before.  The
  // caller makes sure no run still reads the old work lists.
  kleidicv_error_t build(const kleidicv_hip_batch_tuning &t) {
    Plan sizing;
    KCV_TRY(make_plan(prep, t, nullptr, &sizing));
    void *ndev = nullptr;
    Plan nplan;
    if (sizing.bytes != 0) {
      KCV_TRY(from_hip(kcv::device_alloc(&ndev, sizing.bytes, nullptr, false)));
      kleidicv_error_t e = make_plan(prep, t, static_cast<char *>(ndev), &nplan);
      if (e == KLEIDICV_OK) e = upload(nplan, ndev);
      if (e != KLEIDICV_OK) {
        (void)hipFree(ndev);
        return e;
      }
    }
    release();
    dev = ndev;
    plan = std::move(nplan);
    tuning = t;
    return KLEIDICV_OK;
  }
};

extern "C" {

kleidicv_error_t kleidicv_hip_batch_create(const kleidicv_hip_op_desc *ops, size_t count,
                                           const kleidicv_hip_batch_tuning *tuning,
                                           kleidicv_hip_batch **batch,
                                           size_t *failed_index) {
  if (batch == nullptr) return KLEIDICV_ERROR_NULL_POINTER;
  *batch = nullptr;
  if (ops == nullptr && count > 0) return KLEIDICV_ERROR_NULL_POINTER;
  const kleidicv_hip_batch_tuning t = tuning ? *tuning : default_tuning();
  if (!valid_tuning(t)) return KLEIDICV_ERROR_RANGE;
  KCV_TRY(check_supported(t));
  if (!kcv::runtime().ok) return KLEIDICV_ERROR_NOT_IMPLEMENTED;
  auto *b = new (std::nothrow) kleidicv_hip_batch;
  if (b == nullptr) return KLEIDICV_ERROR_ALLOCATION;
  kleidicv_error_t e = prepare_all(ops, count, &b->prep, failed_index);
  if (e != KLEIDICV_OK) {
    delete b;
    return e;
  }
  e = b->build(t);
  if (e != KLEIDICV_OK) {
    b->release();
    delete b;
    return e;
  }
  *batch = b;
  return KLEIDICV_OK;
}

kleidicv_error_t kleidicv_hip_batch_run(kleidicv_hip_batch *batch, void *stream) {
  if (batch == nullptr) return KLEIDICV_ERROR_NULL_POINTER;
  return enqueue(batch->plan, static_cast<hipStream_t>(stream));
}

kleidicv_error_t kleidicv_hip_batch_set_tuning(kleidicv_hip_batch *batch,
                                               const kleidicv_hip_batch_tuning *tuning) {
  if (batch == nullptr || tuning == nullptr) return KLEIDICV_ERROR_NULL_POINTER;
  if (!valid_tuning(*tuning)) return KLEIDICV_ERROR_RANGE;
  KCV_TRY(check_supported(*tuning));
  // Earlier runs may still read the work lists that build() frees.
  KCV_TRY(from_hip(hipDeviceSynchronize()));
  return batch->build(*tuning);
}

kleidicv_error_t kleidicv_hip_batch_get_tuning(const kleidicv_hip_batch *batch,
                                               kleidicv_hip_batch_tuning *tuning) {
  if (batch == nullptr || tuning == nullptr) return KLEIDICV_ERROR_NULL_POINTER;
  *tuning = batch->tuning;
  return KLEIDICV_OK;
}

size_t kleidicv_hip_batch_work_items(const kleidicv_hip_batch *batch) {
  return batch ? batch->plan.items : 0;
}

kleidicv_error_t kleidicv_hip_batch_tune(kleidicv_hip_batch *batch, void *stream,
                                         unsigned iterations,
                                         kleidicv_hip_batch_tuning *best, float *best_ms) {
  if (batch == nullptr) return KLEIDICV_ERROR_NULL_POINTER;
  if (iterations == 0) iterations = 1;
  hipStream_t s = static_cast<hipStream_t>(stream);
  bool has_resize = false;
  for (const Prepared &p : batch->prep) has_resize |= p.family == kcv::kFamResize;

  std::vector<kleidicv_hip_batch_tuning> cands;
  const unsigned bpc[] = {1, 2, 4, 8};
  const unsigned bands_resize[] = {0, 4, 16};
  const unsigned bands_none[] = {0};
  for (unsigned b : bpc)
    for (unsigned x = 0; x <= 1; ++x)
      for (unsigned band : has_resize ? std::vector<unsigned>(bands_resize, bands_resize + 3)
                                      : std::vector<unsigned>(bands_none, bands_none + 1)) {
        kleidicv_hip_batch_tuning t;
        t.blocks_per_cu = b;
        t.xcd_order = x;
        t.resize_band_rows = band;
        t.use_graph = 0;
        cands.push_back(t);
      }

  KCV_TRY(from_hip(hipDeviceSynchronize()));
  hipEvent_t e0, e1;
  KCV_TRY(from_hip(hipEventCreate(&e0)));
  KCV_TRY(from_hip(hipEventCreate(&e1)));
  float best_time = -1.0f;
  kleidicv_hip_batch_tuning best_t = batch->tuning;
  kleidicv_error_t err = KLEIDICV_OK;
  for (const kleidicv_hip_batch_tuning &t : cands) {
    if ((err = batch->build(t)) != KLEIDICV_OK) break;
    if ((err = enqueue(batch->plan, s)) != KLEIDICV_OK) break;  // warm-up
    if ((err = from_hip(hipEventRecord(e0, s))) != KLEIDICV_OK) break;
    for (unsigned i = 0; i < iterations && err == KLEIDICV_OK; ++i)
      err = enqueue(batch->plan, s);
    if (err != KLEIDICV_OK) break;
    if ((err = from_hip(hipEventRecord(e1, s))) != KLEIDICV_OK) break;
    if ((err = from_hip(hipEventSynchronize(e1))) != KLEIDICV_OK) break;
    float ms = 0.0f;
    if ((err = from_hip(hipEventElapsedTime(&ms, e0, e1))) != KLEIDICV_OK) break;
    ms /= static_cast<float>(iterations);
    if (best_time < 0.0f || ms < best_time) {
      best_time = ms;
      best_t = t;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (err == KLEIDICV_OK) err = from_hip(hipDeviceSynchronize());
  const kleidicv_error_t e2 = batch->build(best_t);
  if (err == KLEIDICV_OK) err = e2;
  if (err != KLEIDICV_OK) return err;
  if (best) *best = best_t;
  if (best_ms) *best_ms = best_time;
  return KLEIDICV_OK;
}

kleidicv_error_t kleidicv_hip_batch_workspace_size(const kleidicv_hip_op_desc *ops,
                                                   size_t count,
                                                   const kleidicv_hip_batch_tuning *tuning,
                                                   size_t *bytes, size_t *failed_index) {
  if (bytes == nullptr) return KLEIDICV_ERROR_NULL_POINTER;
  *bytes = 0;
  if (ops == nullptr && count > 0) return KLEIDICV_ERROR_NULL_POINTER;
  const kleidicv_hip_batch_tuning t = tuning ? *tuning : default_tuning();
  if (!valid_tuning(t)) return KLEIDICV_ERROR_RANGE;
  KCV_TRY(check_supported(t));
  if (!kcv::runtime().ok) return KLEIDICV_ERROR_NOT_IMPLEMENTED;
  std::vector<Prepared> prep;
  KCV_TRY(prepare_all(ops, count, &prep, failed_index));
  Plan plan;
  KCV_TRY(make_plan(prep, t, nullptr, &plan));
  *bytes = plan.bytes;
  return KLEIDICV_OK;
}

kleidicv_error_t kleidicv_hip_batch_launch(const kleidicv_hip_op_desc *ops, size_t count,
                                           const kleidicv_hip_batch_tuning *tuning,
                                           void *workspace, size_t workspace_bytes,
                                           void *stream, size_t *required_bytes,
                                           size_t *failed_index) {
  if (required_bytes) *required_bytes = 0;
  if (ops == nullptr && count > 0) return KLEIDICV_ERROR_NULL_POINTER;
  const kleidicv_hip_batch_tuning t = tuning ? *tuning : default_tuning();
  if (!valid_tuning(t)) return KLEIDICV_ERROR_RANGE;
  KCV_TRY(check_supported(t));
  if (!kcv::runtime().ok) return KLEIDICV_ERROR_NOT_IMPLEMENTED;
  std::vector<Prepared> prep;
  KCV_TRY(prepare_all(ops, count, &prep, failed_index));
  Plan plan;
  KCV_TRY(make_plan(prep, t, static_cast<char *>(workspace), &plan));
  if (required_bytes) *required_bytes = plan.bytes;
  if (plan.bytes == 0) return KLEIDICV_OK;
  if (workspace_bytes < plan.bytes) return KLEIDICV_ERROR_RANGE;
  if (workspace == nullptr) return KLEIDICV_ERROR_NULL_POINTER;
  if (!kcv::is_device_pointer(workspace)) return KLEIDICV_ERROR_NOT_IMPLEMENTED;
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (!plan.blob.empty()) {
    const kleidicv_error_t e = staging().copy(workspace, plan.blob.data(), plan.blob.size(), s);
    if (e == KLEIDICV_ERROR_RANGE) {
      KCV_TRY(upload(plan, workspace));  // larger than the ring: synchronous copy
    } else {
      KCV_TRY(e);
    }
  }
  return enqueue(plan, s);
}

void kleidicv_hip_batch_destroy(kleidicv_hip_batch *batch) {
  if (batch == nullptr) return;
  (void)hipDeviceSynchronize();  // runs still in flight read the work lists
  batch->release();
  delete batch;
}

}  // extern "C"
