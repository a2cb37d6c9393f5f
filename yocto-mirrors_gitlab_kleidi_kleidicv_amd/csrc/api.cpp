// C ABI of kleidicv_hip: the drop-in, threaded and device-resident entry
// points.  Validation and argument preparation live in prepare.cpp (shared
// with the batch engine); this file handles host/device pointers and
// dispatches to the HIP launchers.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <deque>
#include <mutex>

#include "kcv_internal.h"
#include "kleidicv_hip/kleidicv.h"
#include "prepare.h"

namespace {

using kcv::Action;
using kcv::from_hip;
using kcv::is_device_pointer;
using kcv::runtime;

// The device API launches on the caller's stream as given; NULL is HIP's
// null stream, as everywhere in HIP.  The drop-in API behaves like the
// reference's synchronous CPU call: it starts after all work already queued
// on the device (a kernel on any stream may still be writing `src`, and the
// library's stream is ordered against none of them), runs on the library's
// own stream and waits for it.
hipStream_t pick_stream(void *stream, bool sync) {
  if (!sync) return static_cast<hipStream_t>(stream);
  (void)hipDeviceSynchronize();
  return runtime().stream;
}

// ---------------------------------------------------------------------------
// Host/device memory handling for the drop-in entry points.  Pointers that
// HIP knows as device or managed memory are used in place; anything else is
// host memory and is staged through stream-ordered device allocations.
size_t extent(size_t stride, size_t row_bytes, size_t rows) {
  return rows == 0 ? 0 : (rows - 1) * stride + row_bytes;
}

// One staged buffer: a host range mirrored in device memory.
struct Staged {
  const void *host = nullptr;  // null when the caller's pointer is used
  void *dev = nullptr;
  size_t bytes = 0;
  bool owned = false;  // a stream-ordered allocation of its own
};

// Device staging memory of the synchronous (host-pointer) calls: one
// grow-only buffer on the library stream, carved up by each call.  A call
// holds the lock from its first copy to its final synchronisation, so calls
// from several host threads never share it at once.  A call that needs more
// than the buffer holds gets stream-ordered allocations for the rest, as
// before, and the buffer then grows to that call's total.  A pair of
// hipMallocAsync / hipFreeAsync per staged range cost about 17 us of host
// time per range.
struct StagingArena {
  std::mutex mutex;
  char *ptr = nullptr;
  size_t bytes = 0;
};

StagingArena &staging_arena() {
  static StagingArena *a = new StagingArena;  // lives until the process ends
  return *a;
}

class Session {
 public:
  // arena: carve device memory from the staging arena (synchronous calls on
  // the library stream only)
  Session(hipStream_t s, bool arena) : stream_(s) {
    if (arena) lock_ = std::unique_lock<std::mutex>(staging_arena().mutex);
  }
  ~Session() {
    for (int i = 0; i < n_; ++i)
      if (bufs_[i].owned) (void)hipFreeAsync(bufs_[i].dev, stream_);
    if (lock_.owns_lock() && total_ > staging_arena().bytes) {
      StagingArena &a = staging_arena();
      void *p = nullptr;
      if (kcv::device_alloc(&p, total_, stream_, true) == hipSuccess) {
        if (a.ptr) (void)hipFreeAsync(a.ptr, stream_);  // after the calls that used it
        a.ptr = static_cast<char *>(p);
        a.bytes = total_;
      }
    }
  }

  // Device view of an input range; copies it in when it is host memory.
  template <typename T>
  kleidicv_error_t input(const T *p, size_t bytes, const T **dev) {
    if (bytes == 0 || is_device_pointer(p)) {
      *dev = p;
      return KLEIDICV_OK;
    }
    void *d = nullptr;
    bool owned = false;
    KCV_TRY(take(bytes, &d, &owned));
    bufs_[n_++] = Staged{p, d, bytes, owned};
    KCV_TRY(from_hip(hipMemcpyAsync(d, p, bytes, hipMemcpyHostToDevice, stream_)));
    *dev = static_cast<const T *>(d);
    return KLEIDICV_OK;
  }

  // Device view of an output range; copied back by finish().  When rows are
  // padded (stride > row bytes) the range is first filled from the host so
  // the padding bytes keep the caller's values.
  template <typename T>
  kleidicv_error_t output(T *p, size_t stride, size_t row_bytes, size_t rows,
                          T **dev) {
    const size_t bytes = extent(stride, row_bytes, rows);
    if (bytes == 0 || is_device_pointer(p)) {
      *dev = p;
      return KLEIDICV_OK;
    }
    void *d = nullptr;
    bool owned = false;
    KCV_TRY(take(bytes, &d, &owned));
    bufs_[n_] = Staged{p, d, bytes, owned};
    out_[n_] = true;
    ++n_;
    if (rows > 1 && stride != row_bytes)
      KCV_TRY(from_hip(hipMemcpyAsync(d, p, bytes, hipMemcpyHostToDevice, stream_)));
    *dev = static_cast<T *>(d);
    return KLEIDICV_OK;
  }

  kleidicv_error_t scratch(size_t bytes, void **dev) {
    *dev = nullptr;
    if (bytes == 0) return KLEIDICV_OK;
    bool owned = false;
    KCV_TRY(take(bytes, dev, &owned));
    bufs_[n_++] = Staged{nullptr, *dev, bytes, owned};
    return KLEIDICV_OK;
  }

  kleidicv_error_t finish() {
    for (int i = 0; i < n_; ++i)
      if (out_[i])
        KCV_TRY(from_hip(hipMemcpyAsync(const_cast<void *>(bufs_[i].host),
                                        bufs_[i].dev, bufs_[i].bytes,
                                        hipMemcpyDeviceToHost, stream_)));
    return from_hip(hipStreamSynchronize(stream_));
  }

 private:
  // Device memory for one staged range: from the arena while it has room,
  // else a stream-ordered allocation of its own.
  kleidicv_error_t take(size_t bytes, void **d, bool *owned) {
    const size_t need = (bytes + 255) & ~size_t{255};
    total_ += need;
    StagingArena &a = staging_arena();
    if (lock_.owns_lock() && used_ + need <= a.bytes) {
      *d = a.ptr + used_;
      used_ += need;
      *owned = false;
      return KLEIDICV_OK;
    }
    KCV_TRY(from_hip(kcv::device_alloc(d, bytes, stream_, true)));
    *owned = true;
    return KLEIDICV_OK;
  }

  hipStream_t stream_;
  std::unique_lock<std::mutex> lock_;
  size_t used_ = 0, total_ = 0;
  Staged bufs_[8];
  bool out_[8] = {};
  int n_ = 0;
};


// ---------------------------------------------------------------------------
// Resize coordinate tables of device-API calls: one grow-only device buffer
// per (device, stream), reused by successive calls on that stream.  A
// stream runs its work in order, so one call's tables kernel never
// overwrites tables that an earlier call's resize kernel still reads.  The
// entry's lock is held while a call enqueues both of its kernels, so calls
// on one stream from several host threads do not interleave them.  A
// per-call hipMallocAsync / hipFreeAsync pair cost about 25 us of host time
// (the free waits), more than a small resize takes on the GPU.  A buffer
// grows stream-ordered (the old one freed after the calls that read it) and
// is kept until the process ends: per stream, the largest table size used.
// Entries are keyed by the stream handle.  This relies on HIP releasing a
// destroyed stream's object only once the work queued on it has drained, so
// that a later stream given the same handle cannot overlap the calls that
// used the buffer.  (The HIP runtime that torch ships has no hipStreamGetId
// to tell the two streams apart.)
class TableCache {
 public:
  struct Lease {
    std::unique_lock<std::mutex> lock;
    void *ptr = nullptr;
  };

  kleidicv_error_t get(hipStream_t s, size_t bytes, Lease *out) {
    int dev = 0;
    KCV_TRY(from_hip(hipGetDevice(&dev)));
    Entry *e = nullptr;
    {
      std::lock_guard<std::mutex> g(mutex_);
      for (Entry &x : entries_)
        if (x.dev == dev && x.stream == s) e = &x;
      if (e == nullptr) {
        entries_.emplace_back();
        e = &entries_.back();
        e->dev = dev;
        e->stream = s;
      }
    }
    std::unique_lock<std::mutex> lock(e->mutex);
    if (e->bytes < bytes) {
      const size_t grow = std::max(bytes, 2 * e->bytes);
      void *p = nullptr;
      KCV_TRY(from_hip(kcv::device_alloc(&p, grow, s, true)));
      if (e->ptr) (void)hipFreeAsync(e->ptr, s);  // after the calls that read it
      e->ptr = p;
      e->bytes = grow;
    }
    out->ptr = e->ptr;
    out->lock = std::move(lock);
    return KLEIDICV_OK;
  }

 private:
  struct Entry {
    int dev = 0;
    hipStream_t stream = nullptr;
    void *ptr = nullptr;
    size_t bytes = 0;
    std::mutex mutex;
  };
  std::mutex mutex_;
  std::deque<Entry> entries_;  // stable addresses
};

TableCache &table_cache() {
  static TableCache *c = new TableCache;  // lives until the process ends
  return *c;
}

// ---------------------------------------------------------------------------
// Entry point bodies: prepare (validate + fill arguments), then either launch
// on the caller's stream (device API) or stage host memory and wait.
kleidicv_error_t gaussian_impl(const uint8_t *src, size_t src_stride, uint8_t *dst,
                               size_t dst_stride, size_t width, size_t height,
                               size_t channels, size_t kernel_width,
                               size_t kernel_height, float sigma_x, float sigma_y,
                               kleidicv_border_type_t border_type, void *stream,
                               bool sync) {
  kcv::SepArgs a;
  kcv::GaussArbArgs arb;
  Action act;
  KCV_TRY(kcv::prepare_gaussian(src, src_stride, dst, dst_stride, width, height, channels,
                                kernel_width, kernel_height, sigma_x, sigma_y, border_type,
                                &a, &act, &arb));
  if (act == Action::kNone) return KLEIDICV_OK;
  hipStream_t s = pick_stream(stream, sync);
  const size_t row_bytes = width * channels;
  if (act == Action::kLaunchArbitrary) {
    // the vertical pass's u8 intermediate, with the destination's stride
    const size_t tmp_bytes = extent(dst_stride, row_bytes, height);
    Session ss(s, sync);
    void *tmp = nullptr;
    KCV_TRY(ss.scratch(tmp_bytes, &tmp));
    arb.tmp = static_cast<uint8_t *>(tmp);
    if (!sync) {
      // scratch is released on the stream after the launch, without a wait
      KCV_TRY(from_hip(kcv::launch_gaussian_arbitrary(arb, s)));
      return KLEIDICV_OK;
    }
    KCV_TRY(ss.input(src, extent(src_stride, row_bytes, height), &arb.src));
    KCV_TRY(ss.output(dst, dst_stride, row_bytes, height, &arb.dst));
    KCV_TRY(from_hip(kcv::launch_gaussian_arbitrary(arb, s)));
    return ss.finish();
  }
  if (!sync) {
    if (act == Action::kCopy)
      return from_hip(hipMemcpy2DAsync(dst, dst_stride, src, src_stride, row_bytes,
                                       height, hipMemcpyDefault, s));
    return from_hip(kcv::launch_separable(a, s));
  }
  Session ss(s, true);
  KCV_TRY(ss.input(src, extent(src_stride, row_bytes, height), &a.src));
  KCV_TRY(ss.output(dst, dst_stride, row_bytes, height, &a.dst));
  if (act == Action::kCopy)
    KCV_TRY(from_hip(hipMemcpy2DAsync(a.dst, dst_stride, a.src, src_stride, row_bytes,
                                      height, hipMemcpyDefault, s)));
  else
    KCV_TRY(from_hip(kcv::launch_separable(a, s)));
  return ss.finish();
}

kleidicv_error_t blur_down_impl(const uint8_t *src, size_t src_stride, size_t width,
                                size_t height, uint8_t *dst, size_t dst_stride,
                                size_t channels, kleidicv_border_type_t border_type,
                                void *stream, bool sync) {
  kcv::SepArgs a;
  Action act;
  KCV_TRY(kcv::prepare_blur_down(src, src_stride, width, height, dst, dst_stride, channels,
                                 border_type, &a, &act));
  if (act == Action::kNone) return KLEIDICV_OK;
  hipStream_t s = pick_stream(stream, sync);
  if (!sync) return from_hip(kcv::launch_blur_down(a, s));
  Session ss(s, true);
  KCV_TRY(ss.input(src, extent(src_stride, width * channels, height), &a.src));
  KCV_TRY(ss.output(dst, dst_stride, (width + 1) / 2 * channels, (height + 1) / 2, &a.dst));
  KCV_TRY(from_hip(kcv::launch_blur_down(a, s)));
  return ss.finish();
}

kleidicv_error_t yuv_impl(const uint8_t *src, size_t src_stride, uint8_t *dst,
                          size_t dst_stride, size_t width, size_t height,
                          kleidicv_color_conversion_t color_format, void *stream, bool sync) {
  kcv::YuvXArgs a;
  Action act;
  KCV_TRY(kcv::prepare_yuv(src, src_stride, dst, dst_stride, width, height, color_format, &a,
                           &act));
  if (act == Action::kNone) return KLEIDICV_OK;
  hipStream_t s = pick_stream(stream, sync);
  if (!sync) return from_hip(kcv::launch_yuv_to_rgb(a, s));
  Session ss(s, true);
  const uint8_t *dsrc = nullptr;
  KCV_TRY(ss.input(src, a.src_extent, &dsrc));
  // the chroma plane pointers move with the luma plane
  a.u = dsrc + (a.u - a.src);
  a.v = dsrc + (a.v - a.src);
  a.src = dsrc;
  KCV_TRY(ss.output(dst, dst_stride, width * (a.alpha ? 4 : 3), height, &a.dst));
  KCV_TRY(from_hip(kcv::launch_yuv_to_rgb(a, s)));
  return ss.finish();
}

kleidicv_error_t rgb_to_yuv_impl(const uint8_t *src, size_t src_stride, uint8_t *dst,
                                 size_t dst_stride, size_t width, size_t height,
                                 kleidicv_color_conversion_t color_format, void *stream,
                                 bool sync) {
  kcv::RgbYuvArgs a;
  Action act;
  KCV_TRY(kcv::prepare_rgb_to_yuv(src, src_stride, dst, dst_stride, width, height,
                                  color_format, &a, &act));
  if (act == Action::kNone) return KLEIDICV_OK;
  hipStream_t s = pick_stream(stream, sync);
  if (!sync) return from_hip(kcv::launch_rgb_to_yuv(a, s));
  Session ss(s, true);
  KCV_TRY(ss.input(src, extent(src_stride, width * a.scn, height), &a.src));
  const size_t out_bytes = a.layout == kcv::kYuv444 ? 3 * width
                           : a.layout == kcv::kYuv422 ? 2 * width : width;
  const size_t out_rows = a.layout == kcv::kYuv420P ? (height * 3 + 1) / 2 : height;
  uint8_t *ddst = nullptr;
  KCV_TRY(ss.output(dst, dst_stride, out_bytes, out_rows, &ddst));
  a.uv = ddst + (a.uv - a.dst);  // planar chroma area moves with the buffer
  a.dst = ddst;
  KCV_TRY(from_hip(kcv::launch_rgb_to_yuv(a, s)));
  return ss.finish();
}

kleidicv_error_t rgb_to_yuv_sp_impl(const uint8_t *src, size_t src_stride, uint8_t *y_dst,
                                    size_t y_stride, uint8_t *uv_dst, size_t uv_stride,
                                    size_t width, size_t height,
                                    kleidicv_color_conversion_t color_format, void *stream,
                                    bool sync) {
  kcv::RgbYuvArgs a;
  Action act;
  KCV_TRY(kcv::prepare_rgb_to_yuv_sp(src, src_stride, y_dst, y_stride, uv_dst, uv_stride,
                                     width, height, color_format, &a, &act));
  if (act == Action::kNone) return KLEIDICV_OK;
  hipStream_t s = pick_stream(stream, sync);
  if (!sync) return from_hip(kcv::launch_rgb_to_yuv(a, s));
  Session ss(s, true);
  KCV_TRY(ss.input(src, extent(src_stride, width * a.scn, height), &a.src));
  KCV_TRY(ss.output(y_dst, y_stride, width, height, &a.dst));
  // (w+1)/2 interleaved pairs per chroma row
  KCV_TRY(ss.output(uv_dst, uv_stride, 2 * ((width + 1) / 2), (height + 1) / 2, &a.uv));
  KCV_TRY(from_hip(kcv::launch_rgb_to_yuv(a, s)));
  return ss.finish();
}

kleidicv_error_t separable_impl(const uint8_t *src, size_t src_stride, uint8_t *dst,
                                size_t dst_stride, size_t width, size_t height,
                                size_t channels, const uint8_t *kernel_x,
                                size_t kernel_width, const uint8_t *kernel_y,
                                size_t kernel_height,
                                kleidicv_border_type_t border_type, void *stream,
                                bool sync) {
  kcv::SepArgs a;
  Action act;
  KCV_TRY(kcv::prepare_separable(src, src_stride, dst, dst_stride, width, height, channels,
                                 kernel_x, kernel_width, kernel_y, kernel_height,
                                 border_type, &a, &act));
  if (act == Action::kNone) return KLEIDICV_OK;
  hipStream_t s = pick_stream(stream, sync);
  if (!sync) return from_hip(kcv::launch_separable(a, s));
  const size_t row_bytes = width * channels;
  Session ss(s, true);
  KCV_TRY(ss.input(src, extent(src_stride, row_bytes, height), &a.src));
  KCV_TRY(ss.output(dst, dst_stride, row_bytes, height, &a.dst));
  KCV_TRY(from_hip(kcv::launch_separable(a, s)));
  return ss.finish();
}

kleidicv_error_t median_impl(const uint8_t *src, size_t src_stride, uint8_t *dst,
                             size_t dst_stride, size_t width, size_t height,
                             size_t channels, size_t kernel_width,
                             size_t kernel_height, kleidicv_border_type_t border_type,
                             void *stream, bool sync) {
  kcv::MedianArgs a;
  Action act;
  KCV_TRY(kcv::prepare_median(src, src_stride, dst, dst_stride, width, height, channels,
                              kernel_width, kernel_height, border_type, &a, &act));
  if (act == Action::kNone) return KLEIDICV_OK;
  hipStream_t s = pick_stream(stream, sync);
  if (!sync) return from_hip(kcv::launch_median(a, s));
  const size_t row_bytes = width * channels;
  Session ss(s, true);
  KCV_TRY(ss.input(src, extent(src_stride, row_bytes, height), &a.src));
  KCV_TRY(ss.output(dst, dst_stride, row_bytes, height, &a.dst));
  KCV_TRY(from_hip(kcv::launch_median(a, s)));
  return ss.finish();
}

kleidicv_error_t resize_impl(const uint8_t *src, size_t src_stride, size_t src_width,
                             size_t src_height, uint8_t *dst, size_t dst_stride,
                             size_t dst_width, size_t dst_height, size_t channels,
                             void *stream, bool sync) {
  kcv::ResizeArgs a;
  Action act;
  KCV_TRY(kcv::prepare_resize(src, src_stride, src_width, src_height, dst, dst_stride,
                              dst_width, dst_height, channels, &a, &act));
  if (act == Action::kNone) return KLEIDICV_OK;
  hipStream_t s = pick_stream(stream, sync);
  const size_t table_bytes = kcv::resize_workspace_bytes(a);
  if (!sync) {
    if (table_bytes == 0) return from_hip(kcv::launch_resize(a, nullptr, s));
    TableCache::Lease lease;
    KCV_TRY(table_cache().get(s, table_bytes, &lease));
    return from_hip(kcv::launch_resize(a, lease.ptr, s));
  }
  Session ss(s, true);
  void *ws = nullptr;
  KCV_TRY(ss.scratch(table_bytes, &ws));
  KCV_TRY(ss.input(src, extent(src_stride, src_width * channels, src_height), &a.src));
  KCV_TRY(ss.output(dst, dst_stride, dst_width * channels, dst_height, &a.dst));
  KCV_TRY(from_hip(kcv::launch_resize(a, ws, s)));
  return ss.finish();
}

kleidicv_error_t remap_s16point5_impl(
    const uint8_t *src, size_t src_stride, size_t src_width, size_t src_height,
    uint8_t *dst, size_t dst_stride, size_t dst_width, size_t dst_height,
    size_t channels, const int16_t *mapxy, size_t mapxy_stride,
    const uint16_t *mapfrac, size_t mapfrac_stride,
    kleidicv_border_type_t border_type, const uint8_t *border_value, void *stream,
    bool sync) {
  kcv::RemapS16Args a;
  Action act;
  KCV_TRY(kcv::prepare_remap_s16point5(src, src_stride, src_width, src_height, dst,
                                       dst_stride, dst_width, dst_height, channels, mapxy,
                                       mapxy_stride, mapfrac, mapfrac_stride, border_type,
                                       border_value, &a, &act));
  if (act == Action::kNone) return KLEIDICV_OK;
  hipStream_t s = pick_stream(stream, sync);
  if (!sync) return from_hip(kcv::launch_remap_s16point5(a, s));
  Session ss(s, true);
  KCV_TRY(ss.input(src, extent(src_stride, src_width * channels, src_height), &a.src));
  KCV_TRY(ss.input(mapxy, extent(mapxy_stride, dst_width * 4, dst_height), &a.mapxy));
  KCV_TRY(ss.input(mapfrac, extent(mapfrac_stride, dst_width * 2, dst_height), &a.mapfrac));
  KCV_TRY(ss.output(dst, dst_stride, dst_width * channels, dst_height, &a.dst));
  KCV_TRY(from_hip(kcv::launch_remap_s16point5(a, s)));
  return ss.finish();
}

kleidicv_error_t remap_f32_impl(const uint8_t *src, size_t src_stride,
                                size_t src_width, size_t src_height, uint8_t *dst,
                                size_t dst_stride, size_t dst_width, size_t dst_height,
                                size_t channels, const float *mapx, size_t mapx_stride,
                                const float *mapy, size_t mapy_stride,
                                kleidicv_interpolation_type_t interpolation,
                                kleidicv_border_type_t border_type,
                                const uint8_t *border_value, void *stream, bool sync) {
  kcv::RemapF32Args a;
  Action act;
  KCV_TRY(kcv::prepare_remap_f32(src, src_stride, src_width, src_height, dst, dst_stride,
                                 dst_width, dst_height, channels, mapx, mapx_stride, mapy,
                                 mapy_stride, interpolation, border_type, border_value, &a,
                                 &act));
  if (act == Action::kNone) return KLEIDICV_OK;
  hipStream_t s = pick_stream(stream, sync);
  if (!sync) return from_hip(kcv::launch_remap_f32(a, s));
  Session ss(s, true);
  KCV_TRY(ss.input(src, extent(src_stride, src_width * channels, src_height), &a.src));
  KCV_TRY(ss.input(mapx, extent(mapx_stride, dst_width * 4, dst_height), &a.mapx));
  KCV_TRY(ss.input(mapy, extent(mapy_stride, dst_width * 4, dst_height), &a.mapy));
  KCV_TRY(ss.output(dst, dst_stride, dst_width * channels, dst_height, &a.dst));
  KCV_TRY(from_hip(kcv::launch_remap_f32(a, s)));
  return ss.finish();
}

kleidicv_error_t warp_impl(const uint8_t *src, size_t src_stride, size_t src_width,
                           size_t src_height, uint8_t *dst, size_t dst_stride,
                           size_t dst_width, size_t dst_height,
                           const float transformation[9], size_t channels,
                           kleidicv_interpolation_type_t interpolation,
                           kleidicv_border_type_t border_type,
                           const uint8_t *border_value, void *stream, bool sync) {
  kcv::WarpArgs a;
  Action act;
  KCV_TRY(kcv::prepare_warp(src, src_stride, src_width, src_height, dst, dst_stride,
                            dst_width, dst_height, transformation, channels, interpolation,
                            border_type, border_value, &a, &act));
  if (act == Action::kNone) return KLEIDICV_OK;
  hipStream_t s = pick_stream(stream, sync);
  if (!sync) return from_hip(kcv::launch_warp_perspective(a, s));
  Session ss(s, true);
  KCV_TRY(ss.input(src, extent(src_stride, src_width * channels, src_height), &a.src));
  KCV_TRY(ss.output(dst, dst_stride, dst_width * channels, dst_height, &a.dst));
  KCV_TRY(from_hip(kcv::launch_warp_perspective(a, s)));
  return ss.finish();
}

kleidicv_error_t yuv_sp_impl(const uint8_t *src_y, size_t src_y_stride,
                             const uint8_t *src_uv, size_t src_uv_stride, uint8_t *dst,
                             size_t dst_stride, size_t width, size_t height,
                             kleidicv_color_conversion_t color_format, void *stream,
                             bool sync) {
  kcv::YuvArgs a;
  Action act;
  KCV_TRY(kcv::prepare_yuv_sp(src_y, src_y_stride, src_uv, src_uv_stride, dst, dst_stride,
                              width, height, color_format, &a, &act));
  if (act == Action::kNone) return KLEIDICV_OK;
  hipStream_t s = pick_stream(stream, sync);
  if (!sync) return from_hip(kcv::launch_yuv_sp_to_rgb(a, s));
  const size_t oc = a.alpha ? 4 : 3;
  Session ss(s, true);
  const size_t uv_bytes = (width + 1) / 2 * 2;
  KCV_TRY(ss.input(src_y, extent(src_y_stride, width, height), &a.y));
  KCV_TRY(ss.input(src_uv, extent(src_uv_stride, uv_bytes, (height + 1) / 2), &a.uv));
  KCV_TRY(ss.output(dst, dst_stride, width * oc, height, &a.dst));
  KCV_TRY(from_hip(kcv::launch_yuv_sp_to_rgb(a, s)));
  return ss.finish();
}

kleidicv_error_t check_mt(kleidicv_thread_multithreading mt) {
  return mt.parallel == nullptr ? KLEIDICV_ERROR_NULL_POINTER : KLEIDICV_OK;
}

// Function-pointer targets for the KLEIDICV_API_DECLARATION symbols.
kleidicv_error_t yuv_sp_entry(const uint8_t *src_y, size_t src_y_stride,
                              const uint8_t *src_uv, size_t src_uv_stride, uint8_t *dst,
                              size_t dst_stride, size_t width, size_t height,
                              kleidicv_color_conversion_t color_format) {
  return yuv_sp_impl(src_y, src_y_stride, src_uv, src_uv_stride, dst, dst_stride, width,
                     height, color_format, nullptr, true);
}

kleidicv_error_t remap_s16point5_entry(
    const uint8_t *src, size_t src_stride, size_t src_width, size_t src_height,
    uint8_t *dst, size_t dst_stride, size_t dst_width, size_t dst_height,
    size_t channels, const int16_t *mapxy, size_t mapxy_stride,
    const uint16_t *mapfrac, size_t mapfrac_stride,
    kleidicv_border_type_t border_type, const uint8_t *border_value) {
  return remap_s16point5_impl(src, src_stride, src_width, src_height, dst, dst_stride,
                              dst_width, dst_height, channels, mapxy, mapxy_stride,
                              mapfrac, mapfrac_stride, border_type, border_value,
                              nullptr, true);
}

kleidicv_error_t remap_f32_entry(const uint8_t *src, size_t src_stride,
                                 size_t src_width, size_t src_height, uint8_t *dst,
                                 size_t dst_stride, size_t dst_width, size_t dst_height,
                                 size_t channels, const float *mapx, size_t mapx_stride,
                                 const float *mapy, size_t mapy_stride,
                                 kleidicv_interpolation_type_t interpolation,
                                 kleidicv_border_type_t border_type,
                                 const uint8_t *border_value) {
  return remap_f32_impl(src, src_stride, src_width, src_height, dst, dst_stride,
                        dst_width, dst_height, channels, mapx, mapx_stride, mapy,
                        mapy_stride, interpolation, border_type, border_value, nullptr,
                        true);
}

}  // namespace

extern "C" {

kleidicv_error_t (*kleidicv_yuv_semiplanar_to_rgb_u8)(
    const uint8_t *, size_t, const uint8_t *, size_t, uint8_t *, size_t, size_t,
    size_t, kleidicv_color_conversion_t) = &yuv_sp_entry;

kleidicv_error_t (*kleidicv_remap_s16point5_u8)(
    const uint8_t *, size_t, size_t, size_t, uint8_t *, size_t, size_t, size_t,
    size_t, const int16_t *, size_t, const uint16_t *, size_t,
    kleidicv_border_type_t, const uint8_t *) = &remap_s16point5_entry;

kleidicv_error_t (*kleidicv_remap_f32_u8)(
    const uint8_t *, size_t, size_t, size_t, uint8_t *, size_t, size_t, size_t,
    size_t, const float *, size_t, const float *, size_t,
    kleidicv_interpolation_type_t, kleidicv_border_type_t,
    const uint8_t *) = &remap_f32_entry;

kleidicv_error_t kleidicv_gaussian_blur_u8(const uint8_t *src, size_t src_stride,
                                           uint8_t *dst, size_t dst_stride,
                                           size_t width, size_t height,
                                           size_t channels, size_t kernel_width,
                                           size_t kernel_height, float sigma_x,
                                           float sigma_y,
                                           kleidicv_border_type_t border_type) {
  return gaussian_impl(src, src_stride, dst, dst_stride, width, height, channels,
                       kernel_width, kernel_height, sigma_x, sigma_y, border_type,
                       nullptr, true);
}

kleidicv_error_t kleidicv_separable_filter_2d_u8(
    const uint8_t *src, size_t src_stride, uint8_t *dst, size_t dst_stride,
    size_t width, size_t height, size_t channels, const uint8_t *kernel_x,
    size_t kernel_width, const uint8_t *kernel_y, size_t kernel_height,
    kleidicv_border_type_t border_type) {
  return separable_impl(src, src_stride, dst, dst_stride, width, height, channels,
                        kernel_x, kernel_width, kernel_y, kernel_height, border_type,
                        nullptr, true);
}

kleidicv_error_t kleidicv_rgb_to_yuv_u8(const uint8_t *src, size_t src_stride, uint8_t *dst,
                                        size_t dst_stride, size_t width, size_t height,
                                        kleidicv_color_conversion_t color_format) {
  return rgb_to_yuv_impl(src, src_stride, dst, dst_stride, width, height, color_format,
                         nullptr, true);
}

kleidicv_error_t kleidicv_rgb_to_yuv_semiplanar_u8(const uint8_t *src, size_t src_stride,
                                                   uint8_t *y_dst, size_t y_stride,
                                                   uint8_t *uv_dst, size_t uv_stride,
                                                   size_t width, size_t height,
                                                   kleidicv_color_conversion_t color_format) {
  return rgb_to_yuv_sp_impl(src, src_stride, y_dst, y_stride, uv_dst, uv_stride, width, height,
                            color_format, nullptr, true);
}

kleidicv_error_t kleidicv_yuv_to_rgb_u8(const uint8_t *src, size_t src_stride, uint8_t *dst,
                                        size_t dst_stride, size_t width, size_t height,
                                        kleidicv_color_conversion_t color_format) {
  return yuv_impl(src, src_stride, dst, dst_stride, width, height, color_format, nullptr, true);
}

kleidicv_error_t kleidicv_blur_and_downsample_u8(const uint8_t *src, size_t src_stride,
                                                 size_t src_width, size_t src_height,
                                                 uint8_t *dst, size_t dst_stride,
                                                 size_t channels,
                                                 kleidicv_border_type_t border_type) {
  return blur_down_impl(src, src_stride, src_width, src_height, dst, dst_stride, channels,
                        border_type, nullptr, true);
}

kleidicv_error_t kleidicv_median_blur_u8(const uint8_t *src, size_t src_stride,
                                         uint8_t *dst, size_t dst_stride, size_t width,
                                         size_t height, size_t channels,
                                         size_t kernel_width, size_t kernel_height,
                                         kleidicv_border_type_t border_type) {
  return median_impl(src, src_stride, dst, dst_stride, width, height, channels,
                     kernel_width, kernel_height, border_type, nullptr, true);
}

kleidicv_error_t kleidicv_resize_linear_u8(const uint8_t *src, size_t src_stride,
                                           size_t src_width, size_t src_height,
                                           uint8_t *dst, size_t dst_stride,
                                           size_t dst_width, size_t dst_height,
                                           size_t channels) {
  return resize_impl(src, src_stride, src_width, src_height, dst, dst_stride, dst_width,
                     dst_height, channels, nullptr, true);
}

kleidicv_error_t kleidicv_warp_perspective_u8(
    const uint8_t *src, size_t src_stride, size_t src_width, size_t src_height,
    uint8_t *dst, size_t dst_stride, size_t dst_width, size_t dst_height,
    const float transformation[9], size_t channels,
    kleidicv_interpolation_type_t interpolation, kleidicv_border_type_t border_type,
    const uint8_t *border_value) {
  return warp_impl(src, src_stride, src_width, src_height, dst, dst_stride, dst_width,
                   dst_height, transformation, channels, interpolation, border_type,
                   border_value, nullptr, true);
}

// Backend-selection twins (kleidicv.h:21-35): same semantics, one backend.
kleidicv_error_t (*kleidicv_yuv_semiplanar_to_rgb_u8_sme)(
    const uint8_t *, size_t, const uint8_t *, size_t, uint8_t *, size_t, size_t,
    size_t, kleidicv_color_conversion_t) = &yuv_sp_entry;

kleidicv_error_t kleidicv_yuv_to_rgb_u8_sme(const uint8_t *src, size_t src_stride,
                                            uint8_t *dst, size_t dst_stride, size_t width,
                                            size_t height,
                                            kleidicv_color_conversion_t color_format) {
  return kleidicv_yuv_to_rgb_u8(src, src_stride, dst, dst_stride, width, height,
                                color_format);
}

kleidicv_error_t kleidicv_rgb_to_yuv_u8_sme(const uint8_t *src, size_t src_stride,
                                            uint8_t *dst, size_t dst_stride, size_t width,
                                            size_t height,
                                            kleidicv_color_conversion_t color_format) {
  return kleidicv_rgb_to_yuv_u8(src, src_stride, dst, dst_stride, width, height,
                                color_format);
}

kleidicv_error_t kleidicv_rgb_to_yuv_semiplanar_u8_sme(
    const uint8_t *src, size_t src_stride, uint8_t *y_dst, size_t y_stride, uint8_t *uv_dst,
    size_t uv_stride, size_t width, size_t height, kleidicv_color_conversion_t color_format) {
  return kleidicv_rgb_to_yuv_semiplanar_u8(src, src_stride, y_dst, y_stride, uv_dst,
                                           uv_stride, width, height, color_format);
}

kleidicv_error_t kleidicv_separable_filter_2d_u8_sme(
    const uint8_t *src, size_t src_stride, uint8_t *dst, size_t dst_stride,
    size_t width, size_t height, size_t channels, const uint8_t *kernel_x,
    size_t kernel_width, const uint8_t *kernel_y, size_t kernel_height,
    kleidicv_border_type_t border_type) {
  return kleidicv_separable_filter_2d_u8(src, src_stride, dst, dst_stride, width, height,
                                         channels, kernel_x, kernel_width, kernel_y,
                                         kernel_height, border_type);
}

kleidicv_error_t kleidicv_gaussian_blur_u8_sme(
    const uint8_t *src, size_t src_stride, uint8_t *dst, size_t dst_stride, size_t width,
    size_t height, size_t channels, size_t kernel_width, size_t kernel_height,
    float sigma_x, float sigma_y, kleidicv_border_type_t border_type) {
  return kleidicv_gaussian_blur_u8(src, src_stride, dst, dst_stride, width, height,
                                   channels, kernel_width, kernel_height, sigma_x,
                                   sigma_y, border_type);
}

kleidicv_error_t kleidicv_blur_and_downsample_u8_sme(const uint8_t *src, size_t src_stride,
                                                     size_t src_width, size_t src_height,
                                                     uint8_t *dst, size_t dst_stride,
                                                     size_t channels,
                                                     kleidicv_border_type_t border_type) {
  return kleidicv_blur_and_downsample_u8(src, src_stride, src_width, src_height, dst,
                                         dst_stride, channels, border_type);
}

kleidicv_error_t kleidicv_median_blur_u8_sme(const uint8_t *src, size_t src_stride,
                                             uint8_t *dst, size_t dst_stride, size_t width,
                                             size_t height, size_t channels,
                                             size_t kernel_width, size_t kernel_height,
                                             kleidicv_border_type_t border_type) {
  return kleidicv_median_blur_u8(src, src_stride, dst, dst_stride, width, height, channels,
                                 kernel_width, kernel_height, border_type);
}

kleidicv_error_t kleidicv_resize_linear_u8_sme(const uint8_t *src, size_t src_stride,
                                               size_t src_width, size_t src_height,
                                               uint8_t *dst, size_t dst_stride,
                                               size_t dst_width, size_t dst_height,
                                               size_t channels) {
  return kleidicv_resize_linear_u8(src, src_stride, src_width, src_height, dst, dst_stride,
                                   dst_width, dst_height, channels);
}

// Threaded API: the stripe scheduler is replaced by the GPU grid.
kleidicv_error_t kleidicv_thread_gaussian_blur_u8(
    const uint8_t *src, size_t src_stride, uint8_t *dst, size_t dst_stride,
    size_t width, size_t height, size_t channels, size_t kernel_width,
    size_t kernel_height, float sigma_x, float sigma_y,
    kleidicv_border_type_t border_type, kleidicv_thread_multithreading mt) {
  KCV_TRY(check_mt(mt));
  return kleidicv_gaussian_blur_u8(src, src_stride, dst, dst_stride, width, height,
                                   channels, kernel_width, kernel_height, sigma_x,
                                   sigma_y, border_type);
}

kleidicv_error_t kleidicv_thread_separable_filter_2d_u8(
    const uint8_t *src, size_t src_stride, uint8_t *dst, size_t dst_stride,
    size_t width, size_t height, size_t channels, const uint8_t *kernel_x,
    size_t kernel_width, const uint8_t *kernel_y, size_t kernel_height,
    kleidicv_border_type_t border_type, kleidicv_thread_multithreading mt) {
  KCV_TRY(check_mt(mt));
  return kleidicv_separable_filter_2d_u8(src, src_stride, dst, dst_stride, width,
                                         height, channels, kernel_x, kernel_width,
                                         kernel_y, kernel_height, border_type);
}

kleidicv_error_t kleidicv_thread_rgb_to_yuv_u8(const uint8_t *src, size_t src_stride,
                                               uint8_t *dst, size_t dst_stride, size_t width,
                                               size_t height,
                                               kleidicv_color_conversion_t color_format,
                                               kleidicv_thread_multithreading mt) {
  KCV_TRY(check_mt(mt));
  return rgb_to_yuv_impl(src, src_stride, dst, dst_stride, width, height, color_format,
                         nullptr, true);
}

kleidicv_error_t kleidicv_thread_rgb_to_yuv_semiplanar_u8(
    const uint8_t *src, size_t src_stride, uint8_t *y_dst, size_t y_stride, uint8_t *uv_dst,
    size_t uv_stride, size_t width, size_t height, kleidicv_color_conversion_t color_format,
    kleidicv_thread_multithreading mt) {
  KCV_TRY(check_mt(mt));
  return rgb_to_yuv_sp_impl(src, src_stride, y_dst, y_stride, uv_dst, uv_stride, width, height,
                            color_format, nullptr, true);
}

kleidicv_error_t kleidicv_thread_yuv_to_rgb_u8(const uint8_t *src, size_t src_stride,
                                               uint8_t *dst, size_t dst_stride, size_t width,
                                               size_t height,
                                               kleidicv_color_conversion_t color_format,
                                               kleidicv_thread_multithreading mt) {
  KCV_TRY(check_mt(mt));
  return yuv_impl(src, src_stride, dst, dst_stride, width, height, color_format, nullptr, true);
}

kleidicv_error_t kleidicv_thread_blur_and_downsample_u8(
    const uint8_t *src, size_t src_stride, size_t src_width, size_t src_height, uint8_t *dst,
    size_t dst_stride, size_t channels, kleidicv_border_type_t border_type,
    kleidicv_thread_multithreading mt) {
  KCV_TRY(check_mt(mt));
  return blur_down_impl(src, src_stride, src_width, src_height, dst, dst_stride, channels,
                        border_type, nullptr, true);
}

kleidicv_error_t kleidicv_thread_median_blur_u8(
    const uint8_t *src, size_t src_stride, uint8_t *dst, size_t dst_stride,
    size_t width, size_t height, size_t channels, size_t kernel_width,
    size_t kernel_height, kleidicv_border_type_t border_type,
    kleidicv_thread_multithreading mt) {
  KCV_TRY(check_mt(mt));
  return kleidicv_median_blur_u8(src, src_stride, dst, dst_stride, width, height,
                                 channels, kernel_width, kernel_height, border_type);
}

kleidicv_error_t kleidicv_thread_resize_linear_u8(
    const uint8_t *src, size_t src_stride, size_t src_width, size_t src_height,
    uint8_t *dst, size_t dst_stride, size_t dst_width, size_t dst_height,
    size_t channels, kleidicv_thread_multithreading mt) {
  KCV_TRY(check_mt(mt));
  return kleidicv_resize_linear_u8(src, src_stride, src_width, src_height, dst,
                                   dst_stride, dst_width, dst_height, channels);
}

kleidicv_error_t kleidicv_thread_remap_s16point5_u8(
    const uint8_t *src, size_t src_stride, size_t src_width, size_t src_height,
    uint8_t *dst, size_t dst_stride, size_t dst_width, size_t dst_height,
    size_t channels, const int16_t *mapxy, size_t mapxy_stride,
    const uint16_t *mapfrac, size_t mapfrac_stride,
    kleidicv_border_type_t border_type, const uint8_t *border_value,
    kleidicv_thread_multithreading mt) {
  KCV_TRY(check_mt(mt));
  return remap_s16point5_impl(src, src_stride, src_width, src_height, dst, dst_stride,
                              dst_width, dst_height, channels, mapxy, mapxy_stride,
                              mapfrac, mapfrac_stride, border_type, border_value,
                              nullptr, true);
}

kleidicv_error_t kleidicv_thread_remap_f32_u8(
    const uint8_t *src, size_t src_stride, size_t src_width, size_t src_height,
    uint8_t *dst, size_t dst_stride, size_t dst_width, size_t dst_height,
    size_t channels, const float *mapx, size_t mapx_stride, const float *mapy,
    size_t mapy_stride, kleidicv_interpolation_type_t interpolation,
    kleidicv_border_type_t border_type, const uint8_t *border_value,
    kleidicv_thread_multithreading mt) {
  KCV_TRY(check_mt(mt));
  return remap_f32_impl(src, src_stride, src_width, src_height, dst, dst_stride,
                        dst_width, dst_height, channels, mapx, mapx_stride, mapy,
                        mapy_stride, interpolation, border_type, border_value, nullptr,
                        true);
}

kleidicv_error_t kleidicv_thread_warp_perspective_u8(
    const uint8_t *src, size_t src_stride, size_t src_width, size_t src_height,
    uint8_t *dst, size_t dst_stride, size_t dst_width, size_t dst_height,
    const float transformation[9], size_t channels,
    kleidicv_interpolation_type_t interpolation, kleidicv_border_type_t border_type,
    const uint8_t *border_value, kleidicv_thread_multithreading mt) {
  KCV_TRY(check_mt(mt));
  return kleidicv_warp_perspective_u8(src, src_stride, src_width, src_height, dst,
                                      dst_stride, dst_width, dst_height, transformation,
                                      channels, interpolation, border_type,
                                      border_value);
}

kleidicv_error_t kleidicv_thread_yuv_semiplanar_to_rgb_u8(
    const uint8_t *src_y, size_t src_y_stride, const uint8_t *src_uv,
    size_t src_uv_stride, uint8_t *dst, size_t dst_stride, size_t width,
    size_t height, kleidicv_color_conversion_t color_format,
    kleidicv_thread_multithreading mt) {
  KCV_TRY(check_mt(mt));
  return yuv_sp_impl(src_y, src_y_stride, src_uv, src_uv_stride, dst, dst_stride, width,
                     height, color_format, nullptr, true);
}

// Device-resident API.
kleidicv_error_t kleidicv_hip_gaussian_blur_u8(
    const uint8_t *src, size_t src_stride, uint8_t *dst, size_t dst_stride,
    size_t width, size_t height, size_t channels, size_t kernel_width,
    size_t kernel_height, float sigma_x, float sigma_y,
    kleidicv_border_type_t border_type, void *stream) {
  return gaussian_impl(src, src_stride, dst, dst_stride, width, height, channels,
                       kernel_width, kernel_height, sigma_x, sigma_y, border_type,
                       stream, false);
}

kleidicv_error_t kleidicv_hip_separable_filter_2d_u8(
    const uint8_t *src, size_t src_stride, uint8_t *dst, size_t dst_stride,
    size_t width, size_t height, size_t channels, const uint8_t *kernel_x,
    size_t kernel_width, const uint8_t *kernel_y, size_t kernel_height,
    kleidicv_border_type_t border_type, void *stream) {
  return separable_impl(src, src_stride, dst, dst_stride, width, height, channels,
                        kernel_x, kernel_width, kernel_y, kernel_height, border_type,
                        stream, false);
}

kleidicv_error_t kleidicv_hip_rgb_to_yuv_u8(const uint8_t *src, size_t src_stride,
                                            uint8_t *dst, size_t dst_stride, size_t width,
                                            size_t height,
                                            kleidicv_color_conversion_t color_format,
                                            void *stream) {
  return rgb_to_yuv_impl(src, src_stride, dst, dst_stride, width, height, color_format, stream,
                         false);
}

kleidicv_error_t kleidicv_hip_rgb_to_yuv_semiplanar_u8(
    const uint8_t *src, size_t src_stride, uint8_t *y_dst, size_t y_stride, uint8_t *uv_dst,
    size_t uv_stride, size_t width, size_t height, kleidicv_color_conversion_t color_format,
    void *stream) {
  return rgb_to_yuv_sp_impl(src, src_stride, y_dst, y_stride, uv_dst, uv_stride, width, height,
                            color_format, stream, false);
}

kleidicv_error_t kleidicv_hip_yuv_to_rgb_u8(const uint8_t *src, size_t src_stride,
                                            uint8_t *dst, size_t dst_stride, size_t width,
                                            size_t height,
                                            kleidicv_color_conversion_t color_format,
                                            void *stream) {
  return yuv_impl(src, src_stride, dst, dst_stride, width, height, color_format, stream, false);
}

kleidicv_error_t kleidicv_hip_blur_and_downsample_u8(const uint8_t *src, size_t src_stride,
                                                     size_t src_width, size_t src_height,
                                                     uint8_t *dst, size_t dst_stride,
                                                     size_t channels,
                                                     kleidicv_border_type_t border_type,
                                                     void *stream) {
  return blur_down_impl(src, src_stride, src_width, src_height, dst, dst_stride, channels,
                        border_type, stream, false);
}

kleidicv_error_t kleidicv_hip_median_blur_u8(
    const uint8_t *src, size_t src_stride, uint8_t *dst, size_t dst_stride,
    size_t width, size_t height, size_t channels, size_t kernel_width,
    size_t kernel_height, kleidicv_border_type_t border_type, void *stream) {
  return median_impl(src, src_stride, dst, dst_stride, width, height, channels,
                     kernel_width, kernel_height, border_type, stream, false);
}

kleidicv_error_t kleidicv_hip_resize_linear_u8(
    const uint8_t *src, size_t src_stride, size_t src_width, size_t src_height,
    uint8_t *dst, size_t dst_stride, size_t dst_width, size_t dst_height,
    size_t channels, void *stream) {
  return resize_impl(src, src_stride, src_width, src_height, dst, dst_stride, dst_width,
                     dst_height, channels, stream, false);
}

kleidicv_error_t kleidicv_hip_remap_s16point5_u8(
    const uint8_t *src, size_t src_stride, size_t src_width, size_t src_height,
    uint8_t *dst, size_t dst_stride, size_t dst_width, size_t dst_height,
    size_t channels, const int16_t *mapxy, size_t mapxy_stride,
    const uint16_t *mapfrac, size_t mapfrac_stride,
    kleidicv_border_type_t border_type, const uint8_t *border_value, void *stream) {
  return remap_s16point5_impl(src, src_stride, src_width, src_height, dst, dst_stride,
                              dst_width, dst_height, channels, mapxy, mapxy_stride,
                              mapfrac, mapfrac_stride, border_type, border_value, stream,
                              false);
}

kleidicv_error_t kleidicv_hip_remap_f32_u8(
    const uint8_t *src, size_t src_stride, size_t src_width, size_t src_height,
    uint8_t *dst, size_t dst_stride, size_t dst_width, size_t dst_height,
    size_t channels, const float *mapx, size_t mapx_stride, const float *mapy,
    size_t mapy_stride, kleidicv_interpolation_type_t interpolation,
    kleidicv_border_type_t border_type, const uint8_t *border_value, void *stream) {
  return remap_f32_impl(src, src_stride, src_width, src_height, dst, dst_stride,
                        dst_width, dst_height, channels, mapx, mapx_stride, mapy,
                        mapy_stride, interpolation, border_type, border_value, stream,
                        false);
}

kleidicv_error_t kleidicv_hip_warp_perspective_u8(
    const uint8_t *src, size_t src_stride, size_t src_width, size_t src_height,
    uint8_t *dst, size_t dst_stride, size_t dst_width, size_t dst_height,
    const float transformation[9], size_t channels,
    kleidicv_interpolation_type_t interpolation, kleidicv_border_type_t border_type,
    const uint8_t *border_value, void *stream) {
  return warp_impl(src, src_stride, src_width, src_height, dst, dst_stride, dst_width,
                   dst_height, transformation, channels, interpolation, border_type,
                   border_value, stream, false);
}

kleidicv_error_t kleidicv_hip_yuv_semiplanar_to_rgb_u8(
    const uint8_t *src_y, size_t src_y_stride, const uint8_t *src_uv,
    size_t src_uv_stride, uint8_t *dst, size_t dst_stride, size_t width,
    size_t height, kleidicv_color_conversion_t color_format, void *stream) {
  return yuv_sp_impl(src_y, src_y_stride, src_uv, src_uv_stride, dst, dst_stride, width,
                     height, color_format, stream, false);
}

void *kleidicv_hip_default_stream(void) { return runtime().stream; }

kleidicv_error_t kleidicv_hip_synchronize(void *stream) {
  if (!runtime().ok) return KLEIDICV_ERROR_NOT_IMPLEMENTED;
  return from_hip(hipStreamSynchronize(static_cast<hipStream_t>(stream)));
}

}  // extern "C"

// Test hook: drops the drop-in staging buffer, so that the next synchronous
// call allocates its staging again (allocation-failure tests).
extern "C" void kcv_test_release_staging() {
  StagingArena &a = staging_arena();
  std::lock_guard<std::mutex> g(a.mutex);
  if (a.ptr) {
    (void)hipFreeAsync(a.ptr, kcv::runtime().stream);
    (void)hipStreamSynchronize(kcv::runtime().stream);
  }
  a.ptr = nullptr;
  a.bytes = 0;
}
