// Persistent stage kernels of the batch engine.  Each kernel walks the work
// items of one segment (tiles of many images, each with its own sizes,
// channels and border) and runs the same tile bodies as the per-image
// kernels.  There is one kernel per (family, variant) so that every
// instantiation keeps the register footprint of its per-image counterpart.
//
// Workgroups loop until the segment's items are exhausted; every workgroup
// reaches the end of the list, and the only barriers are inside the tile
// bodies, which every thread of the workgroup enters.

#include "batch.h"
#include "filters_body.h"
#include "resize_body.h"
#include "transform_body.h"
#include "yuv_body.h"

namespace kcv {
namespace {

constexpr int kThreads = 256;

// Last job whose first item is <= item (first[0] == 0).
__device__ __forceinline__ int find_job(const Segment &s, uint32_t item) {
  int lo = 0, hi = s.njobs - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (s.first[mid] <= item) lo = mid; else hi = mid - 1;
  }
  return lo;
}

// Calls body(job, bx, by) for each item of this workgroup's share of the
// segment: a contiguous run of ceil(nitems / grid) items, so that consecutive
// tiles of an image stay on one CU.  With XCD order the runs of the
// workgroups on one XCD (workgroup b runs on XCD b mod 8) are adjacent, so
// neighbouring tiles and their halos share that XCD's L2.  The job is found
// by one binary search, then followed item by item; jobs are read through
// the constant address space: they are never written while a segment runs,
// so their fields load as scalars.
template <class Body>
__device__ __forceinline__ void for_each_item(const Segment &s, const Body &body) {
  typedef const __attribute__((address_space(4))) Job ConstJob;
  const uint32_t grid = gridDim.x;
  const uint32_t per = (s.nitems + grid - 1) / grid;
  const uint32_t w = s.xcd_order ? static_cast<uint32_t>(xcd_tile(blockIdx.x, grid)) : blockIdx.x;
  uint32_t item = w * per;
  const uint32_t end = min(item + per, s.nitems);
  if (item >= end) return;
  int j = find_job(s, item);
  int bx = 0, by = 0, tiles_x = 1;
  bool fresh = true;
  for (; item < end; ++item) {
    if (j + 1 < s.njobs && s.first[j + 1] <= item) {
      ++j;
      fresh = true;
    }
    const Job &job = *(const Job *)((ConstJob *)s.jobs + j);
    if (fresh) {
      tiles_x = job.tiles_x;
      const int local = static_cast<int>(item - s.first[j]);
      by = local / tiles_x;
      bx = local - by * tiles_x;
      fresh = false;
    }
    body(job, bx, by);
    if (++bx == tiles_x) {
      bx = 0;
      ++by;
    }
  }
}

template <int MODE, int K, int C>
__global__ __launch_bounds__(kThreads) void batch_sep_kernel(Segment s) {
  __shared__ __align__(16) uint8_t smem[filt::separable_smem_bytes<K, kFilterTileH>()];
  for_each_item(s, [&](const Job &job, int bx, int by) {
    filt::separable_tile<MODE, K, C, kFilterTileH>(job.u.sep, bx, by, smem);
  });
}

template <int K, int C>
__global__ __launch_bounds__(kThreads) void batch_median_kernel(Segment s) {
  __shared__ __align__(16) uint8_t smem[filt::median_smem_bytes<K>()];
  for_each_item(s, [&](const Job &job, int bx, int by) {
    filt::median_tile_c<K, C>(job.u.median, bx, by, smem);
  });
}

template <int KIND>
__global__ __launch_bounds__(kThreads) void batch_tables_kernel(Segment s) {
  for_each_item(s, [&](const Job &job, int bx, int) {
    const TablesJob &tj = job.u.tables;
    const int64_t e = static_cast<int64_t>(bx) * kTablesPerItem + threadIdx.x;
    if (KIND == kRsGeneric)
      rsz::resize_tables_elem(tj.plan, tj.cols, tj.rows, e);
    else
      rsz::exact_tables_elem<KIND == kRsExact2 ? 2 : 4>(tj.a, tj.cols, tj.rows,
                                                        static_cast<int>(e));
  });
}

template <int C>
__global__ __launch_bounds__(kThreads) void batch_quarter_kernel(Segment s) {
  for_each_item(s, [&](const Job &job, int bx, int by) {
    const ResizeJob &rj = job.u.resize;
    const int y0 = by * job.band;
    // Rows narrower than the workgroup's span take fewer threads each (a
    // power of two), and the other threads take the band's other rows.
    constexpr int OB = rsz::quarter_thread_bytes(C);
    const int groups = (rj.a.dst_width * C + OB - 1) / OB;
    int lanes = kThreads;
    while (lanes > 1 && lanes / 2 >= groups) lanes /= 2;
    const int rows = kThreads / lanes;
    rsz::resize_quarter_rows<C>(rj.a, bx, threadIdx.x % lanes, lanes,
                                y0 + static_cast<int>(threadIdx.x) / lanes,
                                min(y0 + job.band, rj.a.dst_height), rows);
  });
}

template <int KIND>
__global__ __launch_bounds__(kThreads) void batch_resize_kernel(Segment s) {
  __shared__ __align__(16) uint8_t smem[KIND >= kRsQuarter ? 16 : rsz::kStagedSmem];
  for_each_item(s, [&](const Job &job, int bx, int by) {
    const ResizeJob &rj = job.u.resize;
    if (KIND == kRsUp2x4) {
      rsz::up2x4_group(rj.a, bx, by);
    } else {
      constexpr int kMode = KIND == kRsGeneric  ? rsz::kLerpGeneric
                            : KIND == kRsExact2 ? rsz::kExact2
                                                : rsz::kExact4;
      // narrow rows: fewer threads per row, the rest on the band's other rows
      const int quads = (rj.a.dst_width * rj.a.channels + 3) / 4;
      int lanes = kThreads;
      while (lanes > 1 && lanes / 2 >= quads) lanes /= 2;
      rsz::resize_staged_tile<kMode>(rj.a, rj.cols, rj.rows, job.band, bx, by, smem, lanes);
    }
  });
}

// kRsDirect: generic lerp jobs that read straight from global memory.  No
// LDS, so the kernel's occupancy is not held to the staged window's.
__global__ __launch_bounds__(kThreads) void batch_resize_direct_kernel(Segment s) {
  for_each_item(s, [&](const Job &job, int bx, int by) {
    const ResizeJob &rj = job.u.resize;
    const int lanes = rsz::direct_lanes(rj.a);
    const int y0 = by * job.band;
    rsz::resize_direct_rows<rsz::kLerpGeneric>(
        rj.a, rj.cols, rj.rows, bx * rsz::kTileE + static_cast<int>(threadIdx.x % lanes) * 4,
        y0 + static_cast<int>(threadIdx.x / lanes), min(y0 + job.band, rj.a.dst_height),
        kThreads / lanes);
  });
}

template <int C, int BORDER>
__global__ __launch_bounds__(kThreads) void batch_remap_s16_kernel(Segment s) {
  for_each_item(s, [&](const Job &job, int bx, int by) {
    xform::remap_s16point5_tile<C, BORDER>(job.u.remap_s16, bx, by);
  });
}

template <int C, int INTERP, int BORDER>
__global__ __launch_bounds__(kThreads, float_gather_blocks_per_cu(C)) void batch_remap_f32_kernel(Segment s) {
  for_each_item(s, [&](const Job &job, int bx, int by) {
    xform::remap_f32_tile<C, INTERP, BORDER>(job.u.remap_f32, bx, by);
  });
}

template <int C, int INTERP, int BORDER>
__global__ __launch_bounds__(kThreads, float_gather_blocks_per_cu(C)) void batch_warp_kernel(Segment s) {
  for_each_item(s, [&](const Job &job, int bx, int by) {
    xform::warp_tile<C, INTERP, BORDER>(job.u.warp, bx, by);
  });
}

template <int OC>
__global__ __launch_bounds__(kThreads) void batch_yuv_kernel(Segment s) {
  for_each_item(s, [&](const Job &job, int bx, int by) {
    const int p0 = by * job.band;
    yuv::nv12_rows<OC>(job.u.yuv, bx, p0, p0 + job.band, 1);
  });
}

// Bytes [bx*kCopyTileE, +kCopyTileE) of rows [by*band, +band).
__global__ __launch_bounds__(kThreads) void batch_copy_kernel(Segment s) {
  for_each_item(s, [&](const Job &job, int bx, int by) {
    const CopyArgs &a = job.u.copy;
    const int64_t e0 = static_cast<int64_t>(bx) * kCopyTileE + threadIdx.x * 16;
    const int y0 = by * job.band, y1 = min(y0 + job.band, a.height);
    if (e0 >= a.row_bytes) return;
    const bool vec = e0 + 16 <= a.row_bytes;
    for (int y = y0; y < y1; ++y) {
      const uint8_t *sp = a.src + static_cast<size_t>(y) * a.src_stride + e0;
      uint8_t *dp = a.dst + static_cast<size_t>(y) * a.dst_stride + e0;
      if (vec) {
        const uint4 v = *reinterpret_cast<const u4_any *>(sp);
        if ((reinterpret_cast<uintptr_t>(dp) & 15) == 0)
          store_nt(reinterpret_cast<uint4 *>(dp), v);
        else
          *reinterpret_cast<u4_any *>(dp) = v;
      } else {
        for (int k = 0; k < 16 && e0 + k < a.row_bytes; ++k) dp[k] = sp[k];
      }
    }
  });
}

#define KCV_K(...) reinterpret_cast<const void *>(&__VA_ARGS__)

const void *gather_kernel(bool warp, int variant) {
  const int c = variant >> 3, i = (variant >> 2) & 1, b = variant & 3;
#define KCV_G(C, I, B)                                                   \
  if (c == C && i == I && b == B)                                       \
    return warp ? KCV_K(batch_warp_kernel<C, I, B>)                     \
                : KCV_K(batch_remap_f32_kernel<C, I, B>);
  KCV_G(1, 0, 0) KCV_G(1, 0, 1) KCV_G(1, 1, 0) KCV_G(1, 1, 1)
  KCV_G(2, 0, 0) KCV_G(2, 0, 1) KCV_G(2, 1, 0) KCV_G(2, 1, 1)
  KCV_G(3, 0, 0) KCV_G(3, 0, 1) KCV_G(3, 1, 0) KCV_G(3, 1, 1)
  KCV_G(4, 0, 0) KCV_G(4, 0, 1) KCV_G(4, 1, 0) KCV_G(4, 1, 1)
#undef KCV_G
  return nullptr;
}

}  // namespace

const void *segment_kernel(Family family, int variant) {
  switch (family) {
    case kFamTables:
      return variant == kRsGeneric  ? KCV_K(batch_tables_kernel<kRsGeneric>)
             : variant == kRsExact2 ? KCV_K(batch_tables_kernel<kRsExact2>)
             : variant == kRsExact4 ? KCV_K(batch_tables_kernel<kRsExact4>)
                                    : nullptr;
    case kFamCopy:
      return KCV_K(batch_copy_kernel);
    case kFamSep:
      switch (variant) {
#define KCV_S(MODE, K)                                                  \
  case (MODE * 32 + K) * 8 + 1: return KCV_K(batch_sep_kernel<MODE, K, 1>); \
  case (MODE * 32 + K) * 8 + 2: return KCV_K(batch_sep_kernel<MODE, K, 2>); \
  case (MODE * 32 + K) * 8 + 3: return KCV_K(batch_sep_kernel<MODE, K, 3>); \
  case (MODE * 32 + K) * 8 + 4: return KCV_K(batch_sep_kernel<MODE, K, 4>);
        KCV_S(kSepBinomial, 3) KCV_S(kSepBinomial, 5) KCV_S(kSepBinomial, 7)
        KCV_S(kSepBinomial, 9) KCV_S(kSepSigma, 3) KCV_S(kSepSigma, 5)
        KCV_S(kSepSigma, 7) KCV_S(kSepSigma, 9) KCV_S(kSepSigma, 15)
        KCV_S(kSepSigma, 21) KCV_S(kSepSaturate, 3) KCV_S(kSepSaturate, 5)
        KCV_S(kSepSaturate, 7)
#undef KCV_S
        default: return nullptr;
      }
    case kFamMedian:
      switch (variant) {
#define KCV_M(K, C) \
  case K * 8 + C: return KCV_K(batch_median_kernel<K, C>);
        KCV_M(3, 1) KCV_M(3, 2) KCV_M(3, 3) KCV_M(3, 4)
        KCV_M(5, 1) KCV_M(5, 2) KCV_M(5, 3) KCV_M(5, 4)
        KCV_M(7, 1) KCV_M(7, 2) KCV_M(7, 3) KCV_M(7, 4)
#undef KCV_M
        default: return nullptr;
      }
    case kFamResize:
      return variant == kRsGeneric   ? KCV_K(batch_resize_kernel<kRsGeneric>)
             : variant == kRsExact2  ? KCV_K(batch_resize_kernel<kRsExact2>)
             : variant == kRsExact4  ? KCV_K(batch_resize_kernel<kRsExact4>)
             : variant == kRsQuarter + 8 * 1 ? KCV_K(batch_quarter_kernel<1>)
             : variant == kRsQuarter + 8 * 2 ? KCV_K(batch_quarter_kernel<2>)
             : variant == kRsQuarter + 8 * 3 ? KCV_K(batch_quarter_kernel<3>)
             : variant == kRsQuarter + 8 * 4 ? KCV_K(batch_quarter_kernel<4>)
             : variant == kRsUp2x4   ? KCV_K(batch_resize_kernel<kRsUp2x4>)
             : variant == kRsDirect  ? KCV_K(batch_resize_direct_kernel)
                                     : nullptr;
    case kFamRemapS16:
      return variant == 1 * 8 + kReplicate   ? KCV_K(batch_remap_s16_kernel<1, kReplicate>)
             : variant == 1 * 8 + kConstant  ? KCV_K(batch_remap_s16_kernel<1, kConstant>)
             : variant == 4 * 8 + kReplicate ? KCV_K(batch_remap_s16_kernel<4, kReplicate>)
             : variant == 4 * 8 + kConstant  ? KCV_K(batch_remap_s16_kernel<4, kConstant>)
                                             : nullptr;
    case kFamRemapF32:
      return gather_kernel(false, variant);
    case kFamWarp:
      return gather_kernel(true, variant);
    case kFamYuv:
      return variant ? KCV_K(batch_yuv_kernel<4>) : KCV_K(batch_yuv_kernel<3>);
    default:
      return nullptr;
  }
}

int segment_blocks_per_cu(Family family, int variant) {
  const void *fn = segment_kernel(family, variant);
  int n = 0;
  if (fn == nullptr || hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, fn, kThreads, 0) !=
                           hipSuccess)
    return 0;
  return n;
}

hipError_t launch_segment(Family family, int variant, const Segment &seg, int grid,
                          hipStream_t stream) {
  if (seg.nitems == 0) return hipSuccess;
  const void *fn = segment_kernel(family, variant);
  if (fn == nullptr) return hipErrorInvalidValue;
  Segment arg = seg;
  void *args[] = {&arg};
  return hipLaunchKernel(fn, dim3(static_cast<unsigned>(grid)), dim3(kThreads), args, 0,
                         stream);
}

}  // namespace kcv
