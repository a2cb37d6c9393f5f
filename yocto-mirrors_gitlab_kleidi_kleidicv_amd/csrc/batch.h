// Batch engine internals shared by the persistent stage kernels (batch.hip)
// and the host side that builds their work lists (batch_api.cpp).
//
// A batch is cut into segments, one per (stage, kernel family).  A segment
// is a contiguous run of jobs plus, per job, the index of its first work
// item; item i of the segment belongs to the last job whose first item is
// <= i.  One persistent kernel per segment walks the items: workgroup b of
// G takes items b, b + G, ... or, in XCD order, a contiguous slice of the
// XCD it runs on (the hardware places workgroup b on XCD b % 8).
#ifndef KCV_BATCH_H
#define KCV_BATCH_H

#include "kcv_internal.h"
#include "resize_body.h"

namespace kcv {

enum Family : int {
  kFamTables = 0,  // resize coordinate tables (always run before stage 0)
  kFamCopy,        // plain 2-D copy (Gaussian with a sigma too small to blur)
  kFamSep,         // Gaussian and separable filters
  kFamMedian,
  kFamResize,      // generic, exact 2x / 4x and quarter resize
  kFamRemapS16,
  kFamRemapF32,
  kFamWarp,
  kFamYuv,
  kFamCount,
};

// Resize path of a job (rsz::Path as an int so the job stays trivially
// copyable on both sides).
// kRsDirect is a kernel variant only: a kRsGeneric job whose lerp reads
// straight from global memory (rsz::generic_reads_direct).
enum ResizeKind : int { kRsGeneric = 0, kRsExact2, kRsExact4, kRsQuarter, kRsUp2x4, kRsDirect };

struct CopyArgs {
  const uint8_t *src;
  size_t src_stride;
  uint8_t *dst;
  size_t dst_stride;
  int64_t row_bytes;
  int height;
};

struct TablesJob {
  ResizeArgs a;
  rsz::GenericPlan plan;
  int kind;  // ResizeKind (not kRsQuarter)
  int2 *cols;
  int2 *rows;
};

struct ResizeJob {
  ResizeArgs a;
  int kind;  // ResizeKind
  const int2 *cols;
  const int2 *rows;
};

struct Job {
  int variant;  // kernel variant within the family
  int tiles_x;  // work items per row of tiles
  int band;     // rows per work item (resize, copy, NV12 row pairs)
  int pad;
  union {
    CopyArgs copy;
    SepArgs sep;
    MedianArgs median;
    TablesJob tables;
    ResizeJob resize;
    RemapS16Args remap_s16;
    RemapF32Args remap_f32;
    WarpArgs warp;
    YuvArgs yuv;
  } u;
};

struct Segment {
  const Job *jobs;
  const uint32_t *first;  // first work item of each job, relative
  int njobs;
  uint32_t nitems;
  int xcd_order;
};

// Work-item geometry shared by the host (item counts) and device (tile
// decoding).
constexpr int kFilterTileW = 256;  // bytes
constexpr int kFilterTileH = 16;   // rows
constexpr int kResizeTileE = 1024;            // destination bytes (staged)
constexpr int kUp2x4TilePx = rsz::kUp2x4ThreadPx * 256;  // source pixels (exact 2x, 4 ch)
constexpr int kYuvTileW = 8 * 256;            // pixels
constexpr int kCopyTileE = 16 * 256;          // bytes
constexpr int kTablesPerItem = 256;           // table entries

// Kernel of a (family, variant) pair, or null when the pair does not exist.
// Variants: separable mode * 32 + K, median K, tables and resize ResizeKind,
// remap s16point5 C * 8 + border, remap f32 and warp C * 8 + interpolation *
// 4 + border, NV12 1 for four output channels, copy 0.
const void *segment_kernel(Family family, int variant);
hipError_t launch_segment(Family family, int variant, const Segment &seg, int grid,
                          hipStream_t stream);
// Workgroups of that kernel one compute unit keeps resident at once (its
// register and LDS footprint decide), or 0 when the pair does not exist or
// the runtime cannot say.
int segment_blocks_per_cu(Family family, int variant);

}  // namespace kcv

#endif  // KCV_BATCH_H
