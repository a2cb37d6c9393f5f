// Internal interfaces between the C ABI layer (api.cpp) and the HIP kernel
// launchers.  Launchers take validated arguments and device pointers, enqueue
// on `stream` and return hipSuccess or the launch error.
#ifndef KCV_INTERNAL_H
#define KCV_INTERNAL_H

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace kcv {

enum Border : int {
  kConstant = 0,
  kReplicate = 1,
  kReflect = 2,
  kWrap = 3,
  kReverse = 4,
};

// Separable filter flavours sharing one tiled kernel.
enum SepMode : int {
  kSepBinomial = 0,   // single-rounded 2-D integer convolution, >> shift
  kSepSigma = 1,      // vertical pass rounded to u8, then horizontal
  kSepSaturate = 2,   // min(255, sum) with unsigned taps, no shift
};

// Output rows of a separable-filter tile of a single-image launch for a
// K x K kernel: 8 up to 7x7 (twice the workgroups of 16-row tiles, which a
// 1920x1080 image needs to spread over all CUs; the window's 2R halo rows
// stay few), 16 above.  Batches keep 16-row tiles.
constexpr int sep_tile_rows(int k) { return k <= 7 ? 8 : 16; }

// Tile order for 1-D grids: the hardware dispatches workgroup b to XCD
// b % 8, so tile index = this XCD's contiguous slice of the row-major tile
// list.  Neighbouring tiles (and their shared halo lines) then stay in one
// XCD's L2 instead of being fetched once per XCD.
__device__ __forceinline__ int xcd_tile(int b, int n) {
  const int x = b & 7, slot = b >> 3;
  const int per = n >> 3, rem = n & 7;
  return x * per + (x < rem ? x : rem) + slot;
}

// Output stores.  Each output byte is written once and read by no later
// kernel of the same single-image call, so whole-vector output stores carry
// the non-temporal hint: they stream through the caches instead of
// displacing cached inputs (a kernel with the 4K 2x upscale's traffic: 33.6 ->
// 31.3 us).
__device__ __forceinline__ void store_nt(uint32_t *p, uint32_t v) {
  __builtin_nontemporal_store(v, p);
}
__device__ __forceinline__ void store_nt(uint2 *p, uint2 v) {
  typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));
  __builtin_nontemporal_store(u32x2{v.x, v.y}, reinterpret_cast<u32x2 *>(p));
}
__device__ __forceinline__ void store_nt(uint4 *p, uint4 v) {
  typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
  __builtin_nontemporal_store(u32x4{v.x, v.y, v.z, v.w}, reinterpret_cast<u32x4 *>(p));
}

// Vector accesses at any byte address (gfx950 vector memory runs in
// unaligned mode): rows whose stride is not a multiple of the vector size
// still move in 4-, 8- or 16-byte accesses instead of bytes.
typedef uint32_t __attribute__((aligned(1))) u32_any;
typedef uint2 __attribute__((aligned(1))) u2_any;
typedef uint4 __attribute__((aligned(1))) u4_any;

// Destination tile of the gather kernels (remap s16point5, remap f32, warp
// perspective), per-image and batched: 256 threads of 4 horizontally
// adjacent pixels.  One-channel tiles are 128 x 8 so that each tile row is
// a whole 128-byte line (a 64-pixel row wrote half lines, which reached HBM
// as 1.37x the output bytes); wider pixels keep 64 x 16.
__host__ __device__ constexpr int gather_tile_w(int channels) { return channels == 1 ? 128 : 64; }
__host__ __device__ constexpr int gather_tile_h(int channels) {
  return 1024 / gather_tile_w(channels);
}

// Occupancy floor of the 4-channel float gather kernels (warp perspective,
// remap with float maps): 8 workgroups of 4 waves per CU, i.e. at most 64
// VGPRs and 8 waves per SIMD.  Their exact float lerps keep the VALU busy,
// and more resident waves hide more of the gather latency (4K RGBA warp:
// 26.8 -> 24.7 us against the compiler's 68-VGPR choice).  Fewer channels
// would spill at 64 VGPRs and keep the compiler's choice (1: no floor).
__host__ __device__ constexpr int float_gather_blocks_per_cu(int channels) {
  return channels == 4 ? 8 : 1;
}

// Largest 1-D grid used with xcd_tile (blocks of 256 threads).
constexpr int64_t kMaxLinearTiles = int64_t{1} << 23;

struct SepArgs {
  const uint8_t *src;
  size_t src_stride;
  uint8_t *dst;
  size_t dst_stride;
  int width, height, channels, border;
  int ksize;
  int mode;
  int shift;
  int wy[21];
  int wx[21];
};

hipError_t launch_separable(const SepArgs &a, hipStream_t stream);
// Gaussian blur of any odd size up to 255 (REPLICATE border): a vertical
// pass into an intermediate u8 image `tmp` (same size and stride as dst),
// then a horizontal pass, each (sum + 128) >> 8 with the same symmetric taps.
struct GaussArbArgs {
  const uint8_t *src;
  size_t src_stride;
  uint8_t *dst;
  size_t dst_stride;
  uint8_t *tmp;
  int width, height, channels, ksize;
  uint8_t taps[256];
};

hipError_t launch_gaussian_arbitrary(const GaussArbArgs &a, hipStream_t stream);

// Blur and downsample: SepArgs with the source size in width/height; the
// destination is ((width+1)/2) x ((height+1)/2).
hipError_t launch_blur_down(const SepArgs &a, hipStream_t stream);

struct MedianArgs {
  const uint8_t *src;
  size_t src_stride;
  uint8_t *dst;
  size_t dst_stride;
  int width, height, channels, border, ksize;
};

hipError_t launch_median(const MedianArgs &a, hipStream_t stream);

struct ResizeArgs {
  const uint8_t *src;
  size_t src_stride;
  int src_width, src_height;
  uint8_t *dst;
  size_t dst_stride;
  int dst_width, dst_height, channels;
};

// Path selection follows src/resize/resize_linear_api.cpp:82-216.  The
// generic path needs a per-column table; `workspace` must hold
// resize_workspace_bytes() bytes of device memory.
size_t resize_workspace_bytes(const ResizeArgs &a);
hipError_t launch_resize(const ResizeArgs &a, void *workspace,
                         hipStream_t stream);

struct RemapS16Args {
  const uint8_t *src;
  size_t src_stride;
  int src_width, src_height;
  uint8_t *dst;
  size_t dst_stride;
  int dst_width, dst_height, channels;
  const int16_t *mapxy;
  size_t mapxy_stride;
  const uint16_t *mapfrac;
  size_t mapfrac_stride;
  int border;
  uint8_t border_value[4];
};

hipError_t launch_remap_s16point5(const RemapS16Args &a, hipStream_t stream);

struct RemapF32Args {
  const uint8_t *src;
  size_t src_stride;
  int src_width, src_height;
  uint8_t *dst;
  size_t dst_stride;
  int dst_width, dst_height, channels;
  const float *mapx;
  size_t mapx_stride;
  const float *mapy;
  size_t mapy_stride;
  int interpolation, border;
  uint8_t border_value[4];
};

hipError_t launch_remap_f32(const RemapF32Args &a, hipStream_t stream);

struct WarpArgs {
  const uint8_t *src;
  size_t src_stride;
  int src_width, src_height;
  uint8_t *dst;
  size_t dst_stride;
  int dst_width, dst_height, channels;
  float t[9];
  int interpolation, border;
  int check_zero;  // transform_common.h:17-43 decided on the host
  uint8_t border_value[4];
};

hipError_t launch_warp_perspective(const WarpArgs &a, hipStream_t stream);

struct YuvArgs {
  const uint8_t *y;
  size_t y_stride;
  const uint8_t *uv;
  size_t uv_stride;
  uint8_t *dst;
  size_t dst_stride;
  int width, height;
  int vu, bgr, alpha;
};

hipError_t launch_yuv_sp_to_rgb(const YuvArgs &a, hipStream_t stream);

// One-buffer YUV layouts: planar 4:2:0 (I420 / YV12), interleaved 4:2:2 and
// interleaved 4:4:4.
enum { kYuv420P = 0, kYuv422 = 1, kYuv444 = 2 };

struct YuvXArgs {
  const uint8_t *src;  // luma plane (4:2:0) or the interleaved image
  size_t src_stride;
  size_t src_extent;   // bytes from src that reads may touch
  const uint8_t *u, *v;  // 4:2:0 chroma plane starts
  int u_phase, v_phase;  // 4:2:0: 1 when a plane starts half-way along a row
  uint8_t *dst;
  size_t dst_stride;
  int width, height;
  int layout, bgr, alpha;
  int y_off, u_off, v_off;  // 4:2:2 byte positions within a pixel pair
};

hipError_t launch_yuv_to_rgb(const YuvXArgs &a, hipStream_t stream);

// RGB/BGR(A) to YUV.  4:2:0 writes luma to dst and chroma either to the
// interleaved plane uv (semi-planar) or into the chroma area that follows
// the luma plane (planar, uv = dst + height * dst_stride).
enum { kRgb420SP = 3 };

struct RgbYuvArgs {
  const uint8_t *src;
  size_t src_stride;
  uint8_t *dst;
  size_t dst_stride;
  uint8_t *uv;
  size_t uv_stride;
  int width, height;
  int layout;                 // kYuv420P, kYuv422, kYuv444, kRgb420SP
  int scn, bgr, vu;           // source channels, B first, V first
  int y_off, u_off, v_off;    // 4:2:2 byte positions within a pixel pair
};

hipError_t launch_rgb_to_yuv(const RgbYuvArgs &a, hipStream_t stream);

}  // namespace kcv

#endif  // KCV_INTERNAL_H
