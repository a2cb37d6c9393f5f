// Per-image launches of the linear resize (bodies in resize_body.h), and the
// resize entry point that allocates nothing.

#include <mutex>

#include "kleidicv_hip/kleidicv.h"
#include "prepare.h"
#include "resize_body.h"

namespace kcv {
namespace {

using namespace rsz;

__global__ __launch_bounds__(kThreads) void resize_tables_kernel(GenericPlan p, int2 *cols,
                                                                 int2 *rows) {
  resize_tables_elem(p, cols, rows, static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x);
}

template <int F>
__global__ __launch_bounds__(kThreads) void exact_tables_kernel(ResizeArgs a, int2 *cols,
                                                                int2 *rows) {
  exact_tables_elem<F>(a, cols, rows, blockIdx.x * kThreads + threadIdx.x);
}

template <int MODE>
__global__ __launch_bounds__(kThreads) void resize_staged_kernel(
    ResizeArgs a, const int2 *__restrict__ cols, const int2 *__restrict__ rows, int band) {
  __shared__ __align__(16) uint8_t smem[kStagedSmem];
  resize_staged_tile<MODE>(a, cols, rows, band, blockIdx.x, blockIdx.y, smem);
}

// The generic lerp when it reads straight from global memory: no LDS, so
// registers alone set the occupancy.
__global__ __launch_bounds__(kThreads) void resize_direct_kernel(
    ResizeArgs a, const int2 *__restrict__ cols, const int2 *__restrict__ rows, int band) {
  const int y0 = blockIdx.y * band;
  resize_direct_rows<kLerpGeneric>(a, cols, rows, blockIdx.x * kTileE + threadIdx.x * 4, y0,
                                   min(y0 + band, a.dst_height), 1);
}

__global__ __launch_bounds__(kThreads) void resize_up2x4_kernel(ResizeArgs a) {
  for (int g = blockIdx.y; g <= a.src_height; g += gridDim.y) up2x4_group(a, blockIdx.x, g);
}

template <int C>
__global__ __launch_bounds__(kThreads) void resize_quarter_kernel(ResizeArgs a) {
  resize_quarter_rows<C>(a, blockIdx.x, threadIdx.x, kThreads, blockIdx.y, a.dst_height,
                         gridDim.y);
}

template <int MODE>
hipError_t launch_staged(const ResizeArgs &a, const int2 *cols, const int2 *rows,
                         hipStream_t stream) {
  const int64_t n = int64_t{a.dst_width} * a.channels;
  const bool direct = MODE == kLerpGeneric && generic_reads_direct(a);
  const int band = direct ? direct_band_rows() : band_rows(a);
  const dim3 grid(static_cast<unsigned>((n + kTileE - 1) / kTileE),
                  static_cast<unsigned>((a.dst_height + band - 1) / band));
  if (direct)
    resize_direct_kernel<<<grid, kThreads, 0, stream>>>(a, cols, rows, band);
  else
    resize_staged_kernel<MODE><<<grid, kThreads, 0, stream>>>(a, cols, rows, band);
  return hipGetLastError();
}

}  // namespace

size_t resize_workspace_bytes(const ResizeArgs &a) {
  const Path p = pick(a);
  if (p == Path::kNone || p == Path::kQuarter || p == Path::kUp2x4) return 0;
  return (static_cast<size_t>(a.dst_width) * a.channels + a.dst_height) *
         sizeof(int2);
}

hipError_t launch_resize(const ResizeArgs &a, void *workspace,
                         hipStream_t stream) {
  const Path path = pick(a);
  if (path == Path::kNone || a.dst_width == 0 || a.dst_height == 0) return hipSuccess;
  if (path == Path::kQuarter) {
    const int64_t n = int64_t{a.dst_width} * a.channels;
    const int64_t per = int64_t{quarter_thread_bytes(a.channels)} * kThreads;
    const dim3 grid(static_cast<unsigned>((n + per - 1) / per),
                    static_cast<unsigned>(a.dst_height < 65535 ? a.dst_height : 65535));
    switch (a.channels) {
      case 1: resize_quarter_kernel<1><<<grid, kThreads, 0, stream>>>(a); break;
      case 2: resize_quarter_kernel<2><<<grid, kThreads, 0, stream>>>(a); break;
      case 3: resize_quarter_kernel<3><<<grid, kThreads, 0, stream>>>(a); break;
      default: resize_quarter_kernel<4><<<grid, kThreads, 0, stream>>>(a); break;
    }
    return hipGetLastError();
  }
  if (path == Path::kUp2x4) {
    const int groups = a.src_height + 1;
    const int px = kUp2x4ThreadPx * kThreads;
    const dim3 grid(static_cast<unsigned>((a.src_width + px - 1) / px),
                    static_cast<unsigned>(groups < 65535 ? groups : 65535));
    resize_up2x4_kernel<<<grid, kThreads, 0, stream>>>(a);
    return hipGetLastError();
  }
  const int64_t n = int64_t{a.dst_width} * a.channels;
  int2 *cols = static_cast<int2 *>(workspace);
  int2 *rows = cols + n;
  const int64_t work = n > a.dst_height ? n : a.dst_height;
  const unsigned tgrid = static_cast<unsigned>((work + kThreads - 1) / kThreads);
  if (path == Path::kUp2) {
    exact_tables_kernel<2><<<tgrid, kThreads, 0, stream>>>(a, cols, rows);
    return launch_staged<kExact2>(a, cols, rows, stream);
  }
  if (path == Path::kUp4) {
    exact_tables_kernel<4><<<tgrid, kThreads, 0, stream>>>(a, cols, rows);
    return launch_staged<kExact4>(a, cols, rows, stream);
  }
  const GenericPlan p = make_plan(a);
  resize_tables_kernel<<<tgrid, kThreads, 0, stream>>>(p, cols, rows);
  return launch_staged<kLerpGeneric>(a, cols, rows, stream);
}

}  // namespace kcv

// ---------------------------------------------------------------------------
// No-allocation resize.  The coordinate tables of the generic and exact
// paths live in `scratch`, a fixed array that every GPU holds from module
// load until the process ends.  In host code `scratch` names a host-side
// stand-in, so its device address is looked up once per device and only
// that address is passed to the kernels.  The tables kernel writes every
// entry the lerp kernel reads, on the same stream, so nothing needs
// clearing; two calls that overlap on one GPU (different streams, or
// different host threads) share the array and race.
namespace {

constexpr size_t kNoAllocScratchBytes = KLEIDICV_HIP_NOALLOC_SCRATCH_BYTES;
constexpr int kMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

struct ScratchSlot {
  std::once_flag once;
  void *ptr = nullptr;
  hipError_t err = hipSuccess;
};

// Device address of this device's `scratch` (the current device), or null.
void *scratch_address(int device) {
  static ScratchSlot slots[kMaxDevices];
  if (device < 0 || device >= kMaxDevices) return nullptr;
  ScratchSlot &s = slots[device];
  std::call_once(s.once, [&s] {
    void *p = nullptr;
    s.err = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (s.err == hipSuccess) s.ptr = p;
    else (void)hipGetLastError();
  });
  return s.ptr;
}

}  // namespace

extern "C" kleidicv_error_t kleidicv_hip_resize_linear_u8_noalloc(
    const uint8_t *src, size_t src_stride, size_t src_width, size_t src_height, uint8_t *dst,
    size_t dst_stride, size_t dst_width, size_t dst_height, size_t channels, void *stream) {
  kcv::ResizeArgs a;
  kcv::Action act;
  KCV_TRY(kcv::prepare_resize(src, src_stride, src_width, src_height, dst, dst_stride,
                              dst_width, dst_height, channels, &a, &act));
  if (act == kcv::Action::kNone) return KLEIDICV_OK;
  const size_t need = kcv::resize_workspace_bytes(a);
  if (need > kNoAllocScratchBytes) return KLEIDICV_ERROR_RANGE;
  void *ws = nullptr;
  if (need > 0) {
    int device = 0;
    KCV_TRY(kcv::from_hip(hipGetDevice(&device)));
    ws = scratch_address(device);
    if (ws == nullptr) return KLEIDICV_ERROR_NOT_IMPLEMENTED;
  }
  return kcv::from_hip(kcv::launch_resize(a, ws, static_cast<hipStream_t>(stream)));
}
