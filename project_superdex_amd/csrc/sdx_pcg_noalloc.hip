// Allocation-free variant of the L1 solve (sdx_pcg_bsr3_solve_noalloc).
//
// The same block-Jacobi PCG as sdx_pcg_bsr3_solve, for callers that cannot
// allocate device memory (code under HIP graph capture, tight serving loops,
// harnesses that forbid hipMalloc).  Every device buffer the solve needs (the
// SoA matrix copy, the work vectors, tiles, accumulators and a staging area
// that host arrays pass through) is carved out of `scratch`, a file-scope
// device array of kScratchBytes that each GPU holds from the first call on it
// until the process ends.  Host code never uses the array's name as a device
// address: its per-device address comes from hipGetSymbolAddress once per
// device (with that device current) and only that address is passed to
// kernels and copies.  No hipMalloc, hipFree, stream creation or other
// allocating runtime call is made; the work runs on the null stream.
//
// Calls on one device share that device's array, which keeps the last call's
// contents: everything the solve reads is written by the call first (the CG
// state and accumulators are cleared by kernels).  Calls from threads of one
// process are serialised by a per-device mutex; processes have their own
// copies of the array.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <mutex>
#include <string>

#include "../../include/sdx_fem.h"
#include "sdx_pcg.h"
#include "sdx_world.h"

namespace sdx {
namespace {

constexpr size_t kScratchBytes = size_t(64) << 20;  // per device
constexpr size_t kStageBytes = size_t(8) << 20;     // head of the array: host arrays in transit
constexpr size_t kAlign = 256;
constexpr int kMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_dev_mutex[kMaxDevices];
std::mutex g_addr_mutex;
unsigned char* g_scratch_addr[kMaxDevices] = {};

// Device address of this device's `scratch` (the device must be current).
unsigned char* scratch_address(int device) {
  std::lock_guard<std::mutex> lock(g_addr_mutex);
  if (!g_scratch_addr[device]) {
    void* p = nullptr;
    hip_check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
    if (!p) throw DeviceError("hipGetSymbolAddress(scratch) returned no address");
    g_scratch_addr[device] = static_cast<unsigned char*>(p);
  }
  return g_scratch_addr[device];
}

// Bump allocation inside the solver part of the scratch array; `need` only
// counts when the base is null (the sizing pass).
struct Carve {
  unsigned char* base;
  size_t off = 0;
  template <class U>
  U* take(size_t count) {
    U* p = base ? reinterpret_cast<U*>(base + off) : nullptr;
    off += (std::max<size_t>(count, 1) * sizeof(U) + kAlign - 1) & ~(kAlign - 1);
    return p;
  }
};

__global__ void k_nt_tiles(int32_t n, int32_t rows, int32_t n_tiles, int32_t* env, int32_t* r0, int32_t* r1,
                           int32_t* first) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n_tiles) {
    env[t] = 0;
    r0[t] = t * rows;
    r1[t] = min(t * rows + rows, n);
  }
  if (t == 0) {
    first[0] = 0;
    first[1] = n_tiles;
  }
}

// Slot of each row's diagonal block (slot 0 where a row has none: the
// preconditioner is then off, and the slot is read but unused).
__global__ void k_nt_diag(int32_t n, const int32_t* __restrict__ ptr, const int32_t* __restrict__ col,
                          int32_t* __restrict__ diag) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int d = 0;
  for (int k = ptr[i]; k < ptr[i + 1]; ++k)
    if (col[k] == i) {
      d = k;
      break;
    }
  diag[i] = d;
}

// Blocks [k0, k1) (the blocks of rows [i0, i1)) from the staged caller
// values into 9 SoA planes of T.  Staged values start at block k0; in the
// row-strip layout a row is a row-major 3 x (3 nnz_row) matrix.
template <class V, class T>
__global__ void k_nt_planes(int32_t i0, int32_t i1, const int32_t* __restrict__ ptr, int32_t layout,
                            const V* __restrict__ staged, int32_t nnzb, T* __restrict__ vals) {
  const int k0 = ptr[i0], k1 = ptr[i1];
  const int k = k0 + blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= k1) return;
  int lo = i0, hi = i1 - 1;  // row of block k: the last i with ptr[i] <= k
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (ptr[mid] <= k)
      lo = mid;
    else
      hi = mid - 1;
  }
  const int rs = ptr[lo], nr = ptr[lo + 1] - rs, kk = k - rs;
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const size_t src = layout == SDX_BSR_BLOCK_ROWMAJOR
                             ? size_t(9) * size_t(k - k0) + size_t(3 * r + c)
                             : size_t(9) * size_t(rs - k0) + size_t(r) * size_t(3 * nr) + size_t(3 * kk + c);
      vals[size_t(3 * r + c) * size_t(nnzb) + size_t(k)] = T(staged[src]);
    }
}

template <class S, class D>
void upload_convert(const S* host, size_t n, D* dst, unsigned char* stage) {
  const size_t chunk = kStageBytes / sizeof(S);
  for (size_t i = 0; i < n; i += chunk) {
    const size_t m = std::min(chunk, n - i);
    hip_check(hipMemcpyAsync(stage, host + i, m * sizeof(S), hipMemcpyHostToDevice, nullptr), "hipMemcpyAsync");
    launch_convert<S, D>(nullptr, int64_t(m), reinterpret_cast<const S*>(stage), dst + i);
  }
}

template <class S, class D>
void download_convert(const S* src, size_t n, D* host, unsigned char* stage) {
  const size_t chunk = kStageBytes / sizeof(D);
  for (size_t i = 0; i < n; i += chunk) {
    const size_t m = std::min(chunk, n - i);
    launch_convert<S, D>(nullptr, int64_t(m), src + i, reinterpret_cast<D*>(stage));
    hip_check(hipMemcpyAsync(host + i, stage, m * sizeof(D), hipMemcpyDeviceToHost, nullptr), "hipMemcpyAsync");
  }
  hip_check(hipStreamSynchronize(nullptr), "hipStreamSynchronize");
}

// Carves the system out of `base` (null: only sizes it); returns the bytes.
template <class T, class A>
size_t layout_system(unsigned char* base, int32_t nbr, int32_t nnzb, PcgSystem<T, A>& d, int32_t* tile_ptrs[8]) {
  Carve c{base};
  const size_t n3 = size_t(3) * size_t(nbr);
  const int32_t ntv = (nbr + kVecThreads - 1) / kVecThreads, nta = (nbr + kSpmvRows - 1) / kSpmvRows;
  d.nbr = nbr;
  d.nnzb = nnzb;
  int32_t* ptr = c.take<int32_t>(size_t(nbr) + 1);
  int32_t* col = c.take<int32_t>(size_t(nnzb));
  int32_t* diag = c.take<int32_t>(size_t(nbr));
  T* vals = c.take<T>(size_t(9) * size_t(nnzb));
  T* b = c.take<T>(n3);
  d.ptr = ptr;
  d.col = col;
  d.diag = diag;
  d.vals = vals;
  d.b = b;
  d.x = c.take<A>(n3);
  d.dinv = c.take<T>(size_t(6) * size_t(nbr));
  d.r = c.take<T>(n3);
  d.z = c.take<T>(size_t(4) * size_t(nbr));
  d.p = c.take<T>(n3);
  d.sv = c.take<T>(n3);
  d.w = c.take<T>(n3);
  d.rref = c.take<double>(n3);
  for (int k = 0; k < 3; ++k) tile_ptrs[k] = c.take<int32_t>(size_t(ntv));
  tile_ptrs[3] = c.take<int32_t>(2);
  for (int k = 4; k < 7; ++k) tile_ptrs[k] = c.take<int32_t>(size_t(nta));
  tile_ptrs[7] = c.take<int32_t>(2);
  d.tv = Tiles{tile_ptrs[0], tile_ptrs[1], tile_ptrs[2], tile_ptrs[3], ntv};
  d.ta = Tiles{tile_ptrs[4], tile_ptrs[5], tile_ptrs[6], tile_ptrs[7], nta};
  d.env = c.take<EnvCG>(1);
  d.acc = c.take<A>(size_t(kCgSets) * kSlots * kSlotStride);
  d.num_active = c.take<int32_t>(1);
  d.solve = c.take<int32_t>(1);
  return c.off;
}

template <class T, class A>
sdx_status solve_noalloc(int32_t device, int32_t nbr, int32_t nnzb, const int32_t* ptr, const int32_t* col,
                         const double* values, int32_t layout, const double* b, double* x,
                         const sdx_krylov_params& kp, sdx_linear_status* out) {
  PcgSystem<T, A> d;
  int32_t* tp[8];
  const size_t need = layout_system<T, A>(nullptr, nbr, nnzb, d, tp);
  if (need > kScratchBytes - kStageBytes)
    return set_error(SDX_ERR_INVALID_ARGUMENT,
                     "system too large for the allocation-free solve: it needs " + std::to_string(need) +
                         " bytes of device scratch, the limit is " + std::to_string(kScratchBytes - kStageBytes) +
                         " (use sdx_pcg_bsr3_solve)");
  // rows are staged whole: no row may hold more values than the stage
  for (int i = 0; i < nbr; ++i)
    if (size_t(9) * size_t(ptr[i + 1] - ptr[i]) * sizeof(double) > kStageBytes)
      return set_error(SDX_ERR_INVALID_ARGUMENT, "a block row holds more values than the allocation-free stage");
  hip_check(hipSetDevice(device), "hipSetDevice");
  std::lock_guard<std::mutex> lock(g_dev_mutex[device]);
  unsigned char* base = scratch_address(device);
  unsigned char* stage = base;
  layout_system<T, A>(base + kStageBytes, nbr, nnzb, d, tp);
  hip_check(hipMemcpyAsync(const_cast<int32_t*>(d.ptr), ptr, (size_t(nbr) + 1) * sizeof(int32_t),
                           hipMemcpyHostToDevice, nullptr),
            "hipMemcpyAsync");
  hip_check(hipMemcpyAsync(const_cast<int32_t*>(d.col), col, size_t(nnzb) * sizeof(int32_t), hipMemcpyHostToDevice,
                           nullptr),
            "hipMemcpyAsync");
  hipLaunchKernelGGL(k_nt_tiles, dim3((d.tv.n + 255) / 256), dim3(256), 0, nullptr, nbr, int32_t(kVecThreads),
                     d.tv.n, tp[0], tp[1], tp[2], tp[3]);
  hipLaunchKernelGGL(k_nt_tiles, dim3((d.ta.n + 255) / 256), dim3(256), 0, nullptr, nbr, int32_t(kSpmvRows), d.ta.n,
                     tp[4], tp[5], tp[6], tp[7]);
  hipLaunchKernelGGL(k_nt_diag, dim3((nbr + 255) / 256), dim3(256), 0, nullptr, nbr, d.ptr, d.col,
                     const_cast<int32_t*>(d.diag));
  // the matrix, row range by row range through the stage (both layouts keep
  // a row's values contiguous, at 9 ptr[i])
  const size_t row_cap = kStageBytes / (9 * sizeof(double));  // blocks per staged range
  for (int i0 = 0; i0 < nbr;) {
    int i1 = i0 + 1;
    while (i1 < nbr && size_t(ptr[i1 + 1] - ptr[i0]) <= row_cap) ++i1;
    const int nblk = ptr[i1] - ptr[i0];
    if (nblk > 0) {
      hip_check(hipMemcpyAsync(stage, values + size_t(9) * size_t(ptr[i0]), size_t(9) * size_t(nblk) * sizeof(double),
                               hipMemcpyHostToDevice, nullptr),
                "hipMemcpyAsync");
      hipLaunchKernelGGL((k_nt_planes<double, T>), dim3((nblk + 255) / 256), dim3(256), 0, nullptr, i0, i1, d.ptr,
                         layout, reinterpret_cast<const double*>(stage), nnzb, const_cast<T*>(d.vals));
    }
    i0 = i1;
  }
  const size_t n3 = size_t(3) * size_t(nbr);
  upload_convert<double, T>(b, n3, const_cast<T*>(d.b), stage);
  upload_convert<double, A>(x, n3, d.x, stage);
  hip_check(hipGetLastError(), "kernel launch");
  pcg_run<T, A>(nullptr, d, kp, out);
  download_convert<A, double>(d.x, n3, x, stage);
  return SDX_OK;
}

}  // namespace
}  // namespace sdx

extern "C" sdx_status sdx_pcg_bsr3_solve_noalloc(int32_t device, int32_t n_block_rows, int32_t nnzb,
                                                 const int32_t* block_row_ptr, const int32_t* block_col_idx,
                                                 const double* values, int32_t layout, const double* b, double* x,
                                                 const sdx_krylov_params* params, sdx_linear_status* out) {
  using namespace sdx;
  const sdx_status st =
      pcg_check_args(device, n_block_rows, nnzb, block_row_ptr, block_col_idx, values, layout, b, x, params);
  if (st != SDX_OK) return st;
  if (device >= kMaxDevices) return set_error(SDX_ERR_INVALID_ARGUMENT, "device index beyond the supported range");
  return guarded([&]() -> sdx_status {
    switch (params->precision_mode) {
      case SDX_PREC_MODE_FP32:
        return solve_noalloc<float, float>(device, n_block_rows, nnzb, block_row_ptr, block_col_idx, values, layout,
                                           b, x, *params, out);
      case SDX_PREC_MODE_MIXED:
        return solve_noalloc<float, double>(device, n_block_rows, nnzb, block_row_ptr, block_col_idx, values, layout,
                                            b, x, *params, out);
      default:
        return solve_noalloc<double, double>(device, n_block_rows, nnzb, block_row_ptr, block_col_idx, values,
                                             layout, b, x, *params, out);
    }
  });
}
