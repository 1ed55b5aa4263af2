// scan.hip -- device-wide exclusive prefix sums (stream compaction, CSR
// offsets, counting sorts).  Three phases: per-block sums, scan of the block
// sums (recursive), per-block scan plus block offset.  4096 items per block.
#include "internal.hpp"

namespace phafd {
namespace {

constexpr int ST = 256;            // threads per block
constexpr int SQ = 16;             // items per thread
constexpr int SCHUNK = ST * SQ;    // 4096 items per block

template <class TI>
__global__ __launch_bounds__(ST) void k_reduce(const TI *__restrict__ in, size_t n,
                                               int64_t *__restrict__ bsum) {
  __shared__ int64_t sh[ST];
  const size_t base = (size_t)blockIdx.x * SCHUNK;
  int64_t acc = 0;
#pragma unroll
  for (int q = 0; q < SQ; q++) {
    const size_t i = base + (size_t)q * ST + threadIdx.x;
    if (i < n) acc += (int64_t)in[i];
  }
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int s = ST / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) bsum[blockIdx.x] = sh[0];
}

template <class TI, class TO>
__global__ __launch_bounds__(ST) void k_scan_apply(const TI *__restrict__ in, size_t n,
                                                   const int64_t *__restrict__ boff,
                                                   TO *__restrict__ out) {
  __shared__ int64_t sh[SCHUNK];
  __shared__ int64_t ts[ST];
  const size_t base = (size_t)blockIdx.x * SCHUNK;
#pragma unroll
  for (int q = 0; q < SQ; q++) {
    const size_t i = base + (size_t)q * ST + threadIdx.x;
    sh[q * ST + threadIdx.x] = i < n ? (int64_t)in[i] : 0;
  }
  __syncthreads();
  int64_t s = 0;
  for (int q = 0; q < SQ; q++) {
    const int64_t v = sh[threadIdx.x * SQ + q];
    sh[threadIdx.x * SQ + q] = s;
    s += v;
  }
  ts[threadIdx.x] = s;
  __syncthreads();
  // inclusive Hillis-Steele scan of the per-thread totals
  for (int d = 1; d < ST; d <<= 1) {
    int64_t v = (int)threadIdx.x >= d ? ts[threadIdx.x - d] : 0;
    __syncthreads();
    ts[threadIdx.x] += v;
    __syncthreads();
  }
  const int64_t excl = (threadIdx.x == 0 ? 0 : ts[threadIdx.x - 1]) + (boff ? boff[blockIdx.x] : 0);
  for (int q = 0; q < SQ; q++) sh[threadIdx.x * SQ + q] += excl;
  __syncthreads();
#pragma unroll
  for (int q = 0; q < SQ; q++) {
    const size_t i = base + (size_t)q * ST + threadIdx.x;
    if (i < n) out[i] = (TO)sh[q * ST + threadIdx.x];
  }
  if (blockIdx.x == gridDim.x - 1 && threadIdx.x == ST - 1)
    out[n] = (TO)(ts[ST - 1] + (boff ? boff[blockIdx.x] : 0));
}

// workspace per recursion level (grow-only; one context stream at a time)
DBuf<int64_t> g_ws_sum[4], g_ws_off[4];

template <class TI, class TO>
void scan_impl(const TI *in, TO *out, size_t n, hipStream_t s, int level) {
  if (n == 0) {
    // out[0] = 0
    hipLaunchKernelGGL((k_scan_apply<TI, TO>), dim3(1), dim3(ST), 0, s, in, (size_t)0,
                       (const int64_t *)nullptr, out);
    HIPCHK(hipGetLastError());
    return;
  }
  const size_t nb = (n + SCHUNK - 1) / SCHUNK;
  if (nb == 1) {
    hipLaunchKernelGGL((k_scan_apply<TI, TO>), dim3(1), dim3(ST), 0, s, in, n,
                       (const int64_t *)nullptr, out);
    HIPCHK(hipGetLastError());
    return;
  }
  if (level >= 4) throw std::runtime_error("scan too large");
  g_ws_sum[level].alloc(nb + 1);
  g_ws_off[level].alloc(nb + 1);
  hipLaunchKernelGGL((k_reduce<TI>), dim3((unsigned)nb), dim3(ST), 0, s, in, n, g_ws_sum[level].p);
  HIPCHK(hipGetLastError());
  scan_impl<int64_t, int64_t>(g_ws_sum[level].p, g_ws_off[level].p, nb, s, level + 1);
  hipLaunchKernelGGL((k_scan_apply<TI, TO>), dim3((unsigned)nb), dim3(ST), 0, s, in, n,
                     (const int64_t *)g_ws_off[level].p, out);
  HIPCHK(hipGetLastError());
}

}  // namespace

void scan_exclusive_i32_to_i64(const int *in, int64_t *out, size_t n, hipStream_t s) {
  scan_impl<int, int64_t>(in, out, n, s, 0);
}
void scan_exclusive_i32(const int *in, int *out, size_t n, hipStream_t s) {
  scan_impl<int, int>(in, out, n, s, 0);
}

}  // namespace phafd
