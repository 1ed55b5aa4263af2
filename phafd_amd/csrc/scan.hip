// scan.hip -- device-wide exclusive prefix sums (stream compaction, CSR
// offsets, counting sorts).  Three phases: per-block sums, scan of the block
// sums (recursive), per-block scan plus block offset.  4096 items per block,
// 16 consecutive items per thread.
#include <map>
#include <memory>
#include <mutex>

#include "internal.hpp"

namespace phafd {
namespace {

constexpr int ST = 256;            // threads per block
constexpr int SQ = 16;             // items per thread
constexpr int SCHUNK = ST * SQ;    // 4096 items per block

// Each item owns SQ consecutive entries (vector loads where the run is whole),
// keeps its run in registers, and the block combines item totals with wave
// shuffles plus one LDS exchange between the four waves.
template <class TI>
__device__ __forceinline__ void load_run(const TI *__restrict__ in, size_t base, size_t n, TI *v) {
  if (base + SQ <= n && (sizeof(TI) * SQ) % 16 == 0 && ((uintptr_t)in & 15) == 0) {
    const int4 *p = reinterpret_cast<const int4 *>(in + base);
    constexpr int NV = (int)(sizeof(TI) * SQ / 16);
#pragma unroll
    for (int q = 0; q < NV; q++) reinterpret_cast<int4 *>(v)[q] = p[q];
  } else {
#pragma unroll
    for (int q = 0; q < SQ; q++) v[q] = base + q < n ? in[base + q] : (TI)0;
  }
}

__device__ __forceinline__ int64_t wave_incl_scan(int64_t x) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const int64_t y = __shfl_up(x, d, 64);
    if (lane >= d) x += y;
  }
  return x;
}

template <class TI>
__global__ __launch_bounds__(ST) void k_reduce(const TI *__restrict__ in, size_t n,
                                               int64_t *__restrict__ bsum) {
  __shared__ int64_t ws[ST / 64];
  const size_t base = (size_t)blockIdx.x * SCHUNK + (size_t)threadIdx.x * SQ;
  TI v[SQ];
  load_run(in, base, n, v);
  int64_t acc = 0;
#pragma unroll
  for (int q = 0; q < SQ; q++) acc += (int64_t)v[q];
  acc = wave_incl_scan(acc);
  if ((threadIdx.x & 63) == 63) ws[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    int64_t t = 0;
#pragma unroll
    for (int w = 0; w < ST / 64; w++) t += ws[w];
    bsum[blockIdx.x] = t;
  }
}

template <class TI, class TO>
__global__ __launch_bounds__(ST) void k_scan_apply(const TI *__restrict__ in, size_t n,
                                                   const int64_t *__restrict__ boff,
                                                   TO *__restrict__ out) {
  __shared__ int64_t ws[ST / 64];
  const size_t base = (size_t)blockIdx.x * SCHUNK + (size_t)threadIdx.x * SQ;
  TI v[SQ];
  load_run(in, base, n, v);
  int64_t run = 0;
#pragma unroll
  for (int q = 0; q < SQ; q++) run += (int64_t)v[q];
  const int64_t incl = wave_incl_scan(run);
  const int wid = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 63) ws[wid] = incl;
  __syncthreads();
  int64_t woff = 0, total = 0;
#pragma unroll
  for (int w = 0; w < ST / 64; w++) {
    woff += w < wid ? ws[w] : 0;
    total += ws[w];
  }
  const int64_t b0 = boff ? boff[blockIdx.x] : 0;
  int64_t e = b0 + woff + incl - run;
  if (base + SQ <= n && (sizeof(TO) * SQ) % 16 == 0 && ((uintptr_t)out & 15) == 0) {
    TO o[SQ];
#pragma unroll
    for (int q = 0; q < SQ; q++) {
      o[q] = (TO)e;
      e += (int64_t)v[q];
    }
    int4 *p = reinterpret_cast<int4 *>(out + base);
    constexpr int NV = (int)(sizeof(TO) * SQ / 16);
#pragma unroll
    for (int q = 0; q < NV; q++) p[q] = reinterpret_cast<const int4 *>(o)[q];
  } else {
#pragma unroll
    for (int q = 0; q < SQ; q++) {
      if (base + q < n) out[base + q] = (TO)e;
      e += (int64_t)v[q];
    }
  }
  if (blockIdx.x == gridDim.x - 1 && threadIdx.x == ST - 1) out[n] = (TO)(b0 + total);
}

// Block sums and offsets per recursion level, one workspace per stream:
// scans on one stream run in order and may share it, scans on different
// streams (the members of a batch) must not.  Grow-only, after the stream
// has drained, so no launch still queued reads a freed workspace.
struct ScanWs {
  DBuf<int64_t> sum[4], off[4];
  DBuf<double> red;   // block partials of the fixed-order FP64 reductions
};
std::mutex g_ws_mu;
std::map<hipStream_t, std::unique_ptr<ScanWs>> g_ws;

ScanWs &workspace(hipStream_t s) {
  std::lock_guard<std::mutex> lock(g_ws_mu);
  std::unique_ptr<ScanWs> &w = g_ws[s];
  if (!w) w = std::make_unique<ScanWs>();
  return *w;
}

template <class TI, class TO>
void scan_impl(const TI *in, TO *out, size_t n, hipStream_t s, int level, ScanWs &ws) {
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
  if (ws.sum[level].n < nb + 1 || ws.off[level].n < nb + 1) HIPCHK(hipStreamSynchronize(s));   // before a regrow
  ws.sum[level].alloc(nb + 1);
  ws.off[level].alloc(nb + 1);
  hipLaunchKernelGGL((k_reduce<TI>), dim3((unsigned)nb), dim3(ST), 0, s, in, n, ws.sum[level].p);
  HIPCHK(hipGetLastError());
  scan_impl<int64_t, int64_t>(ws.sum[level].p, ws.off[level].p, nb, s, level + 1, ws);
  hipLaunchKernelGGL((k_scan_apply<TI, TO>), dim3((unsigned)nb), dim3(ST), 0, s, in, n,
                     (const int64_t *)ws.off[level].p, out);
  HIPCHK(hipGetLastError());
}

}  // namespace

void scan_exclusive_i32_to_i64(const int *in, int64_t *out, size_t n, hipStream_t s) {
  scan_impl<int, int64_t>(in, out, n, s, 0, workspace(s));
}
double *reduce_partials(hipStream_t s, size_t n) {
  ScanWs &w = workspace(s);
  if (w.red.n < n) {
    HIPCHK(hipStreamSynchronize(s));   // before a regrow
    w.red.alloc(n);
  }
  return w.red.p;
}

void scan_forget_stream(hipStream_t s) {
  std::unique_ptr<ScanWs> w;
  {
    std::lock_guard<std::mutex> lock(g_ws_mu);
    auto it = g_ws.find(s);
    if (it == g_ws.end()) return;
    w = std::move(it->second);
    g_ws.erase(it);
  }
}
void scan_exclusive_i32(const int *in, int *out, size_t n, hipStream_t s) {
  scan_impl<int, int>(in, out, n, s, 0, workspace(s));
}

}  // namespace phafd
