// Read-only statistics of a state: probability moments (Entropy, GoogleStats) and cost
// histograms (the QAOA GetHistogramFromCostFunction family).  Entry points declared in
// include/iqs_hip.h ("statistics").
//
// Moments: one pass over the state, eleven FP64 sums per workgroup, then a fixed-order sum
// over workgroups on the host (bitwise reproducible).  Histograms: one LDS histogram per
// workgroup (FP64 LDS atomics; the order of additions inside a workgroup is not fixed, so
// bins can differ in the last bits run to run), then a fixed-order sum over workgroups.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <mutex>
#include <vector>

#include "../../include/iqs_hip.h"
#include "iqs_internal.hpp"
#include "iqs_kernels.hpp"

using namespace iqsk;

namespace {

constexpr int kSB = 256;
constexpr unsigned kStatBlocks = 1024;
constexpr int kNStat = 11;         // -sum p ln p, -sum ln p (p > 0), sum p^k for k = 2..10
constexpr int kMaxBins = 4096;     // LDS histogram: 32 KiB of FP64 bins
constexpr unsigned kHistBlocks = 512;

struct StatWork {
  double* part = nullptr;
  size_t have = 0;
  int* flag = nullptr;
  std::vector<double> host;
};
StatWork g_sw;
std::mutex g_sm;

int need_part(size_t count)
{
  if (!g_sw.flag) {
    if (hipMalloc((void**)&g_sw.flag, sizeof(int)) != hipSuccess) return iqsi::fail(IQSH_ERR_ALLOC, "statistics flag");
  }
  if (g_sw.have >= count) return IQSH_OK;
  IQSI_TRY(hipStreamSynchronize(iqsi::stream()));
  if (g_sw.part) IQSI_TRY(hipFree(g_sw.part));
  g_sw.part = nullptr;
  g_sw.have = 0;
  if (hipMalloc((void**)&g_sw.part, count * sizeof(double)) != hipSuccess)
    return iqsi::fail(IQSH_ERR_ALLOC, "statistics buffer of %zu doubles", count);
  g_sw.have = count;
  return IQSH_OK;
}

template <typename T>
__global__ __launch_bounds__(kSB) void k_prob_moments(const typename Vec<T>::t* __restrict__ st, uint64_t N,
                                                      double* __restrict__ part)
{
  double v[kNStat];
#pragma unroll
  for (int r = 0; r < kNStat; ++r) v[r] = 0;
  for (uint64_t i = (uint64_t)blockIdx.x * kSB + threadIdx.x; i < N; i += (uint64_t)gridDim.x * kSB) {
    const auto a = st[i];
    const double p = (double)a.x * a.x + (double)a.y * a.y;
    if (p != 0.0) {
      const double l = log(p);
      v[0] -= p * l;
      v[1] -= l;
    }
    const double p2 = p * p, p3 = p2 * p, p4 = p2 * p2, p5 = p3 * p2;
    v[2] += p2;
    v[3] += p3;
    v[4] += p4;
    v[5] += p5;
    v[6] += p3 * p3;
    v[7] += p4 * p3;
    v[8] += p4 * p4;
    v[9] += p5 * p4;
    v[10] += p5 * p5;
  }
  __shared__ double sh[kNStat][kSB / 64];
#pragma unroll
  for (int r = 0; r < kNStat; ++r)
    for (int off = 32; off > 0; off >>= 1) v[r] += __shfl_xor(v[r], off);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0)
    for (int r = 0; r < kNStat; ++r) sh[r][wid] = v[r];
  __syncthreads();
  if (threadIdx.x < kNStat) {
    double s = 0;
    for (int w = 0; w < kSB / 64; ++w) s += sh[threadIdx.x][w];
    part[(uint64_t)blockIdx.x * kNStat + threadIdx.x] = s;
  }
}

// Bin of one cost value: MODE 0 (int)c, MODE 1 floor(c + 1e-7), MODE 2 floor(c / width + 1e-7).
template <int MODE>
__device__ __forceinline__ int bin_of(double c, double width)
{
  if (MODE == 0) return (int)c;
  if (MODE == 1) return (int)floor(c + 1e-7);
  return (int)floor(c / width + 1e-7);
}

template <typename T, int MODE>
__global__ __launch_bounds__(kSB) void k_cost_hist(const typename Vec<T>::t* __restrict__ psi,
                                                   const typename Vec<T>::t* __restrict__ diag, uint64_t N, int nbins,
                                                   double lo, double hi, double width, double* __restrict__ part,
                                                   int* __restrict__ bad)
{
  __shared__ double bins[kMaxBins];
  for (int b = threadIdx.x; b < nbins; b += kSB) bins[b] = 0;
  __syncthreads();
  int out = 0;
  for (uint64_t i = (uint64_t)blockIdx.x * kSB + threadIdx.x; i < N; i += (uint64_t)gridDim.x * kSB) {
    const double c = (double)diag[i].x;
    const int b = bin_of<MODE>(c, width);
    if (!(c >= lo && c <= hi) || b < 0 || b >= nbins) {
      out = 1;
      continue;
    }
    const auto a = psi[i];
    atomicAdd(&bins[b], (double)a.x * a.x + (double)a.y * a.y);
  }
  if (out) atomicOr(bad, 1);
  __syncthreads();
  for (int b = threadIdx.x; b < nbins; b += kSB) part[(uint64_t)blockIdx.x * nbins + b] = bins[b];
}

template <typename T>
int do_moments(iqsh_handle h, double* out)
{
  const unsigned g = (unsigned)std::min<uint64_t>(kStatBlocks, std::max<uint64_t>(1, h->size() / (kSB * 8)));
  if (int rc = need_part((size_t)g * kNStat)) return rc;
  k_prob_moments<T><<<g, kSB, 0, iqsi::stream()>>>((const typename Vec<T>::t*)h->d, h->size(), g_sw.part);
  IQSI_TRY(hipGetLastError());
  g_sw.host.resize((size_t)g * kNStat);
  IQSI_TRY(hipMemcpyAsync(g_sw.host.data(), g_sw.part, sizeof(double) * g * kNStat, hipMemcpyDeviceToHost,
                          iqsi::stream()));
  IQSI_TRY(hipStreamSynchronize(iqsi::stream()));
  for (int r = 0; r < kNStat; ++r) {
    double s = 0;
    for (unsigned b = 0; b < g; ++b) s += g_sw.host[(size_t)b * kNStat + r];
    out[r] = s;
  }
  return IQSH_OK;
}

template <typename T>
int do_hist(iqsh_handle psi, iqsh_handle diag, int mode, double max_value, double width, double* hist, int nbins)
{
  const unsigned g = (unsigned)std::min<uint64_t>(kHistBlocks, std::max<uint64_t>(1, psi->size() / (kSB * 8)));
  if (int rc = need_part((size_t)g * nbins)) return rc;
  IQSI_TRY(hipMemsetAsync(g_sw.flag, 0, sizeof(int), iqsi::stream()));
  using V = typename Vec<T>::t;
  const double tol = mode == 0 ? 0.0 : 1e-7;
  const double lo = -tol, hi = max_value + tol;
  hipStream_t s = iqsi::stream();
  if (mode == 0)
    k_cost_hist<T, 0><<<g, kSB, 0, s>>>((const V*)psi->d, (const V*)diag->d, psi->size(), nbins, lo, hi, width,
                                        g_sw.part, g_sw.flag);
  else if (mode == 1)
    k_cost_hist<T, 1><<<g, kSB, 0, s>>>((const V*)psi->d, (const V*)diag->d, psi->size(), nbins, lo, hi, width,
                                        g_sw.part, g_sw.flag);
  else
    k_cost_hist<T, 2><<<g, kSB, 0, s>>>((const V*)psi->d, (const V*)diag->d, psi->size(), nbins, lo, hi, width,
                                        g_sw.part, g_sw.flag);
  IQSI_TRY(hipGetLastError());
  int bad = 0;
  g_sw.host.resize((size_t)g * nbins);
  IQSI_TRY(hipMemcpyAsync(g_sw.host.data(), g_sw.part, sizeof(double) * g * nbins, hipMemcpyDeviceToHost, s));
  IQSI_TRY(hipMemcpyAsync(&bad, g_sw.flag, sizeof(int), hipMemcpyDeviceToHost, s));
  IQSI_TRY(hipStreamSynchronize(s));
  if (bad) return iqsi::fail(IQSH_ERR_ARG, "a cost value lies outside [0, max_value]");
  for (int b = 0; b < nbins; ++b) {
    double acc = 0;
    for (unsigned k = 0; k < g; ++k) acc += g_sw.host[(size_t)k * nbins + b];
    hist[b] = acc;
  }
  return IQSH_OK;
}

}  // namespace

extern "C" {

int iqsh_prob_moments(iqsh_handle h, double out[11])
{
  if (!h || !h->d) return iqsi::fail(IQSH_ERR_ARG, "null or released state handle");
  if (!out) return iqsi::fail(IQSH_ERR_ARG, "null output");
  std::lock_guard<std::mutex> lk(g_sm);
  return h->dtype == IQSH_C128 ? do_moments<double>(h, out) : do_moments<float>(h, out);
}

int iqsh_cost_histogram(iqsh_handle psi, iqsh_handle diag, int mode, double max_value, double bin_width,
                        double* hist, int nbins)
{
  if (!psi || !psi->d || !diag || !diag->d) return iqsi::fail(IQSH_ERR_ARG, "null or released state handle");
  if (psi->n != diag->n || psi->dtype != diag->dtype) return iqsi::fail(IQSH_ERR_ARG, "registers differ in size or precision");
  if (mode < 0 || mode > 2) return iqsi::fail(IQSH_ERR_ARG, "histogram mode must be 0, 1 or 2");
  if (!(max_value > 0)) return iqsi::fail(IQSH_ERR_ARG, "max_value must be positive");
  if (mode == 2 && !(bin_width > 0)) return iqsi::fail(IQSH_ERR_ARG, "bin_width must be positive");
  if (!hist || nbins < 1 || nbins > kMaxBins) return iqsi::fail(IQSH_ERR_ARG, "nbins must lie in [1, %d]", kMaxBins);
  std::lock_guard<std::mutex> lk(g_sm);
  return psi->dtype == IQSH_C128 ? do_hist<double>(psi, diag, mode, max_value, bin_width, hist, nbins)
                                 : do_hist<float>(psi, diag, mode, max_value, bin_width, hist, nbins);
}

}  // extern "C"
