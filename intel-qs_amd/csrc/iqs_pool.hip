// Pool of states: the per-state entry points of libiqs_hip.so (include/iqs_hip.h,
// "pool of states").  A pool is one register whose top positions index its states, so
// every ordinary gate kernel already applies a common gate to all states in one launch;
// the kernels here cover what differs from state to state -- initialization, one 2x2 or
// 4x4 matrix per state (noise gates, 1- and 2-qubit channel branches), and reductions with
// one result per state.
//
// Reductions: a workgroup reduces one chunk of C = min(2^state_qubits, 2^14) contiguous
// amplitudes (never straddling two states) in FP64 with a fixed lane/wave order; the
// per-chunk partials are summed per state on the host, in chunk order, so every result is
// bitwise reproducible.
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

constexpr int kPB = 256;              // threads per workgroup
constexpr unsigned kChunkLog = 14;    // at most 2^14 amplitudes per reduction workgroup
constexpr uint64_t kMaxGrid = 1u << 20;
constexpr int kRU = 8;                // loads in flight per thread in the chunk reductions

// Device buffers shared by the pool entry points (guarded; grown on demand).
struct PoolWork {
  void* mats = nullptr;
  size_t mats_bytes = 0;
  double* part = nullptr;
  size_t part_count = 0;
  std::vector<double> host;
};
PoolWork g_pw;
std::mutex g_pm;

int grow(void** p, size_t* have, size_t need)
{
  if (*have >= need) return IQSH_OK;
  // The buffer may still be read by earlier work on the stream.
  IQSI_TRY(hipStreamSynchronize(iqsi::stream()));
  if (*p) IQSI_TRY(hipFree(*p));
  *p = nullptr;
  *have = 0;
  if (hipMalloc(p, need) != hipSuccess) return iqsi::fail(IQSH_ERR_ALLOC, "pool work buffer of %zu bytes", need);
  *have = need;
  return IQSH_OK;
}

unsigned grid_of(uint64_t work) { return (unsigned)std::max<uint64_t>(1, std::min<uint64_t>((work + kPB - 1) / kPB, kMaxGrid)); }

template <typename V>
__device__ __forceinline__ double nrm2(const V& a)
{
  return (double)a.x * (double)a.x + (double)a.y * (double)a.y;
}

// Sum of NV values over the workgroup: lanes by xor-shuffle, then the four waves in order.
template <int NV>
__device__ __forceinline__ void wg_sum(double (&v)[NV], double* out)
{
  __shared__ double sh[NV][kPB / 64];
#pragma unroll
  for (int r = 0; r < NV; ++r)
    for (int off = 32; off > 0; off >>= 1) v[r] += __shfl_xor(v[r], off);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0)
    for (int r = 0; r < NV; ++r) sh[r][wid] = v[r];
  __syncthreads();
  if (threadIdx.x == 0)
    for (int r = 0; r < NV; ++r) {
      double s = 0;
      for (int w = 0; w < kPB / 64; ++w) s += sh[r][w];
      out[r] = s;
    }
}

template <typename T>
__global__ __launch_bounds__(kPB) void k_pool_init(typename Vec<T>::t* __restrict__ st, uint64_t N, unsigned sbits,
                                                   int style, uint64_t index, uint64_t nstates, T amp)
{
  using V = typename Vec<T>::t;
  const uint64_t smask = (uint64_t(1) << sbits) - 1;
  for (uint64_t i = (uint64_t)blockIdx.x * kPB + threadIdx.x; i < N; i += (uint64_t)gridDim.x * kPB) {
    V v;
    v.x = 0;
    v.y = 0;
    if ((i >> sbits) < nstates) {
      if (style == 0) {
        if ((i & smask) == index) v.x = 1;
      } else {
        v.x = amp;
      }
    }
    st[i] = v;
  }
}

// (psi_{b}, psi_{b | 2^pos}) <- M_s (psi_b, psi_{b | 2^pos}) with s = b >> sbits.
template <typename T>
__global__ __launch_bounds__(kPB) void k_pool_1q(typename Vec<T>::t* __restrict__ st, uint64_t items, unsigned pos,
                                                 unsigned sbits, const Mat2<T>* __restrict__ mats, uint64_t nstates)
{
  using V = typename Vec<T>::t;
  const uint64_t hb = uint64_t(1) << pos;
  for (uint64_t k = (uint64_t)blockIdx.x * kPB + threadIdx.x; k < items; k += (uint64_t)gridDim.x * kPB) {
    const uint64_t i0 = ins0(k, pos);
    const uint64_t s = i0 >> sbits;
    if (s >= nstates) continue;
    const Mat2<T> m = mats[s];
    const V a = st[i0], b = st[i0 | hb];
    st[i0] = cmad2(m.m00, a, m.m01, b);
    st[i0 | hb] = cmad2(m.m10, a, m.m11, b);
  }
}

// The same per-state updates in the form of the register's own gate kernels: exactly one
// tile of B * P items per workgroup, tiles XCD-contiguous, both items of a thread loaded
// before either is updated, non-temporal accesses when an item's amplitudes lie in
// different 128-byte lines (cached when they share lines, so the second access finds the
// line in L2).  Items of states at or beyond nstates are neither read nor written.
constexpr int kTB = 128, kTP = 2, kTP2 = 1;
__host__ __device__ constexpr unsigned pool_line_bits(int esz) { return esz == 16 ? 3u : 4u; }

__device__ __forceinline__ uint64_t pool_tile() {
  const uint64_t b = blockIdx.x + (uint64_t)blockIdx.y * gridDim.x, G = (uint64_t)gridDim.x * gridDim.y;
  return (G & 7) ? b : xcd_block(b, G);
}

template <typename T, bool NT, int TB = kTB, int TP = kTP>
__global__ __launch_bounds__(TB) void k_pool_1q_t(typename Vec<T>::t* __restrict__ st, unsigned pos, unsigned sbits,
                                                  const Mat2<T>* __restrict__ mats, uint64_t nstates)
{
  using V = typename Vec<T>::t;
  const uint64_t hb = uint64_t(1) << pos, k0 = pool_tile() * (TB * TP) + threadIdx.x;
  V a[TP], b[TP];
  uint64_t i0[TP];
  bool ok[TP];
#pragma unroll
  for (int j = 0; j < TP; ++j) {
    i0[j] = ins0(k0 + (uint64_t)j * TB, pos);
    ok[j] = (i0[j] >> sbits) < nstates;
    if (ok[j]) {
      a[j] = ldx<NT>(&st[i0[j]]);
      b[j] = ldx<NT>(&st[i0[j] | hb]);
    }
  }
#pragma unroll
  for (int j = 0; j < TP; ++j) {
    if (!ok[j]) continue;
    const Mat2<T> m = mats[i0[j] >> sbits];
    stx<NT>(cmad2(m.m00, a[j], m.m01, b[j]), &st[i0[j]]);
    stx<NT>(cmad2(m.m10, a[j], m.m11, b[j]), &st[i0[j] | hb]);
  }
}

// Per-state 2x2 on a position inside a 128-byte line, in whole runs: thread t of tile b
// holds amplitude i = b * kLB * kLP + j * kLB + t, consecutive lanes consecutive amplitudes,
// and receives its partner i ^ 2^pos (same wavefront, same state) by __shfl_xor.  States of
// at least kLB * kLP amplitudes (the host's condition), so a tile lies in one state and its
// matrix is one uniform load; a tile of a state at or beyond nstates is neither read nor
// written.
constexpr int kLB = 256, kLP = 2;
constexpr unsigned kLaneMinBits = 9;  // 2^9 = kLB * kLP
template <typename T>
__global__ __launch_bounds__(kLB) void k_pool_1q_lane(typename Vec<T>::t* __restrict__ st, unsigned pos,
                                                      unsigned sbits, const Mat2<T>* __restrict__ mats,
                                                      uint64_t nstates)
{
  using V = typename Vec<T>::t;
  const uint64_t t0 = pool_tile() * (kLB * kLP), s = t0 >> sbits;
  if (s >= nstates) return;  // the whole tile: uniform
  const Mat2<T> m = mats[s];
  const uint64_t i0 = t0 + threadIdx.x;
  const int ml = 1 << pos;
  const bool up = (threadIdx.x >> pos) & 1;  // this amplitude's bit pos
  V a[kLP];
#pragma unroll
  for (int j = 0; j < kLP; ++j) a[j] = ldg(&st[i0 + (uint64_t)j * kLB]);
#pragma unroll
  for (int j = 0; j < kLP; ++j) {
    V b;  // the partner amplitude
    b.x = __shfl_xor(a[j].x, ml);
    b.y = __shfl_xor(a[j].y, ml);
    stg(up ? cmad2(m.m10, b, m.m11, a[j]) : cmad2(m.m00, a[j], m.m01, b), &st[i0 + (uint64_t)j * kLB]);
  }
}

// Per-state 4x4 on (ph, pl), both inside a 128-byte line, in whole runs as k_pool_1q_lane:
// a lane holds the amplitude whose row is r = 2 b(ph) + b(pl) and receives rows r ^ 1,
// r ^ 2, r ^ 3 from lanes ^ 2^pl, ^ 2^ph, ^ both; its output is row r of the state's
// matrix, the coefficients m[r][r ^ k] selected per lane from the uniform matrix.
template <typename T>
__global__ __launch_bounds__(kLB) void k_pool_2q_lane(typename Vec<T>::t* __restrict__ st, unsigned ph, unsigned pl,
                                                      unsigned sbits, const Mat4<T>* __restrict__ mats,
                                                      uint64_t nstates)
{
  using V = typename Vec<T>::t;
  const uint64_t t0 = pool_tile() * (kLB * kLP), s = t0 >> sbits;
  if (s >= nstates) return;  // the whole tile: uniform
  // The state's matrix in LDS: per-lane coefficient choices from global memory compiled to
  // per-lane gathers (complex64 (1,0): 4.0 TB/s).
  __shared__ Cx<T> sm[16];
  if (threadIdx.x < 16) sm[threadIdx.x] = mats[s].m[threadIdx.x];
  __syncthreads();
  const int mh = 1 << ph, ml = 1 << pl;
  const int r = (((threadIdx.x >> ph) & 1) << 1) | ((threadIdx.x >> pl) & 1);
  Cx<T> c[4];  // c[k] = m[r][r ^ k]
#pragma unroll
  for (int k = 0; k < 4; ++k) c[k] = sm[4 * r + (r ^ k)];
  const uint64_t i0 = t0 + threadIdx.x;
  V a[kLP];
#pragma unroll
  for (int j = 0; j < kLP; ++j) a[j] = ldg(&st[i0 + (uint64_t)j * kLB]);
#pragma unroll
  for (int j = 0; j < kLP; ++j) {
    V xl, xh, xb;
    xl.x = __shfl_xor(a[j].x, ml);
    xl.y = __shfl_xor(a[j].y, ml);
    xh.x = __shfl_xor(a[j].x, mh);
    xh.y = __shfl_xor(a[j].y, mh);
    xb.x = __shfl_xor(a[j].x, ml | mh);
    xb.y = __shfl_xor(a[j].y, ml | mh);
    V r = cmad2(c[0], a[j], c[1], xl);
    const V u = cmad2(c[2], xh, c[3], xb);
    r.x += u.x;
    r.y += u.y;
    stg(r, &st[i0 + (uint64_t)j * kLB]);
  }
}

template <typename T, bool NT, int TB = kTB>
__global__ __launch_bounds__(TB) void k_pool_2q_t(typename Vec<T>::t* __restrict__ st, unsigned lo, unsigned hi,
                                                  uint64_t dh, uint64_t dl, unsigned sbits,
                                                  const Mat4<T>* __restrict__ mats, uint64_t nstates)
{
  using V = typename Vec<T>::t;
  const uint64_t k0 = pool_tile() * (TB * kTP2) + threadIdx.x;
  const uint64_t off[4] = {0, dl, dh, dh | dl};
  V x[kTP2][4];
  uint64_t i0[kTP2];
  bool ok[kTP2];
#pragma unroll
  for (int j = 0; j < kTP2; ++j) {
    i0[j] = ins0(ins0(k0 + (uint64_t)j * kTB, lo), hi);
    ok[j] = (i0[j] >> sbits) < nstates;
    if (ok[j]) {
#pragma unroll
      for (int c = 0; c < 4; ++c) x[j][c] = ldx<NT>(&st[i0[j] | off[c]]);
    }
  }
#pragma unroll
  for (int j = 0; j < kTP2; ++j) {
    if (!ok[j]) continue;
    const Mat4<T>& m = mats[i0[j] >> sbits];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      V r = cmad2(m.m[4 * t + 0], x[j][0], m.m[4 * t + 1], x[j][1]);
      const V u = cmad2(m.m[4 * t + 2], x[j][2], m.m[4 * t + 3], x[j][3]);
      r.x += u.x;
      r.y += u.y;
      stx<NT>(r, &st[i0[j] | off[t]]);
    }
  }
}

// 2D grid of `tiles` workgroups (at most 65536 in x).
dim3 tile_grid(uint64_t tiles)
{
  const uint64_t gx = tiles < 65536 ? tiles : 65536;
  return dim3((unsigned)gx, (unsigned)((tiles + gx - 1) / gx), 1);
}

// Dense 4x4 per state on (ph, pl), rows 2 b(ph) + b(pl): with lo < hi the two positions in
// ascending order, item k enumerates the indices with both bits clear; s = index >> sbits.
template <typename T>
__global__ __launch_bounds__(kPB) void k_pool_2q(typename Vec<T>::t* __restrict__ st, uint64_t items, unsigned lo,
                                                 unsigned hi, uint64_t dh, uint64_t dl, unsigned sbits,
                                                 const Mat4<T>* __restrict__ mats, uint64_t nstates)
{
  using V = typename Vec<T>::t;
  for (uint64_t k = (uint64_t)blockIdx.x * kPB + threadIdx.x; k < items; k += (uint64_t)gridDim.x * kPB) {
    const uint64_t i0 = ins0(ins0(k, lo), hi);
    const uint64_t s = i0 >> sbits;
    if (s >= nstates) continue;
    const Mat4<T>& m = mats[s];
    const uint64_t off[4] = {0, dl, dh, dh | dl};
    V x[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) x[c] = st[i0 | off[c]];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      V r = cmad2(m.m[4 * t + 0], x[0], m.m[4 * t + 1], x[1]);
      const V u = cmad2(m.m[4 * t + 2], x[2], m.m[4 * t + 3], x[3]);
      r.x += u.x;
      r.y += u.y;
      st[i0 | off[t]] = r;
    }
  }
}

// One partial per chunk of 2^clog amplitudes: KIND 0 sum |psi|^2, 1 over bit `arg` set,
// 2 with the sign (-1)^popcount(i & arg).
template <typename T, int KIND>
__global__ __launch_bounds__(kPB) void k_pool_red(const typename Vec<T>::t* __restrict__ st, unsigned clog,
                                                  uint64_t arg, double* __restrict__ part)
{
  const uint64_t base = (uint64_t)blockIdx.x << clog, C = uint64_t(1) << clog;
  double v[1] = {0};
  auto add = [&](uint64_t i, double p) {
    if (KIND == 0) v[0] += p;
    else if (KIND == 1) v[0] += ((i >> arg) & 1) ? p : 0.0;
    else v[0] += (__popcll(i & arg) & 1) ? -p : p;
  };
  if (C >= kRU * kPB) {
    // C is a power of two: whole batches of kRU loads per thread, in flight together
    for (uint64_t j = threadIdx.x; j < C; j += kRU * kPB) {
      typename Vec<T>::t a[kRU];
#pragma unroll
      for (int u = 0; u < kRU; ++u) a[u] = ldg(&st[base + j + u * kPB]);
#pragma unroll
      for (int u = 0; u < kRU; ++u) add(base + j + u * kPB, nrm2(a[u]));
    }
  } else {
    for (uint64_t j = threadIdx.x; j < C; j += kPB) add(base + j, nrm2(ldg(&st[base + j])));
  }
  wg_sum<1>(v, part + blockIdx.x);
}

// Per chunk: sum conj(ref_j) psi_i with j = i mod 2^sbits.
template <typename T>
__global__ __launch_bounds__(kPB) void k_pool_ovl(const typename Vec<T>::t* __restrict__ st,
                                                  const typename Vec<T>::t* __restrict__ ref, unsigned clog,
                                                  unsigned sbits, double* __restrict__ part)
{
  const uint64_t base = (uint64_t)blockIdx.x << clog, C = uint64_t(1) << clog;
  const uint64_t smask = (uint64_t(1) << sbits) - 1;
  double v[2] = {0, 0};
  using V = typename Vec<T>::t;
  auto add = [&](const V& a, const V& b) {
    v[0] += (double)b.x * a.x + (double)b.y * a.y;
    v[1] += (double)b.x * a.y - (double)b.y * a.x;
  };
  if (C >= kRU * kPB) {
    for (uint64_t j = threadIdx.x; j < C; j += kRU * kPB) {
      V a[kRU], b[kRU];
#pragma unroll
      for (int u = 0; u < kRU; ++u) {
        const uint64_t i = base + j + u * kPB;
        a[u] = ldg(&st[i]);
        b[u] = ref[i & smask];
      }
#pragma unroll
      for (int u = 0; u < kRU; ++u) add(a[u], b[u]);
    }
  } else {
    for (uint64_t j = threadIdx.x; j < C; j += kPB) add(ldg(&st[base + j]), ref[(base + j) & smask]);
  }
  double out[2];
  wg_sum<2>(v, out);
  if (threadIdx.x == 0) {
    part[2 * (uint64_t)blockIdx.x] = out[0];
    part[2 * (uint64_t)blockIdx.x + 1] = out[1];
  }
}

int check_pool(iqsh_handle h, unsigned sbits, uint64_t nstates)
{
  if (!h || !h->d) return iqsi::fail(IQSH_ERR_ARG, "null pool handle");
  if (sbits == 0 || sbits > h->n) return iqsi::fail(IQSH_ERR_ARG, "state_qubits %u outside [1, %u]", sbits, h->n);
  if (nstates == 0 || nstates > (uint64_t(1) << (h->n - sbits)))
    return iqsi::fail(IQSH_ERR_ARG, "nstates %llu outside [1, 2^%u]", (unsigned long long)nstates, h->n - sbits);
  return IQSH_OK;
}

template <typename T>
int do_init(iqsh_handle h, unsigned sbits, int style, uint64_t index, uint64_t nstates)
{
  const T amp = (T)(1.0 / std::sqrt((double)(uint64_t(1) << sbits)));
  k_pool_init<T><<<grid_of(h->size()), kPB, 0, iqsi::stream()>>>((typename Vec<T>::t*)h->d, h->size(), sbits, style,
                                                                   index, nstates, amp);
  IQSI_TRY(hipGetLastError());
  return IQSH_OK;
}

template <typename T>
int do_1q_each(iqsh_handle h, unsigned sbits, unsigned pos, const double* mats, uint64_t nstates)
{
  std::vector<Mat2<T>> m(nstates);
  for (uint64_t s = 0; s < nstates; ++s) {
    const double* p = mats + 8 * s;
    m[s] = Mat2<T>{{(T)p[0], (T)p[1]}, {(T)p[2], (T)p[3]}, {(T)p[4], (T)p[5]}, {(T)p[6], (T)p[7]}};
  }
  const size_t bytes = sizeof(Mat2<T>) * nstates;
  if (int rc = grow(&g_pw.mats, &g_pw.mats_bytes, bytes)) return rc;
  // Earlier launches may still read the matrix buffer: wait for them before refilling it.
  IQSI_TRY(hipStreamSynchronize(iqsi::stream()));
  IQSI_TRY(hipMemcpy(g_pw.mats, m.data(), bytes, hipMemcpyHostToDevice));
  const uint64_t items = h->size() >> 1;
  using V = typename Vec<T>::t;
  if (pos < pool_line_bits(sizeof(V)) && sbits >= kLaneMinBits) {
    k_pool_1q_lane<T><<<tile_grid(h->size() / (kLB * kLP)), kLB, 0, iqsi::stream()>>>(
        (V*)h->d, pos, sbits, (const Mat2<T>*)g_pw.mats, nstates);
  } else if (items % (kTB * kTP) == 0) {
    // items is a power of two: a whole number of tiles, 8 or more from 2^12 amplitudes up
    const dim3 g = tile_grid(items / (kTB * kTP));
    // complex64 from the line width up: 256 x 1 tiles (8 x 28 qubits: position 5 6.25 ->
    // 6.47 TB/s, 13 6.43 -> 6.54, top equal)
    if (pos >= pool_line_bits(sizeof(V)) && sizeof(T) == 4)
      k_pool_1q_t<T, true, 2 * kTB, 1><<<tile_grid(items / (2 * kTB)), 2 * kTB, 0, iqsi::stream()>>>(
          (V*)h->d, pos, sbits, (const Mat2<T>*)g_pw.mats, nstates);
    else if (pos >= pool_line_bits(sizeof(V)))
      k_pool_1q_t<T, true><<<g, kTB, 0, iqsi::stream()>>>((V*)h->d, pos, sbits, (const Mat2<T>*)g_pw.mats, nstates);
    else
      k_pool_1q_t<T, false><<<g, kTB, 0, iqsi::stream()>>>((V*)h->d, pos, sbits, (const Mat2<T>*)g_pw.mats, nstates);
  } else {
    k_pool_1q<T><<<grid_of(items), kPB, 0, iqsi::stream()>>>((V*)h->d, items, pos, sbits, (const Mat2<T>*)g_pw.mats,
                                                              nstates);
  }
  IQSI_TRY(hipGetLastError());
  return IQSH_OK;
}

template <typename T>
int do_2q_each(iqsh_handle h, unsigned sbits, unsigned ph, unsigned pl, const double* mats, uint64_t nstates)
{
  std::vector<Mat4<T>> m(nstates);
  for (uint64_t s = 0; s < nstates; ++s)
    for (int e = 0; e < 16; ++e) m[s].m[e] = Cx<T>{(T)mats[32 * s + 2 * e], (T)mats[32 * s + 2 * e + 1]};
  const size_t bytes = sizeof(Mat4<T>) * nstates;
  if (int rc = grow(&g_pw.mats, &g_pw.mats_bytes, bytes)) return rc;
  IQSI_TRY(hipStreamSynchronize(iqsi::stream()));
  IQSI_TRY(hipMemcpy(g_pw.mats, m.data(), bytes, hipMemcpyHostToDevice));
  const uint64_t items = h->size() >> 2;
  using V = typename Vec<T>::t;
  const unsigned lo = std::min(ph, pl), hi = std::max(ph, pl);
  if (hi < pool_line_bits(sizeof(V)) && sbits >= kLaneMinBits) {
    k_pool_2q_lane<T><<<tile_grid(h->size() / (kLB * kLP)), kLB, 0, iqsi::stream()>>>(
        (V*)h->d, ph, pl, sbits, (const Mat4<T>*)g_pw.mats, nstates);
  } else if (items % (kTB * kTP2) == 0) {
    const dim3 g = tile_grid(items / (kTB * kTP2));
    // complex64 from the line width up: 256-thread tiles (8 x 28 qubits: (27,26) 4.88 ->
    // 5.55 TB/s, (5,20) 5.40 -> 5.52; complex128 within 1% either way)
    if (lo >= pool_line_bits(sizeof(V)) && sizeof(T) == 4 && items % (2 * kTB * kTP2) == 0)
      k_pool_2q_t<T, true, 2 * kTB><<<tile_grid(items / (2 * kTB * kTP2)), 2 * kTB, 0, iqsi::stream()>>>(
          (V*)h->d, lo, hi, uint64_t(1) << ph, uint64_t(1) << pl, sbits, (const Mat4<T>*)g_pw.mats, nstates);
    else if (lo >= pool_line_bits(sizeof(V)))
      k_pool_2q_t<T, true><<<g, kTB, 0, iqsi::stream()>>>((V*)h->d, lo, hi, uint64_t(1) << ph, uint64_t(1) << pl, sbits,
                                                          (const Mat4<T>*)g_pw.mats, nstates);
    else
      k_pool_2q_t<T, false><<<g, kTB, 0, iqsi::stream()>>>((V*)h->d, lo, hi, uint64_t(1) << ph, uint64_t(1) << pl,
                                                           sbits, (const Mat4<T>*)g_pw.mats, nstates);
  } else {
    k_pool_2q<T><<<grid_of(items), kPB, 0, iqsi::stream()>>>((V*)h->d, items, lo, hi, uint64_t(1) << ph,
                                                              uint64_t(1) << pl, sbits, (const Mat4<T>*)g_pw.mats,
                                                              nstates);
  }
  IQSI_TRY(hipGetLastError());
  return IQSH_OK;
}

// Launch `nchunks` chunk reductions with NV values each, then sum them per state on the host.
template <typename Launch>
int reduce_per_state(unsigned sbits, uint64_t nstates, int nv, double* out, Launch launch)
{
  const unsigned clog = std::min(sbits, kChunkLog);
  const uint64_t per_state = uint64_t(1) << (sbits - clog);
  const uint64_t nchunks = nstates * per_state;
  if (nchunks > 0x7fffffffull) return iqsi::fail(IQSH_ERR_NOTSUP, "pool reduction needs %llu workgroups", (unsigned long long)nchunks);
  size_t have = g_pw.part_count * sizeof(double);
  void* p = g_pw.part;
  if (int rc = grow(&p, &have, nchunks * nv * sizeof(double))) return rc;
  g_pw.part = (double*)p;
  g_pw.part_count = have / sizeof(double);
  launch((unsigned)nchunks, clog, g_pw.part);
  IQSI_TRY(hipGetLastError());
  g_pw.host.resize(nchunks * nv);
  IQSI_TRY(hipMemcpyAsync(g_pw.host.data(), g_pw.part, nchunks * nv * sizeof(double), hipMemcpyDeviceToHost,
                          iqsi::stream()));
  IQSI_TRY(hipStreamSynchronize(iqsi::stream()));
  for (uint64_t s = 0; s < nstates; ++s)
    for (int r = 0; r < nv; ++r) {
      double acc = 0;
      for (uint64_t c = 0; c < per_state; ++c) acc += g_pw.host[(s * per_state + c) * nv + r];
      out[s * nv + r] = acc;
    }
  return IQSH_OK;
}

template <typename T>
int do_reduce(iqsh_handle h, unsigned sbits, int kind, uint64_t arg, double* out, uint64_t nstates)
{
  using V = typename Vec<T>::t;
  const V* st = (const V*)h->d;
  return reduce_per_state(sbits, nstates, 1, out, [&](unsigned g, unsigned clog, double* part) {
    hipStream_t s = iqsi::stream();
    if (kind == IQSH_POOL_NORM2) k_pool_red<T, 0><<<g, kPB, 0, s>>>(st, clog, arg, part);
    else if (kind == IQSH_POOL_PROB1) k_pool_red<T, 1><<<g, kPB, 0, s>>>(st, clog, arg, part);
    else k_pool_red<T, 2><<<g, kPB, 0, s>>>(st, clog, arg, part);
  });
}

template <typename T>
int do_overlap(iqsh_handle h, unsigned sbits, iqsh_handle ref, double* out, uint64_t nstates)
{
  using V = typename Vec<T>::t;
  return reduce_per_state(sbits, nstates, 2, out, [&](unsigned g, unsigned clog, double* part) {
    k_pool_ovl<T><<<g, kPB, 0, iqsi::stream()>>>((const V*)h->d, (const V*)ref->d, clog, sbits, part);
  });
}

}  // namespace

extern "C" {

int iqsh_pool_init(iqsh_handle h, unsigned state_qubits, int style, uint64_t index, uint64_t nstates)
{
  if (int rc = check_pool(h, state_qubits, nstates)) return rc;
  if (style != 0 && style != 1) return iqsi::fail(IQSH_ERR_ARG, "pool init style must be 0 (basis) or 1 (uniform)");
  if (style == 0 && index >> state_qubits) return iqsi::fail(IQSH_ERR_ARG, "basis index outside the state");
  return h->dtype == IQSH_C128 ? do_init<double>(h, state_qubits, style, index, nstates)
                               : do_init<float>(h, state_qubits, style, index, nstates);
}

int iqsh_pool_apply_1q_each(iqsh_handle h, unsigned state_qubits, unsigned pos, const double* mats, uint64_t nstates)
{
  if (int rc = check_pool(h, state_qubits, nstates)) return rc;
  if (pos >= state_qubits) return iqsi::fail(IQSH_ERR_ARG, "position %u outside the state (%u qubits)", pos, state_qubits);
  if (!mats) return iqsi::fail(IQSH_ERR_ARG, "null matrices");
  std::lock_guard<std::mutex> lk(g_pm);
  return h->dtype == IQSH_C128 ? do_1q_each<double>(h, state_qubits, pos, mats, nstates)
                               : do_1q_each<float>(h, state_qubits, pos, mats, nstates);
}

int iqsh_pool_apply_2q_each(iqsh_handle h, unsigned state_qubits, unsigned ph, unsigned pl, const double* mats,
                            uint64_t nstates)
{
  if (int rc = check_pool(h, state_qubits, nstates)) return rc;
  if (ph >= state_qubits || pl >= state_qubits)
    return iqsi::fail(IQSH_ERR_ARG, "positions %u, %u outside the state (%u qubits)", ph, pl, state_qubits);
  if (ph == pl) return iqsi::fail(IQSH_ERR_ARG, "2q positions coincide");
  if (!mats) return iqsi::fail(IQSH_ERR_ARG, "null matrices");
  std::lock_guard<std::mutex> lk(g_pm);
  return h->dtype == IQSH_C128 ? do_2q_each<double>(h, state_qubits, ph, pl, mats, nstates)
                               : do_2q_each<float>(h, state_qubits, ph, pl, mats, nstates);
}

int iqsh_pool_reduce(iqsh_handle h, unsigned state_qubits, int kind, uint64_t arg, double* out, uint64_t nstates)
{
  if (int rc = check_pool(h, state_qubits, nstates)) return rc;
  if (!out) return iqsi::fail(IQSH_ERR_ARG, "null output");
  if (kind == IQSH_POOL_PROB1 ? arg >= state_qubits : (kind == IQSH_POOL_ZSTRING ? (arg >> state_qubits) != 0 : kind != IQSH_POOL_NORM2))
    return iqsi::fail(IQSH_ERR_ARG, "bad reduction kind %d or argument", kind);
  std::lock_guard<std::mutex> lk(g_pm);
  return h->dtype == IQSH_C128 ? do_reduce<double>(h, state_qubits, kind, arg, out, nstates)
                               : do_reduce<float>(h, state_qubits, kind, arg, out, nstates);
}

int iqsh_pool_overlap(iqsh_handle h, unsigned state_qubits, iqsh_handle ref, double* out, uint64_t nstates)
{
  if (int rc = check_pool(h, state_qubits, nstates)) return rc;
  if (!ref || !ref->d || ref->n != state_qubits || ref->dtype != h->dtype)
    return iqsi::fail(IQSH_ERR_ARG, "reference register must hold %u qubits of the pool's precision", state_qubits);
  if (!out) return iqsi::fail(IQSH_ERR_ARG, "null output");
  std::lock_guard<std::mutex> lk(g_pm);
  return h->dtype == IQSH_C128 ? do_overlap<double>(h, state_qubits, ref, out, nstates)
                               : do_overlap<float>(h, state_qubits, ref, out, nstates);
}

}  // extern "C"
