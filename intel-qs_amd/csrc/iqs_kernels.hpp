// gfx950 kernels of the state-vector engine.  All kernels are HBM-streaming: every
// amplitude a kernel touches is read once and written at most once per launch.
//
// Index scheme.  A gate on positions {p_0 < ... < p_{K-1}} touches 2^K amplitudes per
// "item"; items are enumerated by k in [0, 2^(n-K)) and the item's base index is k with a
// zero bit inserted at each p_s (ins0, ascending).  Consecutive threads take consecutive
// k, so every load/store instruction of a wavefront covers one contiguous run of the
// state whenever the lowest touched position is >= 6 (DP) / 7 (SP); for lower positions
// the two partner loads of one thread land in the same cache lines and are merged in L2.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace iqsk {

constexpr int kBlock = 256;  // 4 wavefronts of 64

// Amplitudes are clang ext-vector pairs (re, im): one 16-byte (DP) or 8-byte (SP) access.
typedef double dcx __attribute__((ext_vector_type(2)));
typedef float fcx __attribute__((ext_vector_type(2)));
template <typename T> struct Vec;
template <> struct Vec<double> { using t = dcx; };
template <> struct Vec<float> { using t = fcx; };

// Streaming accesses.  IQS_NT=1 (default) marks them non-temporal (the state is touched
// once per launch); IQS_NT=0 uses ordinary cached accesses.
#ifndef IQS_NT
#define IQS_NT 1
#endif
template <typename V>
__device__ __forceinline__ V ldg(const V* p)
{
#if IQS_NT
  return __builtin_nontemporal_load(p);
#else
  return *p;
#endif
}
template <typename V>
__device__ __forceinline__ void stg(const V& v, V* p)
{
#if IQS_NT
  __builtin_nontemporal_store(v, p);
#else
  *p = v;
#endif
}

// Access with a chosen cache policy: NT = true as ldg/stg; NT = false ordinary cached
// accesses, for patterns where several instructions share each cache line (a line fetched
// or written partially by one instruction is completed by the next through L2).
template <bool NT, typename V>
__device__ __forceinline__ V ldx(const V* p)
{
  if (NT) return ldg(p);
  return *p;
}
template <bool NT, typename V>
__device__ __forceinline__ void stx(const V& v, V* p)
{
  if (NT) stg(v, p);
  else *p = v;
}

template <typename T> struct Cx { T r, i; };

template <int K> struct Ins { unsigned p[K > 0 ? K : 1]; };

template <typename T> struct Mat2 { Cx<T> m00, m01, m10, m11; };
template <typename T> struct Mat4 { Cx<T> m[16]; };

__device__ __forceinline__ uint64_t ins0(uint64_t x, unsigned p)
{
  const uint64_t lo = x & ((uint64_t(1) << p) - 1);
  return ((x ^ lo) << 1) | lo;
}

template <int K>
__device__ __forceinline__ uint64_t item_base(uint64_t k, const Ins<K>& ins)
{
#pragma unroll
  for (int s = 0; s < K; ++s) k = ins0(k, ins.p[s]);
  return k;
}

// ---------------------------------------------------------------------------------------
// Work tiling of the streaming kernels (measured on MI355X, 30-qubit complex128 state).
//  * XCD-contiguous blocks: workgroups are dispatched round-robin over the 8 XCDs, so
//    logical block (b % 8) * (G / 8) + b / 8 gives every XCD one contiguous slice of the
//    item space instead of a 1-in-8 interleave.
//  * Split blocks: a block of B*P items takes SUB sub-ranges of B*P/SUB items spaced
//    2^kSubLog items apart.  Keeping a block's partner streams a few tens of KiB apart is
//    what lets the pair update run at 6.3-6.8 TB/s at every target position (a single
//    contiguous range per block stays near 5.6-5.9 TB/s for far-apart partners); the
//    fastest shape for complex128 is 128 threads x 2 items in 4 runs of 64 items.
// Both need a power-of-two item count >= kTiledMin; smaller states use the plain mapping.
constexpr unsigned kSubLog = 11;
constexpr int kSub = 4;
constexpr uint64_t kTiledMin = uint64_t(1) << 14;

__device__ __forceinline__ uint64_t xcd_block(uint64_t b, uint64_t G) { return (b & 7) * (G >> 3) + (b >> 3); }

// Kernels that launch one workgroup per tile use a 2D grid (a dispatch holds fewer than
// 2^32 work-items per dimension); the linear block id runs x fastest, so the dispatch
// order -- and the round-robin over XCDs -- is that of a 1D grid.
__device__ __forceinline__ uint64_t block_id() { return blockIdx.x + (uint64_t)blockIdx.y * gridDim.x; }
__device__ __forceinline__ uint64_t grid_blocks() { return (uint64_t)gridDim.x * gridDim.y; }

// Item index of (block bb, slot t in [0, B*P)) under the split-block tiling.
template <int B, int P, int SUB, unsigned SL = kSubLog>
__device__ __forceinline__ uint64_t tiled_item(uint64_t bb, unsigned t)
{
  if (SUB == 1) return bb * (uint64_t)(B * P) + t;
  constexpr uint64_t sublen = (uint64_t)B * P / SUB;
  constexpr uint64_t span = uint64_t(1) << SL;
  constexpr uint64_t bps = span / sublen;
  static_assert(span % sublen == 0, "sub-range must divide the span");
  const uint64_t win = bb / bps, off = bb % bps;
  return win * span * SUB + off * sublen + (t / sublen) * span + (t % sublen);
}

// a*x + b*y, written in the reference's operand order (m00*in0 + m01*in1).
template <typename T, typename V>
__device__ __forceinline__ V cmad2(const Cx<T>& a, const V& x, const Cx<T>& b, const V& y)
{
  V r;
  r.x = (a.r * x.x - a.i * x.y) + (b.r * y.x - b.i * y.y);
  r.y = (a.r * x.y + a.i * x.x) + (b.r * y.y + b.i * y.x);
  return r;
}

template <typename T, typename V>
__device__ __forceinline__ V cmul(const Cx<T>& a, const V& x)
{
  V r;
  r.x = a.r * x.x - a.i * x.y;
  r.y = a.r * x.y + a.i * x.x;
  return r;
}

// ---------------------------------------------------------------------------------------
// Pair kernel: (st[b|mask0], st[b|mask1]) <- M (st[b|mask0], st[b|mask1]) for every item.
//   1q gate:          K=1 ins={t},        mask0=0,       mask1=1<<t
//   controlled gate:  K=2 ins={c,t} sort, mask0=1<<c,    mask1=1<<c|1<<t
//   multi-controlled: K=1+#c,             mask0=ctrl,    mask1=ctrl|1<<t
//   swap-like:        K=2 ins={p1,p2},    mask0=1<<p1,   mask1=1<<p2
// MOVE=true exchanges the two amplitudes (X-type matrices: a pure, bit-exact move).
// TILED=true: XCD-contiguous, split-block tiling (nitems a power of two >= kTiledMin, grid
// exactly nitems / (B*P)); TILED=false: plain contiguous blocks with bounds checks.
// NT = false: cached accesses, for a lowest position inside a cache line (the two halves of
// every line are read and written by different instructions).
template <typename T, int K, bool MOVE, bool TILED, int B, int P, int SUB = kSub, unsigned SL = kSubLog, bool NT = true>
__global__ __launch_bounds__(B) void k_pair(typename Vec<T>::t* __restrict__ st, uint64_t nitems, Ins<K> ins,
                                            uint64_t mask0, uint64_t mask1, Mat2<T> m)
{
  using V = typename Vec<T>::t;
  V a[P], b[P];
  uint64_t base[P];
  if (TILED) {
    const uint64_t bb = xcd_block(block_id(), grid_blocks());
#pragma unroll
    for (int j = 0; j < P; ++j) {
      base[j] = item_base(tiled_item<B, P, SUB, SL>(bb, j * B + threadIdx.x), ins);
      a[j] = ldx<NT>(&st[base[j] | mask0]);
      b[j] = ldx<NT>(&st[base[j] | mask1]);
    }
#pragma unroll
    for (int j = 0; j < P; ++j) {
      if (MOVE) {
        stx<NT>(b[j], &st[base[j] | mask0]);
        stx<NT>(a[j], &st[base[j] | mask1]);
      } else {
        stx<NT>(cmad2(m.m00, a[j], m.m01, b[j]), &st[base[j] | mask0]);
        stx<NT>(cmad2(m.m10, a[j], m.m11, b[j]), &st[base[j] | mask1]);
      }
    }
  } else {
    for (int j = 0; j < P; ++j) {
      const uint64_t k = block_id() * (B * P) + (uint64_t)j * B + threadIdx.x;
      if (k >= nitems) break;
      const uint64_t bb = item_base(k, ins);
      const V x = st[bb | mask0], y = st[bb | mask1];
      if (MOVE) {
        st[bb | mask0] = y;
        st[bb | mask1] = x;
      } else {
        st[bb | mask0] = cmad2(m.m00, x, m.m01, y);
        st[bb | mask1] = cmad2(m.m10, x, m.m11, y);
      }
    }
  }
}

// 2x2 update for a target below the wavefront width (pos < 6): lanes hold consecutive
// amplitudes (fully coalesced 1 KiB accesses per wavefront instruction) and each lane
// fetches its partner from lane ^ 2^pos.
//   KC > 0: control positions (all >= 6) are inserted as ones, so only the controlled
//           sub-state (N >> KC amplitudes) is read and written;
//   COND:   full pass over the state, the update applied where (i & ctrl) == ctrl (for
//           controls below 6, whose skipped amplitudes share cache lines with updated ones).
// MOVE exchanges partners exactly (X-type matrices).  Requires (N >> KC) a multiple of
// B*P and >= kTiledMin (grid a multiple of 8).
template <typename T, int KC, bool COND, bool MOVE, int B, int P>
__global__ __launch_bounds__(B) void k_pair_lane(typename Vec<T>::t* __restrict__ st, unsigned pos, Ins<KC> cins,
                                                 uint64_t ctrl, Mat2<T> m)
{
  using V = typename Vec<T>::t;
  const uint64_t r0 = xcd_block(block_id(), grid_blocks()) * (uint64_t)(B * P) + threadIdx.x;
  V v[P];
  uint64_t idx[P];
#pragma unroll
  for (int j = 0; j < P; ++j) {
    const uint64_t r = r0 + (uint64_t)j * B;
    idx[j] = KC > 0 ? (item_base(r, cins) | ctrl) : r;
    v[j] = ldg(&st[idx[j]]);
  }
  const int msk = 1 << pos;
  const bool upper = (threadIdx.x >> pos) & 1;
#pragma unroll
  for (int j = 0; j < P; ++j) {
    V p;
    p.x = __shfl_xor(v[j].x, msk);
    p.y = __shfl_xor(v[j].y, msk);
    V o;
    if (MOVE) o = p;
    else o = upper ? cmad2(m.m10, p, m.m11, v[j]) : cmad2(m.m00, v[j], m.m01, p);
    if (COND && (idx[j] & ctrl) != ctrl) o = v[j];
    stg(o, &st[idx[j]]);
  }
}

// Conditional pair update over the target only (controls below 3 share 128-byte lines
// with uncontrolled amplitudes): every pair is read and written back, the update applied
// where (i0 & ctrl) == ctrl.  Tiled mapping as k_pair.
template <typename T, bool MOVE, int B, int P>
__global__ __launch_bounds__(B) void k_pair_cond(typename Vec<T>::t* __restrict__ st, Ins<1> ins, uint64_t ctrl,
                                                 Mat2<T> m)
{
  using V = typename Vec<T>::t;
  const uint64_t bb = xcd_block(block_id(), grid_blocks());
  const uint64_t hb = uint64_t(1) << ins.p[0];
  V a[P], b[P];
  uint64_t base[P];
#pragma unroll
  for (int j = 0; j < P; ++j) {
    base[j] = item_base(tiled_item<B, P, kSub>(bb, j * B + threadIdx.x), ins);
    a[j] = ldg(&st[base[j]]);
    b[j] = ldg(&st[base[j] | hb]);
  }
#pragma unroll
  for (int j = 0; j < P; ++j) {
    V x = a[j], y = b[j];
    if ((base[j] & ctrl) == ctrl) {
      if (MOVE) {
        x = b[j];
        y = a[j];
      } else {
        x = cmad2(m.m00, a[j], m.m01, b[j]);
        y = cmad2(m.m10, a[j], m.m11, b[j]);
      }
    }
    stg(x, &st[base[j]]);
    stg(y, &st[base[j] | hb]);
  }
}

// ---------------------------------------------------------------------------------------
// Scale kernel: st[b|mask] *= f for every item (K=0: the whole state).  Serves diagonal
// gates whose other entries are exactly 1 (Z, S, T, CZ, CPhase, ...) by touching only the
// amplitudes that change.
template <typename T, int K, bool TILED, int B, int P>
__global__ __launch_bounds__(B) void k_scale(typename Vec<T>::t* __restrict__ st, uint64_t nitems, Ins<K> ins,
                                             uint64_t mask, Cx<T> f)
{
  using V = typename Vec<T>::t;
  if (TILED) {
    const uint64_t bb = xcd_block(block_id(), grid_blocks());
    V a[P];
    uint64_t idx[P];
#pragma unroll
    for (int j = 0; j < P; ++j) {
      idx[j] = item_base(tiled_item<B, P, kSub>(bb, j * B + threadIdx.x), ins) | mask;
      a[j] = ldg(&st[idx[j]]);
    }
#pragma unroll
    for (int j = 0; j < P; ++j) stg(cmul(f, a[j]), &st[idx[j]]);
  } else {
    for (int j = 0; j < P; ++j) {
      const uint64_t k = block_id() * (B * P) + (uint64_t)j * B + threadIdx.x;
      if (k >= nitems) break;
      const uint64_t i = item_base(k, ins) | mask;
      st[i] = cmul(f, st[i]);
    }
  }
}

// Elementwise 2-qubit diagonal: st[i] *= d[2*b(p1) + b(p2)].
template <typename T, bool TILED, int B, int P>
__global__ __launch_bounds__(B) void k_diag2(typename Vec<T>::t* __restrict__ st, uint64_t N, unsigned p1, unsigned p2,
                                             Cx<T> d0, Cx<T> d1, Cx<T> d2, Cx<T> d3)
{
  using V = typename Vec<T>::t;
  const uint64_t bb = TILED ? xcd_block(block_id(), grid_blocks()) : block_id();
  V a[P];
  uint64_t idx[P];
#pragma unroll
  for (int j = 0; j < P; ++j) {
    idx[j] = TILED ? tiled_item<B, P, kSub>(bb, j * B + threadIdx.x) : bb * (uint64_t)(B * P) + (uint64_t)j * B + threadIdx.x;
    if (TILED || idx[j] < N) a[j] = ldg(&st[idx[j]]);
  }
#pragma unroll
  for (int j = 0; j < P; ++j) {
    const uint64_t i = idx[j];
    if (TILED || i < N) {
      const bool b1 = (i >> p1) & 1, b2 = (i >> p2) & 1;
      const Cx<T> d = b1 ? (b2 ? d3 : d2) : (b2 ? d1 : d0);
      stg(cmul(d, a[j]), &st[i]);
    }
  }
}

// Dense 4x4 on (ph, pl): indices b, b|dl, b|dh, b|dh|dl; new[t] = sum_c m[t][c] old[c].
template <typename T, bool TILED, int B, int P, int SUB = kSub, unsigned SL = kSubLog>
__global__ __launch_bounds__(B) void k_quad(typename Vec<T>::t* __restrict__ st, uint64_t nitems, Ins<2> ins,
                                            uint64_t dl, uint64_t dh, Mat4<T> m)
{
  using V = typename Vec<T>::t;
  const uint64_t bb = TILED ? xcd_block(block_id(), grid_blocks()) : block_id();
  V x[P][4];
  uint64_t base[P];
  bool ok[P];
#pragma unroll
  for (int j = 0; j < P; ++j) {
    const uint64_t k = TILED ? tiled_item<B, P, SUB, SL>(bb, j * B + threadIdx.x) : bb * (uint64_t)(B * P) + (uint64_t)j * B + threadIdx.x;
    ok[j] = TILED || k < nitems;
    base[j] = item_base(k, ins);
    if (ok[j]) {
      x[j][0] = ldg(&st[base[j]]);
      x[j][1] = ldg(&st[base[j] | dl]);
      x[j][2] = ldg(&st[base[j] | dh]);
      x[j][3] = ldg(&st[base[j] | dh | dl]);
    }
  }
#pragma unroll
  for (int j = 0; j < P; ++j) {
    if (!ok[j]) continue;
    const uint64_t off[4] = {0, dl, dh, dh | dl};
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      V r = cmad2(m.m[4 * t + 0], x[j][0], m.m[4 * t + 1], x[j][1]);
      const V s = cmad2(m.m[4 * t + 2], x[j][2], m.m[4 * t + 3], x[j][3]);
      r.x += s.x;
      r.y += s.y;
      stg(r, &st[base[j] | off[t]]);
    }
  }
}

// Dense 4x4 (or swap-like move) on positions pa (row bit 1) and pb (row bit 0) when the
// lower one is below the wavefront width: every load and store is a contiguous run of the
// state and partners inside a wavefront come through __shfl_xor, instead of 16-byte
// accesses spread over the cache lines.  lo = min(pa, pb) < 6.
//  * hi = max(pa, pb) >= 6: items enumerate indices with bit hi clear; a thread holds the
//    amplitudes with bit hi = 0 and 1 and exchanges both with lane ^ (1 << lo).
//  * hi < 6: a thread holds one amplitude and receives the other three by shuffles.
// MOVE: exact exchange of rows 1 and 2 (SWAP-type), no arithmetic.  Requires the item count
// (N/2 or N) to be a multiple of B*P and >= kTiledMin.
template <typename T, typename V>
__device__ __forceinline__ V quad_row(const Mat4<T>& m, int rw, const V& i0, const V& i1, const V& i2, const V& i3)
{
  V r = cmad2(m.m[4 * rw + 0], i0, m.m[4 * rw + 1], i1);
  const V s = cmad2(m.m[4 * rw + 2], i2, m.m[4 * rw + 3], i3);
  r.x += s.x;
  r.y += s.y;
  return r;
}

// One output row of the 4x4 chosen per lane: row RA, RB, RC or RD for (ul, uh) = (0, 0),
// (1, 0), (0, 1), (1, 1).  The coefficients are selected (uniform values, per-lane choice)
// and one row is formed with quad_row's operation order, instead of forming all four rows
// and keeping one.
template <int RA, int RB, int RC, int RD, typename T, typename V>
__device__ __forceinline__ V quad_row_lane(const Mat4<T>& m, bool ul, bool uh, const V& i0, const V& i1, const V& i2,
                                           const V& i3)
{
  Cx<T> c[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const Cx<T> lo = ul ? m.m[4 * RB + k] : m.m[4 * RA + k];
    const Cx<T> hi = ul ? m.m[4 * RD + k] : m.m[4 * RC + k];
    c[k] = uh ? hi : lo;
  }
  V r = cmad2(c[0], i0, c[1], i1);
  const V s = cmad2(c[2], i2, c[3], i3);
  r.x += s.x;
  r.y += s.y;
  return r;
}

// LOB: the lower position is pb (row bit 0); otherwise pa.  Inputs are named by (bit lo,
// bit hi); rows are row(l, h) = LOB ? 2h + l : 2l + h, all compile-time.
template <typename T, bool MOVE, bool LOB, int B, int P>
__device__ __forceinline__ void quad_lane_body(typename Vec<T>::t* __restrict__ st, unsigned lo, unsigned hi,
                                               const Mat4<T>& m)
{
  using V = typename Vec<T>::t;
  constexpr int R00 = 0, R10 = LOB ? 1 : 2, R01 = LOB ? 2 : 1, R11 = 3;  // row(l, h)
  const uint64_t r0 = xcd_block(block_id(), grid_blocks()) * (uint64_t)(B * P) + threadIdx.x;
  const int ml = 1 << lo;
  const bool ul = (threadIdx.x >> lo) & 1;  // this lane's bit lo
  if (hi >= 6) {
    const uint64_t hb = uint64_t(1) << hi;
    V a[P], b[P];
    uint64_t idx[P];
#pragma unroll
    for (int j = 0; j < P; ++j) {
      idx[j] = ins0(r0 + (uint64_t)j * B, hi);
      a[j] = ldg(&st[idx[j]]);
      b[j] = ldg(&st[idx[j] | hb]);
    }
#pragma unroll
    for (int j = 0; j < P; ++j) {
      V pa0, pb1;  // the partner lane's amplitudes with bit hi = 0 and 1
      pa0.x = __shfl_xor(a[j].x, ml);
      pa0.y = __shfl_xor(a[j].y, ml);
      pb1.x = __shfl_xor(b[j].x, ml);
      pb1.y = __shfl_xor(b[j].y, ml);
      V o0, o1;
      if (MOVE) {
        // rows 1 and 2 are (l, h) = (1, 0) and (0, 1): exchange across lanes and halves.
        o0 = ul ? pb1 : a[j];
        o1 = ul ? b[j] : pa0;
      } else {
        const V x00 = ul ? pa0 : a[j], x10 = ul ? a[j] : pa0;
        const V x01 = ul ? pb1 : b[j], x11 = ul ? b[j] : pb1;
        const V& i1 = LOB ? x10 : x01;
        const V& i2 = LOB ? x01 : x10;
        o0 = quad_row_lane<R00, R10, R00, R10>(m, ul, false, x00, i1, i2, x11);
        o1 = quad_row_lane<R01, R11, R01, R11>(m, ul, false, x00, i1, i2, x11);
      }
      stg(o0, &st[idx[j]]);
      stg(o1, &st[idx[j] | hb]);
    }
  } else {
    const int mh = 1 << hi;
    const bool uh = (threadIdx.x >> hi) & 1;
    V v[P];
    uint64_t idx[P];
#pragma unroll
    for (int j = 0; j < P; ++j) {
      idx[j] = r0 + (uint64_t)j * B;
      v[j] = ldg(&st[idx[j]]);
    }
#pragma unroll
    for (int j = 0; j < P; ++j) {
      V xl, xh, xb;
      xl.x = __shfl_xor(v[j].x, ml);
      xl.y = __shfl_xor(v[j].y, ml);
      xh.x = __shfl_xor(v[j].x, mh);
      xh.y = __shfl_xor(v[j].y, mh);
      xb.x = __shfl_xor(v[j].x, ml | mh);
      xb.y = __shfl_xor(v[j].y, ml | mh);
      V o;
      if (MOVE) {
        o = (ul != uh) ? xb : v[j];  // (1,0) <-> (0,1)
      } else {
        // (l, h) of this lane's amplitude is (ul, uh); name the four by absolute bits.
        const V x00 = ul ? (uh ? xb : xl) : (uh ? xh : v[j]);
        const V x10 = ul ? (uh ? xh : v[j]) : (uh ? xb : xl);
        const V x01 = ul ? (uh ? xl : xb) : (uh ? v[j] : xh);
        const V x11 = ul ? (uh ? v[j] : xh) : (uh ? xl : xb);
        o = quad_row_lane<R00, R10, R01, R11>(m, ul, uh, x00, LOB ? x10 : x01, LOB ? x01 : x10, x11);
      }
      stg(o, &st[idx[j]]);
    }
  }
}

template <typename T, bool MOVE, int B, int P>
__global__ __launch_bounds__(B) void k_quad_lane(typename Vec<T>::t* __restrict__ st, unsigned pa, unsigned pb,
                                                 Mat4<T> m)
{
  const unsigned lo = pa < pb ? pa : pb, hi = pa < pb ? pb : pa;
  if (lo == pb)
    quad_lane_body<T, MOVE, true, B, P>(st, lo, hi, m);
  else
    quad_lane_body<T, MOVE, false, B, P>(st, lo, hi, m);
}

// ---------------------------------------------------------------------------------------
// Dense 2^KQ block on arbitrary positions.  A block stages G groups of 2^KQ amplitudes in
// LDS (gathered through the position masks), then each thread produces one output
// amplitude of one group from the matrix held in LDS.
template <typename T, int KQ>
__global__ __launch_bounds__(kBlock) void k_block(typename Vec<T>::t* __restrict__ st, uint64_t ngroups,
                                                  Ins<KQ> ins, const typename Vec<T>::t* __restrict__ U)
{
  using V = typename Vec<T>::t;
  constexpr int D = 1 << KQ;
  constexpr int G = kBlock / D;  // groups per block
  __shared__ V sU[D * D];
  __shared__ V sx[G][D];
  for (int e = threadIdx.x; e < D * D; e += kBlock) sU[e] = U[e];
  // Per-row offsets: bit j of the row index sets position ins.p[j] (ins sorted ascending,
  // and the host orders the matrix rows accordingly).
  const int g = threadIdx.x / D, r = threadIdx.x % D;
  uint64_t off = 0;
#pragma unroll
  for (int j = 0; j < KQ; ++j)
    if ((r >> j) & 1) off |= uint64_t(1) << ins.p[j];
  const uint64_t grp = block_id() * G + g;
  uint64_t idx = 0;
  if (grp < ngroups) {
    idx = item_base(grp, ins) | off;
    sx[g][r] = st[idx];
  }
  __syncthreads();
  if (grp < ngroups) {
    T yr = 0, yi = 0;
#pragma unroll 8
    for (int c = 0; c < D; ++c) {
      const V u = sU[r * D + c], x = sx[g][c];
      yr += u.x * x.x - u.y * x.y;
      yi += u.x * x.y + u.y * x.x;
    }
    V y;
    y.x = yr;
    y.y = yi;
    st[idx] = y;
  }
}

// ---------------------------------------------------------------------------------------
// Dense 5-qubit block on the matrix cores.  y = U x for every group of 32 amplitudes, in the
// real form [yr; yi] = [[Ur, -Ui], [Ui, Ur]] [xr; xi] (a 64 x 64 real matrix).  One
// wavefront takes a tile of 16 groups: B = the 64 x 16 real tile, D = A B as 4 row blocks
// of 16, K = 64 in 16 steps of mfma 16x16x4 (f32 for complex64, f64 for complex128).
// Operand layouts (checked on gfx950 for every row and column): A lane l <-> A[l%16][l/16],
// B lane l <-> B[l/16][l%16]; accumulator element i of lane l <-> D[4*(l/16)+i][l%16] for
// the f32 instruction but D[(l/16)+4*i][l%16] for the f64 one.  So lane l (c = l/16) reads
// rows c + 4j (j = 0..7) of group l%16 and writes rows {4c+i, 16+4c+i} (f32) or
// {c+4i, 16+c+4i} (f64).  A stays in registers for all tiles of the wavefront; a tile is
// read completely before any of it is written.
template <typename T> struct MfmaAcc;
template <> struct MfmaAcc<float> { typedef float t __attribute__((ext_vector_type(4))); };
template <> struct MfmaAcc<double> { typedef double t __attribute__((ext_vector_type(4))); };

__device__ __forceinline__ MfmaAcc<float>::t mfma16x16x4(float a, float b, MfmaAcc<float>::t d)
{
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, d, 0, 0, 0);
}
__device__ __forceinline__ MfmaAcc<double>::t mfma16x16x4(double a, double b, MfmaAcc<double>::t d)
{
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, d, 0, 0, 0);
}

// NT: non-temporal accesses.  When the lowest block position lies inside a cache line, the
// rows of one group share lines and each load/store instruction covers part of every line
// it touches; with cached accesses L2 completes the lines (30-qubit complex64, positions
// {0,1,9,17,25}: 1.15 -> 4.1 TB/s; {1,9,15,22,29}: 2.9 -> 4.2 TB/s), while spread positions
// run the same either way (5.3 TB/s).
template <typename T, int TPW, bool NT>
__global__ __launch_bounds__(kBlock) void k_block5_mfma(typename Vec<T>::t* __restrict__ st, uint64_t ntiles, Ins<5> ins,
                                                        const typename Vec<T>::t* __restrict__ U)
{
  using V = typename Vec<T>::t;
  using Acc = typename MfmaAcc<T>::t;
  const int lane = threadIdx.x & 63, c = lane >> 4, gl = lane & 15;
  const uint64_t G = grid_blocks();
  const uint64_t bb = (G & 7) ? block_id() : xcd_block(block_id(), G);
  const uint64_t wave = bb * (kBlock / 64) + (threadIdx.x >> 6);
  const uint64_t t0 = wave * TPW;
  if (t0 >= ntiles) return;
  // A operands: a[mb][ks] = A[16 mb + gl][4 ks + c].
  T a[4][16];
#pragma unroll
  for (int mb = 0; mb < 4; ++mb)
#pragma unroll
    for (int ks = 0; ks < 16; ++ks) {
      const int R = 16 * mb + gl, Cc = 4 * ks + c;
      const V u = U[(R & 31) * 32 + (Cc & 31)];
      a[mb][ks] = (R < 32) ? ((Cc < 32) ? u.x : -u.y) : ((Cc < 32) ? u.y : u.x);
    }
  // Row offsets: row bit j sets position ins.p[j].
  uint64_t offin[8], offout[8];
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    const int rin = c + 4 * j;
    const int i = j & 3;
    const int rin4 = (sizeof(T) == 8) ? (c + 4 * i) : (4 * c + i);  // accumulator row of element i
    const int rout = (j < 4) ? rin4 : (16 + rin4);
    uint64_t oi = 0, oo = 0;
#pragma unroll
    for (int b = 0; b < 5; ++b) {
      if ((rin >> b) & 1) oi |= uint64_t(1) << ins.p[b];
      if ((rout >> b) & 1) oo |= uint64_t(1) << ins.p[b];
    }
    offin[j] = oi;
    offout[j] = oo;
  }
  const uint64_t t1 = (t0 + TPW < ntiles) ? t0 + TPW : ntiles;
  // Software pipeline: the next tile's loads are in flight while this tile is multiplied.
  uint64_t base = item_base(t0 * 16 + gl, ins);
  V x[8];
#pragma unroll
  for (int j = 0; j < 8; ++j) x[j] = ldx<NT>(&st[base | offin[j]]);
  for (uint64_t t = t0; t < t1; ++t) {
    const bool more = t + 1 < t1;
    const uint64_t nbase = more ? item_base((t + 1) * 16 + gl, ins) : base;
    V xn[8];
    if (more) {
#pragma unroll
      for (int j = 0; j < 8; ++j) xn[j] = ldx<NT>(&st[nbase | offin[j]]);
    }
    Acc d[4];
#pragma unroll
    for (int mb = 0; mb < 4; ++mb) d[mb] = Acc{0, 0, 0, 0};
#pragma unroll
    for (int ks = 0; ks < 16; ++ks) {
      const T bv = (ks < 8) ? x[ks].x : x[ks - 8].y;
#pragma unroll
      for (int mb = 0; mb < 4; ++mb) d[mb] = mfma16x16x4(a[mb][ks], bv, d[mb]);
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      V y0, y1;
      y0.x = d[0][i];
      y0.y = d[2][i];
      y1.x = d[1][i];
      y1.y = d[3][i];
      stx<NT>(y0, &st[base | offout[i]]);
      stx<NT>(y1, &st[base | offout[4 + i]]);
    }
    if (more) {
#pragma unroll
      for (int j = 0; j < 8; ++j) x[j] = xn[j];
    }
    base = nbase;
  }
}

// The same product with A in LDS instead of registers: the 64 x 64 real matrix is stored
// in fragment order, sA[(16 mb + ks) * 64 + lane] = A[16 mb + lane%16][4 ks + lane/16], so
// every MFMA operand read is one conflict-free 512-byte LDS access, and the 128 VGPRs (f64)
// that held A are free for more resident waves.  PF: load the next tile while this one is
// multiplied.  The MFMA sequence and operands are those of k_block5_mfma (same results).
template <typename T, int TPW, bool NT, bool PF>
__global__ __launch_bounds__(kBlock) void k_block5_mfma_lds(typename Vec<T>::t* __restrict__ st, uint64_t ntiles,
                                                            Ins<5> ins, const typename Vec<T>::t* __restrict__ U)
{
  using V = typename Vec<T>::t;
  using Acc = typename MfmaAcc<T>::t;
  __shared__ T sA[64 * 64];
  for (int e = threadIdx.x; e < 64 * 64; e += kBlock) {
    const int mb = e >> 10, ks = (e >> 6) & 15, l = e & 63;
    const int R = 16 * mb + (l & 15), Cc = 4 * ks + (l >> 4);
    const V u = U[(R & 31) * 32 + (Cc & 31)];
    sA[e] = (R < 32) ? ((Cc < 32) ? u.x : -u.y) : ((Cc < 32) ? u.y : u.x);
  }
  __syncthreads();
  const int lane = threadIdx.x & 63, c = lane >> 4, gl = lane & 15;
  const uint64_t G = grid_blocks();
  const uint64_t bb = (G & 7) ? block_id() : xcd_block(block_id(), G);
  const uint64_t wave = bb * (kBlock / 64) + (threadIdx.x >> 6);
  const uint64_t t0 = wave * TPW;
  if (t0 >= ntiles) return;
  uint64_t offin[8], offout[8];
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    const int rin = c + 4 * j;
    const int i = j & 3;
    const int rin4 = (sizeof(T) == 8) ? (c + 4 * i) : (4 * c + i);
    const int rout = (j < 4) ? rin4 : (16 + rin4);
    uint64_t oi = 0, oo = 0;
#pragma unroll
    for (int b = 0; b < 5; ++b) {
      if ((rin >> b) & 1) oi |= uint64_t(1) << ins.p[b];
      if ((rout >> b) & 1) oo |= uint64_t(1) << ins.p[b];
    }
    offin[j] = oi;
    offout[j] = oo;
  }
  const uint64_t t1 = (t0 + TPW < ntiles) ? t0 + TPW : ntiles;
  uint64_t base = item_base(t0 * 16 + gl, ins);
  V x[8];
  if (PF) {
#pragma unroll
    for (int j = 0; j < 8; ++j) x[j] = ldx<NT>(&st[base | offin[j]]);
  }
  for (uint64_t t = t0; t < t1; ++t) {
    // Keep the A reads inside the loop: hoisted out of it, A would occupy registers again.
    __asm__ volatile("" ::: "memory");
    const bool more = PF && t + 1 < t1;
    const uint64_t nbase = more ? item_base((t + 1) * 16 + gl, ins) : base;
    V xn[PF ? 8 : 1];
    if (PF) {
      if (more) {
#pragma unroll
        for (int j = 0; j < 8; ++j) xn[j] = ldx<NT>(&st[nbase | offin[j]]);
      }
    } else {
      base = item_base(t * 16 + gl, ins);
#pragma unroll
      for (int j = 0; j < 8; ++j) x[j] = ldx<NT>(&st[base | offin[j]]);
    }
    Acc d[4];
#pragma unroll
    for (int mb = 0; mb < 4; ++mb) d[mb] = Acc{0, 0, 0, 0};
#pragma unroll
    for (int ks = 0; ks < 16; ++ks) {
      const T bv = (ks < 8) ? x[ks].x : x[ks - 8].y;
#pragma unroll
      for (int mb = 0; mb < 4; ++mb) d[mb] = mfma16x16x4(sA[(16 * mb + ks) * 64 + lane], bv, d[mb]);
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      V y0, y1;
      y0.x = d[0][i];
      y0.y = d[2][i];
      y1.x = d[1][i];
      y1.y = d[3][i];
      stx<NT>(y0, &st[base | offout[i]]);
      stx<NT>(y1, &st[base | offout[4 + i]]);
    }
    if (PF) {
      if (more) {
#pragma unroll
        for (int j = 0; j < 8; ++j) x[j] = xn[PF ? j : 0];
      }
      base = nbase;
    }
  }
}

// The same block product with three real 32 x 32 products instead of one 64 x 64 one
// (Gauss's three-multiplication form of a complex product):
//   T1 = Ur xr,  T2 = Ui xi,  T3 = (Ur + Ui)(xr + xi);   yr = T1 - T2,  yi = T3 - T1 - T2,
// 48 MFMAs per tile instead of 64 (192 flop per amplitude instead of 256).  The three
// matrices live in LDS in fragment order, sA[((m * 2 + mb) * 8 + ks) * 64 + lane] =
// A_m[16 mb + lane%16][4 ks + lane/16] for A_0 = Ur, A_1 = Ui, A_2 = Ur + Ui; B is row
// c + 4 ks of xr, xi or xr + xi.  Row offsets and accumulator layouts as k_block5_mfma.
// Rounding: yi carries the error of the larger (Ur + Ui)(xr + xi) product, a few units of
// the last place of |U||x| per element instead of one.
template <typename T, int TPW, bool NT>
__global__ __launch_bounds__(kBlock) void k_block5_mfma3(typename Vec<T>::t* __restrict__ st, uint64_t ntiles,
                                                         Ins<5> ins, const typename Vec<T>::t* __restrict__ U)
{
  using V = typename Vec<T>::t;
  using Acc = typename MfmaAcc<T>::t;
  __shared__ T sA[3 * 2 * 8 * 64];
  for (int e = threadIdx.x; e < 3 * 2 * 8 * 64; e += kBlock) {
    const int m = e >> 10, mb = (e >> 9) & 1, ks = (e >> 6) & 7, l = e & 63;
    const V u = U[(16 * mb + (l & 15)) * 32 + 4 * ks + (l >> 4)];
    sA[e] = m == 0 ? u.x : (m == 1 ? u.y : u.x + u.y);
  }
  __syncthreads();
  const int lane = threadIdx.x & 63, c = lane >> 4, gl = lane & 15;
  const uint64_t G = grid_blocks();
  const uint64_t bb = (G & 7) ? block_id() : xcd_block(block_id(), G);
  const uint64_t wave = bb * (kBlock / 64) + (threadIdx.x >> 6);
  const uint64_t t0 = wave * TPW;
  if (t0 >= ntiles) return;
  uint64_t offin[8], offout[8];
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    const int rin = c + 4 * j;
    const int i = j & 3;
    const int rin4 = (sizeof(T) == 8) ? (c + 4 * i) : (4 * c + i);
    const int rout = (j < 4) ? rin4 : (16 + rin4);
    uint64_t oi = 0, oo = 0;
#pragma unroll
    for (int b = 0; b < 5; ++b) {
      if ((rin >> b) & 1) oi |= uint64_t(1) << ins.p[b];
      if ((rout >> b) & 1) oo |= uint64_t(1) << ins.p[b];
    }
    offin[j] = oi;
    offout[j] = oo;
  }
  const uint64_t t1 = (t0 + TPW < ntiles) ? t0 + TPW : ntiles;
  uint64_t base = item_base(t0 * 16 + gl, ins);
  V x[8];
#pragma unroll
  for (int j = 0; j < 8; ++j) x[j] = ldx<NT>(&st[base | offin[j]]);
  for (uint64_t t = t0; t < t1; ++t) {
    // Keep the A reads inside the loop: hoisted out of it, A would occupy registers.
    __asm__ volatile("" ::: "memory");
    const bool more = t + 1 < t1;
    const uint64_t nbase = more ? item_base((t + 1) * 16 + gl, ins) : base;
    V xn[8];
    if (more) {
#pragma unroll
      for (int j = 0; j < 8; ++j) xn[j] = ldx<NT>(&st[nbase | offin[j]]);
    }
    Acc d1[2], d2[2], d3[2];
#pragma unroll
    for (int mb = 0; mb < 2; ++mb) d1[mb] = d2[mb] = d3[mb] = Acc{0, 0, 0, 0};
#pragma unroll
    for (int ks = 0; ks < 8; ++ks) {
      const T br = x[ks].x, bi = x[ks].y, bs = br + bi;
#pragma unroll
      for (int mb = 0; mb < 2; ++mb) {
        d1[mb] = mfma16x16x4(sA[((0 + mb) * 8 + ks) * 64 + lane], br, d1[mb]);
        d2[mb] = mfma16x16x4(sA[((2 + mb) * 8 + ks) * 64 + lane], bi, d2[mb]);
        d3[mb] = mfma16x16x4(sA[((4 + mb) * 8 + ks) * 64 + lane], bs, d3[mb]);
      }
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      V y0, y1;
      y0.x = d1[0][i] - d2[0][i];
      y0.y = d3[0][i] - d1[0][i] - d2[0][i];
      y1.x = d1[1][i] - d2[1][i];
      y1.y = d3[1][i] - d1[1][i] - d2[1][i];
      stx<NT>(y0, &st[base | offout[i]]);
      stx<NT>(y1, &st[base | offout[4 + i]]);
    }
    if (more) {
#pragma unroll
      for (int j = 0; j < 8; ++j) x[j] = xn[j];
    }
    base = nbase;
  }
}

// ---------------------------------------------------------------------------------------
// Fills.
// Loop nest of the reference's Loop_TN (src/highperfkernels.cpp:397-486): for i = c11 +
// ii c13, j = i + c21 + jj c23, k = j + c31 + kk (ii < ni, jj < nj, kk < nk), the 2x2 m on
// the pair (k, k + shift).  The reference's sub-range gate helpers reduce to it (Loop_DN is
// the nest with c21 = 0, nj = 1).  A rare path (ranges of a cache-blocked caller): one item
// per pair with 64-bit divisions, grid-stride.
struct Nest {
  uint64_t c11, c13, c21, c23, c31, ni, nj, nk, shift;
};

template <typename T>
__global__ __launch_bounds__(kBlock) void k_nest(typename Vec<T>::t* __restrict__ st, Nest nt, uint64_t items, Mat2<T> m)
{
  for (uint64_t w = block_id() * kBlock + threadIdx.x; w < items; w += grid_blocks() * kBlock) {
    const uint64_t kk = w % nt.nk, r = w / nt.nk, jj = r % nt.nj, ii = r / nt.nj;
    const uint64_t i0 = nt.c11 + ii * nt.c13 + nt.c21 + jj * nt.c23 + nt.c31 + kk, i1 = i0 + nt.shift;
    const typename Vec<T>::t a = st[i0], b = st[i1];
    st[i0] = cmad2(m.m00, a, m.m01, b);
    st[i1] = cmad2(m.m10, a, m.m11, b);
  }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_fill(typename Vec<T>::t* __restrict__ st, uint64_t N, Cx<T> v)
{
  using V = typename Vec<T>::t;
  V a;
  a.x = v.r;
  a.y = v.i;
  for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < N; i += (uint64_t)gridDim.x * kBlock)
    stg(a, &st[i]);
}

// dst := src in dst's precision
template <typename TD, typename TS>
__global__ __launch_bounds__(kBlock) void k_convert(typename Vec<TD>::t* __restrict__ dst,
                                                    const typename Vec<TS>::t* __restrict__ src, uint64_t N)
{
  for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < N; i += (uint64_t)gridDim.x * kBlock) {
    const auto a = src[i];
    typename Vec<TD>::t b;
    b.x = (TD)a.x;
    b.y = (TD)a.y;
    dst[i] = b;
  }
}

// h += f * x
template <typename T>
__global__ __launch_bounds__(kBlock) void k_axpy(typename Vec<T>::t* __restrict__ h, const typename Vec<T>::t* __restrict__ x,
                                                 uint64_t N, Cx<T> f)
{
  for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < N; i += (uint64_t)gridDim.x * kBlock) {
    auto a = h[i];
    const auto b = cmul(f, x[i]);
    a.x += b.x;
    a.y += b.y;
    h[i] = a;
  }
}

// ---------------------------------------------------------------------------------------
// Byte-table index transforms.  A 64-bit value f(x) that is a XOR (or OR) of per-bit
// contributions c[p] over the set bits p of x is evaluated as the XOR of up to 8 table
// lookups T[byte][x's byte], the tables built in LDS by each block from c[] (kernarg).

struct BitTab {
  uint64_t c[64];   // contribution of each index bit
  unsigned nbytes;  // number of index bytes in use (ceil(n / 8))
};

__device__ __forceinline__ void build_tables(uint64_t (*tab)[256], const BitTab& bt)
{
  for (unsigned e = threadIdx.x; e < bt.nbytes * 256; e += kBlock) {
    const unsigned byte = e >> 8, v = e & 255;
    uint64_t acc = 0;
#pragma unroll
    for (int k = 0; k < 8; ++k)
      if ((v >> k) & 1) acc ^= bt.c[byte * 8 + k];
    tab[byte][v] = acc;
  }
}

__device__ __forceinline__ uint64_t eval_tables(const uint64_t (*tab)[256], unsigned nbytes, uint64_t x)
{
  uint64_t acc = 0;
  for (unsigned b = 0; b < nbytes; ++b) acc ^= tab[b][(x >> (8 * b)) & 255];
  return acc;
}

// Permutation: new[j] = old[src(j)], src(j) = OR of (bit b of j) << src_pos[b].
template <typename T>
__global__ __launch_bounds__(kBlock) void k_permute(typename Vec<T>::t* __restrict__ dst,
                                                    const typename Vec<T>::t* __restrict__ src, uint64_t N, BitTab bt)
{
  __shared__ uint64_t tab[8][256];
  build_tables(tab, bt);
  __syncthreads();
  for (uint64_t j = (uint64_t)blockIdx.x * kBlock + threadIdx.x; j < N; j += (uint64_t)gridDim.x * kBlock)
    stg(src[eval_tables(tab, bt.nbytes, j)], &dst[j]);
}

// Tiled bit permutation, for permutations that move low positions up (where the gather
// above reads with a stride).  A tile is the set of 2^m destination indices that differ only
// in the tile bits: the perm_low<T>() lowest destination bits plus every destination bit
// whose source position is below that (m <= 2 perm_low<T>()).  The tile is read in source order
// (tile bits sorted by source position, so consecutive threads read consecutive source
// amplitudes), kept in LDS in destination order, and written in destination order.
// perm_low: 5 bits, runs of 32 amplitudes (512 bytes complex128, 256 complex64).  Measured on
// MI355X with the loads unrolled: complex64 at 32 qubits 5.5-5.65 TB/s (6 bits, 512-byte runs
// in a 32 KiB tile: 5.40), complex128 at 30 qubits 5.83 TB/s with 512-thread workgroups; with
// one load in flight per thread and iteration the same tiles ran at 1.75 (complex64, 6
// bits) and 4.64 TB/s.
template <typename T>
constexpr unsigned perm_low() { return 5; }
struct PermTile {
  unsigned m;          // tile bits
  uint64_t tj[16];     // destination bits of the tile, ascending: 1 << destination position
  unsigned sd[16];     // i-th tile bit in source order: 1 << its index among the tile bits
  uint64_t ss[16];     // i-th tile bit in source order: 1 << its source position
  unsigned nout;       // destination bits outside the tile
  uint64_t oj[64];     // bit i of the tile index: 1 << its destination position
  uint64_t os[64];     // bit i of the tile index: 1 << its source position
};

// Threads per workgroup of k_permute_tiled: 512 complex128 (two amplitudes per thread; 256
// threads with four ran 4% slower at 30 qubits), 256 complex64 (four; 512 with two ran 15%
// slower at 32 qubits).
template <typename T>
constexpr int perm_threads() { return sizeof(T) == 8 ? 512 : 256; }

template <typename T, int TB>
__global__ __launch_bounds__(TB) void k_permute_tiled(typename Vec<T>::t* __restrict__ dst,
                                                      const typename Vec<T>::t* __restrict__ src, PermTile pt)
{
  using V = typename Vec<T>::t;
  // The host pads every tile to exactly LT bits: each of the TB threads moves M / TB
  // amplitudes, all of its loads in flight at once.  Tile position r = threadIdx.x + k TB:
  // its low LB bits come from the thread, the rest from k (a constant in the unrolled
  // loops).  The per-thread offsets are selected without branches, so that their scalar
  // loads issue together instead of one wait per set bit.
  constexpr unsigned LT = 2 * perm_low<T>(), M = 1u << LT, LB = TB == 512 ? 9 : 8, K = M / TB;
  static_assert(TB == (1 << LB) && M >= TB, "tile of at least one amplitude per thread");
  __shared__ V tile[M];
  const uint64_t t = block_id();
  uint64_t sa = 0, ja = 0;
  unsigned d_t = 0;
#pragma unroll 4
  for (unsigned i = 0; i < pt.nout; ++i) {
    const bool on = (t >> i) & 1;
    const uint64_t oj = pt.oj[i], os = pt.os[i];
    ja |= on ? oj : 0;
    sa |= on ? os : 0;
  }
#pragma unroll
  for (unsigned i = 0; i < LB; ++i) {
    const bool on = (threadIdx.x >> i) & 1;
    const uint64_t ss = pt.ss[i], tj = pt.tj[i];
    const unsigned sd = pt.sd[i];
    sa |= on ? ss : 0;
    ja |= on ? tj : 0;
    d_t |= on ? sd : 0;
  }
  V v[K];
  unsigned d[K];
#pragma unroll
  for (unsigned k = 0; k < K; ++k) {
    uint64_t s = sa;
    d[k] = d_t;
#pragma unroll
    for (unsigned i = LB; i < LT; ++i)
      if ((k >> (i - LB)) & 1) {
        s |= pt.ss[i];
        d[k] |= pt.sd[i];
      }
    v[k] = ldg(&src[s]);
  }
#pragma unroll
  for (unsigned k = 0; k < K; ++k) tile[d[k]] = v[k];
  __syncthreads();
#pragma unroll
  for (unsigned k = 0; k < K; ++k) {
    uint64_t j = ja;
#pragma unroll
    for (unsigned i = LB; i < LT; ++i)
      if ((k >> (i - LB)) & 1) j |= pt.tj[i];
    stg(tile[threadIdx.x + k * TB], &dst[j]);
  }
}

// ---------------------------------------------------------------------------------------
// QAOA MaxCut: edge parity vector E(x) (bit e = [x_u != x_v]) from incidence tables; the cut
// is popcount(E) (unit weights) or sum of weights over set bits of E (byte tables).
struct CutSpec {
  BitTab inc;         // inc.c[p] = mask of edges incident to data position p
  double w[64];       // edge weights
  unsigned ne;        // number of edges (<= 64)
  int weighted;       // 0: integer cut -> phase table; 1: weighted cut
  double ph_re[65];   // unit weights: exp(-i gamma c) for c = 0..ne (host-computed)
  double ph_im[65];
  double gamma;
};

__device__ __forceinline__ void build_weight_tables(double (*wt)[256], const CutSpec& cs)
{
  const unsigned nb = (cs.ne + 7) / 8;
  for (unsigned e = threadIdx.x; e < nb * 256; e += kBlock) {
    const unsigned byte = e >> 8, v = e & 255;
    double acc = 0;
    for (int k = 0; k < 8; ++k)
      if (((v >> k) & 1) && byte * 8 + k < cs.ne) acc += cs.w[byte * 8 + k];
    wt[byte][v] = acc;
  }
}

__device__ __forceinline__ double cut_value(const uint64_t (*tab)[256], const double (*wt)[256], const CutSpec& cs,
                                            uint64_t x, int* icut)
{
  const uint64_t E = eval_tables(tab, cs.inc.nbytes, x);
  if (!cs.weighted) {
    *icut = __popcll(E);
    return (double)*icut;
  }
  double c = 0;
  const unsigned nb = (cs.ne + 7) / 8;
  for (unsigned b = 0; b < nb; ++b) c += wt[b][(E >> (8 * b)) & 255];
  *icut = 0;
  return c;
}

}  // namespace iqsk
