// Layers of one 2x2 gate on every qubit (U^(x)n), forward and adjoint: the QAOA mixer
// exp(-i beta sum_q X_q) = prod_q RX_q(2 beta) and its backward step.  Entry points in
// include/iqs_hip.h ("layers").
//
// Instead of n passes over the state (one per qubit), a layer takes
//   * one lane pass for positions 0..7: each lane holds 4 amplitudes 64 apart, partners at
//     lane ^ 2^q (positions 0-5) come through __shfl_xor, positions 6 and 7 pair the
//     lane's own registers, and every load/store is a contiguous run of 64 amplitudes;
//   * passes of up to five positions each above that (four in the adjoint, which holds psi
//     and lambda): a thread holds the 2^m amplitudes of one group in registers
//     (consecutive threads take consecutive groups, so every instruction is a contiguous
//     run) and applies the m rotations there.
// At 28 qubits: 5 passes instead of 28 forward, 6 adjoint.
//
// Adjoint: the gates of a layer act on different qubits, so each generator G_q commutes
// with every gate of the layer except its own, and inserting G_q after U_q equals inserting
// it after the whole layer.  Each pass therefore takes its qubits' terms
// 2 Im <lam| G_q |psi> on the states it holds, then applies Uinv to them.  Sums in FP64:
// per-workgroup partials, summed in order on the host (bitwise reproducible).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <vector>

#include "../../include/iqs_hip.h"
#include "iqs_internal.hpp"
#include "iqs_kernels.hpp"

using namespace iqsk;

namespace {

constexpr int kLB = 256;          // threads per workgroup
constexpr int kLowQ = 6;          // positions exchanged by lane shuffles in the lane pass (lane bits)
constexpr int kLowP = 4;          // amplitudes per lane in the lane pass
constexpr int kQaoaAdjP = 4;      // amplitudes of psi and of lambda per lane in the QAOA adjoint lane pass
constexpr int kRegQ = 4;          // positions per register pass (adjoint: psi and lambda)
constexpr int kRegQF = 5;         // positions per register pass (forward)
constexpr unsigned kMaxGrid = 2048;    // workgroups of the grid-stride register passes (small states)
// Lane passes: kLowRowsFwd, kLowRowsAdj rows of kLB * P amplitudes per workgroup
// (28-qubit complex128 with the phase, RX/X forms: forward 1.43 ms with one row, 1.48 with
// two, 1.56 with four; adjoint 3.01 ms with two rows, 3.09 with one, 3.21 with four; 2048
// workgroups of 128 rows each streamed 10-20% slower).
constexpr uint64_t kLowRowsFwd = 1, kLowRowsAdj = 2;
constexpr uint64_t kLowMaxGrid = uint64_t(1) << 22;
// Register passes over >= kRegTileMin groups run one group per thread and one tile of kLB
// groups per workgroup (XCD-contiguous, split into kRegSub runs of kLB / kRegSub groups
// 2^kSubLog groups apart): 28-qubit complex128 forward pass 5.31 -> 5.86 TB/s against a
// persistent grid-stride loop over 2048 workgroups.
constexpr int kRegSub = 2;
constexpr uint64_t kRegTileMin = uint64_t(1) << 16;

struct LayerWork {
  double* part = nullptr;
  size_t have = 0;
  std::vector<double> host;
};
LayerWork g_lw;
std::mutex g_lm;

// Partial-sum buffer of at least `count` doubles (2 * kMaxGrid at the least).
int part_buf(size_t count = 0)
{
  count = std::max<size_t>(count, 2 * kMaxGrid);
  if (g_lw.have >= count) return IQSH_OK;
  IQSI_TRY(hipStreamSynchronize(iqsi::stream()));
  if (g_lw.part) IQSI_TRY(hipFree(g_lw.part));
  g_lw.part = nullptr;
  g_lw.have = 0;
  if (hipMalloc((void**)&g_lw.part, sizeof(double) * count) != hipSuccess)
    return iqsi::fail(IQSH_ERR_ALLOC, "layer reduction buffer of %zu doubles", count);
  g_lw.have = count;
  return IQSH_OK;
}

template <typename T, typename V>
__device__ __forceinline__ V mat_row(const Cx<T>& a, const V& x0, const Cx<T>& b, const V& x1)
{
  return cmad2(a, x0, b, x1);
}

// Im(conj(y) * (a x0 + b x1)) with the product in FP64.
template <typename V>
__device__ __forceinline__ double im_term(const V& y, const Cx<double>& a, const V& x0, const Cx<double>& b, const V& x1)
{
  const double gr = (a.r * (double)x0.x - a.i * (double)x0.y) + (b.r * (double)x1.x - b.i * (double)x1.y);
  const double gi = (a.r * (double)x0.y + a.i * (double)x0.x) + (b.r * (double)x1.y + b.i * (double)x1.x);
  return (double)y.x * gi - (double)y.y * gr;
}

// Generator term of one amplitude of a pair (x0, x1): Im(conj(y) (G x)_row), row 1 if hi.
// GX: G = X (the QAOA mixer's generator), (G x)_lo = x1 and (G x)_hi = x0 -- two products
// instead of ten.
template <bool GX, typename V>
__device__ __forceinline__ double gen_term(const V& y, const Mat2<double>& g, const V& x0, const V& x1, bool hi)
{
  if (GX) {
    const V o = hi ? x0 : x1;
    // complex64: the product in FP32 (of FP32 amplitudes), summed in FP64 by the caller
    if constexpr (sizeof(o.x) == 4) return (double)(y.x * o.y - y.y * o.x);
    return (double)y.x * (double)o.y - (double)y.y * (double)o.x;
  }
  return hi ? im_term(y, g.m10, x0, g.m11, x1) : im_term(y, g.m00, x0, g.m01, x1);
}

// Kernel form flags: kFormGX, the generator is X (gen_term's two-product form); kFormRX,
// the 2x2 has a real diagonal and an imaginary off-diagonal (the mixer RX(2 beta) and its
// inverse), so a row of the product is d x_same + i o x_other: four products instead of eight.
constexpr int kFormGX = 1, kFormRX = 2;

// (The lane passes keep the general product written out at the call site: behind this helper
// the compiler gave their forward kernels 86 instead of 56 VGPRs.)
template <int F, typename T, typename V>
__device__ __forceinline__ V urow(const Mat2<T>& u, const V& x0, const V& x1, bool hi)
{
  if constexpr ((F & kFormRX) != 0) {
    const T d = hi ? u.m11.r : u.m00.r, o = hi ? u.m10.i : u.m01.i;
    const V xs = hi ? x1 : x0, xo = hi ? x0 : x1;
    V r;
    r.x = d * xs.x - o * xo.y;
    r.y = d * xs.y + o * xo.x;
    return r;
  } else {
    return hi ? mat_row(u.m10, x0, u.m11, x1) : mat_row(u.m00, x0, u.m01, x1);
  }
}

// Workgroup sum of v (fixed order: lanes by xor-shuffle, then the waves in order) stored
// at out[slot] by thread 0.
__device__ __forceinline__ void wg_sum_store(double v, double* out, uint64_t slot)
{
  __shared__ double sh[kLB / 64];
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) sh[wid] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0;
    for (int w = 0; w < kLB / 64; ++w) s += sh[w];
    out[slot] = s;
  }
}
__device__ __forceinline__ void wg_sum_store(double v, double* out) { wg_sum_store(v, out, blockIdx.x); }

// out[0] = sum of part[0 .. n-1] in a fixed order (one workgroup): thread t adds part[t],
// part[t + kLB], ... in sequence, then the workgroup sum.
__global__ __launch_bounds__(kLB) void k_sum_part(const double* __restrict__ part, uint64_t n, double* __restrict__ out)
{
  double v = 0;
  for (uint64_t i = threadIdx.x; i < n; i += kLB) v += part[i];
  wg_sum_store(v, out, 0);
}

template <typename V>
__device__ __forceinline__ V shfl_v(const V& a, int mask)
{
  V r;
  r.x = __shfl_xor(a.x, mask);
  r.y = __shfl_xor(a.y, mask);
  return r;
}

// Row layout of the lane passes: a row of kLB * P amplitudes, thread t holding in-row
// indices (wave << (6 + log2 P)) | (j << 6) | lane for j < P.  Positions 0-5 are lane bits
// (exchanged by shuffles), positions 6 .. 5 + log2 P register bits (register-local pairs),
// and every load/store instruction covers 64 consecutive amplitudes.
template <int P>
__device__ __forceinline__ unsigned row_index(unsigned tid, int j)
{
  constexpr int LP = P == 1 ? 0 : (P == 2 ? 1 : (P == 4 ? 2 : 3));
  static_assert((1 << LP) == P, "P must be 1, 2, 4 or 8");
  return ((tid >> 6) << (6 + LP)) | ((unsigned)j << 6) | (tid & 63);
}

// The gate on the register positions of a lane pass (bit b of j: position 6 + b), with the
// generator terms of the adjoint taken before each position is undone.
template <typename T, bool ADJ, int P, int F, typename V>
__device__ __forceinline__ void reg_positions(V* x, V* y, const Mat2<T>& u, const Mat2<double>& g, double& acc)
{
#pragma unroll
  for (int b = 0; (1 << b) < P; ++b)
#pragma unroll
    for (int j = 0; j < P; ++j) {
      if ((j >> b) & 1) continue;
      const int j1 = j | (1 << b);
      const V a = x[j], c = x[j1];
      if (ADJ) acc += gen_term<(F & kFormGX) != 0>(y[j], g, a, c, false) + gen_term<(F & kFormGX) != 0>(y[j1], g, a, c, true);
      x[j] = urow<F>(u, a, c, false);
      x[j1] = urow<F>(u, a, c, true);
      if (ADJ) {
        const V ya = y[j], yc = y[j1];
        y[j] = urow<F>(u, ya, yc, false);
        y[j1] = urow<F>(u, ya, yc, true);
      }
    }
}

// Lane pass over positions 0 .. kLowQ + log2(kLowP) - 1 (shuffles for 0-5, registers above).
// N a multiple of kLB * kLowP.
template <typename T, bool ADJ, int F>
__global__ __launch_bounds__(kLB) void k_layer_low(typename Vec<T>::t* __restrict__ psi, typename Vec<T>::t* __restrict__ lam,
                                                   uint64_t N, Mat2<T> u, Mat2<double> g, double* __restrict__ part)
{
  using V = typename Vec<T>::t;
  const int lane = threadIdx.x & 63;
  double acc = 0;
  // Rows of kLB * kLowP amplitudes; one contiguous, XCD-contiguous chunk of rows per
  // workgroup (the bounds are uniform over the workgroup, as the shuffles need).
  const uint64_t rows = N / (kLB * kLowP), G = gridDim.x;
  const uint64_t lb = (G & 7) ? blockIdx.x : xcd_block(blockIdx.x, G);
  const uint64_t chunk = (rows + G - 1) / G;
  const uint64_t b0 = lb * chunk, b1 = (b0 + chunk < rows) ? b0 + chunk : rows;
  for (uint64_t rb = b0; rb < b1; ++rb) {
    const uint64_t r0 = rb * (kLB * kLowP);
    V x[kLowP], y[kLowP];
#pragma unroll
    for (int j = 0; j < kLowP; ++j) {
      x[j] = ldg(&psi[r0 + row_index<kLowP>(threadIdx.x, j)]);
      if (ADJ) y[j] = ldg(&lam[r0 + row_index<kLowP>(threadIdx.x, j)]);
    }
#pragma unroll 1
    for (int q = 0; q < kLowQ; ++q) {
      const int msk = 1 << q;
      const bool hi = (lane >> q) & 1;
#pragma unroll
      for (int j = 0; j < kLowP; ++j) {
        const V xp = shfl_v(x[j], msk);
        const V x0 = hi ? xp : x[j], x1 = hi ? x[j] : xp;
        if (ADJ) {
          acc += gen_term<(F & kFormGX) != 0>(y[j], g, x0, x1, hi);
          const V yp = shfl_v(y[j], msk);
          const V y0 = hi ? yp : y[j], y1 = hi ? y[j] : yp;
          y[j] = (F & kFormRX) ? urow<F>(u, y0, y1, hi) : (hi ? mat_row(u.m10, y0, u.m11, y1) : mat_row(u.m00, y0, u.m01, y1));
        }
        x[j] = (F & kFormRX) ? urow<F>(u, x0, x1, hi) : (hi ? mat_row(u.m10, x0, u.m11, x1) : mat_row(u.m00, x0, u.m01, x1));
      }
    }
    reg_positions<T, ADJ, kLowP, F>(x, y, u, g, acc);
#pragma unroll
    for (int j = 0; j < kLowP; ++j) {
      stg(x[j], &psi[r0 + row_index<kLowP>(threadIdx.x, j)]);
      if (ADJ) stg(y[j], &lam[r0 + row_index<kLowP>(threadIdx.x, j)]);
    }
  }
  if (ADJ) wg_sum_store(acc, part);
}

// The lane pass with a QAOA phase layer fused in: forward, psi_i *= exp(-i gamma cut_i)
// before the rotations; adjoint (run as the last pass of the backward step, after every
// other rotation of the layer is undone), the rotation terms, the undo, then the phase term
// cut_i Im(conj(lam_i) psi_i) and psi, lam *= exp(+i gamma cut) (cs built with -gamma).
// part receives two sums per workgroup (rotations, phase).
//
// The cut edges of index x are E(x) = XOR of inc[b] over the set bits b of x (bit e of
// inc[b]: edge e touches position b), which is linear over GF(2): E(row | lo) =
// E(row) ^ E(lo).  Unit weights (cut = popcount E, phase from a 65-entry table): each thread
// takes E(lo) of its P in-row offsets once, and E(row) comes from the row's set bits, the
// same for the whole workgroup -- no per-workgroup tables, so short workgroups as in the
// plain lane pass.  W: weighted cuts, summed from byte tables of weight sums in LDS.
template <typename T, bool ADJ, bool W, int F>
__global__ __launch_bounds__(kLB) void k_qaoa_low(typename Vec<T>::t* __restrict__ psi, typename Vec<T>::t* __restrict__ lam,
                                                  uint64_t N, Mat2<T> u, Mat2<double> g, CutSpec cs,
                                                  double* __restrict__ part)
{
  using V = typename Vec<T>::t;
  constexpr int P = ADJ ? kQaoaAdjP : kLowP;
  constexpr int RB = 8 + (P == 1 ? 0 : (P == 2 ? 1 : (P == 4 ? 2 : 3)));  // log2 of a row's amplitudes
  static_assert(kLB * P == (1 << RB), "row size");
  __shared__ uint64_t tab[W ? 8 : 1][W ? 256 : 1];
  __shared__ double wt[W ? 8 : 1][W ? 256 : 1];
  __shared__ Cx<T> ph[W ? 1 : 65];
  uint64_t elo[P];
  if constexpr (W) {
    build_tables(tab, cs.inc);
    build_weight_tables(wt, cs);
  } else {
    for (unsigned c = threadIdx.x; c <= cs.ne; c += kLB) ph[c] = Cx<T>{(T)cs.ph_re[c], (T)cs.ph_im[c]};
#pragma unroll
    for (int j = 0; j < P; ++j) {
      const unsigned xl = row_index<P>(threadIdx.x, j);
      uint64_t e = 0;
#pragma unroll
      for (int b = 0; b < RB; ++b)
        if ((xl >> b) & 1) e ^= cs.inc.c[b];
      elo[j] = e;
    }
  }
  __syncthreads();
  const int lane = threadIdx.x & 63;
  double acc = 0, accg = 0;
  const uint64_t rows = N >> RB, G = gridDim.x;
  const uint64_t lb = (G & 7) ? blockIdx.x : xcd_block(blockIdx.x, G);
  const uint64_t chunk = (rows + G - 1) / G;
  const uint64_t b0 = lb * chunk, b1 = (b0 + chunk < rows) ? b0 + chunk : rows;
  for (uint64_t rb = b0; rb < b1; ++rb) {
    const uint64_t r0 = rb << RB;
    uint64_t ehi = 0;
    if constexpr (!W)
      for (uint64_t r = rb; r; r &= r - 1) ehi ^= cs.inc.c[RB + __builtin_ctzll(r)];
    // Phase factor and cut of in-row amplitude j (forward: applied on load; adjoint: the
    // phase term and the undo after the mixer, recomputed there so that they hold no
    // registers across the shuffles).
    auto phase_of = [&](int j, double* cut) -> Cx<T> {
      if constexpr (!W) {
        const int ic = __popcll(ehi ^ elo[j]);
        *cut = (double)ic;
        return ph[ic];
      } else {
        int ic;
        *cut = cut_value(tab, wt, cs, r0 + row_index<P>(threadIdx.x, j), &ic);
        double sn, co;
        sincos(cs.gamma * *cut, &sn, &co);
        return Cx<T>{(T)co, (T)-sn};
      }
    };
    V x[P], y[P];
#pragma unroll
    for (int j = 0; j < P; ++j) {
      x[j] = ldg(&psi[r0 + row_index<P>(threadIdx.x, j)]);
      if (ADJ) y[j] = ldg(&lam[r0 + row_index<P>(threadIdx.x, j)]);
      if (!ADJ) {
        double cut;
        x[j] = cmul(phase_of(j, &cut), x[j]);
      }
    }
#pragma unroll 1
    for (int q = 0; q < kLowQ; ++q) {
      const int msk = 1 << q;
      const bool hi = (lane >> q) & 1;
#pragma unroll
      for (int j = 0; j < P; ++j) {
        const V xp = shfl_v(x[j], msk);
        const V x0 = hi ? xp : x[j], x1 = hi ? x[j] : xp;
        if (ADJ) {
          acc += gen_term<(F & kFormGX) != 0>(y[j], g, x0, x1, hi);
          const V yp = shfl_v(y[j], msk);
          const V y0 = hi ? yp : y[j], y1 = hi ? y[j] : yp;
          y[j] = (F & kFormRX) ? urow<F>(u, y0, y1, hi) : (hi ? mat_row(u.m10, y0, u.m11, y1) : mat_row(u.m00, y0, u.m01, y1));
        }
        x[j] = (F & kFormRX) ? urow<F>(u, x0, x1, hi) : (hi ? mat_row(u.m10, x0, u.m11, x1) : mat_row(u.m00, x0, u.m01, x1));
      }
    }
    reg_positions<T, ADJ, P, F>(x, y, u, g, acc);
#pragma unroll
    for (int j = 0; j < P; ++j) {
      if (ADJ) {
        double cut;
        const Cx<T> f = phase_of(j, &cut);
        accg += cut * ((double)y[j].x * x[j].y - (double)y[j].y * x[j].x);
        x[j] = cmul(f, x[j]);
        y[j] = cmul(f, y[j]);
        stg(y[j], &lam[r0 + row_index<P>(threadIdx.x, j)]);
      }
      stg(x[j], &psi[r0 + row_index<P>(threadIdx.x, j)]);
    }
  }
  if (ADJ) {
    wg_sum_store(acc, part);
    __syncthreads();
    wg_sum_store(accg, part + G);
  }
}

// Register pass over M positions ins (all >= 6): each thread holds the 2^M amplitudes of a
// group.  The generator terms of a qubit are taken before any of this pass's gates.
// TILED: one group per thread, grid = items / kLB workgroups (items a power of two >=
// kRegTileMin), one partial per workgroup at part[block_id()]; otherwise a grid-stride loop
// over one contiguous chunk of groups per workgroup.
template <typename T, bool ADJ, int M, bool TILED, int F>
__global__ __launch_bounds__(kLB) void k_layer_reg(typename Vec<T>::t* __restrict__ psi, typename Vec<T>::t* __restrict__ lam,
                                                   uint64_t items, Ins<M> ins, Mat2<T> u, Mat2<double> g,
                                                   double* __restrict__ part)
{
  using V = typename Vec<T>::t;
  constexpr int D = 1 << M;
  uint64_t off[D];
#pragma unroll
  for (int j = 0; j < D; ++j) {
    uint64_t o = 0;
#pragma unroll
    for (int b = 0; b < M; ++b)
      if ((j >> b) & 1) o |= uint64_t(1) << ins.p[b];
    off[j] = o;
  }
  double acc = 0;
  // Persistent form: one contiguous chunk of groups per workgroup, chunks XCD-contiguous
  // (workgroups are dispatched round-robin over the 8 XCDs).  Tiled form: one group.
  uint64_t k0, k1, step;
  if (TILED) {
    k0 = tiled_item<kLB, 1, kRegSub, kSubLog>(xcd_block(block_id(), grid_blocks()), threadIdx.x);
    k1 = k0 + 1;
    step = 1;
  } else {
    const uint64_t G = gridDim.x;
    const uint64_t lb = (G & 7) ? blockIdx.x : xcd_block(blockIdx.x, G);
    const uint64_t chunk = (items + G - 1) / G;
    k0 = lb * chunk + threadIdx.x;
    k1 = (lb * chunk + chunk < items) ? lb * chunk + chunk : items;
    step = kLB;
  }
  for (uint64_t k = k0; k < k1; k += step) {
    const uint64_t base = item_base(k, ins);
    V x[D], y[D];
#pragma unroll
    for (int j = 0; j < D; ++j) {
      x[j] = ldg(&psi[base | off[j]]);
      if (ADJ) y[j] = ldg(&lam[base | off[j]]);
    }
    if (ADJ) {
#pragma unroll
      for (int b = 0; b < M; ++b)
#pragma unroll
        for (int j = 0; j < D; ++j) {
          if ((j >> b) & 1) continue;
          const int j1 = j | (1 << b);
          acc += gen_term<(F & kFormGX) != 0>(y[j], g, x[j], x[j1], false) +
                 gen_term<(F & kFormGX) != 0>(y[j1], g, x[j], x[j1], true);
        }
    }
#pragma unroll
    for (int b = 0; b < M; ++b)
#pragma unroll
      for (int j = 0; j < D; ++j) {
        if ((j >> b) & 1) continue;
        const int j1 = j | (1 << b);
        const V a = x[j], c = x[j1];
        x[j] = urow<F>(u, a, c, false);
        x[j1] = urow<F>(u, a, c, true);
        if (ADJ) {
          const V ya = y[j], yc = y[j1];
          y[j] = urow<F>(u, ya, yc, false);
          y[j1] = urow<F>(u, ya, yc, true);
        }
      }
#pragma unroll
    for (int j = 0; j < D; ++j) {
      stg(x[j], &psi[base | off[j]]);
      if (ADJ) stg(y[j], &lam[base | off[j]]);
    }
  }
  if (ADJ) wg_sum_store(acc, part, TILED ? block_id() : blockIdx.x);
}

unsigned grid_for_work(uint64_t work_items)
{
  return (unsigned)std::max<uint64_t>(1, std::min<uint64_t>((work_items + kLB - 1) / kLB, kMaxGrid));
}

// Sums of the per-workgroup partials of one adjoint pass, in a fixed order on the device:
// nsum groups of G partials stored one after the other; the sum of group i is added to
// *out[i].  The partial buffer must hold nsum * (G + 1) doubles.
int device_sums(uint64_t G, int nsum, double* const* out)
{
  for (int i = 0; i < nsum; ++i)
    k_sum_part<<<1, kLB, 0, iqsi::stream()>>>(g_lw.part + i * G, G, g_lw.part + nsum * G + i);
  IQSI_TRY(hipGetLastError());
  double s[2] = {0, 0};
  IQSI_TRY(hipMemcpyAsync(s, g_lw.part + nsum * G, sizeof(double) * nsum, hipMemcpyDeviceToHost, iqsi::stream()));
  IQSI_TRY(hipStreamSynchronize(iqsi::stream()));
  for (int i = 0; i < nsum; ++i) *out[i] += s[i];
  return IQSH_OK;
}

// Sum the partials of one adjoint pass (in order) into *acc.
int finish_pass(unsigned grid, double* acc)
{
  IQSI_TRY(hipGetLastError());
  g_lw.host.resize(grid);
  IQSI_TRY(hipMemcpyAsync(g_lw.host.data(), g_lw.part, sizeof(double) * grid, hipMemcpyDeviceToHost, iqsi::stream()));
  IQSI_TRY(hipStreamSynchronize(iqsi::stream()));
  double s = 0;
  for (unsigned b = 0; b < grid; ++b) s += g_lw.host[b];
  *acc += s;
  return IQSH_OK;
}

template <typename T, bool ADJ, int M, int F>
int reg_pass(iqsh_handle h, iqsh_handle lam, unsigned p0, const Mat2<T>& u, const Mat2<double>& g, double* acc)
{
  using V = typename Vec<T>::t;
  Ins<M> ins;
  for (int b = 0; b < M; ++b) ins.p[b] = p0 + b;
  const uint64_t items = h->size() >> M;
  V* lp = ADJ ? (V*)lam->d : nullptr;
  // The complex128 adjoint pass over 4 positions holds 32 amplitudes per thread; in its
  // general form it keeps the grid-stride loop (28 qubits: 3.21 ms against 3.29 ms tiled),
  // in the RX/X form the tiled one (3.05 ms against 3.23).
  const bool tiled = items >= kRegTileMin && !(ADJ && M == 4 && sizeof(T) == 8 && F != (kFormGX | kFormRX));
  if (tiled) {
    const uint64_t G = items / kLB;
    if (ADJ)
      if (int rc = part_buf(G + 1)) return rc;
    const dim3 grid((unsigned)std::min<uint64_t>(G, 65536), (unsigned)((G + 65535) / 65536), 1);
    k_layer_reg<T, ADJ, M, true, F><<<grid, kLB, 0, iqsi::stream()>>>((V*)h->d, lp, items, ins, u, g, g_lw.part);
    IQSI_TRY(hipGetLastError());
    if (!ADJ) return IQSH_OK;
    double* out[1] = {acc};
    return device_sums(G, 1, out);
  }
  const unsigned grid = grid_for_work(items);
  k_layer_reg<T, ADJ, M, false, F><<<grid, kLB, 0, iqsi::stream()>>>((V*)h->d, lp, items, ins, u, g, g_lw.part);
  if (!ADJ) {
    IQSI_TRY(hipGetLastError());
    return IQSH_OK;
  }
  return finish_pass(grid, acc);
}

template <typename T, bool ADJ, int F>
int do_layer(iqsh_handle h, iqsh_handle lam, const double* um, const double* gm, const CutSpec* cs, double* grad,
             double* grad_phase)
{
  using V = typename Vec<T>::t;
  const Mat2<T> u{{(T)um[0], (T)um[1]}, {(T)um[2], (T)um[3]}, {(T)um[4], (T)um[5]}, {(T)um[6], (T)um[7]}};
  Mat2<double> g{{0, 0}, {0, 0}, {0, 0}, {0, 0}};
  if (ADJ) g = Mat2<double>{{gm[0], gm[1]}, {gm[2], gm[3]}, {gm[4], gm[5]}, {gm[6], gm[7]}};
  double acc = 0, accg = 0;
  const unsigned n = h->n;
  // Amplitudes per lane of this lane pass and the positions it covers (0-5 by shuffles,
  // log2 P more in registers).
  const int lowp = (ADJ && cs) ? kQaoaAdjP : kLowP;
  const unsigned lowq = kLowQ + (lowp == 1 ? 0 : (lowp == 2 ? 1 : (lowp == 4 ? 2 : 3)));
  // Forward: the lane pass first (it carries the phase layer, applied before the mixer);
  // adjoint: last (the phase is undone after the whole mixer).
  auto low_pass = [&]() -> int {
    // kLowRowsFwd or kLowRowsAdj rows per workgroup.  The weighted QAOA form builds byte tables of weight sums
    // per workgroup: a persistent grid of <= kMaxGrid workgroups for it.
    const uint64_t rows = h->size() / (kLB * (uint64_t)lowp);
    const unsigned grid = (cs && cs->weighted) ? (unsigned)std::min<uint64_t>(rows, kMaxGrid)
                                               : (unsigned)std::min<uint64_t>(std::max<uint64_t>(1, rows / (ADJ ? kLowRowsAdj : kLowRowsFwd)), kLowMaxGrid);
    if (ADJ)
      if (int rc = part_buf(2 * (uint64_t)grid + 2)) return rc;
    V* lp = ADJ ? (V*)lam->d : nullptr;
    if (cs && cs->weighted)
      k_qaoa_low<T, ADJ, true, F><<<grid, kLB, 0, iqsi::stream()>>>((V*)h->d, lp, h->size(), u, g, *cs, g_lw.part);
    else if (cs)
      k_qaoa_low<T, ADJ, false, F><<<grid, kLB, 0, iqsi::stream()>>>((V*)h->d, lp, h->size(), u, g, *cs, g_lw.part);
    else
      k_layer_low<T, ADJ, F><<<grid, kLB, 0, iqsi::stream()>>>((V*)h->d, ADJ ? (V*)lam->d : nullptr, h->size(), u, g,
                                                             g_lw.part);
    IQSI_TRY(hipGetLastError());
    if (!ADJ) return IQSH_OK;
    // The phase sums (QAOA form) follow the rotation sums in the partial buffer.
    double* out[2] = {&acc, &accg};
    return device_sums(grid, cs ? 2 : 1, out);
  };
  if (!ADJ)
    if (int rc = low_pass()) return rc;
  // Positions above the lane pass in ceil(rest / 4) passes of as equal a size as possible
  // (28 qubits, lane pass over 0-7: 4, 4, 4, 4, 4 -- a pass of two positions streams ~20%
  // slower).
  const unsigned rest = n > lowq ? n - lowq : 0;
  const unsigned perq = ADJ ? kRegQ : kRegQF;
  const unsigned npass = (rest + perq - 1) / perq;
  for (unsigned i = 0, p = lowq; i < npass; ++i) {
    const unsigned m = rest / npass + (i < rest % npass ? 1 : 0);
    int rc = IQSH_OK;
    if (m == 5) rc = reg_pass<T, false, 5, ADJ ? 0 : F>(h, lam, p, u, g, &acc);
    else if (m == 4) rc = reg_pass<T, ADJ, 4, F>(h, lam, p, u, g, &acc);
    else if (m == 3) rc = reg_pass<T, ADJ, 3, F>(h, lam, p, u, g, &acc);
    else if (m == 2) rc = reg_pass<T, ADJ, 2, F>(h, lam, p, u, g, &acc);
    else rc = reg_pass<T, ADJ, 1, F>(h, lam, p, u, g, &acc);
    if (rc) return rc;
    p += m;
  }
  if (ADJ)
    if (int rc = low_pass()) return rc;
  if (ADJ) *grad = 2.0 * acc;
  if (ADJ && grad_phase) *grad_phase = 2.0 * accg;
  return IQSH_OK;
}

// Backward step through a layer: the kernels with G = X when the generator is exactly X,
// and with the RX-form product as well when uinv has a real diagonal and an imaginary
// off-diagonal (the QAOA mixer's inverse).
template <typename T>
int adjoint_layer_t(iqsh_handle psi, iqsh_handle lam, const double* uinv, const double* G, const CutSpec* cs,
                    double* grad, double* grad_phase)
{
  static const double X[8] = {0, 0, 1, 0, 1, 0, 0, 0};
  const bool gx = std::equal(G, G + 8, X);
  const bool rx = uinv[1] == 0 && uinv[2] == 0 && uinv[4] == 0 && uinv[7] == 0;
  if (gx && rx) return do_layer<T, true, kFormGX | kFormRX>(psi, lam, uinv, G, cs, grad, grad_phase);
  if (gx) return do_layer<T, true, kFormGX>(psi, lam, uinv, G, cs, grad, grad_phase);
  return do_layer<T, true, 0>(psi, lam, uinv, G, cs, grad, grad_phase);
}

// Forward layer: the RX-form kernels when m has a real diagonal and an imaginary off-diagonal.
template <typename T>
int forward_layer_t(iqsh_handle h, const double* m, const CutSpec* cs)
{
  const bool rx = m[1] == 0 && m[2] == 0 && m[4] == 0 && m[7] == 0;
  return rx ? do_layer<T, false, kFormRX>(h, nullptr, m, nullptr, cs, nullptr, nullptr)
            : do_layer<T, false, 0>(h, nullptr, m, nullptr, cs, nullptr, nullptr);
}

int adjoint_layer(iqsh_handle psi, iqsh_handle lam, const double* uinv, const double* G, const CutSpec* cs,
                  double* grad, double* grad_phase)
{
  return psi->dtype == IQSH_C128 ? adjoint_layer_t<double>(psi, lam, uinv, G, cs, grad, grad_phase)
                                 : adjoint_layer_t<float>(psi, lam, uinv, G, cs, grad, grad_phase);
}

// The lane pass needs whole rows of kLB * kLowP amplitudes with lanes spanning >= 64.
constexpr unsigned kMinLayerQubits = 10;
static_assert(kMinLayerQubits >= 8 + 2 && kLB * kLowP <= (1 << kMinLayerQubits) && kLB * kQaoaAdjP <= (1 << kMinLayerQubits),
              "a lane pass row must fit the smallest register");

}  // namespace

extern "C" {

int iqsh_apply_1q_all(iqsh_handle h, const double m[8])
{
  if (!h || !h->d) return iqsi::fail(IQSH_ERR_ARG, "null or released state handle");
  if (!m) return iqsi::fail(IQSH_ERR_ARG, "null matrix");
  if (h->n < kMinLayerQubits) return iqsi::fail(IQSH_ERR_NOTSUP, "layer kernels need >= %u qubits", kMinLayerQubits);
  std::lock_guard<std::mutex> lk(g_lm);
  return h->dtype == IQSH_C128 ? forward_layer_t<double>(h, m, nullptr) : forward_layer_t<float>(h, m, nullptr);
}

int iqsh_qaoa_layer(iqsh_handle h, unsigned ne, const unsigned* u, const unsigned* v, const double* w, double gamma,
                    const double mixer[8])
{
  if (!h || !h->d) return iqsi::fail(IQSH_ERR_ARG, "null or released state handle");
  if (!mixer || (ne && (!u || !v))) return iqsi::fail(IQSH_ERR_ARG, "null argument");
  if (h->n < kMinLayerQubits) return iqsi::fail(IQSH_ERR_NOTSUP, "layer kernels need >= %u qubits", kMinLayerQubits);
  CutSpec cs;
  if (int rc = iqsi::build_cut_spec(h, ne, u, v, w, gamma, &cs)) return rc;
  std::lock_guard<std::mutex> lk(g_lm);
  return h->dtype == IQSH_C128 ? forward_layer_t<double>(h, mixer, &cs) : forward_layer_t<float>(h, mixer, &cs);
}

int iqsh_qaoa_layer_adjoint(iqsh_handle psi, iqsh_handle lam, unsigned ne, const unsigned* u, const unsigned* v,
                            const double* w, double gamma, const double mixer_inv[8], const double* G,
                            double* dgamma, double* dmixer)
{
  if (!psi || !psi->d || !lam || !lam->d) return iqsi::fail(IQSH_ERR_ARG, "null or released state handle");
  if (psi->n != lam->n || psi->dtype != lam->dtype) return iqsi::fail(IQSH_ERR_ARG, "registers differ in size or precision");
  if (!mixer_inv || !G || !dgamma || !dmixer || (ne && (!u || !v))) return iqsi::fail(IQSH_ERR_ARG, "null argument");
  if (psi->n < kMinLayerQubits) return iqsi::fail(IQSH_ERR_NOTSUP, "layer kernels need >= %u qubits", kMinLayerQubits);
  CutSpec cs;
  // Un-applying exp(-i gamma C) is applying exp(-i (-gamma) C).
  if (int rc = iqsi::build_cut_spec(psi, ne, u, v, w, -gamma, &cs)) return rc;
  std::lock_guard<std::mutex> lk(g_lm);
  if (int rc = part_buf()) return rc;
  return adjoint_layer(psi, lam, mixer_inv, G, &cs, dmixer, dgamma);
}

int iqsh_adjoint_1q_all(iqsh_handle psi, iqsh_handle lam, const double Uinv[8], const double* G, double* grad)
{
  if (!psi || !psi->d || !lam || !lam->d) return iqsi::fail(IQSH_ERR_ARG, "null or released state handle");
  if (psi->n != lam->n || psi->dtype != lam->dtype) return iqsi::fail(IQSH_ERR_ARG, "registers differ in size or precision");
  if (!Uinv || !G || !grad) return iqsi::fail(IQSH_ERR_ARG, "null argument");
  if (psi->n < kMinLayerQubits) return iqsi::fail(IQSH_ERR_NOTSUP, "layer kernels need >= %u qubits", kMinLayerQubits);
  std::lock_guard<std::mutex> lk(g_lm);
  if (int rc = part_buf()) return rc;
  return adjoint_layer(psi, lam, Uinv, G, nullptr, grad, nullptr);
}

}  // extern "C"
