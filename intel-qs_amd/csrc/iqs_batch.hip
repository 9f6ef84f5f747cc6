// Batches of hardware-efficient-ansatz circuits: forward pass, cost <sum_q Z_q> and the
// adjoint gradient of every angle, for many small registers in one persistent launch.
//
// Circuit of instance b (n qubits, d layers), from |0...0>:
//   layer l:  U_q = RZ(phi[l][q]) RY(theta[l][q]) on every qubit q, then CZ(i, i+1), i < n-1.
// Adjoint method: psi = final state, lam = O psi with O = sum_q Z_q (diagonal, o(x) =
// n - 2 popcount(x)); walking the gates backwards, the rotation exp(-i a P/2) contributes
// dC/da = Im <lam|P|psi> at the point just after it, then it is undone on both vectors.
// The kernel applies a layer as U_l = CZ_l . prod_q RZ_q(phi_lq) . prod_q RY_q(theta_lq):
// real 2x2 rotations per pair for the RYs, and per register phase one diagonal per
// amplitude holding RZ factors and CZ signs (see Tile).  So
//   dC/dtheta = Im <lam|Y_q|psi>  just after the RY (one product per pair),
//   dC/dphi   = Im <lam|Z_q|psi>  at the diagonal that holds the RZ (one per amplitude).
//
// Work decomposition (M = tile qubits, R = amplitudes per thread, NT = 2^M / R threads):
//  * n <= M: the whole instance is one task.  Its state lives in registers (padded to 2^M
//    amplitudes; padding stays zero) and LDS; nothing touches HBM but the results.
//  * n > M (n <= 2M): tile passes.  An A tile holds the low M qubits of one value of the
//    high ones; a B tile holds the n-M high qubits plus the c = 2M-n lowest (so each row
//    of the tile is a contiguous run of 2^c amplitudes).  Pass j gates the qubits local to
//    its locality L_j (A for even j, B for odd j):
//        R_{j-1} on L_j,  CZ ladder of layer j-1,  R_j on L_j
//    so one pass over the state applies about one layer: layer l's rotations are split
//    between passes l and l+1 (they commute), and CZ of layer l sits between them.
//    Forward passes j = 0..d-1, a turnaround pass j = d (forward, cost, lam, backward),
//    backward passes j = d-1..0 on psi and lam: 2d+1 steps of 2^(n-M) tile tasks.  The RZs
//    of the qubits a pass rotates last (its R_j) move to the next pass's first diagonal: an
//    RZ commutes with everything between its qubit's RY of the same layer and the next.
//  * Inside a task the registers cycle through phases: phase k holds a run of log2(R)
//    consecutive local bits (runs overlap by one bit), gates on those bits are register-
//    local, and the layout changes through one LDS transpose per phase.  A CZ pair (i,i+1)
//    is applied in the phase where its upper gated qubit first appears, after the RY_{j-1}
//    of both and before their RY_j (a CZ is diagonal: it only needs the global index).
//  * Scheduling: a global queue of tasks in a host-chosen order; a workgroup takes the
//    next task with one atomic, and a tile task of step s waits until all tiles of step
//    s-1 of its instance have arrived.  Every task waits only on tasks earlier in the
//    queue, which were taken by running workgroups, so the queue always drains; waits are
//    bounded and set an abort flag instead of hanging.
//  * Gradient sums are FP64 (products of FP32 amplitudes are formed in FP32): per gate a DPP
//    wave sum, per tile a fixed-order sum over waves, written as per-tile partials and
//    reduced in tile order by an epilogue kernel.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "../../include/iqs_hip.h"
#include "iqs_internal.hpp"
#include "iqs_kernels.hpp"

using iqsi::fail;

namespace {

template <typename T> using V = typename iqsk::Vec<T>::t;

constexpr uint32_t kGrad = IQSH_HEA_GRAD, kKeep = IQSH_HEA_KEEP_STATE;

struct HeaInst {
  uint32_t n, depth;
  uint32_t whole;   // 1: n <= tile qubits, the instance is a single task
  uint32_t ntl;     // tiled: log2(tiles) = n - tile qubits
  uint64_t psi, lam;  // amplitude offsets of the two vectors in the arena
  uint64_t par;       // first (theta, phi) pair of the instance
  uint64_t part;      // tiled: gradient partials [depth][n][2 (theta, phi)][tiles]
  uint64_t cpart;     // tiled: cost partials [tiles]
  uint32_t arrive;    // tiled: arrival counter of step 0 (one counter per step)
  uint32_t pad;
};

struct HeaTask {
  uint32_t inst;
  uint16_t step;
  uint16_t tile;
};

// One task queue per XCD (see k_hea).
constexpr int kQueues = 8;
constexpr int kCtlAbort = kQueues;       // ctl[0..7]: queue heads, ctl[8]: abort flag
constexpr int kCtlArrive = kQueues + 1;  // ctl[9 + i]: arrival counter i

template <typename T> struct HeaArgs {
  const HeaInst* inst;
  const HeaTask* task;  // the queues, back to back
  const uint32_t* qoff;  // queue q = task[qoff[q] .. qoff[q + 1])
  uint32_t flags;
  uint32_t* ctl;
  V<T>* arena;
  const V<T>* mats;  // 4 entries per pair: (cos, sin) of theta/2 and of phi/2, then 2 unused
  double* grad;
  double* cost;
  double* part;
  double* cpart;
  uint32_t spin_limit;
  uint32_t wait_extra;  // added to every dependency target: 0, or 1 to make waits unsatisfiable
};

constexpr int ilog2(int v) { return v <= 1 ? 0 : 1 + ilog2(v / 2); }

__device__ __forceinline__ uint32_t bit_range(int lo, int hi)
{
  if (hi <= lo) return 0u;
  const uint32_t top = hi >= 32 ? 0xffffffffu : ((1u << hi) - 1u);
  return top & ~((1u << lo) - 1u);
}

// Local index of register k of thread tid when the registers hold local bits [a, a+LR).
template <int LR>
__device__ __forceinline__ uint32_t lidx(uint32_t tid, int k, int a)
{
  return (tid & ((1u << a) - 1u)) | ((uint32_t)k << a) | ((tid >> a) << (a + LR));
}

// LDS slot of local index u: the low LR bits are XOR-ed with the next LR bits so that
// every layout's 16 consecutive lanes hit distinct banks.
template <int LR>
__device__ __forceinline__ uint32_t phys(uint32_t u)
{
  return u ^ ((u >> LR) & ((1u << LR) - 1u));
}

// Pass geometry: gated local bits [la, lb), the tile's local<->global maps and the phases.
template <int M, int LR>
struct Geo {
  int n, kind, la, lb, c;  // kind 0: whole instance, 1: A tile, 2: B tile
  uint32_t t;
  int ne0, K;

  __device__ void init(int n_, int kind_, int la_, int lb_, int c_, uint32_t t_)
  {
    n = n_;
    kind = kind_;
    la = la_;
    lb = lb_;
    c = c_;
    t = t_;
    const int a0 = min(la, M - LR);
    ne0 = min(a0 + LR, lb);
    K = 1 + (lb > ne0 ? (lb - ne0 + LR - 2) / (LR - 1) : 0);
  }
  __device__ int ne(int k) const { return min(ne0 + k * (LR - 1), lb); }
  __device__ int nb(int k) const { return k ? ne(k - 1) : la; }
  __device__ int a(int k) const { return min(k ? nb(k) - 1 : la, M - LR); }
  __device__ int gq(int lbit) const { return (kind == 2 && lbit >= c) ? lbit - c + M : lbit; }
  __device__ uint32_t gidx(uint32_t u) const
  {
    if (kind == 1) return u | (t << M);
    if (kind == 2) return (u & ((1u << c) - 1u)) | (t << c) | ((u >> c) << M);
    return u;
  }
  // CZ pairs (bit i = CZ(i, i+1)) applied in phase k.
  __device__ uint32_t czmask(int k) const
  {
    const int g_ne = gq(ne(k) - 1) + 1;
    uint32_t m = bit_range(k == 0 ? 0 : gq(nb(k)) - 1, g_ne - 1);
    if (k == K - 1) m |= bit_range(gq(lb - 1), n - 1);
    return m & bit_range(0, n - 1);
  }
};

// Coherence of tile passes.  All tiles of one register are produced and consumed by
// workgroups of the same XCD (one queue per XCD, chosen by the XCC_ID hardware register):
// a producer waits for its stores to complete before bumping the arrival counter, and a
// consumer performs an agent-scope acquire (buffer_inv sc1) before loading.  Measured with
// a message-passing test on MI355X (1024 workgroups, 20 rounds per case): same XCD with the
// acquire, 0 stale reads; same XCD with only an L1 invalidate (buffer_inv sc0), every read
// stale; different XCDs without an L2 write-back, every read stale.  Sharing tiles across
// XCDs would need the release's L2 write-back (buffer_wbl2) around every task, which made
// the passes 4x slower.
__device__ __forceinline__ void acquire_tiles() { __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent"); }

// One DPP row operation on a double (both halves moved alike); rows outside ROWMASK read 0.
template <int CTRL, int ROWMASK>
__device__ __forceinline__ double dpp_move(double v)
{
  const uint64_t b = __builtin_bit_cast(uint64_t, v);
  const uint32_t lo = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)b, CTRL, ROWMASK, 0xf, false);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)(b >> 32), CTRL, ROWMASK, 0xf, false);
  return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}

// Sum over the 64 lanes of a wavefront, complete in lane 63: quad swaps, half-row and row
// mirrors, then row broadcasts (lane 15 into rows 1 and 3, lane 31 into rows 2 and 3).
// Fixed order, so bitwise reproducible; VALU moves instead of LDS-routed shuffles.
__device__ __forceinline__ double wave_sum63(double v)
{
  v += dpp_move<0xB1, 0xf>(v);   // quad_perm [1,0,3,2]
  v += dpp_move<0x4E, 0xf>(v);   // quad_perm [2,3,0,1]
  v += dpp_move<0x141, 0xf>(v);  // row_half_mirror
  v += dpp_move<0x140, 0xf>(v);  // row_mirror
  v += dpp_move<0x142, 0xa>(v);  // row_bcast:15
  v += dpp_move<0x143, 0xc>(v);  // row_bcast:31
  return v;
}

template <typename C>
__device__ __forceinline__ C cx_mul(const C& a, const C& b)
{
  C r;
  r.x = a.x * b.x - a.y * b.y;
  r.y = a.x * b.y + a.y * b.x;
  return r;
}

template <typename C>
__device__ __forceinline__ C cx_mulc(const C& a, const C& b)  // a * conj(b)
{
  C r;
  r.x = a.x * b.x + a.y * b.y;
  r.y = a.y * b.x - a.x * b.y;
  return r;
}

// Layer split.  U_l = CZ_l . prod_q RZ_q(phi_lq) . prod_q RY_q(theta_lq): the RYs are real
// 2x2 rotations (half the arithmetic of the complex U = RZ RY per pair), and every RZ and CZ
// factor of a register phase is one diagonal applied per amplitude.  An RZ commutes with
// everything between its qubit's RY of the same layer and that qubit's RY of the next, so
// the RZs of the qubits a pass rotates last (its RY_j) are deferred to the next pass's first
// diagonal.  Gradients: dC/dtheta = Im<lam|Y_q|psi> just after the RY (one sum per pair),
// dC/dphi = Im<lam|Z_q|psi> at the diagonal that holds the RZ (one product per amplitude,
// signed by the qubit's bit), both from partial sums over a thread's registers.
// mats: 4 entries per (layer, qubit): (cos theta/2, sin theta/2), (cos phi/2, sin phi/2), 0, 0.
template <typename T, int M, int R>
struct Tile {
  static constexpr int LR = ilog2(R);
  static constexpr int NT = (1 << M) / R;
  static constexpr int NW = NT / 64;
  static constexpr int KMAX = M;       // register phases of one sweep (at most)
  static constexpr int SLOTS = 4 * M;  // dtheta of two layers' gated bits, dphi of <= 2M qubits
  static_assert(R >= 4 && (R & (R - 1)) == 0, "R must be a power of two >= 4");
  static_assert(NT >= 64 && NT <= 1024, "workgroup of 1..16 wavefronts");
  static_assert(KMAX * R <= NT, "one table entry per thread");
  using G = Geo<M, LR>;
  using C = V<T>;

  static __device__ void relayout(C* v, C* lds, int from, int to)
  {
    const uint32_t tid = threadIdx.x;
#pragma unroll
    for (int k = 0; k < R; ++k) lds[phys<LR>(lidx<LR>(tid, k, from))] = v[k];
    __syncthreads();
#pragma unroll
    for (int k = 0; k < R; ++k) v[k] = lds[phys<LR>(lidx<LR>(tid, k, to))];
    __syncthreads();
  }

  // RY_q (real) on the register bits in rmask; `row` = matrices of qubit 0 of the layer.
  static __device__ void ry_fwd(C* x, uint32_t rmask, const C* row, int a, const G& g)
  {
#pragma unroll
    for (int j = 0; j < LR; ++j) {
      if ((rmask >> j) & 1u) {
        const C cs = row[4 * g.gq(a + j)];
        const T c = cs.x, sn = cs.y;
#pragma unroll
        for (int k = 0; k < R; ++k) {
          if ((k >> j) & 1) continue;
          const int k1 = k | (1 << j);
          const C p0 = x[k], p1 = x[k1];
          x[k].x = c * p0.x - sn * p1.x;
          x[k].y = c * p0.y - sn * p1.y;
          x[k1].x = sn * p0.x + c * p1.x;
          x[k1].y = sn * p0.y + c * p1.y;
        }
      }
    }
  }

  // Backward through RY_q for psi (x) and lam (y): Im<lam|Y_q|psi> (summed over this thread's
  // pairs, then the wavefront) into redw[slot0 + lbit - la], then RY_q^T on both.
  static __device__ void ry_bwd(C* x, C* y, uint32_t rmask, const C* row, int a, const G& g, double* redw, int slot0)
  {
    const bool lane63 = (threadIdx.x & 63) == 63;
#pragma unroll
    for (int j = 0; j < LR; ++j) {
      if ((rmask >> j) & 1u) {
        const C cs = row[4 * g.gq(a + j)];
        const T c = cs.x, sn = cs.y;
        T acc = 0;
#pragma unroll
        for (int k = 0; k < R; ++k) {
          if ((k >> j) & 1) continue;
          const int k1 = k | (1 << j);
          const C p0 = x[k], p1 = x[k1], l0 = y[k], l1 = y[k1];
          // Im(conj(l0) (-i p1) + conj(l1) (i p0)) = Re(conj(l1) p0) - Re(conj(l0) p1)
          acc += (l1.x * p0.x + l1.y * p0.y) - (l0.x * p1.x + l0.y * p1.y);
          x[k].x = c * p0.x + sn * p1.x;
          x[k].y = c * p0.y + sn * p1.y;
          x[k1].x = c * p1.x - sn * p0.x;
          x[k1].y = c * p1.y - sn * p0.y;
          y[k].x = c * l0.x + sn * l1.x;
          y[k].y = c * l0.y + sn * l1.y;
          y[k1].x = c * l1.x - sn * l0.x;
          y[k1].y = c * l1.y - sn * l0.y;
        }
        const double sum = wave_sum63((double)acc);
        if (lane63) redw[slot0 + a + j - g.la] = sum;
      }
    }
  }

  // Register bits of phase k whose RZ is in that phase's diagonal: the pre-rotated bits
  // [nb, ne), and in phase 0 the deferred qubits that are register bits there.
  static __device__ uint32_t zreg(const G& g, int k, bool defer_b)
  {
    const int a = g.a(k);
    uint32_t m = bit_range(g.nb(k) - a, g.ne(k) - a);
    // deferred qubits 0..M-1 of a B pass: local bits [0, c); in phase 0 those >= a are
    // register bits (a = M - LR < c when n < M + LR)
    if (k == 0 && defer_b) m |= bit_range(0, g.c - a);
    return m & ((1u << LR) - 1u);
  }

  // Diagonal tables of one sweep: zt[k][r] = product of the RZ factors of phase k's
  // register bits for register index r (bit 0 -> e^{-i phi/2}, bit 1 -> e^{+i phi/2}).
  // Barriers on both sides: no thread still reads the previous sweep's table.
  static __device__ void build_tables(C* zt, const G& g, const C* row, bool defer_b)
  {
    __syncthreads();
    const uint32_t i = threadIdx.x;
    if (i < (uint32_t)(g.K * R)) {
      const int k = i / R, r = i % R, a = g.a(k);
      const uint32_t m = zreg(g, k, defer_b);
      C f{1, 0};
      for (int j = 0; j < LR; ++j)
        if ((m >> j) & 1u) {
          const C h = row[4 * g.gq(a + j) + 1];
          f = cx_mul(f, C{h.x, ((r >> j) & 1) ? h.y : -h.y});
        }
      zt[i] = f;
    }
    __syncthreads();
  }

  // This thread's factor of the deferred RZs on bits that are not register bits in phase 0
  // (lane bits and tile bits).  defer: global qubits [q0, q1) of the previous pass.
  static __device__ C deferred_factor(const G& g, const C* row, int q0, int q1)
  {
    const int a = g.a(0);
    const uint32_t x0 = g.gidx(lidx<LR>(threadIdx.x, 0, a));  // register bits 0
    C f{1, 0};
    for (int q = q0; q < q1; ++q) {
      // register bit in phase 0: the table has it (only local bits [a, a + LR) of a B tile)
      if (g.kind == 2 && q < g.c && q >= a && q < a + LR) continue;
      const C h = row[4 * q + 1];
      f = cx_mul(f, C{h.x, ((x0 >> q) & 1u) ? h.y : -h.y});
    }
    return f;
  }

  // x *= diag (conj: x *= conj(diag), and y likewise when given); diag = CZ signs (czm) x
  // zt[k][r] (x f0 in phase 0).
  template <bool BOTH>
  static __device__ void diag_apply(C* x, C* y, const C* ztk, const C& f0, bool use_f0, uint32_t czm, int a,
                                    const G& g, bool conj)
  {
    const uint32_t tid = threadIdx.x;
#pragma unroll
    for (int k = 0; k < R; ++k) {
      C f = ztk[k];
      if (use_f0) f = cx_mul(f, f0);
      if (czm) {
        const uint32_t xg = g.gidx(lidx<LR>(tid, k, a));
        if (__popc(xg & (xg >> 1) & czm) & 1) f = C{-f.x, -f.y};
      }
      x[k] = conj ? cx_mulc(x[k], f) : cx_mul(x[k], f);
      if (BOTH) y[k] = conj ? cx_mulc(y[k], f) : cx_mul(y[k], f);
    }
  }

  // Im<lam|Z_q|psi> for the RZ qubits of phase k into redw[slot0 + q] (register bits of
  // zreg, and in phase 0 the deferred qubits [q0, q1) on lane or tile bits).
  static __device__ void dphi(const C* x, const C* y, int k, const G& g, bool defer_b, int q0, int q1, double* redw,
                              int slot0)
  {
    const bool lane63 = (threadIdx.x & 63) == 63;
    const int a = g.a(k);
    T tot = 0, hi[LR];  // hi[j]: registers with bit j set (Z = -1 there)
#pragma unroll
    for (int j = 0; j < LR; ++j) hi[j] = 0;
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const T t = y[r].x * x[r].y - y[r].y * x[r].x;  // Im(conj(lam) psi)
      tot += t;
#pragma unroll
      for (int j = 0; j < LR; ++j)
        if ((r >> j) & 1) hi[j] += t;
    }
    const uint32_t m = zreg(g, k, defer_b);
#pragma unroll
    for (int j = 0; j < LR; ++j) {
      if ((m >> j) & 1u) {
        const double sum = wave_sum63((double)(tot - 2 * hi[j]));
        if (lane63) redw[slot0 + g.gq(a + j)] = sum;
      }
    }
    if (k == 0 && q1 > q0) {
      const uint32_t x0 = g.gidx(lidx<LR>(threadIdx.x, 0, a));
      for (int q = q0; q < q1; ++q) {
        if (g.kind == 2 && q < g.c && q >= a && q < a + LR) continue;  // register bit: done above
        const double sum = wave_sum63((double)(((x0 >> q) & 1u) ? -tot : tot));
        if (lane63) redw[slot0 + q] = sum;
      }
    }
  }

  // Forward sweep: per phase, RY of layer row_pre on the pre bits, the diagonal (RZ of
  // row_pre on those bits and the deferred qubits [q0, q1), CZ pairs), RY of row_post.
  // A negative row skips that layer.  zt: tables built for row_pre.
  static __device__ void sweep_fwd(C* x, C* lds, int& cur, const G& g, const C* mats, int64_t row_pre,
                                   int64_t row_post, const C* zt, const C& f0, bool has_f0)
  {
    for (int k = 0; k < g.K; ++k) {
      const int nb = g.nb(k), ne = g.ne(k), a = g.a(k);
      if (cur != a) {
        relayout(x, lds, cur, a);
        cur = a;
      }
      if (row_pre >= 0) {
        ry_fwd(x, bit_range(nb - a, ne - a), mats + 4 * row_pre, a, g);
        diag_apply<false>(x, x, zt + k * R, f0, k == 0 && has_f0, g.czmask(k), a, g, false);
      }
      if (row_post >= 0) {
        const int lo = max(nb - 1, g.la), hi = (k == g.K - 1) ? g.lb : ne - 1;
        ry_fwd(x, bit_range(lo - a, hi - a), mats + 4 * row_post, a, g);
      }
    }
  }

  // Exact reverse of sweep_fwd on psi and lam with the gradient sums: dtheta of row_post at
  // slots [0, M), of row_pre at [M, 2M) (by local bit - la), dphi of row_pre at 2M + qubit.
  static __device__ void sweep_bwd(C* x, C* y, C* lds, int& cur, const G& g, const C* mats, int64_t row_pre,
                                   int64_t row_post, const C* zt, const C& f0, bool has_f0, bool defer_b, int q0,
                                   int q1, double* redw)
  {
    for (int k = g.K - 1; k >= 0; --k) {
      const int nb = g.nb(k), ne = g.ne(k), a = g.a(k);
      if (cur != a) {
        relayout(x, lds, cur, a);
        relayout(y, lds, cur, a);
        cur = a;
      }
      if (row_post >= 0) {
        const int lo = max(nb - 1, g.la), hi = (k == g.K - 1) ? g.lb : ne - 1;
        ry_bwd(x, y, bit_range(lo - a, hi - a), mats + 4 * row_post, a, g, redw, 0);
      }
      if (row_pre >= 0) {
        dphi(x, y, k, g, defer_b, q0, q1, redw, 2 * M);
        diag_apply<true>(x, y, zt + k * R, f0, k == 0 && has_f0, g.czmask(k), a, g, true);
        ry_bwd(x, y, bit_range(nb - a, ne - a), mats + 4 * row_pre, a, g, redw, M);
      }
    }
  }

  // Cost sum o(x)|psi_x|^2 of this workgroup's amplitudes (FP64, fixed order) and, if
  // lam_out, lam = o psi.  Returns the total on thread 0.
  static __device__ double cost_and_lam(C* x, C* y, bool lam_out, int cur, const G& g, double* redc)
  {
    const uint32_t tid = threadIdx.x;
    double cs = 0;
#pragma unroll
    for (int k = 0; k < R; ++k) {
      const uint32_t xg = g.gidx(lidx<LR>(tid, k, cur));
      const int o = g.n - 2 * __popc(xg);
      const double re = x[k].x, im = x[k].y;
      cs += (double)o * (re * re + im * im);
      if (lam_out) y[k] = x[k] * (T)o;
    }
    cs = wave_sum63(cs);
    if ((tid & 63) == 63) redc[tid >> 6] = cs;
    __syncthreads();
    double tot = 0;
    if (tid == 0)
      for (int w = 0; w < NW; ++w) tot += redc[w];
    return tot;
  }

  static __device__ void run_whole(const HeaArgs<T>& A, const HeaInst& I, uint32_t b, C* lds, C* zt, double* red,
                                   double* redc)
  {
    const uint32_t tid = threadIdx.x;
    const int n = I.n, D = I.depth;
    const bool grad = A.flags & kGrad;
    G g;
    g.init(n, 0, 0, n, 0, 0);
    C x[R], y[R];
    const C one{1, 0};
    int cur = g.a(0);
#pragma unroll
    for (int k = 0; k < R; ++k) {
      x[k] = C{0, 0};
      if (lidx<LR>(tid, k, cur) == 0) x[k].x = 1;
    }
    for (int l = 0; l < D; ++l) {
      const uint64_t row = I.par + (uint64_t)l * n;
      build_tables(zt, g, A.mats + 4 * row, false);
      sweep_fwd(x, lds, cur, g, A.mats, (int64_t)row, -1, zt, one, false);
    }
    const double cost = cost_and_lam(x, y, grad, cur, g, redc);
    if (tid == 0) A.cost[b] = cost;
    if (grad) {
      double* redw = red + (tid >> 6) * SLOTS;
      for (int l = D - 1; l >= 0; --l) {
        const uint64_t row = I.par + (uint64_t)l * n;
        build_tables(zt, g, A.mats + 4 * row, false);
        sweep_bwd(x, y, lds, cur, g, A.mats, (int64_t)row, -1, zt, one, false, false, 0, 0, redw);
        __syncthreads();
        for (int q = tid; q < n; q += NT) {
          double st = 0, sp = 0;
          for (int w = 0; w < NW; ++w) {
            st += red[w * SLOTS + M + q];
            sp += red[w * SLOTS + 2 * M + q];
          }
          A.grad[2 * (row + q)] = st;
          A.grad[2 * (row + q) + 1] = sp;
        }
      }
    }
    if (A.flags & kKeep) {
      C* psi = A.arena + I.psi;
#pragma unroll
      for (int k = 0; k < R; ++k) {
        const uint32_t u = lidx<LR>(tid, k, cur);
        if (u < (1u << n)) psi[u] = x[k];
      }
    }
  }

  static __device__ void run_pass(const HeaArgs<T>& A, const HeaInst& I, int s, uint32_t t, C* lds, C* zt,
                                  double* red, double* redc)
  {
    const uint32_t tid = threadIdx.x;
    const int n = I.n, D = I.depth;
    const bool grad = A.flags & kGrad;
    const int j = s <= D ? s : 2 * D - s;
    const bool isB = j & 1;
    const int c = 2 * M - n;
    G g;
    g.init(n, isB ? 2 : 1, isB ? c : 0, M, c, t);
    const int64_t row_pre = j >= 1 ? (int64_t)(I.par + (uint64_t)(j - 1) * n) : -1;
    const int64_t row_post = j < D ? (int64_t)(I.par + (uint64_t)j * n) : -1;
    // RZs deferred from pass j-1: the other tile kind's qubits (A: 0..M-1, B: M..n-1)
    const int q0 = j >= 1 ? (isB ? 0 : M) : 0, q1 = j >= 1 ? (isB ? M : n) : 0;
    constexpr int L = M - LR;  // load/store layout: consecutive threads, consecutive amplitudes
    C* psi = A.arena + I.psi;
    C* lam = A.arena + I.lam;
    C x[R], y[R];
    if (s == 0) {
#pragma unroll
      for (int k = 0; k < R; ++k) {
        x[k] = C{0, 0};
        if (g.gidx(lidx<LR>(tid, k, L)) == 0) x[k].x = 1;
      }
    } else {
#pragma unroll
      for (int k = 0; k < R; ++k) x[k] = psi[g.gidx(lidx<LR>(tid, k, L))];
    }
    if (s > D) {
#pragma unroll
      for (int k = 0; k < R; ++k) y[k] = lam[g.gidx(lidx<LR>(tid, k, L))];
    }
    C f0{1, 0};
    if (row_pre >= 0) {
      build_tables(zt, g, A.mats + 4 * row_pre, isB);
      f0 = deferred_factor(g, A.mats + 4 * row_pre, q0, q1);
    }
    const bool has_f0 = q1 > q0;
    int cur = L;
    double* redw = red + (tid >> 6) * SLOTS;
    if (s <= D) sweep_fwd(x, lds, cur, g, A.mats, row_pre, row_post, zt, f0, has_f0);
    if (s == D) {
      const double cs = cost_and_lam(x, y, grad, cur, g, redc);
      if (tid == 0) A.cpart[I.cpart + t] = cs;
    }
    if (grad && s >= D)
      sweep_bwd(x, y, lds, cur, g, A.mats, row_pre, row_post, zt, f0, has_f0, isB, q0, q1, redw);
    const bool store_x = s < D || (s == D && (grad || (A.flags & kKeep))) || (s > D && s < 2 * D);
    const bool store_y = grad && s >= D && s < 2 * D;
    if ((store_x || store_y) && cur != L) {
      relayout(x, lds, cur, L);
      if (store_y) relayout(y, lds, cur, L);
      cur = L;
    }
    if (store_x) {
#pragma unroll
      for (int k = 0; k < R; ++k) psi[g.gidx(lidx<LR>(tid, k, L))] = x[k];
    }
    if (store_y) {
#pragma unroll
      for (int k = 0; k < R; ++k) lam[g.gidx(lidx<LR>(tid, k, L))] = y[k];
    }
    if (grad && s >= D) {
      __syncthreads();
      // slots: [0, gg) dtheta of layer j (post), [M, M + gg) dtheta of layer j-1 (pre),
      // [2M, 2M + n) dphi of layer j-1 (every qubit: pre bits and deferred)
      const int gg = g.lb - g.la;
      const int nsl = 2 * gg + (row_pre >= 0 ? n : 0);
      for (int sl = tid; sl < nsl; sl += NT) {
        int layer, q, kind, slot;
        if (sl < 2 * gg) {
          const int which = sl / gg;  // 0: layer j (post), 1: layer j-1 (pre)
          layer = which ? j - 1 : j;
          q = g.gq(g.la + sl % gg);
          kind = 0;
          slot = which * M + sl % gg;
        } else {
          layer = j - 1;
          q = sl - 2 * gg;
          kind = 1;
          slot = 2 * M + q;
        }
        if (layer < 0 || layer >= D) continue;
        double v = 0;
        for (int w = 0; w < NW; ++w) v += red[w * SLOTS + slot];
        A.part[I.part + ((((uint64_t)layer * n + q) * 2 + kind) << I.ntl) + t] = v;
      }
    }
  }
};

template <typename T, int M, int R>
__global__ __launch_bounds__((Tile<T, M, R>::NT), 2) void k_hea(HeaArgs<T> A)
{
  using Tl = Tile<T, M, R>;
  __shared__ V<T> lds[1 << M];
  __shared__ V<T> zt[Tl::KMAX * R];
  __shared__ double red[Tl::NW * Tl::SLOTS];
  __shared__ double redc[Tl::NW];
  __shared__ uint32_t sh[2];
  const uint32_t tid = threadIdx.x;
  // The queue of this workgroup's XCD (hardware register XCC_ID, bits 3:0).
  const uint32_t q = __builtin_amdgcn_s_getreg((20) | (0 << 6) | ((4 - 1) << 11)) % kQueues;
  const HeaTask* task = A.task + A.qoff[q];
  const uint32_t ntask = A.qoff[q + 1] - A.qoff[q];
  uint32_t* head = A.ctl + q;
  // Thread 0 takes queue entries one task ahead and stages the next task's descriptors in
  // LDS while the current task drains its stores, so neither the queue atomic nor the
  // descriptor loads sit between two tasks.
  __shared__ HeaTask sh_tk;
  __shared__ HeaInst sh_in;
  uint32_t next = 0;
  auto stage = [&](uint32_t idx) {
    sh[0] = idx;
    if (idx < ntask) {
      const HeaTask t = task[idx];
      sh_tk = t;
      sh_in = A.inst[t.inst];
    }
  };
  if (tid == 0) {
    stage(atomicAdd(head, 1u));
    next = atomicAdd(head, 1u);
  }
  for (;;) {
    __syncthreads();
    const uint32_t i = sh[0];
    if (i >= ntask) return;
    const HeaTask tk = sh_tk;
    const HeaInst I = sh_in;
    __syncthreads();  // every thread holds its copy before thread 0 stages the next task
    if (I.whole) {
      Tl::run_whole(A, I, tk.inst, lds, zt, red, redc);
      if (tid == 0) {
        stage(next);
        next = atomicAdd(head, 1u);
      }
    } else {
      if (tid == 0) {
        uint32_t ok = 1;
        if (tk.step > 0) {
          uint32_t* ctr = A.ctl + kCtlArrive + I.arrive + tk.step - 1;
          const uint32_t need = (1u << I.ntl) + A.wait_extra;
          uint32_t spins = 0;
          while (__hip_atomic_load(ctr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < need) {
            if (__hip_atomic_load(A.ctl + kCtlAbort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) ||
                ++spins > A.spin_limit) {
              __hip_atomic_store(A.ctl + kCtlAbort, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
              ok = 0;
              break;
            }
            __builtin_amdgcn_s_sleep(1);
          }
        }
        sh[1] = ok;
      }
      __syncthreads();
      const bool ok = sh[1];
      if (ok) {
        if (tk.step > 0) acquire_tiles();
        Tl::run_pass(A, I, tk.step, tk.tile, lds, zt, red, redc);
      }
      __syncthreads();  // sh[1] read by all before it is reused
      if (tid == 0) {
        stage(next);
        next = atomicAdd(head, 1u);
      }
      if (ok) {
        // Publication of the tile: the compiler may not sink the state stores below the
        // wait (the builtin alone does not order memory), the wait completes them in the
        // XCD's L2, and the barrier puts every thread's wait before thread 0's arrival
        // count.  This is the gfx950 form of a release at the L2 of one XCD: a register's
        // tasks all run on one XCD (one queue per XCD), whose L2 is their point of
        // coherence, and the consumer invalidates its L1 (acquire_tiles).  An agent-scope
        // release fence here would also write back L2 for other XCDs, which no consumer
        // needs; it made the 512-register run 2.6x slower.
        __atomic_signal_fence(__ATOMIC_SEQ_CST);
        __builtin_amdgcn_s_waitcnt(0);  // this thread's state stores have completed
        __syncthreads();
        if (tid == 0)
          __hip_atomic_fetch_add(A.ctl + kCtlArrive + I.arrive + tk.step, 1u, __ATOMIC_RELAXED,
                                 __HIP_MEMORY_SCOPE_AGENT);
      }
    }
  }
}

// Which XCDs serve workgroups: bit (XCC_ID % 8) of *mask for every workgroup that runs.
__global__ void k_xcc_probe(uint32_t* mask)
{
  if (threadIdx.x == 0)
    atomicOr(mask, 1u << (__builtin_amdgcn_s_getreg((20) | (0 << 6) | ((4 - 1) << 11)) % kQueues));
}

__global__ void k_hea_zero(uint32_t* p, uint32_t n)
{
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

// Per (layer, qubit) pair: (cos theta/2, sin theta/2) and (cos phi/2, sin phi/2), computed in
// FP64 and rounded once to T (RY = [[c, -s], [s, c]], RZ = diag(e^{-i phi/2}, e^{i phi/2})).
template <typename T>
__global__ void k_hea_mats(const double* __restrict__ ang, V<T>* __restrict__ mats, uint64_t npairs)
{
  const uint64_t p = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  if (p >= npairs) return;
  double s, c, sp, cp;
  sincos(0.5 * ang[2 * p], &s, &c);
  sincos(0.5 * ang[2 * p + 1], &sp, &cp);
  mats[4 * p + 0] = V<T>{(T)c, (T)s};
  mats[4 * p + 1] = V<T>{(T)cp, (T)sp};
  mats[4 * p + 2] = V<T>{0, 0};
  mats[4 * p + 3] = V<T>{0, 0};
}

// One block per (tiled instance, layer): thread q sums the tile partials of qubit q in
// tile order into (d/dtheta, d/dphi); the costs are summed likewise.
__global__ void k_hea_epi(const HeaInst* __restrict__ inst, const uint2* __restrict__ items,
                          const double* __restrict__ part,
                          const double* __restrict__ cpart, double* grad, double* cost, int with_grad)
{
  const uint2 it = items[blockIdx.x];
  const HeaInst I = inst[it.x];
  const uint32_t q = threadIdx.x;
  const uint64_t T = uint64_t(1) << I.ntl;
  if (with_grad && q < I.n) {
    const uint64_t pi = I.par + (uint64_t)it.y * I.n + q;
    const double* p = part + I.part + ((((uint64_t)it.y * I.n + q) * 2) << I.ntl);
    double st = 0, sp = 0;
    for (uint64_t t = 0; t < T; ++t) {
      st += p[t];
      sp += p[T + t];
    }
    grad[2 * pi] = st;
    grad[2 * pi + 1] = sp;
  }
  if (it.y == 0 && q == blockDim.x - 1) {
    double v = 0;
    for (uint64_t t = 0; t < T; ++t) v += cpart[I.cpart + t];
    cost[it.x] = v;
  }
}

// ---- variants ------------------------------------------------------------------------
struct Variant {
  int M, R, test;
};
// Every variant keeps its registers within 256 VGPRs (two waves per SIMD); tiles that
// needed the accumulation registers as spill space ran at one wave per SIMD and are left out.
constexpr Variant kVar64[] = {{13, 16, 0}, {12, 16, 0}, {12, 8, 0}, {8, 4, 1}};
constexpr Variant kVar128[] = {{12, 8, 0}, {11, 8, 0}, {10, 8, 0}, {8, 4, 1}};
constexpr int kNumVar = 4;

const Variant& variant(int dtype, int v) { return dtype == IQSH_C128 ? kVar128[v] : kVar64[v]; }

template <typename T, int M, int R>
hipError_t launch_t(const HeaArgs<T>& a, int workers_per_cu, int cus, hipStream_t st)
{
  constexpr int NT = Tile<T, M, R>::NT;
  int fit = 0;
  hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&fit, k_hea<T, M, R>, NT, 0);
  if (e != hipSuccess) return e;
  fit = std::max(fit, 1);
  const int w = workers_per_cu > 0 ? std::min(workers_per_cu, fit) : fit;
  hipLaunchKernelGGL((k_hea<T, M, R>), dim3(std::max(1, w * cus)), dim3(NT), 0, st, a);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_var(int v, const HeaArgs<T>& a, int wpc, int cus, hipStream_t st);
template <>
hipError_t launch_var<float>(int v, const HeaArgs<float>& a, int wpc, int cus, hipStream_t st)
{
  switch (v) {
    case 0: return launch_t<float, 13, 16>(a, wpc, cus, st);
    case 1: return launch_t<float, 12, 16>(a, wpc, cus, st);
    case 2: return launch_t<float, 12, 8>(a, wpc, cus, st);
    default: return launch_t<float, 8, 4>(a, wpc, cus, st);
  }
}
template <>
hipError_t launch_var<double>(int v, const HeaArgs<double>& a, int wpc, int cus, hipStream_t st)
{
  switch (v) {
    case 0: return launch_t<double, 12, 8>(a, wpc, cus, st);
    case 1: return launch_t<double, 11, 8>(a, wpc, cus, st);
    case 2: return launch_t<double, 10, 8>(a, wpc, cus, st);
    default: return launch_t<double, 8, 4>(a, wpc, cus, st);
  }
}

constexpr uint32_t kSpinLimit = 1u << 20;  // ~1 s of polling before a wait gives up
constexpr int kMaxQubits = 26;
constexpr int kMaxDepth = 10000;

}  // namespace

struct iqsh_hea {
  int dtype = IQSH_C64;
  int count = 0;
  std::vector<int> n, d;
  std::vector<uint64_t> par, soff;
  uint64_t npairs = 0, arena_amps = 0;
  iqsh_hea_tune tune{0, 0, 0, 0};
  int cus = 0;
  uint32_t qmask = 0xffu;     // queues whose XCD serves workgroups (k_xcc_probe)
  uint32_t spin_limit = kSpinLimit, wait_extra = 0;
  hipStream_t cap = nullptr;  // private stream for graph capture

  void* arena = nullptr;
  void* mats = nullptr;
  double* ang = nullptr;
  double* grad = nullptr;
  double* cost = nullptr;
  double* part = nullptr;
  size_t part_cap = 0;
  double* cpart = nullptr;
  size_t cpart_cap = 0;
  HeaInst* inst = nullptr;
  HeaTask* task = nullptr;
  size_t task_cap = 0;
  uint32_t* ctl = nullptr;
  size_t ctl_cap = 0;
  uint32_t* qoff = nullptr;
  size_t qoff_cap = 0;
  std::vector<uint32_t> qlen;  // tasks per queue
  uint2* epi = nullptr;
  size_t epi_cap = 0;

  int built_mode = -1;  // flags (GRAD | KEEP) the schedule was built for
  uint32_t ntask = 0, nepi = 0, nctl = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  int graph_mode = -1;
  bool kept = false;

  size_t esz() const { return dtype == IQSH_C128 ? 16 : 8; }
};

namespace {

void drop_graph(iqsh_hea* h)
{
  if (h->gexec) (void)hipGraphExecDestroy(h->gexec);
  if (h->graph) (void)hipGraphDestroy(h->graph);
  h->gexec = nullptr;
  h->graph = nullptr;
  h->graph_mode = -1;
}

template <typename P>
int grow(P*& p, size_t& cap, size_t need, const char* what)
{
  if (need <= cap && p) return IQSH_OK;
  if (p) (void)hipFree(p);
  p = nullptr;
  cap = 0;
  const size_t cnt = std::max<size_t>(need, 1);
  hipError_t e = hipMalloc((void**)&p, cnt * sizeof(P));
  if (e != hipSuccess) return fail(IQSH_ERR_ALLOC, "hipMalloc(%zu bytes) for %s: %s", cnt * sizeof(P), what,
                                   hipGetErrorString(e));
  cap = cnt;
  return IQSH_OK;
}

int steps_of(int d, bool grad) { return grad ? 2 * d + 1 : d + 1; }

// Instance table, task queue and epilogue items for the current tuning and `mode`.
int build(iqsh_hea* h, int mode)
{
  const Variant& v = variant(h->dtype, h->tune.variant);
  const bool grad = mode & kGrad;
  std::vector<HeaInst> I(h->count);
  uint64_t part = 0, cpart = 0;
  uint32_t arrive = 0;
  std::vector<int> tiled, whole;
  int smax = 0;
  for (int b = 0; b < h->count; ++b) {
    HeaInst& x = I[b];
    std::memset(&x, 0, sizeof x);
    x.n = h->n[b];
    x.depth = h->d[b];
    x.whole = h->n[b] <= v.M;
    x.psi = h->soff[b];
    x.lam = h->soff[b] + (uint64_t(1) << h->n[b]);
    x.par = h->par[b];
    if (x.whole) {
      whole.push_back(b);
      continue;
    }
    if (h->n[b] > 2 * v.M)
      return fail(IQSH_ERR_NOTSUP, "instance %d: %d qubits exceed twice the tile (%d qubits)", b, h->n[b], v.M);
    x.ntl = h->n[b] - v.M;
    x.part = part;
    part += (uint64_t)h->d[b] * h->n[b] * 2 << x.ntl;
    x.cpart = cpart;
    cpart += uint64_t(1) << x.ntl;
    x.arrive = arrive;
    arrive += steps_of(h->d[b], grad);
    smax = std::max(smax, steps_of(h->d[b], grad));
    tiled.push_back(b);
  }
  // Larger registers first within a round; costliest single-tile instances first.
  auto by_size = [&](int a, int b) {
    if (h->n[a] != h->n[b]) return h->n[a] > h->n[b];
    if (h->d[a] != h->d[b]) return h->d[a] > h->d[b];
    return a < b;
  };
  std::stable_sort(tiled.begin(), tiled.end(), by_size);
  std::stable_sort(whole.begin(), whole.end(), by_size);

  // Every register is served by one XCD's queue (its tiles share that XCD's L2).  Registers
  // are dealt longest first to the least loaded queue; the cost is the bytes a register's
  // passes move (single-tile registers stay on chip: weighted by half their padded size).
  auto cost = [&](int b) {
    const double esz = (double)h->esz();
    if (I[b].whole) return 0.5 * h->d[b] * (double)(1u << v.M) * esz * (grad ? 3.0 : 1.0);
    return (double)steps_of(h->d[b], grad) * (double)(uint64_t(1) << h->n[b]) * esz * 3.0;
  };
  std::vector<int> all(h->count);
  for (int b = 0; b < h->count; ++b) all[b] = b;
  std::stable_sort(all.begin(), all.end(), [&](int a, int b) { return cost(a) > cost(b); });
  std::vector<int> qof(h->count);
  // Only queues whose XCD serves workgroups receive registers (a partitioned GPU exposes
  // fewer XCDs).
  double load[kQueues];
  for (int qi = 0; qi < kQueues; ++qi) load[qi] = (h->qmask >> qi) & 1 ? 0.0 : HUGE_VAL;
  for (int b : all) {
    const int qi = (int)(std::min_element(load, load + kQueues) - load);
    qof[b] = qi;
    load[qi] += cost(b);
  }

  std::vector<HeaTask> tasks;
  std::vector<uint32_t> qoff(kQueues + 1, 0);
  const int order = h->tune.order;
  for (int qi = 0; qi < kQueues; ++qi) {
    qoff[qi] = (uint32_t)tasks.size();
    std::vector<int> tq, wq;
    for (int b : tiled)
      if (qof[b] == qi) tq.push_back(b);
    for (int b : whole)
      if (qof[b] == qi) wq.push_back(b);
    auto push_whole = [&](size_t lo, size_t hi) {
      for (size_t i = lo; i < hi; ++i) tasks.push_back(HeaTask{(uint32_t)wq[i], 0, 0});
    };
    // Rounds of a set of tiled registers: round r holds step r - offset of each register.
    auto push_rounds = [&](const std::vector<int>& set, bool align_end, size_t w_lo, size_t w_hi) {
      int sm = 0;
      for (int b : set) sm = std::max(sm, steps_of(h->d[b], grad));
      size_t wn = w_lo;
      for (int r = 0; r < sm; ++r) {
        for (int b : set) {
          const int S = steps_of(h->d[b], grad);
          const int st = r - (align_end ? sm - S : 0);
          if (st < 0 || st >= S) continue;
          const uint32_t T = 1u << I[b].ntl;
          for (uint32_t t = 0; t < T; ++t) tasks.push_back(HeaTask{(uint32_t)b, (uint16_t)st, (uint16_t)t});
        }
        const size_t upto = w_lo + (w_hi - w_lo) * (size_t)(r + 1) / (size_t)std::max(sm, 1);
        push_whole(wn, upto);
        wn = upto;
      }
      push_whole(wn, w_hi);
    };
    if (h->tune.group_mib > 0) {
      // Groups of tiled registers (largest first) within this queue's share of the byte
      // budget; single-tile registers spread evenly over the groups' rounds.
      const double budget = (double)h->tune.group_mib * 1048576.0 / kQueues;
      std::vector<std::vector<int>> groups;
      double bytes = 0;
      for (int b : tq) {
        const double sb = 2.0 * (double)(uint64_t(1) << h->n[b]) * h->esz();
        if (groups.empty() || (bytes + sb > budget && !groups.back().empty())) {
          groups.emplace_back();
          bytes = 0;
        }
        groups.back().push_back(b);
        bytes += sb;
      }
      const size_t G = std::max<size_t>(groups.size(), 1);
      for (size_t gi = 0; gi < groups.size(); ++gi)
        push_rounds(groups[gi], false, wq.size() * gi / G, wq.size() * (gi + 1) / G);
      if (groups.empty()) push_whole(0, wq.size());
    } else {
      if (order == 1) push_whole(0, wq.size());
      if (order == 2)
        push_rounds(tq, false, 0, wq.size());
      else
        push_rounds(tq, order == 3, 0, 0);
      if (order == 0 || order == 3) push_whole(0, wq.size());
    }
  }
  qoff[kQueues] = (uint32_t)tasks.size();
  (void)smax;

  std::vector<uint2> epi;
  for (int b : tiled)
    for (int l = 0; l < (grad ? h->d[b] : 1); ++l) epi.push_back(make_uint2((uint32_t)b, (uint32_t)l));

  int rc;
  if ((rc = grow(h->part, h->part_cap, part, "gradient partials"))) return rc;
  if ((rc = grow(h->cpart, h->cpart_cap, cpart, "cost partials"))) return rc;
  if ((rc = grow(h->task, h->task_cap, tasks.size(), "task queue"))) return rc;
  if ((rc = grow(h->ctl, h->ctl_cap, kCtlArrive + (size_t)arrive, "counters"))) return rc;
  if ((rc = grow(h->qoff, h->qoff_cap, qoff.size(), "queue offsets"))) return rc;
  if ((rc = grow(h->epi, h->epi_cap, epi.size(), "epilogue items"))) return rc;
  hipStream_t st = iqsi::stream();
  IQSI_TRY(hipMemcpyAsync(h->inst, I.data(), sizeof(HeaInst) * I.size(), hipMemcpyHostToDevice, st));
  if (!tasks.empty())
    IQSI_TRY(hipMemcpyAsync(h->task, tasks.data(), sizeof(HeaTask) * tasks.size(), hipMemcpyHostToDevice, st));
  if (!epi.empty()) IQSI_TRY(hipMemcpyAsync(h->epi, epi.data(), sizeof(uint2) * epi.size(), hipMemcpyHostToDevice, st));
  IQSI_TRY(hipMemcpyAsync(h->qoff, qoff.data(), sizeof(uint32_t) * qoff.size(), hipMemcpyHostToDevice, st));
  IQSI_TRY(hipStreamSynchronize(st));
  h->qlen.assign(kQueues, 0);
  for (int qi = 0; qi < kQueues; ++qi) h->qlen[qi] = qoff[qi + 1] - qoff[qi];
  h->ntask = (uint32_t)tasks.size();
  h->nepi = (uint32_t)epi.size();
  h->nctl = kCtlArrive + arrive;
  h->built_mode = mode;
  drop_graph(h);
  return IQSH_OK;
}

template <typename T>
int enqueue_t(iqsh_hea* h, int mode, hipStream_t st)
{
  HeaArgs<T> a;
  a.inst = h->inst;
  a.task = h->task;
  a.qoff = h->qoff;
  a.flags = (uint32_t)mode;
  a.ctl = h->ctl;
  a.arena = (V<T>*)h->arena;
  a.mats = (const V<T>*)h->mats;
  a.grad = h->grad;
  a.cost = h->cost;
  a.part = h->part;
  a.cpart = h->cpart;
  a.spin_limit = h->spin_limit;
  a.wait_extra = h->wait_extra;
  hipLaunchKernelGGL(k_hea_zero, dim3((h->nctl + 255) / 256), dim3(256), 0, st, h->ctl, h->nctl);
  IQSI_TRY(hipGetLastError());
  hipLaunchKernelGGL(k_hea_mats<T>, dim3((unsigned)((h->npairs + 255) / 256)), dim3(256), 0, st, (const double*)h->ang,
                     (V<T>*)h->mats, h->npairs);
  IQSI_TRY(hipGetLastError());
  IQSI_TRY(launch_var<T>(h->tune.variant, a, h->tune.workers_per_cu, h->cus, st));
  if (h->nepi) {
    hipLaunchKernelGGL(k_hea_epi, dim3(h->nepi), dim3(64), 0, st, (const HeaInst*)h->inst, (const uint2*)h->epi,
                       (const double*)h->part, (const double*)h->cpart, h->grad, h->cost,
                       (int)(mode & kGrad));
    IQSI_TRY(hipGetLastError());
  }
  return IQSH_OK;
}

int enqueue(iqsh_hea* h, int mode, hipStream_t st)
{
  return h->dtype == IQSH_C128 ? enqueue_t<double>(h, mode, st) : enqueue_t<float>(h, mode, st);
}

int check_handle(iqsh_hea* h) { return h ? IQSH_OK : fail(IQSH_ERR_ARG, "null batch handle"); }

int check_tune(int dtype, const iqsh_hea_tune& t)
{
  if (t.variant < 0 || t.variant >= kNumVar) return fail(IQSH_ERR_ARG, "bad variant %d", t.variant);
  if (t.workers_per_cu < 0 || t.workers_per_cu > 16) return fail(IQSH_ERR_ARG, "bad workers_per_cu %d", t.workers_per_cu);
  if (t.order < 0 || t.order > 3) return fail(IQSH_ERR_ARG, "bad order %d", t.order);
  if (t.group_mib < 0 || t.group_mib > (1 << 20)) return fail(IQSH_ERR_ARG, "bad group_mib %d", t.group_mib);
  (void)dtype;
  return IQSH_OK;
}

}  // namespace

int iqsh_hea_variants(int dtype, int* count)
{
  if (!count) return fail(IQSH_ERR_ARG, "null output");
  if (dtype != IQSH_C64 && dtype != IQSH_C128) return fail(IQSH_ERR_ARG, "bad dtype %d", dtype);
  *count = kNumVar;
  return IQSH_OK;
}

int iqsh_hea_variant_info(int dtype, int v, int* tile_qubits, int* amps_per_thread, int* threads, int* for_tests)
{
  if (dtype != IQSH_C64 && dtype != IQSH_C128) return fail(IQSH_ERR_ARG, "bad dtype %d", dtype);
  if (v < 0 || v >= kNumVar) return fail(IQSH_ERR_ARG, "bad variant %d", v);
  const Variant& x = variant(dtype, v);
  if (tile_qubits) *tile_qubits = x.M;
  if (amps_per_thread) *amps_per_thread = x.R;
  if (threads) *threads = (1 << x.M) / x.R;
  if (for_tests) *for_tests = x.test;
  return IQSH_OK;
}

int iqsh_hea_create(int dtype, int count, const int* nq, const int* depth, const iqsh_hea_tune* tune,
                    iqsh_hea_handle* out)
{
  if (!out) return fail(IQSH_ERR_ARG, "null output handle");
  *out = nullptr;
  if (dtype != IQSH_C64 && dtype != IQSH_C128) return fail(IQSH_ERR_ARG, "bad dtype %d", dtype);
  if (count < 1 || !nq || !depth) return fail(IQSH_ERR_ARG, "empty batch");
  const iqsh_hea_tune t = tune ? *tune : iqsh_hea_tune{0, 0, 0, 0};
  if (int rc = check_tune(dtype, t)) return rc;
  iqsh_hea* h = new iqsh_hea;
  h->dtype = dtype;
  h->count = count;
  h->tune = t;
  h->cus = std::max(1, iqsi::device_cus());
  for (int b = 0; b < count; ++b) {
    if (nq[b] < 1 || nq[b] > kMaxQubits || depth[b] < 1 || depth[b] > kMaxDepth) {
      delete h;
      return fail(IQSH_ERR_ARG, "instance %d: %d qubits, %d layers (need 1..%d qubits, 1..%d layers)", b, nq[b],
                  depth[b], kMaxQubits, kMaxDepth);
    }
    h->n.push_back(nq[b]);
    h->d.push_back(depth[b]);
    h->par.push_back(h->npairs);
    h->npairs += (uint64_t)nq[b] * depth[b];
    h->soff.push_back(h->arena_amps);
    h->arena_amps += uint64_t(2) << nq[b];  // psi and lam
  }
  auto bail = [&](int rc) {
    iqsh_hea_destroy(h);
    return rc;
  };
  auto alloc = [&](void** p, size_t bytes, const char* what) -> int {
    hipError_t e = hipMalloc(p, std::max<size_t>(bytes, 1));
    return e == hipSuccess ? IQSH_OK
                           : fail(IQSH_ERR_ALLOC, "hipMalloc(%zu bytes) for %s: %s", bytes, what, hipGetErrorString(e));
  };
  int rc;
  if ((rc = alloc(&h->arena, h->arena_amps * h->esz(), "batch states"))) return bail(rc);
  if ((rc = alloc(&h->mats, h->npairs * 4 * h->esz(), "gate matrices"))) return bail(rc);
  if ((rc = alloc((void**)&h->ang, h->npairs * 2 * sizeof(double), "angles"))) return bail(rc);
  if ((rc = alloc((void**)&h->grad, h->npairs * 2 * sizeof(double), "gradients"))) return bail(rc);
  if ((rc = alloc((void**)&h->cost, (size_t)count * sizeof(double), "costs"))) return bail(rc);
  if ((rc = alloc((void**)&h->inst, (size_t)count * sizeof(HeaInst), "instance table"))) return bail(rc);
  if (hipMemset(h->ang, 0, h->npairs * 2 * sizeof(double)) != hipSuccess ||
      hipMemset(h->grad, 0, h->npairs * 2 * sizeof(double)) != hipSuccess ||
      hipMemset(h->cost, 0, (size_t)count * sizeof(double)) != hipSuccess)
    return bail(fail(IQSH_ERR_HIP, "clearing batch buffers failed"));
  if (hipStreamCreateWithFlags(&h->cap, hipStreamNonBlocking) != hipSuccess)
    return bail(fail(IQSH_ERR_HIP, "hipStreamCreateWithFlags failed"));
  {
    // Probe which XCDs serve workgroups: 8 workgroups per CU reach every XCD, since
    // workgroups are dispatched round-robin over them.
    uint32_t* m = nullptr;
    if ((rc = alloc((void**)&m, sizeof(uint32_t), "XCD probe"))) return bail(rc);
    uint32_t found = 0;
    bool okp = hipMemset(m, 0, sizeof(uint32_t)) == hipSuccess;
    if (okp) {
      hipLaunchKernelGGL(k_xcc_probe, dim3(8 * h->cus), dim3(64), 0, nullptr, m);
      okp = hipGetLastError() == hipSuccess && hipMemcpy(&found, m, sizeof found, hipMemcpyDeviceToHost) == hipSuccess;
    }
    (void)hipFree(m);
    if (!okp || !found) return bail(fail(IQSH_ERR_HIP, "XCD probe failed"));
    h->qmask = found;
  }
  if ((rc = build(h, kGrad))) return bail(rc);
  *out = h;
  return IQSH_OK;
}

int iqsh_hea_destroy(iqsh_hea_handle h)
{
  if (!h) return IQSH_OK;
  (void)hipStreamSynchronize(iqsi::stream());
  drop_graph(h);
  void* bufs[] = {h->arena, h->mats, h->ang, h->grad, h->cost, h->part, h->cpart, h->inst, h->task, h->ctl, h->epi,
                  h->qoff};
  for (void* p : bufs)
    if (p) (void)hipFree(p);
  if (h->cap) (void)hipStreamDestroy(h->cap);
  delete h;
  return IQSH_OK;
}

int iqsh_hea_num_params(iqsh_hea_handle h, uint64_t* np)
{
  if (int rc = check_handle(h)) return rc;
  if (!np) return fail(IQSH_ERR_ARG, "null output");
  *np = 2 * h->npairs;
  return IQSH_OK;
}

int iqsh_hea_set_tune(iqsh_hea_handle h, const iqsh_hea_tune* t)
{
  if (int rc = check_handle(h)) return rc;
  if (!t) return fail(IQSH_ERR_ARG, "null tune");
  if (int rc = check_tune(h->dtype, *t)) return rc;
  const bool same = t->variant == h->tune.variant && t->workers_per_cu == h->tune.workers_per_cu &&
                    t->order == h->tune.order && t->group_mib == h->tune.group_mib;
  if (same) return IQSH_OK;
  const iqsh_hea_tune old = h->tune;
  h->tune = *t;
  const int rc = build(h, h->built_mode < 0 ? (int)kGrad : h->built_mode);
  if (rc) {
    h->tune = old;
    build(h, h->built_mode < 0 ? (int)kGrad : h->built_mode);
  }
  return rc;
}

int iqsh_hea_get_tune(iqsh_hea_handle h, iqsh_hea_tune* t)
{
  if (int rc = check_handle(h)) return rc;
  if (!t) return fail(IQSH_ERR_ARG, "null tune");
  *t = h->tune;
  return IQSH_OK;
}

int iqsh_hea_set_angles(iqsh_hea_handle h, const double* a)
{
  if (int rc = check_handle(h)) return rc;
  if (!a) return fail(IQSH_ERR_ARG, "null angles");
  IQSI_TRY(hipMemcpyAsync(h->ang, a, h->npairs * 2 * sizeof(double), hipMemcpyHostToDevice, iqsi::stream()));
  IQSI_TRY(hipStreamSynchronize(iqsi::stream()));
  return IQSH_OK;
}

int iqsh_hea_set_angles_device(iqsh_hea_handle h, const double* a)
{
  if (int rc = check_handle(h)) return rc;
  if (!a) return fail(IQSH_ERR_ARG, "null angles");
  IQSI_TRY(hipMemcpyAsync(h->ang, a, h->npairs * 2 * sizeof(double), hipMemcpyDeviceToDevice, iqsi::stream()));
  return IQSH_OK;
}

int iqsh_hea_run(iqsh_hea_handle h, int flags)
{
  if (int rc = check_handle(h)) return rc;
  if (flags & ~(IQSH_HEA_GRAD | IQSH_HEA_KEEP_STATE | IQSH_HEA_NO_GRAPH)) return fail(IQSH_ERR_ARG, "bad flags %d", flags);
  if ((flags & kGrad) && (flags & kKeep)) return fail(IQSH_ERR_ARG, "IQSH_HEA_KEEP_STATE needs a forward-only run");
  const int mode = flags & (kGrad | kKeep);
  if (h->built_mode != mode)
    if (int rc = build(h, mode)) return rc;
  hipStream_t st = iqsi::stream();
  if (flags & IQSH_HEA_NO_GRAPH) {
    if (int rc = enqueue(h, mode, st)) return rc;
  } else {
    if (h->graph_mode != mode) {
      drop_graph(h);
      IQSI_TRY(hipStreamBeginCapture(h->cap, hipStreamCaptureModeThreadLocal));
      const int rc = enqueue(h, mode, h->cap);
      hipGraph_t g = nullptr;
      const hipError_t e = hipStreamEndCapture(h->cap, &g);
      if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
      }
      if (e != hipSuccess) return fail(IQSH_ERR_HIP, "hipStreamEndCapture: %s", hipGetErrorString(e));
      h->graph = g;
      IQSI_TRY(hipGraphInstantiate(&h->gexec, h->graph, nullptr, nullptr, 0));
      h->graph_mode = mode;
    }
    IQSI_TRY(hipGraphLaunch(h->gexec, st));
  }
  h->kept = (mode & kKeep) != 0;
  return IQSH_OK;
}

namespace {
int check_ctl(const iqsh_hea* h, const uint32_t* ctl)
{
  if (ctl[kCtlAbort]) return fail(IQSH_ERR_HIP, "batch run aborted: a tile dependency wait timed out");
  for (int qi = 0; qi < kQueues; ++qi)
    if (ctl[qi] < h->qlen[qi])
      return fail(IQSH_ERR_HIP, "batch run incomplete: queue %d served %u of %u tasks (no workgroup on its XCD)", qi,
                  ctl[qi], h->qlen[qi]);
  return IQSH_OK;
}
}  // namespace

int iqsh_hea_set_wait_limit(iqsh_hea_handle h, uint32_t max_polls, int unmet)
{
  if (int rc = check_handle(h)) return rc;
  if (max_polls < 1) return fail(IQSH_ERR_ARG, "max_polls must be at least 1");
  h->spin_limit = max_polls;
  h->wait_extra = unmet ? 1u : 0u;
  drop_graph(h);  // the captured launch holds the old limits
  return IQSH_OK;
}

int iqsh_hea_results(iqsh_hea_handle h, double* cost, double* grad)
{
  if (int rc = check_handle(h)) return rc;
  hipStream_t st = iqsi::stream();
  uint32_t ctl[kCtlArrive] = {0};
  IQSI_TRY(hipMemcpyAsync(ctl, h->ctl, sizeof ctl, hipMemcpyDeviceToHost, st));
  if (cost) IQSI_TRY(hipMemcpyAsync(cost, h->cost, (size_t)h->count * sizeof(double), hipMemcpyDeviceToHost, st));
  if (grad) IQSI_TRY(hipMemcpyAsync(grad, h->grad, h->npairs * 2 * sizeof(double), hipMemcpyDeviceToHost, st));
  IQSI_TRY(hipStreamSynchronize(st));
  return check_ctl(h, ctl);
}

int iqsh_hea_results_device(iqsh_hea_handle h, double* cost, double* grad)
{
  if (int rc = check_handle(h)) return rc;
  hipStream_t st = iqsi::stream();
  // The run's control words decide whether the results are valid: read them first (this
  // synchronizes the stream, as the host form does).
  uint32_t ctl[kCtlArrive] = {0};
  IQSI_TRY(hipMemcpyAsync(ctl, h->ctl, sizeof ctl, hipMemcpyDeviceToHost, st));
  IQSI_TRY(hipStreamSynchronize(st));
  if (int rc = check_ctl(h, ctl)) return rc;
  if (cost) IQSI_TRY(hipMemcpyAsync(cost, h->cost, (size_t)h->count * sizeof(double), hipMemcpyDeviceToDevice, st));
  if (grad) IQSI_TRY(hipMemcpyAsync(grad, h->grad, h->npairs * 2 * sizeof(double), hipMemcpyDeviceToDevice, st));
  return IQSH_OK;
}

int iqsh_hea_state(iqsh_hea_handle h, int b, void* host)
{
  if (int rc = check_handle(h)) return rc;
  if (b < 0 || b >= h->count || !host) return fail(IQSH_ERR_ARG, "bad instance %d or null buffer", b);
  if (!h->kept) return fail(IQSH_ERR_ARG, "no state kept: run with IQSH_HEA_KEEP_STATE first");
  hipStream_t st = iqsi::stream();
  IQSI_TRY(hipMemcpyAsync(host, (const char*)h->arena + h->soff[b] * h->esz(), (size_t(1) << h->n[b]) * h->esz(),
                          hipMemcpyDeviceToHost, st));
  IQSI_TRY(hipStreamSynchronize(st));
  return IQSH_OK;
}

int iqsh_hea_traffic(iqsh_hea_handle h, int flags, double* bytes)
{
  if (int rc = check_handle(h)) return rc;
  if (!bytes) return fail(IQSH_ERR_ARG, "null output");
  const Variant& v = variant(h->dtype, h->tune.variant);
  const bool grad = flags & kGrad, keep = flags & kKeep;
  double tot = 0;
  for (int b = 0; b < h->count; ++b) {
    const double S = (double)(uint64_t(1) << h->n[b]) * h->esz();
    const int D = h->d[b];
    if (h->n[b] <= v.M) {
      if (keep) tot += S;
      continue;
    }
    // Forward: step 0 writes psi, steps 1..D-1 read and write it.
    tot += S + 2.0 * S * (D - 1);
    if (grad) {
      tot += 3.0 * S;                // turnaround: read psi, write psi and lam
      tot += 4.0 * S * (D - 1);      // backward steps 1..D-1: read and write both
      tot += 2.0 * S;                // last backward step: read both
    } else {
      tot += S + (keep ? S : 0.0);   // last forward step: read psi (and keep it)
    }
  }
  *bytes = tot;
  return IQSH_OK;
}

int iqsh_hea_autotune(iqsh_hea_handle h, int flags, int reps, iqsh_hea_trial* trials, int max_trials, int* ntrials,
                      double* best_ms)
{
  if (int rc = check_handle(h)) return rc;
  reps = std::max(1, reps);
  const iqsh_hea_tune start = h->tune;
  int nmax = 0;
  for (int b = 0; b < h->count; ++b) nmax = std::max(nmax, h->n[b]);
  hipEvent_t e0, e1;
  IQSI_TRY(hipEventCreate(&e0));
  IQSI_TRY(hipEventCreate(&e1));
  hipStream_t st = iqsi::stream();
  double best = 1e300;
  iqsh_hea_tune bt = start;
  int nt = 0, rc = IQSH_OK;
  for (int v = 0; v < kNumVar && !rc; ++v) {
    const Variant& x = variant(h->dtype, v);
    if (x.test || nmax > 2 * x.M) continue;
    for (int w = 1; w <= 2 && !rc; ++w) {
      for (int o = 0; o < 6 && !rc; ++o) {
        const iqsh_hea_tune t = o < 4 ? iqsh_hea_tune{v, w, o, 0} : iqsh_hea_tune{v, w, 0, o == 4 ? 64 : 160};
        if ((rc = iqsh_hea_set_tune(h, &t))) break;
        if ((rc = iqsh_hea_run(h, flags))) break;  // warm-up (and graph capture)
        if (hipEventRecord(e0, st) != hipSuccess) { rc = fail(IQSH_ERR_HIP, "hipEventRecord failed"); break; }
        for (int r = 0; r < reps && !rc; ++r) rc = iqsh_hea_run(h, flags);
        if (rc) break;
        if (hipEventRecord(e1, st) != hipSuccess || hipEventSynchronize(e1) != hipSuccess) {
          rc = fail(IQSH_ERR_HIP, "event timing failed");
          break;
        }
        float ms = 0;
        (void)hipEventElapsedTime(&ms, e0, e1);
        if ((rc = iqsh_hea_results(h, nullptr, nullptr))) break;  // aborted runs are errors
        const double per = ms / reps;
        if (trials && nt < max_trials) trials[nt] = iqsh_hea_trial{t, per};
        ++nt;
        if (per < best) {
          best = per;
          bt = t;
        }
      }
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (ntrials) *ntrials = nt;
  if (rc) {
    iqsh_hea_set_tune(h, &start);
    return rc;
  }
  if (nt == 0) return fail(IQSH_ERR_NOTSUP, "no kernel variant fits instances of %d qubits", nmax);
  if (best_ms) *best_ms = best;
  return iqsh_hea_set_tune(h, &bt);
}
