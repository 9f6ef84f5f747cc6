// C-ABI implementation of libiqs_hip.so: state lifetime, launch dispatch, reductions,
// QAOA and adjoint kernels.  See include/iqs_hip.h for the contract of each entry point.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <vector>

#include "../../include/iqs_hip.h"
#include "iqs_kernels.hpp"
#include "iqs_internal.hpp"

using namespace iqsk;


namespace {

thread_local char g_err[512] = "";
hipStream_t g_stream = nullptr;
int g_block_mfma = 1;  // 5-qubit block kernel (iqsh_set_block_mfma): 0 VALU, 1 MFMA 3 products, 2 MFMA 4 products

int fail(int code, const char* fmt, ...)
{
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof g_err, fmt, ap);
  va_end(ap);
  return code;
}

#define HIP_TRY(expr)                                                                   \
  do {                                                                                  \
    hipError_t e_ = (expr);                                                             \
    if (e_ != hipSuccess) return fail(IQSH_ERR_HIP, "%s: %s", #expr, hipGetErrorString(e_)); \
  } while (0)

#define CHECK_LAUNCH() HIP_TRY(hipGetLastError())

// ---- reduction workspace: partial sums [NV][kMaxBlocks] + result [NV], pinned host copy ----
constexpr int kMaxRedBlocks = 2048;
constexpr int kMaxRedVals = 4;
struct RedWork {
  double* partial = nullptr;
  double* result = nullptr;
  double* host = nullptr;
};
RedWork g_red;
std::mutex g_mu;
// One reduction at a time uses g_red: a reduction holds this from its first launch until
// red_finish has synchronized the stream, so host threads sharing the library do not race
// on the partial sums.
std::mutex g_red_mu;

int red_ws(RedWork** out)
{
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_red.partial) {
    HIP_TRY(hipMalloc(&g_red.partial, sizeof(double) * kMaxRedVals * kMaxRedBlocks));
    HIP_TRY(hipMalloc(&g_red.result, sizeof(double) * kMaxRedVals));
    HIP_TRY(hipHostMalloc((void**)&g_red.host, sizeof(double) * kMaxRedVals, hipHostMallocDefault));
  }
  *out = &g_red;
  return IQSH_OK;
}

// ---- device matrix ring for block gates: pinned staging + device slots + events ----
constexpr int kRing = 32;
constexpr size_t kSlotBytes = 32 * 32 * 16;
struct Ring {
  char* host = nullptr;
  char* dev = nullptr;
  hipEvent_t ev[kRing];
  bool used[kRing] = {};
  int next = 0;
};
Ring g_ring;

int ring_put(const void* src, size_t bytes, void** dptr, int* slot)
{
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_ring.host) {
    HIP_TRY(hipHostMalloc((void**)&g_ring.host, kSlotBytes * kRing, hipHostMallocDefault));
    HIP_TRY(hipMalloc((void**)&g_ring.dev, kSlotBytes * kRing));
    for (int i = 0; i < kRing; ++i) HIP_TRY(hipEventCreateWithFlags(&g_ring.ev[i], hipEventDisableTiming));
  }
  const int s = g_ring.next;
  g_ring.next = (s + 1) % kRing;
  if (g_ring.used[s]) HIP_TRY(hipEventSynchronize(g_ring.ev[s]));
  std::memcpy(g_ring.host + s * kSlotBytes, src, bytes);
  HIP_TRY(hipMemcpyAsync(g_ring.dev + s * kSlotBytes, g_ring.host + s * kSlotBytes, bytes, hipMemcpyHostToDevice,
                         g_stream));
  *dptr = g_ring.dev + s * kSlotBytes;
  *slot = s;
  return IQSH_OK;
}

int ring_done(int slot)
{
  HIP_TRY(hipEventRecord(g_ring.ev[slot], g_stream));
  g_ring.used[slot] = true;
  return IQSH_OK;
}

inline unsigned grid_for(uint64_t items, int per_block)
{
  const uint64_t g = (items + per_block - 1) / per_block;
  return (unsigned)std::max<uint64_t>(g, 1);
}

// 2D grid of `nblocks` workgroups (x fastest, at most 65536 in x).
inline dim3 grid2(uint64_t nblocks)
{
  const uint64_t gx = nblocks < 65536 ? (nblocks ? nblocks : 1) : 65536;
  return dim3((unsigned)gx, (unsigned)((nblocks + gx - 1) / gx), 1);
}

inline unsigned red_grid(uint64_t items)
{
  // Fixed function of the problem size (deterministic summation order).
  const uint64_t per = (uint64_t)kBlock * 8;
  return (unsigned)std::min<uint64_t>(std::max<uint64_t>((items + per - 1) / per, 1), kMaxRedBlocks);
}

int check_h(iqsh_handle h)
{
  if (!h || !h->d) return fail(IQSH_ERR_ARG, "null or released state handle");
  return IQSH_OK;
}

int check_pos(iqsh_handle h, unsigned p)
{
  if (p >= h->n) return fail(IQSH_ERR_ARG, "position %u out of range for %u qubits", p, h->n);
  return IQSH_OK;
}

template <typename T>
Cx<T> cx(const double* a)
{
  return Cx<T>{(T)a[0], (T)a[1]};
}

template <typename T>
Mat2<T> mat2(const double* m)
{
  return Mat2<T>{cx<T>(m), cx<T>(m + 2), cx<T>(m + 4), cx<T>(m + 6)};
}

bool is_c(const double* a, double re, double im) { return a[0] == re && a[1] == im; }

// ---------------------------------------------------------------------------------------
// Launchers

// Launch configurations (see the tiling notes in iqs_kernels.hpp).
constexpr int kTB = 1024, kTP = 2;  // pair and 4x4 kernels for complex64 (tiled)
// Pair and 4x4 kernels, per precision (30q sweep of block size x items per thread x split
// count): complex128 streams best with 128-thread blocks of 2 items split 4 ways (64-item
// runs, one wavefront access each): dense 1q 6.15 -> 6.53 TB/s for targets >= 11, controlled
// 5.88 -> 6.17, dense 2q 5.80 -> 6.18; complex64 keeps 1024 x 2 (smaller blocks are no
// faster there).
// QB x QP: the dense 4x4 kernel with a position above 10 (256 x 1 for both precisions:
// 30q complex128 at (5,20), (8,25), (12,13), (20,21), (24,25), (28,29), (15,29) averaged
// 6.14 -> 6.27 TB/s against 128 x 2, 32q complex64 6.16 -> 6.30 against 1024 x 1).
// CB x CP: pairs with two or more inserted bits (controlled 2x2, swap family) from the line
// width up and above position 10: complex64 256 x 1 (32q: 12 controlled and swap cases
// averaged 6.07 -> 6.27 TB/s against 1024 x 2, every case faster); complex128 256 x 1 for
// moves (SWAP, CNOT: 30q swaps 6.15 -> 6.39 against 128 x 2) and 128 x 2 for products
// (256 x 1 ran controlled SU(2) gates 2-4% slower).
// SB x SP: dense 1q with the target from position 6 up, outside the complex128 <= 10 rule
// (complex64 256 x 1: 32q targets 6-31 averaged 6.43 -> 6.63 TB/s against 1024 x 2, every
// target faster; complex128 128 x 2 is the headline kernel, at the in-place scale rate).
template <typename T> struct PairTile {
  static constexpr int B = 1024, P = 2, QB = 256, QP = 1, CB = 256, CP = 1, MB = 256, MP = 1, SB = 256, SP = 1;
};
template <> struct PairTile<double> {
  static constexpr int B = 128, P = 2, QB = 256, QP = 1, CB = 128, CP = 2, MB = 256, MP = 1, SB = 128, SP = 2;
};
constexpr int kLB = 256, kLP = 2;   // lane-shuffle 1q kernel
// Lane-shuffle 1q kernel and 2q diagonal, per precision (30q sweeps): complex128 lane kernel
// 64 x 2 (one wavefront per workgroup) 6.36 -> 6.60 TB/s against 256 x 2, complex64 keeps
// 256 x 2 (6.71); diagonal 128 x 2 for both: 6.61 -> 6.77 (complex128) against 1024 x 2.
template <typename T> struct LaneTile { static constexpr int B = kLB, P = kLP; };
template <> struct LaneTile<double> { static constexpr int B = 64, P = 2; };
// Diagonal and partial-state scale kernels: 128 x 2 for both precisions (2q diagonal,
// complex64 6.22 -> 6.54 TB/s; Z / CZ scale 6.31 -> 6.48 complex128, 6.43 -> 6.60 complex64).
template <typename T> struct DiagTile { static constexpr int B = 128, P = 2; };
// Lane-shuffle 4x4, for a higher position >= 6 and below it: complex64 64 x 4 everywhere
// (30q: (1,0) 4.98 -> 5.63 TB/s, (29,0) 5.86 -> 6.14); complex128 64 x 2 above ((0,29)
// 5.83 -> 6.08, swap (1,28) 2.74 -> 3.02 algorithmic), 256 x 2 below (5.45 against 5.15).
// B x P: the lane 4x4 with its higher position >= 6; complex64 256 x 1 (32q: dense (0,31),
// (3,29) and swaps (0,31), (1,30), (2,29) 1-8% faster than 64 x 4, (5,20) equal).
template <typename T> struct QuadLaneTile { static constexpr int B = 256, P = 1, BL = 64, PL = 4; };
template <> struct QuadLaneTile<double> { static constexpr int B = 64, P = 2, BL = kLB, PL = kLP; };

// Index bits inside one 128-byte cache line: 3 for complex128, 4 for complex64.
template <typename T>
constexpr unsigned line_bits() { return sizeof(T) == 8 ? 3u : 4u; }
constexpr int kSB = 256, kSP = 4;   // plain kernels for small states

template <typename T, int K, bool MOVE>
int launch_pair_m(iqsh_handle h, const Ins<K>& ins, uint64_t mask0, uint64_t mask1, const double* m)
{
  using V = typename Vec<T>::t;
  const uint64_t items = h->size() >> K;
  constexpr int TB = PairTile<T>::B, TP = PairTile<T>::P;
  const dim3 grid = grid2(items / (TB * TP));
  if (items >= kTiledMin && K >= 2 && ins.p[0] < line_bits<T>())
    // Lowest position inside a cache line (a control above it): cached accesses let L2
    // complete the lines that the two halves of each pair share.
    k_pair<T, K, MOVE, true, TB, TP, kSub, kSubLog, false><<<grid, TB, 0, g_stream>>>((V*)h->d, items, ins, mask0,
                                                                                           mask1, mat2<T>(m));
  else if (items >= kTiledMin && sizeof(T) == 8 && ins.p[K - 1] <= 10)
    // Complex128 with every touched position <= 10: sub-ranges 2^10 items apart (30q sweep,
    // dense 1q targets 6-10: 6.44 TB/s at 2^11, 6.57 at 2^10).
    k_pair<T, K, MOVE, true, TB, TP, kSub, 10><<<grid, TB, 0, g_stream>>>((V*)h->d, items, ins, mask0, mask1,
                                                                            mat2<T>(m));
  else if (items >= kTiledMin && K >= 2 && sizeof(T) == 8 && ins.p[0] < 6)
    // Complex128 with the lowest position 3-5 (inside a wavefront, from the line width up):
    // 128 x 1 (30q: controlled (3,20) 5.42 -> 6.16 TB/s, swap (4,25) 5.82 -> 6.57).
    k_pair<T, K, MOVE, true, 128, 1><<<grid2(items / 128), 128, 0, g_stream>>>((V*)h->d, items, ins, mask0, mask1,
                                                                               mat2<T>(m));
  else if (items >= kTiledMin && K >= 2) {
    constexpr int GB = MOVE ? PairTile<T>::MB : PairTile<T>::CB, GP = MOVE ? PairTile<T>::MP : PairTile<T>::CP;
    k_pair<T, K, MOVE, true, GB, GP><<<grid2(items / (GB * GP)), GB, 0, g_stream>>>((V*)h->d, items, ins, mask0, mask1,
                                                                                    mat2<T>(m));
  }
  else if (items >= kTiledMin) {
    // one inserted bit here (two or more took the branch above)
    constexpr int GB = PairTile<T>::SB, GP = PairTile<T>::SP;
    k_pair<T, K, MOVE, true, GB, GP><<<grid2(items / (GB * GP)), GB, 0, g_stream>>>((V*)h->d, items, ins, mask0, mask1,
                                                                                    mat2<T>(m));
  } else
    k_pair<T, K, MOVE, false, kSB, kSP><<<grid2(grid_for(items, kSB * kSP)), kSB, 0, g_stream>>>((V*)h->d, items, ins, mask0,
                                                                                          mask1, mat2<T>(m));
  CHECK_LAUNCH();
  return IQSH_OK;
}

template <typename T, int K>
int launch_pair(iqsh_handle h, const Ins<K>& ins, uint64_t mask0, uint64_t mask1, const double* m, bool move)
{
  return move ? launch_pair_m<T, K, true>(h, ins, mask0, mask1, m) : launch_pair_m<T, K, false>(h, ins, mask0, mask1, m);
}

template <typename T, int KC, bool COND, bool MOVE>
int launch_lane(iqsh_handle h, unsigned t, const Ins<KC>& cins, uint64_t ctrl, const double* m)
{
  using V = typename Vec<T>::t;
  const uint64_t items = h->size() >> KC;
  constexpr int LB = LaneTile<T>::B, LP = LaneTile<T>::P;
  k_pair_lane<T, KC, COND, MOVE, LB, LP><<<grid2(items / (LB * LP)), LB, 0, g_stream>>>((V*)h->d, t, cins, ctrl,
                                                                                                  mat2<T>(m));
  CHECK_LAUNCH();
  return IQSH_OK;
}

template <typename T, int KC>
int lane_dispatch(iqsh_handle h, unsigned t, const Ins<KC>& cins, uint64_t ctrl, const double* m, bool move, bool cond)
{
  if (cond) return move ? launch_lane<T, KC, true, true>(h, t, cins, ctrl, m) : launch_lane<T, KC, true, false>(h, t, cins, ctrl, m);
  return move ? launch_lane<T, KC, false, true>(h, t, cins, ctrl, m) : launch_lane<T, KC, false, false>(h, t, cins, ctrl, m);
}

template <typename T>
int launch_pair_cond(iqsh_handle h, unsigned t, uint64_t ctrl, const double* m, bool move)
{
  using V = typename Vec<T>::t;
  Ins<1> ins;
  ins.p[0] = t;
  const uint64_t items = h->size() >> 1;
  // complex128: the pair kernel's 128 x 2 shape (30q: 6.0 -> 6.5 TB/s of full-pass traffic).
  constexpr int TB = PairTile<T>::B, TP = PairTile<T>::P;
  const uint64_t grid = items / (TB * TP);
  if (move)
    k_pair_cond<T, true, TB, TP><<<grid2(grid), TB, 0, g_stream>>>((V*)h->d, ins, ctrl, mat2<T>(m));
  else
    k_pair_cond<T, false, TB, TP><<<grid2(grid), TB, 0, g_stream>>>((V*)h->d, ins, ctrl, mat2<T>(m));
  CHECK_LAUNCH();
  return IQSH_OK;
}

template <typename T>
int launch_diag2(iqsh_handle h, unsigned p1, unsigned p2, Cx<T> d0, Cx<T> d1, Cx<T> d2, Cx<T> d3);

template <typename T, int K>
int launch_scale(iqsh_handle h, const Ins<K>& ins, uint64_t mask, const double* f)
{
  using V = typename Vec<T>::t;
  const uint64_t items = h->size() >> K;
  if (K >= 1 && K <= 2 && ins.p[0] < line_bits<T>() && h->size() >= kTiledMin) {
    // A position inside a 128-byte line: the touched amplitudes share every line with
    // untouched ones, and partial-line reads and writes cost more than a full elementwise
    // pass (30q complex128, Z on position 0: 8.6 ms scaled vs 5.2 ms full).  Same result:
    // the other amplitudes are multiplied by exactly 1.
    const Cx<T> one{1, 0}, fv = cx<T>(f);
    const unsigned pa = ins.p[0], pb = ins.p[K - 1];
    Cx<T> d[4];
    for (int v = 0; v < 4; ++v) {
      // entry v = 2 b(pb) + b(pa) for K = 2; for K = 1 entries 0 and 3 (pa = pb)
      const uint64_t bits = (((uint64_t)(v >> 1)) << pb) | (((uint64_t)(v & 1)) << pa);
      d[v] = (bits == mask) ? fv : one;
    }
    if (K == 1) return launch_diag2<T>(h, pb, pa, d[0], one, one, d[3]);
    return launch_diag2<T>(h, pb, pa, d[0], d[1], d[2], d[3]);
  }
  if (items >= kTiledMin)
    k_scale<T, K, true, DiagTile<T>::B, DiagTile<T>::P><<<grid2(items / (DiagTile<T>::B * DiagTile<T>::P)),
                                                          DiagTile<T>::B, 0, g_stream>>>((V*)h->d, items, ins, mask,
                                                                                         cx<T>(f));
  else
    k_scale<T, K, false, kSB, kSP><<<grid2(grid_for(items, kSB * kSP)), kSB, 0, g_stream>>>((V*)h->d, items, ins, mask, cx<T>(f));
  CHECK_LAUNCH();
  return IQSH_OK;
}

template <typename T>
int launch_diag2(iqsh_handle h, unsigned p1, unsigned p2, Cx<T> d0, Cx<T> d1, Cx<T> d2, Cx<T> d3)
{
  using V = typename Vec<T>::t;
  const uint64_t N = h->size();
  if (N >= kTiledMin)
    k_diag2<T, true, DiagTile<T>::B, DiagTile<T>::P><<<grid2(N / (DiagTile<T>::B * DiagTile<T>::P)), DiagTile<T>::B, 0,
                                                       g_stream>>>((V*)h->d, N, p1, p2, d0, d1, d2, d3);
  else
    k_diag2<T, false, kSB, kSP><<<grid2(grid_for(N, kSB * kSP)), kSB, 0, g_stream>>>((V*)h->d, N, p1, p2, d0, d1, d2, d3);
  CHECK_LAUNCH();
  return IQSH_OK;
}

template <int K>
Ins<K> sorted_ins(const unsigned* p)
{
  Ins<K> ins;
  for (int i = 0; i < K; ++i) ins.p[i] = p[i];
  std::sort(ins.p, ins.p + K);
  return ins;
}

// Dispatch a controlled / multi-controlled pair update: items carry K inserted zero bits.
template <typename T>
int pair_dispatch(iqsh_handle h, const std::vector<unsigned>& ins_pos, uint64_t mask0, uint64_t mask1,
                  const double* m, bool move)
{
  const unsigned* p = ins_pos.data();
  switch (ins_pos.size()) {
    case 1: return launch_pair<T, 1>(h, sorted_ins<1>(p), mask0, mask1, m, move);
    case 2: return launch_pair<T, 2>(h, sorted_ins<2>(p), mask0, mask1, m, move);
    case 3: return launch_pair<T, 3>(h, sorted_ins<3>(p), mask0, mask1, m, move);
    case 4: return launch_pair<T, 4>(h, sorted_ins<4>(p), mask0, mask1, m, move);
    default: return fail(IQSH_ERR_NOTSUP, "at most 3 controls are supported");
  }
}

template <typename T>
int scale_dispatch(iqsh_handle h, const std::vector<unsigned>& ins_pos, uint64_t mask, const double* f)
{
  const unsigned* p = ins_pos.data();
  switch (ins_pos.size()) {
    case 0: return launch_scale<T, 0>(h, Ins<0>{}, 0, f);
    case 1: return launch_scale<T, 1>(h, sorted_ins<1>(p), mask, f);
    case 2: return launch_scale<T, 2>(h, sorted_ins<2>(p), mask, f);
    case 3: return launch_scale<T, 3>(h, sorted_ins<3>(p), mask, f);
    case 4: return launch_scale<T, 4>(h, sorted_ins<4>(p), mask, f);
    default: return fail(IQSH_ERR_NOTSUP, "at most 3 controls are supported");
  }
}

// A (possibly controlled) 2x2 update on target t with control mask `ctrl` (bits all 1).
// Matrix structure is used exactly: identity -> nothing; diag(1, d) -> scale the bit-t = 1
// amplitudes; diag(a, b) -> two scales; X -> move; otherwise the dense pair kernel.
template <typename T>
int controlled_2x2(iqsh_handle h, uint64_t ctrl, unsigned t, const double* m)
{
  std::vector<unsigned> ins;
  for (unsigned b = 0; b < h->n; ++b)
    if ((ctrl >> b) & 1) ins.push_back(b);
  ins.push_back(t);
  const uint64_t tb = uint64_t(1) << t;
  const bool offzero = is_c(m + 2, 0, 0) && is_c(m + 4, 0, 0);
  if (offzero) {
    const bool one0 = is_c(m, 1, 0), one1 = is_c(m + 6, 1, 0);
    if (one0 && one1) return IQSH_OK;
    if (!one0 && !one1 && ctrl == 0) {
      // diag(a, b) on one position: one elementwise pass, d[2*b(t) + b(t)] in {a, b}.
      return launch_diag2<T>(h, t, t, cx<T>(m), Cx<T>{0, 0}, Cx<T>{0, 0}, cx<T>(m + 6));
    }
    int rc = IQSH_OK;
    if (!one1) rc = scale_dispatch<T>(h, ins, ctrl | tb, m + 6);
    if (rc == IQSH_OK && !one0) rc = scale_dispatch<T>(h, ins, ctrl, m);
    return rc;
  }
  const bool xmove = is_c(m, 0, 0) && is_c(m + 6, 0, 0) && is_c(m + 2, 1, 0) && is_c(m + 4, 1, 0);
  const int nc = __builtin_popcountll(ctrl);
  const unsigned lowc = ctrl ? (unsigned)__builtin_ctzll(ctrl) : 64u;
  if (h->size() >= (kTiledMin << 3)) {
    if (t < 6) {
      // Target inside a wavefront: lane-shuffle kernels.
      if (lowc >= 6) {
        std::vector<unsigned> cpos(ins.begin(), ins.end() - 1);
        switch (nc) {
          case 0: return lane_dispatch<T, 0>(h, t, Ins<0>{}, 0, m, xmove, false);
          case 1: return lane_dispatch<T, 1>(h, t, sorted_ins<1>(cpos.data()), ctrl, m, xmove, false);
          case 2: return lane_dispatch<T, 2>(h, t, sorted_ins<2>(cpos.data()), ctrl, m, xmove, false);
          case 3: return lane_dispatch<T, 3>(h, t, sorted_ins<3>(cpos.data()), ctrl, m, xmove, false);
          default: break;
        }
      } else if (lowc >= line_bits<T>()) {
        // Control=1 amplitudes fill whole 128-byte lines: touch only those lines.
        return pair_dispatch<T>(h, ins, ctrl, ctrl | tb, m, xmove);
      } else {
        return lane_dispatch<T, 0>(h, t, Ins<0>{}, ctrl, m, xmove, true);
      }
    } else if (lowc < 3) {
      return launch_pair_cond<T>(h, t, ctrl, m, xmove);
    }
  }
  return pair_dispatch<T>(h, ins, ctrl, ctrl | tb, m, xmove);
}

// ---- SpecializeV2 matrices (src/spec_kernels.cpp:172-178, 264-270, 356-361): the gate-typed
// bodies compute the same 2x2 action; their constants are evaluated here in the register's
// precision and applied through the structured 2x2 path.
template <typename T>
void spec_matrix(int spec, double angle, double* m)
{
  const T theta = (T)angle;
  const T isqrt2 = (T)(1 / std::sqrt(2.0));
  const T c2 = std::cos(theta / 2), s2 = std::sin(theta / 2);
  for (int i = 0; i < 8; ++i) m[i] = 0;
  switch (spec) {
    case IQSH_SPEC_H: m[0] = m[2] = m[4] = isqrt2; m[6] = -isqrt2; break;
    case IQSH_SPEC_RX: m[0] = m[6] = c2; m[3] = m[5] = -s2; break;
    case IQSH_SPEC_RY: m[0] = m[6] = c2; m[2] = -s2; m[4] = s2; break;
    case IQSH_SPEC_RZ: m[0] = m[6] = c2; m[1] = -s2; m[7] = s2; break;
    case IQSH_SPEC_X: m[2] = m[4] = 1; break;
    case IQSH_SPEC_Y: m[3] = -1; m[5] = 1; break;
    case IQSH_SPEC_Z: m[0] = 1; m[6] = -1; break;
    default: break;
  }
}

// ---------------------------------------------------------------------------------------
// Reductions: NV running sums per thread (or a max), block partials, one final block.

template <int NV>
__device__ __forceinline__ void block_sum_store(double (&v)[NV], double* partial)
{
  __shared__ double sh[NV][kBlock / 64];
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
      for (int w = 0; w < kBlock / 64; ++w) s += sh[r][w];
      partial[r * kMaxRedBlocks + blockIdx.x] = s;
    }
}

__device__ __forceinline__ void block_max_store(double v, double* partial)
{
  __shared__ double sh[kBlock / 64];
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off));
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) sh[wid] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = sh[0];
    for (int w = 1; w < kBlock / 64; ++w) s = fmax(s, sh[w]);
    partial[blockIdx.x] = s;
  }
}

// Second stage: out[r] = sum_b partial[r][b] (or max), fixed order.
__global__ __launch_bounds__(kBlock) void k_red_final(const double* __restrict__ partial, unsigned nblocks,
                                                      unsigned nv, int is_max, double* __restrict__ out)
{
  __shared__ double sh[kBlock];
  for (unsigned r = 0; r < nv; ++r) {
    double s = is_max ? -1.0 : 0.0;
    for (unsigned b = threadIdx.x; b < nblocks; b += kBlock) {
      const double x = partial[r * kMaxRedBlocks + b];
      s = is_max ? fmax(s, x) : s + x;
    }
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
      if (threadIdx.x < w) sh[threadIdx.x] = is_max ? fmax(sh[threadIdx.x], sh[threadIdx.x + w]) : sh[threadIdx.x] + sh[threadIdx.x + w];
      __syncthreads();
    }
    if (threadIdx.x == 0) out[r] = sh[0];
    __syncthreads();
  }
}

template <typename V>
__device__ __forceinline__ double norm2(const V& a)
{
  return (double)a.x * (double)a.x + (double)a.y * (double)a.y;
}

// Every k < items owned by this thread, U at a time with their loads in flight together
// (one load per lane in flight left complex64 norms at 5.5 TB/s): by default (U = 0) 8
// items of 16 bytes or more, 16 of 8 bytes.  Workgroup b takes chunks of U * kBlock items at
// b, b + gridDim.x, ...; thread t takes k = chunk U kBlock + u kBlock + t.  `load(k)` reads
// item k, `use(k, x)` accumulates it, in ascending k per thread, so the summation order is
// a function of the grid and the item count only.
template <int U = 0, typename L, typename F>
__device__ __forceinline__ void for_items(uint64_t items, L load, F use)
{
  using X = decltype(load(uint64_t(0)));
  constexpr int kRedU = U ? U : (sizeof(X) <= 8 ? 16 : 8);
  const uint64_t span = (uint64_t)kRedU * kBlock, step = (uint64_t)gridDim.x * span;
  uint64_t k0 = (uint64_t)blockIdx.x * span + threadIdx.x;
  for (; k0 + (uint64_t)(kRedU - 1) * kBlock < items; k0 += step) {
    X x[kRedU];
#pragma unroll
    for (int u = 0; u < kRedU; ++u) x[u] = load(k0 + (uint64_t)u * kBlock);
#pragma unroll
    for (int u = 0; u < kRedU; ++u) use(k0 + (uint64_t)u * kBlock, x[u]);
  }
#pragma unroll
  for (int u = 0; u < kRedU; ++u) {
    const uint64_t k = k0 + (uint64_t)u * kBlock;
    if (k < items) use(k, load(k));
  }
}

template <typename V>
struct Two {
  V a, b;
};

// sum |psi_i|^2 over items (K inserted bits | mask), K = 0 or 1.
template <typename T, int K>
__global__ __launch_bounds__(kBlock) void k_red_norm(const typename Vec<T>::t* __restrict__ st, uint64_t items,
                                                     Ins<K> ins, uint64_t mask, double* partial)
{
  double v[1] = {0};
  for_items(items, [&](uint64_t k) { return ldg(&st[item_base(k, ins) | mask]); },
            [&](uint64_t, const typename Vec<T>::t& a) { v[0] += norm2(a); });
  block_sum_store<1>(v, partial);
}

// max |psi_i|^2 over items (one inserted bit | mask)
template <typename T>
__global__ __launch_bounds__(kBlock) void k_red_maxnorm(const typename Vec<T>::t* __restrict__ st, uint64_t items,
                                                        Ins<1> ins, uint64_t mask, double* partial)
{
  double m = -1.0;
  for_items(items, [&](uint64_t k) { return ldg(&st[item_base(k, ins) | mask]); },
            [&](uint64_t, const typename Vec<T>::t& a) { m = fmax(m, norm2(a)); });
  block_max_store(m, partial);
}

// Pauli string expectation + norm in one pass (see iqsh_expval_pauli).
// FORM: 0 no X/Y, 1 every X/Y position below 6, 2 otherwise; one
// instantiation per form, so that each has the registers of its own path.
template <typename T, int FORM>
__global__ __launch_bounds__(kBlock) void k_red_pauli(const typename Vec<T>::t* __restrict__ st, uint64_t N,
                                                      uint64_t xmask, uint64_t zmask, unsigned ny, double* partial)
{
  using V = typename Vec<T>::t;
  double v[2] = {0, 0};
  if (FORM == 0) {
    for_items(N, [&](uint64_t i) { return ldg(&st[i]); },
              [&](uint64_t i, const V& a) {
                const double p = norm2(a);
                v[0] += (__popcll(i & zmask) & 1) ? -p : p;
                v[1] += p;
              });
  } else {
    // Pairs (j, j ^ xmask) enumerated with a zero inserted at the highest X/Y position: both
    // j and j ^ xmask then run through whole cache lines across a wavefront.  (Inserting it
    // at the lowest position made each access half a line whose other half is read far
    // away: 2.9 TB/s for X on positions 2 and 26 at 30 qubits, against 5.6 for one X.)
    const unsigned msb = 63 - __clzll((long long)xmask);
    // <P> = sum_k conj(psi_{k^x}) psi_k i^ny (-1)^popc(k & z).  For the pair (j, j^x) with
    // a = psi_j, b = psi_{j^x}, t = conj(a) b, s = (-1)^popc(j & z):
    //   term = i^ny s [conj(t) + (-1)^ny t]
    //        = (-1)^(ny/2) s 2 Re t            (ny even)
    //        = (-1)^((ny-1)/2) s 2 Im t        (ny odd)
    const double sgn_even = ((ny / 2) & 1) ? -1.0 : 1.0;
    const double sgn_odd = (((ny - 1) / 2) & 1) ? -1.0 : 1.0;
    auto use = [&](uint64_t k, const Two<V>& t) {
      const V &a = t.a, &b = t.b;
      const uint64_t j = ins0(k, msb);
      const double tr = (double)a.x * b.x + (double)a.y * b.y;  // Re conj(a) b
      const double ti = (double)a.x * b.y - (double)a.y * b.x;  // Im conj(a) b
      const double s = (__popcll(j & zmask) & 1) ? -1.0 : 1.0;
      v[0] += (ny & 1) ? sgn_odd * s * 2.0 * ti : sgn_even * s * 2.0 * tr;
      v[1] += norm2(a) + norm2(b);
    };
    if (FORM == 1) {
      // Every X/Y position among the lane bits: item k sits in lane k & 63 of its wavefront
      // (kBlock is a multiple of 64) and its partner k ^ xmask in lane (k & 63) ^ xmask of
      // the same wavefront and iteration, so each item is read once, coalesced, and the
      // partner arrives by a lane shuffle.  The per-item form of the sum above:
      // Re[conj(psi_{k^x}) psi_k i^ny] (-1)^popc(k & z).  Items below N are whole
      // wavefronts, or for N < 64 lanes whose partners are also below N.
      const int xl = (int)xmask;
      for_items<8>(N, [&](uint64_t i) { return ldg(&st[i]); },
                [&](uint64_t i, const V& a) {
                  const double bx = __shfl_xor(a.x, xl), by = __shfl_xor(a.y, xl);
                  const double tr = bx * a.x + by * a.y;  // Re conj(b) a
                  const double ti = bx * a.y - by * a.x;  // Im conj(b) a
                  const double re = (ny & 1) ? ((ny & 2) ? ti : -ti) : ((ny & 2) ? -tr : tr);
                  v[0] += (__popcll(i & zmask) & 1) ? -re : re;
                  v[1] += norm2(a);
                });
    } else {
      // Partners at least 64 items apart, each access instruction covering whole cache
      // lines: non-temporal loads (X on position 29: 7.0 TB/s against 6.2 cached).
      for_items(
          N / 2,
          [&](uint64_t k) {
            const uint64_t j = ins0(k, msb);
            return Two<V>{ldg(&st[j]), ldg(&st[j ^ xmask])};
          },
          use);
    }
  }
  block_sum_store<2>(v, partial);
}

// The form of k_red_pauli for these masks, on grid g (red_grid of N / 2 items when an X/Y
// position is 6 or above, else of N).
template <typename T>
void launch_pauli(unsigned g, const typename Vec<T>::t* st, uint64_t N, uint64_t x, uint64_t z, unsigned ny,
                  double* partial)
{
  if (x == 0)
    k_red_pauli<T, 0><<<g, kBlock, 0, g_stream>>>(st, N, x, z, ny, partial);
  else if ((x >> 6) == 0)
    k_red_pauli<T, 1><<<g, kBlock, 0, g_stream>>>(st, N, x, z, ny, partial);
  else
    k_red_pauli<T, 2><<<g, kBlock, 0, g_stream>>>(st, N, x, z, ny, partial);
}

// <x|h> = sum conj(x) h ; l2: sum |h - x|^2 ; maxabs: max |h - s x|
template <typename T>
__global__ __launch_bounds__(kBlock) void k_red_overlap(const typename Vec<T>::t* __restrict__ h,
                                                        const typename Vec<T>::t* __restrict__ x, uint64_t N,
                                                        double* partial)
{
  using V = typename Vec<T>::t;
  double v[2] = {0, 0};
  for_items(N, [&](uint64_t i) { return Two<V>{ldg(&h[i]), ldg(&x[i])}; },
            [&](uint64_t, const Two<V>& t) {
              const V &a = t.a, &b = t.b;
              v[0] += (double)b.x * a.x + (double)b.y * a.y;
              v[1] += (double)b.x * a.y - (double)b.y * a.x;
            });
  block_sum_store<2>(v, partial);
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_red_l2diff(const typename Vec<T>::t* __restrict__ h,
                                                       const typename Vec<T>::t* __restrict__ x, uint64_t N,
                                                       double* partial)
{
  using V = typename Vec<T>::t;
  double v[1] = {0};
  for_items(N, [&](uint64_t i) { return Two<V>{ldg(&h[i]), ldg(&x[i])}; },
            [&](uint64_t, const Two<V>& t) {
              const double dr = (double)t.a.x - t.b.x, di = (double)t.a.y - t.b.y;
              v[0] += dr * dr + di * di;
            });
  block_sum_store<1>(v, partial);
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_red_maxabs(const typename Vec<T>::t* __restrict__ h,
                                                       const typename Vec<T>::t* __restrict__ x, uint64_t N,
                                                       double sr, double si, double* partial)
{
  using V = typename Vec<T>::t;
  double m = -1.0;
  for_items(N, [&](uint64_t i) { return Two<V>{ldg(&h[i]), ldg(&x[i])}; },
            [&](uint64_t, const Two<V>& t) {
              const V &a = t.a, &b = t.b;
              const double dr = (double)a.x - (sr * b.x - si * b.y), di = (double)a.y - (sr * b.y + si * b.x);
              m = fmax(m, sqrt(dr * dr + di * di));
            });
  block_max_store(m, partial);
}

// ---- QAOA kernels ----
template <typename T>
__global__ __launch_bounds__(kBlock) void k_qaoa_phase(typename Vec<T>::t* __restrict__ st, uint64_t N, CutSpec cs)
{
  __shared__ uint64_t tab[8][256];
  __shared__ double wt[8][256];
  __shared__ Cx<T> ph[65];
  build_tables(tab, cs.inc);
  if (cs.weighted) build_weight_tables(wt, cs);
  for (unsigned c = threadIdx.x; c <= cs.ne && c < 65; c += kBlock) ph[c] = Cx<T>{(T)cs.ph_re[c], (T)cs.ph_im[c]};
  __syncthreads();
  for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < N; i += (uint64_t)gridDim.x * kBlock) {
    int ic;
    const double c = cut_value(tab, wt, cs, i, &ic);
    Cx<T> f;
    if (!cs.weighted) f = ph[ic];
    else {
      double s, co;
      sincos(cs.gamma * c, &s, &co);
      f = Cx<T>{(T)co, (T)-s};
    }
    st[i] = cmul(f, st[i]);
  }
}

// Unit-weight cut diagonal without tables: rows of kBlock * 4 amplitudes, thread t holding
// in-row offsets t + j kBlock.  The cut edges E(x) are linear over GF(2) in the bits of x,
// E(row | lo) = E(row) ^ E(lo): each thread takes E(lo) of its four offsets once and E(row)
// from the row's set bits, and its four loads of a row are in flight together.  MODE 0:
// psi *= ph[cut] (as k_qaoa_phase); MODE 1: the sum of cut |psi|^2 and, if write_lam,
// lam := cut psi (as k_qaoa_cost).  One contiguous, XCD-contiguous chunk of rows per
// workgroup; N >= kBlock * 4.
template <typename T, int MODE>
__global__ __launch_bounds__(kBlock) void k_qaoa_diag_rows(typename Vec<T>::t* __restrict__ st,
                                                           typename Vec<T>::t* __restrict__ lam, uint64_t N, CutSpec cs,
                                                           int write_lam, double* __restrict__ partial)
{
  using V = typename Vec<T>::t;
  constexpr int U = 4, RB = 10;
  static_assert(kBlock * U == (1 << RB), "row size");
  __shared__ Cx<T> ph[65];
  if (MODE == 0)
    for (unsigned c = threadIdx.x; c <= cs.ne && c < 65; c += kBlock) ph[c] = Cx<T>{(T)cs.ph_re[c], (T)cs.ph_im[c]};
  uint64_t elo[U];
#pragma unroll
  for (int j = 0; j < U; ++j) {
    const unsigned xl = threadIdx.x + j * kBlock;
    uint64_t e = 0;
#pragma unroll
    for (int b = 0; b < RB; ++b)
      if ((xl >> b) & 1) e ^= cs.inc.c[b];
    elo[j] = e;
  }
  __syncthreads();
  const uint64_t rows = N >> RB, G = gridDim.x;
  const uint64_t lb = (G & 7) ? blockIdx.x : xcd_block(blockIdx.x, G);
  const uint64_t chunk = (rows + G - 1) / G;
  const uint64_t b0 = lb * chunk, b1 = (b0 + chunk < rows) ? b0 + chunk : rows;
  double v[1] = {0};
  for (uint64_t rb = b0; rb < b1; ++rb) {
    uint64_t ehi = 0;
    for (uint64_t r = rb; r; r &= r - 1) ehi ^= cs.inc.c[RB + __builtin_ctzll(r)];
    const uint64_t r0 = (rb << RB) + threadIdx.x;
    V a[U];
#pragma unroll
    for (int j = 0; j < U; ++j) a[j] = ldg(&st[r0 + (uint64_t)j * kBlock]);
#pragma unroll
    for (int j = 0; j < U; ++j) {
      const int ic = __popcll(ehi ^ elo[j]);
      if (MODE == 0) {
        stg(cmul(ph[ic], a[j]), &st[r0 + (uint64_t)j * kBlock]);
      } else {
        const double c = (double)ic;
        v[0] += c * norm2(a[j]);
        if (write_lam) {
          V l;
          l.x = (T)(c * a[j].x);
          l.y = (T)(c * a[j].y);
          stg(l, &lam[r0 + (uint64_t)j * kBlock]);
        }
      }
    }
  }
  if (MODE == 1) block_sum_store<1>(v, partial);
}

// mode 0: sum cut |psi|^2 ; mode 1: also lam := cut * psi
template <typename T>
__global__ __launch_bounds__(kBlock) void k_qaoa_cost(const typename Vec<T>::t* __restrict__ st,
                                                      typename Vec<T>::t* __restrict__ lam, uint64_t N, CutSpec cs,
                                                      int write_lam, double* partial)
{
  __shared__ uint64_t tab[8][256];
  __shared__ double wt[8][256];
  build_tables(tab, cs.inc);
  if (cs.weighted) build_weight_tables(wt, cs);
  __syncthreads();
  double v[1] = {0};
  for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < N; i += (uint64_t)gridDim.x * kBlock) {
    int ic;
    const double c = cut_value(tab, wt, cs, i, &ic);
    const auto a = st[i];
    v[0] += c * norm2(a);
    if (write_lam) {
      typename Vec<T>::t l;
      l.x = (T)(c * a.x);
      l.y = (T)(c * a.y);
      lam[i] = l;
    }
  }
  block_sum_store<1>(v, partial);
}

// grad partial: sum Im(conj(lam) c psi); then psi, lam *= exp(+i gamma c)
template <typename T>
__global__ __launch_bounds__(kBlock) void k_qaoa_phase_adj(typename Vec<T>::t* __restrict__ st,
                                                           typename Vec<T>::t* __restrict__ lam, uint64_t N, CutSpec cs,
                                                           double* partial)
{
  __shared__ uint64_t tab[8][256];
  __shared__ double wt[8][256];
  __shared__ Cx<T> ph[65];
  build_tables(tab, cs.inc);
  if (cs.weighted) build_weight_tables(wt, cs);
  for (unsigned c = threadIdx.x; c <= cs.ne && c < 65; c += kBlock) ph[c] = Cx<T>{(T)cs.ph_re[c], (T)cs.ph_im[c]};
  __syncthreads();
  double v[1] = {0};
  for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < N; i += (uint64_t)gridDim.x * kBlock) {
    int ic;
    const double c = cut_value(tab, wt, cs, i, &ic);
    const auto a = st[i], l = lam[i];
    v[0] += c * ((double)l.x * a.y - (double)l.y * a.x);
    Cx<T> f;
    if (!cs.weighted) f = ph[ic];
    else {
      double s, co;
      sincos(cs.gamma * c, &s, &co);
      f = Cx<T>{(T)co, (T)-s};
    }
    st[i] = cmul(f, a);
    lam[i] = cmul(f, l);
  }
  block_sum_store<1>(v, partial);
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_qaoa_diag(typename Vec<T>::t* __restrict__ st,
                                                      const typename Vec<T>::t* __restrict__ diag, uint64_t N, T gamma)
{
  for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < N; i += (uint64_t)gridDim.x * kBlock) {
    const T x = gamma * diag[i].x;
    T s, c;
    if (sizeof(T) == 8) {
      double sd, cd;
      sincos((double)x, &sd, &cd);
      s = (T)sd;
      c = (T)cd;
    } else {
      float sf, cf;
      sincosf((float)x, &sf, &cf);
      s = (T)sf;
      c = (T)cf;
    }
    st[i] = cmul(Cx<T>{c, -s}, st[i]);
  }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_qaoa_diag_expect(const typename Vec<T>::t* __restrict__ st,
                                                             const typename Vec<T>::t* __restrict__ diag, uint64_t N,
                                                             int squared, double* partial)
{
  using V = typename Vec<T>::t;
  double v[1] = {0};
  for_items(N, [&](uint64_t i) { return Two<V>{ldg(&st[i]), ldg(&diag[i])}; },
            [&](uint64_t, const Two<V>& t) {
              const double d = t.b.x;
              v[0] += (squared ? d * d : d) * norm2(t.a);
            });
  block_sum_store<1>(v, partial);
}

// diag_i := cut_i over the edges of cs (accumulate: diag_i += that), and the maximum of
// the values written, per workgroup.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_maxcut_fill(typename Vec<T>::t* __restrict__ diag, uint64_t N, CutSpec cs,
                                                        int accumulate, double* partial)
{
  __shared__ uint64_t tab[8][256];
  __shared__ double wt[8][256];
  build_tables(tab, cs.inc);
  if (cs.weighted) build_weight_tables(wt, cs);
  __syncthreads();
  double m = -1.0;
  for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < N; i += (uint64_t)gridDim.x * kBlock) {
    int ic;
    const double c = cut_value(tab, wt, cs, i, &ic);
    const double tot = accumulate ? c + (double)diag[i].x : c;
    typename Vec<T>::t d;
    d.x = (T)tot;
    d.y = 0;
    diag[i] = d;
    m = fmax(m, tot);
  }
  block_max_store(m, partial);
}

// ---- adjoint 1q step on two states ----
template <typename T>
__global__ __launch_bounds__(kBlock) void k_adj_1q(typename Vec<T>::t* __restrict__ st, typename Vec<T>::t* __restrict__ lam,
                                                   uint64_t items, unsigned pos, Mat2<T> uinv, Mat2<double> g,
                                                   int with_grad, double* partial)
{
  double v[1] = {0};
  const uint64_t hb = uint64_t(1) << pos;
  for (uint64_t k = (uint64_t)blockIdx.x * kBlock + threadIdx.x; k < items; k += (uint64_t)gridDim.x * kBlock) {
    const uint64_t i0 = ins0(k, pos), i1 = i0 | hb;
    const auto p0 = st[i0], p1 = st[i1], l0 = lam[i0], l1 = lam[i1];
    if (with_grad) {
      // (G psi) components in double
      const double g0r = (g.m00.r * p0.x - g.m00.i * p0.y) + (g.m01.r * p1.x - g.m01.i * p1.y);
      const double g0i = (g.m00.r * p0.y + g.m00.i * p0.x) + (g.m01.r * p1.y + g.m01.i * p1.x);
      const double g1r = (g.m10.r * p0.x - g.m10.i * p0.y) + (g.m11.r * p1.x - g.m11.i * p1.y);
      const double g1i = (g.m10.r * p0.y + g.m10.i * p0.x) + (g.m11.r * p1.y + g.m11.i * p1.x);
      v[0] += ((double)l0.x * g0i - (double)l0.y * g0r) + ((double)l1.x * g1i - (double)l1.y * g1r);
    }
    st[i0] = cmad2(uinv.m00, p0, uinv.m01, p1);
    st[i1] = cmad2(uinv.m10, p0, uinv.m11, p1);
    lam[i0] = cmad2(uinv.m00, l0, uinv.m01, l1);
    lam[i1] = cmad2(uinv.m10, l0, uinv.m11, l1);
  }
  block_sum_store<1>(v, partial);
}

// Finish a reduction: second stage + copy nv doubles to host (synchronizes the stream).
int red_finish(RedWork* w, unsigned nblocks, unsigned nv, int is_max, double* out)
{
  k_red_final<<<1, kBlock, 0, g_stream>>>(w->partial, nblocks, nv, is_max, w->result);
  CHECK_LAUNCH();
  HIP_TRY(hipMemcpyAsync(w->host, w->result, sizeof(double) * nv, hipMemcpyDeviceToHost, g_stream));
  HIP_TRY(hipStreamSynchronize(g_stream));
  for (unsigned r = 0; r < nv; ++r) out[r] = w->host[r];
  return IQSH_OK;
}

int build_cut(iqsh_handle h, unsigned ne, const unsigned* u, const unsigned* v, const double* w, double gamma,
              CutSpec* cs)
{
  if (ne > 64) return fail(IQSH_ERR_NOTSUP, "at most 64 edges are supported (got %u)", ne);
  std::memset(cs, 0, sizeof *cs);
  cs->ne = ne;
  cs->gamma = gamma;
  cs->inc.nbytes = (h->n + 7) / 8;
  bool weighted = false;
  for (unsigned e = 0; e < ne; ++e) {
    if (u[e] >= h->n || v[e] >= h->n || u[e] == v[e]) return fail(IQSH_ERR_ARG, "bad edge %u: (%u, %u)", e, u[e], v[e]);
    cs->inc.c[u[e]] ^= uint64_t(1) << e;
    cs->inc.c[v[e]] ^= uint64_t(1) << e;
    const double we = w ? w[e] : 1.0;
    cs->w[e] = we;
    weighted |= (we != 1.0);
  }
  cs->weighted = weighted;
  // exp(-i gamma c) for integer cuts, evaluated as the reference does per amplitude:
  // (cos(gamma * c), -sin(gamma * c)) in the register's precision.
  for (unsigned c = 0; c <= ne; ++c) {
    if (h->dtype == IQSH_C128) {
      cs->ph_re[c] = std::cos(gamma * (double)c);
      cs->ph_im[c] = -std::sin(gamma * (double)c);
    } else {
      const float x = (float)gamma * (float)c;
      cs->ph_re[c] = std::cos(x);
      cs->ph_im[c] = -std::sin(x);
    }
  }
  return IQSH_OK;
}

}  // namespace

// =======================================================================================
// C ABI

#define DISPATCH(h, CALL)                                   \
  ((h)->dtype == IQSH_C128 ? CALL<double> : CALL<float>)

namespace iqsi {
int fail(int code, const char* fmt, ...)
{
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof g_err, fmt, ap);
  va_end(ap);
  return code;
}
hipStream_t stream() { return g_stream; }
int device_cus()
{
  int dev = 0, cus = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
  return cus;
}
int build_cut_spec(iqsh_handle h, unsigned ne, const unsigned* u, const unsigned* v, const double* w, double gamma,
                   iqsk::CutSpec* cs)
{
  return build_cut(h, ne, u, v, w, gamma, cs);
}
}  // namespace iqsi

// Definitions take C linkage from their declarations in iqs_hip.h.

const char* iqsh_last_error(void) { return g_err; }

int iqsh_set_stream(void* stream)
{
  g_stream = (hipStream_t)stream;
  return IQSH_OK;
}

int iqsh_set_block_mfma(int mode)
{
  if (mode < 0 || mode > 2) return fail(IQSH_ERR_ARG, "block kernel mode %d (0, 1 or 2)", mode);
  g_block_mfma = mode;
  return IQSH_OK;
}

int iqsh_sync(void)
{
  HIP_TRY(hipStreamSynchronize(g_stream));
  return IQSH_OK;
}

int iqsh_device_cus(int* cus)
{
  int dev = 0;
  HIP_TRY(hipGetDevice(&dev));
  HIP_TRY(hipDeviceGetAttribute(cus, hipDeviceAttributeMultiprocessorCount, dev));
  return IQSH_OK;
}

static hipEvent_t g_t0 = nullptr, g_t1 = nullptr;

int iqsh_timer_start(void)
{
  if (!g_t0) {
    HIP_TRY(hipEventCreate(&g_t0));
    HIP_TRY(hipEventCreate(&g_t1));
  }
  HIP_TRY(hipEventRecord(g_t0, g_stream));
  return IQSH_OK;
}

int iqsh_timer_stop(double* ms)
{
  if (!g_t0 || !ms) return fail(IQSH_ERR_ARG, "timer not started");
  HIP_TRY(hipEventRecord(g_t1, g_stream));
  HIP_TRY(hipEventSynchronize(g_t1));
  float f = 0;
  HIP_TRY(hipEventElapsedTime(&f, g_t0, g_t1));
  *ms = f;
  return IQSH_OK;
}

int iqsh_create(unsigned num_qubits, int dtype, iqsh_handle* out)
{
  if (!out) return fail(IQSH_ERR_ARG, "null output handle");
  *out = nullptr;
  if (num_qubits > 40) return fail(IQSH_ERR_ARG, "num_qubits %u out of range [0, 40]", num_qubits);
  if (dtype != IQSH_C64 && dtype != IQSH_C128) return fail(IQSH_ERR_ARG, "bad dtype %d", dtype);
  auto* h = new iqsh_state;
  h->n = num_qubits;
  h->dtype = dtype;
  const size_t bytes = h->size() * h->esz();
  hipError_t e = hipMalloc(&h->d, bytes);
  if (e != hipSuccess) {
    delete h;
    (void)hipGetLastError();
    return fail(IQSH_ERR_ALLOC, "hipMalloc(%zu bytes) for %u qubits: %s", bytes, num_qubits, hipGetErrorString(e));
  }
  *out = h;
  return IQSH_OK;
}

int iqsh_wrap(unsigned num_qubits, int dtype, void* device_ptr, iqsh_handle* out)
{
  if (!out || !device_ptr) return fail(IQSH_ERR_ARG, "null pointer");
  if (num_qubits < 1 || num_qubits > 40) return fail(IQSH_ERR_ARG, "num_qubits %u out of range", num_qubits);
  if (dtype != IQSH_C64 && dtype != IQSH_C128) return fail(IQSH_ERR_ARG, "bad dtype %d", dtype);
  auto* h = new iqsh_state;
  h->n = num_qubits;
  h->dtype = dtype;
  h->d = device_ptr;
  h->owned = false;
  *out = h;
  return IQSH_OK;
}

int iqsh_destroy(iqsh_handle h)
{
  if (!h) return IQSH_OK;
  HIP_TRY(hipStreamSynchronize(g_stream));
  if (h->owned && h->d) HIP_TRY(hipFree(h->d));
  if (h->spare) HIP_TRY(hipFree(h->spare));
  delete h;
  return IQSH_OK;
}

int iqsh_info(iqsh_handle h, unsigned* nq, int* dtype, void** ptr)
{
  if (int rc = check_h(h)) return rc;
  if (nq) *nq = h->n;
  if (dtype) *dtype = h->dtype;
  if (ptr) *ptr = h->d;
  return IQSH_OK;
}

int iqsh_upload(iqsh_handle h, const void* host, uint64_t offset, uint64_t count)
{
  if (int rc = check_h(h)) return rc;
  if (!host || offset > h->size() || count > h->size() - offset) return fail(IQSH_ERR_ARG, "upload range out of bounds");
  HIP_TRY(hipMemcpyAsync((char*)h->d + offset * h->esz(), host, count * h->esz(), hipMemcpyHostToDevice, g_stream));
  HIP_TRY(hipStreamSynchronize(g_stream));
  return IQSH_OK;
}

int iqsh_download(iqsh_handle h, void* host, uint64_t offset, uint64_t count)
{
  if (int rc = check_h(h)) return rc;
  if (!host || offset > h->size() || count > h->size() - offset) return fail(IQSH_ERR_ARG, "download range out of bounds");
  HIP_TRY(hipMemcpyAsync(host, (const char*)h->d + offset * h->esz(), count * h->esz(), hipMemcpyDeviceToHost, g_stream));
  HIP_TRY(hipStreamSynchronize(g_stream));
  return IQSH_OK;
}

int iqsh_copy(iqsh_handle dst, iqsh_handle src)
{
  if (int rc = check_h(dst)) return rc;
  if (int rc = check_h(src)) return rc;
  if (dst->n != src->n || dst->dtype != src->dtype) return fail(IQSH_ERR_ARG, "copy between different registers");
  if (dst == src) return IQSH_OK;
  HIP_TRY(hipMemcpyAsync(dst->d, src->d, src->size() * src->esz(), hipMemcpyDeviceToDevice, g_stream));
  return IQSH_OK;
}

namespace {
template <typename T>
int do_fill(iqsh_handle h, double re, double im)
{
  const unsigned grid = (unsigned)std::min<uint64_t>(grid_for(h->size(), kBlock), 65536);
  k_fill<T><<<grid, kBlock, 0, g_stream>>>((typename Vec<T>::t*)h->d, h->size(), Cx<T>{(T)re, (T)im});
  CHECK_LAUNCH();
  return IQSH_OK;
}
}  // namespace

int iqsh_convert(iqsh_handle dst, iqsh_handle src)
{
  if (int rc = check_h(dst)) return rc;
  if (int rc = check_h(src)) return rc;
  if (dst->n != src->n) return fail(IQSH_ERR_ARG, "convert between registers of different size");
  if (dst->dtype == src->dtype) return iqsh_copy(dst, src);
  const unsigned grid = (unsigned)std::min<uint64_t>(grid_for(dst->size(), kBlock), 65536);
  if (dst->dtype == IQSH_C128)
    k_convert<double, float><<<grid, kBlock, 0, g_stream>>>((Vec<double>::t*)dst->d, (const Vec<float>::t*)src->d,
                                                            dst->size());
  else
    k_convert<float, double><<<grid, kBlock, 0, g_stream>>>((Vec<float>::t*)dst->d, (const Vec<double>::t*)src->d,
                                                            dst->size());
  CHECK_LAUNCH();
  return IQSH_OK;
}

int iqsh_init_uniform(iqsh_handle h, double re, double im)
{
  if (int rc = check_h(h)) return rc;
  return DISPATCH(h, do_fill)(h, re, im);
}

int iqsh_init_basis(iqsh_handle h, uint64_t idx)
{
  if (int rc = check_h(h)) return rc;
  if (idx >= h->size()) return fail(IQSH_ERR_ARG, "basis index out of range");
  if (int rc = iqsh_init_uniform(h, 0, 0)) return rc;
  const double one[2] = {1.0, 0.0};
  const float onef[2] = {1.0f, 0.0f};
  // Small pageable copy, ordered after the fill on the same stream and waited for.
  HIP_TRY(hipMemcpyAsync((char*)h->d + idx * h->esz(), h->dtype == IQSH_C128 ? (const void*)one : (const void*)onef,
                         h->esz(), hipMemcpyHostToDevice, g_stream));
  HIP_TRY(hipStreamSynchronize(g_stream));
  return IQSH_OK;
}

int iqsh_apply_1q(iqsh_handle h, unsigned pos, const double m[8])
{
  if (int rc = check_h(h)) return rc;
  if (int rc = check_pos(h, pos)) return rc;
  if (!m) return fail(IQSH_ERR_ARG, "null matrix");
  return DISPATCH(h, controlled_2x2)(h, 0, pos, m);
}

int iqsh_spec_matrix(int dtype, int spec, int controlled, double angle, double m[8])
{
  if (dtype != IQSH_C64 && dtype != IQSH_C128) return fail(IQSH_ERR_ARG, "bad dtype %d", dtype);
  if (!m) return fail(IQSH_ERR_ARG, "null matrix");
  if (spec == IQSH_SPEC_T_OR_CPHASE && !controlled) {
    const double mt[8] = {1, 0, 0, 0, 0, 0, std::cos(M_PI / 4), std::sin(M_PI / 4)};
    std::memcpy(m, mt, 8 * sizeof(double));
  } else if (spec == IQSH_SPEC_T_OR_CPHASE) {
    // CPhase: pexp = (cos(theta), sin(theta)) in the register precision.
    if (dtype == IQSH_C128) {
      const double mp[8] = {1, 0, 0, 0, 0, 0, std::cos(angle), std::sin(angle)};
      std::memcpy(m, mp, 8 * sizeof(double));
    } else {
      const float th = (float)angle;
      const double mp[8] = {1, 0, 0, 0, 0, 0, (double)std::cos(th), (double)std::sin(th)};
      std::memcpy(m, mp, 8 * sizeof(double));
    }
  } else if (spec >= IQSH_SPEC_H && spec <= IQSH_SPEC_Z) {
    if (dtype == IQSH_C128) spec_matrix<double>(spec, angle, m);
    else spec_matrix<float>(spec, angle, m);
  } else
    return fail(IQSH_ERR_ARG, "bad spec %d", spec);
  return IQSH_OK;
}

int iqsh_apply_1q_spec(iqsh_handle h, unsigned pos, int spec, double angle)
{
  if (int rc = check_h(h)) return rc;
  if (int rc = check_pos(h, pos)) return rc;
  double m[8];
  if (int rc = iqsh_spec_matrix(h->dtype, spec, 0, angle, m)) return rc;
  return DISPATCH(h, controlled_2x2)(h, 0, pos, m);
}

namespace {
template <typename T>
int do_nest(iqsh_handle h, const Nest& nt, uint64_t items, const double* m)
{
  const Mat2<T> mm{{(T)m[0], (T)m[1]}, {(T)m[2], (T)m[3]}, {(T)m[4], (T)m[5]}, {(T)m[6], (T)m[7]}};
  k_nest<T><<<grid2(std::min<uint64_t>(grid_for(items, kBlock), uint64_t(1) << 20)), kBlock, 0, g_stream>>>(
      (typename Vec<T>::t*)h->d, nt, items, mm);
  CHECK_LAUNCH();
  return IQSH_OK;
}
}  // namespace

int iqsh_apply_nest(iqsh_handle h, const uint64_t c[9], const double m[8])
{
  if (int rc = check_h(h)) return rc;
  if (!c || !m) return fail(IQSH_ERR_ARG, "null argument");
  const uint64_t c11 = c[0], c12 = c[1], c13 = c[2], c21 = c[3], c22 = c[4], c23 = c[5], c31 = c[6], c32 = c[7],
                 shift = c[8];
  if (!c13 || !c23 || !shift) return fail(IQSH_ERR_ARG, "zero stride or shift");
  if (c12 <= c11 || c22 <= c21 || c32 <= c31) return IQSH_OK;  // empty nest: nothing to do
  Nest nt{c11, c13, c21, c23, c31, (c12 - c11 + c13 - 1) / c13, (c22 - c21 + c23 - 1) / c23, c32 - c31, shift};
  // The last index touched must lie in the register; overflow-checked in long double.
  const long double last = (long double)c11 + (long double)(nt.ni - 1) * c13 + c21 + (long double)(nt.nj - 1) * c23 +
                           c32 - 1 + shift;
  if (last >= (long double)h->size()) return fail(IQSH_ERR_ARG, "loop nest reaches index %.0Lf of %llu", last,
                                                   (unsigned long long)h->size());
  if ((long double)nt.ni * nt.nj * nt.nk > 1.9e19L) return fail(IQSH_ERR_ARG, "loop nest too large");
  return DISPATCH(h, do_nest)(h, nt, nt.ni * nt.nj * nt.nk, m);
}

int iqsh_apply_c1q(iqsh_handle h, unsigned c, unsigned t, const double m[8])
{
  if (int rc = check_h(h)) return rc;
  if (int rc = check_pos(h, c)) return rc;
  if (int rc = check_pos(h, t)) return rc;
  if (c == t) return fail(IQSH_ERR_ARG, "control and target coincide");
  if (!m) return fail(IQSH_ERR_ARG, "null matrix");
  return DISPATCH(h, controlled_2x2)(h, uint64_t(1) << c, t, m);
}

int iqsh_apply_c1q_spec(iqsh_handle h, unsigned c, unsigned t, int spec, double angle)
{
  if (int rc = check_h(h)) return rc;
  if (int rc = check_pos(h, c)) return rc;
  if (int rc = check_pos(h, t)) return rc;
  if (c == t) return fail(IQSH_ERR_ARG, "control and target coincide");
  double m[8];
  if (int rc = iqsh_spec_matrix(h->dtype, spec, 1, angle, m)) return rc;
  return DISPATCH(h, controlled_2x2)(h, uint64_t(1) << c, t, m);
}

int iqsh_apply_mc1q(iqsh_handle h, uint64_t ctrl, unsigned t, const double m[8])
{
  if (int rc = check_h(h)) return rc;
  if (int rc = check_pos(h, t)) return rc;
  if (ctrl >> h->n) return fail(IQSH_ERR_ARG, "control mask exceeds register");
  if ((ctrl >> t) & 1) return fail(IQSH_ERR_ARG, "target is also a control");
  if (__builtin_popcountll(ctrl) > 3) return fail(IQSH_ERR_NOTSUP, "at most 3 controls are supported");
  if (!m) return fail(IQSH_ERR_ARG, "null matrix");
  return DISPATCH(h, controlled_2x2)(h, ctrl, t, m);
}

namespace {
// Dense 4x4 with row index 2 bit(pa) + bit(pb) through the lane-shuffle kernel; needs
// min(pa, pb) < 6 and a tiled-size register.
template <typename T>
int launch_quad_lane(iqsh_handle h, unsigned pa, unsigned pb, const Mat4<T>& M, bool move)
{
  using V = typename Vec<T>::t;
  const unsigned hi = std::max(pa, pb);
  const uint64_t items = hi >= 6 ? h->size() >> 1 : h->size();
  using Q = QuadLaneTile<T>;
  if (hi >= 6) {
    const uint64_t grid = items / (Q::B * Q::P);
    if (move)
      k_quad_lane<T, true, Q::B, Q::P><<<grid2(grid), Q::B, 0, g_stream>>>((V*)h->d, pa, pb, M);
    else
      k_quad_lane<T, false, Q::B, Q::P><<<grid2(grid), Q::B, 0, g_stream>>>((V*)h->d, pa, pb, M);
    CHECK_LAUNCH();
    return IQSH_OK;
  }
  const uint64_t grid = items / (Q::BL * Q::PL);
  if (move)
    k_quad_lane<T, true, Q::BL, Q::PL><<<grid2(grid), Q::BL, 0, g_stream>>>((V*)h->d, pa, pb, M);
  else
    k_quad_lane<T, false, Q::BL, Q::PL><<<grid2(grid), Q::BL, 0, g_stream>>>((V*)h->d, pa, pb, M);
  CHECK_LAUNCH();
  return IQSH_OK;
}

template <typename T>
int do_swaplike(iqsh_handle h, unsigned p1, unsigned p2, const double* m)
{
  const unsigned ins[2] = {p1, p2};
  const bool xmove = is_c(m, 0, 0) && is_c(m + 6, 0, 0) && is_c(m + 2, 1, 0) && is_c(m + 4, 1, 0);
  if (is_c(m, 1, 0) && is_c(m + 6, 1, 0) && is_c(m + 2, 0, 0) && is_c(m + 4, 0, 0)) return IQSH_OK;
  if (std::min(p1, p2) < line_bits<T>() && h->size() >= (kTiledMin << 3)) {
    // A position inside a 128-byte line: every line holds moved and unmoved amplitudes, so
    // all of them are read and written anyway; the lane-shuffle 4x4 does it in whole runs
    // instead of 16-byte accesses spread over the lines (measured 30q complex128: 1.9 ->
    // 2.8 TB/s algorithmic, whose ceiling is half the stream rate).  Rows 2 bit(p1) +
    // bit(p2): m acts on (row 2, row 1), the other two rows are the identity.
    Mat4<T> M;
    for (int i = 0; i < 16; ++i) M.m[i] = Cx<T>{0, 0};
    M.m[0] = M.m[15] = Cx<T>{1, 0};
    M.m[4 * 2 + 2] = cx<T>(m);
    M.m[4 * 2 + 1] = cx<T>(m + 2);
    M.m[4 * 1 + 2] = cx<T>(m + 4);
    M.m[4 * 1 + 1] = cx<T>(m + 6);
    return launch_quad_lane<T>(h, p1, p2, M, xmove);
  }
  return launch_pair<T, 2>(h, sorted_ins<2>(ins), uint64_t(1) << p1, uint64_t(1) << p2, m, xmove);
}
}  // namespace

int iqsh_apply_swaplike(iqsh_handle h, unsigned p1, unsigned p2, const double m[8])
{
  if (int rc = check_h(h)) return rc;
  if (int rc = check_pos(h, p1)) return rc;
  if (int rc = check_pos(h, p2)) return rc;
  if (p1 == p2) return fail(IQSH_ERR_ARG, "swap positions coincide");
  if (!m) return fail(IQSH_ERR_ARG, "null matrix");
  if (p1 > p2) std::swap(p1, p2);
  return DISPATCH(h, do_swaplike)(h, p1, p2, m);
}

namespace {
template <typename T>
int do_diag2(iqsh_handle h, unsigned p1, unsigned p2, const double* d)
{
  // Entries exactly 1 leave their amplitudes untouched: with one non-unit entry only a
  // quarter of the state is read and written.
  int nonunit = 0, which = -1;
  for (int k = 0; k < 4; ++k)
    if (!is_c(d + 2 * k, 1, 0)) {
      ++nonunit;
      which = k;
    }
  if (nonunit == 0) return IQSH_OK;
  if (nonunit == 1) {
    const uint64_t mask = (((uint64_t)(which >> 1)) << p1) | (((uint64_t)(which & 1)) << p2);
    const unsigned ins[2] = {p1, p2};
    return launch_scale<T, 2>(h, sorted_ins<2>(ins), mask, d + 2 * which);
  }
  return launch_diag2<T>(h, p1, p2, cx<T>(d), cx<T>(d + 2), cx<T>(d + 4), cx<T>(d + 6));
}
}  // namespace

int iqsh_apply_diag2(iqsh_handle h, unsigned p1, unsigned p2, const double d[8])
{
  if (int rc = check_h(h)) return rc;
  if (int rc = check_pos(h, p1)) return rc;
  if (int rc = check_pos(h, p2)) return rc;
  if (p1 == p2) return fail(IQSH_ERR_ARG, "diag positions coincide");
  if (!d) return fail(IQSH_ERR_ARG, "null diagonal");
  return DISPATCH(h, do_diag2)(h, p1, p2, d);
}

namespace {
template <typename T>
int do_2q(iqsh_handle h, unsigned ph, unsigned pl, const double* m)
{
  using V = typename Vec<T>::t;
  Mat4<T> M;
  for (int i = 0; i < 16; ++i) M.m[i] = cx<T>(m + 2 * i);
  // A position inside a 128-byte line: lane-shuffle 4x4 (30q complex128, (0,1): 1.6 -> 5.5 TB/s);
  // from the line width up the quad kernel's own accesses fill whole lines.
  if (std::min(ph, pl) < line_bits<T>() && h->size() >= (kTiledMin << 3)) return launch_quad_lane<T>(h, ph, pl, M, false);
  const unsigned ins[2] = {ph, pl};
  const uint64_t items = h->size() >> 2;
  constexpr int TB = PairTile<T>::QB, TP = PairTile<T>::QP;
  if (items >= kTiledMin && std::max(ph, pl) <= 10)
    // Both positions <= 10: 1024 x 1 blocks (30q complex128: 6.07-6.23 TB/s against 6.02-6.08
    // with the small blocks, which are faster above).
    k_quad<T, true, kTB, 1><<<grid2(items / kTB), kTB, 0, g_stream>>>((V*)h->d, items, sorted_ins<2>(ins),
                                                                           uint64_t(1) << pl, uint64_t(1) << ph, M);
  else if (items >= kTiledMin)
    k_quad<T, true, TB, TP><<<grid2(items / (TB * TP)), TB, 0, g_stream>>>((V*)h->d, items, sorted_ins<2>(ins),
                                                                            uint64_t(1) << pl, uint64_t(1) << ph, M);
  else
    k_quad<T, false, kSB, 2><<<grid2(grid_for(items, kSB * 2)), kSB, 0, g_stream>>>((V*)h->d, items, sorted_ins<2>(ins),
                                                                            uint64_t(1) << pl, uint64_t(1) << ph, M);
  CHECK_LAUNCH();
  return IQSH_OK;
}
}  // namespace

int iqsh_apply_2q(iqsh_handle h, unsigned ph, unsigned pl, const double m[32])
{
  if (int rc = check_h(h)) return rc;
  if (int rc = check_pos(h, ph)) return rc;
  if (int rc = check_pos(h, pl)) return rc;
  if (ph == pl) return fail(IQSH_ERR_ARG, "2q positions coincide");
  if (!m) return fail(IQSH_ERR_ARG, "null matrix");
  return DISPATCH(h, do_2q)(h, ph, pl, m);
}

namespace {
template <typename T, int KQ>
int do_block_k(iqsh_handle h, const unsigned* pos, const double* U)
{
  using V = typename Vec<T>::t;
  constexpr int D = 1 << KQ;
  // Reorder rows/columns so that row bit j corresponds to the j-th smallest position.
  unsigned order[KQ];
  for (int j = 0; j < KQ; ++j) order[j] = j;
  std::sort(order, order + KQ, [&](unsigned a, unsigned b) { return pos[a] < pos[b]; });
  Ins<KQ> ins;
  for (int j = 0; j < KQ; ++j) ins.p[j] = pos[order[j]];
  // caller row index r has bit j <-> pos[j]; sorted index s has bit j <-> ins.p[j] = pos[order[j]]
  auto to_caller = [&](int s) {
    int r = 0;
    for (int j = 0; j < KQ; ++j)
      if ((s >> j) & 1) r |= 1 << order[j];
    return r;
  };
  std::vector<T> buf(2 * D * D);
  for (int rs = 0; rs < D; ++rs)
    for (int cs = 0; cs < D; ++cs) {
      const int r = to_caller(rs), c = to_caller(cs);
      buf[2 * (rs * D + cs)] = (T)U[2 * (r * D + c)];
      buf[2 * (rs * D + cs) + 1] = (T)U[2 * (r * D + c) + 1];
    }
  void* dU;
  int slot;
  if (int rc = ring_put(buf.data(), buf.size() * sizeof(T), &dU, &slot)) return rc;
  const uint64_t groups = h->size() >> KQ;
  if (KQ == 5 && groups >= 16 && g_block_mfma) {
    // Matrix-core path: tiles of 16 groups, kTilesPerWave tiles per wavefront.
    constexpr int kTilesPerWave = 8;
    const uint64_t tiles = groups / 16;
    const uint64_t waves = (tiles + kTilesPerWave - 1) / kTilesPerWave;
    Ins<5> ins5;
    for (int j = 0; j < 5; ++j) ins5.p[j] = ins.p[j < KQ ? j : 0];
    const dim3 grid = grid2((waves + 3) / 4);
    const bool nt = ins5.p[0] >= line_bits<T>();
    // Default: three real products, the matrices in LDS.  Mode 2, the 64 x 64 real form:
    // complex128 with the matrix in LDS and the next tile prefetched (160 VGPRs instead of
    // 224), complex64 with the matrix in registers (136 VGPRs).
    if (g_block_mfma == 1 && nt)
      k_block5_mfma3<T, kTilesPerWave, true><<<grid, kBlock, 0, g_stream>>>((V*)h->d, tiles, ins5, (const V*)dU);
    else if (g_block_mfma == 1)
      k_block5_mfma3<T, kTilesPerWave, false><<<grid, kBlock, 0, g_stream>>>((V*)h->d, tiles, ins5, (const V*)dU);
    else if (sizeof(T) == 8 && nt)
      k_block5_mfma_lds<T, kTilesPerWave, true, true><<<grid, kBlock, 0, g_stream>>>((V*)h->d, tiles, ins5, (const V*)dU);
    else if (sizeof(T) == 8)
      k_block5_mfma_lds<T, kTilesPerWave, false, true><<<grid, kBlock, 0, g_stream>>>((V*)h->d, tiles, ins5, (const V*)dU);
    else if (nt)
      k_block5_mfma<T, kTilesPerWave, true><<<grid, kBlock, 0, g_stream>>>((V*)h->d, tiles, ins5, (const V*)dU);
    else
      k_block5_mfma<T, kTilesPerWave, false><<<grid, kBlock, 0, g_stream>>>((V*)h->d, tiles, ins5, (const V*)dU);
  } else {
    constexpr int G = kBlock / D;
    k_block<T, KQ><<<grid2(grid_for(groups, G)), kBlock, 0, g_stream>>>((V*)h->d, groups, ins, (const V*)dU);
  }
  CHECK_LAUNCH();
  return ring_done(slot);
}

template <typename T>
int do_block(iqsh_handle h, unsigned k, const unsigned* pos, const double* U)
{
  // 3- and 4-qubit blocks are padded to 5 qubits, U (x) I on the highest free positions,
  // and run on the matrix cores: the vector-ALU gather kernel streams ~0.55 TB/s, the
  // MFMA kernel 4-5 TB/s, and both move 2S bytes.
  if ((k == 3 || k == 4) && g_block_mfma && h->n >= 9) {
    unsigned p5[5];
    uint64_t used = 0;
    for (unsigned j = 0; j < k; ++j) {
      p5[j] = pos[j];
      used |= uint64_t(1) << pos[j];
    }
    unsigned m = k;
    for (int q = (int)h->n - 1; q >= 0 && m < 5; --q)
      if (!((used >> q) & 1)) p5[m++] = (unsigned)q;
    const int dk = 1 << k;
    std::vector<double> U5(2 * 32 * 32, 0.0);
    for (int r = 0; r < 32; ++r)
      for (int c = 0; c < 32; ++c)
        if ((r >> k) == (c >> k)) {
          U5[2 * (r * 32 + c)] = U[2 * ((r & (dk - 1)) * dk + (c & (dk - 1)))];
          U5[2 * (r * 32 + c) + 1] = U[2 * ((r & (dk - 1)) * dk + (c & (dk - 1))) + 1];
        }
    return do_block_k<T, 5>(h, p5, U5.data());
  }
  switch (k) {
    case 1: return do_block_k<T, 1>(h, pos, U);
    case 2: return do_block_k<T, 2>(h, pos, U);
    case 3: return do_block_k<T, 3>(h, pos, U);
    case 4: return do_block_k<T, 4>(h, pos, U);
    case 5: return do_block_k<T, 5>(h, pos, U);
    default: return fail(IQSH_ERR_NOTSUP, "block size %u not supported", k);
  }
}
}  // namespace

int iqsh_apply_kq(iqsh_handle h, unsigned k, const unsigned* pos, const double* U)
{
  if (int rc = check_h(h)) return rc;
  if (!pos || !U) return fail(IQSH_ERR_ARG, "null argument");
  if (k < 1 || k > 5 || k > h->n) return fail(IQSH_ERR_ARG, "block size %u out of range", k);
  uint64_t seen = 0;
  for (unsigned j = 0; j < k; ++j) {
    if (int rc = check_pos(h, pos[j])) return rc;
    if ((seen >> pos[j]) & 1) return fail(IQSH_ERR_ARG, "repeated position %u", pos[j]);
    seen |= uint64_t(1) << pos[j];
  }
  return DISPATCH(h, do_block)(h, k, pos, U);
}

int iqsh_scale(iqsh_handle h, double re, double im)
{
  if (int rc = check_h(h)) return rc;
  if (re == 1.0 && im == 0.0) return IQSH_OK;
  const double f[2] = {re, im};
  std::vector<unsigned> none;
  return DISPATCH(h, scale_dispatch)(h, none, 0, f);
}

namespace {
template <typename T>
int do_axpy(iqsh_handle h, iqsh_handle x, double re, double im)
{
  using V = typename Vec<T>::t;
  const unsigned grid = (unsigned)std::min<uint64_t>(grid_for(h->size(), kBlock), 65536);
  k_axpy<T><<<grid, kBlock, 0, g_stream>>>((V*)h->d, (const V*)x->d, h->size(), Cx<T>{(T)re, (T)im});
  CHECK_LAUNCH();
  return IQSH_OK;
}
}  // namespace

int iqsh_axpy(iqsh_handle h, iqsh_handle x, double re, double im)
{
  if (int rc = check_h(h)) return rc;
  if (int rc = check_h(x)) return rc;
  if (h->n != x->n || h->dtype != x->dtype) return fail(IQSH_ERR_ARG, "axpy between different registers");
  return DISPATCH(h, do_axpy)(h, x, re, im);
}

int iqsh_collapse(iqsh_handle h, unsigned pos, int value)
{
  if (int rc = check_h(h)) return rc;
  if (int rc = check_pos(h, pos)) return rc;
  const double zero[2] = {0, 0};
  std::vector<unsigned> ins{pos};
  const uint64_t mask = value ? 0 : (uint64_t(1) << pos);
  return DISPATCH(h, scale_dispatch)(h, ins, mask, zero);
}

namespace {
template <typename T>
int do_permute(iqsh_handle h, const unsigned* src_pos)
{
  using V = typename Vec<T>::t;
  if (!h->owned) return fail(IQSH_ERR_NOTSUP, "permutation of a wrapped (non-owned) state");
  if (!h->spare) {
    hipError_t e = hipMalloc(&h->spare, h->size() * h->esz());
    if (e != hipSuccess) {
      (void)hipGetLastError();
      return fail(IQSH_ERR_ALLOC, "permutation buffer: %s", hipGetErrorString(e));
    }
  }
  // Tile bits: the L = perm_low<T>() lowest destination bits (runs of 32 amplitudes) and every
  // destination bit whose source position is below L.  If that adds bits (low positions
  // move up), the tiled transpose reads and writes whole runs; otherwise the gather's reads
  // are already runs.
  constexpr unsigned L = perm_low<T>();
  const unsigned n = h->n;
  bool in_tile[64] = {};
  for (unsigned b = 0; b < n; ++b) in_tile[b] = b < L || src_pos[b] < L;
  unsigned m = 0;
  for (unsigned b = 0; b < n; ++b) m += in_tile[b];
  const bool tiled = n >= 16 && m > L;
  // Pad the tile to 2L bits with the lowest remaining destination bits, so every workgroup
  // moves 2^(2L) amplitudes (a 6-bit complex128 tile of 64 left 3/4 of the threads idle
  // and ran at 0.85 TB/s).
  for (unsigned b = 0; tiled && b < n && m < 2 * L; ++b)
    if (!in_tile[b]) {
      in_tile[b] = true;
      ++m;
    }
  if (tiled) {
    // k_permute_tiled moves exactly 2^(2L) amplitudes per workgroup (n >= 16 >= 2L).
    if (m != 2 * L) return fail(IQSH_ERR_ARG, "permutation tile of %u bits (%u expected)", m, 2 * L);
    PermTile pt;
    std::memset(&pt, 0, sizeof pt);
    pt.m = m;
    unsigned k = 0, o = 0, tb[16];
    for (unsigned b = 0; b < n; ++b) {
      if (in_tile[b]) {
        tb[k] = b;
        pt.tj[k++] = uint64_t(1) << b;
      } else {
        pt.oj[o] = uint64_t(1) << b;
        pt.os[o++] = uint64_t(1) << src_pos[b];
      }
    }
    pt.nout = o;
    unsigned order[16];
    for (unsigned i = 0; i < m; ++i) order[i] = i;
    std::sort(order, order + m, [&](unsigned a, unsigned b) { return src_pos[tb[a]] < src_pos[tb[b]]; });
    for (unsigned i = 0; i < m; ++i) {
      pt.sd[i] = 1u << order[i];
      pt.ss[i] = uint64_t(1) << src_pos[tb[order[i]]];
    }
    constexpr int TB = perm_threads<T>();
    k_permute_tiled<T, TB><<<grid2(uint64_t(1) << (n - m)), TB, 0, g_stream>>>((V*)h->spare, (const V*)h->d, pt);
  } else {
    BitTab bt;
    std::memset(&bt, 0, sizeof bt);
    bt.nbytes = (n + 7) / 8;
    for (unsigned b = 0; b < n; ++b) bt.c[b] = uint64_t(1) << src_pos[b];
    const unsigned grid = (unsigned)std::min<uint64_t>(grid_for(h->size(), kBlock * 8), 65536);
    k_permute<T><<<grid, kBlock, 0, g_stream>>>((V*)h->spare, (const V*)h->d, h->size(), bt);
  }
  CHECK_LAUNCH();
  std::swap(h->d, h->spare);
  return IQSH_OK;
}
}  // namespace

int iqsh_permute(iqsh_handle h, const unsigned* src_pos)
{
  if (int rc = check_h(h)) return rc;
  if (!src_pos) return fail(IQSH_ERR_ARG, "null permutation");
  uint64_t seen = 0;
  bool ident = true;
  for (unsigned b = 0; b < h->n; ++b) {
    if (int rc = check_pos(h, src_pos[b])) return rc;
    if ((seen >> src_pos[b]) & 1) return fail(IQSH_ERR_ARG, "not a permutation");
    seen |= uint64_t(1) << src_pos[b];
    ident &= (src_pos[b] == b);
  }
  if (ident) return IQSH_OK;
  return DISPATCH(h, do_permute)(h, src_pos);
}

// ---- reductions ----
namespace {
template <typename T>
int do_norm_sub(iqsh_handle h, int K, unsigned pos, double* out)
{
  RedWork* w;
  std::lock_guard<std::mutex> red_lk(g_red_mu);  // held until red_finish has synchronized
  if (int rc = red_ws(&w)) return rc;
  using V = typename Vec<T>::t;
  if (K == 0) {
    const unsigned g = red_grid(h->size());
    k_red_norm<T, 0><<<g, kBlock, 0, g_stream>>>((const V*)h->d, h->size(), Ins<0>{}, 0, w->partial);
    CHECK_LAUNCH();
    return red_finish(w, g, 1, 0, out);
  }
  Ins<1> ins;
  ins.p[0] = pos;
  const unsigned g = red_grid(h->size() / 2);
  k_red_norm<T, 1><<<g, kBlock, 0, g_stream>>>((const V*)h->d, h->size() / 2, ins, uint64_t(1) << pos, w->partial);
  CHECK_LAUNCH();
  return red_finish(w, g, 1, 0, out);
}

template <typename T>
int do_pauli(iqsh_handle h, uint64_t x, uint64_t z, unsigned ny, double* out2)
{
  RedWork* w;
  std::lock_guard<std::mutex> red_lk(g_red_mu);  // held until red_finish has synchronized
  if (int rc = red_ws(&w)) return rc;
  using V = typename Vec<T>::t;
  const unsigned g = red_grid((x >> 6) ? h->size() / 2 : h->size());
  launch_pauli<T>(g, (const V*)h->d, h->size(), x, z, ny, w->partial);
  CHECK_LAUNCH();
  return red_finish(w, g, 2, 0, out2);
}

template <typename T>
int do_two(iqsh_handle h, iqsh_handle x, int kind, double sr, double si, double* out)
{
  RedWork* w;
  std::lock_guard<std::mutex> red_lk(g_red_mu);  // held until red_finish has synchronized
  if (int rc = red_ws(&w)) return rc;
  using V = typename Vec<T>::t;
  const unsigned g = red_grid(h->size());
  if (kind == 0) {
    k_red_overlap<T><<<g, kBlock, 0, g_stream>>>((const V*)h->d, (const V*)x->d, h->size(), w->partial);
    CHECK_LAUNCH();
    return red_finish(w, g, 2, 0, out);
  } else if (kind == 1) {
    k_red_l2diff<T><<<g, kBlock, 0, g_stream>>>((const V*)h->d, (const V*)x->d, h->size(), w->partial);
    CHECK_LAUNCH();
    return red_finish(w, g, 1, 0, out);
  }
  k_red_maxabs<T><<<g, kBlock, 0, g_stream>>>((const V*)h->d, (const V*)x->d, h->size(), sr, si, w->partial);
  CHECK_LAUNCH();
  return red_finish(w, g, 1, 1, out);
}
}  // namespace

int iqsh_prob1(iqsh_handle h, unsigned pos, double* out)
{
  if (int rc = check_h(h)) return rc;
  if (int rc = check_pos(h, pos)) return rc;
  if (!out) return fail(IQSH_ERR_ARG, "null output");
  return DISPATCH(h, do_norm_sub)(h, 1, pos, out);
}

namespace {
template <typename T>
int do_maxnorm_half(iqsh_handle h, unsigned pos, int value, double* out)
{
  RedWork* w;
  std::lock_guard<std::mutex> red_lk(g_red_mu);  // held until red_finish has synchronized
  if (int rc = red_ws(&w)) return rc;
  Ins<1> ins;
  ins.p[0] = pos;
  const unsigned g = red_grid(h->size() / 2);
  k_red_maxnorm<T><<<g, kBlock, 0, g_stream>>>((const typename Vec<T>::t*)h->d, h->size() / 2, ins,
                                               value ? (uint64_t(1) << pos) : 0, w->partial);
  CHECK_LAUNCH();
  return red_finish(w, g, 1, 1, out);
}
}  // namespace

int iqsh_maxnorm_half(iqsh_handle h, unsigned pos, int value, double* out)
{
  if (int rc = check_h(h)) return rc;
  if (int rc = check_pos(h, pos)) return rc;
  if (!out) return fail(IQSH_ERR_ARG, "null output");
  return DISPATCH(h, do_maxnorm_half)(h, pos, value, out);
}

int iqsh_norm2(iqsh_handle h, double* out)
{
  if (int rc = check_h(h)) return rc;
  if (!out) return fail(IQSH_ERR_ARG, "null output");
  return DISPATCH(h, do_norm_sub)(h, 0, 0, out);
}

int iqsh_expval_pauli(iqsh_handle h, uint64_t xmask, uint64_t zmask, unsigned ny, double* value, double* nrm2)
{
  if (int rc = check_h(h)) return rc;
  if ((xmask | zmask) >> h->n) return fail(IQSH_ERR_ARG, "Pauli mask exceeds register");
  if ((unsigned)__builtin_popcountll(xmask & zmask) != ny) return fail(IQSH_ERR_ARG, "ny must equal popcount(xmask & zmask)");
  double r[2];
  if (int rc = DISPATCH(h, do_pauli)(h, xmask, zmask, ny, r)) return rc;
  if (value) *value = r[0];
  if (nrm2) *nrm2 = r[1];
  return IQSH_OK;
}

// ---- the expectation value without allocation (iqsh_expval_pauli_noalloc) ----
// Its partial sums and result live in `scratch`, a device array of fixed size (one copy
// per GPU, held until the process ends): kMaxRedVals rows of kMaxRedBlocks partial sums,
// then kMaxRedVals results.  Host code never uses the array's name as an address: on its
// first call on a device, with that device current, the entry point takes the device
// address with hipGetSymbolAddress and keeps it for that device.
namespace {
constexpr size_t kScratchDoubles = size_t(kMaxRedVals) * kMaxRedBlocks + kMaxRedVals;  // 64 KiB + 32 B
__device__ __align__(256) double scratch[kScratchDoubles];
constexpr int kMaxDevices = 64;
double* g_scratch_ptr[kMaxDevices] = {};
std::mutex g_scratch_mu;  // calls of this process take turns on the one array per GPU

int scratch_ptr(double** out)
{
  int dev = 0;
  HIP_TRY(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return fail(IQSH_ERR_NOTSUP, "device %d beyond %d", dev, kMaxDevices);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_scratch_ptr[dev] = (double*)p;
  }
  *out = g_scratch_ptr[dev];
  return IQSH_OK;
}

template <typename T>
int do_pauli_noalloc(iqsh_handle h, uint64_t x, uint64_t z, unsigned ny, double* out2)
{
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  double* scr;
  if (int rc = scratch_ptr(&scr)) return rc;
  const unsigned g = red_grid((x >> 6) ? h->size() / 2 : h->size());
  if (g > kMaxRedBlocks) return fail(IQSH_ERR_ARG, "%u partial sums exceed the scratch array's %d", g, kMaxRedBlocks);
  double* result = scr + size_t(kMaxRedVals) * kMaxRedBlocks;
  launch_pauli<T>(g, (const typename Vec<T>::t*)h->d, h->size(), x, z, ny, scr);
  CHECK_LAUNCH();
  k_red_final<<<1, kBlock, 0, g_stream>>>(scr, g, 2, 0, result);
  CHECK_LAUNCH();
  HIP_TRY(hipMemcpyAsync(out2, result, 2 * sizeof(double), hipMemcpyDeviceToHost, g_stream));
  HIP_TRY(hipStreamSynchronize(g_stream));
  return IQSH_OK;
}
}  // namespace

int iqsh_expval_pauli_noalloc(iqsh_handle h, uint64_t xmask, uint64_t zmask, unsigned ny, double* value, double* nrm2)
{
  if (int rc = check_h(h)) return rc;
  if ((xmask | zmask) >> h->n) return fail(IQSH_ERR_ARG, "Pauli mask exceeds register");
  if ((unsigned)__builtin_popcountll(xmask & zmask) != ny) return fail(IQSH_ERR_ARG, "ny must equal popcount(xmask & zmask)");
  double r[2];
  if (int rc = DISPATCH(h, do_pauli_noalloc)(h, xmask, zmask, ny, r)) return rc;
  if (value) *value = r[0];
  if (nrm2) *nrm2 = r[1];
  return IQSH_OK;
}

static int two_check(iqsh_handle h, iqsh_handle x)
{
  if (int rc = check_h(h)) return rc;
  if (int rc = check_h(x)) return rc;
  if (h->n != x->n || h->dtype != x->dtype) return fail(IQSH_ERR_ARG, "registers differ in size or precision");
  return IQSH_OK;
}

int iqsh_overlap(iqsh_handle h, iqsh_handle x, double* re, double* im)
{
  if (int rc = two_check(h, x)) return rc;
  double r[2];
  if (int rc = DISPATCH(h, do_two)(h, x, 0, 0, 0, r)) return rc;
  if (re) *re = r[0];
  if (im) *im = r[1];
  return IQSH_OK;
}

int iqsh_l2diff2(iqsh_handle h, iqsh_handle x, double* out)
{
  if (int rc = two_check(h, x)) return rc;
  if (!out) return fail(IQSH_ERR_ARG, "null output");
  return DISPATCH(h, do_two)(h, x, 1, 0, 0, out);
}

int iqsh_maxabsdiff(iqsh_handle h, iqsh_handle x, double sr, double si, double* out)
{
  if (int rc = two_check(h, x)) return rc;
  if (!out) return fail(IQSH_ERR_ARG, "null output");
  return DISPATCH(h, do_two)(h, x, 2, sr, si, out);
}

// ---- QAOA ----
namespace {
unsigned stream_grid(uint64_t N) { return (unsigned)std::min<uint64_t>(grid_for(N, kBlock * 4), 8192); }

// Unit weights on registers of >= 1024 amplitudes: the table-free row kernel (28-qubit
// complex128: phase 1.68 -> 1.45 ms, cost 0.98 -> 0.65 ms); weighted cuts and small registers:
// the byte-table kernels.
constexpr uint64_t kDiagRow = (uint64_t)kBlock * 4;

template <typename T>
int do_qaoa_phase(iqsh_handle h, const CutSpec& cs)
{
  using V = typename Vec<T>::t;
  if (!cs.weighted && h->size() >= kDiagRow) {
    const uint64_t rows = h->size() / kDiagRow;
    k_qaoa_diag_rows<T, 0><<<(unsigned)std::min<uint64_t>(rows, uint64_t(1) << 22), kBlock, 0, g_stream>>>(
        (V*)h->d, nullptr, h->size(), cs, 0, nullptr);
  } else {
    k_qaoa_phase<T><<<stream_grid(h->size()), kBlock, 0, g_stream>>>((V*)h->d, h->size(), cs);
  }
  CHECK_LAUNCH();
  return IQSH_OK;
}

template <typename T>
int do_qaoa_cost(iqsh_handle h, iqsh_handle lam, const CutSpec& cs, double* out)
{
  RedWork* w;
  std::lock_guard<std::mutex> red_lk(g_red_mu);  // held until red_finish has synchronized
  if (int rc = red_ws(&w)) return rc;
  const unsigned g = red_grid(h->size());
  if (!cs.weighted && h->size() >= kDiagRow)
    k_qaoa_diag_rows<T, 1><<<g, kBlock, 0, g_stream>>>((typename Vec<T>::t*)h->d,
                                                       lam ? (typename Vec<T>::t*)lam->d : nullptr, h->size(), cs,
                                                       lam != nullptr, w->partial);
  else
    k_qaoa_cost<T><<<g, kBlock, 0, g_stream>>>((const typename Vec<T>::t*)h->d,
                                               lam ? (typename Vec<T>::t*)lam->d : nullptr, h->size(), cs, lam != nullptr,
                                               w->partial);
  CHECK_LAUNCH();
  return red_finish(w, g, 1, 0, out);
}

template <typename T>
int do_qaoa_phase_adj(iqsh_handle h, iqsh_handle lam, const CutSpec& cs, double* out)
{
  RedWork* w;
  std::lock_guard<std::mutex> red_lk(g_red_mu);  // held until red_finish has synchronized
  if (int rc = red_ws(&w)) return rc;
  const unsigned g = red_grid(h->size());
  k_qaoa_phase_adj<T><<<g, kBlock, 0, g_stream>>>((typename Vec<T>::t*)h->d, (typename Vec<T>::t*)lam->d, h->size(), cs,
                                                  w->partial);
  CHECK_LAUNCH();
  double s;
  if (int rc = red_finish(w, g, 1, 0, &s)) return rc;
  *out = 2.0 * s;
  return IQSH_OK;
}

template <typename T>
int do_maxcut_fill(iqsh_handle h, const CutSpec& cs, int accumulate, double* out)
{
  RedWork* w;
  std::lock_guard<std::mutex> red_lk(g_red_mu);  // held until red_finish has synchronized
  if (int rc = red_ws(&w)) return rc;
  const unsigned g = red_grid(h->size());
  k_maxcut_fill<T><<<g, kBlock, 0, g_stream>>>((typename Vec<T>::t*)h->d, h->size(), cs, accumulate, w->partial);
  CHECK_LAUNCH();
  return red_finish(w, g, 1, 1, out);
}

template <typename T>
int do_qaoa_diag(iqsh_handle psi, iqsh_handle diag, double gamma)
{
  k_qaoa_diag<T><<<stream_grid(psi->size()), kBlock, 0, g_stream>>>((typename Vec<T>::t*)psi->d,
                                                                    (const typename Vec<T>::t*)diag->d, psi->size(),
                                                                    (T)gamma);
  CHECK_LAUNCH();
  return IQSH_OK;
}

template <typename T>
int do_qaoa_diag_expect(iqsh_handle psi, iqsh_handle diag, int squared, double* out)
{
  RedWork* w;
  std::lock_guard<std::mutex> red_lk(g_red_mu);  // held until red_finish has synchronized
  if (int rc = red_ws(&w)) return rc;
  const unsigned g = red_grid(psi->size());
  k_qaoa_diag_expect<T><<<g, kBlock, 0, g_stream>>>((const typename Vec<T>::t*)psi->d,
                                                    (const typename Vec<T>::t*)diag->d, psi->size(), squared, w->partial);
  CHECK_LAUNCH();
  return red_finish(w, g, 1, 0, out);
}

template <typename T>
int do_adj_1q(iqsh_handle h, iqsh_handle lam, unsigned pos, const double* uinv, const double* G, double* out)
{
  RedWork* w;
  std::lock_guard<std::mutex> red_lk(g_red_mu);  // held until red_finish has synchronized
  if (int rc = red_ws(&w)) return rc;
  const uint64_t items = h->size() / 2;
  const unsigned g = red_grid(items);
  const double zero[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  k_adj_1q<T><<<g, kBlock, 0, g_stream>>>((typename Vec<T>::t*)h->d, (typename Vec<T>::t*)lam->d, items, pos,
                                          mat2<T>(uinv), mat2<double>(G ? G : zero), G != nullptr, w->partial);
  CHECK_LAUNCH();
  if (!G) return IQSH_OK;
  double s;
  if (int rc = red_finish(w, g, 1, 0, &s)) return rc;
  *out = 2.0 * s;
  return IQSH_OK;
}
}  // namespace

int iqsh_qaoa_phase(iqsh_handle h, unsigned ne, const unsigned* u, const unsigned* v, const double* w, double gamma)
{
  if (int rc = check_h(h)) return rc;
  if (ne && (!u || !v)) return fail(IQSH_ERR_ARG, "null edge list");
  CutSpec cs;
  if (int rc = build_cut(h, ne, u, v, w, gamma, &cs)) return rc;
  return DISPATCH(h, do_qaoa_phase)(h, cs);
}

int iqsh_qaoa_expect(iqsh_handle h, unsigned ne, const unsigned* u, const unsigned* v, const double* w, double* out)
{
  if (int rc = check_h(h)) return rc;
  if (!out || (ne && (!u || !v))) return fail(IQSH_ERR_ARG, "null argument");
  CutSpec cs;
  if (int rc = build_cut(h, ne, u, v, w, 0.0, &cs)) return rc;
  return DISPATCH(h, do_qaoa_cost)(h, nullptr, cs, out);
}

int iqsh_qaoa_cost_apply(iqsh_handle psi, iqsh_handle lam, unsigned ne, const unsigned* u, const unsigned* v,
                         const double* w, double* cost)
{
  if (int rc = two_check(psi, lam)) return rc;
  if (!cost || (ne && (!u || !v))) return fail(IQSH_ERR_ARG, "null argument");
  CutSpec cs;
  if (int rc = build_cut(psi, ne, u, v, w, 0.0, &cs)) return rc;
  return DISPATCH(psi, do_qaoa_cost)(psi, lam, cs, cost);
}

int iqsh_qaoa_phase_adjoint(iqsh_handle psi, iqsh_handle lam, unsigned ne, const unsigned* u, const unsigned* v,
                            const double* w, double gamma, double* grad)
{
  if (int rc = two_check(psi, lam)) return rc;
  if (!grad || (ne && (!u || !v))) return fail(IQSH_ERR_ARG, "null argument");
  CutSpec cs;
  // Un-applying exp(-i gamma C) is applying exp(-i (-gamma) C).
  if (int rc = build_cut(psi, ne, u, v, w, -gamma, &cs)) return rc;
  return DISPATCH(psi, do_qaoa_phase_adj)(psi, lam, cs, grad);
}

int iqsh_maxcut_diag(iqsh_handle diag, unsigned ne, const unsigned* u, const unsigned* v, const double* w,
                     double* max_cut)
{
  if (int rc = check_h(diag)) return rc;
  if (ne && (!u || !v)) return fail(IQSH_ERR_ARG, "null edge list");
  // Any number of edges: one pass per 64 edges, the later passes adding to the register
  // (the maximum is that of the last pass, over the complete cuts).
  double m = 0;
  unsigned e0 = 0;
  do {
    const unsigned k = std::min(ne - e0, 64u);
    CutSpec cs;
    if (int rc = build_cut(diag, k, u + e0, v + e0, w ? w + e0 : nullptr, 0.0, &cs)) return rc;
    if (int rc = DISPATCH(diag, do_maxcut_fill)(diag, cs, e0 > 0, &m)) return rc;
    e0 += k;
  } while (e0 < ne);
  if (max_cut) *max_cut = m;
  return IQSH_OK;
}

int iqsh_qaoa_phase_diag(iqsh_handle psi, iqsh_handle diag, double gamma)
{
  if (int rc = two_check(psi, diag)) return rc;
  return DISPATCH(psi, do_qaoa_diag)(psi, diag, gamma);
}

int iqsh_qaoa_expect_diag(iqsh_handle psi, iqsh_handle diag, int squared, double* out)
{
  if (int rc = two_check(psi, diag)) return rc;
  if (!out) return fail(IQSH_ERR_ARG, "null output");
  return DISPATCH(psi, do_qaoa_diag_expect)(psi, diag, squared, out);
}

int iqsh_adjoint_1q(iqsh_handle psi, iqsh_handle lam, unsigned pos, const double uinv[8], const double* G, double* grad)
{
  if (int rc = two_check(psi, lam)) return rc;
  if (int rc = check_pos(psi, pos)) return rc;
  if (!uinv || (G && !grad)) return fail(IQSH_ERR_ARG, "null argument");
  return DISPATCH(psi, do_adj_1q)(psi, lam, pos, uinv, G, grad);
}
