/* Typed body of the CPU oracle; included by iqs_oracle.c once per precision with
 *   CT  complex amplitude type (double _Complex | float _Complex)
 *   BT  base type (double | float)
 *   SFX name suffix (dp | sp)
 *   MK  complex constructor (CMPLX | CMPLXF)
 *   COSB/SINB/SQRTB/CABS/CONJ  the libm call that std::cos/sin/sqrt/abs/conj resolve to
 *   for a BT (or complex<BT>) argument
 * TEST INFRASTRUCTURE ONLY (see iqs_oracle.c).  Every loop restates the reference's loop
 * and operation order so that, on x86-64 without FMA contraction, elementwise results are
 * bit-identical to the compiled reference; reductions follow its single-thread order. */

#define CAT_(a, b) a##_##b
#define CAT(a, b) CAT_(a, b)
#define F(name) CAT(name, SFX)
#define RE(z) (__real__(z))
#define IM(z) (__imag__(z))

/* ---- 2x2 update on pairs (i, i+2^pos), bit pos of i clear.
 * Loop_DN general body, highperfkernels.cpp:326-361 (s0 = m00*in0 + m01*in1, ...). */
void F(orc_1q)(CT *st, int n, unsigned pos, const CT *m)
{
  const u64 N = (u64)1 << n, h = (u64)1 << pos;
  const CT m00 = m[0], m01 = m[1], m10 = m[2], m11 = m[3];
  for (u64 group = 0; group < N; group += 2 * h)
    for (u64 i0 = group; i0 < group + h; ++i0) {
      const u64 i1 = i0 + h;
      const CT in0 = st[i0], in1 = st[i1];
      st[i0] = m00 * in0 + m01 * in1;
      st[i1] = m10 * in0 + m11 * in1;
    }
}

/* ---- controlled 2x2: pairs with bit c = 1, bit t = 0 / 1.
 * Loop_TN general body, highperfkernels.cpp:397-486, index nests of
 * qureg_applyctrl1qubitgate.cpp:312-346 (both nests enumerate the same pair set). */
void F(orc_c1q)(CT *st, int n, unsigned c, unsigned t, const CT *m)
{
  const u64 N = (u64)1 << n, dc = (u64)1 << c, dt = (u64)1 << t;
  const CT m00 = m[0], m01 = m[1], m10 = m[2], m11 = m[3];
  for (u64 i0 = 0; i0 < N; ++i0) {
    if (!(i0 & dc) || (i0 & dt)) continue;
    const u64 i1 = i0 + dt;
    const CT in0 = st[i0], in1 = st[i1];
    st[i0] = m00 * in0 + m01 * in1;
    st[i1] = m10 * in0 + m11 * in1;
  }
}

/* ---- swap-like 2x2 on (|p1=1,p2=0>, |p1=0,p2=1>), p1 < p2.
 * ApplySwap_helper's Loop_TN call, qureg_applyswap.cpp:203-206. */
void F(orc_swaplike)(CT *st, int n, unsigned p1, unsigned p2, const CT *m)
{
  const u64 N = (u64)1 << n, d1 = (u64)1 << p1, d2 = (u64)1 << p2;
  const CT m00 = m[0], m01 = m[1], m10 = m[2], m11 = m[3];
  for (u64 i0 = 0; i0 < N; ++i0) {
    if (!(i0 & d1) || (i0 & d2)) continue;
    const u64 i1 = i0 - d1 + d2;
    const CT in0 = st[i0], in1 = st[i1];
    st[i0] = m00 * in0 + m01 * in1;
    st[i1] = m10 * in0 + m11 * in1;
  }
}

/* ---- 2-qubit diagonal: amplitude *= d[2*b(p1) + b(p2)].
 * ApplyDiag local branch, qureg_applydiag.cpp:157-174 ("state *= d"). */
void F(orc_diag)(CT *st, int n, unsigned p1, unsigned p2, const CT *d)
{
  const u64 N = (u64)1 << n;
  for (u64 i = 0; i < N; ++i) {
    const unsigned k = 2u * (unsigned)((i >> p1) & 1) + (unsigned)((i >> p2) & 1);
    st[i] *= d[k];
  }
}

/* ---- dense 4x4 on (ph, pl); row index t = 2*b(ph) + b(pl).
 * Apply2QubitGate, qureg_apply2qubitgate.cpp:53-62: idx = {t, t+dl, t+dh, t+dh+dl},
 * new[t] = ((m[t][0]*uu + m[t][1]*ud) + m[t][2]*du) + m[t][3]*dd.  The pair set is
 * enumerated by position (the reference's loop bounds compare program qubits, which
 * coincide with positions under the identity permutation). */
void F(orc_2q)(CT *st, int n, unsigned ph, unsigned pl, const CT *m)
{
  const u64 N = (u64)1 << n, dh = (u64)1 << ph, dl = (u64)1 << pl;
  for (u64 b = 0; b < N; ++b) {
    if ((b & dh) || (b & dl)) continue;
    const u64 idx[4] = {b, b + dl, b + dh, b + dh + dl};
    const CT uu = st[idx[0]], ud = st[idx[1]], du = st[idx[2]], dd = st[idx[3]];
    for (int t = 0; t < 4; ++t)
      st[idx[t]] = m[4 * t + 0] * uu + m[4 * t + 1] * ud + m[4 * t + 2] * du + m[4 * t + 3] * dd;
  }
}

/* ---- SpecializeV2 bodies for 1q gates, spec_kernels.cpp:36-102 with the constants of
 * Loop_DN spec, spec_kernels.cpp:264-270 (cos/sin in BT precision). */
void F(orc_1q_spec)(CT *st, int n, unsigned pos, int spec, double angle)
{
  const u64 N = (u64)1 << n, h = (u64)1 << pos;
  const BT theta = (BT)angle;
  const BT isqrt2 = (BT)(1 / sqrt(2.0));
  const BT cos_2 = COSB(theta / 2), sin_2 = SINB(theta / 2), msin_2 = -sin_2;
  const CT texp = MK((BT)cos(M_PI / 4), (BT)sin(M_PI / 4));
  for (u64 group = 0; group < N; group += 2 * h)
    for (u64 i0 = group; i0 < group + h; ++i0) {
      const u64 i1 = i0 + h;
      const CT in0 = st[i0], in1 = st[i1];
      switch (spec) {
        case SPEC_H: st[i0] = (in0 + in1) * isqrt2; st[i1] = (in0 - in1) * isqrt2; break;
        case SPEC_RX:
          st[i0] = cos_2 * in0 + MK(sin_2 * IM(in1), msin_2 * RE(in1));
          st[i1] = MK(sin_2 * IM(in0), msin_2 * RE(in0)) + cos_2 * in1; break;
        case SPEC_RY: st[i0] = cos_2 * in0 + msin_2 * in1; st[i1] = sin_2 * in0 + cos_2 * in1; break;
        case SPEC_RZ: st[i0] = MK(cos_2, msin_2) * in0; st[i1] = MK(cos_2, sin_2) * in1; break;
        case SPEC_X: st[i0] = in1; st[i1] = in0; break;
        case SPEC_Y: st[i0] = MK(IM(in1), -RE(in1)); st[i1] = MK(-IM(in0), RE(in0)); break;
        case SPEC_Z: st[i1] = -in1; break;
        case SPEC_T: st[i1] = in1 * texp; break;
        default: abort();
      }
    }
}

/* ---- SpecializeV2 bodies for controlled gates, spec_kernels.cpp:104-153 with the
 * constants of Loop_TN spec, spec_kernels.cpp:356-361 (CPhase: pexp = e^{i theta}). */
void F(orc_c1q_spec)(CT *st, int n, unsigned c, unsigned t, int spec, double angle)
{
  const u64 N = (u64)1 << n, dc = (u64)1 << c, dt = (u64)1 << t;
  const BT theta = (BT)angle;
  const BT isqrt2 = (BT)(1 / sqrt(2.0));
  const BT cos_2 = COSB(theta / 2), sin_2 = SINB(theta / 2), msin_2 = -sin_2;
  const CT pexp = MK(COSB(theta), SINB(theta));
  for (u64 i0 = 0; i0 < N; ++i0) {
    if (!(i0 & dc) || (i0 & dt)) continue;
    const u64 i1 = i0 + dt;
    const CT in0 = st[i0], in1 = st[i1];
    switch (spec) {
      case SPEC_H: st[i0] = (in0 + in1) * isqrt2; st[i1] = (in0 - in1) * isqrt2; break;
      case SPEC_RX:
        st[i0] = cos_2 * in0 + MK(sin_2 * IM(in1), msin_2 * RE(in1));
        st[i1] = MK(sin_2 * IM(in0), msin_2 * RE(in0)) + cos_2 * in1; break;
      case SPEC_RY: st[i0] = cos_2 * in0 + msin_2 * in1; st[i1] = sin_2 * in0 + cos_2 * in1; break;
      case SPEC_RZ: st[i0] = MK(cos_2, msin_2) * in0; st[i1] = MK(cos_2, sin_2) * in1; break;
      case SPEC_X: st[i0] = in1; st[i1] = in0; break;
      case SPEC_Y: st[i0] = MK(IM(in1), -RE(in1)); st[i1] = MK(-IM(in0), RE(in0)); break;
      case SPEC_Z: st[i1] = -in1; break;
      case SPEC_CPHASE: st[i1] = in1 * pexp; break;
      default: abort();
    }
  }
}

/* ---- physical reordering: state[P_new.program2data_(P_old.data2program_(i))] = old[i].
 * PermuteLocalQubits, qureg_permute.cpp:89-100; index maps permutation.hpp:249-264. */
void F(orc_permute)(CT *st, int n, const u64 *old_map, const u64 *new_imap)
{
  const u64 N = (u64)1 << n;
  CT *old = (CT *)malloc(N * sizeof(CT));
  memcpy(old, st, N * sizeof(CT));
  for (u64 i = 0; i < N; ++i) {
    u64 prog = 0, to = 0;
    for (int k = 0; k < n; ++k) prog |= ((i >> old_map[k]) & 1) << k;   /* data2program_ */
    for (int k = 0; k < n; ++k) to |= ((prog >> new_imap[k]) & 1) << k; /* program2data_ */
    st[to] = old[i];
  }
  free(old);
}

/* ---- P(bit pos = 1), serial BT accumulation: GetProbability, qureg_measure.cpp:149-157. */
double F(orc_prob1)(const CT *st, int n, unsigned pos)
{
  const u64 N = (u64)1 << n, d = (u64)1 << pos;
  BT p = 0;
  for (u64 i = d; i < N; i += 2 * d)
    for (u64 j = 0; j < d; ++j) {
      const CT z = st[i + j];
      p += RE(z) * RE(z) + IM(z) * IM(z);
    }
  return (double)p;
}

/* ---- sqrt(sum |psi|^2) in BT: ComputeNorm, qureg_utils.cpp:236-256. */
double F(orc_norm)(const CT *st, int n)
{
  const u64 N = (u64)1 << n;
  BT s = 0;
  for (u64 i = 0; i < N; ++i) s += RE(st[i]) * RE(st[i]) + IM(st[i]) * IM(st[i]);
  return (double)SQRTB(s);
}

/* ---- sum (-1)^{popcount(i & y)} |psi_i|^2 in BT: ExpectationValue reduction,
 * qureg_expectval.cpp:173-185 (after its basis change). */
double F(orc_parity_norm)(const CT *st, int n, u64 y)
{
  const u64 N = (u64)1 << n;
  BT v = 0;
  for (u64 i = 0; i < N; ++i) {
    const BT z2 = RE(st[i]) * RE(st[i]) + IM(st[i]) * IM(st[i]);
    if (__builtin_popcountll(i & y) & 1) v -= z2; else v += z2;
  }
  return (double)v;
}

/* ---- <b|a> = sum conj(b_i) * a_i, separate BT sums: ComputeOverlap, qureg_utils.cpp:267-300. */
void F(orc_overlap)(const CT *a, const CT *b, int n, double *re, double *im)
{
  const u64 N = (u64)1 << n;
  BT sr = 0, si = 0;
  for (u64 i = 0; i < N; ++i) {
    const CT o = CONJ(b[i]) * a[i];
    sr += RE(o);
    si += IM(o);
  }
  *re = (double)sr;
  *im = (double)si;
}

/* ---- max |a_i - s*b_i|: MaxAbsDiff, qureg_utils.cpp:40-68 (starts at -1). */
double F(orc_maxabsdiff)(const CT *a, const CT *b, int n, CT s)
{
  const u64 N = (u64)1 << n;
  BT mx = -1.0;
  for (u64 i = 0; i < N; ++i) {
    const BT d = CABS(a[i] - s * b[i]);
    if (d > mx) mx = d;
  }
  return (double)mx;
}

/* ---- state *= f: AmplitudeWiseScalarMultiplication, qureg_utils.cpp:191-201. */
void F(orc_scale)(CT *st, int n, CT f)
{
  const u64 N = (u64)1 << n;
  for (u64 i = 0; i < N; ++i) st[i] = st[i] * f;
}

/* ---- CollapseQubit, qureg_measure.cpp:93-126: zero the half with bit != value. */
void F(orc_collapse)(CT *st, int n, unsigned pos, int value)
{
  const u64 N = (u64)1 << n, d = (u64)1 << pos;
  for (u64 i = value ? 0 : d; i < N; i += 2 * d)
    for (u64 j = 0; j < d; ++j) st[i + j] = 0;
}

/* ---- psi_i *= (cos(g*c_i), -sin(g*c_i)), c_i = diag_i.real():
 * ImplementQaoaLayerBasedOnCostFunction, qaoa_features.cpp:248-264. */
void F(orc_qaoa_phase)(CT *psi, const CT *diag, int n, BT gamma)
{
  const u64 N = (u64)1 << n;
  for (u64 i = 0; i < N; ++i) {
    const BT x = gamma * RE(diag[i]);
    psi[i] *= MK(COSB(x), -SINB(x));
  }
}

/* ---- sum c_i |psi_i|^2 in BT: GetExpectationValueFromCostFunction, qaoa_features.cpp:273-295. */
double F(orc_qaoa_expect)(const CT *psi, const CT *diag, int n)
{
  const u64 N = (u64)1 << n;
  BT e = 0;
  for (u64 i = 0; i < N; ++i)
    e += RE(diag[i]) * (RE(psi[i]) * RE(psi[i]) + IM(psi[i]) * IM(psi[i]));
  return (double)e;
}

/* ---- MaxCut cost diagonal in DATA order: InitializeVectorAsMaxCutCostFunction and its
 * weighted variant, qaoa_features.cpp:73-239 (x^T.ADJ.x over +-1 colourings). */
void F(orc_maxcut_diag)(CT *diag, int n, const u64 *map, const BT *adj, int weighted)
{
  const u64 N = (u64)1 << n;
  BT total = 0;
  int num_edges = 0;
  for (int v1 = 0; v1 < n; ++v1)
    for (int v2 = v1 + 1; v2 < n; ++v2)
      if (adj[v1 * n + v2] != 0) { num_edges += 1; total += adj[v1 * n + v2]; }
  for (u64 i = 0; i < N; ++i) {
    u64 x = 0;
    for (int k = 0; k < n; ++k) x |= ((i >> map[k]) & 1) << k; /* data2program_ */
    if (!weighted) {
      int cut = 0;
      for (int v = 0; v < n; ++v)
        for (int u = 0; u < n; ++u) {
          const int xv = ((x >> v) & 1) ? 1 : -1, xu = ((x >> u) & 1) ? 1 : -1;
          cut += (int)adj[v * n + u] * xv * xu;
        }
      cut = (num_edges - cut / 2) / 2;
      diag[i] = MK((BT)cut, 0);
    } else {
      BT cut = 0;
      for (int v = 0; v < n; ++v)
        for (int u = 0; u < n; ++u) {
          const BT xv = ((x >> v) & 1) ? 1 : -1, xu = ((x >> u) & 1) ? 1 : -1;
          cut += adj[v * n + u] * xv * xu;
        }
      cut = (total - cut / 2);
      cut /= 2;
      diag[i] = MK(cut, 0);
    }
  }
}

#undef F
#undef RE
#undef IM
#undef CAT
#undef CAT_
