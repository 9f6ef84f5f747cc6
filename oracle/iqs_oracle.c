/* CPU ORACLE for the gate-application path of intel-qs -- TEST INFRASTRUCTURE ONLY.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load this
 * library, and only as the checker.  The product (intel-qs_amd/) never links it.
 *
 * It restates, in plain C, the reference's loops (iqs_oracle_impl.h, once per precision)
 * and its QubitRegister front end (this file): named-gate matrices, the program-qubit ->
 * data-position map, SpecializeV2 dispatch, the ExpectationValue basis-change sequence,
 * the QAOA cost diagonal.  Each function cites the reference file:line it follows.
 *
 * Pinning (how this restatement is checked, see tests/test_oracle_*.py):
 *   - bit-identical to the compiled reference (oracle/_ref, built by oracle/Makefile) on
 *     the committed fixtures in tests/golden/ (elementwise ops, DP and SP);
 *   - reductions equal the reference's single-thread results (fixtures are made with
 *     OMP_NUM_THREADS=1);
 *   - brute-force Kronecker-product matrices and closed forms on small registers.
 * Deviations kept on purpose: TurnOnFusion/TurnOffFusion are accepted and ignored (the
 * reference's fusion is cache blocking; its unfused result is the same arithmetic).
 */
#include <complex.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

typedef uint64_t u64;
/* complex<double>(re, im) from any arithmetic arguments */
#define CX(a, b) CMPLX((double)(a), (double)(b))

enum { SPEC_H = 0, SPEC_RX, SPEC_RY, SPEC_RZ, SPEC_X, SPEC_Y, SPEC_Z, SPEC_T, SPEC_CPHASE, SPEC_NONE };

#define CT double _Complex
#define BT double
#define SFX dp
#define MK(a, b) CMPLX((double)(a), (double)(b))
#define COSB cos
#define SINB sin
#define SQRTB sqrt
#define CABS cabs
#define CONJ conj
#include "iqs_oracle_impl.h"
#undef CT
#undef BT
#undef SFX
#undef MK
#undef COSB
#undef SINB
#undef SQRTB
#undef CABS
#undef CONJ

#define CT float _Complex
#define BT float
#define SFX sp
#define MK(a, b) CMPLXF((float)(a), (float)(b))
#define COSB cosf
#define SINB sinf
#define SQRTB sqrtf
#define CABS cabsf
#define CONJ conjf
#include "iqs_oracle_impl.h"
#undef CT
#undef BT
#undef SFX
#undef MK
#undef COSB
#undef SINB
#undef SQRTB
#undef CABS
#undef CONJ

/* ------------------------------------------------------------------------------------ */
/* Register front end.  Matrices are kept as double _Complex and narrowed to the register's
 * precision exactly where the reference narrows (TinyMatrix<Type> element assignment). */

typedef struct {
  int n, sp;
  void *st;
  u64 map[64], imap[64];
  int spec2;
  void *diag; /* QAOA cost diagonal, same layout as st */
  int ne;
  int eu[256], ev[256];
  double ew[256];
  char *out;
  size_t out_len, out_cap;
} Reg;

static void emit(Reg *r, const char *fmt, ...) __attribute__((format(printf, 2, 3)));
#include <stdarg.h>
static void emit(Reg *r, const char *fmt, ...)
{
  va_list ap;
  va_start(ap, fmt);
  char buf[8192];
  int k = vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  if (k < 0) return;
  if (r->out_len + (size_t)k + 1 > r->out_cap) {
    r->out_cap = 2 * (r->out_len + (size_t)k + 1);
    r->out = (char *)realloc(r->out, r->out_cap);
  }
  memcpy(r->out + r->out_len, buf, (size_t)k + 1);
  r->out_len += (size_t)k;
}

/* Narrow a double-precision value as the reference's Type(...) constructor does. */
static double _Complex nar(const Reg *r, double re, double im)
{
  if (r->sp) return CX((double)(float)re, (double)(float)im);
  return CX(re, im);
}

static void m2_to(const Reg *r, const double _Complex m[4], void *dst)
{
  if (r->sp) { float _Complex *d = (float _Complex *)dst; for (int i = 0; i < 4; ++i) d[i] = CMPLXF((float)creal(m[i]), (float)cimag(m[i])); }
  else { double _Complex *d = (double _Complex *)dst; for (int i = 0; i < 4; ++i) d[i] = m[i]; }
}

static void gate1(Reg *r, unsigned q, const double _Complex m[4], int spec, double angle)
{
  /* Apply1QubitGate_helper, qureg_apply1qubitgate.cpp:173-230 (single rank: Loop_DN). */
  const unsigned pos = (unsigned)r->map[q];
  if (r->spec2 && spec != SPEC_NONE) {
    if (r->sp) orc_1q_spec_sp((float _Complex *)r->st, r->n, pos, spec, angle);
    else orc_1q_spec_dp((double _Complex *)r->st, r->n, pos, spec, angle);
    return;
  }
  double _Complex mm[4];
  for (int i = 0; i < 4; ++i) mm[i] = nar(r, creal(m[i]), cimag(m[i]));
  if (r->sp) { float _Complex t[4]; m2_to(r, mm, t); orc_1q_sp((float _Complex *)r->st, r->n, pos, t); }
  else orc_1q_dp((double _Complex *)r->st, r->n, pos, mm);
}

static void cgate1(Reg *r, unsigned c, unsigned t, const double _Complex m[4], int spec, double angle)
{
  /* ApplyControlled1QubitGate_helper, qureg_applyctrl1qubitgate.cpp:228-406 (C,T < M). */
  const unsigned cp = (unsigned)r->map[c], tp = (unsigned)r->map[t];
  if (r->spec2 && spec != SPEC_NONE) {
    if (r->sp) orc_c1q_spec_sp((float _Complex *)r->st, r->n, cp, tp, spec, angle);
    else orc_c1q_spec_dp((double _Complex *)r->st, r->n, cp, tp, spec, angle);
    return;
  }
  double _Complex mm[4];
  for (int i = 0; i < 4; ++i) mm[i] = nar(r, creal(m[i]), cimag(m[i]));
  if (r->sp) { float _Complex tt[4]; m2_to(r, mm, tt); orc_c1q_sp((float _Complex *)r->st, r->n, cp, tp, tt); }
  else orc_c1q_dp((double _Complex *)r->st, r->n, cp, tp, mm);
}

static void swaplike(Reg *r, unsigned a, unsigned b, const double _Complex m[4])
{
  /* ApplySwap_helper, qureg_applyswap.cpp:114-234: order positions so p1 < p2. */
  unsigned p1 = (unsigned)r->map[a], p2 = (unsigned)r->map[b];
  if (p1 > p2) { unsigned t = p1; p1 = p2; p2 = t; }
  double _Complex mm[4];
  for (int i = 0; i < 4; ++i) mm[i] = nar(r, creal(m[i]), cimag(m[i]));
  if (r->sp) { float _Complex t[4]; m2_to(r, mm, t); orc_swaplike_sp((float _Complex *)r->st, r->n, p1, p2, t); }
  else orc_swaplike_dp((double _Complex *)r->st, r->n, p1, p2, mm);
}

static double rbt(const Reg *r, double x) { return r->sp ? (double)(float)x : x; }
/* std::sin / std::cos of a BaseType argument: sinf/cosf for ComplexSP. */
static double sinb(const Reg *r, double x) { return r->sp ? (double)sinf((float)x) : sin(x); }
static double cosb(const Reg *r, double x) { return r->sp ? (double)cosf((float)x) : cos(x); }

/* Named 1q gates: qureg_apply1qubitgate.cpp:277-501. theta is first narrowed to BaseType
 * (the API takes BaseType), then `theta / 2.` is evaluated in double. */
static void named1(Reg *r, const char *name, unsigned q, double th, double ph)
{
  double _Complex m[4];
  th = rbt(r, th);
  ph = rbt(r, ph);
  const double s = sin(th / 2.), c = cos(th / 2.);
  if (!strcmp(name, "ApplyRotationX")) {
    m[0] = CX(c, 0); m[1] = CX(0, -s); m[2] = CX(0, -s); m[3] = CX(c, 0);
    gate1(r, q, m, SPEC_RX, th);
  } else if (!strcmp(name, "ApplyRotationY")) {
    m[0] = CX(c, 0); m[1] = CX(-s, 0); m[2] = CX(s, 0); m[3] = CX(c, 0);
    gate1(r, q, m, SPEC_RY, th);
  } else if (!strcmp(name, "ApplyRotationZ")) {
    m[0] = CX(c, -s); m[1] = 0; m[2] = 0; m[3] = CX(c, s);
    gate1(r, q, m, SPEC_RZ, th);
  } else if (!strcmp(name, "ApplyRotationXY")) {
    /* :472-481; std::sin(phi)/std::cos(phi) take a BaseType argument. */
    m[0] = CX(c, 0);
    m[1] = CX(-s * sinb(r, ph), -s * cosb(r, ph));
    m[2] = CX(s * sinb(r, ph), -s * cosb(r, ph));
    m[3] = CX(c, 0);
    gate1(r, q, m, SPEC_NONE, 0);
  } else if (!strcmp(name, "ApplyPauliX")) {
    m[0] = 0; m[1] = 1; m[2] = 1; m[3] = 0; gate1(r, q, m, SPEC_X, 0);
  } else if (!strcmp(name, "ApplyPauliY")) {
    m[0] = 0; m[1] = CX(0, -1); m[2] = CX(0, 1); m[3] = 0; gate1(r, q, m, SPEC_Y, 0);
  } else if (!strcmp(name, "ApplyPauliZ")) {
    m[0] = 1; m[1] = 0; m[2] = 0; m[3] = -1; gate1(r, q, m, SPEC_Z, 0);
  } else if (!strcmp(name, "ApplyPauliSqrtX")) {
    m[0] = CX(.5, .5); m[1] = CX(.5, -.5); m[2] = CX(.5, -.5); m[3] = CX(.5, .5);
    gate1(r, q, m, SPEC_NONE, 0);
  } else if (!strcmp(name, "ApplyPauliSqrtY")) {
    m[0] = CX(.5, .5); m[1] = CX(-.5, -.5); m[2] = CX(.5, .5); m[3] = CX(.5, .5);
    gate1(r, q, m, SPEC_NONE, 0);
  } else if (!strcmp(name, "ApplyPauliSqrtZ")) {
    m[0] = 1; m[1] = 0; m[2] = 0; m[3] = CX(0, 1); gate1(r, q, m, SPEC_NONE, 0);
  } else if (!strcmp(name, "ApplyT")) {
    m[0] = 1; m[1] = 0; m[2] = 0; m[3] = CX(cos(M_PI / 4.0), sin(M_PI / 4.0));
    gate1(r, q, m, SPEC_T, 0);
  } else if (!strcmp(name, "ApplyHadamard")) {
    const double f = rbt(r, 1. / sqrt(2.));
    m[0] = f; m[1] = f; m[2] = f; m[3] = -f; gate1(r, q, m, SPEC_H, 0);
  } else {
    fprintf(stderr, "oracle: unknown 1q op %s\n", name);
    abort();
  }
}

/* Named controlled gates: qureg_applyctrl1qubitgate.cpp:458-618. */
static void named2(Reg *r, const char *name, unsigned c, unsigned t, double th)
{
  double _Complex m[4];
  th = rbt(r, th);
  const double s = sin(th / 2.), co = cos(th / 2.);
  if (!strcmp(name, "ApplyCRotationX")) {
    m[0] = CX(co, 0); m[1] = CX(0, -s); m[2] = CX(0, -s); m[3] = CX(co, 0);
    cgate1(r, c, t, m, SPEC_RX, th);
  } else if (!strcmp(name, "ApplyCRotationY")) {
    m[0] = CX(co, 0); m[1] = CX(-s, 0); m[2] = CX(s, 0); m[3] = CX(co, 0);
    cgate1(r, c, t, m, SPEC_RY, th);
  } else if (!strcmp(name, "ApplyCRotationZ")) {
    m[0] = CX(co, -s); m[1] = 0; m[2] = 0; m[3] = CX(co, s);
    cgate1(r, c, t, m, SPEC_RZ, th);
  } else if (!strcmp(name, "ApplyCPauliX")) {
    m[0] = 0; m[1] = 1; m[2] = 1; m[3] = 0; cgate1(r, c, t, m, SPEC_X, 0);
  } else if (!strcmp(name, "ApplyCPauliY")) {
    m[0] = 0; m[1] = CX(0, -1); m[2] = CX(0, 1); m[3] = 0; cgate1(r, c, t, m, SPEC_Y, 0);
  } else if (!strcmp(name, "ApplyCPauliZ")) {
    m[0] = 1; m[1] = 0; m[2] = 0; m[3] = -1; cgate1(r, c, t, m, SPEC_Z, 0);
  } else if (!strcmp(name, "ApplyCPauliSqrtZ")) {
    m[0] = 1; m[1] = 0; m[2] = 0; m[3] = CX(0, 1); cgate1(r, c, t, m, SPEC_NONE, 0);
  } else if (!strcmp(name, "ApplyCHadamard")) {
    const double f = rbt(r, 1. / sqrt(2.));
    m[0] = f; m[1] = f; m[2] = f; m[3] = -f; cgate1(r, c, t, m, SPEC_H, 0);
  } else if (!strcmp(name, "ApplyCPhaseRotation")) {
    /* :610-618: Type(std::cos(theta), std::sin(theta)) with a BaseType theta. */
    m[0] = 1; m[1] = 0; m[2] = 0; m[3] = CX(cosb(r, th), sinb(r, th));
    cgate1(r, c, t, m, SPEC_CPHASE, th);
  } else {
    fprintf(stderr, "oracle: unknown 2q op %s\n", name);
    abort();
  }
}

static double prob(Reg *r, unsigned q)
{
  const unsigned pos = (unsigned)r->map[q];
  return r->sp ? orc_prob1_sp((float _Complex *)r->st, r->n, pos) : orc_prob1_dp((double _Complex *)r->st, r->n, pos);
}

/* ExpectationValue, qureg_expectval.cpp:18-74 and :102-213. */
static double expect(Reg *r, double coeff, int k, const unsigned *qs, const unsigned *os)
{
  const double f = rbt(r, 1. / sqrt(2.));
  const double _Complex G[4] = {CX(f, 0), CX(0, -f), CX(f, 0), CX(0, f)};
  const double _Complex Gi[4] = {CX(f, 0), CX(f, 0), CX(0, f), CX(0, -f)};
  coeff = rbt(r, coeff);
  if (k == 0) return coeff;
  if (k == 1) {
    double e;
    if (os[0] == 1) { named1(r, "ApplyHadamard", qs[0], 0, 0); e = rbt(r, 1. - 2. * prob(r, qs[0])); named1(r, "ApplyHadamard", qs[0], 0, 0); }
    else if (os[0] == 2) { gate1(r, qs[0], G, SPEC_NONE, 0); e = rbt(r, 1. - 2. * prob(r, qs[0])); gate1(r, qs[0], Gi, SPEC_NONE, 0); }
    else e = rbt(r, 1. - 2. * prob(r, qs[0]));
    return rbt(r, coeff * e);
  }
  for (int i = 0; i < k; ++i) {
    if (os[i] == 1) named1(r, "ApplyHadamard", qs[i], 0, 0);
    else if (os[i] == 2) gate1(r, qs[i], G, SPEC_NONE, 0);
  }
  u64 y = 0;
  for (int i = 0; i < k; ++i) y += (u64)(long)(1 << r->map[qs[i]]); /* int shift, as :170 */
  const double v = r->sp ? orc_parity_norm_sp((float _Complex *)r->st, r->n, y) : orc_parity_norm_dp((double _Complex *)r->st, r->n, y);
  for (int i = 0; i < k; ++i) {
    if (os[i] == 1) named1(r, "ApplyHadamard", qs[i], 0, 0);
    else if (os[i] == 2) gate1(r, qs[i], Gi, SPEC_NONE, 0);
  }
  return rbt(r, coeff * v);
}

static void ensure_diag(Reg *r)
{
  if (r->diag) return;
  const u64 N = (u64)1 << r->n;
  const int n = r->n;
  int weighted = 0;
  for (int e = 0; e < r->ne; ++e) weighted |= (r->ew[e] != 1.0);
  r->diag = calloc(N, r->sp ? sizeof(float _Complex) : sizeof(double _Complex));
  if (r->sp) {
    float *adj = (float *)calloc((size_t)n * n, sizeof(float));
    for (int e = 0; e < r->ne; ++e) adj[r->eu[e] * n + r->ev[e]] = adj[r->ev[e] * n + r->eu[e]] = (float)r->ew[e];
    orc_maxcut_diag_sp((float _Complex *)r->diag, n, r->map, adj, weighted);
    free(adj);
  } else {
    double *adj = (double *)calloc((size_t)n * n, sizeof(double));
    for (int e = 0; e < r->ne; ++e) adj[r->eu[e] * n + r->ev[e]] = adj[r->ev[e] * n + r->eu[e]] = r->ew[e];
    orc_maxcut_diag_dp((double _Complex *)r->diag, n, r->map, adj, weighted);
    free(adj);
  }
}

static int parse_m(char **sp, int cnt, double _Complex *m)
{
  for (int i = 0; i < cnt; ++i) {
    char *e;
    double re = strtod(*sp, &e); *sp = e;
    double im = strtod(*sp, &e); *sp = e;
    m[i] = CX(re, im);
  }
  return 0;
}

static unsigned pu(char **s) { char *e; unsigned long v = strtoul(*s, &e, 10); *s = e; return (unsigned)v; }
static double pd(char **s) { char *e; double v = strtod(*s, &e); *s = e; return v; }

static void run_op(Reg *r, const char *name, char *s)
{
  const int n = r->n;
  if (!strncmp(name, "ApplyRotationXY", 32)) { unsigned q = pu(&s); double ph = pd(&s), th = pd(&s); named1(r, name, q, th, ph); }
  else if (!strcmp(name, "ApplyRotationX") || !strcmp(name, "ApplyRotationY") || !strcmp(name, "ApplyRotationZ")) {
    unsigned q = pu(&s); double th = pd(&s); named1(r, name, q, th, 0);
  }
  else if (!strcmp(name, "ApplyPauliX") || !strcmp(name, "ApplyPauliY") || !strcmp(name, "ApplyPauliZ") ||
           !strcmp(name, "ApplyPauliSqrtX") || !strcmp(name, "ApplyPauliSqrtY") || !strcmp(name, "ApplyPauliSqrtZ") ||
           !strcmp(name, "ApplyT") || !strcmp(name, "ApplyHadamard")) { unsigned q = pu(&s); named1(r, name, q, 0, 0); }
  else if (!strcmp(name, "Apply1QubitGate")) { unsigned q = pu(&s); double _Complex m[4]; parse_m(&s, 4, m); gate1(r, q, m, SPEC_NONE, 0); }
  else if (!strcmp(name, "ApplyControlled1QubitGate")) { unsigned c = pu(&s), t = pu(&s); double _Complex m[4]; parse_m(&s, 4, m); cgate1(r, c, t, m, SPEC_NONE, 0); }
  else if (!strcmp(name, "ApplyCRotationX") || !strcmp(name, "ApplyCRotationY") || !strcmp(name, "ApplyCRotationZ") ||
           !strcmp(name, "ApplyCPhaseRotation")) { unsigned c = pu(&s), t = pu(&s); double th = pd(&s); named2(r, name, c, t, th); }
  else if (!strcmp(name, "ApplyCPauliX") || !strcmp(name, "ApplyCPauliY") || !strcmp(name, "ApplyCPauliZ") ||
           !strcmp(name, "ApplyCPauliSqrtZ") || !strcmp(name, "ApplyCHadamard")) { unsigned c = pu(&s), t = pu(&s); named2(r, name, c, t, 0); }
  else if (!strcmp(name, "ApplySwap")) { unsigned a = pu(&s), b = pu(&s); double _Complex m[4] = {0, 1, 1, 0}; swaplike(r, a, b, m); }
  else if (!strcmp(name, "ApplyISwap")) { unsigned a = pu(&s), b = pu(&s); double _Complex m[4] = {0, CX(0, 1), CX(0, 1), 0}; swaplike(r, a, b, m); }
  else if (!strcmp(name, "ApplySqrtISwap")) {
    unsigned a = pu(&s), b = pu(&s); const double f = rbt(r, 1. / sqrt(2.));
    double _Complex m[4] = {CX(f, 0), CX(0, f), CX(0, f), CX(f, 0)}; swaplike(r, a, b, m);
  }
  else if (!strcmp(name, "Apply4thRootISwap")) {
    /* qureg_applyswap.cpp:93-108 with std::polar(rho, th) = (rho cos th, rho sin th). */
    unsigned a = pu(&s), b = pu(&s);
    const double _Complex A = CX(.5 * cos(M_PI / 8.), .5 * sin(M_PI / 8.));
    const double _Complex B = CX(.5 * cos(7. * M_PI / 8.), .5 * sin(7. * M_PI / 8.));
    const double _Complex f0 = A - B, f1 = A + B;
    double _Complex m[4] = {f0, f1, f1, f0}; swaplike(r, a, b, m);
  }
  else if (!strcmp(name, "ApplyISwapRotation")) { unsigned a = pu(&s), b = pu(&s); double _Complex m[4]; parse_m(&s, 4, m); swaplike(r, a, b, m); }
  else if (!strcmp(name, "ApplyToffoli")) {
    /* qureg_applytoffoli.cpp:24-47. */
    unsigned c1 = pu(&s), c2 = pu(&s), t = pu(&s);
    const double _Complex V[4] = {CX(.5, -.5), CX(.5, .5), CX(.5, .5), CX(.5, -.5)};
    const double _Complex Vd[4] = {CX(.5, .5), CX(.5, -.5), CX(.5, -.5), CX(.5, .5)};
    cgate1(r, c1, t, V, SPEC_NONE, 0);
    named2(r, "ApplyCPauliX", c2, c1, 0);
    cgate1(r, c1, t, Vd, SPEC_NONE, 0);
    named2(r, "ApplyCPauliX", c2, c1, 0);
    cgate1(r, c2, t, V, SPEC_NONE, 0);
  }
  else if (!strcmp(name, "ApplyDiag")) {
    unsigned a = pu(&s), b = pu(&s); double _Complex d[4]; parse_m(&s, 4, d);
    const unsigned p1 = (unsigned)r->map[a], p2 = (unsigned)r->map[b];
    if (r->sp) { float _Complex t[4]; for (int i = 0; i < 4; ++i) t[i] = CMPLXF((float)creal(d[i]), (float)cimag(d[i])); orc_diag_sp((float _Complex *)r->st, n, p1, p2, t); }
    else orc_diag_dp((double _Complex *)r->st, n, p1, p2, d);
  }
  else if (!strcmp(name, "Apply2QubitGate")) {
    unsigned a = pu(&s), b = pu(&s); double _Complex m[16]; parse_m(&s, 16, m);
    const unsigned ph = (unsigned)r->map[a], pl = (unsigned)r->map[b];
    if (r->sp) { float _Complex t[16]; for (int i = 0; i < 16; ++i) t[i] = CMPLXF((float)creal(m[i]), (float)cimag(m[i])); orc_2q_sp((float _Complex *)r->st, n, ph, pl, t); }
    else orc_2q_dp((double _Complex *)r->st, n, ph, pl, m);
  }
  else if (!strcmp(name, "PermuteQubits")) {
    while (*s == ' ') ++s;
    const int inverse = !strncmp(s, "inverse", 7);
    s += inverse ? 7 : 6;
    u64 m[64], nmap[64], nimap[64];
    for (int i = 0; i < n; ++i) m[i] = pu(&s);
    if (inverse) { for (int i = 0; i < n; ++i) { nimap[i] = m[i]; nmap[m[i]] = (u64)i; } }
    else { for (int i = 0; i < n; ++i) { nmap[i] = m[i]; nimap[m[i]] = (u64)i; } }
    int same = 1;
    for (int i = 0; i < n; ++i) same &= (r->imap[i] == nimap[i]);
    if (!same) {
      if (r->sp) orc_permute_sp((float _Complex *)r->st, n, r->map, nimap);
      else orc_permute_dp((double _Complex *)r->st, n, r->map, nimap);
    }
    memcpy(r->map, nmap, sizeof nmap);
    memcpy(r->imap, nimap, sizeof nimap);
  }
  else if (!strcmp(name, "EmulateSwap")) {
    /* qureg_permute.cpp:45-53 / permutation.hpp:193-201. */
    unsigned a = pu(&s), b = pu(&s);
    const u64 pa = r->map[a], pb = r->map[b];
    r->map[a] = pb; r->map[b] = pa; r->imap[pa] = b; r->imap[pb] = a;
  }
  else if (!strcmp(name, "CollapseQubit")) {
    unsigned q = pu(&s), v = pu(&s);
    if (r->sp) orc_collapse_sp((float _Complex *)r->st, n, (unsigned)r->map[q], (int)v);
    else orc_collapse_dp((double _Complex *)r->st, n, (unsigned)r->map[q], (int)v);
  }
  else if (!strcmp(name, "Normalize")) {
    /* qureg_utils.cpp:162-168: Type(1/global_norm, 0). */
    if (r->sp) { float nr = (float)orc_norm_sp((float _Complex *)r->st, n); orc_scale_sp((float _Complex *)r->st, n, CMPLXF(1 / nr, 0)); }
    else { double nr = orc_norm_dp((double _Complex *)r->st, n); orc_scale_dp((double _Complex *)r->st, n, CX(1 / nr, 0)); }
  }
  else if (!strcmp(name, "TurnOnSpecializeV2")) r->spec2 = 1;
  else if (!strcmp(name, "TurnOffSpecializeV2")) r->spec2 = 0;
  else if (!strcmp(name, "TurnOnSpecialize") || !strcmp(name, "TurnOffSpecialize") ||
           !strcmp(name, "TurnOnFusion") || !strcmp(name, "TurnOffFusion")) { /* same arithmetic */ }
  else if (!strcmp(name, "QaoaPhase")) {
    double g = pd(&s);
    ensure_diag(r);
    if (r->sp) orc_qaoa_phase_sp((float _Complex *)r->st, (float _Complex *)r->diag, n, (float)g);
    else orc_qaoa_phase_dp((double _Complex *)r->st, (double _Complex *)r->diag, n, g);
  }
  else if (!strcmp(name, "GetProbability")) { unsigned q = pu(&s); emit(r, "R GetProbability %u %.17g\n", q, prob(r, q)); }
  else if (!strcmp(name, "ExpectationValue")) {
    double coeff = pd(&s); unsigned k = pu(&s), qs[64], os[64];
    for (unsigned i = 0; i < k; ++i) { qs[i] = pu(&s); os[i] = pu(&s); }
    emit(r, "R ExpectationValue %.17g\n", expect(r, coeff, (int)k, qs, os));
  }
  else if (!strcmp(name, "ComputeNorm")) {
    emit(r, "R ComputeNorm %.17g\n", r->sp ? orc_norm_sp((float _Complex *)r->st, n) : orc_norm_dp((double _Complex *)r->st, n));
  }
  else if (!strcmp(name, "GetGlobalAmplitude")) {
    char *e; unsigned long long k = strtoull(s, &e, 10);
    u64 d = 0;
    for (int i = 0; i < n; ++i) d |= ((k >> r->imap[i]) & 1) << i; /* program2data_ */
    double re, im;
    if (r->sp) { float _Complex z = ((float _Complex *)r->st)[d]; re = crealf(z); im = cimagf(z); }
    else { double _Complex z = ((double _Complex *)r->st)[d]; re = creal(z); im = cimag(z); }
    emit(r, "R GetGlobalAmplitude %llu %.17g %.17g\n", k, re, im);
  }
  else if (!strcmp(name, "QaoaExpectation")) {
    ensure_diag(r);
    emit(r, "R QaoaExpectation %.17g\n", r->sp ? orc_qaoa_expect_sp((float _Complex *)r->st, (float _Complex *)r->diag, n)
                                                : orc_qaoa_expect_dp((double _Complex *)r->st, (double _Complex *)r->diag, n));
  }
  else if (!strcmp(name, "ZAll")) {
    emit(r, "R ZAll");
    for (int q = 0; q < n; ++q) emit(r, " %.17g", rbt(r, 1. - 2. * prob(r, (unsigned)q)));
    emit(r, "\n");
  }
  else { fprintf(stderr, "oracle: unknown op '%s'\n", name); abort(); }
}

/* Run a circuit in the text format of iqs_amd/circuits.py.
 * state_out receives the final state (2^n complex of the circuit's precision) if capacity
 * (in bytes) suffices; *results receives a malloc'ed string of "R ..." lines (free with
 * orc_free).  Returns 0 on success. */
int orc_run(const char *text, void *state_out, u64 state_cap, char **results)
{
  Reg r;
  memset(&r, 0, sizeof r);
  char *buf = strdup(text);
  char *save = NULL;
  char *line = strtok_r(buf, "\n", &save);
  int init_plus = 0;
  u64 base = 0;
  for (; line; line = strtok_r(NULL, "\n", &save)) {
    char tag[64] = {0};
    if (sscanf(line, "%63s", tag) != 1) continue;
    char *s = line + strspn(line, " ") + strlen(tag);
    if (!strcmp(tag, "begin")) break;
    if (!strcmp(tag, "qubits")) r.n = (int)pu(&s);
    else if (!strcmp(tag, "precision")) { while (*s == ' ') ++s; r.sp = !strncmp(s, "sp", 2); }
    else if (!strcmp(tag, "init")) { while (*s == ' ') ++s; if (!strncmp(s, "plus", 4)) init_plus = 1; else { s += 4; base = strtoull(s, NULL, 10); } }
    else if (!strcmp(tag, "graph")) {
      pu(&s); r.ne = (int)pu(&s);
      for (int e = 0; e < r.ne; ++e) { r.eu[e] = (int)pu(&s); r.ev[e] = (int)pu(&s); r.ew[e] = pd(&s); }
    }
  }
  const u64 N = (u64)1 << r.n;
  const size_t esz = r.sp ? sizeof(float _Complex) : sizeof(double _Complex);
  r.st = calloc(N, esz);
  for (int i = 0; i < r.n; ++i) r.map[i] = r.imap[i] = (u64)i;
  if (init_plus) {
    /* Initialize("++++"), qureg_init.cpp:344-348: BaseType(1/sqrt(GlobalSize())). */
    const double a = 1. / sqrt((double)N);
    for (u64 i = 0; i < N; ++i) {
      if (r.sp) ((float _Complex *)r.st)[i] = CMPLXF((float)a, 0);
      else ((double _Complex *)r.st)[i] = CX(a, 0);
    }
  } else {
    if (r.sp) ((float _Complex *)r.st)[base] = 1; else ((double _Complex *)r.st)[base] = 1;
  }
  r.out = (char *)malloc(1024);
  r.out[0] = 0;
  r.out_cap = 1024;
  for (line = strtok_r(NULL, "\n", &save); line; line = strtok_r(NULL, "\n", &save)) {
    char name[64] = {0};
    if (line[0] == '#' || sscanf(line, "%63s", name) != 1) continue;
    char *s = line + strspn(line, " ") + strlen(name);
    run_op(&r, name, s);
  }
  if (state_out && state_cap >= N * esz) memcpy(state_out, r.st, N * esz);
  *results = r.out;
  free(r.st);
  free(r.diag);
  free(buf);
  return 0;
}

void orc_free(void *p) { free(p); }
