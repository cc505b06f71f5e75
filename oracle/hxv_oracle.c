/*
 * ORACLE -- test infrastructure, not part of the product.
 *
 * A CPU FP64 restatement of the reference's zero-, one- and two-site
 * effective-Hamiltonian application, written to check the HIP engine and to
 * serve as the timed CPU baseline.  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load it.  It follows the reference literally:
 *
 *   two-site basis      Qbasis::pullData / combine / outer_num / inner_num /
 *                       leftOffset            tensors/Qbasis.h:228-254,335-350,
 *                                             378-388,440-462,580-590,686-798
 *   block structure     precalc_blockStructure (two-site)
 *                                             tensors/DmrgIndexGymnastics.h:698-856
 *                       AAWWAA                tensors/DmrgIndexGymnastics.h:198-262
 *                       precalc_blockStructure (one-site) + LAWA
 *                                             tensors/DmrgIndexGymnastics.h:485-590, 32-77
 *   one-site OxV        pivot/DmrgPivotMatrix1.h:131-332
 *   zero-site HxV       pivot/DmrgPivotMatrix0.h:31-96
 *   coefficients        coeff_tensorProd / coeff_Apair / coeff_HPsi / coeff_dot
 *                                             symmetry/SU2.h:284-290,388-397,455-466,508-519
 *                       phase                 symmetry/functions.h:50-55
 *                       6j / 9j (GSL, 2j = q-1) symbol values   symmetry/SU2Wrappers.h:71-94
 *   OxV                 per-term  Vout[s1s3][q13] += f*W12(a,w)*W34(w,b) * (L[qL][a]*Vin*R[qR][b])
 *                                             pivot/DmrgPivotMatrix2.h:85-273
 *   dot                 sum tr(V1^T V2) coeff_dot(out)   pivot/DmrgPivotVector.h:296-339
 *
 * The builder entry points have the same signatures as the engine's
 * (include/vmps_hxv.h) with the prefix oracle_ instead of vmps_.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define OK 0
#define ERR_ARG 1
#define ERR_LAYOUT 2
#define ERR_OOM 3
#define ERR_STATE 6

static __thread char g_err[512];
static int fail(int st, const char* msg) {
  snprintf(g_err, sizeof g_err, "%s", msg);
  return st;
}
const char* oracle_last_error(void) { return g_err; }

/* ------------------------------------------------------------------ */
typedef struct {
  int v[2];
} qn_t;

typedef struct {
  int nq;
  int kind[2]; /* 1 = U(1), 2 = SU(2) */
} sym_t;

static int qeq(qn_t a, qn_t b) { return a.v[0] == b.v[0] && a.v[1] == b.v[1]; }
static int qcmp(qn_t a, qn_t b) {
  if (a.v[0] != b.v[0]) return a.v[0] < b.v[0] ? -1 : 1;
  if (a.v[1] != b.v[1]) return a.v[1] < b.v[1] ? -1 : 1;
  return 0;
}
static qn_t qread(const int* p, int nq) {
  qn_t q;
  q.v[0] = p[0];
  q.v[1] = nq > 1 ? p[1] : 0;
  return q;
}

/* reduceSilent: U1 -> {a+b}; SU2 -> |a-b|+1 .. a+b-1 step 2; products:
 * Cartesian, first component outer (symmetry/U1.h:171-177, SU2.h:176-183,
 * S1xS2.h:195-210).  Returns the count written to out (at most 64). */
static int reduce(const sym_t* S, qn_t a, qn_t b, qn_t* out) {
  int c0[64], c1[64], n0 = 0, n1 = 0;
  for (int i = 0; i < S->nq; ++i) {
    int* c = i == 0 ? c0 : c1;
    int* n = i == 0 ? &n0 : &n1;
    if (S->kind[i] == 1) {
      c[(*n)++] = a.v[i] + b.v[i];
    } else {
      int lo = abs(a.v[i] - b.v[i]) + 1, hi = abs(a.v[i] + b.v[i]) - 1;
      for (int q = lo; q <= hi && *n < 64; q += 2) c[(*n)++] = q;
    }
  }
  int cnt = 0;
  if (S->nq == 1) {
    for (int i = 0; i < n0; ++i) {
      out[cnt].v[0] = c0[i];
      out[cnt].v[1] = 0;
      ++cnt;
    }
  } else {
    for (int i = 0; i < n0; ++i)
      for (int j = 0; j < n1 && cnt < 64; ++j) {
        out[cnt].v[0] = c0[i];
        out[cnt].v[1] = c1[j];
        ++cnt;
      }
  }
  return cnt;
}

static int validate3(const sym_t* S, qn_t a, qn_t b, qn_t c) {
  for (int i = 0; i < S->nq; ++i) {
    if (S->kind[i] == 1) {
      if (a.v[i] + b.v[i] != c.v[i]) return 0;
    } else {
      if (!(c.v[i] - 1 >= abs(a.v[i] - b.v[i]) && c.v[i] - 1 <= a.v[i] + b.v[i] - 2)) return 0;
    }
  }
  return 1;
}

static int non_abelian(const sym_t* S) {
  for (int i = 0; i < S->nq; ++i)
    if (S->kind[i] == 2) return 1;
  return 0;
}

/* ---- Wigner symbols, Racah formula in long double ------------------- */
static long double g_fact[1200];
static void init_fact(void) {
  if (g_fact[0] == 1.0L) return;
  long double f = 1.0L;
  g_fact[0] = 1.0L;
  for (int i = 1; i < 1200; ++i) {
    f *= (long double)i;
    g_fact[i] = f;
  }
}
static int tri(int a, int b, int c) {
  if (a < 0 || b < 0 || c < 0 || ((a + b + c) & 1)) return 0;
  return c >= abs(a - b) && c <= a + b;
}
static long double tri_delta(int a, int b, int c) {
  return sqrtl(g_fact[(a + b - c) / 2] * g_fact[(a - b + c) / 2] * g_fact[(b + c - a) / 2] / g_fact[(a + b + c) / 2 + 1]);
}
/* {j1 j2 j3; j4 j5 j6} with arguments 2j */
double oracle_wigner6j(int a, int b, int c, int d, int e, int f) {
  init_fact();
  if (!tri(a, b, c) || !tri(a, e, f) || !tri(d, b, f) || !tri(d, e, c)) return 0.0;
  int t1 = (a + b + c) / 2, t2 = (a + e + f) / 2, t3 = (d + b + f) / 2, t4 = (d + e + c) / 2;
  int u1 = (a + b + d + e) / 2, u2 = (b + c + e + f) / 2, u3 = (a + c + d + f) / 2;
  int lo = t1;
  if (t2 > lo) lo = t2;
  if (t3 > lo) lo = t3;
  if (t4 > lo) lo = t4;
  int hi = u1;
  if (u2 < hi) hi = u2;
  if (u3 < hi) hi = u3;
  long double s = 0.0L;
  for (int z = lo; z <= hi; ++z) {
    long double term = g_fact[z + 1] / (g_fact[z - t1] * g_fact[z - t2] * g_fact[z - t3] * g_fact[z - t4] *
                                        g_fact[u1 - z] * g_fact[u2 - z] * g_fact[u3 - z]);
    s += (z % 2) ? -term : term;
  }
  return (double)(s * tri_delta(a, b, c) * tri_delta(a, e, f) * tri_delta(d, b, f) * tri_delta(d, e, c));
}
/* {j1 j2 j3; j4 j5 j6; j7 j8 j9} with arguments 2j: sum over x of
 * (-1)^{2x} (2x+1) {j1 j4 j7; j8 j9 x} {j2 j5 j8; j4 x j6} {j3 j6 j9; x j1 j2} */
double oracle_wigner9j(int a, int b, int c, int d, int e, int f, int g, int h, int i) {
  if (!tri(a, b, c) || !tri(d, e, f) || !tri(g, h, i) || !tri(a, d, g) || !tri(b, e, h) || !tri(c, f, i)) return 0.0;
  int lo = abs(a - i), hi = a + i;
  if (abs(d - h) > lo) lo = abs(d - h);
  if (abs(b - f) > lo) lo = abs(b - f);
  if (d + h < hi) hi = d + h;
  if (b + f < hi) hi = b + f;
  long double s = 0.0L;
  for (int x = lo; x <= hi; x += 2) {
    long double p = (long double)oracle_wigner6j(a, d, g, h, i, x) * (long double)oracle_wigner6j(b, e, h, d, x, f) *
                    (long double)oracle_wigner6j(c, f, i, x, a, b) * (long double)(x + 1);
    s += (x % 2) ? -p : p;
  }
  return (double)s;
}

static double phase(int q) { return (q % 2) ? -1.0 : 1.0; }

/* q = 2S+1 arguments */
static double c9(const sym_t* S, const qn_t* q, int buildR) {
  double out = 1.0;
  for (int k = 0; k < S->nq; ++k) {
    if (S->kind[k] != 2) continue;
    int a[9];
    for (int j = 0; j < 9; ++j) a[j] = q[j].v[k];
    double w = oracle_wigner9j(a[0] - 1, a[1] - 1, a[2] - 1, a[3] - 1, a[4] - 1, a[5] - 1, a[6] - 1, a[7] - 1, a[8] - 1);
    w *= sqrt((double)(a[6] * a[7] * a[2] * a[5]));
    if (buildR) w *= (double)a[8] / (double)a[6];
    out *= w;
  }
  return out;
}
double oracle_coeff_HPsi(int nq, const int* kinds, const int* q9) {
  sym_t S = {nq, {kinds[0], nq > 1 ? kinds[1] : 1}};
  qn_t q[9];
  for (int j = 0; j < 9; ++j) q[j] = qread(q9 + j * nq, nq);
  return c9(&S, q, 0);
}
double oracle_coeff_buildR(int nq, const int* kinds, const int* q9) {
  sym_t S = {nq, {kinds[0], nq > 1 ? kinds[1] : 1}};
  qn_t q[9];
  for (int j = 0; j < 9; ++j) q[j] = qread(q9 + j * nq, nq);
  return c9(&S, q, 1);
}
/* coeff_buildL (symmetry/SU2.h:481-492) has the same form as coeff_HPsi */
double oracle_coeff_buildL(int nq, const int* kinds, const int* q9) { return oracle_coeff_HPsi(nq, kinds, q9); }
static double c_apair(const sym_t* S, qn_t q1, qn_t q2, qn_t q3, qn_t q4, qn_t q5, qn_t q6) {
  double out = 1.0;
  for (int k = 0; k < S->nq; ++k) {
    if (S->kind[k] != 2) continue;
    int a = q1.v[k], b = q2.v[k], c = q3.v[k], d = q4.v[k], e = q5.v[k], f = q6.v[k];
    out *= oracle_wigner6j(a - 1, b - 1, c - 1, d - 1, e - 1, f - 1) * sqrt((double)(c * f)) *
           phase((a + b + d + e - 4) / 2);
  }
  return out;
}
double oracle_coeff_Apair(int nq, const int* kinds, const int* q6) {
  sym_t S = {nq, {kinds[0], nq > 1 ? kinds[1] : 1}};
  qn_t q[6];
  for (int j = 0; j < 6; ++j) q[j] = qread(q6 + j * nq, nq);
  return c_apair(&S, q[0], q[1], q[2], q[3], q[4], q[5]);
}
static double c_dot(const sym_t* S, qn_t q) {
  double w = 1.0;
  for (int k = 0; k < S->nq; ++k)
    if (S->kind[k] == 2) w *= (double)q.v[k];
  return w;
}

/* ------------------------------------------------------------------ */
typedef struct {
  int64_t rows, cols;
  double* data; /* NULL when absent */
} mat_t;
typedef struct {
  qn_t in, out, mid;
  int adim;
  mat_t* blk;
} env_t;
typedef struct {
  int n, cap;
  env_t* s;
} envlist_t;
typedef struct {
  qn_t in, out;
  int rows, cols;
  int64_t nnz;
  int *row, *col; /* column-major order */
  double* val;
} wblk_t;
typedef struct {
  int n, cap;
  wblk_t* b;
} wbiped_t;
typedef struct {
  qn_t in, out;
  int64_t rows, cols, off;
} vblk_t;
typedef struct {
  int n, cap;
  vblk_t* b;
} vbiped_t;

typedef struct {
  qn_t q1, q2;
  int64_t dim1, dim2;
} src_t;
typedef struct {
  qn_t q;
  int64_t outer;
  int nsrc;
  src_t src[64];
} hist_t;

typedef struct {
  sym_t S;
  int nsite; /* 2: PivotMatrix2, 1: PivotMatrix1 (W, qloc, qOp in the 12 slots), 0: PivotMatrix0 */
  int nloc12, nloc34, nop12, nop34;
  qn_t *qloc12, *qloc34, *qop12, *qop34;
  envlist_t L, R;
  wbiped_t *W12, *W34;
  /* two-site basis */
  int ns;
  qn_t* qs;
  int nh;
  hist_t* hist;
  int64_t *inner12, *inner34;
  vbiped_t* V;
  int64_t vsize;
  /* block structure */
  int has_bs;
  int64_t nlhs, nrhs;
  int64_t* qlhs;      /* 2 per lhs */
  int64_t* rhs_begin; /* nlhs+1 */
  int64_t* qrhs;      /* 12 per rhs */
  double* fac;
  /* projectors */
  int nproj;
  double Epen;
  double* A0;
} opivot_t;

#define W12AT(P, s1, s2, k) (&(P)->W12[((s1) * (P)->nloc12 + (s2)) * (P)->nop12 + (k)])
#define W34AT(P, s1, s2, k) (&(P)->W34[((s1) * (P)->nloc34 + (s2)) * (P)->nop34 + (k)])

static void* xrealloc(void* p, size_t n) {
  void* q = realloc(p, n ? n : 1);
  if (!q) {
    fprintf(stderr, "oracle: out of memory\n");
    abort();
  }
  return q;
}

/* Qbasis::pullData of a list of quantum numbers and inner_num(s) */
static void pull_inner(const qn_t* qs, int n, int64_t* inner, int64_t* plain_of) {
  qn_t eq[256];
  int64_t enm[256], epl[256];
  int ne = 0;
  int64_t curr = 0;
  for (int s = 0; s < n; ++s) {
    int f = -1;
    for (int e = 0; e < ne; ++e)
      if (qeq(eq[e], qs[s])) f = e;
    if (f < 0) {
      eq[ne] = qs[s];
      enm[ne] = curr;
      epl[ne] = 1;
      ++ne;
    } else {
      epl[f] += 1;
    }
    curr += 1;
  }
  for (int s = 0; s < n; ++s) {
    inner[s] = 0;
    int done = 0;
    for (int e = 0; e < ne && !done; ++e)
      for (int64_t i = 0; i < epl[e] && !done; ++i)
        if ((int64_t)s == enm[e] + i) {
          inner[s] = i;
          done = 1;
        }
    for (int e = 0; e < ne; ++e)
      if (qeq(eq[e], qs[s])) plain_of[s] = epl[e];
  }
}

static int build_basis(opivot_t* P) {
  int64_t* plain12 = (int64_t*)xrealloc(NULL, sizeof(int64_t) * P->nloc12);
  int64_t* plain34 = (int64_t*)xrealloc(NULL, sizeof(int64_t) * P->nloc34);
  P->inner12 = (int64_t*)xrealloc(NULL, sizeof(int64_t) * P->nloc12);
  P->inner34 = (int64_t*)xrealloc(NULL, sizeof(int64_t) * P->nloc34);
  pull_inner(P->qloc12, P->nloc12, P->inner12, plain12);
  pull_inner(P->qloc34, P->nloc34, P->inner34, plain34);
  P->nh = 0;
  P->hist = NULL;
  qn_t red[64];
  for (int s1 = 0; s1 < P->nloc12; ++s1) {
    /* only the first state of each local sector enters (one entry per sector) */
    int first = 1;
    for (int t = 0; t < s1; ++t)
      if (qeq(P->qloc12[t], P->qloc12[s1])) first = 0;
    if (!first) continue;
    for (int s3 = 0; s3 < P->nloc34; ++s3) {
      int first3 = 1;
      for (int t = 0; t < s3; ++t)
        if (qeq(P->qloc34[t], P->qloc34[s3])) first3 = 0;
      if (!first3) continue;
      int nr = reduce(&P->S, P->qloc12[s1], P->qloc34[s3], red);
      for (int r = 0; r < nr; ++r) {
        int h = -1;
        for (int j = 0; j < P->nh; ++j)
          if (qeq(P->hist[j].q, red[r])) h = j;
        if (h < 0) {
          P->hist = (hist_t*)xrealloc(P->hist, sizeof(hist_t) * (P->nh + 1));
          h = P->nh++;
          P->hist[h].q = red[r];
          P->hist[h].nsrc = 0;
        }
        hist_t* H = &P->hist[h];
        int have = 0;
        for (int j = 0; j < H->nsrc; ++j)
          if (qeq(H->src[j].q1, P->qloc12[s1]) && qeq(H->src[j].q2, P->qloc34[s3])) have = 1;
        if (!have && H->nsrc < 64) {
          H->src[H->nsrc].q1 = P->qloc12[s1];
          H->src[H->nsrc].q2 = P->qloc34[s3];
          H->src[H->nsrc].dim1 = plain12[s1];
          H->src[H->nsrc].dim2 = plain34[s3];
          H->nsrc++;
        }
      }
    }
  }
  /* sort sources of each sector by (q1, q2); sectors by q (insertion sorts) */
  for (int h = 0; h < P->nh; ++h) {
    hist_t* H = &P->hist[h];
    for (int i = 1; i < H->nsrc; ++i)
      for (int j = i; j > 0; --j) {
        int c = qcmp(H->src[j - 1].q1, H->src[j].q1);
        if (c == 0) c = qcmp(H->src[j - 1].q2, H->src[j].q2);
        if (c > 0) {
          src_t t = H->src[j];
          H->src[j] = H->src[j - 1];
          H->src[j - 1] = t;
        }
      }
  }
  for (int i = 1; i < P->nh; ++i)
    for (int j = i; j > 0 && qcmp(P->hist[j - 1].q, P->hist[j].q) > 0; --j) {
      hist_t t = P->hist[j];
      P->hist[j] = P->hist[j - 1];
      P->hist[j - 1] = t;
    }
  int64_t num = 0;
  P->ns = 0;
  P->qs = NULL;
  for (int h = 0; h < P->nh; ++h) {
    P->hist[h].outer = num;
    int64_t d = 0;
    for (int j = 0; j < P->hist[h].nsrc; ++j) d += P->hist[h].src[j].dim1 * P->hist[h].src[j].dim2;
    P->qs = (qn_t*)xrealloc(P->qs, sizeof(qn_t) * (P->ns + d));
    for (int64_t i = 0; i < d; ++i) P->qs[P->ns++] = P->hist[h].q;
    num += d;
  }
  free(plain12);
  free(plain34);
  return OK;
}

/* outer_num(qmerge) + leftOffset(qmerge, {q1,q3}, {inner(s1), inner(s3)}) */
static int64_t basis_index(const opivot_t* P, int s1, int s3, qn_t qm) {
  for (int h = 0; h < P->nh; ++h) {
    if (!qeq(P->hist[h].q, qm)) continue;
    int64_t out = P->hist[h].outer;
    for (int j = 0; j < P->hist[h].nsrc; ++j) {
      const src_t* S = &P->hist[h].src[j];
      if (!(qeq(S->q1, P->qloc12[s1]) && qeq(S->q2, P->qloc34[s3]))) {
        out += S->dim1 * S->dim2;
      } else {
        for (int64_t k = 0; k < S->dim1 * S->dim2; ++k) {
          int64_t t = k / S->dim1;
          if (!(k - S->dim1 * t == P->inner12[s1] && t == P->inner34[s3]))
            out += 1;
          else
            break;
        }
        return out;
      }
    }
  }
  return -1;
}

int oracle_pivot2_create(int nq, const int* kinds, int nloc12, const int* qloc12, int nloc34, const int* qloc34,
                         int nop12, const int* qop12, int nop34, const int* qop34, void** out) {
  if (!out || nq < 1 || nq > 2 || nloc12 <= 0 || nloc34 <= 0 || nop12 <= 0 || nop34 <= 0)
    return fail(ERR_ARG, "bad arguments");
  opivot_t* P = (opivot_t*)calloc(1, sizeof(opivot_t));
  P->S.nq = nq;
  P->S.kind[0] = kinds[0];
  P->S.kind[1] = nq > 1 ? kinds[1] : 1;
  P->nloc12 = nloc12;
  P->nloc34 = nloc34;
  P->nop12 = nop12;
  P->nop34 = nop34;
  P->qloc12 = (qn_t*)xrealloc(NULL, sizeof(qn_t) * nloc12);
  P->qloc34 = (qn_t*)xrealloc(NULL, sizeof(qn_t) * nloc34);
  P->qop12 = (qn_t*)xrealloc(NULL, sizeof(qn_t) * nop12);
  P->qop34 = (qn_t*)xrealloc(NULL, sizeof(qn_t) * nop34);
  for (int i = 0; i < nloc12; ++i) P->qloc12[i] = qread(qloc12 + i * nq, nq);
  for (int i = 0; i < nloc34; ++i) P->qloc34[i] = qread(qloc34 + i * nq, nq);
  for (int i = 0; i < nop12; ++i) P->qop12[i] = qread(qop12 + i * nq, nq);
  for (int i = 0; i < nop34; ++i) P->qop34[i] = qread(qop34 + i * nq, nq);
  P->W12 = (wbiped_t*)calloc((size_t)nloc12 * nloc12 * nop12, sizeof(wbiped_t));
  P->W34 = (wbiped_t*)calloc((size_t)nloc34 * nloc34 * nop34, sizeof(wbiped_t));
  P->nsite = 2;
  build_basis(P);
  P->V = (vbiped_t*)calloc(P->ns ? P->ns : 1, sizeof(vbiped_t));
  *out = P;
  return OK;
}

/* PivotMatrix1 (pivot/DmrgPivotMatrix1.h:11-66): one site, PivotVector
 * components = local states */
int oracle_pivot1_create(int nq, const int* kinds, int nloc, const int* qloc, int nop, const int* qop, void** out) {
  if (!out || nq < 1 || nq > 2 || nloc <= 0 || nop <= 0) return fail(ERR_ARG, "bad arguments");
  opivot_t* P = (opivot_t*)calloc(1, sizeof(opivot_t));
  P->S.nq = nq;
  P->S.kind[0] = kinds[0];
  P->S.kind[1] = nq > 1 ? kinds[1] : 1;
  P->nsite = 1;
  P->nloc12 = nloc;
  P->nop12 = nop;
  P->qloc12 = (qn_t*)xrealloc(NULL, sizeof(qn_t) * nloc);
  P->qop12 = (qn_t*)xrealloc(NULL, sizeof(qn_t) * nop);
  for (int i = 0; i < nloc; ++i) P->qloc12[i] = qread(qloc + i * nq, nq);
  for (int i = 0; i < nop; ++i) P->qop12[i] = qread(qop + i * nq, nq);
  P->W12 = (wbiped_t*)calloc((size_t)nloc * nloc * nop, sizeof(wbiped_t));
  P->ns = nloc;
  P->qs = (qn_t*)xrealloc(NULL, sizeof(qn_t) * nloc);
  memcpy(P->qs, P->qloc12, sizeof(qn_t) * nloc);
  P->V = (vbiped_t*)calloc(nloc, sizeof(vbiped_t));
  *out = P;
  return OK;
}

/* PivotMatrix0 (pivot/DmrgPivotMatrix0.h:10-27): L and R only, one vector
 * component */
int oracle_pivot0_create(int nq, const int* kinds, void** out) {
  if (!out || nq < 1 || nq > 2) return fail(ERR_ARG, "bad arguments");
  opivot_t* P = (opivot_t*)calloc(1, sizeof(opivot_t));
  P->S.nq = nq;
  P->S.kind[0] = kinds[0];
  P->S.kind[1] = nq > 1 ? kinds[1] : 1;
  P->nsite = 0;
  P->ns = 1;
  P->qs = (qn_t*)calloc(1, sizeof(qn_t));
  for (int i = 0; i < nq; ++i) P->qs[0].v[i] = P->S.kind[i] == 2 ? 1 : 0;
  P->V = (vbiped_t*)calloc(1, sizeof(vbiped_t));
  *out = P;
  return OK;
}

static void free_env(envlist_t* E) {
  for (int i = 0; i < E->n; ++i) {
    for (int a = 0; a < E->s[i].adim; ++a) free(E->s[i].blk[a].data);
    free(E->s[i].blk);
  }
  free(E->s);
}

void oracle_pivot2_destroy(void* h) {
  opivot_t* P = (opivot_t*)h;
  if (!P) return;
  free_env(&P->L);
  free_env(&P->R);
  int n12 = P->nloc12 * P->nloc12 * P->nop12, n34 = P->nloc34 * P->nloc34 * P->nop34;
  for (int i = 0; i < n12 + n34; ++i) {
    wbiped_t* W = i < n12 ? &P->W12[i] : &P->W34[i - n12];
    for (int j = 0; j < W->n; ++j) {
      free(W->b[j].row);
      free(W->b[j].col);
      free(W->b[j].val);
    }
    free(W->b);
  }
  free(P->W12);
  free(P->W34);
  for (int s = 0; s < P->ns; ++s) free(P->V[s].b);
  free(P->V);
  free(P->qs);
  free(P->hist);
  free(P->inner12);
  free(P->inner34);
  free(P->qloc12);
  free(P->qloc34);
  free(P->qop12);
  free(P->qop34);
  free(P->qlhs);
  free(P->rhs_begin);
  free(P->qrhs);
  free(P->fac);
  free(P->A0);
  free(P);
}

int oracle_pivot2_local_basis(const void* h, int* ns, int* qloc) {
  const opivot_t* P = (const opivot_t*)h;
  *ns = P->ns;
  if (qloc)
    for (int s = 0; s < P->ns; ++s)
      for (int i = 0; i < P->S.nq; ++i) qloc[s * P->S.nq + i] = P->qs[s].v[i];
  return OK;
}

int oracle_pivot2_add_env(void* h, int which, const int* qin, const int* qout, const int* qmid, int adim,
                          const int64_t* rows, const int64_t* cols, const double* data) {
  opivot_t* P = (opivot_t*)h;
  if (which != 0 && which != 1) return fail(ERR_ARG, "which");
  envlist_t* E = which == 0 ? &P->L : &P->R;
  if (E->n == E->cap) {
    E->cap = E->cap ? 2 * E->cap : 64;
    E->s = (env_t*)xrealloc(E->s, sizeof(env_t) * E->cap);
  }
  env_t* e = &E->s[E->n++];
  e->in = qread(qin, P->S.nq);
  e->out = qread(qout, P->S.nq);
  e->mid = qread(qmid, P->S.nq);
  e->adim = adim;
  e->blk = (mat_t*)calloc(adim ? adim : 1, sizeof(mat_t));
  int64_t src = 0;
  for (int a = 0; a < adim; ++a) {
    e->blk[a].rows = rows[a];
    e->blk[a].cols = cols[a];
    int64_t n = rows[a] * cols[a];
    if (n > 0) {
      e->blk[a].data = (double*)xrealloc(NULL, sizeof(double) * n);
      memcpy(e->blk[a].data, data + src, sizeof(double) * n);
      src += n;
    }
  }
  return OK;
}

int oracle_pivot2_add_W(void* h, int which, int s1, int s2, int k, const int* qin, const int* qout, int rows, int cols,
                        int64_t nnz, const int* row, const int* col, const double* val) {
  opivot_t* P = (opivot_t*)h;
  wbiped_t* W;
  if (which == 12) {
    if (s1 < 0 || s1 >= P->nloc12 || s2 < 0 || s2 >= P->nloc12 || k < 0 || k >= P->nop12) return fail(ERR_ARG, "idx");
    W = W12AT(P, s1, s2, k);
  } else if (which == 34) {
    if (s1 < 0 || s1 >= P->nloc34 || s2 < 0 || s2 >= P->nloc34 || k < 0 || k >= P->nop34) return fail(ERR_ARG, "idx");
    W = W34AT(P, s1, s2, k);
  } else {
    return fail(ERR_ARG, "which");
  }
  if (W->n == W->cap) {
    W->cap = W->cap ? 2 * W->cap : 8;
    W->b = (wblk_t*)xrealloc(W->b, sizeof(wblk_t) * W->cap);
  }
  wblk_t* b = &W->b[W->n++];
  b->in = qread(qin, P->S.nq);
  b->out = qread(qout, P->S.nq);
  b->rows = rows;
  b->cols = cols;
  b->nnz = nnz;
  b->row = (int*)xrealloc(NULL, sizeof(int) * (nnz ? nnz : 1));
  b->col = (int*)xrealloc(NULL, sizeof(int) * (nnz ? nnz : 1));
  b->val = (double*)xrealloc(NULL, sizeof(double) * (nnz ? nnz : 1));
  /* column-major (InnerIterator) order: selection by (col, row) */
  int64_t* idx = (int64_t*)xrealloc(NULL, sizeof(int64_t) * (nnz ? nnz : 1));
  for (int64_t i = 0; i < nnz; ++i) idx[i] = i;
  for (int64_t i = 1; i < nnz; ++i)
    for (int64_t j = i; j > 0; --j) {
      int64_t x = idx[j - 1], y = idx[j];
      if (col[x] > col[y] || (col[x] == col[y] && row[x] > row[y])) {
        idx[j - 1] = y;
        idx[j] = x;
      } else {
        break;
      }
    }
  for (int64_t i = 0; i < nnz; ++i) {
    b->row[i] = row[idx[i]];
    b->col[i] = col[idx[i]];
    b->val[i] = val[idx[i]];
  }
  free(idx);
  return OK;
}

int oracle_pivot2_add_V_block(void* h, int s, const int* qin, const int* qout, int64_t rows, int64_t cols,
                              int64_t* offset) {
  opivot_t* P = (opivot_t*)h;
  if (s < 0 || s >= P->ns) return fail(ERR_ARG, "s out of range");
  vbiped_t* V = &P->V[s];
  if (V->n == V->cap) {
    V->cap = V->cap ? 2 * V->cap : 16;
    V->b = (vblk_t*)xrealloc(V->b, sizeof(vblk_t) * V->cap);
  }
  vblk_t* b = &V->b[V->n++];
  b->in = qread(qin, P->S.nq);
  b->out = qread(qout, P->S.nq);
  b->rows = rows;
  b->cols = cols;
  b->off = P->vsize;
  P->vsize += rows * cols;
  if (offset) *offset = b->off;
  P->has_bs = 0;
  return OK;
}

int oracle_pivot2_vsize(const void* h, int64_t* size) {
  *size = ((const opivot_t*)h)->vsize;
  return OK;
}

int oracle_pivot2_set_projectors(void* h, double Epenalty, int n, const double* A0) {
  opivot_t* P = (opivot_t*)h;
  free(P->A0);
  P->nproj = n;
  P->Epen = Epenalty;
  P->A0 = (double*)xrealloc(NULL, sizeof(double) * (size_t)(n * P->vsize + 1));
  memcpy(P->A0, A0, sizeof(double) * (size_t)n * P->vsize);
  return OK;
}

static int vfind(const vbiped_t* V, qn_t in, qn_t out) {
  for (int q = 0; q < V->n; ++q)
    if (qeq(V->b[q].in, in) && qeq(V->b[q].out, out)) return q;
  return -1;
}
static int wfind(const wbiped_t* W, qn_t in, qn_t out) {
  for (int q = 0; q < W->n; ++q)
    if (qeq(W->b[q].in, in) && qeq(W->b[q].out, out)) return q;
  return -1;
}
static int efind(const envlist_t* E, qn_t in, qn_t out, qn_t mid) {
  for (int q = 0; q < E->n; ++q)
    if (qeq(E->s[q].in, in) && qeq(E->s[q].out, out) && qeq(E->s[q].mid, mid)) return q;
  return -1;
}

/* growable list of (key = (s1s3, q13), val[12], factor) entries */
typedef struct {
  int64_t key0, key1;
  int64_t val[12];
  double f;
  int64_t seq;
} entry_t;

static int entry_cmp(const void* a, const void* b) {
  const entry_t* x = (const entry_t*)a;
  const entry_t* y = (const entry_t*)b;
  if (x->key0 != y->key0) return x->key0 < y->key0 ? -1 : 1;
  if (x->key1 != y->key1) return x->key1 < y->key1 ? -1 : 1;
  return x->seq < y->seq ? -1 : (x->seq > y->seq);
}

#define CGC_EPS 1e-14

static void push_entry(entry_t** ent, int64_t* nent, int64_t* cap, int64_t k0, int64_t k1, const int64_t* v, double f) {
  if (*nent == *cap) {
    *cap = *cap ? 2 * *cap : 1024;
    *ent = (entry_t*)xrealloc(*ent, sizeof(entry_t) * *cap);
  }
  entry_t* e = &(*ent)[*nent];
  e->key0 = k0;
  e->key1 = k1;
  memcpy(e->val, v, sizeof(int64_t) * 12);
  e->f = f;
  e->seq = *nent;
  ++*nent;
}

/* sort the entries by key (stable in visiting order) into qlhs/qrhs/fac */
static void store_entries(opivot_t* P, entry_t* ent, int64_t nent) {
  qsort(ent, (size_t)nent, sizeof(entry_t), entry_cmp);
  free(P->qlhs);
  free(P->rhs_begin);
  free(P->qrhs);
  free(P->fac);
  int64_t nl = 0;
  for (int64_t i = 0; i < nent; ++i)
    if (i == 0 || ent[i].key0 != ent[i - 1].key0 || ent[i].key1 != ent[i - 1].key1) ++nl;
  P->nlhs = nl;
  P->nrhs = nent;
  P->qlhs = (int64_t*)xrealloc(NULL, sizeof(int64_t) * (2 * nl + 1));
  P->rhs_begin = (int64_t*)xrealloc(NULL, sizeof(int64_t) * (nl + 1));
  P->qrhs = (int64_t*)xrealloc(NULL, sizeof(int64_t) * (12 * nent + 1));
  P->fac = (double*)xrealloc(NULL, sizeof(double) * (nent + 1));
  int64_t l = -1;
  for (int64_t i = 0; i < nent; ++i) {
    if (i == 0 || ent[i].key0 != ent[i - 1].key0 || ent[i].key1 != ent[i - 1].key1) {
      ++l;
      P->qlhs[2 * l] = ent[i].key0;
      P->qlhs[2 * l + 1] = ent[i].key1;
      P->rhs_begin[l] = i;
    }
    memcpy(&P->qrhs[12 * i], ent[i].val, sizeof(int64_t) * 12);
    P->fac[i] = ent[i].f;
  }
  P->rhs_begin[nl] = nent;
  P->has_bs = 1;
}

/* One-site precalc_blockStructure (tensors/DmrgIndexGymnastics.h:485-590)
 * with the LAWA matcher (:32-77) on Abra = Aket = Vin. */
static int precalc_one_site(opivot_t* P) {
  const sym_t* S = &P->S;
  const int NA = non_abelian(S);
  entry_t* ent = NULL;
  int64_t nent = 0, cap = 0;
  qn_t Routs[64], Rins[64], Rmids[64];
  for (int s1 = 0; s1 < P->nloc12; ++s1)
    for (int s2 = 0; s2 < P->nloc12; ++s2)
      for (int k = 0; k < P->nop12; ++k) {
        if (!validate3(S, P->qloc12[s2], P->qop12[k], P->qloc12[s1])) continue;
        const wbiped_t* W = W12AT(P, s1, s2, k);
        for (int qL = 0; qL < P->L.n; ++qL) {
          const env_t* Le = &P->L.s[qL];
          int nRo = reduce(S, Le->in, P->qloc12[s1], Routs);
          for (int a = 0; a < nRo; ++a) {
            int q1 = vfind(&P->V[s1], Le->in, Routs[a]);
            if (q1 < 0) continue;
            int nRi = reduce(S, Le->out, P->qloc12[s2], Rins);
            for (int b = 0; b < nRi; ++b) {
              int q2 = vfind(&P->V[s2], Le->out, Rins[b]);
              if (q2 < 0) continue;
              int nRm = reduce(S, Le->mid, P->qop12[k], Rmids);
              for (int c = 0; c < nRm; ++c) {
                int qW = wfind(W, Le->mid, Rmids[c]);
                if (qW < 0) continue;
                if (!validate3(S, Rins[b], Rmids[c], Routs[a])) continue;
                int qR = efind(&P->R, Rins[b], Routs[a], Rmids[c]);
                if (qR < 0) continue;
                const wblk_t* w = &W->b[qW];
                int empty = 1;
                for (int64_t i = 0; i < w->nnz; ++i) {
                  if (w->row[i] >= Le->adim || w->col[i] >= P->R.s[qR].adim) return fail(ERR_LAYOUT, "W index");
                  const mat_t* La = &Le->blk[w->row[i]];
                  const mat_t* Rb = &P->R.s[qR].blk[w->col[i]];
                  if (La->rows * La->cols > 0 && Rb->rows * Rb->cols > 0) empty = 0;
                }
                if (empty) continue;
                double f = 1.0;
                if (NA) {
                  qn_t a9[9] = {P->V[s2].b[q2].in, P->qloc12[s2], P->V[s2].b[q2].out,
                                Le->mid,           P->qop12[k],   Rmids[c],
                                P->V[s1].b[q1].in, P->qloc12[s1], P->V[s1].b[q1].out};
                  f = c9(S, a9, 0);
                }
                if (fabs(f) < CGC_EPS) continue;
                int64_t v[12] = {s2, q2, qL, qR, k, qW, 0, 0, 0, 0, 0, 0};
                push_entry(&ent, &nent, &cap, s1, q1, v, f);
              }
            }
          }
        }
      }
  store_entries(P, ent, nent);
  free(ent);
  return OK;
}

int oracle_pivot2_precalc(void* h) {
  opivot_t* P = (opivot_t*)h;
  if (P->nsite == 1) return precalc_one_site(P);
  if (P->nsite == 0) {
    P->nlhs = P->nrhs = 0;
    P->has_bs = 1;
    return OK;
  }
  const sym_t* S = &P->S;
  const int NA = non_abelian(S);
  entry_t* ent = NULL;
  int64_t nent = 0, cap = 0;
  qn_t qOps[64], qm13[64], qm24[64], qRouts[64], qRins[64], qWs[64], qRmids[64];
  for (int s1 = 0; s1 < P->nloc12; ++s1)
    for (int s2 = 0; s2 < P->nloc12; ++s2)
      for (int k12 = 0; k12 < P->nop12; ++k12) {
        if (!validate3(S, P->qloc12[s2], P->qop12[k12], P->qloc12[s1])) continue;
        for (int s3 = 0; s3 < P->nloc34; ++s3)
          for (int s4 = 0; s4 < P->nloc34; ++s4)
            for (int k34 = 0; k34 < P->nop34; ++k34) {
              if (!validate3(S, P->qloc34[s4], P->qop34[k34], P->qloc34[s3])) continue;
              int nOps = reduce(S, P->qop12[k12], P->qop34[k34], qOps);
              for (int io = 0; io < nOps; ++io) {
                qn_t qOp = qOps[io];
                int n13 = reduce(S, P->qloc12[s1], P->qloc34[s3], qm13);
                int n24 = reduce(S, P->qloc12[s2], P->qloc34[s4], qm24);
                for (int i13 = 0; i13 < n13; ++i13)
                  for (int i24 = 0; i24 < n24; ++i24) {
                    qn_t qmerge13 = qm13[i13], qmerge24 = qm24[i24];
                    int64_t s1s3 = basis_index(P, s1, s3, qmerge13);
                    int64_t s2s4 = basis_index(P, s2, s4, qmerge24);
                    if (s1s3 < 0 || s2s4 < 0) continue;
                    double cgc9 = 1.0;
                    if (NA) {
                      qn_t a9[9] = {P->qloc12[s2], P->qloc34[s4], qmerge24, P->qop12[k12], P->qop34[k34],
                                    qOp,           P->qloc12[s1], P->qloc34[s3], qmerge13};
                      cgc9 = c9(S, a9, 0);
                    }
                    if (fabs(cgc9) < CGC_EPS) continue;
                    const vbiped_t* A13 = &P->V[s1s3];
                    const vbiped_t* A24 = &P->V[s2s4];
                    const wbiped_t* W12 = W12AT(P, s1, s2, k12);
                    const wbiped_t* W34 = W34AT(P, s3, s4, k34);
                    for (int qL = 0; qL < P->L.n; ++qL) {
                      const env_t* Le = &P->L.s[qL];
                      /* AAWWAA */
                      int nRo = reduce(S, Le->in, qmerge13, qRouts);
                      int nRi = reduce(S, Le->out, qmerge24, qRins);
                      int nW = reduce(S, Le->mid, P->qop12[k12], qWs);
                      for (int a = 0; a < nRo; ++a) {
                        int q13 = vfind(A13, Le->in, qRouts[a]);
                        if (q13 < 0) continue;
                        for (int b = 0; b < nRi; ++b) {
                          int q24 = vfind(A24, Le->out, qRins[b]);
                          if (q24 < 0) continue;
                          for (int c = 0; c < nW; ++c) {
                            int qW12 = wfind(W12, Le->mid, qWs[c]);
                            if (qW12 < 0) continue;
                            int nRm = reduce(S, qWs[c], P->qop34[k34], qRmids);
                            for (int d = 0; d < nRm; ++d) {
                              int qW34 = wfind(W34, qWs[c], qRmids[d]);
                              if (qW34 < 0) continue;
                              if (!validate3(S, qRins[b], qRmids[d], qRouts[a])) continue;
                              /* hit: (qRin, qRout, qRmid), qW, q13, q24, qW12, qW34 */
                              int qR = efind(&P->R, qRins[b], qRouts[a], qRmids[d]);
                              double cgc6 = NA ? c_apair(S, Le->mid, P->qop12[k12], qWs[c], P->qop34[k34],
                                                         qRmids[d], qOp)
                                               : 1.0;
                              if (fabs(cgc6) < CGC_EPS) continue;
                              if (qR < 0) continue;
                              double cgcH = 1.0;
                              if (NA) {
                                qn_t a9[9] = {A24->b[q24].in, qmerge24, A24->b[q24].out,
                                              Le->mid,        qOp,      P->R.s[qR].mid,
                                              A13->b[q13].in, qmerge13, A13->b[q13].out};
                                cgcH = c9(S, a9, 0);
                              }
                              int64_t v[12] = {s2s4, q24, qL, qR, s1, s2, k12, qW12, s3, s4, k34, qW34};
                              push_entry(&ent, &nent, &cap, s1s3, q13, v, cgc6 * cgc9 * cgcH);
                            }
                          }
                        }
                      }
                    }
                  }
              }
            }
      }
  store_entries(P, ent, nent);
  free(ent);
  return OK;
}

int oracle_pivot2_set_blockstructure(void* h, int64_t nlhs, const int64_t* qlhs, const int64_t* rhs_begin,
                                     const int64_t* qrhs, const double* factors) {
  opivot_t* P = (opivot_t*)h;
  int64_t nrhs = nlhs ? rhs_begin[nlhs] : 0;
  free(P->qlhs);
  free(P->rhs_begin);
  free(P->qrhs);
  free(P->fac);
  P->nlhs = nlhs;
  P->nrhs = nrhs;
  P->qlhs = (int64_t*)xrealloc(NULL, sizeof(int64_t) * (2 * nlhs + 1));
  P->rhs_begin = (int64_t*)xrealloc(NULL, sizeof(int64_t) * (nlhs + 1));
  P->qrhs = (int64_t*)xrealloc(NULL, sizeof(int64_t) * (12 * nrhs + 1));
  P->fac = (double*)xrealloc(NULL, sizeof(double) * (nrhs + 1));
  memcpy(P->qlhs, qlhs, sizeof(int64_t) * 2 * nlhs);
  memcpy(P->rhs_begin, rhs_begin, sizeof(int64_t) * (nlhs + 1));
  memcpy(P->qrhs, qrhs, sizeof(int64_t) * 12 * nrhs);
  memcpy(P->fac, factors, sizeof(double) * nrhs);
  P->has_bs = 1;
  return OK;
}

int oracle_pivot2_get_blockstructure(const void* h, int64_t* nlhs, int64_t* nrhs, int64_t* qlhs, int64_t* rhs_begin,
                                     int64_t* qrhs, double* factors) {
  const opivot_t* P = (const opivot_t*)h;
  if (!P->has_bs) return fail(ERR_STATE, "no block structure");
  *nlhs = P->nlhs;
  *nrhs = P->nrhs;
  if (qlhs && rhs_begin && qrhs && factors) {
    memcpy(qlhs, P->qlhs, sizeof(int64_t) * 2 * P->nlhs);
    memcpy(rhs_begin, P->rhs_begin, sizeof(int64_t) * (P->nlhs + 1));
    memcpy(qrhs, P->qrhs, sizeof(int64_t) * 12 * P->nrhs);
    memcpy(factors, P->fac, sizeof(double) * P->nrhs);
  }
  return OK;
}

/* ---- dense kernels: C(m x n) = A(m x k) . B(k x n), column-major --------
 * Blocked over k and n so the inner loop runs down contiguous columns of A
 * and C; vectorised by the compiler at -O3 with -march x86-64-v3 / v4. */
static void dgemm_nn(int64_t m, int64_t n, int64_t k, const double* A, int64_t lda, const double* B, int64_t ldb,
                     double* C, int64_t ldc) {
  for (int64_t j = 0; j < n; ++j) {
    double* c = C + j * ldc;
    for (int64_t i = 0; i < m; ++i) c[i] = 0.0;
  }
  const int64_t KB = 256, NB = 64;
  for (int64_t k0 = 0; k0 < k; k0 += KB) {
    int64_t k1 = k0 + KB < k ? k0 + KB : k;
    for (int64_t j0 = 0; j0 < n; j0 += NB) {
      int64_t j1 = j0 + NB < n ? j0 + NB : n;
      for (int64_t j = j0; j < j1; ++j) {
        double* __restrict c = C + j * ldc;
        const double* b = B + j * ldb;
        int64_t p = k0;
        for (; p + 3 < k1; p += 4) {
          const double b0 = b[p], b1 = b[p + 1], b2 = b[p + 2], b3 = b[p + 3];
          const double* __restrict a0 = A + p * lda;
          const double* __restrict a1 = a0 + lda;
          const double* __restrict a2 = a1 + lda;
          const double* __restrict a3 = a2 + lda;
          for (int64_t i = 0; i < m; ++i) c[i] += a0[i] * b0 + a1[i] * b1 + a2[i] * b2 + a3[i] * b3;
        }
        for (; p < k1; ++p) {
          const double bp = b[p];
          const double* __restrict a = A + p * lda;
          for (int64_t i = 0; i < m; ++i) c[i] += a[i] * bp;
        }
      }
    }
  }
}

/* GFLOP/s of dgemm_nn on n x n x n (strength of the CPU baseline) */
double oracle_dgemm_gflops(int n, int reps) {
  double* A = (double*)xrealloc(NULL, sizeof(double) * n * n);
  double* B = (double*)xrealloc(NULL, sizeof(double) * n * n);
  double* C = (double*)xrealloc(NULL, sizeof(double) * n * n);
  for (int64_t i = 0; i < (int64_t)n * n; ++i) {
    A[i] = (double)(i % 17) * 0.01;
    B[i] = (double)(i % 13) * 0.02;
  }
  struct timespec t0, t1;
  clock_gettime(CLOCK_MONOTONIC, &t0);
  for (int r = 0; r < reps; ++r) dgemm_nn(n, n, n, A, n, B, n, C, n);
  clock_gettime(CLOCK_MONOTONIC, &t1);
  double s = (t1.tv_sec - t0.tv_sec) + 1e-9 * (t1.tv_nsec - t0.tv_nsec);
  double g = 2.0 * n * (double)n * n * reps / s * 1e-9;
  free(A);
  free(B);
  free(C);
  return g;
}

static double weighted_dot(const opivot_t* P, const double* x, const double* y) {
  double res = 0.0;
  for (int s = 0; s < P->ns; ++s)
    for (int q = 0; q < P->V[s].n; ++q) {
      const vblk_t* b = &P->V[s].b[q];
      double t = 0.0;
      for (int64_t i = 0; i < b->rows * b->cols; ++i) t += x[b->off + i] * y[b->off + i];
      if (b->rows * b->cols > 0) res += t * c_dot(&P->S, b->out);
    }
  return res;
}

int oracle_dot(const void* h, const double* x, const double* y, double* r) {
  *r = weighted_dot((const opivot_t*)h, x, y);
  return OK;
}

/* OxV for the lhs entries [l0, l1) into vout (which the caller zeroed) */
static int oxv_range(const opivot_t* P, const double* vin, double* vout, int64_t l0, int64_t l1) {
  double *t1 = NULL, *t2 = NULL;
  int64_t c1 = 0, c2 = 0;
  for (int64_t q = l0; q < l1; ++q) {
    int64_t s1s3 = P->qlhs[2 * q], q13 = P->qlhs[2 * q + 1];
    if (s1s3 < 0 || s1s3 >= P->ns || q13 < 0 || q13 >= P->V[s1s3].n) return ERR_LAYOUT;
    const vblk_t* O = &P->V[s1s3].b[q13];
    for (int64_t p = P->rhs_begin[q]; p < P->rhs_begin[q + 1]; ++p) {
      const int64_t* e = &P->qrhs[12 * p];
      int64_t s2s4 = e[0], q24 = e[1], qL = e[2], qR = e[3], s1 = e[4], s2 = e[5], k12 = e[6], qW12 = e[7], s3 = e[8],
              s4 = e[9], k34 = e[10], qW34 = e[11];
      if (s2s4 < 0 || s2s4 >= P->ns || q24 < 0 || q24 >= P->V[s2s4].n || qL < 0 || qL >= P->L.n || qR < 0 ||
          qR >= P->R.n)
        return ERR_LAYOUT;
      const wblk_t* w12 = &W12AT(P, s1, s2, k12)->b[qW12];
      const wblk_t* w34 = &W34AT(P, s3, s4, k34)->b[qW34];
      const vblk_t* Vb = &P->V[s2s4].b[q24];
      const env_t* Le = &P->L.s[qL];
      const env_t* Re = &P->R.s[qR];
      for (int64_t i = 0; i < w12->nnz; ++i)
        for (int64_t j = 0; j < w34->nnz; ++j) {
          int a = w12->row[i], w = w12->col[i], w2 = w34->row[j], b = w34->col[j];
          if (a >= Le->adim || b >= Re->adim) return ERR_LAYOUT;
          const mat_t* La = &Le->blk[a];
          const mat_t* Rb = &Re->blk[b];
          if (!(La->rows * La->cols > 0 && Rb->rows * Rb->cols > 0 && Vb->rows * Vb->cols > 0 && w == w2)) continue;
          if (La->cols != Vb->rows || Vb->cols != Rb->rows || La->rows != O->rows || Rb->cols != O->cols)
            return ERR_LAYOUT;
          double f = P->fac[p] * w12->val[i] * w34->val[j];
          int64_t m = La->rows, k = La->cols, n = Vb->cols, n2 = Rb->cols;
          if (m * n > c1) {
            c1 = m * n;
            t1 = (double*)xrealloc(t1, sizeof(double) * c1);
          }
          if (m * n2 > c2) {
            c2 = m * n2;
            t2 = (double*)xrealloc(t2, sizeof(double) * c2);
          }
          dgemm_nn(m, n, k, La->data, m, vin + Vb->off, k, t1, m);
          dgemm_nn(m, n2, n, t1, m, Rb->data, n, t2, m);
          double* o = vout + O->off;
          for (int64_t x = 0; x < m * n2; ++x) o[x] += f * t2[x];
        }
    }
  }
  free(t1);
  free(t2);
  return OK;
}

/* one literal term  o += f * (L . V) . R  with heap temporaries, as Eigen's
 * left-to-right evaluation */
static int term_lvr(const mat_t* La, const vblk_t* Vb, const mat_t* Rb, const vblk_t* O, double f, const double* vin,
                    double* vout) {
  if (La->cols != Vb->rows || Vb->cols != Rb->rows || La->rows != O->rows || Rb->cols != O->cols) return ERR_LAYOUT;
  int64_t m = La->rows, k = La->cols, n = Vb->cols, n2 = Rb->cols;
  double* t1 = (double*)xrealloc(NULL, sizeof(double) * (m * n + 1));
  double* t2 = (double*)xrealloc(NULL, sizeof(double) * (m * n2 + 1));
  dgemm_nn(m, n, k, La->data, m, vin + Vb->off, k, t1, m);
  dgemm_nn(m, n2, n, t1, m, Rb->data, n, t2, m);
  double* o = vout + O->off;
  for (int64_t x = 0; x < m * n2; ++x) o[x] += f * t2[x];
  free(t1);
  free(t2);
  return OK;
}

/* One-site OxV (pivot/DmrgPivotMatrix1.h:218-262) for lhs entries [l0, l1):
 * Vout[s1][q1] += f * W(a,b) * L[qL][a] . Vin[s2][q2] . R[qR][b] */
static int oxv1_range(const opivot_t* P, const double* vin, double* vout, int64_t l0, int64_t l1) {
  for (int64_t q = l0; q < l1; ++q) {
    int64_t s1 = P->qlhs[2 * q], q1 = P->qlhs[2 * q + 1];
    if (s1 < 0 || s1 >= P->ns || q1 < 0 || q1 >= P->V[s1].n) return ERR_LAYOUT;
    const vblk_t* O = &P->V[s1].b[q1];
    for (int64_t p = P->rhs_begin[q]; p < P->rhs_begin[q + 1]; ++p) {
      const int64_t* e = &P->qrhs[12 * p];
      int64_t s2 = e[0], q2 = e[1], qL = e[2], qR = e[3], k = e[4], qW = e[5];
      if (s2 < 0 || s2 >= P->ns || q2 < 0 || q2 >= P->V[s2].n || qL < 0 || qL >= P->L.n || qR < 0 || qR >= P->R.n ||
          k < 0 || k >= P->nop12 || qW < 0 || qW >= W12AT(P, s1, s2, k)->n)
        return ERR_LAYOUT;
      const wblk_t* w = &W12AT(P, s1, s2, k)->b[qW];
      const vblk_t* Vb = &P->V[s2].b[q2];
      const env_t* Le = &P->L.s[qL];
      const env_t* Re = &P->R.s[qR];
      for (int64_t i = 0; i < w->nnz; ++i) {
        if (w->row[i] >= Le->adim || w->col[i] >= Re->adim) return ERR_LAYOUT;
        const mat_t* La = &Le->blk[w->row[i]];
        const mat_t* Rb = &Re->blk[w->col[i]];
        if (!(La->rows * La->cols > 0 && Rb->rows * Rb->cols > 0 && Vb->rows * Vb->cols > 0)) continue;
        int st = term_lvr(La, Vb, Rb, O, P->fac[p] * w->val[i], vin, vout);
        if (st) return st;
      }
    }
  }
  return OK;
}

/* Zero-site HxV (pivot/DmrgPivotMatrix0.h:31-96):
 * Vout[0](L.in, L.in) += sum_a L[qL][a] . Vin[0](L.out, L.out) . R[{L.out, L.in, L.mid}][a] */
static int oxv0(const opivot_t* P, const double* vin, double* vout) {
  for (int qL = 0; qL < P->L.n; ++qL) {
    const env_t* Le = &P->L.s[qL];
    int qR = efind(&P->R, Le->out, Le->in, Le->mid);
    if (qR < 0) continue;
    int qi = vfind(&P->V[0], Le->out, Le->out);
    if (qi < 0) continue;
    const env_t* Re = &P->R.s[qR];
    int qo = vfind(&P->V[0], Re->out, Re->out);
    if (qo < 0) continue;
    if (Le->adim != Re->adim) return ERR_LAYOUT;
    const vblk_t* Vb = &P->V[0].b[qi];
    if (Vb->rows * Vb->cols == 0) continue;
    for (int a = 0; a < Le->adim; ++a) {
      const mat_t* La = &Le->blk[a];
      const mat_t* Rb = &Re->blk[a];
      if (!(La->rows * La->cols > 0 && Rb->rows * Rb->cols > 0)) continue;
      int st = term_lvr(La, Vb, Rb, &P->V[0].b[qo], 1.0, vin, vout);
      if (st) return st;
    }
  }
  return OK;
}

static int oxv_any(const opivot_t* P, const double* vin, double* vout, int64_t l0, int64_t l1) {
  if (P->nsite == 1) return oxv1_range(P, vin, vout, l0, l1);
  return oxv_range(P, vin, vout, l0, l1);
}

static int finish(const opivot_t* P, const double* vin, double* vout) {
  for (int n = 0; n < P->nproj; ++n) {
    const double* a0 = P->A0 + (size_t)n * P->vsize;
    double ov = weighted_dot(P, a0, vin);
    for (int64_t i = 0; i < P->vsize; ++i) vout[i] += P->Epen * ov * a0[i];
  }
  return OK;
}

/* HxV(H, Vin, Vout), single thread */
int oracle_hxv(const void* h, const double* vin, double* vout) {
  const opivot_t* P = (const opivot_t*)h;
  if (!P->has_bs) return fail(ERR_STATE, "no block structure");
  memset(vout, 0, sizeof(double) * P->vsize);
  int st = P->nsite == 0 ? oxv0(P, vin, vout) : oxv_any(P, vin, vout, 0, P->nlhs);
  if (st) return fail(st, "layout mismatch in OxV");
  return finish(P, vin, vout);
}

/* The same with OpenMP over output blocks (DMRG_PIVOT2_PARALLELIZE,
 * DmrgPivotMatrix2.h:179-182); distinct lhs entries write distinct blocks. */
int oracle_hxv_omp(const void* h, const double* vin, double* vout, int nthreads) {
  const opivot_t* P = (const opivot_t*)h;
  if (!P->has_bs) return fail(ERR_STATE, "no block structure");
  if (P->nsite == 0) return oracle_hxv(h, vin, vout);
  memset(vout, 0, sizeof(double) * P->vsize);
  int bad = 0;
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic) num_threads(nthreads) reduction(| : bad)
#endif
  for (int64_t q = 0; q < P->nlhs; ++q) bad |= oxv_any(P, vin, vout, q, q + 1);
  (void)nthreads;
  if (bad) return fail(ERR_LAYOUT, "layout mismatch in OxV");
  return finish(P, vin, vout);
}

/* Single-threaded literal loop over the output entries q0 <= q < q1 only
 * (a bounded sample of one HxV for the single-core CPU rate); vout must be
 * zeroed by the caller and holds only those entries' blocks afterwards. */
int oracle_hxv_range(const void* h, const double* vin, double* vout, int64_t q0, int64_t q1) {
  const opivot_t* P = (const opivot_t*)h;
  if (!P->has_bs) return fail(ERR_STATE, "no block structure");
  if (P->nsite != 2) return fail(ERR_STATE, "range form implemented for two-site pivots");
  if (q0 < 0 || q1 > P->nlhs || q0 > q1) return fail(ERR_ARG, "bad output range");
  if (oxv_any(P, vin, vout, q0, q1)) return fail(ERR_LAYOUT, "layout mismatch in OxV");
  return OK;
}

int64_t oracle_nlhs(const void* h) { return ((const opivot_t*)h)->nlhs; }

static int flops_ref_range(const void* h, int64_t q0, int64_t q1, double* flops);

/* work counts of the literal loop (for reporting the CPU baseline rate) */
int oracle_flops_ref(const void* h, double* flops) {
  return flops_ref_range(h, 0, ((const opivot_t*)h)->nlhs, flops);
}

int oracle_flops_ref_range(const void* h, int64_t q0, int64_t q1, double* flops) {
  return flops_ref_range(h, q0, q1, flops);
}

static int flops_ref_range(const void* h, int64_t q0, int64_t q1, double* flops) {
  const opivot_t* P = (const opivot_t*)h;
  double f = 0.0;
  if (P->nsite != 2) return fail(ERR_STATE, "flop count implemented for two-site pivots");
  for (int64_t q = q0; q < q1; ++q)
    for (int64_t p = P->rhs_begin[q]; p < P->rhs_begin[q + 1]; ++p) {
      const int64_t* e = &P->qrhs[12 * p];
      const wblk_t* w12 = &W12AT(P, e[4], e[5], e[6])->b[e[7]];
      const wblk_t* w34 = &W34AT(P, e[8], e[9], e[10])->b[e[11]];
      const vblk_t* Vb = &P->V[e[0]].b[e[1]];
      const env_t* Le = &P->L.s[e[2]];
      const env_t* Re = &P->R.s[e[3]];
      for (int64_t i = 0; i < w12->nnz; ++i)
        for (int64_t j = 0; j < w34->nnz; ++j) {
          const mat_t* La = &Le->blk[w12->row[i]];
          const mat_t* Rb = &Re->blk[w34->col[j]];
          if (La->rows * La->cols > 0 && Rb->rows * Rb->cols > 0 && Vb->rows * Vb->cols > 0 &&
              w12->col[i] == w34->row[j] && P->fac[p] * w12->val[i] * w34->val[j] != 0.0)
            f += 2.0 * La->rows * La->cols * Vb->cols + 2.0 * La->rows * Vb->cols * Rb->cols;
        }
    }
  *flops = f;
  return OK;
}
