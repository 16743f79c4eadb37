/*
 * fm_oracle.c -- CPU restatement of the reference's fast-matmul hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.  The
 * product path (fast-matmul_amd/) never links, loads or calls it.
 *
 * What it restates (citations are into the reference tree):
 *   fmo_fastmatmul_{d,s}  <alg>::FastMatmul -> FastMatmulRecursive
 *                         (algorithms/strassen.hpp:378-722; every generated
 *                         header has the same shape, e.g.
 *                         algorithms/fast424_26_206.hpp:2947-3007), as emitted
 *                         by codegen/gen.py:574-657 and :910-1006:
 *     - multiplier folding C.mult *= A.mult * B.mult      (strassen.hpp:380-383)
 *     - base case MatMul with alpha = C.mult, beta = 0     (strassen.hpp:385-388,
 *                                                          linalg/linalg.hpp:261-269)
 *     - floor-sized block views                           (linalg/linalg.hpp:97-102,179-184)
 *     - CSE temporaries A_X/B_X before the products       (fast424_26_206.hpp:2107-2131)
 *     - per product: write-once S/T adds when the column has >1 nonzero,
 *       else the block view itself with the single coefficient folded into
 *       M_r's multiplier                                  (codegen/gen.py:600-621)
 *     - recursion with beta = 0                           (strassen.hpp:459)
 *     - M_X temporaries, then C_ij = sum W M (+ beta C)   (fast424_26_206.hpp:2893-2934)
 *     - dynamic peeling                                   (linalg/linalg.hpp:201-244)
 *   Add expressions keep the reference's term order and operator form
 *   (codegen/gen.py:216-233, 444-523): first term x | -x | c*x, then
 *   "+ x", "- x", "+ c*x" left to right, then "+ beta*C".
 *   fmo_quad_*             QuadMatmul (linalg/quadmatmul.hpp:8-26)
 *   fmo_max_rel_diff etc.  MaxRelativeDiff / MaxAbsDiff / FrobeniusDiff
 *                          (linalg/linalg.hpp:280-345)
 *
 * Leaf products go through the Fortran BLAS dgemm_/sgemm_ symbol, like
 * linalg/blas_wrap.hpp:109-121.  fmo_set_blas() installs the BLAS (the test
 * harness opens the image's MKL runtime, the same library the reference is
 * linked to); without one a plain triple loop is used.
 *
 * Build with -ffp-contract=off: the reference's adds are separate multiply
 * and add roundings on x86-64 without FMA.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <quadmath.h>

/* Built with -fopenmp -DFMO_OMP_ADDS (liboracle_par.so, for the CPU
 * baseline only), the write-once adds split their columns over OpenMP
 * threads: the reference's _PARALLEL_ DFS mode runs its add loops as OpenMP
 * parallel-for loops (algorithms/strassen.hpp:17-19).  Each element's sum
 * is unchanged, so results are identical either way. */
#ifdef FMO_OMP_ADDS
#define FMO_PAR_FOR _Pragma("omp parallel for schedule(static)")
#else
#define FMO_PAR_FOR
#endif

typedef void (*dgemm_fn)(const char*, const char*, const int*, const int*, const int*,
                         const double*, const double*, const int*, const double*, const int*,
                         const double*, double*, const int*);
typedef void (*sgemm_fn)(const char*, const char*, const int*, const int*, const int*,
                         const float*, const float*, const int*, const float*, const int*,
                         const float*, float*, const int*);

static dgemm_fn g_dgemm = NULL;
static sgemm_fn g_sgemm = NULL;

void fmo_set_blas(void* dgemm, void* sgemm) {
  g_dgemm = (dgemm_fn)dgemm;
  g_sgemm = (sgemm_fn)sgemm;
}

int fmo_has_blas(void) { return g_dgemm != NULL; }

/* Algorithm description.  Row-major coefficient arrays:
 *   U  : (m*k + nsx) x R     V  : (k*n + ntx) x R
 *   W  : (m*n) x (R + nmx)   SX : nsx x (m*k)   TX : ntx x (k*n)   MX : nmx x R */
typedef struct {
  int m, k, n, R, nsx, ntx, nmx;
  const double *U, *V, *W, *SX, *TX, *MX;
} fmo_table;

#define DEFINE_ORACLE(T, SUF, GEMMPTR)                                                       \
  typedef struct {                                                                           \
    T* data;                                                                                 \
    int stride, m, n;                                                                        \
    T mult;                                                                                  \
  } mat_##SUF;                                                                               \
                                                                                             \
  static mat_##SUF view_##SUF(T* d, int stride, int m, int n, T mult) {                       \
    mat_##SUF r;                                                                             \
    r.data = d; r.stride = stride; r.m = m; r.n = n; r.mult = mult;                          \
    return r;                                                                                \
  }                                                                                          \
                                                                                             \
  static mat_##SUF alloc_##SUF(int m, int n, T mult) {                                       \
    size_t cnt = (size_t)m * (size_t)n;                                                      \
    T* d = (T*)malloc((cnt ? cnt : 1) * sizeof(T));                                         \
    return view_##SUF(d, m, m, n, mult);                                                     \
  }                                                                                          \
                                                                                             \
  /* C <- alpha A B + beta C (column-major, 'N','N'); alpha = C.mult (linalg.hpp:266). */     \
  static void matmul_##SUF(mat_##SUF A, mat_##SUF B, mat_##SUF C, T beta) {                  \
    T alpha = C.mult;                                                                        \
    if (C.m <= 0 || C.n <= 0) return;                                                        \
    if (GEMMPTR) {                                                                           \
      char tr = 'N';                                                                         \
      int m = A.m, n = B.n, k = A.n, lda = A.stride, ldb = B.stride, ldc = C.stride;         \
      if (k <= 0) {                                                                          \
        /* BLAS with k=0 scales C by beta */                                                 \
      }                                                                                      \
      GEMMPTR(&tr, &tr, &m, &n, &k, &alpha, A.data, &lda, B.data, &ldb, &beta, C.data, &ldc);\
      return;                                                                                \
    }                                                                                        \
    for (int j = 0; j < C.n; ++j) {                                                          \
      for (int i = 0; i < C.m; ++i) {                                                        \
        T acc = 0;                                                                           \
        for (int l = 0; l < A.n; ++l)                                                        \
          acc += A.data[i + (size_t)l * A.stride] * B.data[l + (size_t)j * B.stride];        \
        T* c = &C.data[i + (size_t)j * C.stride];                                            \
        *c = (beta == (T)0) ? alpha * acc : alpha * acc + beta * *c;                         \
      }                                                                                      \
    }                                                                                        \
  }                                                                                          \
                                                                                             \
  /* dst = sum_t coef_t * src_t in the reference's write-once form. */                      \
  static void add_##SUF(int nterm, mat_##SUF* src, const double* coef, mat_##SUF dst,        \
                        int with_beta, T beta) {                                             \
    FMO_PAR_FOR                                                                              \
    for (int j = 0; j < dst.n; ++j) {                                                        \
      for (int i = 0; i < dst.m; ++i) {                                                      \
        T acc = 0;                                                                           \
        for (int t = 0; t < nterm; ++t) {                                                    \
          T v = src[t].data[i + (size_t)j * src[t].stride];                                  \
          double c = coef[t];                                                                \
          if (t == 0) {                                                                      \
            acc = (c == 1.0) ? v : (c == -1.0) ? -v : (T)c * v;                              \
          } else {                                                                           \
            acc = (c == 1.0) ? acc + v : (c == -1.0) ? acc - v : acc + (T)c * v;             \
          }                                                                                  \
        }                                                                                    \
        T* d = &dst.data[i + (size_t)j * dst.stride];                                        \
        if (with_beta) acc = acc + beta * *d;                                                \
        *d = acc;                                                                            \
      }                                                                                      \
    }                                                                                        \
  }                                                                                          \
                                                                                             \
  static mat_##SUF subblock_##SUF(mat_##SUF X, int nbr, int nbc, int bi, int bj) {           \
    int rs = X.m / nbr, cs = X.n / nbc;                                                      \
    return view_##SUF(X.data + (size_t)(rs * bi) + (size_t)(cs * bj) * X.stride, X.stride,   \
                      rs, cs, X.mult);                                                       \
  }                                                                                          \
                                                                                             \
  static void peel_##SUF(mat_##SUF A, mat_##SUF B, mat_##SUF C, int d1, int d2, int d3,      \
                         T beta) {                                                           \
    int extra_rows_A = A.m - (A.m / d1) * d1;                                                \
    int extra_cols_A = A.n - (A.n / d2) * d2;                                                \
    int extra_rows_B = B.m - (B.m / d2) * d2;                                                \
    int extra_cols_B = B.n - (B.n / d3) * d3;                                                \
    if (extra_cols_A > 0) {                                                                  \
      int row_dim = A.m - extra_rows_A, col_dim = B.n - extra_cols_B;                        \
      int kk = A.n - extra_cols_A;                                                           \
      mat_##SUF Ae = view_##SUF(A.data + (size_t)kk * A.stride, A.stride, row_dim,           \
                                extra_cols_A, A.mult);                                       \
      mat_##SUF Be = view_##SUF(B.data + kk, B.stride, extra_rows_B, col_dim, B.mult);       \
      mat_##SUF Ce = view_##SUF(C.data, C.stride, row_dim, col_dim, C.mult);                 \
      matmul_##SUF(Ae, Be, Ce, (T)1.0);                                                      \
    }                                                                                        \
    if (extra_cols_B > 0) {                                                                  \
      int s = B.n - extra_cols_B;                                                            \
      mat_##SUF Be = view_##SUF(B.data + (size_t)s * B.stride, B.stride, B.m, extra_cols_B,  \
                                B.mult);                                                     \
      mat_##SUF Ce = view_##SUF(C.data + (size_t)s * C.stride, C.stride, C.m, extra_cols_B,  \
                                C.mult);                                                     \
      matmul_##SUF(A, Be, Ce, beta);                                                         \
    }                                                                                        \
    if (extra_rows_A > 0) {                                                                  \
      int s = A.m - extra_rows_A, nc = B.n - extra_cols_B;                                   \
      mat_##SUF Ae = view_##SUF(A.data + s, A.stride, extra_rows_A, A.n, A.mult);            \
      mat_##SUF Be = view_##SUF(B.data, B.stride, B.m, nc, B.mult);                          \
      mat_##SUF Ce = view_##SUF(C.data + s, C.stride, extra_rows_A, nc, C.mult);             \
      matmul_##SUF(Ae, Be, Ce, beta);                                                        \
    }                                                                                        \
  }                                                                                          \
                                                                                             \
  static void rec_##SUF(const fmo_table* t, mat_##SUF* A, mat_##SUF* B, mat_##SUF* C,        \
                        int steps_left, T beta) {                                            \
    C->mult *= A->mult;                                                                      \
    C->mult *= B->mult;                                                                      \
    A->mult = (T)1.0;                                                                        \
    B->mult = (T)1.0;                                                                        \
    if (steps_left == 0) {                                                                   \
      matmul_##SUF(*A, *B, *C, (T)0.0);                                                      \
      return;                                                                                \
    }                                                                                        \
    const int m = t->m, k = t->k, n = t->n, R = t->R;                                        \
    const int mk = m * k, kn = k * n, mn = m * n;                                            \
    mat_##SUF* Ab = (mat_##SUF*)malloc(sizeof(mat_##SUF) * (mk + t->nsx));                   \
    mat_##SUF* Bb = (mat_##SUF*)malloc(sizeof(mat_##SUF) * (kn + t->ntx));                   \
    mat_##SUF* Cb = (mat_##SUF*)malloc(sizeof(mat_##SUF) * mn);                              \
    mat_##SUF* M = (mat_##SUF*)malloc(sizeof(mat_##SUF) * (R + t->nmx));                     \
    mat_##SUF* src = (mat_##SUF*)malloc(sizeof(mat_##SUF) * (R + t->nmx + mk + kn + 8));     \
    double* cf = (double*)malloc(sizeof(double) * (R + t->nmx + mk + kn + 8));               \
    for (int i = 0; i < m; ++i)                                                              \
      for (int j = 0; j < k; ++j) Ab[i * k + j] = subblock_##SUF(*A, m, k, i, j);            \
    for (int i = 0; i < k; ++i)                                                              \
      for (int j = 0; j < n; ++j) Bb[i * n + j] = subblock_##SUF(*B, k, n, i, j);            \
    for (int i = 0; i < m; ++i)                                                              \
      for (int j = 0; j < n; ++j) Cb[i * n + j] = subblock_##SUF(*C, m, n, i, j);            \
    for (int r = 0; r < R; ++r) M[r] = alloc_##SUF(Cb[0].m, Cb[0].n, C->mult);              \
    /* input CSE temporaries (codegen/gen.py:710-739) */                                     \
    for (int x = 0; x < t->nsx; ++x) {                                                       \
      int nt = 0;                                                                            \
      for (int p = 0; p < mk; ++p)                                                           \
        if (t->SX[x * mk + p] != 0.0) { src[nt] = Ab[p]; cf[nt++] = t->SX[x * mk + p]; }     \
      Ab[mk + x] = alloc_##SUF(Ab[0].m, Ab[0].n, (T)1.0);                                    \
      add_##SUF(nt, src, cf, Ab[mk + x], 0, (T)0);                                           \
    }                                                                                        \
    for (int x = 0; x < t->ntx; ++x) {                                                       \
      int nt = 0;                                                                            \
      for (int p = 0; p < kn; ++p)                                                           \
        if (t->TX[x * kn + p] != 0.0) { src[nt] = Bb[p]; cf[nt++] = t->TX[x * kn + p]; }     \
      Bb[kn + x] = alloc_##SUF(Bb[0].m, Bb[0].n, (T)1.0);                                    \
      add_##SUF(nt, src, cf, Bb[kn + x], 0, (T)0);                                           \
    }                                                                                        \
    for (int r = 0; r < R; ++r) {                                                            \
      mat_##SUF Sop, Top;                                                                    \
      int sown = 0, town = 0;                                                                \
      int nt = 0;                                                                            \
      double single = 1.0;                                                                   \
      for (int p = 0; p < mk + t->nsx; ++p)                                                  \
        if (t->U[p * R + r] != 0.0) { src[nt] = Ab[p]; cf[nt++] = t->U[p * R + r]; }         \
      if (nt > 1) {                                                                          \
        Sop = alloc_##SUF(Ab[0].m, Ab[0].n, (T)1.0);                                         \
        add_##SUF(nt, src, cf, Sop, 0, (T)0);                                                \
        sown = 1;                                                                            \
      } else {                                                                               \
        Sop = src[0];                                                                        \
        single = cf[0];                                                                      \
        M[r].mult *= (T)single;                                                              \
      }                                                                                      \
      nt = 0;                                                                                \
      for (int p = 0; p < kn + t->ntx; ++p)                                                  \
        if (t->V[p * R + r] != 0.0) { src[nt] = Bb[p]; cf[nt++] = t->V[p * R + r]; }         \
      if (nt > 1) {                                                                          \
        Top = alloc_##SUF(Bb[0].m, Bb[0].n, (T)1.0);                                         \
        add_##SUF(nt, src, cf, Top, 0, (T)0);                                                \
        town = 1;                                                                            \
      } else {                                                                               \
        Top = src[0];                                                                        \
        M[r].mult *= (T)cf[0];                                                               \
      }                                                                                      \
      rec_##SUF(t, &Sop, &Top, &M[r], steps_left - 1, (T)0.0);                               \
      if (sown) free(Sop.data);                                                              \
      if (town) free(Top.data);                                                              \
    }                                                                                        \
    /* output CSE temporaries (codegen/gen.py:742-781) */                                    \
    for (int x = 0; x < t->nmx; ++x) {                                                       \
      int nt = 0;                                                                            \
      for (int q = 0; q < R; ++q)                                                            \
        if (t->MX[x * R + q] != 0.0) { src[nt] = M[q]; cf[nt++] = t->MX[x * R + q]; }        \
      M[R + x] = alloc_##SUF(Cb[0].m, Cb[0].n, (T)1.0);                                      \
      if (nt > 0) add_##SUF(nt, src, cf, M[R + x], 0, (T)0);                                 \
    }                                                                                        \
    const int WC = R + t->nmx;                                                               \
    for (int c = 0; c < mn; ++c) {                                                           \
      int nt = 0;                                                                            \
      for (int q = 0; q < WC; ++q)                                                           \
        if (t->W[c * WC + q] != 0.0) { src[nt] = M[q]; cf[nt++] = t->W[c * WC + q]; }        \
      add_##SUF(nt, src, cf, Cb[c], beta != (T)0.0, beta);                                   \
    }                                                                                        \
    for (int q = 0; q < R + t->nmx; ++q) free(M[q].data);                                    \
    for (int x = 0; x < t->nsx; ++x) free(Ab[mk + x].data);                                  \
    for (int x = 0; x < t->ntx; ++x) free(Bb[kn + x].data);                                  \
    peel_##SUF(*A, *B, *C, m, k, n, beta);                                                   \
    free(Ab); free(Bb); free(Cb); free(M); free(src); free(cf);                              \
  }                                                                                          \
                                                                                             \
  /* C := alpha A B + beta C with num_steps levels; the stateless form of       */          \
  /* <alg>::FastMatmul (strassen.hpp:662-722).                                  */          \
  void fmo_fastmatmul_##SUF(const fmo_table* t, int M_, int K_, int N_, const T* A, int lda,  \
                            const T* B, int ldb, T* C, int ldc, int steps, T alpha, T beta) {\
    mat_##SUF Am = view_##SUF((T*)A, lda, M_, K_, alpha);                                    \
    mat_##SUF Bm = view_##SUF((T*)B, ldb, K_, N_, (T)1.0);                                   \
    mat_##SUF Cm = view_##SUF(C, ldc, M_, N_, (T)1.0);                                       \
    rec_##SUF(t, &Am, &Bm, &Cm, steps, beta);                                                \
  }                                                                                          \
                                                                                             \
  /* Plain C <- alpha A B + beta C through the same leaf (linalg.hpp:261-269). */           \
  void fmo_gemm_##SUF(int M_, int K_, int N_, const T* A, int lda, const T* B, int ldb, T* C, \
                      int ldc, T alpha, T beta) {                                            \
    mat_##SUF Am = view_##SUF((T*)A, lda, M_, K_, (T)1.0);                                   \
    mat_##SUF Bm = view_##SUF((T*)B, ldb, K_, N_, (T)1.0);                                   \
    mat_##SUF Cm = view_##SUF(C, ldc, M_, N_, alpha);                                        \
    matmul_##SUF(Am, Bm, Cm, beta);                                                          \
  }

DEFINE_ORACLE(double, d, g_dgemm)
DEFINE_ORACLE(float, s, g_sgemm)

/* ---- quad-precision oracle (linalg/quadmatmul.hpp:8-26) -------------------- */

/* C = A B accumulated in __float128, rounded to double at the end. */
void fmo_quad_matmul(int M_, int K_, int N_, const double* A, int lda, const double* B,
                     int ldb, double* C, int ldc) {
  for (int j = 0; j < N_; ++j) {
    for (int i = 0; i < M_; ++i) {
      __float128 acc = 0;
      for (int l = 0; l < K_; ++l)
        acc += (__float128)A[i + (size_t)l * lda] * (__float128)B[l + (size_t)j * ldb];
      C[i + (size_t)j * ldc] = (double)acc;
    }
  }
}

/* Same, for a list of (row, col) entries only, so large products can be
 * checked on a sample.  Also returns sum_l |a_il||b_lj| per entry in absab
 * (may be NULL) for componentwise error bounds. */
void fmo_quad_entries(int K_, const double* A, int lda, const double* B, int ldb,
                      const int* rows, const int* cols, int count, double* out,
                      double* absab) {
  for (int e = 0; e < count; ++e) {
    __float128 acc = 0, aa = 0;
    int i = rows[e], j = cols[e];
    for (int l = 0; l < K_; ++l) {
      __float128 p = (__float128)A[i + (size_t)l * lda] * (__float128)B[l + (size_t)j * ldb];
      acc += p;
      aa += p < 0 ? -p : p;
    }
    out[e] = (double)acc;
    if (absab) absab[e] = (double)aa;
  }
}

/* ---- error metrics (linalg/linalg.hpp:280-345) ----------------------------- */

/* max_ij |a - b| / |a|, dividing by the FIRST argument (linalg.hpp:288). */
double fmo_max_rel_diff(int m, int n, const double* A, int lda, const double* B, int ldb) {
  double worst = 0;
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < m; ++i) {
      double a = A[i + (size_t)j * lda], b = B[i + (size_t)j * ldb];
      double d = fabs(a - b) / fabs(a);
      if (d > worst) worst = d;
    }
  return worst;
}

double fmo_max_abs_diff(int m, int n, const double* A, int lda, const double* B, int ldb) {
  double worst = 0;
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < m; ++i) {
      double d = fabs(A[i + (size_t)j * lda] - B[i + (size_t)j * ldb]);
      if (d > worst) worst = d;
    }
  return worst;
}

double fmo_frobenius_diff(int m, int n, const double* A, int lda, const double* B, int ldb) {
  double acc = 0;
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < m; ++i) {
      double d = A[i + (size_t)j * lda] - B[i + (size_t)j * ldb];
      acc += d * d;
    }
  return sqrt(acc);
}
