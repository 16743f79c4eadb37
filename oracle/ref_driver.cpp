// ref_driver.cpp -- C entry points over the reference's own header-only
// fast-matmul path, compiled from the sources where they lie (see
// oracle/Makefile).  TEST INFRASTRUCTURE ONLY: it builds oracle/_ref/libfmref.so,
// which tests and bench.py's cpu_baseline leg load to pin the oracle and to
// time the reference on the host.  Nothing here is part of the product.
//
// Each entry wraps caller memory in reference Matrix views
// (linalg/linalg.hpp:52-54) and calls <alg>::FastMatmul
// (e.g. algorithms/strassen.hpp:662-722), returning its elapsed ms.
#include <cstring>
#include <string>

#include "all_algorithms.hpp"
#include "fast423_20_130.hpp"
#include "fast423_20_134.hpp"
#include "fast423_20_138.hpp"
#include "fast423_20_156.hpp"
#include "smirnov333_20_182_approx.hpp"
#include "tichavsky_kovac336_40_960.hpp"
#include "tichavsky_kovac363_40_960.hpp"
#include "quadmatmul.hpp"

#define FM_ALGS(X)                                                             \
  X(strassen) X(classical222_8_24) X(classical333_27_81) X(classical423_24_72) \
  X(bini322_10_52_approx) X(schonhage333_21_117_approx)                        \
  X(smirnov333_20_182_approx) X(grey234_20_144) X(grey243_20_144)              \
  X(grey252_18_99) X(grey322_11_50) X(grey323_15_103) X(grey323_15_89)         \
  X(grey324_20_144) X(grey332_15_103) X(grey333_23_125) X(grey333_23_152)      \
  X(grey333_23_221) X(grey342_20_144) X(grey343_29_234) X(grey422_14_84)       \
  X(fast423_130) X(fast423_134) X(fast423_138) X(fast423_156)                  \
  X(grey423_20_144) X(grey424_26_206) X(grey424_26_257) X(grey432_20_144)      \
  X(grey433_29_234) X(grey442_26_257) X(grey522_18_99) X(hk323_15_84)          \
  X(hk323_15_94) X(hk332_15_94) X(smirnov333_23_128) X(smirnov333_23_139)      \
  X(smirnov336_40_960) X(smirnov363_40_960) X(smirnov633_40_960)               \
  X(tichavsky_kovac336_40_960) X(tichavsky_kovac363_40_960)

template <typename Scalar>
static double run(const char* name, int m, int k, int n, Scalar* A, int lda, Scalar* B,
                  int ldb, Scalar* C, int ldc, int steps, double x, Scalar alpha,
                  Scalar beta) {
  Matrix<Scalar> Am(A, lda, m, k);
  Matrix<Scalar> Bm(B, ldb, k, n);
  Matrix<Scalar> Cm(C, ldc, m, n);
#define FM_CASE(ns) \
  if (std::strcmp(name, #ns) == 0) return ns::FastMatmul(Am, Bm, Cm, steps, x, alpha, beta);
  FM_ALGS(FM_CASE)
#undef FM_CASE
  return -1.0;
}

extern "C" {
double fmref_fastmatmul_d(const char* name, int m, int k, int n, double* A, int lda,
                          double* B, int ldb, double* C, int ldc, int steps, double x,
                          double alpha, double beta) {
  return run<double>(name, m, k, n, A, lda, B, ldb, C, ldc, steps, x, alpha, beta);
}

double fmref_fastmatmul_s(const char* name, int m, int k, int n, float* A, int lda,
                          float* B, int ldb, float* C, int ldc, int steps, double x,
                          float alpha, float beta) {
  return run<float>(name, m, k, n, A, lda, B, ldb, C, ldc, steps, x, alpha, beta);
}

// Classical product through the reference's MatMul (linalg/linalg.hpp:261-269).
double fmref_matmul_d(int m, int k, int n, double* A, int lda, double* B, int ldb,
                      double* C, int ldc, double alpha, double beta) {
  Matrix<double> Am(A, lda, m, k);
  Matrix<double> Bm(B, ldb, k, n);
  Matrix<double> Cm(C, ldc, m, n, alpha);
  return Time([&] { MatMul(Am, Bm, Cm, beta); });
}

// QuadMatmul (linalg/quadmatmul.hpp:8-26), result rounded to double.
void fmref_quad_matmul(int m, int k, int n, double* A, int lda, double* B, int ldb,
                       double* C, int ldc) {
  Matrix<double> Am(A, lda, m, k);
  Matrix<double> Bm(B, ldb, k, n);
  Matrix<__float128> Q(m, n);
  QuadMatmul(Am, Bm, Q);
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < m; ++i) C[i + (size_t)j * ldc] = (double)Q(i, j);
}

double fmref_max_rel_diff(int m, int n, double* A, int lda, double* B, int ldb) {
  Matrix<double> Am(A, lda, m, n);
  Matrix<double> Bm(B, ldb, m, n);
  return MaxRelativeDiff(Am, Bm);
}
}
