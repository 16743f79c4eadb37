// A host program written the way the reference's tests are
// (tests/matmul_tests.cpp:19-27): Matrix<double>, <alg>::FastMatmul, MatMul,
// built against include/fastmm/fast_matmul.hpp instead of the reference's
// headers.  Integer inputs make the fast and classical products exact, so
// the program checks bit equality and exits non-zero on any difference.
#include <cstdio>
#include <cstdlib>
#include <stdexcept>

#include "fastmm/fast_matmul.hpp"

template <typename Scalar>
static void fill(Matrix<Scalar>& A, unsigned seed) {
  for (int j = 0; j < A.n(); ++j)
    for (int i = 0; i < A.m(); ++i) {
      seed = seed * 1103515245u + 12345u;
      A(i, j) = Scalar((int)((seed >> 16) % 17) - 8);
    }
}

template <typename Scalar>
static int single(const char* name, int m, int k, int n, int steps,
                  double (*fast)(Matrix<Scalar>&, Matrix<Scalar>&, Matrix<Scalar>&, int, double,
                                 Scalar, Scalar)) {
  Matrix<Scalar> A(m, k), B(k, n), C1(m, n), C2(m, n);
  fill(A, 1u);
  fill(B, 2u);
  double ms = fast(A, B, C1, steps, 1e-8, Scalar(1), Scalar(0));
  MatMul(A, B, C2);
  int bad = 0;
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < m; ++i) bad += C1(i, j) != C2(i, j);
  std::printf("%s_%d: %d, %d, %d  mismatches %d  (%.3f ms)\n", name, steps, m, k, n, bad, ms);
  return bad;
}

int main() {
  int bad = 0;
  bad += single<double>("STRASSEN", 257, 500, 55, 1, strassen::FastMatmul<double>);
  bad += single<double>("FAST424_26_206", 402, 500, 562, 2, grey424_26_206::FastMatmul<double>);
  bad += single<double>("SMIRNOV333_23_139", 366, 500, 200, 2,
                        smirnov333_23_139::FastMatmul<double>);
  bad += single<float>("STRASSEN", 300, 200, 100, 2, strassen::FastMatmul<float>);
  // views: block (2,1) of a 2x2 split times a submatrix
  Matrix<double> A(300, 200), B(200, 100);
  fill(A, 3u);
  fill(B, 4u);
  Matrix<double> Ab = A.Subblock(2, 2, 2, 1), Bs = B.Submatrix(0, 0, 100, 100);
  Matrix<double> C1(150, 100), C2(150, 100);
  grey423_20_144::FastMatmul(Ab, Bs, C1, 1);
  MatMul(Ab, Bs, C2);
  for (int j = 0; j < 100; ++j)
    for (int i = 0; i < 150; ++i) bad += C1(i, j) != C2(i, j);
  std::printf("views: mismatches so far %d\n", bad);
  // inconsistent dimensions are refused before anything is staged or run
  bool threw = false;
  Matrix<double> Cw(150, 99);
  try {
    MatMul(Ab, Bs, Cw);
  } catch (const std::invalid_argument&) {
    threw = true;
  }
  if (!threw) {
    std::printf("MatMul accepted a C of the wrong width\n");
    ++bad;
  }
  threw = false;
  try {
    strassen::FastMatmul(Ab, Bs, Cw, 1);
  } catch (const std::invalid_argument&) {
    threw = true;
  }
  if (!threw) {
    std::printf("FastMatmul accepted a C of the wrong width\n");
    ++bad;
  }
  return bad ? 1 : 0;
}
