// fast_matmul.hpp -- the reference's C++ spelling over the fastmm C ABI.
//
// Re-exposes, for host code written against the reference:
//   * Matrix<Scalar>: column-major matrix / view with stride and multiplier
//     (linalg/linalg.hpp:41-192): Matrix(m, n), Matrix(data, stride, m, n),
//     Matrix(m, n, multiplier), copy/move, operator()(i, j), Subblock,
//     Submatrix, m(), n(), stride(), data(), multiplier(), set_multiplier,
//     UpdateMultiplier;
//   * namespace <alg> { double FastMatmul(Matrix<Scalar>& A, Matrix<Scalar>& B,
//     Matrix<Scalar>& C, int num_steps, double x = 1e-8, Scalar alpha = 1,
//     Scalar beta = 0); } for every registered algorithm
//     (algorithms/strassen.hpp:662-722);
//   * MatMul(A, B, C, beta) (linalg/linalg.hpp:261-269).
//
// Host matrices are staged to the current HIP device for each call; the
// returned time is the device time of the multiply alone (like the
// reference's FastMatmul, which returns its own elapsed ms).  The leaf
// multiplier is alpha * B.multiplier() * C.multiplier() as in the
// reference's recursion (algorithms/strassen.hpp:669, 380-383), but no
// Matrix is modified (the reference leaves C's multiplier at alpha).
#ifndef FASTMM_FAST_MATMUL_HPP_
#define FASTMM_FAST_MATMUL_HPP_

#include <hip/hip_runtime_api.h>

#include <cstring>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>

#include "../fastmm.h"

namespace fastmm {

template <typename Scalar>
class Matrix {
 public:
  Matrix(Scalar* data, int stride, int m, int n, Scalar multiplier, bool is_view)
      : data_(data), stride_(stride), m_(m), n_(n), is_view_(is_view), multiplier_(multiplier) {
    if (!is_view_) allocate();
  }
  Matrix(Scalar* data, int stride, int m, int n, Scalar multiplier)
      : Matrix(data, stride, m, n, multiplier, true) {}
  Matrix(Scalar* data, int stride, int m, int n) : Matrix(data, stride, m, n, Scalar(1)) {}
  Matrix(int m, int n, Scalar multiplier) : Matrix(nullptr, m, m, n, multiplier, false) {}
  Matrix(int m, int n) : Matrix(m, n, Scalar(1)) {}
  explicit Matrix(int n = 0) : Matrix(n, n) {}

  Matrix(const Matrix& that) : Matrix(nullptr, that.m_, that.m_, that.n_, that.multiplier_, false) {
    for (int j = 0; j < n_; ++j)
      for (int i = 0; i < m_; ++i) data_[i + (size_t)j * stride_] = that(i, j);
  }
  Matrix(Matrix&& that) noexcept : Matrix() { swap(*this, that); }
  Matrix& operator=(Matrix that) {
    swap(*this, that);
    return *this;
  }
  ~Matrix() {
    if (!is_view_) delete[] data_;
  }
  friend void swap(Matrix& a, Matrix& b) noexcept {
    using std::swap;
    swap(a.data_, b.data_);
    swap(a.stride_, b.stride_);
    swap(a.m_, b.m_);
    swap(a.n_, b.n_);
    swap(a.is_view_, b.is_view_);
    swap(a.multiplier_, b.multiplier_);
  }

  // block (row_ind, col_ind), 1-indexed, of floor size (linalg/linalg.hpp:97-102, 179-184)
  Matrix Subblock(int num_block_rows, int num_block_cols, int row_ind, int col_ind) {
    const int rs = m_ / num_block_rows, cs = n_ / num_block_cols;
    return Submatrix(rs * (row_ind - 1), cs * (col_ind - 1), rs, cs);
  }
  Matrix Submatrix(int start_row, int start_col, int num_rows, int num_cols) {
    return Matrix(data(start_row, start_col), stride_, num_rows, num_cols, multiplier_);
  }

  Scalar* data() { return data_; }
  const Scalar* data() const { return data_; }
  Scalar* data(int i, int j) { return data_ + i + (size_t)j * stride_; }
  int stride() const { return stride_; }
  int m() const { return m_; }
  int n() const { return n_; }
  int height() const { return m_; }
  int width() const { return n_; }
  const Scalar& operator()(int i, int j) const { return data_[i + (size_t)j * stride_]; }
  Scalar& operator()(int i, int j) { return data_[i + (size_t)j * stride_]; }
  void UpdateMultiplier(Scalar v) { multiplier_ *= v; }
  void set_multiplier(Scalar v) { multiplier_ = v; }
  Scalar multiplier() const { return multiplier_; }
  void deallocate() {
    if (!is_view_) delete[] data_;
    data_ = nullptr;
  }

 private:
  void allocate() {
    data_ = (m_ > 0 && n_ > 0) ? new Scalar[(size_t)m_ * n_] : nullptr;
  }
  Scalar* data_;
  int stride_;
  int m_;
  int n_;
  bool is_view_;
  Scalar multiplier_;
};

namespace detail {

inline void check(fm_status s) {
  if (s != FM_OK)
    throw std::runtime_error(std::string("fastmm: ") + fm_status_string(s) + ": " +
                             fm_last_error());
}
inline void hcheck(hipError_t e) {
  if (e != hipSuccess) throw std::runtime_error(std::string("fastmm: ") + hipGetErrorString(e));
}

// Device staging buffers, one per operand slot, kept per thread and grown on
// demand: calls in a loop reuse them instead of allocating per call.  They
// live until the thread ends (the HIP runtime reclaims device memory at
// process exit; they are not freed from a thread-exit destructor, which may
// run after the runtime has shut down).
struct StagingSlot {
  void* p = nullptr;
  size_t bytes = 0;
};
inline void* staging(int slot, size_t bytes) {
  thread_local StagingSlot slots[3];
  StagingSlot& s = slots[slot];
  if (s.bytes < bytes) {
    if (s.p) {
      hcheck(hipDeviceSynchronize());
      (void)hipFree(s.p);
      s.p = nullptr;
      s.bytes = 0;
    }
    hcheck(hipMalloc(&s.p, bytes));
    s.bytes = bytes;
  }
  return s.p;
}

// Device copy of a host view (packed, ld = m) in staging slot `slot`.
template <typename Scalar>
struct DevMat {
  Scalar* p = nullptr;
  int m, n;
  DevMat(int slot, const Matrix<Scalar>& h, bool upload) : m(h.m()), n(h.n()) {
    p = static_cast<Scalar*>(
        staging(slot, sizeof(Scalar) * (size_t)(m > 0 ? m : 1) * (n > 0 ? n : 1)));
    if (upload && m > 0 && n > 0)
      hcheck(hipMemcpy2D(p, sizeof(Scalar) * m, h.data(), sizeof(Scalar) * h.stride(),
                         sizeof(Scalar) * m, n, hipMemcpyHostToDevice));
  }
  void download(Matrix<Scalar>& h) const {
    if (m > 0 && n > 0)
      hcheck(hipMemcpy2D(h.data(), sizeof(Scalar) * h.stride(), p, sizeof(Scalar) * m,
                         sizeof(Scalar) * m, n, hipMemcpyDeviceToHost));
  }
};

template <typename Scalar>
constexpr fm_dtype dtype_of() {
  static_assert(std::is_same<Scalar, double>::value || std::is_same<Scalar, float>::value,
                "fastmm supports double and float");
  return std::is_same<Scalar, double>::value ? FM_F64 : FM_F32;
}

template <typename Scalar>
void check_dims(const Matrix<Scalar>& A, const Matrix<Scalar>& B, const Matrix<Scalar>& C) {
  if (A.n() != B.m() || A.m() != C.m() || B.n() != C.n())
    throw std::invalid_argument("fastmm: inconsistent matrix dimensions");
}

template <typename Scalar>
double run(const char* alg, Matrix<Scalar>& A, Matrix<Scalar>& B, Matrix<Scalar>& C,
           int num_steps, double x, Scalar alpha, Scalar beta) {
  check_dims(A, B, C);
  const Scalar eff = alpha * B.multiplier() * C.multiplier();
  DevMat<Scalar> dA(0, A, true), dB(1, B, true), dC(2, C, beta != Scalar(0));
  float ms = 0.f;
  check(fm_fastmatmul(alg, dtype_of<Scalar>(), FM_MODE_PLAIN, A.m(), A.n(), B.n(), dA.p,
                      dA.m > 0 ? dA.m : 1, dB.p, dB.m > 0 ? dB.m : 1, dC.p,
                      dC.m > 0 ? dC.m : 1, num_steps, x, (double)eff, (double)beta, nullptr,
                      nullptr, 0, &ms));
  dC.download(C);
  return ms;
}

}  // namespace detail

// C <- C.multiplier() * A B + beta C (linalg/linalg.hpp:261-269).
template <typename Scalar>
void MatMul(Matrix<Scalar>& A, Matrix<Scalar>& B, Matrix<Scalar>& C, Scalar beta = Scalar(0)) {
  detail::check_dims(A, B, C);
  detail::DevMat<Scalar> dA(0, A, true), dB(1, B, true), dC(2, C, beta != Scalar(0));
  detail::check(fm_gemm(detail::dtype_of<Scalar>(), FM_MODE_PLAIN, A.m(), A.n(), B.n(), dA.p,
                        dA.m > 0 ? dA.m : 1, dB.p, dB.m > 0 ? dB.m : 1, dC.p,
                        dC.m > 0 ? dC.m : 1, (double)C.multiplier(), (double)beta, nullptr));
  detail::hcheck(hipDeviceSynchronize());
  dC.download(C);
}

}  // namespace fastmm

// One namespace per algorithm, as the reference's generated headers declare
// them (namespace name = table name, codegen/gen.py:902-906).
#define FM_ALG(NS)                                                                     \
  namespace NS {                                                                       \
  template <typename Scalar>                                                           \
  double FastMatmul(fastmm::Matrix<Scalar>& A, fastmm::Matrix<Scalar>& B,              \
                    fastmm::Matrix<Scalar>& C, int num_steps, double x = 1e-8,         \
                    Scalar alpha = Scalar(1.0), Scalar beta = Scalar(0.0)) {           \
    return fastmm::detail::run<Scalar>(#NS, A, B, C, num_steps, x, alpha, beta);       \
  }                                                                                    \
  }
#include "../../fast-matmul_amd/csrc/gen/fm_algorithms.inc"
#undef FM_ALG

// The reference declares Matrix and MatMul at global scope.
#ifndef FASTMM_NO_GLOBAL_NAMES
using fastmm::Matrix;
using fastmm::MatMul;
#endif

#endif  // FASTMM_FAST_MATMUL_HPP_
