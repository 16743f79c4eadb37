// scaling.hip -- diagonal scaling around the fast multiply
// (scaling/scaling.hpp:105-196 of the reference).
//
// Outer scaling divides the rows of A by their largest magnitude x_i and the
// columns of B by theirs y_j (the factors accumulate in r and s); inner
// scaling multiplies column k of A by 1/z_k and row k of B by z_k with
// z_k = sqrt(max|A[:,k]| / max|B[k,:]|).  After the multiply,
// PostProcessScaling multiplies row i of the product by r_i and column j by
// s_j.  Each pass is the reference's: the factor vector is formed first
// (1/x rounded, sqrt rounded), then every element is multiplied by it, one
// rounding per pass, in the reference's pass order.  A row or column that is
// all zeros keeps factor 1 (the reference would divide by zero).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "scaling.hpp"

namespace {

template <typename T>
struct Bits;
template <>
struct Bits<double> {
  typedef unsigned long long U;
  __device__ static U of(double v) { return (U)__double_as_longlong(v); }
  __device__ static double val(U u) { return __longlong_as_double((long long)u); }
};
template <>
struct Bits<float> {
  typedef unsigned int U;
  __device__ static U of(float v) { return __float_as_uint(v); }
  __device__ static float val(U u) { return __uint_as_float(u); }
};

// separately rounded products and sums (no contraction into FMA)
__device__ __forceinline__ double mul_rn(double a, double b) { return __dmul_rn(a, b); }
__device__ __forceinline__ float mul_rn(float a, float b) { return __fmul_rn(a, b); }
__device__ __forceinline__ double add_rn(double a, double b) { return __dadd_rn(a, b); }
__device__ __forceinline__ float add_rn(float a, float b) { return __fadd_rn(a, b); }

// zero a vector of n unsigned words (a kernel, so the sequence stays
// graph-capturable)
template <typename U>
__global__ void fm_zero_kernel(U* p, int n) {
  for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) p[i] = 0;
}

// out[i] = max_j |X[i, j]| for a column-major rows x cols matrix; grid.y
// splits the columns, partial maxima meet in an integer atomicMax on the
// bit pattern (non-negative floats order like their bits).
template <typename T>
__global__ void fm_rowmax_kernel(const T* X, int64_t ld, int rows, int cols, int chunk,
                                 typename Bits<T>::U* out) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= rows) return;
  const int j0 = blockIdx.y * chunk, j1 = min(cols, j0 + chunk);
  T mx = 0;
  for (int j = j0; j < j1; ++j) mx = fmax(mx, fabs(X[i + (int64_t)j * ld]));
  atomicMax(out + i, Bits<T>::of(mx));
}

// out[j] = max_i |X[i, j]|: one work group per column
template <typename T>
__global__ void fm_colmax_kernel(const T* X, int64_t ld, int rows, typename Bits<T>::U* out) {
  const int j = blockIdx.x;
  const T* c = X + (int64_t)j * ld;
  T mx = 0;
  for (int i = threadIdx.x; i < rows; i += 256) mx = fmax(mx, fabs(c[i]));
  __shared__ T red[256];
  red[threadIdx.x] = mx;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) out[j] = Bits<T>::of(red[0]);
}

// Factor arithmetic follows the reference's types: the maxima are doubles
// (MaxAbsRow / MaxAbsCol return double), 1/x and sqrt(a/b) are evaluated in
// double and stored as Scalar (InverseVector, InnerScaling), and the running
// products r, s are Scalar (UpdateVals).

// outer-scaling factors: f[i] = 1 / x[i] (1 for x = 0), r[i] *= x[i] (x = 0: unchanged)
template <typename T>
__global__ void fm_outer_factors(const typename Bits<T>::U* xb, int n, T* f, T* r) {
  for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
    const T x = Bits<T>::val(xb[i]);
    if (x > T(0)) {
      f[i] = (T)(1.0 / (double)x);
      r[i] = mul_rn(r[i], x);
    } else {
      f[i] = T(1);
    }
  }
}

// inner-scaling factors: z = sqrt(ca / rb) (1 where either is 0), zi = 1 / z
template <typename T>
__global__ void fm_inner_factors(const typename Bits<T>::U* ca, const typename Bits<T>::U* rb,
                                 int n, T* z, T* zi) {
  for (int k = blockIdx.x * 256 + threadIdx.x; k < n; k += gridDim.x * 256) {
    const double a = (double)Bits<T>::val(ca[k]), b = (double)Bits<T>::val(rb[k]);
    T v = T(1);
    if (a > 0.0 && b > 0.0) v = (T)sqrt(a / b);
    z[k] = v;
    zi[k] = (T)(1.0 / (double)v);
  }
}

template <typename T>
__global__ void fm_fill_one(T* p, int n) {
  for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) p[i] = T(1);
}

// Y[i, j] = X[i, j] * rf[i] (if rf) * cf[j] (if cf), one rounding per factor;
// Y may alias X.
template <typename T>
__global__ void fm_scale_kernel(const T* X, int64_t ldx, T* Y, int64_t ldy, int rows, int cols,
                                const T* rf, const T* cf) {
  const int64_t total = (int64_t)rows * cols;
  for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * 256) {
    const int j = (int)(e / rows), i = (int)(e - (int64_t)j * rows);
    T v = X[i + (int64_t)j * ldx];
    if (rf) v = mul_rn(v, rf[i]);
    if (cf) v = mul_rn(v, cf[j]);
    Y[i + (int64_t)j * ldy] = v;
  }
}

// C[i, j] = alpha * ((P[i, j] * r[i]) * s[j]) + beta * C[i, j]
template <typename T>
__global__ void fm_post_kernel(const T* P, int64_t ldp, T* C, int64_t ldc, int rows, int cols,
                               const T* r, const T* s, T alpha, T beta) {
  const int64_t total = (int64_t)rows * cols;
  for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * 256) {
    const int j = (int)(e / rows), i = (int)(e - (int64_t)j * rows);
    T v = mul_rn(P[i + (int64_t)j * ldp], r[i]);
    v = mul_rn(v, s[j]);
    if (alpha != T(1)) v = mul_rn(alpha, v);
    T* c = C + i + (int64_t)j * ldc;
    if (beta != T(0)) v = add_rn(v, mul_rn(beta, *c));
    *c = v;
  }
}

inline unsigned blocks_for(int64_t n) {
  int64_t b = (n + 255) / 256;
  if (b > 65536) b = 65536;
  return (unsigned)(b < 1 ? 1 : b);
}

}  // namespace

template <typename T>
static hipError_t run_scaling(const FmScaleArgs& a, hipStream_t s) {
  typedef typename Bits<T>::U U;
  T* A = reinterpret_cast<T*>(a.A);
  T* B = reinterpret_cast<T*>(a.B);
  T* r = reinterpret_cast<T*>(a.r);
  T* sv = reinterpret_cast<T*>(a.s);
  T* f = reinterpret_cast<T*>(a.f);    // length max(m, k, n)
  T* f2 = reinterpret_cast<T*>(a.f2);  // length max(m, k, n)
  U* u1 = reinterpret_cast<U*>(a.u1);
  U* u2 = reinterpret_cast<U*>(a.u2);
  const int m = a.m, k = a.k, n = a.n;
  const int64_t lda = a.lda, ldb = a.ldb;
  auto rowmax = [&](const T* X, int64_t ld, int rows, int cols, U* out) {
    hipLaunchKernelGGL(fm_zero_kernel<U>, dim3(blocks_for(rows)), dim3(256), 0, s, out, rows);
    const int chunk = 256;
    hipLaunchKernelGGL(fm_rowmax_kernel<T>, dim3((rows + 255) / 256, (cols + chunk - 1) / chunk),
                       dim3(256), 0, s, X, ld, rows, cols, chunk, out);
  };
  auto colmax = [&](const T* X, int64_t ld, int rows, int cols, U* out) {
    hipLaunchKernelGGL(fm_colmax_kernel<T>, dim3(cols), dim3(256), 0, s, X, ld, rows, out);
  };
  auto scale = [&](T* X, int64_t ld, int rows, int cols, const T* rf, const T* cf) {
    hipLaunchKernelGGL(fm_scale_kernel<T>, dim3(blocks_for((int64_t)rows * cols)), dim3(256), 0,
                       s, X, ld, X, ld, rows, cols, rf, cf);
  };
  auto outer = [&]() {
    rowmax(A, lda, m, k, u1);
    hipLaunchKernelGGL(fm_outer_factors<T>, dim3(blocks_for(m)), dim3(256), 0, s, u1, m, f, r);
    scale(A, lda, m, k, f, nullptr);
    colmax(B, ldb, k, n, u2);
    hipLaunchKernelGGL(fm_outer_factors<T>, dim3(blocks_for(n)), dim3(256), 0, s, u2, n, f2, sv);
    scale(B, ldb, k, n, nullptr, f2);
  };
  auto inner = [&]() {
    colmax(A, lda, m, k, u1);
    rowmax(B, ldb, k, n, u2);
    hipLaunchKernelGGL(fm_inner_factors<T>, dim3(blocks_for(k)), dim3(256), 0, s, u1, u2, k, f,
                       f2);
    scale(A, lda, m, k, nullptr, f2);  // ColScale(A, 1 / z)
    scale(B, ldb, k, n, f, nullptr);   // RowScale(B, z)
  };
  hipLaunchKernelGGL(fm_fill_one<T>, dim3(blocks_for(m)), dim3(256), 0, s, r, m);
  hipLaunchKernelGGL(fm_fill_one<T>, dim3(blocks_for(n)), dim3(256), 0, s, sv, n);
  for (int it = 0; it < a.steps; ++it) {
    switch (a.type) {
      case 0: outer(); inner(); break;  // OUTSIDE_INSIDE
      case 1: inner(); outer(); break;  // INSIDE_OUTSIDE
      case 2: inner(); break;           // INSIDE
      default: outer(); break;          // OUTSIDE
    }
  }
  return hipGetLastError();
}

hipError_t fm_scaling_run(int dtype_f64, const FmScaleArgs& a, hipStream_t s) {
  return dtype_f64 ? run_scaling<double>(a, s) : run_scaling<float>(a, s);
}

// copy X (rows x cols, ldx) to Y (ldy)
template <typename T>
static void copy_mat(const void* X, int64_t ldx, void* Y, int64_t ldy, int rows, int cols,
                     hipStream_t s) {
  hipLaunchKernelGGL(fm_scale_kernel<T>, dim3(blocks_for((int64_t)rows * cols)), dim3(256), 0, s,
                     reinterpret_cast<const T*>(X), ldx, reinterpret_cast<T*>(Y), ldy, rows,
                     cols, nullptr, nullptr);
}

hipError_t fm_scaling_copy(int dtype_f64, const void* X, int64_t ldx, void* Y, int64_t ldy,
                           int rows, int cols, hipStream_t s) {
  if (dtype_f64)
    copy_mat<double>(X, ldx, Y, ldy, rows, cols, s);
  else
    copy_mat<float>(X, ldx, Y, ldy, rows, cols, s);
  return hipGetLastError();
}

hipError_t fm_scaling_post(int dtype_f64, const void* P, int64_t ldp, void* C, int64_t ldc,
                           int rows, int cols, const void* r, const void* sv, double alpha,
                           double beta, hipStream_t s) {
  const unsigned g = blocks_for((int64_t)rows * cols);
  if (dtype_f64)
    hipLaunchKernelGGL(fm_post_kernel<double>, dim3(g), dim3(256), 0, s,
                       reinterpret_cast<const double*>(P), ldp, reinterpret_cast<double*>(C), ldc,
                       rows, cols, reinterpret_cast<const double*>(r),
                       reinterpret_cast<const double*>(sv), alpha, beta);
  else
    hipLaunchKernelGGL(fm_post_kernel<float>, dim3(g), dim3(256), 0, s,
                       reinterpret_cast<const float*>(P), ldp, reinterpret_cast<float*>(C), ldc,
                       rows, cols, reinterpret_cast<const float*>(r),
                       reinterpret_cast<const float*>(sv), (float)alpha, (float)beta);
  return hipGetLastError();
}
