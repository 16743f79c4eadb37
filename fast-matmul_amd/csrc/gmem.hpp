// gmem.hpp -- global-memory pointers for device code.
//
// Operand addresses reach the kernels through descriptor tables, so the
// compiler cannot tell that they point to global memory and emits flat
// loads and stores.  A flat access counts against the LDS wait counter as
// well as the vector-memory one, and may complete out of order with LDS
// reads, so every wait for an LDS read then also waits for all global
// loads in flight: a GEMM's register prefetch of the next K slice stalls
// the first MFMAs that follow it.  Casting to the global address space
// gives global_load/global_store, which only the vector-memory counter
// tracks.
#pragma once

#include <hip/hip_runtime.h>

template <typename T>
using fm_gptr = __attribute__((address_space(1))) T*;

template <typename T>
__device__ __forceinline__ fm_gptr<T> fm_global(T* p) {
  return (fm_gptr<T>)p;
}

// fm_ptr<G, T>: T* in the global address space when G, generic otherwise
// (for kernels whose measured speed decides the form per element type).
template <bool G, typename T>
struct fm_ptr_sel {
  using type = T*;
};
template <typename T>
struct fm_ptr_sel<true, T> {
  using type = fm_gptr<T>;
};
template <bool G, typename T>
using fm_ptr = typename fm_ptr_sel<G, T>::type;
