// Global-memory views of device pointers.
//
// Pointers that a kernel reads out of a descriptor in memory are generic to
// the compiler, which then emits flat loads and stores.  A flat access counts
// against both the vector-memory and the LDS/scalar-memory wait counters, so
// every wait for an LDS access or a scalar constant load (the round constants
// of the hash permutations) also waits for the HBM accesses in flight.
// Viewed through the global address space the same accesses become global_*
// instructions, which count against the vector-memory counter only.
#pragma once
#include <hip/hip_runtime.h>

namespace zkx {

template <class T>
using gptr = __attribute__((address_space(1))) T *;

template <class T>
__device__ __forceinline__ gptr<T> gmem(T *p)
{
    return (gptr<T>)p;
}

// Read-only data that no kernel writes while it runs (descriptor lists the
// host fills before the launch), viewed through the constant address space:
// wave-uniform reads become scalar loads into SGPRs even after the kernel
// has stored to global memory, which would otherwise keep the compiler on
// vector loads.
template <class T>
using cptr = __attribute__((address_space(4))) const T *;

template <class T>
__device__ __forceinline__ cptr<T> cmem(const T *p)
{
    return (cptr<T>)p;
}

}  // namespace zkx
