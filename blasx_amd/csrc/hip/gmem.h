// gmem.h -- loads and stores through the global address space.
//
// A pointer that a kernel reads from a descriptor table (rather than
// receives as a kernel argument) is generic to the compiler, which then
// emits FLAT memory instructions for it.  On CDNA a FLAT load counts
// against the LDS counter as well as the vector-memory counter, so every
// wait for an LDS read in the MFMA loop would also wait for the prefetch of
// the next K tile from HBM, exposing its full latency once per K tile.
// Casting to address space 1 makes the compiler emit GLOBAL instructions,
// which only the vector-memory counter tracks.
//
// The access goes through a scalar or vector type of the same size: a
// struct copied out of an address-space-1 pointer is copied by its (generic)
// copy constructor, which turns the access back into a FLAT one.
#pragma once
#include <stdint.h>

namespace gmem_detail {
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));
template <int BYTES> struct Word;
template <> struct Word<4> { typedef uint32_t type; };
template <> struct Word<8> { typedef u32x2 type; };
template <> struct Word<16> { typedef u32x4 type; };
} // namespace gmem_detail

template <typename T> __device__ __forceinline__ T gload(const T *p)
{
    typedef typename gmem_detail::Word<sizeof(T)>::type W;
    const W w = *(const __attribute__((address_space(1))) W *)p;
    return __builtin_bit_cast(T, w);
}

template <typename T> __device__ __forceinline__ void gstore(T *p, const T &v)
{
    typedef typename gmem_detail::Word<sizeof(T)>::type W;
    *(__attribute__((address_space(1))) W *)p = __builtin_bit_cast(W, v);
}
