// axs_vec.hpp -- small device helpers shared by the streaming stage kernels
// (axs_rk.hip, axs_pair.hip): periodic index wrapping, the Star type of a
// precision, and VX-word vector loads and stores.
#pragma once

#include <hip/hip_runtime.h>

#include "axs_stage_math.hpp"

namespace axs
{

__device__ __forceinline__ int wrap_lo(int i, int n) { return i < 0 ? i + n : i; }
__device__ __forceinline__ int wrap_hi(int i, int n) { return i >= n ? i - n : i; }

template <typename T> struct StarOf;
template <> struct StarOf<double>
{
    using type = Star;
};
template <> struct StarOf<float>
{
    using type = StarF;
};

// VX consecutive x cells per lane: vector loads/stores of VX words
template <typename T, int VX> struct VecOf
{
    typedef T type __attribute__((ext_vector_type(VX)));
};

template <typename T, int VX>
__device__ __forceinline__ void vload(const T *p, T (&o)[VX])
{
    if constexpr (VX == 1)
        o[0] = p[0];
    else
    {
        const typename VecOf<T, VX>::type v =
            *reinterpret_cast<const typename VecOf<T, VX>::type *>(p);
#pragma unroll
        for (int i = 0; i < VX; ++i)
            o[i] = v[i];
    }
}

// the same with the non-temporal cache policy: words read once, at the own
// cell, which no neighbouring tile reads (the L2 keeps the stencil fields)
template <typename T, int VX>
__device__ __forceinline__ void vload_nt(const T *p, T (&o)[VX])
{
    if constexpr (VX == 1)
        o[0] = __builtin_nontemporal_load(p);
    else
    {
        const typename VecOf<T, VX>::type v = __builtin_nontemporal_load(
            reinterpret_cast<const typename VecOf<T, VX>::type *>(p));
#pragma unroll
        for (int i = 0; i < VX; ++i)
            o[i] = v[i];
    }
}

template <typename T, int VX>
__device__ __forceinline__ void vstore(T *p, const T (&x)[VX], bool nt)
{
    if constexpr (VX == 1)
    {
        if (nt)
            __builtin_nontemporal_store(x[0], p);
        else
            p[0] = x[0];
    }
    else
    {
        typename VecOf<T, VX>::type v;
#pragma unroll
        for (int i = 0; i < VX; ++i)
            v[i] = x[i];
        auto *q = reinterpret_cast<typename VecOf<T, VX>::type *>(p);
        if (nt)
            __builtin_nontemporal_store(v, q);
        else
            *q = v;
    }
}

} // namespace axs
