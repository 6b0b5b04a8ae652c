// axs_lattice.hpp -- device-side lattice indexing shared by the kernels.
//
// Layout (reference src/common/common.hpp:28, 81-119):
//   2D  m = i + N*j            axis 0 = i (stride 1), axis 1 = j (stride N)
//   3D  m = (i*N + j)*N + k    axis 0 = i (stride N*N), axis 1 = j (stride N),
//                              axis 2 = k (stride 1)
// "Axis order" is the reference's COO insertion order of neighbours
// (s_integrate.cpp:50-75, 105-142): for each distance l, +l then -l along
// axis 0, then axis 1, then axis 2.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace axs {

// Storage vector types.  Device state is interleaved per site:
//   FP64 form  PHI = (phi1, phi2), VEL = (phidot1, phidot2), KER = (ker1, ker2),
//              one dvec2 array each;
//   FP32 form  one fvec4 array (phi1, phi2, u1, u2), u = phidot + dt/2 ker,
//              with fvec2 (ker1, ker2) and (phidot1, phidot2) read-back buffers.
typedef double dvec2 __attribute__((ext_vector_type(2)));
typedef float fvec4 __attribute__((ext_vector_type(4)));
typedef float fvec2 __attribute__((ext_vector_type(2)));

// read-only views of the field and its velocity, widened to double
struct PhiF64 {
    const dvec2 *__restrict__ p;
    __device__ __forceinline__ void get(uint32_t m, double &a, double &b) const
    {
        dvec2 v = p[m];
        a = v.x;
        b = v.y;
    }
};
struct PhiF32 {
    const fvec4 *__restrict__ p;
    __device__ __forceinline__ void get(uint32_t m, double &a, double &b) const
    {
        fvec4 v = p[m];
        a = (double)v.x;
        b = (double)v.y;
    }
};
struct VelF64 {
    const dvec2 *__restrict__ p;
    __device__ __forceinline__ void get(uint32_t m, double &a, double &b) const
    {
        dvec2 v = p[m];
        a = v.x;
        b = v.y;
    }
};
struct VelF32 {
    const fvec2 *__restrict__ p;
    __device__ __forceinline__ void get(uint32_t m, double &a, double &b) const
    {
        fvec2 v = p[m];
        a = (double)v.x;
        b = (double)v.y;
    }
};

template <typename V>
__device__ __forceinline__ void store_nt(V *p, V v)
{
    __builtin_nontemporal_store(v, p);
}

// Value of the next (lane + 1) or previous (lane - 1) lane of the wave, by
// the whole-wave DPP shifts (wave_shl:1, wave_shr:1): a register operand,
// no LDS crossbar, no address.  The end lanes (63, 0) read zero.
__device__ __forceinline__ uint32_t dpp_next_u32(uint32_t v)
{
    return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x130, 0xf, 0xf, true);
}
__device__ __forceinline__ uint32_t dpp_prev_u32(uint32_t v)
{
    return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x138, 0xf, 0xf, true);
}
__device__ __forceinline__ float lane_next(float v)
{
    return __builtin_bit_cast(float, dpp_next_u32(__builtin_bit_cast(uint32_t, v)));
}
__device__ __forceinline__ float lane_prev(float v)
{
    return __builtin_bit_cast(float, dpp_prev_u32(__builtin_bit_cast(uint32_t, v)));
}
__device__ __forceinline__ double lane_next(double v)
{
    const uint64_t b = __builtin_bit_cast(uint64_t, v);
    const uint64_t lo = dpp_next_u32((uint32_t)b), hi = dpp_next_u32((uint32_t)(b >> 32));
    return __builtin_bit_cast(double, lo | (hi << 32));
}
__device__ __forceinline__ double lane_prev(double v)
{
    const uint64_t b = __builtin_bit_cast(uint64_t, v);
    const uint64_t lo = dpp_prev_u32((uint32_t)b), hi = dpp_prev_u32((uint32_t)(b >> 32));
    return __builtin_bit_cast(double, lo | (hi << 32));
}

// The FP32 form's arithmetic, shared by every FP32 kernel so that one-step
// and multi-step launches give identical bits.  The force is regrouped as
//     ker = alpha (c0 x + sum_m c_m n_m) + cl x S
//         = alpha * a + x * (alpha c0 + cl S),    a = sum_m c_m n_m
// (the neighbour sum a starts from -0.0, the additive identity), and every
// product-sum is a fused multiply-add: one rounding each and a third fewer
// instructions per site than the reference's expression tree.  The step
// folds dt into the scalars, u' = u + dt ker = (u + x tdt) + dta a with
// tdt = dtcl S + dtac0 (dta = dt alpha, dtcl = dt cl, dtac0 = dt alpha c0),
// two multiply-adds per component where force then u + dt ker take three; the
// read-back of ker keeps the unscaled force.  The FP64 form keeps the
// reference's tree and separate roundings, which its bit-for-bit parity
// needs.
struct F32Math {
    // drifted field x = phi + dt u
    static __device__ __forceinline__ float drift(float phi, float u, float dt)
    {
        return __builtin_fmaf(dt, u, phi);
    }
    // Laplacian accumulation a + c v (c = 1 for the 5/7-point stencil: the
    // product is exact, so this is the plain sum a + v)
    static __device__ __forceinline__ float acc(float a, float c, float v)
    {
        return __builtin_fmaf(c, v, a);
    }
    // potential factor S = x^2 + y^2 + pot0, pot0 = T0^2/3 - tau^2
    static __device__ __forceinline__ float pot(float x, float y, float pot0)
    {
        return __builtin_fmaf(x, x, __builtin_fmaf(y, y, pot0));
    }
    // the site's own factor t = alpha c0 + cl S (shared by both components)
    static __device__ __forceinline__ float own(float cl, float S, float ac0)
    {
        return __builtin_fmaf(cl, S, ac0);
    }
    // ker = alpha a + x t
    static __device__ __forceinline__ float force(float alpha, float a, float x, float t)
    {
        return __builtin_fmaf(alpha, a, x * t);
    }
    // the step's kick with dt folded in: u' = (u + x tdt) + dta a, where
    // tdt = own(dtcl, S, dtac0) is dt t for both components
    static __device__ __forceinline__ float kick_dt(float u, float dta, float a, float x,
                                                    float tdt)
    {
        return __builtin_fmaf(dta, a, __builtin_fmaf(x, tdt, u));
    }
};

// The dispatcher deals the work-groups of a launch to the eight XCDs in
// turn (work-group b runs on XCD b % 8).  xcd_order(b, nb) gives XCD x the
// contiguous logical range [x q + min(x, r), ...) of q = nb / 8 indices, one
// more for the first r = nb % 8 XCDs: a bijection of [0, nb) for any nb, so
// tiles that are neighbours in the logical order share one XCD's L2.
__device__ __forceinline__ uint32_t xcd_order(uint32_t b, uint32_t nb)
{
    const uint32_t x = b & 7u, q = nb >> 3, r = nb & 7u;
    return x * q + (x < r ? x : r) + (b >> 3);
}

__device__ __forceinline__ int wrap_coord(int x, int N)
{
    // valid for -N <= x < 2N, which holds for every stencil radius <= 4 on
    // lattices with N >= 4 (checked at context creation)
    x = (x >= N) ? x - N : x;
    return (x < 0) ? x + N : x;
}

template <int ND>
struct Site {
    int c[3];        // coordinates in axis order
    uint32_t m;      // linear index
};

template <int ND>
__device__ __forceinline__ Site<ND> decompose(uint32_t m, int N)
{
    Site<ND> s;
    s.m = m;
    if (ND == 2) {
        s.c[0] = (int)(m % (uint32_t)N);
        s.c[1] = (int)(m / (uint32_t)N);
        s.c[2] = 0;
    } else {
        uint32_t NN = (uint32_t)N * (uint32_t)N;
        s.c[0] = (int)(m / NN);
        uint32_t layer = m - (uint32_t)s.c[0] * NN;
        s.c[1] = (int)(layer / (uint32_t)N);
        s.c[2] = (int)(layer % (uint32_t)N);
    }
    return s;
}

template <int ND>
__device__ __forceinline__ uint32_t axis_stride(int axis, int N)
{
    if (ND == 2) return axis == 0 ? 1u : (uint32_t)N;
    return axis == 0 ? (uint32_t)N * (uint32_t)N : (axis == 1 ? (uint32_t)N : 1u);
}

// index of the site displaced by d along `axis` (periodic)
template <int ND>
__device__ __forceinline__ uint32_t shifted(const Site<ND> &s, int axis, int d, int N)
{
    int x = wrap_coord(s.c[axis] + d, N);
    int delta = x - s.c[axis];
    return (uint32_t)((int64_t)s.m + (int64_t)delta * (int64_t)axis_stride<ND>(axis, N));
}

}  // namespace axs
