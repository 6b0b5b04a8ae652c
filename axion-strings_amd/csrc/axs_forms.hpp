// axs_forms.hpp -- the per-site arithmetic of one time step, as used by the
// multi-step (temporally blocked) kernels, for the two storage forms.
//
// A kernel level computes, at a site whose drifted field x and neighbour
// sum `a` (centre coefficient times x, then the neighbours in COO order) it
// already has:
//     ker = alpha * a + cl * x * S,   S = x.x^2 + x.y^2 - tau2 + T03
// (a starts from the form's centre term), then the form's kick and the
// next step's drift:
//   F32Form  (phi, u) in one float4, u = phidot + dt/2 ker: a is the
//            neighbour sum alone and ker = alpha a + x (alpha c0 + cl S);
//            u' = u + dt ker;   x' = x + dt u'  (fused multiply-adds of
//            F32Math, as in every FP32 kernel)
//   F64Form  (phi, phidot, ker) in three double2 arrays, the reference's
//            velocity-Verlet (s_integrate.cpp:311-347):
//            v' = v + 0.5 (k + ker) dt;  x' = x + dt (v' + 0.5 ker dt)
//            (Hamiltonian form, :368-369: x' = x + dt/tau^2 (v' + 0.5f dt ker))
// so a pass of D levels gives the bits of D one-step launches.
#pragma once

#include "axs_internal.hpp"
#include "axs_lattice.hpp"

namespace axs {

template <class Sc, int D>
struct LevelScalars {
    Sc s[D];       // step l+1 uses s[l]
};

struct F32Form {
    using T = float;
    using Sc = StepScalarsF;
    struct In {
        const fvec4 *__restrict__ st;
    };
    struct Out {
        fvec4 *__restrict__ st;
    };
    using Raw = fvec4;
    struct Own {
        float u1, u2;
    };
    // own data of a site whose value is never used: any bits (no moves)
    static __device__ __forceinline__ Own unset()
    {
        return Own{__builtin_nondeterministic_value(0.0f), __builtin_nondeterministic_value(0.0f)};
    }
    // values a level reads from vector registers, set once per work-group:
    // a multiply-add reads one scalar register at most, and dtac0 is the
    // same at every level of a launch (dt, alpha and c0 do not change with
    // tau); the empty asm keeps it in a vector register instead of a copy
    // from the scalar one at every use
    struct Vc {
        float dtac0;
    };
    static __device__ __forceinline__ Vc vconst(const Sc &s)
    {
        Vc v{s.dtac0};
        __asm__ volatile("" : "+v"(v.dtac0));
        return v;
    }
    static __device__ __forceinline__ void fetch(const In &a, uint32_t q, Raw &r) { r = a.st[q]; }
    // Laplacian accumulation: start (the centre joins the potential), a + c v
    static __device__ __forceinline__ T centre(const Sc &, T) { return -0.0f; }
    static __device__ __forceinline__ T acc(T a, T c, T v) { return F32Math::acc(a, c, v); }
    static __device__ __forceinline__ void drift0(const Raw &r, const Sc &s, T &x, T &y, Own &o)
    {
        x = F32Math::drift(r.x, r.z, s.dt);
        y = F32Math::drift(r.y, r.w, s.dt);
        o.u1 = r.z;
        o.u2 = r.w;
    }
    // one step at a site; nx, ny: the next step's drift (sn: its scalars)
    static __device__ __forceinline__ void level(T x, T y, T a, T c, const Own &o, const Sc &s,
                                                 const Sc &sn, const Vc &vc, Own &on, T &nx,
                                                 T &ny)
    {
        const T tdt = F32Math::own(s.dtcl, F32Math::pot(x, y, s.pot0), vc.dtac0);
        on.u1 = F32Math::kick_dt(o.u1, s.dta, a, x, tdt);
        on.u2 = F32Math::kick_dt(o.u2, s.dta, c, y, tdt);
        nx = F32Math::drift(x, on.u1, sn.dt);
        ny = F32Math::drift(y, on.u2, sn.dt);
    }
    static __device__ __forceinline__ void store(const Out &o, uint32_t q, T x, T y, const Own &w)
    {
        fvec4 v;
        v.x = x;
        v.y = y;
        v.z = w.u1;
        v.w = w.u2;
        store_nt(o.st + q, v);
    }
};

template <bool HAM>
struct F64FormT {
    using T = double;
    using Sc = StepScalars;
    struct In {
        const dvec2 *__restrict__ phi, *__restrict__ vel, *__restrict__ ker;
    };
    struct Out {
        dvec2 *__restrict__ phi, *__restrict__ vel, *__restrict__ ker;
    };
    struct Raw {
        dvec2 p, v, k;
    };
    struct Own {
        dvec2 v, k;
    };
    struct Vc {
    };
    static __device__ __forceinline__ Vc vconst(const Sc &) { return Vc{}; }
    // own data of a site whose value is never used: any bits (no moves)
    static __device__ __forceinline__ Own unset()
    {
        Own o;
        o.v.x = __builtin_nondeterministic_value(0.0);
        o.v.y = __builtin_nondeterministic_value(0.0);
        o.k.x = __builtin_nondeterministic_value(0.0);
        o.k.y = __builtin_nondeterministic_value(0.0);
        return o;
    }
    static __device__ __forceinline__ void fetch(const In &a, uint32_t q, Raw &r)
    {
        r.p = a.phi[q];
        r.v = a.vel[q];
        r.k = a.ker[q];
    }
    // Laplacian accumulation: c0 x, then a + c v, products and sums rounded
    // separately (s_integrate.cpp:322-324 in COO row order)
    static __device__ __forceinline__ T centre(const Sc &s, T x) { return s.c[0] * x; }
    static __device__ __forceinline__ T acc(T a, T c, T v) { return a + c * v; }
    static __device__ __forceinline__ void drift(T p, T v, T k, const Sc &s, T &x)
    {
        if (HAM)
            x = p + s.dcoef * (v + s.hdt * k);
        else
            x = p + s.dt * (v + 0.5 * k * s.dt);
    }
    static __device__ __forceinline__ void drift0(const Raw &r, const Sc &s, T &x, T &y, Own &o)
    {
        drift(r.p.x, r.v.x, r.k.x, s, x);
        drift(r.p.y, r.v.y, r.k.y, s, y);
        o.v = r.v;
        o.k = r.k;
    }
    static __device__ __forceinline__ void level(T x, T y, T a, T c, const Own &o, const Sc &s,
                                                 const Sc &sn, const Vc &, Own &on, T &nx,
                                                 T &ny)
    {
        const T lap1 = s.alpha * a, lap2 = s.alpha * c;
        const T S = x * x + y * y - s.tau2 + s.T03;
        on.k.x = lap1 + s.cl * x * S;
        on.k.y = lap2 + s.cl * y * S;
        on.v.x = o.v.x + 0.5 * (o.k.x + on.k.x) * s.dt;
        on.v.y = o.v.y + 0.5 * (o.k.y + on.k.y) * s.dt;
        drift(x, on.v.x, on.k.x, sn, nx);
        drift(y, on.v.y, on.k.y, sn, ny);
    }
    static __device__ __forceinline__ void store(const Out &o, uint32_t q, T x, T y, const Own &w)
    {
        dvec2 p;
        p.x = x;
        p.y = y;
        store_nt(o.phi + q, p);
        store_nt(o.vel + q, w.v);
        store_nt(o.ker + q, w.k);
    }
};

}  // namespace axs
