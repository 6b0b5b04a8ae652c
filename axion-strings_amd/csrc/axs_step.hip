// axs_step.hip -- the velocity-Verlet step of vvsl_field_rescaled
// (reference src/standard/s_integrate.cpp:305-348) as fused gfx950 kernels.
//
// The reference makes five sweeps per step (drift, two sparse-matrix
// Laplacians, potential, kick, copy).  Here one kernel does the whole step:
// each work-item recomputes the drift of the 2*ND*r neighbours it needs
// (bit-identical to its owner's drift, since the expression is the same),
// applies the stencil directly (no matrix), adds the potential force, kicks,
// and writes the new state to the other half of a ping-pong pair, so the
// reference's ker copy becomes a pointer swap.
//
// FP64 form: every expression keeps the reference's tree and the library is
// built with -ffp-contract=off, so each result is the correctly rounded value
// of the same operation sequence as the CPU path -- bit-identical fields.
// FP32 form: phi and u = phidot + dt/2 ker are stored in FP32; one step is
//   phi' = phi + dt u;  ker' = F(phi', tau');  u' = u + dt ker'
// which is the same velocity-Verlet map written in half-step velocities.
#include "axs_internal.hpp"
#include "axs_lattice.hpp"

namespace axs {

struct In64 {
    const double *__restrict__ p1, *__restrict__ p2, *__restrict__ v1, *__restrict__ v2,
        *__restrict__ k1, *__restrict__ k2;
};
struct Out64 {
    double *__restrict__ p1, *__restrict__ p2, *__restrict__ v1, *__restrict__ v2,
        *__restrict__ k1, *__restrict__ k2;
};
struct In32 {
    const float *__restrict__ p1, *__restrict__ p2, *__restrict__ u1, *__restrict__ u2;
};
struct Out32 {
    float *__restrict__ p1, *__restrict__ p2, *__restrict__ u1, *__restrict__ u2;
};

// drift, s_integrate.cpp:313-314: phi += dt * (phidot + 0.5f * ker * dt)
__device__ __forceinline__ void drift64(const In64 &a, uint32_t q, double dt, double &x1,
                                        double &x2)
{
    x1 = a.p1[q] + dt * (a.v1[q] + 0.5 * a.k1[q] * dt);
    x2 = a.p2[q] + dt * (a.v2[q] + 0.5 * a.k2[q] * dt);
}

template <int ND>
__global__ void __launch_bounds__(256)
step_fp64_kernel(int N, uint32_t L, In64 in, Out64 out, StepScalars s)
{
    uint32_t m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= L) return;
    Site<ND> site = decompose<ND>(m, N);

    double x1, x2;
    drift64(in, m, s.dt, x1, x2);

    // Laplacian in COO row order (centre first), then alpha (s_integrate.cpp:322-324)
    double s1 = s.c[0] * x1, s2 = s.c[0] * x2;
    for (int l = 1; l < s.ncoef; ++l) {
        double cl = s.c[l];
#pragma unroll
        for (int a = 0; a < ND; ++a) {
            double y1, y2;
            drift64(in, shifted<ND>(site, a, l, N), s.dt, y1, y2);
            s1 = s1 + cl * y1;
            s2 = s2 + cl * y2;
            drift64(in, shifted<ND>(site, a, -l, N), s.dt, y1, y2);
            s1 = s1 + cl * y1;
            s2 = s2 + cl * y2;
        }
    }
    double lap1 = s.alpha * s1, lap2 = s.alpha * s2;

    // potential, s_integrate.cpp:333-334
    double S = x1 * x1 + x2 * x2 - s.tau2 + s.T03;
    double kn1 = lap1 + s.cl * x1 * S;
    double kn2 = lap2 + s.cl * x2 * S;

    // kick, s_integrate.cpp:339-340
    double kc1 = in.k1[m], kc2 = in.k2[m];
    out.v1[m] = in.v1[m] + 0.5 * (kc1 + kn1) * s.dt;
    out.v2[m] = in.v2[m] + 0.5 * (kc2 + kn2) * s.dt;
    out.p1[m] = x1;
    out.p2[m] = x2;
    out.k1[m] = kn1;
    out.k2[m] = kn2;
}

template <int ND>
__global__ void __launch_bounds__(256)
step_fp32_kernel(int N, uint32_t L, In32 in, Out32 out, StepScalarsF s)
{
    uint32_t m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= L) return;
    Site<ND> site = decompose<ND>(m, N);

    float x1 = in.p1[m] + s.dt * in.u1[m];
    float x2 = in.p2[m] + s.dt * in.u2[m];
    float s1 = s.c[0] * x1, s2 = s.c[0] * x2;
    for (int l = 1; l < s.ncoef; ++l) {
        float cl = s.c[l];
#pragma unroll
        for (int a = 0; a < ND; ++a) {
            uint32_t qp = shifted<ND>(site, a, l, N);
            uint32_t qm = shifted<ND>(site, a, -l, N);
            s1 = s1 + cl * (in.p1[qp] + s.dt * in.u1[qp]);
            s2 = s2 + cl * (in.p2[qp] + s.dt * in.u2[qp]);
            s1 = s1 + cl * (in.p1[qm] + s.dt * in.u1[qm]);
            s2 = s2 + cl * (in.p2[qm] + s.dt * in.u2[qm]);
        }
    }
    float S = x1 * x1 + x2 * x2 - s.tau2 + s.T03;
    float k1 = s.alpha * s1 + s.cl * x1 * S;
    float k2 = s.alpha * s2 + s.cl * x2 * S;
    out.u1[m] = in.u1[m] + s.dt * k1;
    out.u2[m] = in.u2[m] + s.dt * k2;
    out.p1[m] = x1;
    out.p2[m] = x2;
}

// FP32: ker = F(phi, tau) at the current state and phidot = u - dt/2 ker
template <int ND>
__global__ void __launch_bounds__(256)
canonical_fp32_kernel(int N, uint32_t L, In32 in, float *__restrict__ k1o,
                      float *__restrict__ k2o, float *__restrict__ v1o,
                      float *__restrict__ v2o, StepScalarsF s, int recompute)
{
    uint32_t m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= L) return;
    float k1, k2;
    if (recompute) {
        Site<ND> site = decompose<ND>(m, N);
        float x1 = in.p1[m], x2 = in.p2[m];
        float s1 = s.c[0] * x1, s2 = s.c[0] * x2;
        for (int l = 1; l < s.ncoef; ++l) {
            float cl = s.c[l];
#pragma unroll
            for (int a = 0; a < ND; ++a) {
                uint32_t qp = shifted<ND>(site, a, l, N);
                uint32_t qm = shifted<ND>(site, a, -l, N);
                s1 = s1 + cl * in.p1[qp];
                s2 = s2 + cl * in.p2[qp];
                s1 = s1 + cl * in.p1[qm];
                s2 = s2 + cl * in.p2[qm];
            }
        }
        float S = x1 * x1 + x2 * x2 - s.tau2 + s.T03;
        k1 = s.alpha * s1 + s.cl * x1 * S;
        k2 = s.alpha * s2 + s.cl * x2 * S;
        k1o[m] = k1;
        k2o[m] = k2;
    } else {
        k1 = k1o[m];
        k2 = k2o[m];
    }
    v1o[m] = in.u1[m] - s.half_dt * k1;
    v2o[m] = in.u2[m] - s.half_dt * k2;
}

static inline unsigned grid_for(uint32_t L, unsigned block) { return (L + block - 1) / block; }

int launch_step(Context *c, const StepScalars &s)
{
    const int src = c->cur, dst = c->cur ^ 1;
    for (Member &m : c->members) {
        uint32_t L = (uint32_t)m.L;
        unsigned grid = grid_for(L, 256);
        if (c->tune.kernel != AXS_KERNEL_SIMPLE && march3d_supported(c, m.N)) {
            AXS_TRY(launch_march3d(c, m, src, dst, s));
            continue;
        }
        if (c->prec == AXS_FP64) {
            In64 in{(const double *)m.state[src][PHI1], (const double *)m.state[src][PHI2],
                    (const double *)m.state[src][VEL1], (const double *)m.state[src][VEL2],
                    (const double *)m.state[src][KER1], (const double *)m.state[src][KER2]};
            Out64 out{(double *)m.state[dst][PHI1], (double *)m.state[dst][PHI2],
                      (double *)m.state[dst][VEL1], (double *)m.state[dst][VEL2],
                      (double *)m.state[dst][KER1], (double *)m.state[dst][KER2]};
            if (c->nd == 2)
                hipLaunchKernelGGL(step_fp64_kernel<2>, dim3(grid), dim3(256), 0, c->stream, m.N,
                                   L, in, out, s);
            else
                hipLaunchKernelGGL(step_fp64_kernel<3>, dim3(grid), dim3(256), 0, c->stream, m.N,
                                   L, in, out, s);
        } else {
            StepScalarsF f = to_float(s);
            In32 in{(const float *)m.state[src][PHI1], (const float *)m.state[src][PHI2],
                    (const float *)m.state[src][VEL1], (const float *)m.state[src][VEL2]};
            Out32 out{(float *)m.state[dst][PHI1], (float *)m.state[dst][PHI2],
                      (float *)m.state[dst][VEL1], (float *)m.state[dst][VEL2]};
            if (c->nd == 2)
                hipLaunchKernelGGL(step_fp32_kernel<2>, dim3(grid), dim3(256), 0, c->stream, m.N,
                                   L, in, out, f);
            else
                hipLaunchKernelGGL(step_fp32_kernel<3>, dim3(grid), dim3(256), 0, c->stream, m.N,
                                   L, in, out, f);
        }
    }
    AXS_HIP_TRY(hipGetLastError());
    c->cur = dst;
    return AXS_OK;
}

int launch_fp32_canonical(Context *c, Member &m, const StepScalars &s, bool recompute_ker)
{
    StepScalarsF f = to_float(s);
    In32 in{(const float *)m.state[c->cur][PHI1], (const float *)m.state[c->cur][PHI2],
            (const float *)m.state[c->cur][VEL1], (const float *)m.state[c->cur][VEL2]};
    uint32_t L = (uint32_t)m.L;
    unsigned grid = grid_for(L, 256);
    auto kern = (c->nd == 2) ? canonical_fp32_kernel<2> : canonical_fp32_kernel<3>;
    hipLaunchKernelGGL(kern, dim3(grid), dim3(256), 0, c->stream, m.N, L, in,
                       (float *)m.aux_ker[0], (float *)m.aux_ker[1], (float *)m.aux_vel[0],
                       (float *)m.aux_vel[1], f, recompute_ker ? 1 : 0);
    AXS_HIP_TRY(hipGetLastError());
    return AXS_OK;
}

}  // namespace axs
