// axs_snapshot.hip -- asynchronous snapshots of a member's state.
//
// The reference writes snapshots synchronously from host arrays
// (fileio.cpp:15-46 called at s_run.cpp:166-199).  Here a snapshot is taken
// in two stages so that stepping continues while the data travels:
//   1. on the context's stream: the requested fields of the first `count`
//      sites are copied (FP32: widened from the float4 state and the
//      canonical velocity/kernel buffers) into a device staging buffer --
//      a device-to-device copy, milliseconds even at 1024^3;
//   2. on a second stream (default priority), once stage 1 is done: one
//      device-to-host copy of the staging buffer into pinned host memory.
// The next step only waits for stage 1 (it is on the same stream); stage 2
// overlaps with it.  snapshot_end waits for stage 2 and unpacks the pinned
// copy into the caller's arrays; it may run on another host thread while
// axs_step runs.  One snapshot is in flight per context at a time.
#include <string.h>

#include "axs_internal.hpp"
#include "axs_lattice.hpp"

namespace axs {

// staging layout: for each requested field in the order phi, phidot, ker,
// `count` pairs (double pairs in the FP64 form, float pairs in FP32)
static int nfields(int fields)
{
    return ((fields & AXS_SNAP_PHI) ? 1 : 0) + ((fields & AXS_SNAP_VEL) ? 1 : 0) +
           ((fields & AXS_SNAP_KER) ? 1 : 0);
}

__global__ void __launch_bounds__(256)
pack_phi32_kernel(uint32_t n, const fvec4 *__restrict__ st, fvec2 *__restrict__ out)
{
    uint32_t m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m < n) {
        fvec4 v = st[m];
        fvec2 o;
        o.x = v.x;
        o.y = v.y;
        out[m] = o;
    }
}

void destroy_snapshot(Context *c)
{
    if (c->copy_stream) (void)hipStreamSynchronize(c->copy_stream);
    if (c->snap_dev) (void)hipFree(c->snap_dev);
    if (c->snap_host) (void)hipHostFree(c->snap_host);
    if (c->snap_staged) (void)hipEventDestroy(c->snap_staged);
    if (c->snap_done) (void)hipEventDestroy(c->snap_done);
    if (c->copy_stream) (void)hipStreamDestroy(c->copy_stream);
    c->snap_dev = c->snap_host = nullptr;
    c->copy_stream = nullptr;
}

int snapshot_begin(Context *c, Member &m, int fields, int64_t count)
{
    if (!m.has_state) {
        set_error("member has no state to snapshot");
        return AXS_ERR_STATE;
    }
    const int nf = nfields(fields);
    const uint32_t n = (uint32_t)count;
    const size_t pair = (c->prec == AXS_FP64) ? 16 : 8;
    const size_t per_field = (size_t)n * pair;
    const size_t bytes = per_field * (size_t)nf;
    int expected = 0;
    if (!c->snap_pending.compare_exchange_strong(expected, 1)) {
        set_error("a snapshot is already in flight: call axs_snapshot_end first");
        return AXS_ERR_STATE;
    }
    auto fail = [&](int rc) {
        c->snap_pending.store(0);
        return rc;
    };
    if (!c->copy_stream) {
        if (hipStreamCreateWithFlags(&c->copy_stream, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&c->snap_staged, hipEventDisableTiming) != hipSuccess ||
            hipEventCreateWithFlags(&c->snap_done, hipEventDisableTiming) != hipSuccess) {
            set_error("snapshot stream/event creation failed");
            return fail(AXS_ERR_HIP);
        }
    }
    // buffers grow on demand; the previous snapshot's copy has completed
    if (bytes > c->snap_dev_bytes) {
        if (c->snap_dev) {
            (void)hipStreamSynchronize(c->stream);
            (void)hipFree(c->snap_dev);
            c->snap_dev = nullptr;
            c->snap_dev_bytes = 0;
        }
        if (hipMalloc(&c->snap_dev, bytes) != hipSuccess) {
            set_error("snapshot staging of %zu bytes failed", bytes);
            return fail(AXS_ERR_OUT_OF_MEMORY);
        }
        c->snap_dev_bytes = bytes;
    }
    if (bytes > c->snap_host_bytes) {
        if (c->snap_host) {
            (void)hipHostFree(c->snap_host);
            c->snap_host = nullptr;
            c->snap_host_bytes = 0;
        }
        if (hipHostMalloc(&c->snap_host, bytes, hipHostMallocDefault) != hipSuccess) {
            set_error("pinned snapshot buffer of %zu bytes failed", bytes);
            return fail(AXS_ERR_OUT_OF_MEMORY);
        }
        c->snap_host_bytes = bytes;
    }
    // stage 1 on the compute stream
    char *dst = (char *)c->snap_dev;
    if (c->prec == AXS_FP64) {
        const int slots[3] = {PHI, VEL, KER};
        const int bits[3] = {AXS_SNAP_PHI, AXS_SNAP_VEL, AXS_SNAP_KER};
        for (int f = 0; f < 3; ++f) {
            if (!(fields & bits[f])) continue;
            if (hipMemcpyAsync(dst, m.state[c->cur][slots[f]], per_field, hipMemcpyDeviceToDevice,
                               c->stream) != hipSuccess) {
                set_error("snapshot staging copy failed");
                return fail(AXS_ERR_HIP);
            }
            dst += per_field;
        }
    } else {
        if (fields & (AXS_SNAP_VEL | AXS_SNAP_KER)) {
            // canonical phidot and ker of the current state and tau
            int rc = fp32_refresh_aux(c, m);
            if (rc != AXS_OK) return fail(rc);
        }
        if (fields & AXS_SNAP_PHI) {
            hipLaunchKernelGGL(pack_phi32_kernel, dim3((n + 255) / 256), dim3(256), 0, c->stream,
                               n, (const fvec4 *)m.state[c->cur][PHI], (fvec2 *)dst);
            dst += per_field;
        }
        if ((fields & AXS_SNAP_VEL) &&
            hipMemcpyAsync(dst, m.aux_vel, per_field, hipMemcpyDeviceToDevice, c->stream) !=
                hipSuccess) {
            set_error("snapshot staging copy failed");
            return fail(AXS_ERR_HIP);
        }
        if (fields & AXS_SNAP_VEL) dst += per_field;
        if ((fields & AXS_SNAP_KER) &&
            hipMemcpyAsync(dst, m.aux_ker, per_field, hipMemcpyDeviceToDevice, c->stream) !=
                hipSuccess) {
            set_error("snapshot staging copy failed");
            return fail(AXS_ERR_HIP);
        }
    }
    // stage 2 on the copy stream, after stage 1
    if (hipGetLastError() != hipSuccess || hipEventRecord(c->snap_staged, c->stream) != hipSuccess ||
        hipStreamWaitEvent(c->copy_stream, c->snap_staged, 0) != hipSuccess ||
        (bytes && hipMemcpyAsync(c->snap_host, c->snap_dev, bytes, hipMemcpyDeviceToHost,
                                 c->copy_stream) != hipSuccess) ||
        hipEventRecord(c->snap_done, c->copy_stream) != hipSuccess) {
        set_error("snapshot copy could not be queued");
        (void)hipStreamSynchronize(c->copy_stream);
        return fail(AXS_ERR_HIP);
    }
    c->snap_fields = fields;
    c->snap_count = count;
    c->snap_tau = c->tau;
    return AXS_OK;
}

struct Unpack {
    const void *src;
    int f64;
    double *a, *b;
};

static void unpack_fn(size_t lo, size_t hi, void *arg)
{
    Unpack *u = (Unpack *)arg;
    if (u->f64) {
        const double *s = (const double *)u->src;
        for (size_t i = lo; i < hi; ++i) {
            if (u->a) u->a[i] = s[2 * i];
            if (u->b) u->b[i] = s[2 * i + 1];
        }
    } else {
        const float *s = (const float *)u->src;
        for (size_t i = lo; i < hi; ++i) {
            if (u->a) u->a[i] = (double)s[2 * i];
            if (u->b) u->b[i] = (double)s[2 * i + 1];
        }
    }
}

int snapshot_end(Context *c, double *out[6], double *tau)
{
    if (c->snap_pending.load() != 1) {
        set_error("no snapshot in flight: call axs_snapshot_begin first");
        return AXS_ERR_STATE;
    }
    hipError_t e = hipEventSynchronize(c->snap_done);
    if (e != hipSuccess) {
        c->snap_pending.store(0);
        set_error("snapshot copy failed: %s", hipGetErrorString(e));
        return AXS_ERR_HIP;
    }
    const size_t pair = (c->prec == AXS_FP64) ? 16 : 8;
    const size_t per_field = (size_t)c->snap_count * pair;
    const char *src = (const char *)c->snap_host;
    const int bits[3] = {AXS_SNAP_PHI, AXS_SNAP_VEL, AXS_SNAP_KER};
    for (int f = 0; f < 3; ++f) {
        if (!(c->snap_fields & bits[f])) continue;
        Unpack u{src, c->prec == AXS_FP64, out[2 * f], out[2 * f + 1]};
        if (u.a || u.b) parallel_for((size_t)c->snap_count, unpack_fn, &u);
        src += per_field;
    }
    if (tau) *tau = c->snap_tau;
    c->snap_pending.store(0);
    return AXS_OK;
}

}  // namespace axs
