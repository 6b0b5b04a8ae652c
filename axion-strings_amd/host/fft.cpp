// fft.cpp -- see fft.hpp.
#include "fft.hpp"

#include <algorithm>
#include <cmath>
#include <thread>
#include <vector>

namespace axs_host {

namespace {

struct Plan {
    int n = 0;
    int sign = -1;
    std::vector<int> radices;
    std::vector<cplx> tw;     // tw[t] = exp(sign 2 pi i t / n)
};

Plan make_plan(int n, int sign)
{
    Plan p;
    p.n = n;
    p.sign = sign;
    int r = n;
    while (r % 4 == 0) {
        p.radices.push_back(4);
        r /= 4;
    }
    while (r % 2 == 0) {
        p.radices.push_back(2);
        r /= 2;
    }
    for (int f = 3; r > 1;) {
        if (r % f == 0) {
            p.radices.push_back(f);
            r /= f;
        } else {
            f += 2;
        }
    }
    p.tw.resize(n);
    for (int t = 0; t < n; ++t) {
        double a = sign * 2.0 * M_PI * (double)t / (double)n;
        p.tw[t] = cplx(std::cos(a), std::sin(a));
    }
    return p;
}

// One Stockham stage over s interleaved sequences of the current length n:
// x[q + s*(pp + r*m)] -> y[q + s*(p*pp + u)], with the radix-p DFT over r and
// the twiddle w_n^(pp*u).  The global table holds roots of the full length
// `ntot`, so w_n^e = tw[e * ntot/n].
void stage(const Plan &P, int n, int p, int64_t s, const cplx *x, cplx *y)
{
    const int m = n / p;
    const int64_t step = P.n / n;
    const double sg = (double)P.sign;
    for (int pp = 0; pp < m; ++pp) {
        for (int64_t q = 0; q < s; ++q) {
            const cplx *xi = x + q + s * pp;
            cplx *yo = y + q + s * (int64_t)p * pp;
            if (p == 2) {
                cplx a = xi[0], b = xi[s * m];
                yo[0] = a + b;
                yo[s] = (a - b) * P.tw[(pp * step) % P.n];
            } else if (p == 4) {
                cplx a0 = xi[0], a1 = xi[s * m], a2 = xi[2 * s * m], a3 = xi[3 * s * m];
                cplx t0 = a0 + a2, t1 = a0 - a2, t2 = a1 + a3, t3 = a1 - a3;
                // multiply by sign*i
                cplx it3(-sg * t3.imag(), sg * t3.real());
                cplx b0 = t0 + t2, b1 = t1 + it3, b2 = t0 - t2, b3 = t1 - it3;
                yo[0] = b0;
                yo[s] = b1 * P.tw[((int64_t)pp * 1 * step) % P.n];
                yo[2 * s] = b2 * P.tw[((int64_t)pp * 2 * step) % P.n];
                yo[3 * s] = b3 * P.tw[((int64_t)pp * 3 * step) % P.n];
            } else if (p == 3) {
                cplx a0 = xi[0], a1 = xi[s * m], a2 = xi[2 * s * m];
                const double c3 = -0.5, s3 = sg * std::sqrt(3.0) / 2.0;
                cplx t = a1 + a2;
                cplx d = a1 - a2;
                cplx b0 = a0 + t;
                cplx base = a0 + c3 * t;
                cplx rot(-s3 * d.imag(), s3 * d.real());
                cplx b1 = base + rot, b2 = base - rot;
                yo[0] = b0;
                yo[s] = b1 * P.tw[((int64_t)pp * 1 * step) % P.n];
                yo[2 * s] = b2 * P.tw[((int64_t)pp * 2 * step) % P.n];
            } else {
                // generic radix: direct DFT of size p with roots w_p = w_n^m
                cplx small[64];
                std::vector<cplx> big;
                cplx *a = small;
                if (p > 64) {
                    big.resize(p);
                    a = big.data();
                }
                for (int r = 0; r < p; ++r) a[r] = xi[s * (int64_t)r * m];
                for (int u = 0; u < p; ++u) {
                    cplx acc = a[0];
                    for (int r = 1; r < p; ++r)
                        acc += a[r] * P.tw[(((int64_t)r * u) % p) * m * step % P.n];
                    yo[s * u] = acc * P.tw[((int64_t)pp * u * step) % P.n];
                }
            }
        }
    }
}

// transform s interleaved sequences of length P.n held in x (x[q + s*t]);
// y is scratch of the same size; result left in x
void run(const Plan &P, int64_t s, cplx *x, cplx *y)
{
    int n = P.n;
    int64_t ss = s;
    cplx *a = x, *b = y;
    for (int p : P.radices) {
        stage(P, n, p, ss, a, b);
        n /= p;
        ss *= p;
        std::swap(a, b);
    }
    if (a != x) std::copy(a, a + (int64_t)P.n * s, x);
}

}  // namespace

void fft_1d(cplx *data, int n, int sign)
{
    Plan P = make_plan(n, sign);
    std::vector<cplx> w(n);
    run(P, 1, data, w.data());
}

void fft_nd(cplx *data, int ndims, int N, int sign, int threads)
{
    if (threads <= 0) {
        unsigned hw = std::thread::hardware_concurrency();
        threads = (int)std::min(hw ? hw : 4u, 16u);
    }
    Plan P = make_plan(N, sign);
    int64_t L = 1;
    for (int d = 0; d < ndims; ++d) L *= N;
    for (int axis = 0; axis < ndims; ++axis) {
        // stride of this axis; last axis is contiguous
        int64_t st = 1;
        for (int d = axis + 1; d < ndims; ++d) st *= N;
        const int64_t outer = L / (st * N);          // index above the axis
        const int B = (st == 1) ? 1 : (int)std::min<int64_t>(st, 16);
        const int64_t blocks_per_outer = (st + B - 1) / B;
        const int64_t nblocks = outer * blocks_per_outer;
        auto work = [&](int tid) {
            std::vector<cplx> buf((size_t)N * B), scratch((size_t)N * B);
            for (int64_t blk = tid; blk < nblocks; blk += threads) {
                int64_t o = blk / blocks_per_outer;
                int64_t lo0 = (blk % blocks_per_outer) * B;
                int nb = (int)std::min<int64_t>(B, st - lo0);
                cplx *base = data + o * st * N + lo0;
                for (int t = 0; t < N; ++t)
                    for (int b = 0; b < nb; ++b) buf[(size_t)t * nb + b] = base[t * st + b];
                run(P, nb, buf.data(), scratch.data());
                for (int t = 0; t < N; ++t)
                    for (int b = 0; b < nb; ++b) base[t * st + b] = buf[(size_t)t * nb + b];
            }
        };
        std::vector<std::thread> pool;
        for (int t = 0; t < threads; ++t) pool.emplace_back(work, t);
        for (auto &t : pool) t.join();
    }
}

}  // namespace axs_host
