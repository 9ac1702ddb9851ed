// kappa.cpp -- FP64 sparse kernels of the kappa check (host code, OpenMP):
// y = A x and the Jacobi-preconditioned conjugate-gradient solve of A x = b
// for a symmetric positive definite CSR matrix (a shifted graph Laplacian).
// The Lanczos iteration of the kappa check (dygrass_amd/kappa.py) applies the
// sparsifier's Laplacian inverse through this solve where a sparse LU
// factorisation would fill in (graphs with long-range edges).
//
// Reductions are summed over fixed blocks in block order, so every result is
// the same for any thread count.
#include <omp.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <vector>

namespace {

constexpr int64_t kBlock = 1 << 14;  // rows per reduction block

int64_t nblocks(int64_t n) { return (n + kBlock - 1) / kBlock; }

// sum over blocks of f(lo, hi), each block's partial summed in block order
template <class F>
double blocked_sum(int64_t n, int threads, F f) {
    const int64_t nb = nblocks(n);
    std::vector<double> part(nb);
#pragma omp parallel for schedule(static) num_threads(threads)
    for (int64_t b = 0; b < nb; ++b) part[b] = f(b * kBlock, std::min(n, (b + 1) * kBlock));
    double s = 0.0;
    for (int64_t b = 0; b < nb; ++b) s += part[b];
    return s;
}

void spmv(int64_t n, const int64_t* ip, const int32_t* ix, const double* a, const double* x,
          double* y, int threads) {
#pragma omp parallel for schedule(static, 4096) num_threads(threads)
    for (int64_t i = 0; i < n; ++i) {
        double s = 0.0;
        for (int64_t p = ip[i]; p < ip[i + 1]; ++p) s += a[p] * x[ix[p]];
        y[i] = s;
    }
}

}  // namespace

extern "C" {

// y = A x for an n-row CSR matrix (indptr ip[n+1], indices ix, values a).
void dyg_kappa_spmv(int64_t n, const int64_t* ip, const int32_t* ix, const double* a,
                    const double* x, double* y, int threads) {
    spmv(n, ip, ix, a, x, y, threads > 0 ? threads : 1);
}

// Solves A x = b by conjugate gradients preconditioned with A's diagonal,
// from x = 0, until ||b - A x|| <= tol * ||b|| or maxit iterations.
// Returns the iterations taken, or -1 when maxit was reached first, -2 on a
// non-positive diagonal or curvature (A not positive definite).
int dyg_kappa_pcg(int64_t n, const int64_t* ip, const int32_t* ix, const double* a,
                  const double* b, double* x, double tol, int maxit, int threads) {
    threads = threads > 0 ? threads : 1;
    std::vector<double> dinv(n), r(n), z(n), p(n), q(n);
    bool bad = false;
#pragma omp parallel for schedule(static) num_threads(threads) reduction(|| : bad)
    for (int64_t i = 0; i < n; ++i) {
        double d = 0.0;
        for (int64_t k = ip[i]; k < ip[i + 1]; ++k)
            if (ix[k] == i) d += a[k];
        if (!(d > 0.0)) bad = true;
        dinv[i] = d > 0.0 ? 1.0 / d : 0.0;
        x[i] = 0.0;
        r[i] = b[i];
        z[i] = dinv[i] * b[i];
        p[i] = z[i];
    }
    if (bad) return -2;
    const double bb = blocked_sum(n, threads, [&](int64_t lo, int64_t hi) {
        double s = 0.0;
        for (int64_t i = lo; i < hi; ++i) s += b[i] * b[i];
        return s;
    });
    if (bb == 0.0) return 0;
    const double stop = tol * tol * bb;
    double rz = blocked_sum(n, threads, [&](int64_t lo, int64_t hi) {
        double s = 0.0;
        for (int64_t i = lo; i < hi; ++i) s += r[i] * z[i];
        return s;
    });
    for (int it = 1; it <= maxit; ++it) {
        spmv(n, ip, ix, a, p.data(), q.data(), threads);
        const double pq = blocked_sum(n, threads, [&](int64_t lo, int64_t hi) {
            double s = 0.0;
            for (int64_t i = lo; i < hi; ++i) s += p[i] * q[i];
            return s;
        });
        if (!(pq > 0.0)) return -2;
        const double alpha = rz / pq;
        // x += alpha p, r -= alpha q, z = D^-1 r; the norms of r and r.z
        // summed in the same pass
        const int64_t nb = nblocks(n);
        std::vector<double> rr_part(nb), rz_part(nb);
#pragma omp parallel for schedule(static) num_threads(threads)
        for (int64_t blk = 0; blk < nb; ++blk) {
            const int64_t lo = blk * kBlock, hi = std::min(n, lo + kBlock);
            double s_rr = 0.0, s_rz = 0.0;
            for (int64_t i = lo; i < hi; ++i) {
                x[i] += alpha * p[i];
                r[i] -= alpha * q[i];
                z[i] = dinv[i] * r[i];
                s_rr += r[i] * r[i];
                s_rz += r[i] * z[i];
            }
            rr_part[blk] = s_rr;
            rz_part[blk] = s_rz;
        }
        double rr = 0.0, rz_new = 0.0;
        for (int64_t blk = 0; blk < nb; ++blk) {
            rr += rr_part[blk];
            rz_new += rz_part[blk];
        }
        if (rr <= stop) return it;
        const double beta = rz_new / rz;
        rz = rz_new;
#pragma omp parallel for schedule(static) num_threads(threads)
        for (int64_t i = 0; i < n; ++i) p[i] = z[i] + beta * p[i];
    }
    return -1;
}

}  // extern "C"
