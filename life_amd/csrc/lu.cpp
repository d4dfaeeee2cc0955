#include "lu.h"

#include <cmath>
#include <cstring>

namespace life {

int lu_solve_sparse(int n, int bc, const double* A, const double* b, double* x, std::vector<double>& work,
                    std::vector<int>& iwork) {
    const int m = n - bc;
    work.resize((size_t)m * m + m);
    iwork.resize(m);
    double* B = work.data();
    double* y = B + (size_t)m * m;
    for (int i = 0; i < m; i++) {
        std::memcpy(B + (size_t)i * m, A + (size_t)(i + bc) * n + bc, sizeof(double) * m);
        y[i] = b[i + bc];
    }
    int info = 0;
    int* cols = iwork.data();
    for (int k = 0; k < m; k++) {
        int p = k;
        double best = std::fabs(B[(size_t)k * m + k]);
        for (int i = k + 1; i < m; i++) {
            const double a = std::fabs(B[(size_t)i * m + k]);
            if (a > best) {
                best = a;
                p = i;
            }
        }
        if (best == 0.0) {
            if (!info) info = k + 1;
            continue;
        }
        if (p != k) {
            double* rk = B + (size_t)k * m;
            double* rp = B + (size_t)p * m;
            for (int j = 0; j < m; j++) {
                const double t = rk[j];
                rk[j] = rp[j];
                rp[j] = t;
            }
            const double t = y[k];
            y[k] = y[p];
            y[p] = t;
        }
        const double* rk = B + (size_t)k * m;
        const double piv = rk[k];
        int nc = 0;
        for (int j = k + 1; j < m; j++)
            if (rk[j] != 0.0) cols[nc++] = j;
        for (int i = k + 1; i < m; i++) {
            double* ri = B + (size_t)i * m;
            if (ri[k] == 0.0) continue;
            const double l = ri[k] / piv;
            ri[k] = l;
            for (int q = 0; q < nc; q++) {
                const int j = cols[q];
                ri[j] = ri[j] - l * rk[j];
            }
            y[i] = y[i] - l * y[k];
        }
    }
    for (int i = m - 1; i >= 0; i--) {
        const double* ri = B + (size_t)i * m;
        double s = y[i];
        for (int j = i + 1; j < m; j++)
            if (ri[j] != 0.0) s = s - ri[j] * x[j + bc];
        x[i + bc] = s / ri[i];
    }
    for (int i = 0; i < bc; i++) x[i] = 0.0;
    return info;
}

void BandMatrix::resize(int n_, int kl_, int ku_) {
    n = n_;
    kl = kl_;
    ku = ku_;
    w = 2 * kl + ku + 1;
    data.assign((size_t)n * w, 0.0);
}

void BandMatrix::zero() { std::memset(data.data(), 0, sizeof(double) * data.size()); }

namespace {
// row_i[j] -= l * row_k[j] for j in [lo, hi] (rows k != i: no aliasing)
inline void axpy_row(double* __restrict ri, const double* __restrict rk, double l, int lo, int hi) {
    for (int j = lo; j <= hi; j++) ri[j] = ri[j] - l * rk[j];
}
}  // namespace

int lu_solve_band(const BandMatrix& A, int bc, const double* b, double* x, BandMatrix& W, std::vector<double>& y) {
    const int n = A.n, kl = A.kl, ku = A.ku;
    const int m = n - bc;
    if (W.n != m || W.kl != kl || W.ku != ku) W.resize(m, kl, ku);
    W.zero();
    // row(r)[c] is entry (r, c) for c in [r - kl, r + kl + ku]
    const int w = W.w;
    double* const base = W.data.data();
    auto row = [&](int r) { return base + (ptrdiff_t)r * w + kl - r; };
    const double* const abase = A.data.data();
    auto arow = [&](int r) { return abase + (ptrdiff_t)r * A.w + kl - r; };
    // copy the trailing block; entries beyond the original upper band start at zero
    for (int r = 0; r < m; r++) {
        const int lo = r - kl < 0 ? 0 : r - kl;
        const int hi = r + ku > m - 1 ? m - 1 : r + ku;
        double* wr = row(r);
        const double* ar = arow(r + bc) + bc;
        for (int c = lo; c <= hi; c++) wr[c] = ar[c];
    }
    y.resize(m);
    double* const yv = y.data();
    for (int r = 0; r < m; r++) yv[r] = b[r + bc];
    const int kext = kl + ku;
    int info = 0;
    for (int k = 0; k < m; k++) {
        const int ilast = k + kl > m - 1 ? m - 1 : k + kl;
        int p = k;
        double best = std::fabs(row(k)[k]);
        for (int i = k + 1; i <= ilast; i++) {
            const double a = std::fabs(row(i)[k]);
            if (a > best) {
                best = a;
                p = i;
            }
        }
        if (best == 0.0) {
            if (!info) info = k + 1;
            continue;
        }
        const int jlast = k + kext > m - 1 ? m - 1 : k + kext;
        double* rk = row(k);
        if (p != k) {
            double* rp = row(p);
            for (int j = k; j <= jlast; j++) {
                const double t = rk[j];
                rk[j] = rp[j];
                rp[j] = t;
            }
            const double t = yv[k];
            yv[k] = yv[p];
            yv[p] = t;
        }
        const double piv = rk[k];
        for (int i = k + 1; i <= ilast; i++) {
            double* ri = row(i);
            const double l = ri[k] / piv;
            ri[k] = l;
            axpy_row(ri, rk, l, k + 1, jlast);
            yv[i] = yv[i] - l * yv[k];
        }
    }
    for (int i = m - 1; i >= 0; i--) {
        const double* ri = row(i);
        double s = yv[i];
        const int jlast = i + kext > m - 1 ? m - 1 : i + kext;
        for (int j = i + 1; j <= jlast; j++) s = s - ri[j] * x[j + bc];
        x[i + bc] = s / ri[i];
    }
    for (int i = 0; i < bc; i++) x[i] = 0.0;
    return info;
}

}  // namespace life
