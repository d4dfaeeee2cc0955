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

// Row operations touch only the columns a row can hold non-zeros in:
// last[r] is the last such column of row r (its band, widened by the fill
// of row swaps and updates).  Entries past it are exact zeros, so the
// skipped operations (a - l * 0, s - 0 * x) would leave every value as it
// is.  Without row swaps (the diagonally dominant FEM and epsilon systems)
// each update covers ku columns instead of kl + ku.
int lu_solve_band(const BandMatrix& A, int bc, const double* b, double* x, BandMatrix& W, std::vector<double>& y) {
    const int n = A.n, kl = A.kl, ku = A.ku;
    const int m = n - bc;
    if (W.n != m || W.kl != kl || W.ku != ku) W.resize(m, kl, ku);
    // row(r)[c] is entry (r, c) for c in [r - kl, r + kl + ku]
    const int w = W.w;
    double* const base = W.data.data();
    auto row = [&](int r) { return base + (ptrdiff_t)r * w + kl - r; };
    const double* const abase = A.data.data();
    auto arow = [&](int r) { return abase + (ptrdiff_t)r * A.w + kl - r; };
    y.resize(2 * (size_t)m);
    double* const yv = y.data();
    int* const last = reinterpret_cast<int*>(yv + m);  // m ints in the second half
    const int kext = kl + ku;
    // copy the trailing block; the fill columns past the original upper band start at zero
    for (int r = 0; r < m; r++) {
        const int lo = r - kl < 0 ? 0 : r - kl;
        const int hi = r + ku > m - 1 ? m - 1 : r + ku;
        const int top = r + kext > m - 1 ? m - 1 : r + kext;
        double* wr = row(r);
        const double* ar = arow(r + bc) + bc;
        for (int c = lo; c <= hi; c++) wr[c] = ar[c];
        for (int c = hi + 1; c <= top; c++) wr[c] = 0.0;
        last[r] = hi;
        yv[r] = b[r + bc];
    }
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
        double* rk = row(k);
        if (p != k) {
            double* rp = row(p);
            const int jl = last[k] > last[p] ? last[k] : last[p];
            for (int j = k; j <= jl; j++) {
                const double t = rk[j];
                rk[j] = rp[j];
                rp[j] = t;
            }
            const int tl = last[k];
            last[k] = last[p];
            last[p] = tl;
            const double t = yv[k];
            yv[k] = yv[p];
            yv[p] = t;
        }
        const double piv = rk[k];
        const int jlast = last[k];
        for (int i = k + 1; i <= ilast; i++) {
            double* ri = row(i);
            const double l = ri[k] / piv;
            ri[k] = l;
            axpy_row(ri, rk, l, k + 1, jlast);
            if (last[i] < jlast) last[i] = jlast;
            yv[i] = yv[i] - l * yv[k];
        }
    }
    for (int i = m - 1; i >= 0; i--) {
        const double* ri = row(i);
        double s = yv[i];
        const int jlast = last[i];
        for (int j = i + 1; j <= jlast; j++) s = s - ri[j] * x[j + bc];
        x[i + bc] = s / ri[i];
    }
    for (int i = 0; i < bc; i++) x[i] = 0.0;
    return info;
}

}  // namespace life
