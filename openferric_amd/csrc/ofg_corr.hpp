// Correlation matrices on the host: validation, Higham's nearest-correlation
// repair and the PSD Cholesky factor (src/math/correlation.rs:220-389), shared
// by the basket and structured-product entry points.
#pragma once

#include <algorithm>
#include <cmath>
#include <vector>

#include "ofg_host.hpp"

namespace ofg_corr {

using ofg_host::fail;

constexpr double PSD_TOL = 1.0e-10;    // PsdProjectionConfig::default (correlation.rs:26-33)
constexpr int PSD_MAX_ITER = 100;

using Mat = std::vector<double>;   // row-major n x n

// ---------------------------------------------------------------------------
// Symmetric eigen-decomposition by cyclic Jacobi rotations (n <= 64). The
// reference uses nalgebra's SymmetricEigen; both converge to the exact
// decomposition, so they differ only by rounding.
// ---------------------------------------------------------------------------
inline void jacobi_eigen(const Mat &a_in, int n, std::vector<double> &w, Mat &v) {
    Mat a = a_in;
    v.assign((size_t)n * n, 0.0);
    for (int i = 0; i < n; ++i) v[(size_t)i * n + i] = 1.0;
    for (int sweep = 0; sweep < 100; ++sweep) {
        double off = 0.0, diag = 0.0;
        for (int i = 0; i < n; ++i) {
            diag += a[(size_t)i * n + i] * a[(size_t)i * n + i];
            for (int j = i + 1; j < n; ++j) off += a[(size_t)i * n + j] * a[(size_t)i * n + j];
        }
        if (off == 0.0 || off <= 1e-36 * diag) break;
        for (int p = 0; p < n; ++p)
            for (int q = p + 1; q < n; ++q) {
                const double apq = a[(size_t)p * n + q];
                if (apq == 0.0) continue;
                const double app = a[(size_t)p * n + p], aqq = a[(size_t)q * n + q];
                const double theta = (aqq - app) / (2.0 * apq);
                const double t = (theta >= 0.0 ? 1.0 : -1.0) /
                                 (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
                const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
                for (int k = 0; k < n; ++k) {   // columns p, q
                    const double akp = a[(size_t)k * n + p], akq = a[(size_t)k * n + q];
                    a[(size_t)k * n + p] = c * akp - s * akq;
                    a[(size_t)k * n + q] = s * akp + c * akq;
                }
                for (int k = 0; k < n; ++k) {   // rows p, q
                    const double apk = a[(size_t)p * n + k], aqk = a[(size_t)q * n + k];
                    a[(size_t)p * n + k] = c * apk - s * aqk;
                    a[(size_t)q * n + k] = s * apk + c * aqk;
                }
                a[(size_t)p * n + q] = a[(size_t)q * n + p] = 0.0;
                for (int k = 0; k < n; ++k) {
                    const double vkp = v[(size_t)k * n + p], vkq = v[(size_t)k * n + q];
                    v[(size_t)k * n + p] = c * vkp - s * vkq;
                    v[(size_t)k * n + q] = s * vkp + c * vkq;
                }
            }
    }
    w.resize((size_t)n);
    for (int i = 0; i < n; ++i) w[(size_t)i] = a[(size_t)i * n + i];
}

inline Mat symmetrize(const Mat &m, int n) {
    Mat out((size_t)n * n);
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j)
            out[(size_t)i * n + j] = 0.5 * (m[(size_t)i * n + j] + m[(size_t)j * n + i]);
    return out;
}

inline double min_eigenvalue(const Mat &m, int n) {
    std::vector<double> w;
    Mat v;
    jacobi_eigen(m, n, w, v);
    return *std::min_element(w.begin(), w.end());
}

// project_psd (correlation.rs:608-617): V max(D, 0) V^T, symmetrized.
inline Mat project_psd(const Mat &m, int n) {
    std::vector<double> w;
    Mat v;
    jacobi_eigen(symmetrize(m, n), n, w, v);
    Mat out((size_t)n * n, 0.0);
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
            double s = 0.0;
            for (int k = 0; k < n; ++k)
                s += v[(size_t)i * n + k] * std::max(w[(size_t)k], 0.0) * v[(size_t)j * n + k];
            out[(size_t)i * n + j] = s;
        }
    return symmetrize(out, n);
}

// validate_correlation_matrix (correlation.rs:220-248); throws.
inline void validate_corr(const Mat &c, int n) {
    for (int i = 0; i < n; ++i) {
        const double di = c[(size_t)i * n + i];
        if (!std::isfinite(di) || std::fabs(di - 1.0) > 1.0e-10)
            fail(OFG_INVALID_INPUT, "correlation matrix diagonal must be 1");
        for (int j = 0; j < n; ++j) {
            const double rho = c[(size_t)i * n + j];
            if (!std::isfinite(rho) || !(rho >= -1.0 && rho <= 1.0))
                fail(OFG_INVALID_INPUT, "correlation entries must be finite and in [-1, 1]");
            if (std::fabs(rho - c[(size_t)j * n + i]) > 1.0e-10)
                fail(OFG_INVALID_INPUT, "correlation matrix must be symmetric");
        }
    }
}

// nearest_correlation_matrix_higham (correlation.rs:271-335).
inline Mat nearest_correlation(const Mat &m, int n) {
    Mat y = symmetrize(m, n);
    for (int i = 0; i < n; ++i) y[(size_t)i * n + i] = 1.0;
    Mat ds((size_t)n * n, 0.0);
    for (int it = 0; it < PSD_MAX_ITER; ++it) {
        Mat r((size_t)n * n);
        for (size_t k = 0; k < r.size(); ++k) r[k] = y[k] - ds[k];
        r = symmetrize(r, n);
        Mat x = project_psd(r, n);
        for (size_t k = 0; k < r.size(); ++k) ds[k] = x[k] - r[k];
        for (int i = 0; i < n; ++i) x[(size_t)i * n + i] = 1.0;
        x = symmetrize(x, n);
        double diff = 0.0;
        for (size_t k = 0; k < x.size(); ++k) diff += (x[k] - y[k]) * (x[k] - y[k]);
        y = x;
        if (std::sqrt(diff) < PSD_TOL) break;
    }
    y = project_psd(symmetrize(y, n), n);
    for (int i = 0; i < n; ++i) y[(size_t)i * n + i] = 1.0;
    y = symmetrize(y, n);
    for (int i = 0; i < n; ++i) {
        y[(size_t)i * n + i] = 1.0;
        for (int j = i + 1; j < n; ++j) {
            const double v = std::min(1.0, std::max(-1.0, y[(size_t)i * n + j]));
            y[(size_t)i * n + j] = y[(size_t)j * n + i] = v;
        }
    }
    return y;
}

// validate_or_repair_correlation_matrix (correlation.rs:340-358); throws.
inline Mat validate_or_repair(const Mat &c, int n, bool &repaired) {
    validate_corr(c, n);
    repaired = false;
    if (min_eigenvalue(c, n) >= -PSD_TOL) return c;
    Mat out = nearest_correlation(c, n);
    validate_corr(out, n);
    if (!(min_eigenvalue(out, n) >= -1.0e-8))
        fail(OFG_INVALID_INPUT, "nearest-correlation projection did not produce PSD output");
    repaired = true;
    return out;
}

// cholesky_lower_psd (correlation.rs:362-389) with tolerance 1e-12.
inline bool cholesky_psd(const Mat &a, int n, Mat &l) {
    const double tol = 1.0e-12;
    l.assign((size_t)n * n, 0.0);
    for (int i = 0; i < n; ++i)
        for (int j = 0; j <= i; ++j) {
            double sum = a[(size_t)i * n + j];
            for (int k = 0; k < j; ++k) sum -= l[(size_t)i * n + k] * l[(size_t)j * n + k];
            if (i == j) {
                if (sum < -tol) return false;
                l[(size_t)i * n + j] = std::sqrt(std::fmax(sum, tol));
            } else if (l[(size_t)j * n + j] > tol) {
                l[(size_t)i * n + j] = sum / l[(size_t)j * n + j];
            }
        }
    return true;
}

}  // namespace ofg_corr
