// The matd_t algebra of the upstream API (common/matd.h:72-441 of the
// reference) that callers of the pose functions use on their results:
// element access, arithmetic, products, transpose, determinant, inverse
// and the vector helpers. Host code: small dense matrices, row-major.
// A matrix with no rows or no columns is a scalar (one element), as in
// the reference (common/matd.h:38-41).
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "common/matd.h"

namespace {

bool is_scalar(const matd_t *m) { return m->nrows == 0 || m->ncols == 0; }
size_t count(const matd_t *m) { return is_scalar(m) ? 1 : (size_t)m->nrows * m->ncols; }

matd_t *same_shape(const matd_t *a) {
    return is_scalar(a) ? matd_create_scalar(0.0) : matd_create((int)a->nrows, (int)a->ncols);
}

// LU with partial pivoting of the n x n row-major a (in place); returns
// the permutation's sign, 0 when a pivot is exactly zero.
int lu_in_place(std::vector<double> &a, std::vector<int> &perm, int n) {
    int sign = 1;
    perm.resize((size_t)n);
    for (int i = 0; i < n; i++) perm[(size_t)i] = i;
    for (int k = 0; k < n; k++) {
        int piv = k;
        for (int r = k + 1; r < n; r++)
            if (std::fabs(a[(size_t)(r * n + k)]) > std::fabs(a[(size_t)(piv * n + k)])) piv = r;
        if (a[(size_t)(piv * n + k)] == 0.0) return 0;
        if (piv != k) {
            for (int c = 0; c < n; c++) std::swap(a[(size_t)(k * n + c)], a[(size_t)(piv * n + c)]);
            std::swap(perm[(size_t)k], perm[(size_t)piv]);
            sign = -sign;
        }
        for (int r = k + 1; r < n; r++) {
            const double f = a[(size_t)(r * n + k)] / a[(size_t)(k * n + k)];
            a[(size_t)(r * n + k)] = f;
            for (int c = k + 1; c < n; c++) a[(size_t)(r * n + c)] -= f * a[(size_t)(k * n + c)];
        }
    }
    return sign;
}

}  // namespace

extern "C" {

matd_t *matd_create_scalar(double v) {
    matd_t *m = (matd_t *)calloc(1, sizeof(matd_t));
    m->nrows = 0;
    m->ncols = 0;
    m->data = (double *)calloc(1, sizeof(double));
    m->data[0] = v;
    return m;
}

matd_t *matd_create_dataf(int rows, int cols, const float *data) {
    matd_t *m = matd_create(rows, cols);
    for (size_t i = 0; i < count(m); i++) m->data[i] = (double)data[i];
    return m;
}

double matd_get(const matd_t *m, unsigned int row, unsigned int col) { return MATD_EL(m, row, col); }
void matd_put(matd_t *m, unsigned int row, unsigned int col, double value) {
    if (is_scalar(m)) {
        m->data[0] = value;
        return;
    }
    MATD_EL(m, row, col) = value;
}
double matd_get_scalar(const matd_t *m) { return m->data[0]; }
void matd_put_scalar(matd_t *m, double value) { m->data[0] = value; }

// rows r0..r1 and columns c0..c1, inclusive
matd_t *matd_select(const matd_t *a, unsigned int r0, int r1, unsigned int c0, int c1) {
    if (r1 < (int)r0 || c1 < (int)c0) return matd_create(0, 0);
    matd_t *m = matd_create(r1 - (int)r0 + 1, c1 - (int)c0 + 1);
    for (int r = (int)r0; r <= r1; r++)
        for (int c = (int)c0; c <= c1; c++) MATD_EL(m, r - (int)r0, c - (int)c0) = MATD_EL(a, r, c);
    return m;
}

void matd_print(const matd_t *m, const char *fmt) {
    if (is_scalar(m)) {
        std::printf(fmt, m->data[0]);
        std::printf("\n");
        return;
    }
    for (unsigned r = 0; r < m->nrows; r++) {
        for (unsigned c = 0; c < m->ncols; c++) std::printf(fmt, MATD_EL(m, r, c));
        std::printf("\n");
    }
}

void matd_print_transpose(const matd_t *m, const char *fmt) {
    if (is_scalar(m)) {
        matd_print(m, fmt);
        return;
    }
    for (unsigned c = 0; c < m->ncols; c++) {
        for (unsigned r = 0; r < m->nrows; r++) std::printf(fmt, MATD_EL(m, r, c));
        std::printf("\n");
    }
}

matd_t *matd_add(const matd_t *a, const matd_t *b) {
    matd_t *m = same_shape(a);
    for (size_t i = 0; i < count(a); i++) m->data[i] = a->data[i] + b->data[i];
    return m;
}
void matd_add_inplace(matd_t *a, const matd_t *b) {
    for (size_t i = 0; i < count(a); i++) a->data[i] += b->data[i];
}
matd_t *matd_subtract(const matd_t *a, const matd_t *b) {
    matd_t *m = same_shape(a);
    for (size_t i = 0; i < count(a); i++) m->data[i] = a->data[i] - b->data[i];
    return m;
}
void matd_subtract_inplace(matd_t *a, const matd_t *b) {
    for (size_t i = 0; i < count(a); i++) a->data[i] -= b->data[i];
}
matd_t *matd_scale(const matd_t *a, double s) {
    matd_t *m = same_shape(a);
    for (size_t i = 0; i < count(a); i++) m->data[i] = a->data[i] * s;
    return m;
}
void matd_scale_inplace(matd_t *a, double s) {
    for (size_t i = 0; i < count(a); i++) a->data[i] *= s;
}

// a b; a scalar (or 1x1) operand scales the other
matd_t *matd_multiply(const matd_t *a, const matd_t *b) {
    if (matd_is_scalar(a)) return matd_scale(b, a->data[0]);
    if (matd_is_scalar(b)) return matd_scale(a, b->data[0]);
    if (a->ncols != b->nrows) return nullptr;
    matd_t *m = matd_create((int)a->nrows, (int)b->ncols);
    for (unsigned r = 0; r < a->nrows; r++)
        for (unsigned c = 0; c < b->ncols; c++) {
            double s = 0.0;
            for (unsigned k = 0; k < a->ncols; k++) s += MATD_EL(a, r, k) * MATD_EL(b, k, c);
            MATD_EL(m, r, c) = s;
        }
    return m;
}

matd_t *matd_transpose(const matd_t *a) {
    if (matd_is_scalar(a)) return matd_create_scalar(a->data[0]);
    matd_t *m = matd_create((int)a->ncols, (int)a->nrows);
    for (unsigned r = 0; r < a->nrows; r++)
        for (unsigned c = 0; c < a->ncols; c++) MATD_EL(m, c, r) = MATD_EL(a, r, c);
    return m;
}

double matd_det(const matd_t *a) {
    const int n = (int)a->nrows;
    if (n == 1) return a->data[0];
    if (n == 2) return a->data[0] * a->data[3] - a->data[1] * a->data[2];
    std::vector<double> lu(a->data, a->data + (size_t)n * (size_t)n);
    std::vector<int> perm;
    const int sign = lu_in_place(lu, perm, n);
    if (sign == 0) return 0.0;
    double d = sign;
    for (int i = 0; i < n; i++) d *= lu[(size_t)(i * n + i)];
    return d;
}

// NULL when the matrix is exactly singular (as the reference)
matd_t *matd_inverse(const matd_t *a) {
    if (matd_is_scalar(a)) return a->data[0] == 0.0 ? nullptr : matd_create_scalar(1.0 / a->data[0]);
    const int n = (int)a->nrows;
    if ((int)a->ncols != n) return nullptr;
    std::vector<double> lu(a->data, a->data + (size_t)n * (size_t)n);
    std::vector<int> perm;
    if (lu_in_place(lu, perm, n) == 0) return nullptr;
    matd_t *m = matd_create(n, n);
    std::vector<double> x((size_t)n);
    for (int col = 0; col < n; col++) {
        // solve L U x = P e_col
        for (int i = 0; i < n; i++) {
            double s = perm[(size_t)i] == col ? 1.0 : 0.0;
            for (int k = 0; k < i; k++) s -= lu[(size_t)(i * n + k)] * x[(size_t)k];
            x[(size_t)i] = s;
        }
        for (int i = n - 1; i >= 0; i--) {
            double s = x[(size_t)i];
            for (int k = i + 1; k < n; k++) s -= lu[(size_t)(i * n + k)] * x[(size_t)k];
            x[(size_t)i] = s / lu[(size_t)(i * n + i)];
        }
        for (int i = 0; i < n; i++) MATD_EL(m, i, col) = x[(size_t)i];
    }
    return m;
}

double matd_vec_dot_product(const matd_t *a, const matd_t *b) {
    double s = 0.0;
    for (size_t i = 0; i < count(a); i++) s += a->data[i] * b->data[i];
    return s;
}
double matd_vec_mag(const matd_t *a) { return std::sqrt(matd_vec_dot_product(a, a)); }
double matd_vec_dist_n(const matd_t *a, const matd_t *b, int n) {
    double s = 0.0;
    for (int i = 0; i < n; i++) {
        const double d = a->data[i] - b->data[i];
        s += d * d;
    }
    return std::sqrt(s);
}
double matd_vec_dist(const matd_t *a, const matd_t *b) { return matd_vec_dist_n(a, b, (int)count(a)); }
matd_t *matd_vec_normalize(const matd_t *a) {
    const double mag = matd_vec_mag(a);
    matd_t *m = same_shape(a);
    for (size_t i = 0; i < count(a); i++) m->data[i] = a->data[i] / mag;
    return m;
}
matd_t *matd_crossproduct(const matd_t *a, const matd_t *b) {
    matd_t *m = same_shape(a);
    const double *x = a->data, *y = b->data;
    m->data[0] = x[1] * y[2] - x[2] * y[1];
    m->data[1] = x[2] * y[0] - x[0] * y[2];
    m->data[2] = x[0] * y[1] - x[1] * y[0];
    return m;
}
// the largest absolute elementwise difference
double matd_err_inf(const matd_t *a, const matd_t *b) {
    double e = 0.0;
    for (size_t i = 0; i < count(a); i++) e = std::fmax(e, std::fabs(a->data[i] - b->data[i]));
    return e;
}
double matd_max(matd_t *m) {
    double v = -HUGE_VAL;
    for (size_t i = 0; i < count(m); i++) v = std::fmax(v, m->data[i]);
    return v;
}

}  // extern "C"
