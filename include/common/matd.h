/* matd_t: the dense double matrix of the upstream API (common/matd.h:45-49
 * in the reference), as apriltag_detection_t.H and the pose results use
 * it, with the reference's element access and algebra (common/matd.h:
 * 72-441; not its SVD, Cholesky, LU objects or expression evaluator). A
 * matrix with no rows or no columns is a scalar of one element. */
#ifndef ATG_COMMON_MATD_H
#define ATG_COMMON_MATD_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    unsigned int nrows, ncols;
    double *data;
} matd_t;

#define MATD_EL(m, row, col) (m)->data[((row) * (m)->ncols + (col))]
#define MATD_EPS 1e-8

matd_t *matd_create(int rows, int cols);                               /* zeros */
matd_t *matd_create_data(int rows, int cols, const double *data);
matd_t *matd_create_dataf(int rows, int cols, const float *data);
matd_t *matd_create_scalar(double v);
matd_t *matd_copy(const matd_t *m);
matd_t *matd_identity(int dim);
void matd_destroy(matd_t *m);

double matd_get(const matd_t *m, unsigned int row, unsigned int col);
void matd_put(matd_t *m, unsigned int row, unsigned int col, double value);
double matd_get_scalar(const matd_t *m);
void matd_put_scalar(matd_t *m, double value);
matd_t *matd_select(const matd_t *a, unsigned int r0, int r1, unsigned int c0, int c1);  /* inclusive */
void matd_print(const matd_t *m, const char *fmt);
void matd_print_transpose(const matd_t *m, const char *fmt);

matd_t *matd_add(const matd_t *a, const matd_t *b);
void matd_add_inplace(matd_t *a, const matd_t *b);
matd_t *matd_subtract(const matd_t *a, const matd_t *b);
void matd_subtract_inplace(matd_t *a, const matd_t *b);
matd_t *matd_scale(const matd_t *a, double s);
void matd_scale_inplace(matd_t *a, double s);
matd_t *matd_multiply(const matd_t *a, const matd_t *b);  /* a scalar operand scales the other */
matd_t *matd_transpose(const matd_t *a);
double matd_det(const matd_t *a);
matd_t *matd_inverse(const matd_t *a);  /* NULL when exactly singular */

static inline int matd_is_scalar(const matd_t *a) { return a->ncols <= 1 && a->nrows <= 1; }
static inline int matd_is_vector(const matd_t *a) { return a->ncols == 1 || a->nrows == 1; }
static inline int matd_is_vector_len(const matd_t *a, int len) {
    return (a->ncols == 1 && a->nrows == (unsigned int)len) || (a->ncols == (unsigned int)len && a->nrows == 1);
}
double matd_vec_mag(const matd_t *a);
double matd_vec_dist(const matd_t *a, const matd_t *b);
double matd_vec_dist_n(const matd_t *a, const matd_t *b, int n);
double matd_vec_dot_product(const matd_t *a, const matd_t *b);
matd_t *matd_vec_normalize(const matd_t *a);
matd_t *matd_crossproduct(const matd_t *a, const matd_t *b);
double matd_err_inf(const matd_t *a, const matd_t *b);
double matd_max(matd_t *m);

#ifdef __cplusplus
}
#endif

#endif
