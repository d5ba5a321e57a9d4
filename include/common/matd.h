/* matd_t: the dense double matrix of the upstream API (common/matd.h:45-49
 * in the reference), as apriltag_detection_t.H uses it. Only what that
 * needs: create, copy, identity, destroy and element access. */
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

matd_t *matd_create(int rows, int cols);                               /* zeros */
matd_t *matd_create_data(int rows, int cols, const double *data);
matd_t *matd_copy(const matd_t *m);
matd_t *matd_identity(int dim);
void matd_destroy(matd_t *m);

#ifdef __cplusplus
}
#endif

#endif
