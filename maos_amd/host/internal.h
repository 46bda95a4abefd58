/*
 * Helpers shared by the setup library's source files; not part of its C ABI
 * (include/maos_setup.h).
 */
#ifndef MXS_INTERNAL_H
#define MXS_INTERNAL_H
/* Symmetric eigen-decomposition by cyclic Jacobi: A (n x n, destroyed) =
 * V diag(d) V^T, V column-major. */
void mxs_sym_eig(double *A, int n, double *d, double *V);
#endif
