// ORACLE _ref driver — test infrastructure only.
//
// Compiles the reference's own header-only linear-algebra templates
// (linearalgebra/slap.h: Matrix, vec, pcg, flatten, unflatten,
// diagonalPreconditioner) from the reference tree, so that the oracle's
// matrix-free PCG restatement can be checked bit for bit against the
// reference's dense pcg().  The dense 5-point A and the boundary-zero RHS are
// assembled the way FieldSolve::pcgsolver does it (field.cpp:212-264).
// The rest of the reference's physics sources include domain.h -> output.h
// -> H5Cpp.h, which this image does not have, so they are not built.
#include "slap.h"

extern "C" {

// rhs: (nx*ny) doubles already in pcgsolver's rhom form, (i, j) -> i*ny + j.
// Writes phi (same layout) and the residual norm ||A x - b|| / n.
int ref_pcg(int nx, int ny, const double* rhs, double* phi, int max_iteration, double tolerance,
            double* error) {
  int n = nx * ny;
  Matrix<double> A(n, n), rhom(nx, ny);
  for (int i = 0; i < n; ++i) {
    int row = i / nx;
    int col = i % nx;
    A(i, i) = -4;
    if (col > 0) A(i, i - 1) = 1;
    if (col < nx - 1) A(i, i + 1) = 1;
    if (row > 0) A(i, i - nx) = 1;
    if (row < ny - 1) A(i, i + nx) = 1;
  }
  for (int i = 0; i < nx; i++)
    for (int j = 0; j < ny; j++) rhom(i, j) = rhs[i * ny + j];
  vec<double> rhovec = flatten(rhom);
  vec<double> x(n);
  x = pcg(A, x, rhovec, max_iteration, tolerance);
  vec<double> err = A * x - rhovec;
  if (error) *error = err.norm() / n;
  Matrix<double> xarr = unflatten(x, nx, ny);
  for (int i = 0; i < nx; i++)
    for (int j = 0; j < ny; j++) phi[i * ny + j] = xarr(i, j);
  return 0;
}

// FieldSolve::nrpcgsolver (field.cpp:141-210) on the reference's dense
// Matrix, vec and cg(): ten Newton steps, each cg(J, 0, F, 10000, 1e-4) with
// J = A - diag(exp(x)); rhs is rho itself (the driver forms rhom the way the
// solver does).  s = L^2/LDe^2, dx2 = dx^2.
int ref_nrpcg(int nx, int ny, const double* rho, double dx2, double s, double* phi) {
  int n = nx * ny;
  Matrix<double> A(n, n), rhom(nx, ny);
  for (int i = 0; i < n; ++i) {
    int row = i / nx;
    int col = i % nx;
    A(i, i) = -4;
    if (col > 0) A(i, i - 1) = 1;
    if (col < nx - 1) A(i, i + 1) = 1;
    if (row > 0) A(i, i - nx) = 1;
    if (row < ny - 1) A(i, i + nx) = 1;
  }
  for (int i = 0; i < nx; i++)
    for (int j = 0; j < ny; j++) rhom(i, j) = -rho[i * ny + j] * dx2 * s;
  for (int i = 0; i < nx; i++) {
    rhom(i, 0) = 0;
    rhom(i, ny - 1) = 0;
  }
  for (int i = 0; i < ny; i++) {
    rhom(0, i) = 0;
    rhom(nx - 1, i) = 0;
  }
  vec<double> rhovec = flatten(rhom);
  vec<double> rhovecx(n);
  vec<double> x(n);
  for (int iter = 0; iter < 10; iter++) {
    for (int i = 0; i < n; i++) rhovecx(i) = exp(x(i)) * s;
    rhovec = rhovec + rhovecx;
    Matrix<double> J = A;
    vec<double> F = A * x - rhovec;
    for (int i = 0; i < n; i++) J(i, i) = J(i, i) - exp(x(i));
    vec<double> x_ini(n);
    vec<double> y = cg(J, x_ini, F, 10000, 1e-4);
    x = x - y;
  }
  Matrix<double> xarr = unflatten(x, nx, ny);
  for (int i = 0; i < nx; i++)
    for (int j = 0; j < ny; j++) phi[i * ny + j] = xarr(i, j);
  return 0;
}

}  // extern "C"
