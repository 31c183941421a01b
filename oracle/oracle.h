/*
 * oracle.h -- CPU restatement of depSolver's BEM hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: tests/, the smoke
 * entry point and bench.py's cpu_baseline leg are the only callers.  The
 * product (libdepsolver_amd.so) never links, loads or calls it.
 *
 * Every function restates one reference function with the same loop order
 * and the same arithmetic order, so that with the reference's compiler flags
 * (gcc -O3 -msse2 -mfpmath=sse, no FMA contraction) results are bitwise equal
 * to the reference build.  Conventions differ from the reference only in
 * plumbing: flat row-major arrays instead of row-pointer matrices, 0-based
 * connectivity, explicit sizes instead of globals, error codes instead of
 * exit(1).
 *
 * Layouts
 *   xyz   : N x 3 node coordinates
 *   conn  : E x npe 0-based node ids (npe = 3 for tria3, 6 for tria6)
 *   bctype: 2N bc types (Re rows then Im rows), as read from bcs.bem
 *   bc    : 2N x 3 bc values (bc[i*3+1] holds the 1-based interface id of an
 *           interface row)
 *   ifaces: I x 2 1-based material ids;  mats: M x 2 (sigma, eps_r)
 *   sigma : 2N solution (Re block then Im block), as the reference's vB
 */
#ifndef DEPSOLVER_ORACLE_H
#define DEPSOLVER_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* element library (shapeFunctions.c, globalToLocal.c, normals.c) */
void orc_shape_tria3(const double *L, double *N);
void orc_shape_tria6(const double *L, double *N);
double orc_x2l_tria3(const double *X, double *XL, const double *L, const double *N);
double orc_x2l_tria6(const double *X, double *XL, const double *L, const double *N);
void orc_local_normal_tria3(const double *X, double *n);
void orc_local_normal_tria6(const double *L, const double *X, double *n);
void orc_node_normal(int npe, int node, const double *xyz, int E, const int *conn, double *n);

/* quadrature kernels (kernelIntegrals.c).  X is npe x 3 and is not modified. */
void orc_intG(int npe, const double *X, const double *xe, double *out);
void orc_intH(int npe, int flag, const double *X, const double *xe, double *out, const double *n);
void orc_intSingularG(int npe, int sin_node, const double *X, const double *xe, double *out);
void orc_intSingularH(int npe, int flag, int sin_node, const double *X, const double *xe,
                      double *out, const double *n);

/* interface weights A_k, B_k (assembly_el.c:496-512) */
void orc_iface_weights(int nI, const unsigned *ifaces, const double *mats, double freq,
                       double *A, double *B);

/* Assembly (assembly_el.c).  full=1: 2N x 2N LU layout; full=0: N x 2N GMRES
 * layout.  Only rows [row0,row1) are assembled (row1<0: all); mA must be
 * zeroed by the caller (the reference allocates it with calloc). */
int orc_assemble(int npe, int full, int N, const double *xyz, int E, const int *conn,
                 const unsigned *bctype, const double *bc, int nI, const unsigned *ifaces,
                 const double *mats, double freq, double *mA, double *vB, int row0, int row1);

/* Gaussian elimination + back-substitution (gaussBksb.c).  A is n x n
 * row-major and is destroyed; B is the RHS on entry and the solution on exit.
 * max_pivots >= 0 stops the elimination after that many pivots (timing
 * samples only; B is then not a solution).  Returns 0, or -1 on singularity. */
int orc_gauss_bksb(int n, double *A, double *B, int max_pivots);

/* GMRES on the N x 2N layout (gmres_el.c).  tol_rel replaces the reference's
 * hard-coded 1e-6.  x0 (2N) may be NULL (zero initial guess).  The residual
 * history is written to log_path in the reference's gmres.log format when
 * log_path is not NULL.  Returns 0 on convergence, 1 on no convergence,
 * 2 on stagnation; *iters gets the iteration count. */
int orc_gmres_el(int N, int precond, const double *x0, double *A, double *B, double tol_rel,
                 int max_iters, const char *log_path, int *iters);

/* Post-processing at one point (potential.c, field.c) */
void orc_potential(int npe, int N, const double *xyz, int E, const int *conn,
                   const double *sigma, const double *x, double *pot);
void orc_field(int npe, int N, const double *xyz, int E, const int *conn,
               const double *sigma, const double *x, double *ef);

/* Maxwell-stress-tensor DEP force (force.c forceMST_*, forceIntegrals.c intF_*).
 * Fce, Xce: E x 3 (rows of non bc-6 elements are left untouched).
 * max_field_points >= 0 limits the number of bc-6 nodes whose field is
 * evaluated (timing samples only). */
int orc_force_mst(int npe, int N, const double *xyz, int E, const int *conn,
                  const unsigned *bctype, const double *sigma, double eps,
                  double *Fce, double *Xce, double *Fcm, double *Xcm, int max_field_points);

/* Multipolar DEP force on a sphere at x (force.c forceMultipole_*,
 * forceIntegrals.c intDE_*), orders 1..3. */
void orc_force_multipole(int npe, int order, double R, const double *x, int N,
                         const double *xyz, int E, const int *conn, const double *sigma,
                         const double *mats, double freq, double *F);

/* Depolarisation factors of an ellipsoid of semi-axes axis[3]
 * (depolarization.c:44-115, the 64-point rule on the truncated integral). */
int orc_depolarization(const double *axis, double *Ld);

/* Dipolar DEP force on an ellipsoid at each of m points (force.c
 * forceEllipsoid_*), analysis 10 (homogeneous) or 11 (single-shell oblate,
 * which scales the particle row of mats by C per point, in place). */
void orc_force_ellipsoid(int npe, int analysis, const double *axis, int m, const double *pts,
                         int N, const double *xyz, int E, const int *conn, const double *sigma,
                         double *mats, double freq, double *F);

#ifdef __cplusplus
}
#endif
#endif
