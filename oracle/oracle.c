/*
 * oracle.c -- CPU restatement of depSolver's BEM hot path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP path (see oracle.h).
 * Each function cites the reference file:line it restates.  Loop order and
 * the order of every floating-point expression follow the reference, so that
 * compiled with the reference's flags the results are bitwise equal to the
 * reference build (checked in tests/test_oracle_vs_ref.py).
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"
#include "orc_tables.h"

/* ------------------------------------------------------------------------ */
/* Element library                                                          */
/* ------------------------------------------------------------------------ */

/* src/shapeFunctions.c:53-62 */
void orc_shape_tria3(const double *L, double *N)
{
    N[0] = L[0];
    N[1] = L[1];
    N[2] = 1.0 - L[0] - L[1];
}

/* src/shapeFunctions.c:68-87 (node order corner, mid, corner, mid, corner, mid) */
void orc_shape_tria6(const double *L, double *N)
{
    const double a = L[0], b = L[1], c = 1.0 - a - b;
    N[0] = a * (2.0 * a - 1.0);
    N[1] = 4.0 * a * b;
    N[2] = b * (2.0 * b - 1.0);
    N[3] = 4.0 * b * c;
    N[4] = c * (2.0 * c - 1.0);
    N[5] = 4.0 * a * c;
}

#define XC(X, k, c) ((X)[(k) * 3 + (c)])

/* src/globalToLocal.c:75-101: x(L) and |dx/dL1 x dx/dL2| for a flat triangle */
double orc_x2l_tria3(const double *X, double *XL, const double *L, const double *N)
{
    (void)L;
    for (int c = 0; c < 3; c++)
        XL[c] = N[0] * XC(X, 0, c) + N[1] * XC(X, 1, c) + N[2] * XC(X, 2, c);
    const double ux = XC(X, 0, 0) - XC(X, 2, 0), vx = XC(X, 1, 0) - XC(X, 2, 0);
    const double uy = XC(X, 0, 1) - XC(X, 2, 1), vy = XC(X, 1, 1) - XC(X, 2, 1);
    const double uz = XC(X, 0, 2) - XC(X, 2, 2), vz = XC(X, 1, 2) - XC(X, 2, 2);
    const double j1 = uy * vz - vy * uz;
    const double j2 = vx * uz - ux * vz;
    const double j3 = ux * vy - vx * uy;
    return sqrt(j1 * j1 + j2 * j2 + j3 * j3);
}

/* Tangent vectors of the quadratic map at L (globalToLocal.c:125-140,
 * normals.c:91-102 use the same closed form). */
static void tria6_tangents(const double *X, double a, double b, double *t1, double *t2)
{
    for (int c = 0; c < 3; c++) {
        const double s = XC(X, 1, c) - XC(X, 3, c) + XC(X, 4, c) - XC(X, 5, c);
        t1[c] = 4.0 * (a * (XC(X, 0, c) + XC(X, 4, c) - 2.0 * XC(X, 5, c)) + b * s + XC(X, 5, c))
                - XC(X, 0, c) - 3.0 * XC(X, 4, c);
        t2[c] = 4.0 * (a * s + b * (XC(X, 2, c) + XC(X, 4, c) - 2.0 * XC(X, 3, c)) + XC(X, 3, c))
                - XC(X, 2, c) - 3.0 * XC(X, 4, c);
    }
}

/* src/globalToLocal.c:115-148 */
double orc_x2l_tria6(const double *X, double *XL, const double *L, const double *N)
{
    for (int c = 0; c < 3; c++)
        XL[c] = N[0] * XC(X, 0, c) + N[1] * XC(X, 1, c) + N[2] * XC(X, 2, c)
                + N[3] * XC(X, 3, c) + N[4] * XC(X, 4, c) + N[5] * XC(X, 5, c);
    double t1[3], t2[3];
    tria6_tangents(X, L[0], L[1], t1, t2);
    const double j1 = t1[1] * t2[2] - t2[1] * t1[2];
    const double j2 = t2[0] * t1[2] - t1[0] * t2[2];
    const double j3 = t1[0] * t2[1] - t2[0] * t1[1];
    return sqrt(j1 * j1 + j2 * j2 + j3 * j3);
}

static void unit_cross(const double *t1, const double *t2, double *n)
{
    n[0] = t1[1] * t2[2] - t2[1] * t1[2];
    n[1] = t2[0] * t1[2] - t1[0] * t2[2];
    n[2] = t1[0] * t2[1] - t2[0] * t1[1];
    const double inv = 1.0 / sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    n[0] = n[0] * inv;
    n[1] = n[1] * inv;
    n[2] = n[2] * inv;
}

/* src/normals.c:45-70 */
void orc_local_normal_tria3(const double *X, double *n)
{
    double t1[3], t2[3];
    for (int c = 0; c < 3; c++) {
        t1[c] = XC(X, 0, c) - XC(X, 2, c);
        t2[c] = XC(X, 1, c) - XC(X, 2, c);
    }
    unit_cross(t1, t2, n);
}

/* src/normals.c:79-112 */
void orc_local_normal_tria6(const double *L, const double *X, double *n)
{
    double t1[3], t2[3];
    tria6_tangents(X, L[0], L[1], t1, t2);
    unit_cross(t1, t2, n);
}

/* Local coordinates of the six tria6 nodes (normals.c:244-267) */
static const double TRIA6_NODE_L[6][2] = {
    {1.0, 0.0}, {0.5, 0.5}, {0.0, 1.0}, {0.0, 0.5}, {0.0, 0.0}, {0.5, 0.0}};

static void gather_X(int npe, const double *xyz, const int *conn, int e, double *X)
{
    for (int k = 0; k < npe; k++) {
        const int v = conn[e * npe + k];
        X[k * 3 + 0] = xyz[v * 3 + 0];
        X[k * 3 + 1] = xyz[v * 3 + 1];
        X[k * 3 + 2] = xyz[v * 3 + 2];
    }
}

/* src/normals.c:120-199 (tria3) and :207-319 (tria6): unit normals of every
 * element containing the node, visited in element order (first matching
 * position in each element), summed and normalised. */
void orc_node_normal(int npe, int node, const double *xyz, int E, const int *conn, double *n)
{
    double acc[3] = {0.0, 0.0, 0.0};
    double X[18], loc[3];
    for (int e = 0; e < E; e++) {
        int pos = -1;
        for (int k = 0; k < npe; k++)
            if (conn[e * npe + k] == node) { pos = k; break; }
        if (pos < 0) continue;
        gather_X(npe, xyz, conn, e, X);
        if (npe == 3) orc_local_normal_tria3(X, loc);
        else orc_local_normal_tria6(TRIA6_NODE_L[pos], X, loc);
        acc[0] += loc[0];
        acc[1] += loc[1];
        acc[2] += loc[2];
    }
    const double inv = 1.0 / sqrt(acc[0] * acc[0] + acc[1] * acc[1] + acc[2] * acc[2]);
    n[0] = acc[0] * inv;
    n[1] = acc[1] * inv;
    n[2] = acc[2] * inv;
}

/* ------------------------------------------------------------------------ */
/* Quadrature kernels                                                       */
/* ------------------------------------------------------------------------ */

static double shape_and_map(int npe, const double *X, const double *L, double *N, double *XL)
{
    if (npe == 3) {
        orc_shape_tria3(L, N);
        return orc_x2l_tria3(X, XL, L, N);
    }
    orc_shape_tria6(L, N);
    return orc_x2l_tria6(X, XL, L, N);
}

/* src/kernelIntegrals.c:61-95 (tria3), :109-142 (tria6):
 * out[k] = sum_g W J N_k / |x - y_g| */
void orc_intG(int npe, const double *X, const double *xe, double *out)
{
    double L[2], N[6], XL[3];
    for (int k = 0; k < npe; k++) out[k] = 0.0;
    for (int g = 0; g < 7; g++) {
        L[0] = ORC_T7[g][0];
        L[1] = ORC_T7[g][1];
        const double W = ORC_T7[g][2];
        const double J = shape_and_map(npe, X, L, N, XL);
        const double dx = xe[0] - XL[0], dy = xe[1] - XL[1], dz = xe[2] - XL[2];
        const double A = W * J / sqrt(dx * dx + dy * dy + dz * dz);
        for (int k = 0; k < npe; k++) out[k] += A * N[k];
    }
}

/* src/kernelIntegrals.c:160-213 (tria3), :231-282 (tria6).
 * flag == 0: out[k] -= sum W J N_k (d.n)/r^3;  flag != 0: out[3k+c] += sum W J N_k d_c/r^3 */
void orc_intH(int npe, int flag, const double *X, const double *xe, double *out, const double *n)
{
    double L[2], N[6], XL[3];
    const int nout = flag ? 3 * npe : npe;
    for (int k = 0; k < nout; k++) out[k] = 0.0;
    for (int g = 0; g < 7; g++) {
        L[0] = ORC_T7[g][0];
        L[1] = ORC_T7[g][1];
        const double W = ORC_T7[g][2];
        const double J = shape_and_map(npe, X, L, N, XL);
        const double dx = xe[0] - XL[0], dy = xe[1] - XL[1], dz = xe[2] - XL[2];
        const double rsq = dx * dx + dy * dy + dz * dz;
        double A = J * W / pow(rsq, 1.5);
        if (!flag) {
            const double nd = dx * n[0] + dy * n[1] + dz * n[2];
            A = A * nd;
            for (int k = 0; k < npe; k++) out[k] -= A * N[k];
        } else {
            const double Cz = A * dz, By = A * dy, Ax = A * dx;
            for (int k = 0; k < npe; k++) {
                out[k * 3] += Ax * N[k];
                out[k * 3 + 1] += By * N[k];
                out[k * 3 + 2] += Cz * N[k];
            }
        }
    }
}

/* src/kernelIntegrals.c:299-350 (tria3), :367-454 (tria6): Duffy-collapsed
 * square with Gauss-Legendre x Gauss-Jacobi(0,1); a midside singular node
 * splits the element into two sub-triangles (weight factor 1/16). */
void orc_intSingularG(int npe, int sin_node, const double *X, const double *xe, double *out)
{
    double L1[2], L2[2], N1[6], N2[6], XL[3];
    for (int k = 0; k < npe; k++) out[k] = 0.0;
    int two = 0;
    double B = 0.125;
    if (npe == 6 && sin_node % 2 == 0) {
        two = 1;
        B = 0.0625;
    }
    for (int i = 0; i < 8; i++) {
        const double U1 = ORC_GL8[i][0], W1 = ORC_GL8[i][1];
        for (int j = 0; j < 8; j++) {
            const double U2 = ORC_GJ8[j][0], W2 = ORC_GJ8[j][1];
            if (npe == 3) {
                if (sin_node == 1) {
                    L1[0] = 0.5 * (1.0 - U2);
                    L1[1] = 0.25 * (1.0 + U1) * (1.0 + U2);
                } else if (sin_node == 2) {
                    L1[0] = 0.25 * (1.0 - U1) * (1.0 + U2);
                    L1[1] = 0.5 * (1.0 - U2);
                } else {
                    L1[0] = 0.25 * (1.0 + U1) * (1.0 + U2);
                    L1[1] = 0.25 * (1.0 - U1) * (1.0 + U2);
                }
                orc_shape_tria3(L1, N1);
                const double J = orc_x2l_tria3(X, XL, L1, N1);
                const double dx = xe[0] - XL[0], dy = xe[1] - XL[1], dz = xe[2] - XL[2];
                const double A = W1 * W2 * J * 0.125 / sqrt(dx * dx + dy * dy + dz * dz);
                for (int k = 0; k < 3; k++) out[k] += A * N1[k];
                continue;
            }
            switch (sin_node) {
            case 1:
                L1[0] = 0.5 * (1.0 - U2);
                L1[1] = 0.25 * (1.0 + U1) * (1.0 + U2);
                break;
            case 2:
                L1[0] = 0.25 * (1.0 - U2);
                L1[1] = 0.5 + 0.25 * U1 * (1.0 + U2);
                L2[0] = 0.5 - 0.25 * U1 * (1.0 + U2);
                L2[1] = 0.25 * (1.0 - U2);
                break;
            case 3:
                L1[0] = 0.25 * (1.0 - U1) * (1.0 + U2);
                L1[1] = 0.5 * (1.0 - U2);
                break;
            case 4:
                L1[0] = 0.25 * (1.0 + U1) * (1.0 + U2);
                L1[1] = 0.5 - 0.25 * U1 * (1.0 + U2);
                L2[0] = 0.25 * (1.0 - U1) * (1.0 + U2);
                L2[1] = 0.25 * (1.0 - U2);
                break;
            case 5:
                L1[0] = 0.25 * (1.0 + U1) * (1.0 + U2);
                L1[1] = 0.25 * (1.0 - U1) * (1.0 + U2);
                break;
            default:
                L1[0] = 0.5 + 0.25 * U1 * (1.0 + U2);
                L1[1] = 0.25 * (1.0 - U1) * (1.0 + U2);
                L2[0] = 0.25 * (1 - U2);
                L2[1] = 0.25 * (1.0 + U1) * (1.0 + U2);
                break;
            }
            orc_shape_tria6(L1, N1);
            double J = orc_x2l_tria6(X, XL, L1, N1);
            double dx = xe[0] - XL[0], dy = xe[1] - XL[1], dz = xe[2] - XL[2];
            double A = W1 * W2 * J * B / sqrt(dx * dx + dy * dy + dz * dz);
            for (int k = 0; k < 6; k++) out[k] += A * N1[k];
            if (two) {
                orc_shape_tria6(L2, N2);
                J = orc_x2l_tria6(X, XL, L2, N2);
                dx = xe[0] - XL[0];
                dy = xe[1] - XL[1];
                dz = xe[2] - XL[2];
                A = W1 * W2 * J * B / sqrt(dx * dx + dy * dy + dz * dz);
                for (int k = 0; k < 6; k++) out[k] += A * N2[k];
            }
        }
    }
}

/* Cyclic move a <- b <- c <- a, as the reference's shift macro (constants.h:58) */
static void rot3(double *a, double *b, double *c)
{
    const double d = *a;
    *a = *b;
    *b = *c;
    *c = d;
}

/* The three sub-triangles of the first subdivision level, in local (L1,L2)
 * coordinates; the singular corner is at L = (0,0) (kernelIntegrals.c:508-518). */
static const double SUBTRI0[3][3][2] = {
    {{1.0, 0.0}, {0.5, 0.5}, {0.5, 0.0}},
    {{0.5, 0.0}, {0.5, 0.5}, {0.0, 0.5}},
    {{0.5, 0.5}, {0.0, 1.0}, {0.0, 0.5}}};

static void accum_H(int npe, int flag, double A, double dx, double dy, double dz,
                    const double *N, const double *n, double *out)
{
    if (!flag) {
        const double nd = dx * n[0] + dy * n[1] + dz * n[2];
        A = A * nd;
        for (int k = 0; k < npe; k++) out[k] -= A * N[k];
    } else {
        const double Cz = A * dz, By = A * dy, Ax = A * dx;
        for (int k = 0; k < npe; k++) {
            out[k * 3] += Ax * N[k];
            out[k * 3 + 1] += By * N[k];
            out[k * 3 + 2] += Cz * N[k];
        }
    }
}

/* src/kernelIntegrals.c:472-593 (tria3), :611-805 (tria6): strongly singular
 * integral by NSUBDIVISIONS = 8 levels of 3 sub-triangles x 7 points.  A
 * corner singular node is rotated to the last corner (node 3 / node 5), a
 * midside one is handled by splitting every sub-triangle point in two.  The
 * z-component of the midside gradient (flag != 0) adds B2*N2 where C2*N2
 * belongs (kernelIntegrals.c:762); that is reproduced here. */
void orc_intSingularH(int npe, int flag, int sin_node, const double *Xin, const double *xe,
                      double *out, const double *n)
{
    double X[18];
    memcpy(X, Xin, sizeof(double) * 3 * npe);
    const int nout = flag ? 3 * npe : npe;
    for (int k = 0; k < nout; k++) out[k] = 0.0;

    if (npe == 3) {
        if (sin_node == 1)
            for (int c = 0; c < 3; c++) rot3(&X[2 * 3 + c], &X[0 * 3 + c], &X[1 * 3 + c]);
        else if (sin_node == 2)
            for (int c = 0; c < 3; c++) rot3(&X[2 * 3 + c], &X[1 * 3 + c], &X[0 * 3 + c]);
    } else {
        if (sin_node == 1) {
            for (int c = 0; c < 3; c++) {
                rot3(&X[4 * 3 + c], &X[0 * 3 + c], &X[2 * 3 + c]);
                rot3(&X[5 * 3 + c], &X[1 * 3 + c], &X[3 * 3 + c]);
            }
        } else if (sin_node == 3) {
            for (int c = 0; c < 3; c++) {
                rot3(&X[4 * 3 + c], &X[2 * 3 + c], &X[0 * 3 + c]);
                rot3(&X[3 * 3 + c], &X[1 * 3 + c], &X[5 * 3 + c]);
            }
        }
    }

    double xs[3][3][2];
    memcpy(xs, SUBTRI0, sizeof(xs));
    const int corner = (npe == 3) || (sin_node % 2 == 1);
    double J = 1.0;
    double L[2], L1[2], L2[2], N[6], N1[6], N2[6], XL[3], XL1[3], XL2[3];
    for (int m = 0; m < 8; m++) {
        J = J * 0.25;
        for (int i = 0; i < 3; i++) {
            for (int j = 0; j < 7; j++) {
                const double M1 = ORC_T7[j][0], M2 = ORC_T7[j][1];
                const double M3 = 1.0 - M1 - M2;
                const double W = ORC_T7[j][2];
                if (corner) {
                    L[0] = xs[i][0][0] * M1 + xs[i][1][0] * M2 + xs[i][2][0] * M3;
                    L[1] = xs[i][0][1] * M1 + xs[i][1][1] * M2 + xs[i][2][1] * M3;
                    const double JL = shape_and_map(npe, X, L, N, XL);
                    const double dx = xe[0] - XL[0], dy = xe[1] - XL[1], dz = xe[2] - XL[2];
                    double A;
                    if (npe == 3) {
                        const double rsq = dx * dx + dy * dy + dz * dz;
                        A = J * JL * W / pow(rsq, 1.5);
                    } else {
                        const double dr = sqrt(dx * dx + dy * dy + dz * dz);
                        A = J * JL * W / (dr * dr * dr);
                    }
                    accum_H(npe, flag, A, dx, dy, dz, N, n, out);
                } else {
                    double U1, U2, U3;
                    if (sin_node == 2) {
                        U1 = 0.5 * M3; U2 = M1 + 0.5 * M3; U3 = 1.0 - U1 - U2;
                        L1[0] = xs[i][0][0] * U1 + xs[i][1][0] * U2 + xs[i][2][0] * U3;
                        L1[1] = xs[i][0][1] * U1 + xs[i][1][1] * U2 + xs[i][2][1] * U3;
                        U1 = M1 + 0.5 * M3; U2 = M3; U3 = 1.0 - U1 - U2;
                        L2[0] = xs[i][0][0] * U1 + xs[i][1][0] * U2 + xs[i][2][0] * U3;
                        L2[1] = xs[i][0][1] * U1 + xs[i][1][1] * U2 + xs[i][2][1] * U3;
                    } else if (sin_node == 4) {
                        U1 = M1; U2 = 0.5 * M3; U3 = 1.0 - U1 - U2;
                        L1[0] = xs[i][0][0] * U1 + xs[i][1][0] * U2 + xs[i][2][0] * U3;
                        L1[1] = xs[i][0][1] * U1 + xs[i][1][1] * U2 + xs[i][2][1] * U3;
                        U1 = M1; U2 = M2 + 0.5 * M3; U3 = 1.0 - U1 - U2;
                        L2[0] = xs[i][0][0] * U1 + xs[i][1][0] * U2 + xs[i][2][0] * U3;
                        L2[1] = xs[i][0][1] * U1 + xs[i][1][1] * U2 + xs[i][2][1] * U3;
                    } else {
                        U1 = 0.5 * M3; U2 = M1; U3 = 1.0 - U1 - U2;
                        L1[0] = xs[i][0][0] * U1 + xs[i][1][0] * U2 + xs[i][2][0] * U3;
                        L1[1] = xs[i][0][1] * U1 + xs[i][1][1] * U2 + xs[i][2][1] * U3;
                        U1 = M1 + 0.5 * M3; U2 = M2; U3 = 1.0 - U1 - U2;
                        L2[0] = xs[i][0][0] * U1 + xs[i][1][0] * U2 + xs[i][2][0] * U3;
                        L2[1] = xs[i][0][1] * U1 + xs[i][1][1] * U2 + xs[i][2][1] * U3;
                    }
                    orc_shape_tria6(L1, N1);
                    orc_shape_tria6(L2, N2);
                    const double JL1 = orc_x2l_tria6(X, XL1, L1, N1);
                    const double JL2 = orc_x2l_tria6(X, XL2, L2, N2);
                    const double dx1 = xe[0] - XL1[0], dy1 = xe[1] - XL1[1], dz1 = xe[2] - XL1[2];
                    const double dx2 = xe[0] - XL2[0], dy2 = xe[1] - XL2[1], dz2 = xe[2] - XL2[2];
                    const double dr1 = sqrt(dx1 * dx1 + dy1 * dy1 + dz1 * dz1);
                    const double dr2 = sqrt(dx2 * dx2 + dy2 * dy2 + dz2 * dz2);
                    double A1 = 0.5 * J * JL1 * W / (dr1 * dr1 * dr1);
                    double A2 = 0.5 * J * JL2 * W / (dr2 * dr2 * dr2);
                    if (!flag) {
                        A1 = A1 * (dx1 * n[0] + dy1 * n[1] + dz1 * n[2]);
                        A2 = A2 * (dx2 * n[0] + dy2 * n[1] + dz2 * n[2]);
                        for (int k = 0; k < 6; k++) out[k] -= A1 * N1[k] + A2 * N2[k];
                    } else {
                        const double C1 = A1 * dz1, B1 = A1 * dy1, Ax1 = A1 * dx1;
                        const double B2 = A2 * dy2, Ax2 = A2 * dx2;
                        for (int k = 0; k < 6; k++) {
                            out[k * 3] += Ax1 * N1[k] + Ax2 * N2[k];
                            out[k * 3 + 1] += B1 * N1[k] + B2 * N2[k];
                            out[k * 3 + 2] += C1 * N1[k] + B2 * N2[k];
                        }
                    }
                }
            }
        }
        for (int i = 0; i < 3; i++)
            for (int v = 0; v < 3; v++) {
                xs[i][v][0] = xs[i][v][0] * 0.5;
                xs[i][v][1] = xs[i][v][1] * 0.5;
            }
    }

    /* put the node contributions back in element order */
    if (npe == 3) {
        if (!flag) {
            if (sin_node == 1) rot3(&out[1], &out[0], &out[2]);
            else if (sin_node == 2) rot3(&out[0], &out[1], &out[2]);
        } else {
            if (sin_node == 1)
                for (int c = 0; c < 3; c++) rot3(&out[3 + c], &out[c], &out[6 + c]);
            else if (sin_node == 2)
                for (int c = 0; c < 3; c++) rot3(&out[c], &out[3 + c], &out[6 + c]);
        }
    } else {
        if (!flag) {
            if (sin_node == 1) {
                rot3(&out[2], &out[0], &out[4]);
                rot3(&out[3], &out[1], &out[5]);
            } else if (sin_node == 3) {
                rot3(&out[0], &out[2], &out[4]);
                rot3(&out[5], &out[1], &out[3]);
            }
        } else {
            if (sin_node == 1) {
                for (int c = 0; c < 3; c++) {
                    rot3(&out[6 + c], &out[c], &out[12 + c]);
                    rot3(&out[9 + c], &out[3 + c], &out[15 + c]);
                }
            } else if (sin_node == 3) {
                for (int c = 0; c < 3; c++) {
                    rot3(&out[c], &out[6 + c], &out[12 + c]);
                    rot3(&out[15 + c], &out[3 + c], &out[9 + c]);
                }
            }
        }
    }
}

/* ------------------------------------------------------------------------ */
/* Assembly                                                                 */
/* ------------------------------------------------------------------------ */

/* src/assembly_el.c:496-512: per-interface complex weight K = A - iB scaled
 * by 1/(4 pi eps0), from (sigma, eps_r) of the two materials. */
void orc_iface_weights(int nI, const unsigned *ifaces, const double *mats, double freq,
                       double *A, double *B)
{
    const double C = 1.0 / (ORC_PI4 * ORC_EPS0);
    const double W = freq * ORC_PI2 * ORC_EPS0;
    for (int i = 0; i < nI; i++) {
        const int m1 = (int)ifaces[i * 2] - 1, m2 = (int)ifaces[i * 2 + 1] - 1;
        const double SigTot = mats[m1 * 2] + mats[m2 * 2];
        const double SigDif = mats[m1 * 2] - mats[m2 * 2];
        const double EpsTot = W * (mats[m1 * 2 + 1] + mats[m2 * 2 + 1]);
        const double EpsDif = W * (mats[m1 * 2 + 1] - mats[m2 * 2 + 1]);
        const double D = C / (SigTot * SigTot + EpsTot * EpsTot);
        A[i] = D * (SigTot * SigDif + EpsTot * EpsDif);
        B[i] = D * (SigDif * EpsTot - SigTot * EpsDif);
    }
}

/* src/assembly_el.c:470-589 (full, tria6), :329-448 (full, tria3),
 * :197-307 (el, tria6), :63-174 (el, tria3).  Row i loops over every element
 * j; the pair is singular when node i is one of j's nodes. */
int orc_assemble(int npe, int full, int N, const double *xyz, int E, const int *conn,
                 const unsigned *bctype, const double *bc, int nI, const unsigned *ifaces,
                 const double *mats, double freq, double *mA, double *vB, int row0, int row1)
{
    const double C = 1.0 / (ORC_PI4 * ORC_EPS0);
    const long ld = full ? 2L * N : 2L * N; /* both layouts have 2N columns */
    double *wA = NULL, *wB = NULL;
    if (nI > 0) {
        wA = (double *)malloc(sizeof(double) * nI);
        wB = (double *)malloc(sizeof(double) * nI);
        if (!wA || !wB) { free(wA); free(wB); return -1; }
        orc_iface_weights(nI, ifaces, mats, freq, wA, wB);
    }
    if (row1 < 0 || row1 > N) row1 = N;
    double X[18], sub[6], normal[3], xe[3];
    int iface = 0;
    for (int i = row0; i < row1; i++) {
        xe[0] = xyz[i * 3];
        xe[1] = xyz[i * 3 + 1];
        xe[2] = xyz[i * 3 + 2];
        const int dirichlet = (bctype[i] == 1);
        if (bctype[i] == 0 || bctype[i] == 6) {
            iface = (int)bc[i * 3 + 1] - 1;
            orc_node_normal(npe, i, xyz, E, conn, normal);
            mA[(long)i * ld + i] -= 0.5 / ORC_EPS0;
            if (full) mA[(long)(i + N) * ld + (i + N)] -= 0.5 / ORC_EPS0;
        }
        for (int j = 0; j < E; j++) {
            int sin_node = 0;
            for (int k = 0; k < npe; k++) {
                const int v = conn[j * npe + k];
                X[k * 3] = xyz[v * 3];
                X[k * 3 + 1] = xyz[v * 3 + 1];
                X[k * 3 + 2] = xyz[v * 3 + 2];
                if (v == i) sin_node = k + 1;
            }
            if (dirichlet) {
                if (sin_node == 0) orc_intG(npe, X, xe, sub);
                else orc_intSingularG(npe, sin_node, X, xe, sub);
                for (int k = 0; k < npe; k++) {
                    const int c = conn[j * npe + k];
                    if (full) {
                        const double F = C * sub[k];
                        mA[(long)i * ld + c] += F;
                        mA[(long)(i + N) * ld + (c + N)] += F;
                    } else {
                        mA[(long)i * ld + c] += C * sub[k];
                    }
                }
            } else {
                if (sin_node == 0) orc_intH(npe, 0, X, xe, sub, normal);
                else orc_intSingularH(npe, 0, sin_node, X, xe, sub, normal);
                for (int k = 0; k < npe; k++) {
                    const int c = conn[j * npe + k];
                    if (full) {
                        const double F = wA[iface] * sub[k];
                        const double G = wB[iface] * sub[k];
                        mA[(long)i * ld + c] += F;
                        mA[(long)i * ld + (c + N)] += G;
                        mA[(long)(i + N) * ld + c] -= G;
                        mA[(long)(i + N) * ld + (c + N)] += F;
                    } else {
                        mA[(long)i * ld + c] += wA[iface] * sub[k];
                        mA[(long)i * ld + (c + N)] += wB[iface] * sub[k];
                    }
                }
            }
        }
        vB[i] = bc[i * 3];
        vB[i + N] = bc[(i + N) * 3];
    }
    free(wA);
    free(wB);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Dense solvers                                                            */
/* ------------------------------------------------------------------------ */

/* src/gaussBksb.c:39-108: elimination with row normalisation by the pivot;
 * rows are exchanged only when |pivot| <= 1e-9 (the reference's exchange loop
 * is reproduced, including its exit on the first later row at or below the
 * tolerance).  The determinant the reference computes inside the
 * back-substitution loop is never used and is not computed here. */
int orc_gauss_bksb(int n, double *A, double *B, int max_pivots)
{
    const double TOL = 1.0E-9;
    const int n1 = n - 1;
    int npiv = (max_pivots >= 0 && max_pivots < n1) ? max_pivots : n1;
    for (int i = 0; i < npiv; i++) {
        double *Ai = A + (long)i * n;
        double C = Ai[i];
        if ((fabs(C) - TOL) <= 0) {
            for (int j = i + 1; j < n; j++) {
                double *Aj = A + (long)j * n;
                if (fabs(Aj[i]) - TOL > 0) {
                    for (int k = i; k < n; k++) {
                        C = Ai[k];
                        Ai[k] = Aj[k];
                        Aj[k] = C;
                    }
                    C = B[i];
                    B[i] = B[j];
                    B[j] = C;
                } else {
                    return -1;
                }
            }
        }
        C = Ai[i];
        for (int j = i + 1; j < n; j++) Ai[j] = Ai[j] / C;
        B[i] = B[i] / C;
        for (int j = i + 1; j < n; j++) {
            double *Aj = A + (long)j * n;
            const double Cj = Aj[i];
            for (int k = i + 1; k < n; k++) Aj[k] = Aj[k] - Cj * Ai[k];
            B[j] = B[j] - Cj * B[i];
        }
    }
    if (npiv < n1) return 0;
    if (fabs(A[(long)n1 * n + n1]) > TOL) {
        B[n1] = B[n1] / A[(long)n1 * n + n1];
        for (int i = 0; i < n1; i++) {
            const int k = n1 - i - 1;
            for (int j = k + 1; j < n; j++) B[k] = B[k] - A[(long)k * n + j] * B[j];
        }
        return 0;
    }
    return -1;
}

static double dot_prod(int n, const double *x, const double *y)
{
    double p = 0.0;
    for (int i = 0; i < n; i++) p += x[i] * y[i];
    return p;
}

static double l2norm(int n, const double *x)
{
    double s = 0.0;
    for (int i = 0; i < n; i++) s += x[i] * x[i];
    return sqrt(s);
}

/* src/gmres_el.c:323-337: complex matvec on the N x 2N layout */
static void matvec_el(int N, const double *A, const double *x, double *y)
{
    const long ld = 2L * N;
    for (int i = 0; i < N; i++) {
        const double *a = A + (long)i * ld;
        y[i] = 0.0;
        y[i + N] = 0.0;
        for (int j = 0; j < N; j++) {
            y[i] += a[j] * x[j] + a[j + N] * x[j + N];
            y[i + N] += a[j] * x[j + N] - a[j + N] * x[j];
        }
    }
}

/* src/gmres_el.c:66-128 (driver), :149-260 (iteration), :272-294 (initial
 * residual).  Unrestarted GMRES with modified Gram-Schmidt and Givens
 * rotations; the Jacobi option scales row i of A and B[i] but not B[i+N]
 * (gmres_el.c:87-93). */
int orc_gmres_el(int N, int precond, const double *x0, double *A, double *B, double tol_rel,
                 int max_iters, const char *log_path, int *iters_out)
{
    const int size = 2 * N;
    const long ld = 2L * N;
    FILE *fg = NULL;
    if (log_path) fg = fopen(log_path, "w");
    double *x = (double *)calloc(size, sizeof(double));
    double *s = (double *)calloc(max_iters + 1, sizeof(double));
    double *c = (double *)calloc(max_iters + 1, sizeof(double));
    double *res = (double *)calloc(max_iters + 1, sizeof(double));
    double *h = (double *)calloc((size_t)(max_iters + 1) * (max_iters + 1), sizeof(double));
    double *V = (double *)malloc(sizeof(double) * (size_t)size * (max_iters + 1));
    double *y = (double *)malloc(sizeof(double) * size);
    int rc = 1, iters = 0;
    if (!x || !s || !c || !res || !h || !V || !y) { rc = -1; goto done; }
#define H(r, col) h[(size_t)(r) * (max_iters + 1) + (col)]
    if (x0) memcpy(x, x0, sizeof(double) * size);
    if (precond == 1) {
        for (int i = 0; i < N; i++) {
            const double diag = 1.0 / A[(long)i * ld + i];
            B[i] = B[i] * diag;
            for (int j = 0; j < size; j++) A[(long)i * ld + j] = diag * A[(long)i * ld + j];
        }
    }
    const double tol = l2norm(size, B) * tol_rel;
    if (fg) {
        fprintf(fg, "\n\nRELATIVE TOLERANCE = %10.4e\n", tol);
        fprintf(fg, "\n%10s%21s\n", "ITER #    ", "RELATIVE RESIDUAL");
        fprintf(fg, "%10s%21s\n", "***************", "*******************");
    }
    /* initial residual (gmres_el.c:272-294) */
    double *v0 = V;
    if (l2norm(size, x) != 0.0) {
        matvec_el(N, A, x, y);
        for (int i = 0; i < size; i++) v0[i] = B[i] - y[i];
    } else {
        for (int i = 0; i < size; i++) v0[i] = B[i];
    }
    double norm = l2norm(size, v0);
    for (int i = 0; i < size; i++) v0[i] = v0[i] / norm;
    res[0] = norm;
    double err0 = norm, err1;
    if (err0 < tol) { rc = 0; goto solved; }
    for (int i = 0; i < max_iters; i++) {
        iters++;
        double *vi = V + (size_t)i * size, *w = V + (size_t)(i + 1) * size;
        matvec_el(N, A, vi, w);
        norm = l2norm(size, w);
        for (int j = 0; j <= i; j++) {
            const double *vj = V + (size_t)j * size;
            H(j, i) = dot_prod(size, vj, w);
            for (int k = 0; k < size; k++) w[k] = w[k] - H(j, i) * vj[k];
        }
        H(i + 1, i) = l2norm(size, w);
        if (H(i + 1, i) != 0.0)
            for (int j = 0; j < size; j++) w[j] = w[j] / H(i + 1, i);
        for (int j = 0; j <= i - 1; j++) {
            const double h1 = c[j] * H(j, i) + s[j] * H(j + 1, i);
            const double h2 = -s[j] * H(j, i) + c[j] * H(j + 1, i);
            H(j, i) = h1;
            H(j + 1, i) = h2;
        }
        const double htmp = H(i + 1, i) / H(i, i);
        if (norm != 0.0) {
            c[i] = 1.0 / sqrt(1.0 + htmp * htmp);
            s[i] = c[i] * htmp;
            H(i, i) = c[i] * H(i, i) + s[i] * H(i + 1, i);
            H(i + 1, i) = 0.0;
            res[i + 1] = -s[i] * res[i];
            res[i] = c[i] * res[i];
        }
        err1 = fabs(res[i + 1]);
        if (fg) fprintf(fg, "%8d%25.4e\n", iters, err1);
        if (err1 < tol) { rc = 0; break; }
        else if (fabs((err1 - err0) / err0) < 1.0e-6 * tol) { rc = 2; break; }
        else err0 = err1;
    }
    if (rc == 2) goto done;
    /* back-substitution for the Krylov weights (gmres_el.c:241-244) */
    for (int i = iters - 1; i >= 0; i--) {
        for (int j = i + 1; j <= iters; j++) res[i] -= H(i, j) * res[j];
        res[i] = res[i] / H(i, i);
    }
    for (int i = 0; i < size; i++)
        for (int j = 0; j < iters; j++) x[i] += res[j] * V[(size_t)j * size + i];
solved:
    if (rc == 0) {
        if (fg) fprintf(fg, "\n\n***PROBLEM CONVERGES IN %d ITERATIONS***\n\n", iters);
        memcpy(B, x, sizeof(double) * size);
    } else if (fg) {
        fprintf(fg, "Problem fails to vonverge to given tolerance in %d iterations\n\n", max_iters);
    }
#undef H
done:
    if (fg) fclose(fg);
    if (iters_out) *iters_out = iters;
    free(x); free(s); free(c); free(res); free(h); free(V); free(y);
    return rc;
}

/* ------------------------------------------------------------------------ */
/* Post-processing                                                          */
/* ------------------------------------------------------------------------ */

/* Singular test of potential.c:146-165 / field.c:163-181: exact coordinate
 * equality with an element node; returns the 1-based position of the first
 * match, 0 if none. */
static int coincident_node(int npe, const double *X, const double *x)
{
    for (int k = 0; k < npe; k++) {
        const double dx = X[k * 3] - x[0], dy = X[k * 3 + 1] - x[1], dz = X[k * 3 + 2] - x[2];
        if (dx == 0 && dy == 0.0 && dz == 0.0) return k + 1;
    }
    return 0;
}

/* src/potential.c:55-108 (tria3), :126-179 (tria6) */
void orc_potential(int npe, int N, const double *xyz, int E, const int *conn,
                   const double *sigma, const double *x, double *pot)
{
    const double A = 1.0 / (ORC_PI4 * ORC_EPS0);
    double X[18], sub[6];
    pot[0] = pot[1] = 0.0;
    for (int e = 0; e < E; e++) {
        gather_X(npe, xyz, conn, e, X);
        const int sn = coincident_node(npe, X, x);
        if (sn) orc_intSingularG(npe, sn, X, x, sub);
        else orc_intG(npe, X, x, sub);
        for (int k = 0; k < npe; k++) {
            const int v = conn[e * npe + k];
            pot[0] += A * sub[k] * sigma[v];
            pot[1] += A * sub[k] * sigma[v + N];
        }
    }
}

/* src/field.c:57-120 (tria3), :137-200 (tria6).  For a point on a node, the
 * jump term B*sigma*n is added once per incident element with sigma of that
 * element's FIRST node and the averaged node normal (field.c:189-195). */
void orc_field(int npe, int N, const double *xyz, int E, const int *conn,
               const double *sigma, const double *x, double *ef)
{
    const double B = 0.5 / ORC_EPS0;
    const double A = B / ORC_PI2;
    double X[18], sub[18], normal[3] = {0.0, 0.0, 0.0};
    for (int k = 0; k < 6; k++) ef[k] = 0.0;
    for (int e = 0; e < E; e++) {
        gather_X(npe, xyz, conn, e, X);
        const int sn = coincident_node(npe, X, x);
        if (sn) {
            orc_node_normal(npe, conn[e * npe + sn - 1], xyz, E, conn, normal);
            orc_intSingularH(npe, 1, sn, X, x, sub, normal);
        } else {
            orc_intH(npe, 1, X, x, sub, normal);
        }
        int jump = sn ? 1 : 0;
        for (int j = 0; j < npe; j++) {
            const int v = conn[e * npe + j];
            for (int c = 0; c < 3; c++) {
                ef[c * 2] += A * sub[j * 3 + c] * sigma[v];
                ef[c * 2 + 1] += A * sub[j * 3 + c] * sigma[v + N];
            }
            if (jump) {
                for (int c = 0; c < 3; c++) {
                    ef[c * 2] += B * sigma[v] * normal[c];
                    ef[c * 2 + 1] += B * sigma[v + N] * normal[c];
                }
                jump = 0;
            }
        }
    }
}

/* src/forceIntegrals.c:632-671 (tria3), :684-723 (tria6) */
static void intF(int npe, const double *X, const double *Ef, double *out)
{
    double L[2], N[6], XL[3], n[3];
    out[0] = out[1] = out[2] = 0.0;
    for (int g = 0; g < 7; g++) {
        L[0] = ORC_T7[g][0];
        L[1] = ORC_T7[g][1];
        const double W = ORC_T7[g][2];
        const double J = shape_and_map(npe, X, L, N, XL);
        if (npe == 3) orc_local_normal_tria3(X, n);
        else orc_local_normal_tria6(L, X, n);
        double e[3];
        for (int c = 0; c < 3; c++) {
            e[c] = Ef[0 * 3 + c] * N[0] + Ef[1 * 3 + c] * N[1] + Ef[2 * 3 + c] * N[2];
            if (npe == 6)
                e[c] = Ef[0 * 3 + c] * N[0] + Ef[1 * 3 + c] * N[1] + Ef[2 * 3 + c] * N[2]
                       + Ef[3 * 3 + c] * N[3] + Ef[4 * 3 + c] * N[4] + Ef[5 * 3 + c] * N[5];
        }
        const double NE = e[0] * n[0] + e[1] * n[1] + e[2] * n[2];
        const double Esq = e[0] * e[0] + e[1] * e[1] + e[2] * e[2];
        const double A = J * W;
        out[0] += A * (e[0] * NE - 0.5 * Esq * n[0]);
        out[1] += A * (e[1] * NE - 0.5 * Esq * n[1]);
        out[2] += A * (e[2] * NE - 0.5 * Esq * n[2]);
    }
}

/* src/force.c:348-437 (tria3), :463-552 (tria6).  Re E at every bc-6 node,
 * then 1/2 eps [E(E.n) - 1/2 |E|^2 n] over every element whose first node is
 * bc 6; Xcm is the mean of the bc-6 nodes. */
int orc_force_mst(int npe, int N, const double *xyz, int E, const int *conn,
                  const unsigned *bctype, const double *sigma, double eps,
                  double *Fce, double *Xce, double *Fcm, double *Xcm, int max_field_points)
{
    double *Ef = (double *)calloc((size_t)N * 3, sizeof(double));
    if (!Ef) return -1;
    const double A = 0.5 * eps;
    const double Lc[2] = {0.25, 0.25};
    int cnt = 0, nfield = 0;
    double ef[6], X[18], EE[18], M[6], xc[3], mF[3];
    Xcm[0] = Xcm[1] = Xcm[2] = 0.0;
    Fcm[0] = Fcm[1] = Fcm[2] = 0.0;
    for (int i = 0; i < N; i++) {
        if (bctype[i] != 6) continue;
        Xcm[0] += xyz[i * 3];
        Xcm[1] += xyz[i * 3 + 1];
        Xcm[2] += xyz[i * 3 + 2];
        cnt++;
        if (max_field_points >= 0 && nfield >= max_field_points) continue;
        orc_field(npe, N, xyz, E, conn, sigma, xyz + i * 3, ef);
        nfield++;
        Ef[i * 3] = ef[0];
        Ef[i * 3 + 1] = ef[2];
        Ef[i * 3 + 2] = ef[4];
    }
    Xcm[0] = Xcm[0] / cnt;
    Xcm[1] = Xcm[1] / cnt;
    Xcm[2] = Xcm[2] / cnt;
    for (int e = 0; e < E; e++) {
        if (bctype[conn[e * npe]] != 6) continue;
        for (int k = 0; k < npe; k++) {
            const int v = conn[e * npe + k];
            X[k * 3] = xyz[v * 3];
            X[k * 3 + 1] = xyz[v * 3 + 1];
            X[k * 3 + 2] = xyz[v * 3 + 2];
            EE[k * 3] = Ef[v * 3];
            EE[k * 3 + 1] = Ef[v * 3 + 1];
            EE[k * 3 + 2] = Ef[v * 3 + 2];
        }
        intF(npe, X, EE, mF);
        shape_and_map(npe, X, Lc, M, xc);
        for (int c = 0; c < 3; c++) {
            Fce[e * 3 + c] = A * mF[c];
            Xce[e * 3 + c] = xc[c];
        }
        Fcm[0] += Fce[e * 3];
        Fcm[1] += Fce[e * 3 + 1];
        Fcm[2] += Fce[e * 3 + 2];
    }
    free(Ef);
    return 0;
}

/* Index of a field-derivative slot E[p][q][r] in a 7x7x7 cube */
#define D3(p, q, r) (((p) * 7 + (q)) * 7 + (r))

/* src/forceIntegrals.c:61-329 (tria3), :352-619 (tria6), orders 1..3: the
 * derivatives of the Green's-function gradient, integrated against N_k. */
static void intDE(int npe, int order, const double *X, const double *xe, double *Ef /* npe*343 */)
{
    double L[2], N[6], XL[3], DE[31];
    memset(Ef, 0, sizeof(double) * npe * 343);
    for (int i = 0; i < 31; i++) DE[i] = 0.0;
    for (int g = 0; g < 7; g++) {
        L[0] = ORC_T7[g][0];
        L[1] = ORC_T7[g][1];
        const double W = ORC_T7[g][2];
        const double J = shape_and_map(npe, X, L, N, XL);
        const double dx = xe[0] - XL[0], dy = xe[1] - XL[1], dz = xe[2] - XL[2];
        const double dx2 = dx * dx, dy2 = dy * dy, dz2 = dz * dz;
        const double rsq = dx2 + dy2 + dz2;
        const double A = J * W / pow(rsq, 1.5);
        const double B = 3.0 * A / rsq;
        DE[0] = A - B * dx2;
        DE[1] = -B * dx * dy;
        DE[2] = -B * dx * dz;
        DE[3] = A - B * dy2;
        DE[4] = -B * dy * dz;
        DE[5] = A - B * dz2;
        if (order > 1) {
            const double C = 5.0 * B / rsq;
            DE[6] = (C * dx2 - 3.0 * B) * dx;
            DE[7] = (C * dx2 - B) * dy;
            DE[8] = (C * dx2 - B) * dz;
            DE[9] = (C * dy2 - B) * dx;
            DE[10] = C * dx * dy * dz;
            DE[11] = (C * dz2 - B) * dx;
            DE[12] = (C * dy2 - 3.0 * B) * dy;
            DE[13] = (C * dy2 - B) * dz;
            DE[14] = (C * dz2 - B) * dy;
            DE[15] = (C * dz2 - 3.0 * B) * dz;
            if (order > 2) {
                const double D = 7.0 * C / rsq;
                DE[16] = (-D * dx2 + 6.0 * C) * dx2 - 3.0 * B;
                DE[17] = (-D * dx2 + 3.0 * C) * dx * dy;
                DE[18] = (-D * dx2 + 3.0 * C) * dx * dz;
                DE[19] = -D * dx2 * dy2 + C * (dx2 + dy2) - B;
                DE[20] = (-D * dx2 + C) * dy * dz;
                DE[21] = -D * dx2 * dz2 + C * (dx2 + dz2) - B;
                DE[22] = (-D * dy2 + 3.0 * C) * dx * dy;
                DE[23] = (-D * dy2 + C) * dx * dz;
                DE[24] = (-D * dz2 + C) * dx * dy;
                DE[25] = (-D * dz2 + 3.0 * C) * dx * dz;
                DE[26] = (-D * dy2 + 6.0 * C) * dy2 - 3.0 * B;
                DE[27] = (-D * dy2 + 3.0 * C) * dy * dz;
                DE[28] = -D * dy2 * dz2 + C * (dy2 + dz2) - B;
                DE[29] = (-D * dz2 + 3.0 * C) * dy * dz;
                DE[30] = (-D * dz2 + 6.0 * C) * dz2 - 3.0 * B;
            }
        }
        static const int slot1[6][3] = {{2,0,0},{1,1,0},{1,0,1},{0,2,0},{0,1,1},{0,0,2}};
        static const int slot2[10][3] = {{3,0,0},{2,1,0},{2,0,1},{1,2,0},{1,1,1},{1,0,2},
                                         {0,3,0},{0,2,1},{0,1,2},{0,0,3}};
        static const int slot3[15][3] = {{4,0,0},{3,1,0},{3,0,1},{2,2,0},{2,1,1},{2,0,2},
                                         {1,3,0},{1,2,1},{1,1,2},{1,0,3},{0,4,0},{0,3,1},
                                         {0,2,2},{0,1,3},{0,0,4}};
        for (int j = 0; j < npe; j++) {
            double *Ej = Ef + j * 343;
            for (int t = 0; t < 6; t++) Ej[D3(slot1[t][0], slot1[t][1], slot1[t][2])] += DE[t] * N[j];
            if (order > 1) {
                for (int t = 0; t < 10; t++)
                    Ej[D3(slot2[t][0], slot2[t][1], slot2[t][2])] += DE[6 + t] * N[j];
                if (order > 2)
                    for (int t = 0; t < 15; t++)
                        Ej[D3(slot3[t][0], slot3[t][1], slot3[t][2])] += DE[16 + t] * N[j];
            }
        }
    }
}

/* src/force.c:572-688 (tria3), :709-825 (tria6).  Re sigma only; orders above
 * 3 contribute nothing to F in the reference, so order is clamped to 3. */
void orc_force_multipole(int npe, int order, double R, const double *x, int N,
                         const double *xyz, int E, const int *conn, const double *sigma,
                         const double *mats, double freq, double *F)
{
    if (order > 3) order = 3;
    const double w = freq * ORC_PI2 * ORC_EPS0;
    const double Eps = mats[1] * ORC_EPS0;
    const double B = 1.0 / (ORC_PI4 * ORC_EPS0);
    double Cc[4] = {0.0, 0.0, 0.0, 0.0}, ef[6], X[18];
    double *Ecube = (double *)calloc(343, sizeof(double));
    double *intE = (double *)malloc(sizeof(double) * 6 * 343);
    F[0] = F[1] = F[2] = 0.0;
    if (!Ecube || !intE) { free(Ecube); free(intE); return; }
    orc_field(npe, N, xyz, E, conn, sigma, x, ef);
    Ecube[D3(1, 0, 0)] = ef[0];
    Ecube[D3(0, 1, 0)] = ef[2];
    Ecube[D3(0, 0, 1)] = ef[4];
    for (int e = 0; e < E; e++) {
        gather_X(npe, xyz, conn, e, X);
        intDE(npe, order, X, x, intE);
        for (int j = 0; j < npe; j++) {
            const double s = sigma[conn[e * npe + j]];
            for (int t = 0; t < 343; t++) Ecube[t] += B * intE[j * 343 + t] * s;
        }
    }
    const double EpsDif = w * (mats[3] - mats[1]);
    double EpsPlus = w * (mats[3] + 2.0 * mats[1]);
    const double SigDif = mats[2] - mats[0];
    double SigPlus = mats[2] + 2.0 * mats[0];
    const double CM1 = (EpsDif * EpsPlus + SigDif * SigPlus) / (EpsPlus * EpsPlus + SigPlus * SigPlus);
    Cc[1] = ORC_PI2 * Eps * R * R * R;
    if (order > 1) {
        EpsPlus = w * (2.0 * mats[3] + 3.0 * mats[1]);
        SigPlus = 2.0 * mats[2] + 3.0 * mats[0];
        const double CM2 = (EpsDif * EpsPlus + SigDif * SigPlus) / (EpsPlus * EpsPlus + SigPlus * SigPlus);
        Cc[2] = (npe == 6) ? Cc[1] * R * R * 0.333333333333333333 : Cc[1] * R * R / 3.0;
        if (order > 2) {
            EpsPlus = w * (3.0 * mats[3] + 4.0 * mats[1]);
            SigPlus = 3.0 * mats[2] + 4.0 * mats[0];
            const double CM3 = (EpsDif * EpsPlus + SigDif * SigPlus) / (EpsPlus * EpsPlus + SigPlus * SigPlus);
            Cc[3] = (npe == 6) ? Cc[2] * R * R * 0.1 : Cc[2] * R * R / 10.0;
            Cc[3] = Cc[3] * CM3;
        }
        Cc[2] = Cc[2] * CM2;
    }
    Cc[1] = Cc[1] * CM1;
#define EC(p, q, r) Ecube[D3(p, q, r)]
    F[0] = Cc[1] * (EC(1,0,0) * EC(2,0,0) + EC(0,1,0) * EC(1,1,0) + EC(0,0,1) * EC(1,0,1));
    F[1] = Cc[1] * (EC(1,0,0) * EC(1,1,0) + EC(0,1,0) * EC(0,2,0) + EC(0,0,1) * EC(0,1,1));
    F[2] = Cc[1] * (EC(1,0,0) * EC(1,0,1) + EC(0,1,0) * EC(0,1,1) + EC(0,0,1) * EC(0,0,2));
    if (order > 1) {
        F[0] += Cc[2] * (EC(2,0,0) * EC(3,0,0) + EC(0,2,0) * EC(1,2,0) + EC(0,0,2) * EC(1,0,2)
                + 2.0 * (EC(1,1,0) * EC(2,1,0) + EC(1,0,1) * EC(2,0,1) + EC(0,1,1) * EC(1,1,1)));
        F[1] += Cc[2] * (EC(2,0,0) * EC(2,1,0) + EC(0,2,0) * EC(0,3,0) + EC(0,0,2) * EC(0,1,2)
                + 2.0 * (EC(1,1,0) * EC(1,2,0) + EC(1,0,1) * EC(1,1,1) + EC(0,1,1) * EC(0,2,1)));
        F[2] += Cc[2] * (EC(2,0,0) * EC(2,0,1) + EC(0,2,0) * EC(0,2,1) + EC(0,0,2) * EC(0,0,3)
                + 2.0 * (EC(1,1,0) * EC(1,1,1) + EC(1,0,1) * EC(1,0,2) + EC(0,1,1) * EC(0,1,2)));
        if (order > 2) {
            F[0] += Cc[3] * (EC(3,0,0) * EC(4,0,0) + EC(0,3,0) * EC(1,3,0) + EC(0,0,3) * EC(1,0,3)
                    + 3.0 * (EC(2,1,0) * EC(3,1,0) + EC(2,0,1) * EC(3,0,1) + EC(1,2,0) * EC(2,2,0)
                             + EC(1,0,2) * EC(2,0,2) + EC(0,2,1) * EC(1,2,1) + EC(0,1,2) * EC(1,1,2))
                    + 6.0 * EC(1,1,1) * EC(2,1,1));
            F[1] += Cc[3] * (EC(3,0,0) * EC(3,1,0) + EC(0,3,0) * EC(0,4,0) + EC(0,0,3) * EC(0,1,3)
                    + 3.0 * (EC(2,1,0) * EC(2,2,0) + EC(2,0,1) * EC(2,1,1) + EC(1,2,0) * EC(1,3,0)
                             + EC(1,0,2) * EC(1,1,2) + EC(0,2,1) * EC(0,3,1) + EC(0,1,2) * EC(0,2,2))
                    + 6.0 * EC(1,1,1) * EC(1,2,1));
            F[2] += Cc[3] * (EC(3,0,0) * EC(3,0,1) + EC(0,3,0) * EC(0,3,1) + EC(0,0,3) * EC(0,0,4)
                    + 3.0 * (EC(2,1,0) * EC(2,1,1) + EC(2,0,1) * EC(2,0,2) + EC(1,2,0) * EC(1,2,1)
                             + EC(1,0,2) * EC(1,0,3) + EC(0,2,1) * EC(0,2,2) + EC(0,1,2) * EC(0,1,3))
                    + 6.0 * EC(1,1,1) * EC(1,1,2));
        }
    }
#undef EC
    free(Ecube);
    free(intE);
}

/* 64-point Gauss-Legendre rule on (-1, 1), ascending nodes: Newton's method on
 * P_64 from the Chebyshev-like start cos(pi (i + 3/4) / (n + 1/2)), weights
 * 2 / ((1 - x^2) P'_64(x)^2).  The reference tabulates the same rule to 15
 * digits (src/gaussData.h:76, GaussH); computed here from its definition. */
static void gauss_legendre64(double *x, double *w)
{
    const int n = 64;
    for (int i = 0; i < n / 2; i++) {
        double z = cos(3.14159265358979323846 * (i + 0.75) / (n + 0.5)), dp = 0.0;
        for (int it = 0; it < 100; it++) {
            double p0 = 1.0, p1 = z;
            for (int k = 2; k <= n; k++) {
                const double p2 = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p0) / k;
                p0 = p1;
                p1 = p2;
            }
            dp = n * (z * p1 - p0) / (z * z - 1.0);
            const double dz = p1 / dp;
            z -= dz;
            if (fabs(dz) < 1e-17) break;
        }
        {
            double p0 = 1.0, p1 = z;
            for (int k = 2; k <= n; k++) {
                const double p2 = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p0) / k;
                p0 = p1;
                p1 = p2;
            }
            dp = n * (z * p1 - p0) / (z * z - 1.0);
        }
        const double wt = 2.0 / ((1.0 - z * z) * dp * dp);
        x[i] = -z;
        w[i] = wt;
        x[n - 1 - i] = z;
        w[n - 1 - i] = wt;
    }
}

/* Depolarisation factors of an ellipsoid (src/depolarization.c:44-115):
 * L_k = (abc/2) int_0^{100 a_k^2} ds / ((s + a_k^2) sqrt((s+a^2)(s+b^2)(s+c^2)))
 * by the 64-point rule on the segment [0, 100 a_k^2] (the reference truncates
 * the infinite integral there).  Returns -1 if a factor is negative (the
 * reference stops with an error, depolarization.c:106-108). */
int orc_depolarization(const double *axis, double *Ld)
{
    double gx[64], gw[64];
    gauss_legendre64(gx, gw);
    const double A = 0.5 * axis[0] * axis[1] * axis[2];
    const double a2[3] = {axis[0] * axis[0], axis[1] * axis[1], axis[2] * axis[2]};
    for (int k = 0; k < 3; k++) {
        const double top = 100.0 * a2[k];
        const double J = 0.5 * top;   /* X2L_line2 on the segment (0,0)-(top,0) */
        double acc = 0.0;
        for (int i = 0; i < 64; i++) {
            const double s = 0.5 * (1.0 + gx[i]) * top;
            const double R = J * gw[i] / sqrt((s + a2[0]) * (s + a2[1]) * (s + a2[2]));
            acc += R / (s + a2[k]);
        }
        Ld[k] = A * acc;
    }
    for (int k = 0; k < 3; k++)
        if (Ld[k] < 0.0) return -1;
    return 0;
}

/* Dipolar DEP force on an ellipsoid centred at each of m points
 * (src/force.c:65-183 forceEllipsoid_tria3, :204-322 forceEllipsoid_tria6):
 * Re E at the point and its first derivatives from intDE of order 1, the
 * generalised Clausius-Mossotti factor per axis, then
 * F = (2 pi eps_m abc / 3) sum_k K_k E_k dE_k/dx_j.
 * analysis 10: homogeneous ellipsoid; 11: single-shell oblate ellipsoid, whose
 * particle (sigma, eps_r) the reference scales by C = ac / (2c + a) in its
 * material table on every call (force.c:148-155), so point k sees them scaled
 * by C^(k+1); mats (nM x 2) is updated in place the same way. */
void orc_force_ellipsoid(int npe, int analysis, const double *axis, int m, const double *pts,
                         int N, const double *xyz, int E, const int *conn, const double *sigma,
                         double *mats, double freq, double *F)
{
    const double w = freq * ORC_PI2 * ORC_EPS0;
    const double B = 1.0 / (ORC_PI4 * ORC_EPS0);
    const double C = axis[0] * axis[2] / (2.0 * axis[2] + axis[0]);
    double *Ecube = (double *)malloc(sizeof(double) * 343);
    double *intE = (double *)malloc(sizeof(double) * 6 * 343);
    double ef[6], X[18], Ld[3];
    if (!Ecube || !intE) { free(Ecube); free(intE); return; }
    for (int p = 0; p < m; p++) {
        const double *x = pts + 3 * p;
        double *Fp = F + 3 * p;
        const double Eps = mats[1] * ORC_EPS0;
        const double A = ORC_PI2 * Eps * axis[0] * axis[1] * axis[2] / 3.0;
        double K[3] = {0.0, 0.0, 0.0};
        memset(Ecube, 0, sizeof(double) * 343);
        orc_field(npe, N, xyz, E, conn, sigma, x, ef);
        Ecube[D3(1, 0, 0)] = ef[0];
        Ecube[D3(0, 1, 0)] = ef[2];
        Ecube[D3(0, 0, 1)] = ef[4];
        for (int e = 0; e < E; e++) {
            gather_X(npe, xyz, conn, e, X);
            intDE(npe, 1, X, x, intE);
            for (int j = 0; j < npe; j++) {
                const double s = sigma[conn[e * npe + j]];
                for (int t = 0; t < 343; t++) Ecube[t] += B * intE[j * 343 + t] * s;
            }
        }
        orc_depolarization(axis, Ld);
        if (analysis == 10 || analysis == 11) {
            double ep[3], sg[3];
            if (analysis == 10) {
                for (int k = 0; k < 3; k++) { ep[k] = mats[3]; sg[k] = mats[2]; }
            } else {
                mats[3] = mats[3] * C;
                mats[2] = mats[2] * C;
                for (int k = 0; k < 3; k++) { ep[k] = mats[3] / Ld[k]; sg[k] = mats[2] / Ld[k]; }
            }
            for (int k = 0; k < 3; k++) {
                const double EpsDif = w * (ep[k] - mats[1]);
                const double SigDif = sg[k] - mats[0];
                const double EpsPlus = w * (mats[1] + (ep[k] - mats[1]) * Ld[k]);
                const double SigPlus = mats[0] + (sg[k] - mats[0]) * Ld[k];
                K[k] = (EpsDif * EpsPlus + SigDif * SigPlus) / (EpsPlus * EpsPlus + SigPlus * SigPlus);
            }
        }
#define EC(p, q, r) Ecube[D3(p, q, r)]
        Fp[0] = A * (K[0] * EC(1,0,0) * EC(2,0,0) + K[1] * EC(0,1,0) * EC(1,1,0) + K[2] * EC(0,0,1) * EC(1,0,1));
        Fp[1] = A * (K[0] * EC(1,0,0) * EC(1,1,0) + K[1] * EC(0,1,0) * EC(0,2,0) + K[2] * EC(0,0,1) * EC(0,1,1));
        Fp[2] = A * (K[0] * EC(1,0,0) * EC(1,0,1) + K[1] * EC(0,1,0) * EC(0,1,1) + K[2] * EC(0,0,1) * EC(0,0,2));
#undef EC
    }
    free(Ecube);
    free(intE);
}
