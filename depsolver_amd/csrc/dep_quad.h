// dep_quad.h -- device quadrature for singular (node-on-element) pairs.
//
// Restates the weakly and strongly singular rules of the reference
// (src/kernelIntegrals.c:299-454 intSingularG_*, :472-805 intSingularH_*)
// for one (collocation point, element) pair per thread.  Regular pairs never
// come here: the regular 7-point sums are evaluated in tiles from the
// precomputed quadrature points (dep_assemble.hip, dep_post.hip).
#pragma once

#include "dep_common.h"

// Rotation that brings a corner singular node to the element's last corner
// (kernelIntegrals.c:492-497 tria3, :630-641 tria6): rotated position p holds
// original node (p + s) % NPE.
template <int NPE>
__device__ __forceinline__ int singular_rotation(int sn)
{
    if constexpr (NPE == 3) return sn % 3;
    else return (sn % 2 == 1) ? (sn + 1) % 6 : 0;
}

template <int NPE>
__device__ __forceinline__ void rotate_nodes(const double (*X)[3], int s, double (*Xr)[3])
{
#pragma unroll
    for (int p = 0; p < NPE; p++)
#pragma unroll
        for (int c = 0; c < 3; c++) {
            double v = X[p][c];
            if constexpr (NPE == 3) {
                v = (s == 1) ? X[(p + 1) % 3][c] : v;
                v = (s == 2) ? X[(p + 2) % 3][c] : v;
            } else {
                v = (s == 2) ? X[(p + 2) % 6][c] : v;
                v = (s == 4) ? X[(p + 4) % 6][c] : v;
            }
            Xr[p][c] = v;
        }
}

// out[(p + s) % NPE] = rot[p], per component block of size W
template <int NPE, int W>
__device__ __forceinline__ void unrotate(const double* rot, int s, double* out)
{
#pragma unroll
    for (int q = 0; q < NPE; q++)
#pragma unroll
        for (int c = 0; c < W; c++) {
            // original node q came from rotated position (q - s) mod NPE
            double v = rot[q * W + c];
            if constexpr (NPE == 3) {
                v = (s == 1) ? rot[((q + 2) % 3) * W + c] : v;
                v = (s == 2) ? rot[((q + 1) % 3) * W + c] : v;
            } else {
                v = (s == 2) ? rot[((q + 4) % 6) * W + c] : v;
                v = (s == 4) ? rot[((q + 2) % 6) * W + c] : v;
            }
            out[q * W + c] = v;
        }
}

// Weakly singular G: Duffy-collapsed square, 8 Gauss-Legendre x 8
// Gauss-Jacobi(0,1) points; a tria6 midside node splits the element in two
// (weight factor 1/16).  out[k] = sum W1 W2 J B N_k / r.
template <int NPE>
__device__ void singular_G(int sn, const double (*X)[3], const double* xe, double* out)
{
#pragma unroll
    for (int k = 0; k < NPE; k++) out[k] = 0.0;
    const bool two = (NPE == 6) && (sn % 2 == 0);
    const double B = two ? 0.0625 : 0.125;
    for (int i = 0; i < 8; i++) {
        const double U1 = kGL8[i][0], W1 = kGL8[i][1];
        for (int j = 0; j < 8; j++) {
            const double U2 = kGJ8[j][0], W2 = kGJ8[j][1];
            double a1, b1, a2 = 0.0, b2 = 0.0;
            if constexpr (NPE == 3) {
                if (sn == 1) { a1 = 0.5 * (1.0 - U2); b1 = 0.25 * (1.0 + U1) * (1.0 + U2); }
                else if (sn == 2) { a1 = 0.25 * (1.0 - U1) * (1.0 + U2); b1 = 0.5 * (1.0 - U2); }
                else { a1 = 0.25 * (1.0 + U1) * (1.0 + U2); b1 = 0.25 * (1.0 - U1) * (1.0 + U2); }
            } else {
                switch (sn) {
                case 1: a1 = 0.5 * (1.0 - U2); b1 = 0.25 * (1.0 + U1) * (1.0 + U2); break;
                case 2:
                    a1 = 0.25 * (1.0 - U2); b1 = 0.5 + 0.25 * U1 * (1.0 + U2);
                    a2 = 0.5 - 0.25 * U1 * (1.0 + U2); b2 = 0.25 * (1.0 - U2);
                    break;
                case 3: a1 = 0.25 * (1.0 - U1) * (1.0 + U2); b1 = 0.5 * (1.0 - U2); break;
                case 4:
                    a1 = 0.25 * (1.0 + U1) * (1.0 + U2); b1 = 0.5 - 0.25 * U1 * (1.0 + U2);
                    a2 = 0.25 * (1.0 - U1) * (1.0 + U2); b2 = 0.25 * (1.0 - U2);
                    break;
                case 5: a1 = 0.25 * (1.0 + U1) * (1.0 + U2); b1 = 0.25 * (1.0 - U1) * (1.0 + U2); break;
                default:
                    a1 = 0.5 + 0.25 * U1 * (1.0 + U2); b1 = 0.25 * (1.0 - U1) * (1.0 + U2);
                    a2 = 0.25 * (1.0 - U2); b2 = 0.25 * (1.0 + U1) * (1.0 + U2);
                    break;
                }
            }
            double N[NPE], xl[3];
            shape_fn<NPE>(a1, b1, N);
            double J = map_point<NPE>(X, a1, b1, N, xl);
            double dx = xe[0] - xl[0], dy = xe[1] - xl[1], dz = xe[2] - xl[2];
            double A = W1 * W2 * J * B / sqrt(dx * dx + dy * dy + dz * dz);
#pragma unroll
            for (int k = 0; k < NPE; k++) out[k] += A * N[k];
            if (two) {
                shape_fn<NPE>(a2, b2, N);
                J = map_point<NPE>(X, a2, b2, N, xl);
                dx = xe[0] - xl[0]; dy = xe[1] - xl[1]; dz = xe[2] - xl[2];
                A = W1 * W2 * J * B / sqrt(dx * dx + dy * dy + dz * dz);
#pragma unroll
                for (int k = 0; k < NPE; k++) out[k] += A * N[k];
            }
        }
    }
}

// Strongly singular H by 8 levels of subdivision (3 sub-triangles x 7 points
// per level, the innermost triangle dropped).  FLAG == 0: out[k] -= W J N_k
// (d.n)/r^3 (NPE values); FLAG == 1: out[3k+c] += W J N_k d_c / r^3.  For a
// tria6 midside node the z-component adds B2*N2 where C2*N2 belongs, as the
// reference does (kernelIntegrals.c:762).
template <int NPE, int FLAG>
__device__ void singular_H(int sn, const double (*Xin)[3], const double* xe, const double* nrm,
                           double* out)
{
    constexpr int W = FLAG ? 3 : 1;
    double X[NPE][3];
    const int s = singular_rotation<NPE>(sn);
    rotate_nodes<NPE>(Xin, s, X);
    double acc[NPE * W];
#pragma unroll
    for (int k = 0; k < NPE * W; k++) acc[k] = 0.0;
    const bool corner = (NPE == 3) || (sn % 2 == 1);
    // first-level sub-triangles (kernelIntegrals.c:508-518), singular corner at L = (0,0)
    const double T0[3][3][2] = {{{1.0, 0.0}, {0.5, 0.5}, {0.5, 0.0}},
                                {{0.5, 0.0}, {0.5, 0.5}, {0.0, 0.5}},
                                {{0.5, 0.5}, {0.0, 1.0}, {0.0, 0.5}}};
    double J = 1.0, scale = 1.0;
    for (int m = 0; m < DEP_NSUB; m++) {
        J = J * 0.25;
        for (int i = 0; i < 3; i++) {
            const double x0a = T0[i][0][0] * scale, x0b = T0[i][0][1] * scale;
            const double x1a = T0[i][1][0] * scale, x1b = T0[i][1][1] * scale;
            const double x2a = T0[i][2][0] * scale, x2b = T0[i][2][1] * scale;
            for (int j = 0; j < 7; j++) {
                const double M1 = kT7[j][0], M2 = kT7[j][1];
                const double M3 = 1.0 - M1 - M2;
                const double Wt = kT7[j][2];
                if (corner) {
                    const double a = x0a * M1 + x1a * M2 + x2a * M3;
                    const double b = x0b * M1 + x1b * M2 + x2b * M3;
                    double N[NPE], xl[3];
                    shape_fn<NPE>(a, b, N);
                    const double JL = map_point<NPE>(X, a, b, N, xl);
                    const double dx = xe[0] - xl[0], dy = xe[1] - xl[1], dz = xe[2] - xl[2];
                    const double rsq = dx * dx + dy * dy + dz * dz;
                    const double r = sqrt(rsq);
                    double A = J * JL * Wt / (rsq * r);
                    if constexpr (FLAG == 0) {
                        A = A * (dx * nrm[0] + dy * nrm[1] + dz * nrm[2]);
#pragma unroll
                        for (int k = 0; k < NPE; k++) acc[k] -= A * N[k];
                    } else {
                        const double Cz = A * dz, By = A * dy, Ax = A * dx;
#pragma unroll
                        for (int k = 0; k < NPE; k++) {
                            acc[3 * k] += Ax * N[k];
                            acc[3 * k + 1] += By * N[k];
                            acc[3 * k + 2] += Cz * N[k];
                        }
                    }
                } else {
                    double U1, U2, U3, V1, V2, V3;
                    if (sn == 2) {
                        U1 = 0.5 * M3; U2 = M1 + 0.5 * M3;
                        V1 = M1 + 0.5 * M3; V2 = M3;
                    } else if (sn == 4) {
                        U1 = M1; U2 = 0.5 * M3;
                        V1 = M1; V2 = M2 + 0.5 * M3;
                    } else {
                        U1 = 0.5 * M3; U2 = M1;
                        V1 = M1 + 0.5 * M3; V2 = M2;
                    }
                    U3 = 1.0 - U1 - U2;
                    V3 = 1.0 - V1 - V2;
                    const double a1 = x0a * U1 + x1a * U2 + x2a * U3;
                    const double b1 = x0b * U1 + x1b * U2 + x2b * U3;
                    const double a2 = x0a * V1 + x1a * V2 + x2a * V3;
                    const double b2 = x0b * V1 + x1b * V2 + x2b * V3;
                    double N1[NPE], N2[NPE], xl1[3], xl2[3];
                    shape_fn<NPE>(a1, b1, N1);
                    shape_fn<NPE>(a2, b2, N2);
                    const double JL1 = map_point<NPE>(X, a1, b1, N1, xl1);
                    const double JL2 = map_point<NPE>(X, a2, b2, N2, xl2);
                    const double dx1 = xe[0] - xl1[0], dy1 = xe[1] - xl1[1], dz1 = xe[2] - xl1[2];
                    const double dx2 = xe[0] - xl2[0], dy2 = xe[1] - xl2[1], dz2 = xe[2] - xl2[2];
                    const double dr1 = sqrt(dx1 * dx1 + dy1 * dy1 + dz1 * dz1);
                    const double dr2 = sqrt(dx2 * dx2 + dy2 * dy2 + dz2 * dz2);
                    double A1 = 0.5 * J * JL1 * Wt / (dr1 * dr1 * dr1);
                    double A2 = 0.5 * J * JL2 * Wt / (dr2 * dr2 * dr2);
                    if constexpr (FLAG == 0) {
                        A1 = A1 * (dx1 * nrm[0] + dy1 * nrm[1] + dz1 * nrm[2]);
                        A2 = A2 * (dx2 * nrm[0] + dy2 * nrm[1] + dz2 * nrm[2]);
#pragma unroll
                        for (int k = 0; k < NPE; k++) acc[k] -= A1 * N1[k] + A2 * N2[k];
                    } else {
                        const double C1 = A1 * dz1, B1 = A1 * dy1, Ax1 = A1 * dx1;
                        const double B2 = A2 * dy2, Ax2 = A2 * dx2;
#pragma unroll
                        for (int k = 0; k < NPE; k++) {
                            acc[3 * k] += Ax1 * N1[k] + Ax2 * N2[k];
                            acc[3 * k + 1] += B1 * N1[k] + B2 * N2[k];
                            acc[3 * k + 2] += C1 * N1[k] + B2 * N2[k];
                        }
                    }
                }
            }
        }
        scale *= 0.5;
    }
    unrotate<NPE, W>(acc, s, out);
}

// Averaged node normal: unit normals of the incident elements at the node,
// summed in element order and normalised (src/normals.c:120-199, :207-319).
template <int NPE>
__device__ void node_normal(const double* xyz, const int* conn, const int* ek, int n_inc,
                            double* nrm)
{
    double s0 = 0.0, s1 = 0.0, s2 = 0.0;
    for (int t = 0; t < n_inc; t++) {
        const int e = ek[t] >> 3, k = ek[t] & 7;
        double X[NPE][3];
#pragma unroll
        for (int p = 0; p < NPE; p++) {
            const int v = conn[e * NPE + p];
            X[p][0] = xyz[3 * v];
            X[p][1] = xyz[3 * v + 1];
            X[p][2] = xyz[3 * v + 2];
        }
        double a, b, n[3];
        node_local<NPE>(k, &a, &b);
        local_normal<NPE>(X, a, b, n);
        s0 += n[0];
        s1 += n[1];
        s2 += n[2];
    }
    const double inv = 1.0 / sqrt(s0 * s0 + s1 * s1 + s2 * s2);
    nrm[0] = s0 * inv;
    nrm[1] = s1 * inv;
    nrm[2] = s2 * inv;
}
