/*
 * dropin_main.c -- calls the reference-signature entry points of
 * libdepsolver_amd.so the way src/depSolver.c:284-310 does: globals nNodes,
 * nElems, nInterfaces and file_log defined here, row-pointer matrices,
 * 1-based connectivity, then electricForm[Full]A_tria{3,6} and gaussBksb or
 * solverGMRES_el, then postProcess_tria* ("STD" files) and, when internal
 * points are given, vtkPostProcess_tria* on an nInternalPoints x 1 x 1 grid
 * (results.vtk).  Input: a text problem written by the test; output: the
 * solution vector (2N values), one per line.
 *
 *   dropin_main problem.txt solution.txt lu|gmres
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../include/depsolver_amd.h"

FILE *file_log;
unsigned int nNodes, nElems, nInterfaces, nInternalPoints, nFPoints;
unsigned int NX, NY, NZ;            /* the VTK grid (src/depSolver.c globals) */
double deltaX, deltaY, deltaZ;

static double **dmat(unsigned r, unsigned c)
{
    double **m = malloc(sizeof(double *) * r);
    for (unsigned i = 0; i < r; i++) m[i] = calloc(c, sizeof(double));
    return m;
}

int main(int argc, char **argv)
{
    if (argc < 4) return 2;
    FILE *f = fopen(argv[1], "r");
    unsigned npe, nM;
    double freq;
    if (fscanf(f, "%u %u %u %u %u %le", &nNodes, &nElems, &npe, &nM, &nInterfaces, &freq) != 6) return 3;
    const unsigned N = nNodes;
    double **nodes = dmat(N, 3), **mats = dmat(nM, 2), **bc = dmat(2 * N, 3);
    unsigned **el = malloc(sizeof(unsigned *) * nElems), **ifc = malloc(sizeof(unsigned *) * (nInterfaces + 1));
    unsigned *bct = calloc(2 * N, sizeof(unsigned));
    for (unsigned i = 0; i < N; i++)
        if (fscanf(f, "%le %le %le", &nodes[i][0], &nodes[i][1], &nodes[i][2]) != 3) return 4;
    for (unsigned e = 0; e < nElems; e++) {
        el[e] = malloc(sizeof(unsigned) * npe);
        for (unsigned k = 0; k < npe; k++)
            if (fscanf(f, "%u", &el[e][k]) != 1) return 5;
    }
    for (unsigned m = 0; m < nM; m++)
        if (fscanf(f, "%le %le", &mats[m][0], &mats[m][1]) != 2) return 6;
    for (unsigned i = 0; i < nInterfaces; i++) {
        ifc[i] = malloc(sizeof(unsigned) * 2);
        if (fscanf(f, "%u %u", &ifc[i][0], &ifc[i][1]) != 2) return 7;
    }
    for (unsigned j = 0; j < 2 * N; j++)
        if (fscanf(f, "%u %le %le %le", &bct[j], &bc[j][0], &bc[j][1], &bc[j][2]) != 4) return 8;
    double **Xin = NULL;
    if (fscanf(f, "%u", &nInternalPoints) == 1 && nInternalPoints) {
        Xin = dmat(nInternalPoints, 3);
        for (unsigned i = 0; i < nInternalPoints; i++)
            if (fscanf(f, "%le %le %le", &Xin[i][0], &Xin[i][1], &Xin[i][2]) != 3) return 9;
    }
    fclose(f);
    file_log = fopen("/dev/null", "w");
    const int lu = strcmp(argv[3], "lu") == 0;
    double **A = dmat(lu ? 2 * N : N, 2 * N);
    double *B = calloc(2 * N, sizeof(double));
    double prob[1] = {freq};
    if (npe == 6) {
        if (lu) electricFormFullA_tria6(nodes, el, bct, ifc, mats, prob, A, bc, B);
        else electricFormA_tria6(nodes, el, bct, ifc, mats, prob, A, bc, B);
    } else {
        if (lu) electricFormFullA_tria3(nodes, el, bct, ifc, mats, prob, A, bc, B);
        else electricFormA_tria3(nodes, el, bct, ifc, mats, prob, A, bc, B);
    }
    if (lu) gaussBksb(2 * N, A, B);
    else solverGMRES_el(1, 0, A, B);
    FILE *o = fopen(argv[2], "w");
    for (unsigned j = 0; j < 2 * N; j++) fprintf(o, "%.17e\n", B[j]);
    fclose(o);
    double axis[3] = {1.0, 1.0, 1.0};
    char std_type[] = "STD";
    if (npe == 6) postProcess_tria6(3, std_type, axis, NULL, Xin, nodes, el, mats, prob, bct, B);
    else postProcess_tria3(3, std_type, axis, NULL, Xin, nodes, el, mats, prob, bct, B);
    if (nInternalPoints) {
        char vtk_type[] = "VTK";
        NX = nInternalPoints;
        NY = NZ = 1;
        deltaX = 0.5;
        deltaY = deltaZ = 1.0;
        if (npe == 6) vtkPostProcess_tria6(3, vtk_type, axis, NULL, Xin, nodes, el, mats, prob, bct, B);
        else vtkPostProcess_tria3(3, vtk_type, axis, NULL, Xin, nodes, el, mats, prob, bct, B);
    }
    return 0;
}
