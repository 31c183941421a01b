/*
 * orc_tables.h -- quadrature tables and physical constants for the oracle.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * The decimal literals are the reference's (src/gaussData.h:38-74,
 * src/constants.h:42-55).  They must be the same numbers for the oracle to be
 * bitwise comparable with the reference build; tests/test_oracle_tables.py
 * rebuilds every table from its closed form (Golub-Welsch for Gauss-Legendre
 * and Gauss-Jacobi(0,1), the Radon 7-point degree-5 rule) and checks them.
 */
#ifndef DEPSOLVER_ORC_TABLES_H
#define DEPSOLVER_ORC_TABLES_H

#define ORC_EPS0 8.854187817E-12
#define ORC_PI2  6.283185307179586E+00
#define ORC_PI4  1.256637061435917E+01

/* 8-point Gauss-Legendre on (-1,1): abscissa, weight */
static const double ORC_GL8[8][2] = {
    {-9.60289856497536E-01, 1.01228536290370E-01},
    {-7.96666477413627E-01, 2.22381034453376E-01},
    {-5.25532409916329E-01, 3.13706645877887E-01},
    {-1.83434642495650E-01, 3.62683783378362E-01},
    { 1.83434642495650E-01, 3.62683783378362E-01},
    { 5.25532409916329E-01, 3.13706645877887E-01},
    { 7.96666477413627E-01, 2.22381034453376E-01},
    { 9.60289856497536E-01, 1.01228536290370E-01}};

/* 8-point Gauss-Jacobi on (-1,1) with weight (1+x): abscissa, weight */
static const double ORC_GJ8[8][2] = {
    {-9.10732089420060E-01, 1.31807657689955E-02},
    {-7.11267485915709E-01, 7.13716106239446E-02},
    {-4.26350485711139E-01, 1.81757278018796E-01},
    {-9.03733696068533E-02, 3.16798397969277E-01},
    { 2.56135670833455E-01, 4.24189437743720E-01},
    { 5.71383041208738E-01, 4.50023197883551E-01},
    { 8.17352784200412E-01, 3.64476094545495E-01},
    { 9.64440169705273E-01, 1.78203217446225E-01}};

/* 7-point degree-5 triangle rule (L1, L2, weight); weights sum to 1/2.
 * The reference uses the same table for regular and singular cubature. */
static const double ORC_T7[7][3] = {
    {3.33333333333333E-01, 3.33333333333333E-01, 1.1250000000000E-01},
    {4.70142064105115E-01, 4.70142064105115E-01, 6.6197076394253E-02},
    {5.97158717897700E-02, 4.70142064105115E-01, 6.6197076394253E-02},
    {4.70142064105115E-01, 5.97158717897700E-02, 6.6197076394253E-02},
    {1.01286507323456E-01, 1.01286507323456E-01, 6.2969590272414E-02},
    {7.97426985353088E-01, 1.01286507323456E-01, 6.2969590272414E-02},
    {1.01286507323456E-01, 7.97426985353088E-01, 6.2969590272414E-02}};

#endif
