/*
 * wd2_oracle.c -- CPU restatement of the wolfd2 per-timestep hot path.
 *
 * TEST INFRASTRUCTURE ONLY (the parity checker; see wd2_oracle.h).  Built with
 * gcc -O2 -ffp-contract=off so that every expression is evaluated in the
 * reference's order with IEEE binary64 rounding and no fused multiply-add.
 *
 * Indexing: F(a,i,j) is Fortran a(i,j) of an array dimensioned (0:mnx,0:mny)
 * with ld = mnx+1.  Region-table helpers take Fortran (1-based) indices.
 *
 * Local work arrays of the reference (gfortran/flang static storage) are
 * calloc'ed per call.  Entries that the reference reads without writing in
 * the same call are never written by any call on the same grid, so they are
 * zero there too (cj1(nx+1,j), cj2(i,ny+1), ...).
 */
#include "wd2_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define F(a, i, j) (a)[(size_t)(i) + (size_t)ld * (size_t)(j)]

/* wolfd2.h:15-42 */
enum { RM_BLOCKG = 0, RM_INTERN = 1, RM_POROUS = 2 };
enum { BM_INTERN = 0, BM_WALL1 = 1, BM_WALL2 = 2, BM_INLET = 3, BM_OUTLT1 = 4, BM_OUTLT2 = 5 };
enum { WEST = 1, EAST = 2, SOUTH = 3, NORTH = 4 };
enum { VU = 1, VV = 2, VP = 3, VT = 4 };

#define ERR_FLAG (-100)

static inline int brd(const wd2o_regions *r, int ir, int jr, int side) {
    return r->nRegBrd[(ir - 1) + r->mgri * ((jr - 1) + r->mgrj * (side - 1))];
}
static inline int btp(const wd2o_regions *r, int ir, int jr, int side) {
    return r->nMomBdTp[(ir - 1) + r->mgri * ((jr - 1) + r->mgrj * (side - 1))];
}
static inline int rtype(const wd2o_regions *r, int ir, int jr) {
    return r->nRegType[(ir - 1) + r->mgri * (jr - 1)];
}
static inline double bcv(const wd2o_regions *r, int ir, int jr, int side, int var) {
    return r->dBCVal[(ir - 1) + r->mgri * ((jr - 1) + r->mgrj * ((side - 1) + 4 * (var - 1)))];
}
/* Boundary value on grid line k of a side (k = j on the west and east sides,
 * i on the south and north ones): the side's profile when it has one (an
 * extension, not in the reference), else the reference's constant. */
static inline double bcvk(const wd2o_regions *r, int ir, int jr, int side, int var, int k) {
    if (r->poff) {
        const int o = r->poff[(ir - 1) + r->mgri * ((jr - 1) + r->mgrj * ((side - 1) + 4 * (var - 1)))];
        if (o >= 0) return r->pval[o + k];
    }
    return bcv(r, ir, jr, side, var);
}
static inline double tab2(const wd2o_regions *r, const double *t, int ir, int jr) {
    return t[(ir - 1) + r->mgri * (jr - 1)];
}

static double *zalloc(size_t n) {
    double *p = (double *)calloc(n ? n : 1, sizeof(double));
    if (!p) {
        fprintf(stderr, "wd2_oracle: out of memory\n");
        abort();
    }
    return p;
}
#define FIELD_N(ld, ny) ((size_t)(ld) * (size_t)((ny) + 2))

static inline double dmax(double a, double b) { return a > b ? a : b; }

/* ------------------------------------------------------------------ grid.f */

void wd2o_grid(int nx, int ny, int ld, double dlref, double *gx, double *gy,
               double *rau, double *rbu, double *rbv, double *rgv,
               double *ran, double *rbn, double *rgn,
               double *rac, double *rbc, double *rgc,
               double *dju, double *djv, double *djc, double *djn,
               double *xen, double *yen, double *xzn, double *yzn,
               double *xec, double *yec, double *xzc, double *yzc,
               double *xeu, double *yeu, double *xzv, double *yzv,
               double *xzu, double *yzu, double *xev, double *yev)
{
    int i, j;
    const double dTwo = 2.0, dHalf = 0.5, dOne = 1.0, dQrtr = 0.25;
    size_t n = FIELD_N(ld, ny);
    double *xu = zalloc(n), *yu = zalloc(n), *xv = zalloc(n), *yv = zalloc(n);
    double *xc = zalloc(n), *yc = zalloc(n);
    double *x = gx, *y = gy;

    /* grid.f:Grid 100-106: scale with respect to the reference length */
    for (j = 1; j <= ny; j++)
        for (i = 1; i <= nx; i++) {
            F(x, i, j) = F(x, i, j) / dlref;
            F(y, i, j) = F(y, i, j) / dlref;
        }

    /* grid.f:MirrorPts 258-306 */
    for (i = 1; i <= nx; i++) {
        F(x, i, 0) = dTwo * F(x, i, 1) - F(x, i, 2);
        F(x, i, ny + 1) = dTwo * F(x, i, ny) - F(x, i, ny - 1);
        F(y, i, 0) = dTwo * F(y, i, 1) - F(y, i, 2);
        F(y, i, ny + 1) = dTwo * F(y, i, ny) - F(y, i, ny - 1);
    }
    for (j = 1; j <= ny; j++) {
        F(x, 0, j) = dTwo * F(x, 1, j) - F(x, 2, j);
        F(x, nx + 1, j) = dTwo * F(x, nx, j) - F(x, nx - 1, j);
        F(y, 0, j) = dTwo * F(y, 1, j) - F(y, 2, j);
        F(y, nx + 1, j) = dTwo * F(y, nx, j) - F(y, nx - 1, j);
    }
    F(x, 0, 0) = dTwo * F(x, 1, 1) - F(x, 2, 2);
    F(y, 0, 0) = dTwo * F(y, 1, 1) - F(y, 2, 2);
    F(x, 0, ny + 1) = dTwo * F(x, 1, ny) - F(x, 2, ny - 1);
    F(y, 0, ny + 1) = dTwo * F(y, 1, ny) - F(y, 2, ny - 1);
    F(x, nx + 1, 0) = dTwo * F(x, nx, 1) - F(x, nx - 1, 2);
    F(y, nx + 1, 0) = dTwo * F(y, nx, 1) - F(y, nx - 1, 2);
    F(x, nx + 1, ny + 1) = dTwo * F(x, nx, ny) - F(x, nx - 1, ny - 1);
    F(y, nx + 1, ny + 1) = dTwo * F(y, nx, ny) - F(y, nx - 1, ny - 1);

    /* grid.f:FullGrid 312-362 */
    for (j = 1; j <= ny + 1; j++)
        for (i = 0; i <= nx + 1; i++) {
            F(xu, i, j) = dHalf * (F(x, i, j) + F(x, i, j - 1));
            F(yu, i, j) = dHalf * (F(y, i, j) + F(y, i, j - 1));
        }
    for (j = 0; j <= ny + 1; j++)
        for (i = 1; i <= nx + 1; i++) {
            F(xv, i, j) = dHalf * (F(x, i, j) + F(x, i - 1, j));
            F(yv, i, j) = dHalf * (F(y, i, j) + F(y, i - 1, j));
        }
    for (j = 1; j <= ny + 1; j++)
        for (i = 1; i <= nx + 1; i++) {
            F(xc, i, j) = dHalf * (F(xu, i, j) + F(xu, i - 1, j));
            F(yc, i, j) = dHalf * (F(yv, i, j) + F(yv, i, j - 1));
        }

    /* grid.f:Metric 368-535 */
    for (j = 1; j <= ny; j++)
        for (i = 1; i <= nx; i++) {
            double g11, g12, g22;
            F(xzn, i, j) = F(xv, i + 1, j) - F(xv, i, j);
            F(xen, i, j) = F(xu, i, j + 1) - F(xu, i, j);
            F(yzn, i, j) = F(yv, i + 1, j) - F(yv, i, j);
            F(yen, i, j) = F(yu, i, j + 1) - F(yu, i, j);
            F(djn, i, j) = dOne / (F(xzn, i, j) * F(yen, i, j) - F(xen, i, j) * F(yzn, i, j));
            g11 = F(xzn, i, j) * F(xzn, i, j) + F(yzn, i, j) * F(yzn, i, j);
            g12 = F(xzn, i, j) * F(xen, i, j) + F(yzn, i, j) * F(yen, i, j);
            g22 = F(xen, i, j) * F(xen, i, j) + F(yen, i, j) * F(yen, i, j);
            F(ran, i, j) = F(djn, i, j) * g22;
            F(rbn, i, j) = -F(djn, i, j) * g12 * dQrtr;
            F(rgn, i, j) = F(djn, i, j) * g11;
        }
    for (j = 1; j <= ny; j++)
        for (i = 1; i <= nx; i++) {
            double g12, g22;
            F(xzu, i, j) = F(xc, i + 1, j) - F(xc, i, j);
            F(xeu, i, j) = F(x, i, j) - F(x, i, j - 1);
            F(yzu, i, j) = F(yc, i + 1, j) - F(yc, i, j);
            F(yeu, i, j) = F(y, i, j) - F(y, i, j - 1);
            F(dju, i, j) = dOne / (F(xzu, i, j) * F(yeu, i, j) - F(xeu, i, j) * F(yzu, i, j));
            g12 = F(xzu, i, j) * F(xeu, i, j) + F(yzu, i, j) * F(yeu, i, j);
            g22 = F(xeu, i, j) * F(xeu, i, j) + F(yeu, i, j) * F(yeu, i, j);
            F(rau, i, j) = F(dju, i, j) * g22;
            F(rbu, i, j) = -F(dju, i, j) * g12 * dQrtr;
        }
    for (j = 1; j <= ny; j++)
        for (i = 1; i <= nx; i++) {
            double g11, g12;
            F(xzv, i, j) = F(x, i, j) - F(x, i - 1, j);
            F(xev, i, j) = F(xc, i, j + 1) - F(xc, i, j);
            F(yzv, i, j) = F(y, i, j) - F(y, i - 1, j);
            F(yev, i, j) = F(yc, i, j + 1) - F(yc, i, j);
            F(djv, i, j) = dOne / (F(xzv, i, j) * F(yev, i, j) - F(xev, i, j) * F(yzv, i, j));
            g11 = F(xzv, i, j) * F(xzv, i, j) + F(yzv, i, j) * F(yzv, i, j);
            g12 = F(xzv, i, j) * F(xev, i, j) + F(yzv, i, j) * F(yev, i, j);
            F(rbv, i, j) = -F(djv, i, j) * g12 * dQrtr;
            F(rgv, i, j) = F(djv, i, j) * g11;
        }
    for (j = 1; j <= ny; j++)
        for (i = 1; i <= nx; i++) {
            double g11, g12, g22;
            F(xzc, i, j) = F(xu, i, j) - F(xu, i - 1, j);
            F(xec, i, j) = F(xv, i, j) - F(xv, i, j - 1);
            F(yzc, i, j) = F(yu, i, j) - F(yu, i - 1, j);
            F(yec, i, j) = F(yv, i, j) - F(yv, i, j - 1);
            F(djc, i, j) = dOne / (F(xzc, i, j) * F(yec, i, j) - F(xec, i, j) * F(yzc, i, j));
            g11 = F(xzc, i, j) * F(xzc, i, j) + F(yzc, i, j) * F(yzc, i, j);
            g12 = F(xzc, i, j) * F(xec, i, j) + F(yzc, i, j) * F(yec, i, j);
            g22 = F(xec, i, j) * F(xec, i, j) + F(yec, i, j) * F(yec, i, j);
            F(rac, i, j) = F(djc, i, j) * g22;
            F(rbc, i, j) = -F(djc, i, j) * g12 * dQrtr;
            F(rgc, i, j) = F(djc, i, j) * g11;
        }
    free(xu); free(yu); free(xv); free(yv); free(xc); free(yc);
}

/* -------------------------------------------------------------- momentum.f */

/* momentum.f:AltTridLU 1307-1339.  a(3,n) column-major: a[3*(k-1)+0..2] =
 * a(1..3,k).  Row 1's super-diagonal is divided by row 2's diagonal. */
void wd2o_alt_trid_lu(int n, double *a, double *b)
{
#define A(r, k) a[3 * ((k) - 1) + ((r) - 1)]
#define B(k) b[(k) - 1]
    int i;
    A(3, 1) = A(3, 1) / A(2, 2);
    B(1) = B(1) / A(2, 1);
    for (i = 2; i <= n - 1; i++) {
        A(2, i) = A(2, i) - (A(1, i) * A(3, i - 1));
        A(3, i) = A(3, i) / A(2, i);
        B(i) = (B(i) - A(1, i) * B(i - 1)) / A(2, i);
    }
    A(2, n) = A(2, n) - (A(1, n) * A(3, n - 1));
    B(n) = (B(n) - A(1, n) * B(n - 1)) / A(2, n);
    for (i = n - 1; i >= 1; i--)
        B(i) = B(i) - (A(3, i) * B(i + 1));
#undef A
#undef B
}

/* momentum.f:ConvCoef 864-981 */
int wd2o_conv_coef(int nx, int ny, int ld, int ncomp, int njacob,
                   const double *xzi, const double *xet, const double *yzi, const double *yet,
                   const double *u, const double *v, double *cc1, double *cc2)
{
    int i, j;
    const double dHalf = 0.5, dFour = 4.0;
    double djac = 1.0;
    if (njacob == 1) djac = 2.0;

    switch (ncomp) {
    case 1:
        for (j = 1; j <= ny + 1; j++)
            for (i = 1; i <= nx + 1; i++)
                F(cc1, i, j) = (djac * F(yet, i, j) * (F(u, i, j) + F(u, i - 1, j))
                                - F(xet, i, j) * (F(v, i, j) + F(v, i, j - 1))) * dHalf;
        for (j = 1; j <= ny; j++)
            for (i = 1; i <= nx; i++)
                F(cc2, i, j) = (F(xzi, i, j) * (F(v, i + 1, j) + F(v, i, j))
                                - djac * F(yzi, i, j) * (F(u, i, j + 1) + F(u, i, j))) * dHalf;
        break;
    case 2:
        for (j = 1; j <= ny; j++)
            for (i = 1; i <= nx; i++)
                F(cc1, i, j) = (F(yet, i, j) * (F(u, i, j + 1) + F(u, i, j))
                                - djac * F(xet, i, j) * (F(v, i + 1, j) + F(v, i, j))) * dHalf;
        for (j = 1; j <= ny + 1; j++)
            for (i = 1; i <= nx + 1; i++)
                F(cc2, i, j) = (djac * F(xzi, i, j) * (F(v, i, j) + F(v, i, j - 1))
                                - F(yzi, i, j) * (F(u, i, j) + F(u, i - 1, j))) * dHalf;
        break;
    case 3:
        for (j = 1; j <= ny; j++)
            for (i = 1; i <= nx; i++) {
                F(cc1, i, j) = (F(yet, i, j) * F(u, i, j)
                                - F(xet, i, j) * (F(v, i + 1, j) + F(v, i, j) + F(v, i + 1, j - 1) + F(v, i, j - 1)) / dFour) * dHalf;
                F(cc2, i, j) = (F(xzi, i, j) * F(v, i, j)
                                - F(yzi, i, j) * (F(u, i, j + 1) + F(u, i - 1, j + 1) + F(u, i, j) + F(u, i - 1, j)) / dFour) * dHalf;
            }
        break;
    case 4:
        for (j = 1; j <= ny; j++)
            for (i = 0; i <= nx; i++) {
                F(cc1, i, j) = djac * F(yet, i, j) * F(u, i, j)
                               - F(xet, i, j) * (F(v, i + 1, j) + F(v, i, j) + F(v, i + 1, j - 1) + F(v, i, j - 1)) / dFour;
                F(cc2, i, j) = F(xzi, i, j) * (F(v, i + 1, j) + F(v, i, j) + F(v, i + 1, j - 1) + F(v, i, j - 1)) / dFour
                               - djac * F(yzi, i, j) * F(u, i, j);
            }
        break;
    case 5:
        for (j = 0; j <= ny; j++)
            for (i = 1; i <= nx; i++) {
                F(cc1, i, j) = F(yet, i, j) * (F(u, i, j + 1) + F(u, i - 1, j + 1) + F(u, i, j) + F(u, i - 1, j)) / dFour
                               - djac * F(xet, i, j) * F(v, i, j);
                F(cc2, i, j) = djac * F(xzi, i, j) * F(v, i, j)
                               - F(yzi, i, j) * (F(u, i, j + 1) + F(u, i - 1, j + 1) + F(u, i, j) + F(u, i - 1, j)) / dFour;
            }
        break;
    case 6:
        for (j = 1; j <= ny + 1; j++)
            for (i = 1; i <= nx + 1; i++) {
                F(cc1, i, j) = (F(yet, i, j) * (F(u, i, j) + F(u, i - 1, j))
                                - F(xet, i, j) * (F(v, i, j) + F(v, i, j - 1))) * dHalf;
                F(cc2, i, j) = (F(xzi, i, j) * (F(v, i, j) + F(v, i, j - 1))
                                - F(yzi, i, j) * (F(u, i, j) + F(u, i - 1, j))) * dHalf;
            }
        break;
    default:
        fprintf(stderr, "Error: Wrong component flag passed to ConvCoef: %d\n", ncomp);
        return ERR_FLAG;
    }
    return 0;
}

/* momentum.f:DConvU 987-1009 */
static void dconv_u(int nx, int ny, int ld, const double *c1, const double *c2,
                    const double *u, double *c)
{
    int i, j;
    for (j = 2; j <= ny; j++)
        for (i = 1; i <= nx; i++)
            F(c, i, j) = -F(c2, i, j - 1) * F(u, i, j - 1) - F(c1, i, j) * F(u, i - 1, j)
                         + (F(c1, i + 1, j) - F(c1, i, j) + F(c2, i, j) - F(c2, i, j - 1)) * F(u, i, j)
                         + F(c1, i + 1, j) * F(u, i + 1, j) + F(c2, i, j) * F(u, i, j + 1);
}

/* momentum.f:DDiffU 1015-1045 */
static void ddiff_u(int nx, int ny, int ld, const double *ac, const double *bc,
                    const double *bn, const double *gn, const double *u, double *d)
{
    int i, j;
    for (j = 2; j <= ny; j++)
        for (i = 1; i <= nx; i++) {
            double s1, s2;
            s1 = F(ac, i + 1, j) * (F(u, i + 1, j) - F(u, i, j))
                 - F(ac, i, j) * (F(u, i, j) - F(u, i - 1, j))
                 + F(bc, i + 1, j) * (F(u, i + 1, j + 1) + F(u, i, j + 1) - F(u, i + 1, j - 1) - F(u, i, j - 1))
                 - F(bc, i, j) * (F(u, i, j + 1) + F(u, i - 1, j + 1) - F(u, i, j - 1) - F(u, i - 1, j - 1));
            s2 = F(bn, i, j) * (F(u, i + 1, j + 1) + F(u, i + 1, j) - F(u, i - 1, j + 1) - F(u, i - 1, j))
                 - F(bn, i, j - 1) * (F(u, i + 1, j) + F(u, i + 1, j - 1) - F(u, i - 1, j) - F(u, i - 1, j - 1))
                 + F(gn, i, j) * (F(u, i, j + 1) - F(u, i, j))
                 - F(gn, i, j - 1) * (F(u, i, j) - F(u, i, j - 1));
            F(d, i, j) = s1 + s2;
        }
}

/* momentum.f:DConvV 1051-1073 */
static void dconv_v(int nx, int ny, int ld, const double *c1, const double *c2,
                    const double *v, double *c)
{
    int i, j;
    for (j = 1; j <= ny; j++)
        for (i = 2; i <= nx; i++)
            F(c, i, j) = -F(c2, i, j) * F(v, i, j - 1) - F(c1, i - 1, j) * F(v, i - 1, j)
                         + (F(c1, i, j) - F(c1, i - 1, j) + F(c2, i, j + 1) - F(c2, i, j)) * F(v, i, j)
                         + F(c1, i, j) * F(v, i + 1, j) + F(c2, i, j + 1) * F(v, i, j + 1);
}

/* momentum.f:DDiffV 1079-1109 */
static void ddiff_v(int nx, int ny, int ld, const double *an, const double *bc,
                    const double *bn, const double *gc, const double *v, double *d)
{
    int i, j;
    for (j = 1; j <= ny; j++)
        for (i = 2; i <= nx; i++) {
            double s1, s2;
            s1 = F(an, i, j) * (F(v, i + 1, j) - F(v, i, j))
                 - F(an, i - 1, j) * (F(v, i, j) - F(v, i - 1, j))
                 + F(bn, i, j) * (F(v, i + 1, j + 1) + F(v, i, j + 1) - F(v, i + 1, j - 1) - F(v, i, j - 1))
                 - F(bn, i - 1, j) * (F(v, i, j + 1) + F(v, i - 1, j + 1) - F(v, i, j - 1) - F(v, i - 1, j - 1));
            s2 = F(bc, i, j + 1) * (F(v, i + 1, j + 1) + F(v, i + 1, j) - F(v, i - 1, j + 1) - F(v, i - 1, j))
                 - F(bc, i, j) * (F(v, i + 1, j) + F(v, i + 1, j - 1) - F(v, i - 1, j) - F(v, i - 1, j - 1))
                 + F(gc, i, j + 1) * (F(v, i, j + 1) - F(v, i, j))
                 - F(gc, i, j) * (F(v, i, j) - F(v, i, j - 1));
            F(d, i, j) = s1 + s2;
        }
}

/* momentum.f:PorosCoef 1115-1226 */
static int poros_coef(const wd2o_regions *r, int ncomp, int njacob,
                      const double *u, const double *v, double *cp)
{
    int ld = r->ld, ir, jr, i, j;
    const double dFour = 4.0;
    for (jr = 1; jr <= r->nReg[1]; jr++)
        for (ir = 1; ir <= r->nReg[0]; ir++) {
            int iW = brd(r, ir, jr, WEST), iE = brd(r, ir, jr, EAST);
            int jS = brd(r, ir, jr, SOUTH), jN = brd(r, ir, jr, NORTH);
            double porc1, porc2;
            if (rtype(r, ir, jr) != RM_POROUS) {
                for (j = jS; j <= jN; j++)
                    for (i = iW; i <= iE; i++) F(cp, i, j) = 0.0;
                continue;
            }
            porc1 = tab2(r, r->dPRporc1, ir, jr);
            porc2 = tab2(r, r->dPRporc2, ir, jr);
            switch (ncomp) {
            case 1:
                for (j = jS + 1; j <= jN; j++)
                    for (i = iW; i <= iE; i++) {
                        /* (v+v+v+v/4)**2 exactly as written (precedence of /) */
                        double w = F(v, i, j) + F(v, i + 1, j) + F(v, i, j - 1) + F(v, i + 1, j - 1) / dFour;
                        double unorm = sqrt(F(u, i, j) * F(u, i, j) + w * w);
                        double unrm1 = 0.0;
                        if (unorm > 1.e-8) unrm1 = (F(u, i, j) * F(u, i, j)) / unorm;
                        if (njacob == 1) unorm = unrm1 + unorm;
                        F(cp, i, j) = porc1 + porc2 * unorm;
                    }
                break;
            case 2:
                for (j = jS; j <= jN; j++)
                    for (i = iW + 1; i <= iE; i++) {
                        double w = (F(u, i - 1, j + 1) + F(u, i, j + 1) + F(u, i - 1, j) + F(u, i, j)) / dFour;
                        double unorm = sqrt(w * w + F(v, i, j) * F(v, i, j));
                        double unrm1 = 0.0;
                        if (unorm > 1.e-8) unrm1 = (F(v, i, j) * F(v, i, j)) / unorm;
                        if (njacob == 1) unorm = unrm1 + unorm;
                        F(cp, i, j) = porc1 + porc2 * unorm;
                    }
                break;
            default:
                fprintf(stderr, "Error: Wrong ncomp flag passed to PorosCoef: %d\n", ncomp);
                return ERR_FLAG;
            }
        }
    return 0;
}

/* Identity row a=(0,1,0), b=0 in a chain (1-based ind). */
static inline void ident_row(double *a, double *b, int ind)
{
    a[3 * (ind - 1) + 0] = 0.0;
    a[3 * (ind - 1) + 1] = 1.0;
    a[3 * (ind - 1) + 2] = 0.0;
    b[ind - 1] = 0.0;
}

/* momentum.f:XMomentum 199-514 */
int wd2o_x_momentum(const wd2o_regions *r, double dk, double re, const wd2o_metrics *m,
                    const double *us, const double *vs, const double *un, const double *vn,
                    double *dus)
{
    int nx = r->nx, ny = r->ny, ld = r->ld, i, j, ind, ir, jr, st = 0;
    const double dZero = 0.0, dOne = 1.0, dHalf = 0.5;
    size_t n = FIELD_N(ld, ny);
    int nchain = nx * (ny - 1);
    double re1 = dOne / re, dk2 = dk * dHalf;
    double *cj1 = zalloc(n), *cj2 = zalloc(n), *c1s = zalloc(n), *c2s = zalloc(n);
    double *c1n = zalloc(n), *c2n = zalloc(n), *cps = zalloc(n), *cpn = zalloc(n), *cpj = zalloc(n);
    double *cnvs = zalloc(n), *difs = zalloc(n), *cnvn = zalloc(n), *difn = zalloc(n);
    double *a = zalloc(3 * (size_t)nchain), *b = zalloc((size_t)nchain);

    wd2o_conv_coef(nx, ny, ld, 4, 1, m->xzu, m->xeu, m->yzu, m->yeu, us, vs, cj1, cj2);
    wd2o_conv_coef(nx, ny, ld, 1, 0, m->xzn, m->xec, m->yzn, m->yec, us, vs, c1s, c2s);
    wd2o_conv_coef(nx, ny, ld, 1, 0, m->xzn, m->xec, m->yzn, m->yec, un, vn, c1n, c2n);

    /* momentum.f:295-324: porous regions divide the convective terms */
    for (jr = 1; jr <= r->nReg[1]; jr++)
        for (ir = 1; ir <= r->nReg[0]; ir++) {
            int iW = brd(r, ir, jr, WEST), iE = brd(r, ir, jr, EAST);
            int jS = brd(r, ir, jr, SOUTH), jN = brd(r, ir, jr, NORTH);
            if (rtype(r, ir, jr) == RM_POROUS) {
                double por = tab2(r, r->dPRporos, ir, jr);
                for (j = jS + 1; j <= jN; j++)
                    for (i = iW; i <= iE; i++) {
                        F(cj1, i, j) = F(cj1, i, j) / por;
                        F(cj2, i, j) = F(cj2, i, j) / por;
                        F(c1s, i, j) = F(c1s, i, j) / por;
                        F(c2s, i, j) = F(c2s, i, j) / por;
                        F(c1n, i, j) = F(c1n, i, j) / por;
                        F(c2n, i, j) = F(c2n, i, j) / por;
                    }
            }
        }

    dconv_u(nx, ny, ld, c1s, c2s, us, cnvs);
    dconv_u(nx, ny, ld, c1n, c2n, un, cnvn);
    ddiff_u(nx, ny, ld, m->rac, m->rbc, m->rbn, m->rgn, us, difs);
    ddiff_u(nx, ny, ld, m->rac, m->rbc, m->rbn, m->rgn, un, difn);
    if ((st = poros_coef(r, 1, 1, us, vs, cpj)) != 0) goto out;
    if ((st = poros_coef(r, 1, 0, us, vs, cps)) != 0) goto out;
    if ((st = poros_coef(r, 1, 0, un, vn, cpn)) != 0) goto out;

    /* First split step, momentum.f:350-384 */
    for (j = 2; j <= ny; j++)
        for (i = 1; i <= nx; i++) {
            double rkj, *ar;
            ind = (j - 2) * nx + i;
            rkj = dk2 * F(m->dju, i, j);
            ar = &a[3 * (ind - 1)];
            if (F(cj1, i, j) >= dZero) {
                ar[0] = rkj * (-F(cj1, i - 1, j) - re1 * F(m->rac, i, j));
                ar[1] = dOne + rkj * (F(cj1, i, j) + re1 * (F(m->rac, i + 1, j) + F(m->rac, i, j))) + dk2 * F(cpj, i, j);
                ar[2] = rkj * (-re1 * F(m->rac, i + 1, j));
            } else {
                ar[0] = rkj * (-re1 * F(m->rac, i, j));
                ar[1] = dOne + rkj * (-F(cj1, i, j) + re1 * (F(m->rac, i + 1, j) + F(m->rac, i, j))) + dk2 * F(cpj, i, j);
                ar[2] = rkj * (F(cj1, i + 1, j) - re1 * F(m->rac, i + 1, j));
            }
            b[ind - 1] = F(un, i, j) - F(us, i, j) + rkj * (-F(cnvs, i, j) - F(cnvn, i, j))
                         + rkj * re1 * (F(difs, i, j) + F(difn, i, j))
                         - dk2 * (F(cps, i, j) * F(us, i, j) + F(cpn, i, j) * F(un, i, j));
        }
    wd2o_alt_trid_lu(nchain, a, b);

    /* Second split step, momentum.f:396-428 (y coefficients, x-ordered chain) */
    for (j = 2; j <= ny; j++)
        for (i = 1; i <= nx; i++) {
            double rkj, *ar;
            ind = (j - 2) * nx + i;
            rkj = dk2 * F(m->dju, i, j);
            ar = &a[3 * (ind - 1)];
            if (F(cj2, i, j) >= dZero) {
                ar[0] = rkj * (-F(cj2, i, j - 1) - re1 * F(m->rgn, i, j - 1));
                ar[1] = dOne + rkj * (F(cj2, i, j) + re1 * (F(m->rgn, i, j) + F(m->rgn, i, j - 1))) + dk2 * F(cpj, i, j);
                ar[2] = rkj * (-re1 * F(m->rgn, i, j));
            } else {
                ar[0] = rkj * (-re1 * F(m->rgn, i, j - 1));
                ar[1] = dOne + rkj * (-F(cj2, i, j) + re1 * (F(m->rgn, i, j) + F(m->rgn, i, j - 1))) + dk2 * F(cpj, i, j);
                ar[2] = rkj * (F(cj2, i, j + 1) - re1 * F(m->rgn, i, j));
            }
        }

    /* Blockages and wall/inlet rows, momentum.f:432-496 */
    for (jr = 1; jr <= r->nReg[1]; jr++)
        for (ir = 1; ir <= r->nReg[0]; ir++) {
            int iW = brd(r, ir, jr, WEST), iE = brd(r, ir, jr, EAST);
            int jS = brd(r, ir, jr, SOUTH), jN = brd(r, ir, jr, NORTH);
            if (rtype(r, ir, jr) == RM_BLOCKG)
                for (j = jS + 1; j <= jN; j++)
                    for (i = iW; i <= iE; i++) ident_row(a, b, (j - 2) * nx + i);
            switch (btp(r, ir, jr, WEST)) {
            case BM_INTERN: case BM_OUTLT1: case BM_OUTLT2: break;
            case BM_WALL1: case BM_WALL2: case BM_INLET:
                for (j = jS + 1; j <= jN; j++) ident_row(a, b, (j - 2) * nx + iW);
                break;
            default:
                fprintf(stderr, "Wrong nBdTypeW flag in region %d,%d\n", ir, jr);
                st = ERR_FLAG; goto out;
            }
            switch (btp(r, ir, jr, EAST)) {
            case BM_INTERN: case BM_OUTLT1: case BM_OUTLT2: break;
            case BM_WALL1: case BM_WALL2: case BM_INLET:
                for (j = jS + 1; j <= jN; j++) ident_row(a, b, (j - 2) * nx + iE);
                break;
            default:
                fprintf(stderr, "Wrong nBdTypeE flag in region %d,%d\n", ir, jr);
                st = ERR_FLAG; goto out;
            }
        }
    wd2o_alt_trid_lu(nchain, a, b);

    for (j = 2; j <= ny; j++)
        for (i = 1; i <= nx; i++) F(dus, i, j) = b[(j - 2) * nx + i - 1];
out:
    free(cj1); free(cj2); free(c1s); free(c2s); free(c1n); free(c2n);
    free(cps); free(cpn); free(cpj); free(cnvs); free(difs); free(cnvn); free(difn);
    free(a); free(b);
    return st;
}

/* momentum.f:YMomentum 520-838 */
int wd2o_y_momentum(const wd2o_regions *r, double dk, double re, double fr,
                    const wd2o_metrics *m, const double *d, const double *dn,
                    const double *us, const double *vs, const double *un, const double *vn,
                    double *dvs)
{
    int nx = r->nx, ny = r->ny, ld = r->ld, i, j, ind, ir, jr, st = 0;
    const double dZero = 0.0, dOne = 1.0, dFour = 4.0, dHalf = 0.5;
    size_t n = FIELD_N(ld, ny);
    int nchain = (nx - 1) * ny;
    double re1 = dOne / re, dk2 = dk * dHalf;
    double *cj1 = zalloc(n), *cj2 = zalloc(n), *c1s = zalloc(n), *c2s = zalloc(n);
    double *c1n = zalloc(n), *c2n = zalloc(n), *cps = zalloc(n), *cpn = zalloc(n), *cpj = zalloc(n);
    double *cnvs = zalloc(n), *difs = zalloc(n), *cnvn = zalloc(n), *difn = zalloc(n);
    double *a = zalloc(3 * (size_t)nchain), *b = zalloc((size_t)nchain);

    wd2o_conv_coef(nx, ny, ld, 5, 1, m->xzv, m->xev, m->yzv, m->yev, us, vs, cj1, cj2);
    wd2o_conv_coef(nx, ny, ld, 2, 0, m->xzc, m->xen, m->yzc, m->yen, us, vs, c1s, c2s);
    wd2o_conv_coef(nx, ny, ld, 2, 0, m->xzc, m->xen, m->yzc, m->yen, un, vn, c1n, c2n);

    /* momentum.f:618-647 */
    for (jr = 1; jr <= r->nReg[1]; jr++)
        for (ir = 1; ir <= r->nReg[0]; ir++) {
            int iW = brd(r, ir, jr, WEST), iE = brd(r, ir, jr, EAST);
            int jS = brd(r, ir, jr, SOUTH), jN = brd(r, ir, jr, NORTH);
            if (rtype(r, ir, jr) == RM_POROUS) {
                double por = tab2(r, r->dPRporos, ir, jr);
                for (j = jS; j <= jN; j++)
                    for (i = iW + 1; i <= iE; i++) {
                        F(cj1, i, j) = F(cj1, i, j) / por;
                        F(cj2, i, j) = F(cj2, i, j) / por;
                        F(c1s, i, j) = F(c1s, i, j) / por;
                        F(c2s, i, j) = F(c2s, i, j) / por;
                        F(c1n, i, j) = F(c1n, i, j) / por;
                        F(c2n, i, j) = F(c2n, i, j) / por;
                    }
            }
        }

    dconv_v(nx, ny, ld, c1s, c2s, vs, cnvs);
    dconv_v(nx, ny, ld, c1n, c2n, vn, cnvn);
    ddiff_v(nx, ny, ld, m->ran, m->rbc, m->rbn, m->rgc, vs, difs);
    ddiff_v(nx, ny, ld, m->ran, m->rbc, m->rbn, m->rgc, vn, difn);
    if ((st = poros_coef(r, 2, 1, us, vs, cpj)) != 0) goto out;
    if ((st = poros_coef(r, 2, 0, us, vs, cps)) != 0) goto out;
    if ((st = poros_coef(r, 2, 0, un, vn, cpn)) != 0) goto out;

    /* First split step, momentum.f:675-711 */
    for (j = 1; j <= ny; j++)
        for (i = 2; i <= nx; i++) {
            double rkj, buoy, *ar;
            ind = (j - 1) * (nx - 1) + i - 1;
            rkj = dk2 * F(m->djv, i, j);
            ar = &a[3 * (ind - 1)];
            if (F(cj1, i, j) >= dZero) {
                ar[0] = rkj * (-F(cj1, i - 1, j) - re1 * F(m->ran, i - 1, j));
                ar[1] = rkj * (F(cj1, i, j) + re1 * (F(m->ran, i, j) + F(m->ran, i - 1, j))) + dk2 * F(cpj, i, j) + dOne;
                ar[2] = rkj * (-re1 * F(m->ran, i, j));
            } else {
                ar[0] = rkj * (-re1 * F(m->ran, i - 1, j));
                ar[1] = rkj * (-F(cj1, i, j) + re1 * (F(m->ran, i, j) + F(m->ran, i - 1, j))) + dk2 * F(cpj, i, j) + dOne;
                ar[2] = rkj * (F(cj1, i + 1, j) - re1 * F(m->ran, i, j));
            }
            buoy = dk * (F(d, i, j + 1) + F(d, i, j) + F(dn, i, j + 1) + F(dn, i, j)) / (dFour * fr);
            b[ind - 1] = F(vn, i, j) - F(vs, i, j) + rkj * (-F(cnvs, i, j) - F(cnvn, i, j))
                         + rkj * re1 * (F(difs, i, j) + F(difn, i, j))
                         - dk2 * (F(cps, i, j) * F(vs, i, j) + F(cpn, i, j) * F(vn, i, j))
                         - buoy;
        }
    wd2o_alt_trid_lu(nchain, a, b);

    /* Second split step, momentum.f:723-754 */
    for (j = 1; j <= ny; j++)
        for (i = 2; i <= nx; i++) {
            double rkj, *ar;
            ind = (j - 1) * (nx - 1) + i - 1;
            rkj = dk2 * F(m->djv, i, j);
            ar = &a[3 * (ind - 1)];
            if (F(cj2, i, j) >= dZero) {
                ar[0] = rkj * (-F(cj2, i, j - 1) - re1 * F(m->rgc, i, j));
                ar[1] = rkj * (F(cj2, i, j) + re1 * (F(m->rgc, i, j + 1) + F(m->rgc, i, j))) + dk2 * F(cpj, i, j) + dOne;
                ar[2] = rkj * (-re1 * F(m->rgc, i, j + 1));
            } else {
                ar[0] = rkj * (-re1 * F(m->rgc, i, j));
                ar[1] = rkj * (-F(cj2, i, j) + re1 * (F(m->rgc, i, j + 1) + F(m->rgc, i, j))) + dk2 * F(cpj, i, j) + dOne;
                ar[2] = rkj * (F(cj2, i, j + 1) - re1 * F(m->rgc, i, j + 1));
            }
        }

    /* momentum.f:758-821 */
    for (jr = 1; jr <= r->nReg[1]; jr++)
        for (ir = 1; ir <= r->nReg[0]; ir++) {
            int iW = brd(r, ir, jr, WEST), iE = brd(r, ir, jr, EAST);
            int jS = brd(r, ir, jr, SOUTH), jN = brd(r, ir, jr, NORTH);
            if (rtype(r, ir, jr) == RM_BLOCKG)
                for (j = jS; j <= jN; j++)
                    for (i = iW + 1; i <= iE; i++) ident_row(a, b, (j - 1) * (nx - 1) + i - 1);
            switch (btp(r, ir, jr, SOUTH)) {
            case BM_INTERN: case BM_OUTLT1: case BM_OUTLT2: break;
            case BM_WALL1: case BM_WALL2: case BM_INLET:
                for (i = iW + 1; i <= iE; i++) ident_row(a, b, (jS - 1) * (nx - 1) + i - 1);
                break;
            default:
                fprintf(stderr, "Wrong nBdTypeS flag in region %d,%d\n", ir, jr);
                st = ERR_FLAG; goto out;
            }
            switch (btp(r, ir, jr, NORTH)) {
            case BM_INTERN: case BM_OUTLT1: case BM_OUTLT2: break;
            case BM_WALL1: case BM_WALL2: case BM_INLET:
                for (i = iW + 1; i <= iE; i++) ident_row(a, b, (jN - 1) * (nx - 1) + i - 1);
                break;
            default:
                fprintf(stderr, "Wrong nBdTypeN flag in region %d,%d\n", ir, jr);
                st = ERR_FLAG; goto out;
            }
        }
    wd2o_alt_trid_lu(nchain, a, b);

    for (j = 1; j <= ny; j++)
        for (i = 2; i <= nx; i++) F(dvs, i, j) = b[(j - 1) * (nx - 1) + i - 2];
out:
    free(cj1); free(cj2); free(c1s); free(c2s); free(c1n); free(c2n);
    free(cps); free(cpn); free(cpj); free(cnvs); free(difs); free(cnvn); free(difn);
    free(a); free(b);
    return st;
}

/* --------------------------------------------------------------- utility.f */

double wd2o_dmax_norm(int nx, int ny, int ld, const double *u)
{
    int i, j;
    double diff = fabs(F(u, 5, 5));
    for (j = 2; j <= ny - 1; j++)
        for (i = 2; i <= nx - 1; i++) diff = dmax(diff, fabs(F(u, i, j)));
    return diff;
}

double wd2o_diff_max_norm(int nx, int ny, int ld, const double *un, const double *u)
{
    int i, j;
    double diff = fabs(F(un, 2, 2) - F(u, 2, 2));
    for (j = 2; j <= ny - 1; j++)
        for (i = 2; i <= nx - 1; i++) diff = dmax(diff, fabs(F(un, i, j) - F(u, i, j)));
    return diff;
}

/* momentum.f:nAuxMomentum 33-193 */
int wd2o_naux_momentum(const wd2o_regions *r, int mqiter, double dk, double re, double fr,
                       double qtol, const wd2o_metrics *m, const double *d, const double *dn,
                       const double *un, const double *vn, double *us, double *vs)
{
    int nx = r->nx, ny = r->ny, ld = r->ld, i, j, it, st, ret = -1;
    size_t n = FIELD_N(ld, ny);
    double *dus = zalloc(n), *dvs = zalloc(n);

    for (j = 1; j <= ny + 1; j++)
        for (i = 1; i <= nx + 1; i++) {
            F(us, i, j) = F(un, i, j);
            F(vs, i, j) = F(vn, i, j);
        }
    for (it = 1; it <= mqiter; it++) {
        double d1, d2, difmax;
        if ((st = wd2o_vel_outflow_bcs(r, us, vs)) != 0) { ret = st; break; }
        for (j = 0; j <= ny + 1; j++)
            for (i = 0; i <= nx + 1; i++) {
                F(dus, i, j) = 0.0;
                F(dvs, i, j) = 0.0;
            }
        if ((st = wd2o_x_momentum(r, dk, re, m, us, vs, un, vn, dus)) != 0) { ret = st; break; }
        if ((st = wd2o_y_momentum(r, dk, re, fr, m, d, dn, us, vs, un, vn, dvs)) != 0) { ret = st; break; }
        for (j = 1; j <= ny; j++)
            for (i = 1; i <= nx; i++) {
                F(us, i, j) = F(us, i, j) + F(dus, i, j);
                F(vs, i, j) = F(vs, i, j) + F(dvs, i, j);
            }
        d1 = wd2o_dmax_norm(nx, ny, ld, dus);
        d2 = wd2o_dmax_norm(nx, ny, ld, dvs);
        difmax = dmax(d1, d2);
        if (difmax <= qtol) { ret = it; break; }
    }
    free(dus); free(dvs);
    return ret;
}

/* ------------------------------------------------------------ bound_cond.f */

int wd2o_vel_bound_cond(const wd2o_regions *r, double *u, double *v)
{
    int ld = r->ld, ir, jr, i, j;
    const double dTwo = 2.0, dThree = 3.0, dFour = 4.0, dFive = 5.0, dEight = 8.0;
    for (jr = 1; jr <= r->nReg[1]; jr++)
        for (ir = 1; ir <= r->nReg[0]; ir++) {
            int iW = brd(r, ir, jr, WEST), iE = brd(r, ir, jr, EAST);
            int jS = brd(r, ir, jr, SOUTH), jN = brd(r, ir, jr, NORTH);
            switch (btp(r, ir, jr, WEST)) {
            case BM_INTERN: break;
            case BM_WALL1:
                for (j = jS; j <= jN; j++) F(u, iW, j) = 0.0;
                for (j = jS + 1; j <= jN; j++) F(v, iW, j) = dTwo * bcvk(r, ir, jr, WEST, VV, j) - F(v, iW + 1, j);
                break;
            case BM_WALL2:
                for (j = jS; j <= jN; j++) F(u, iW, j) = 0.0;
                for (j = jS + 1; j <= jN; j++) F(v, iW, j) = F(v, iW + 1, j);
                break;
            case BM_INLET:
                for (j = jS; j <= jN; j++) F(u, iW, j) = bcvk(r, ir, jr, WEST, VU, j);
                for (j = jS + 1; j <= jN; j++) F(v, iW, j) = dTwo * bcvk(r, ir, jr, WEST, VV, j) - F(v, iW + 1, j);
                break;
            case BM_OUTLT1:
                for (j = jS; j <= jN; j++) F(u, iW - 1, j) = bcvk(r, ir, jr, WEST, VU, j) + F(u, iW, j);
                for (j = jS + 1; j <= jN; j++) F(v, iW, j) = -F(v, iW + 1, j);
                break;
            case BM_OUTLT2:
                for (j = jS + 1; j <= jN; j++) F(u, iW, j) = F(u, iW + 1, j) + F(v, iW + 1, j) - F(v, iW + 1, j - 1);
                for (j = jS + 1; j <= jN; j++)
                    F(v, iW, j) = -F(v, iW, j - 1) + dFive * (F(v, iW + 1, j) - F(v, iW + 1, j - 1))
                                  + dEight * (F(u, iW + 1, j) - F(u, iW, j));
                break;
            default:
                fprintf(stderr, "Wrong nBdTypeW flag in region %d,%d\n", ir, jr);
                return ERR_FLAG;
            }
            switch (btp(r, ir, jr, EAST)) {
            case BM_INTERN: break;
            case BM_WALL1:
                for (j = jS; j <= jN; j++) F(u, iE, j) = 0.0;
                for (j = jS + 1; j <= jN; j++) F(v, iE + 1, j) = dTwo * bcvk(r, ir, jr, EAST, VV, j) - F(v, iE, j);
                break;
            case BM_WALL2:
                for (j = jS; j <= jN; j++) F(u, iE, j) = 0.0;
                for (j = jS + 1; j <= jN; j++) F(v, iE + 1, j) = F(v, iE, j);
                break;
            case BM_INLET:
                for (j = jS; j <= jN; j++) F(u, iE, j) = bcvk(r, ir, jr, EAST, VU, j);
                for (j = jS + 1; j <= jN; j++) F(v, iE + 1, j) = dTwo * bcvk(r, ir, jr, EAST, VV, j) - F(v, iE, j);
                break;
            case BM_OUTLT1:
                for (j = jS; j <= jN; j++) F(u, iE + 1, j) = bcvk(r, ir, jr, EAST, VU, j) + F(u, iE, j);
                for (j = jS + 1; j <= jN; j++) F(v, iE + 1, j) = -F(v, iE, j);
                break;
            case BM_OUTLT2:
                for (j = jS + 1; j <= jN; j++) F(u, iE, j) = F(u, iE - 1, j) - (F(v, iE, j) - F(v, iE, j - 1));
                for (j = jS + 1; j <= jN - 1; j++)
                    F(v, iE + 1, j) = F(v, iE + 1, j - 1) + dThree * (F(v, iE, j - 1) - F(v, iE, j))
                                      - dFour * (F(u, iE, j) - F(u, iE - 1, j));
                break;
            default:
                fprintf(stderr, "Wrong nBdTypeE flag in region %d,%d\n", ir, jr);
                return ERR_FLAG;
            }
            switch (btp(r, ir, jr, SOUTH)) {
            case BM_INTERN: break;
            case BM_WALL1:
                for (i = iW + 1; i <= iE; i++) F(u, i, jS) = dTwo * bcvk(r, ir, jr, SOUTH, VU, i) - F(u, i, jS + 1);
                for (i = iW; i <= iE; i++) F(v, i, jS) = 0.0;
                break;
            case BM_WALL2:
                for (i = iW + 1; i <= iE; i++) F(u, i, jS) = F(u, i, jS + 1);
                for (i = iW; i <= iE; i++) F(v, i, jS) = 0.0;
                break;
            case BM_INLET:
                for (i = iW + 1; i <= iE; i++) F(u, i, jS) = dTwo * bcvk(r, ir, jr, SOUTH, VU, i) - F(u, i, jS + 1);
                for (i = iW; i <= iE; i++) F(v, i, jS) = bcvk(r, ir, jr, SOUTH, VV, i);
                break;
            case BM_OUTLT1:
                for (i = iW + 1; i <= iE; i++) F(u, i, jS) = -F(u, i, jS + 1);
                for (i = iW; i <= iE; i++) F(v, i, jS) = bcvk(r, ir, jr, SOUTH, VV, i) + F(v, i, jS);
                break;
            case BM_OUTLT2:
                for (i = iW + 1; i <= iE; i++) F(v, i, jS) = F(v, i, jS + 1) + (F(u, i, jS + 1) - F(u, i - 1, jS + 1));
                for (i = iW + 1; i <= iE - 1; i++)
                    F(u, i, jS) = F(u, i - 1, jS) + dThree * (F(u, i - 1, jS + 1) - F(u, i, jS + 1))
                                  - dFour * (F(v, i, jS + 1) - F(v, i, jS));
                break;
            default:
                fprintf(stderr, "Wrong nBdTypeS flag in region %d,%d\n", ir, jr);
                return ERR_FLAG;
            }
            switch (btp(r, ir, jr, NORTH)) {
            case BM_INTERN: break;
            case BM_WALL1:
                for (i = iW + 1; i <= iE; i++) F(u, i, jN + 1) = dTwo * bcvk(r, ir, jr, NORTH, VU, i) - F(u, i, jN);
                for (i = iW; i <= iE; i++) F(v, i, jN) = 0.0;
                break;
            case BM_WALL2:
                for (i = iW + 1; i <= iE; i++) F(u, i, jN + 1) = F(u, i, jN);
                for (i = iW; i <= iE; i++) F(v, i, jN) = 0.0;
                break;
            case BM_INLET:
                for (i = iW + 1; i <= iE; i++) F(u, i, jN + 1) = dTwo * bcvk(r, ir, jr, NORTH, VU, i) - F(u, i, jN);
                for (i = iW; i <= iE; i++) F(v, i, jN) = bcvk(r, ir, jr, NORTH, VV, i);
                break;
            case BM_OUTLT1:
                for (i = iW + 1; i <= iE; i++) F(u, i, jN + 1) = -F(u, i, jN);
                for (i = iW; i <= iE; i++) F(v, i, jN + 1) = bcvk(r, ir, jr, NORTH, VV, i) + F(v, i, jN);
                break;
            case BM_OUTLT2:
                for (i = iW; i <= iE; i++) F(v, i, jN) = F(v, i, jN - 1) - (F(u, i, jN) - F(u, i - 1, jN));
                for (i = iW + 1; i <= iE - 1; i++)
                    F(u, i, jN + 1) = F(u, i - 1, jN + 1) + dThree * (F(u, i - 1, jN) - F(u, i, jN))
                                      - dFour * (F(v, i, jN) - F(v, i, jN - 1));
                break;
            default:
                fprintf(stderr, "Wrong nBdTypeN flag in region %d,%d\n", ir, jr);
                return ERR_FLAG;
            }
        }
    return 0;
}

int wd2o_pres_bound_cond(const wd2o_regions *r, double *p)
{
    int ld = r->ld, ir, jr, i, j;
    const double dTwo = 2.0;
    for (jr = 1; jr <= r->nReg[1]; jr++)
        for (ir = 1; ir <= r->nReg[0]; ir++) {
            int iW = brd(r, ir, jr, WEST), iE = brd(r, ir, jr, EAST);
            int jS = brd(r, ir, jr, SOUTH), jN = brd(r, ir, jr, NORTH);
            if (rtype(r, ir, jr) == RM_BLOCKG) {
                for (j = jS + 1; j <= jN; j++)
                    for (i = iW + 1; i <= iE; i++) F(p, i, j) = 0.0;
                for (j = jS + 1; j <= jN; j++) F(p, iW + 1, j) = bcvk(r, ir, jr, WEST, VP, j) + F(p, iW, j);
                for (j = jS + 1; j <= jN; j++) F(p, iE, j) = bcvk(r, ir, jr, EAST, VP, j) + F(p, iE + 1, j);
                for (i = iW + 1; i <= iE; i++) F(p, i, jS + 1) = bcvk(r, ir, jr, SOUTH, VP, i) + F(p, i, jS);
                for (i = iW + 1; i <= iE; i++) F(p, i, jN) = bcvk(r, ir, jr, NORTH, VP, i) + F(p, i, jN + 1);
                continue;
            }
            switch (btp(r, ir, jr, WEST)) {
            case BM_INTERN: break;
            case BM_WALL1: case BM_WALL2: case BM_INLET:
                for (j = jS + 1; j <= jN; j++) F(p, iW, j) = bcvk(r, ir, jr, WEST, VP, j) + F(p, iW + 1, j);
                break;
            case BM_OUTLT1: case BM_OUTLT2:
                for (j = jS + 1; j <= jN; j++) F(p, iW, j) = dTwo * bcvk(r, ir, jr, WEST, VP, j) - F(p, iW + 1, j);
                break;
            default:
                fprintf(stderr, "Wrong nBdTypeW flag in region %d,%d\n", ir, jr);
                return ERR_FLAG;
            }
            switch (btp(r, ir, jr, EAST)) {
            case BM_INTERN: break;
            case BM_WALL1: case BM_WALL2: case BM_INLET:
                for (j = jS + 1; j <= jN; j++) F(p, iE + 1, j) = bcvk(r, ir, jr, EAST, VP, j) + F(p, iE, j);
                break;
            case BM_OUTLT1: case BM_OUTLT2:
                for (j = jS + 1; j <= jN; j++) F(p, iE + 1, j) = dTwo * bcvk(r, ir, jr, EAST, VP, j) - F(p, iE, j);
                break;
            default:
                fprintf(stderr, "Wrong nBdTypeE flag in region %d,%d\n", ir, jr);
                return ERR_FLAG;
            }
            switch (btp(r, ir, jr, SOUTH)) {
            case BM_INTERN: break;
            case BM_WALL1: case BM_WALL2: case BM_INLET:
                for (i = iW + 1; i <= iE; i++) F(p, i, jS) = bcvk(r, ir, jr, SOUTH, VP, i) + F(p, i, jS + 1);
                break;
            case BM_OUTLT1: case BM_OUTLT2:
                for (i = iW + 1; i <= iE; i++) F(p, i, jS) = dTwo * bcvk(r, ir, jr, SOUTH, VP, i) - F(p, i, jS + 1);
                break;
            default:
                fprintf(stderr, "Wrong nBdTypeS flag in region %d,%d\n", ir, jr);
                return ERR_FLAG;
            }
            switch (btp(r, ir, jr, NORTH)) {
            case BM_INTERN: break;
            case BM_WALL1: case BM_WALL2: case BM_INLET:
                for (i = iW + 1; i <= iE; i++) F(p, i, jN + 1) = bcvk(r, ir, jr, NORTH, VP, i) + F(p, i, jN);
                break;
            case BM_OUTLT1: case BM_OUTLT2:
                for (i = iW + 1; i <= iE; i++) F(p, i, jN + 1) = dTwo * bcvk(r, ir, jr, NORTH, VP, i) - F(p, i, jN);
                break;
            default:
                fprintf(stderr, "Wrong nBdTypeN flag in region %d,%d\n", ir, jr);
                return ERR_FLAG;
            }
        }
    return 0;
}

int wd2o_vel_outflow_bcs(const wd2o_regions *r, double *u, double *v)
{
    int ld = r->ld, ir, jr, i, j;
    const double dThree = 3.0, dFour = 4.0, dFive = 5.0, dEight = 8.0;
    for (jr = 1; jr <= r->nReg[1]; jr++)
        for (ir = 1; ir <= r->nReg[0]; ir++) {
            int iW = brd(r, ir, jr, WEST), iE = brd(r, ir, jr, EAST);
            int jS = brd(r, ir, jr, SOUTH), jN = brd(r, ir, jr, NORTH);
            switch (btp(r, ir, jr, WEST)) {
            case BM_INTERN: case BM_WALL1: case BM_WALL2: case BM_INLET: break;
            case BM_OUTLT1:
                for (j = jS; j <= jN; j++) F(u, iW - 1, j) = bcvk(r, ir, jr, WEST, VU, j) + F(u, iW, j);
                for (j = jS + 1; j <= jN; j++) F(v, iW, j) = -F(v, iW + 1, j);
                break;
            case BM_OUTLT2:
                /* note the sign difference with VelBoundCond (bound_cond.f:1727) */
                for (j = jS + 1; j <= jN; j++) F(u, iW, j) = F(u, iW + 1, j) - F(v, iW + 1, j) + F(v, iW + 1, j - 1);
                for (j = jS + 1; j <= jN; j++)
                    F(v, iW, j) = -F(v, iW, j - 1) + dFive * (F(v, iW + 1, j) - F(v, iW + 1, j - 1))
                                  + dEight * (F(u, iW + 1, j) - F(u, iW, j));
                break;
            default:
                fprintf(stderr, "Wrong nBdTypeW flag in region %d,%d\n", ir, jr);
                return ERR_FLAG;
            }
            switch (btp(r, ir, jr, EAST)) {
            case BM_INTERN: case BM_WALL1: case BM_WALL2: case BM_INLET: break;
            case BM_OUTLT1:
                for (j = jS; j <= jN; j++) F(u, iE + 1, j) = bcvk(r, ir, jr, EAST, VU, j) + F(u, iE, j);
                for (j = jS + 1; j <= jN; j++) F(v, iE + 1, j) = -F(v, iE, j);
                break;
            case BM_OUTLT2:
                for (j = jS + 1; j <= jN; j++) F(u, iE, j) = F(u, iE - 1, j) - (F(v, iE, j) - F(v, iE, j - 1));
                for (j = jS + 1; j <= jN - 1; j++)
                    F(v, iE + 1, j) = F(v, iE + 1, j - 1) + dThree * (F(v, iE, j - 1) - F(v, iE, j))
                                      - dFour * (F(u, iE, j) - F(u, iE - 1, j));
                break;
            default:
                fprintf(stderr, "Wrong nBdTypeE flag in region %d,%d\n", ir, jr);
                return ERR_FLAG;
            }
            switch (btp(r, ir, jr, SOUTH)) {
            case BM_INTERN: case BM_WALL1: case BM_WALL2: case BM_INLET: break;
            case BM_OUTLT1:
                for (i = iW + 1; i <= iE; i++) F(u, i, jS) = -F(u, i, jS + 1);
                for (i = iW; i <= iE; i++) F(v, i, jS) = bcvk(r, ir, jr, SOUTH, VV, i) + F(v, i, jS);
                break;
            case BM_OUTLT2:
                for (i = iW + 1; i <= iE; i++) F(v, i, jS) = F(v, i, jS + 1) + (F(u, i, jS + 1) - F(u, i - 1, jS + 1));
                for (i = iW + 1; i <= iE - 1; i++)
                    F(u, i, jS) = F(u, i - 1, jS) + dThree * (F(u, i - 1, jS + 1) - F(u, i, jS + 1))
                                  - dFour * (F(v, i, jS + 1) - F(v, i, jS));
                break;
            default:
                fprintf(stderr, "Wrong nBdTypeS flag in region %d,%d\n", ir, jr);
                return ERR_FLAG;
            }
            switch (btp(r, ir, jr, NORTH)) {
            case BM_INTERN: case BM_WALL1: case BM_WALL2: case BM_INLET: break;
            case BM_OUTLT1:
                for (i = iW + 1; i <= iE; i++) F(u, i, jN + 1) = -F(u, i, jN);
                for (i = iW; i <= iE; i++) F(v, i, jN + 1) = bcvk(r, ir, jr, NORTH, VV, i) + F(v, i, jN);
                break;
            case BM_OUTLT2:
                for (i = iW; i <= iE; i++) F(v, i, jN) = F(v, i, jN - 1) - (F(u, i, jN) - F(u, i - 1, jN));
                for (i = iW + 1; i <= iE - 1; i++)
                    F(u, i, jN + 1) = F(u, i - 1, jN + 1) + dThree * (F(u, i - 1, jN) - F(u, i, jN))
                                      - dFour * (F(v, i, jN) - F(v, i, jN - 1));
                break;
            default:
                fprintf(stderr, "Wrong nBdTypeN flag in region %d,%d\n", ir, jr);
                return ERR_FLAG;
            }
        }
    return 0;
}

/* -------------------------------------------------------------- pressure.f */

int wd2o_divergence(int nx, int ny, int ld, int nloc,
                    const double *xet, const double *yet, const double *xzi, const double *yzi,
                    const double *u, const double *v, double *div)
{
    int i, j;
    switch (nloc) {
    case 1:
        for (j = 1; j <= ny; j++)
            for (i = 1; i <= nx; i++) {
                double ucij, uci1j, vcij, vcij1;
                ucij = F(yet, i, j) * F(u, i, j)
                       - F(xet, i, j) * (F(v, i + 1, j) + F(v, i, j) + F(v, i + 1, j - 1) + F(v, i, j - 1)) / 4.0;
                uci1j = F(yet, i - 1, j) * F(u, i - 1, j)
                        - F(xet, i - 1, j) * (F(v, i, j) + F(v, i - 1, j) + F(v, i, j - 1) + F(v, i - 1, j - 1)) / 4.0;
                vcij = F(xzi, i, j) * F(v, i, j)
                       - F(yzi, i, j) * (F(u, i, j + 1) + F(u, i - 1, j + 1) + F(u, i, j) + F(u, i - 1, j)) / 4.0;
                vcij1 = F(xzi, i, j - 1) * F(v, i, j - 1)
                        - F(yzi, i, j - 1) * (F(u, i, j) + F(u, i - 1, j) + F(u, i, j - 1) + F(u, i - 1, j - 1)) / 4.0;
                F(div, i, j) = ucij - uci1j + vcij - vcij1;
            }
        break;
    case 2:
        for (j = 1; j <= ny; j++)
            for (i = 1; i <= nx; i++) {
                double uci1j, ucij, vcij1, vcij;
                uci1j = F(yet, i + 1, j) * (F(u, i, j + 1) + F(u, i + 1, j + 1) + F(u, i, j) + F(u, i + 1, j)) / 4.0
                        - F(xet, i + 1, j) * F(v, i + 1, j);
                ucij = F(yet, i, j) * (F(u, i - 1, j + 1) + F(u, i, j + 1) + F(u, i - 1, j) + F(u, i, j)) / 4.0
                       - F(xet, i, j) * F(v, i, j);
                vcij1 = F(xzi, i, j + 1) * (F(v, i, j + 1) + F(v, i + 1, j + 1) + F(v, i, j) + F(v, i + 1, j)) / 4.0
                        - F(yzi, i, j + 1) * F(u, i, j + 1);
                vcij = F(xzi, i, j) * (F(v, i, j) + F(v, i + 1, j) + F(v, i, j - 1) + F(v, i + 1, j - 1)) / 4.0
                       - F(yzi, i, j) * F(u, i, j);
                F(div, i, j) = uci1j - ucij + vcij1 - vcij;
            }
        break;
    default:
        fprintf(stderr, "Error: Wrong location flag passed to Divergence: %d\n", nloc);
        return ERR_FLAG;
    }
    return 0;
}

/* pressure.f:RhsPpe 329-378; b is 0-based here (b[ind-1]). */
static void rhs_ppe(int nx, int ny, int ld, int cart, double dk, const double *rbu,
                    const double *rbv, const double *div, const double *p, double *b)
{
    int i, j, ind;
    for (j = 2; j <= ny; j++)
        for (i = 2; i <= nx; i++) {
            ind = (j - 2) * (nx - 1) + i - 1;
            b[ind - 1] = F(div, i, j) / dk;
        }
    if (!cart)
        for (j = 2; j <= ny; j++)
            for (i = 2; i <= nx; i++) {
                ind = (j - 2) * (nx - 1) + i - 1;
                b[ind - 1] = b[ind - 1]
                    - (F(rbu, i, j) * (F(p, i + 1, j + 1) + F(p, i, j + 1) - F(p, i + 1, j - 1) - F(p, i, j - 1))
                       - F(rbu, i - 1, j) * (F(p, i, j + 1) + F(p, i - 1, j + 1) - F(p, i, j - 1) - F(p, i - 1, j - 1))
                       + F(rbv, i, j) * (F(p, i + 1, j + 1) + F(p, i + 1, j) - F(p, i - 1, j + 1) - F(p, i - 1, j))
                       - F(rbv, i, j - 1) * (F(p, i + 1, j) + F(p, i + 1, j - 1) - F(p, i - 1, j) - F(p, i - 1, j - 1)));
            }
}

/* PPE matrix a(mn,5): column c (1..5) of row ind (1-based) */
#define PA(ind, c) pa[(size_t)((c) - 1) * (size_t)nrow + (size_t)((ind) - 1)]

/* One point update of Sor/SorRB/SorRBP (pressure.f:425-431); returns |s|. */
static inline double sor_point(int nx, int ld, int nrow, const double *pa, const double *b,
                               double sorrel, double *p, int i, int j)
{
    int ind = (j - 2) * (nx - 1) + i - 1;
    double sum = b[ind - 1] - PA(ind, 1) * F(p, i, j - 1) - PA(ind, 2) * F(p, i - 1, j)
                 - PA(ind, 4) * F(p, i + 1, j) - PA(ind, 5) * F(p, i, j + 1);
    sum = sum / PA(ind, 3) - F(p, i, j);
    F(p, i, j) = F(p, i, j) + sorrel * sum;
    return fabs(sum);
}

/* Solve line j=nl of the Slor family (pressure.f:746-777): fills bline with
 * the line solution (ncomps = nx-1 values, 0-based). */
static void slor_line(int nx, int ld, int nrow, const double *pa, const double *b,
                      const double *p, int nl, double *aline, double *bline)
{
    int nc, ncomps = nx - 1;
    for (nc = 2; nc <= ncomps + 1; nc++) {
        int ind = (nl - 2) * ncomps + nc - 1;
        double plprev = F(p, nc, nl - 1), plnext = F(p, nc, nl + 1);
        aline[3 * (nc - 2) + 0] = PA(ind, 2);
        aline[3 * (nc - 2) + 1] = PA(ind, 3);
        aline[3 * (nc - 2) + 2] = PA(ind, 4);
        bline[nc - 2] = b[ind - 1] - (PA(ind, 1) * plprev + PA(ind, 5) * plnext);
    }
    wd2o_alt_trid_lu(ncomps, aline, bline);
}

int wd2o_ppe(const wd2o_regions *r, int lCartesGrid, int nPpeSolver, int msorit,
             int *nSorConv, double dk, double sortol, double sorrel,
             const wd2o_metrics *m, const double *u, const double *v, double *p)
{
    int nx = r->nx, ny = r->ny, ld = r->ld, i, j, ind, ir, jr, it, nl, nc;
    int nrow = (nx - 1) * (ny - 1), nconv = 0, cart = lCartesGrid != 0;
    size_t n = FIELD_N(ld, ny);
    double *div = zalloc(n), *pa = zalloc(5 * (size_t)nrow), *b = zalloc((size_t)nrow);
    double *aline = zalloc(3 * (size_t)nx), *bline = zalloc((size_t)nx), *plold = zalloc((size_t)nx);
    double *pn = NULL;

    wd2o_divergence(nx, ny, ld, 1, m->xeu, m->yeu, m->xzv, m->yzv, u, v, div);

    /* pressure.f:97-106 */
    for (j = 2; j <= ny; j++)
        for (i = 2; i <= nx; i++) {
            ind = (j - 2) * (nx - 1) + i - 1;
            PA(ind, 1) = F(m->rgv, i, j - 1);
            PA(ind, 2) = F(m->rau, i - 1, j);
            PA(ind, 3) = -F(m->rau, i, j) - F(m->rau, i - 1, j) - F(m->rgv, i, j) - F(m->rgv, i, j - 1);
            PA(ind, 4) = F(m->rau, i, j);
            PA(ind, 5) = F(m->rgv, i, j);
        }

    /* pressure.f:109-196: blockages */
    for (jr = 1; jr <= r->nReg[1]; jr++)
        for (ir = 1; ir <= r->nReg[0]; ir++) {
            if (rtype(r, ir, jr) != RM_BLOCKG) continue;
            int iW = brd(r, ir, jr, WEST), iE = brd(r, ir, jr, EAST);
            int jS = brd(r, ir, jr, SOUTH), jN = brd(r, ir, jr, NORTH);
#define IDROW(ix) do { PA(ix,1)=0.0; PA(ix,2)=0.0; PA(ix,3)=1.0; PA(ix,4)=0.0; PA(ix,5)=0.0; } while (0)
            for (j = jS + 2; j <= jN - 1; j++)
                for (i = iW + 2; i <= iE - 1; i++) {
                    ind = (j - 2) * (nx - 1) + i - 1;
                    IDROW(ind);
                    F(div, i, j) = 0.0;
                }
            if (ir > 1 && rtype(r, ir - 1, jr) == RM_BLOCKG)
                for (j = jS + 1; j <= jN; j++) {
                    ind = (j - 2) * (nx - 1) + iW;
                    IDROW(ind);
                    F(div, iW + 1, j) = 0.0;
                }
            if (ir < r->nReg[0] && rtype(r, ir + 1, jr) == RM_BLOCKG)
                for (j = jS + 1; j <= jN; j++) {
                    ind = (j - 2) * (nx - 1) + iE - 1;
                    IDROW(ind);
                    F(div, iE, j) = 0.0;
                }
            if (jr > 1 && rtype(r, ir, jr - 1) == RM_BLOCKG)
                for (i = iW + 1; i <= iE; i++) {
                    ind = (jS - 1) * (nx - 1) + i - 1;
                    IDROW(ind);
                    F(div, i, jS + 1) = 0.0;
                }
            if (jr < r->nReg[1] && rtype(r, ir, jr + 1) == RM_BLOCKG)
                for (i = iW + 1; i <= iE; i++) {
                    ind = (jN - 2) * (nx - 1) + i - 1;
                    IDROW(ind);
                    F(div, i, jN) = 0.0;
                }
#undef IDROW
        }

    switch (nPpeSolver) {
    case 1:   /* pressure.f:Sor 384-450 */
        if (cart) rhs_ppe(nx, ny, ld, cart, dk, m->rbu, m->rbv, div, p, b);
        for (it = 1; it <= msorit; it++) {
            double dif = 0.0;
            if (!cart) rhs_ppe(nx, ny, ld, cart, dk, m->rbu, m->rbv, div, p, b);
            for (j = 2; j <= ny; j++)
                for (i = 2; i <= nx; i++) dif = dmax(dif, sor_point(nx, ld, nrow, pa, b, sorrel, p, i, j));
            if (it > 1 && dif < sortol) { nconv = it; break; }
        }
        break;
    case 5:   /* pressure.f:SorRB 457-541 */
    case 6:   /* pressure.f:SorRBP 548-656 (same iterates; per-point dif array) */
        if (cart) rhs_ppe(nx, ny, ld, cart, dk, m->rbu, m->rbv, div, p, b);
        for (it = 1; it <= msorit; it++) {
            double dif = 0.0;
            if (!cart) rhs_ppe(nx, ny, ld, cart, dk, m->rbu, m->rbv, div, p, b);
            for (j = 2; j <= ny; j++)
                for (i = 2 + j % 2; i <= nx; i += 2) dif = dmax(dif, sor_point(nx, ld, nrow, pa, b, sorrel, p, i, j));
            for (j = 2; j <= ny; j++)
                for (i = 2 + (j + 1) % 2; i <= nx; i += 2) dif = dmax(dif, sor_point(nx, ld, nrow, pa, b, sorrel, p, i, j));
            if (it > 1 && dif < sortol) { nconv = it; break; }
        }
        break;
    case 2:   /* pressure.f:Slor 673-802, ndir=1 (j=const lines) */
        if (cart) rhs_ppe(nx, ny, ld, cart, dk, m->rbu, m->rbv, div, p, b);
        for (it = 1; it <= msorit; it++) {
            double dif = 0.0;
            if (!cart) rhs_ppe(nx, ny, ld, cart, dk, m->rbu, m->rbv, div, p, b);
            for (nl = 2; nl <= ny; nl++) {
                for (nc = 2; nc <= nx; nc++) plold[nc - 2] = F(p, nc, nl);
                slor_line(nx, ld, nrow, pa, b, p, nl, aline, bline);
                for (nc = 1; nc <= nx - 1; nc++) {
                    double sum = bline[nc - 1] - plold[nc - 1];
                    bline[nc - 1] = plold[nc - 1] + sorrel * sum;
                    dif = dmax(dif, fabs(sum));
                }
                for (nc = 2; nc <= nx; nc++) F(p, nc, nl) = bline[nc - 2];
            }
            if (it > 1 && dif < sortol) { nconv = it; break; }
        }
        break;
    case 3:   /* pressure.f:SlorRB 819-959, ndir=1 */
        if (cart) rhs_ppe(nx, ny, ld, cart, dk, m->rbu, m->rbv, div, p, b);
        for (it = 1; it <= msorit; it++) {
            double dif = 0.0;
            int k;
            for (k = 2; k <= 3; k++) {
                if (!cart) rhs_ppe(nx, ny, ld, cart, dk, m->rbu, m->rbv, div, p, b);
                for (nl = k; nl <= ny; nl += 2) {
                    for (nc = 2; nc <= nx; nc++) plold[nc - 2] = F(p, nc, nl);
                    slor_line(nx, ld, nrow, pa, b, p, nl, aline, bline);
                    for (nc = 1; nc <= nx - 1; nc++) {
                        double sum = bline[nc - 1] - plold[nc - 1];
                        bline[nc - 1] = plold[nc - 1] + sorrel * sum;
                        dif = dmax(dif, fabs(sum));
                    }
                    for (nc = 2; nc <= nx; nc++) F(p, nc, nl) = bline[nc - 2];
                }
            }
            if (it > 1 && dif < sortol) { nconv = it; break; }
        }
        break;
    case 4:   /* pressure.f:SlorRBP 976-1138 */
        pn = zalloc(n);
        if (cart) rhs_ppe(nx, ny, ld, cart, dk, m->rbu, m->rbv, div, p, b);
        for (it = 1; it <= msorit; it++) {
            double dif = 0.0;
            int k;
            for (nl = 1; nl <= ny; nl++)
                for (nc = 1; nc <= nx; nc++) F(pn, nc, nl) = F(p, nc, nl);
            for (k = 2; k <= 3; k++) {
                if (!cart) rhs_ppe(nx, ny, ld, cart, dk, m->rbu, m->rbv, div, p, b);
                for (nl = k; nl <= ny; nl += 2) {
                    slor_line(nx, ld, nrow, pa, b, p, nl, aline, bline);
                    for (nc = 2; nc <= nx; nc++) F(p, nc, nl) = bline[nc - 2];
                }
                for (nl = k; nl <= ny; nl += 2)
                    for (nc = 2; nc <= nx; nc++)
                        F(p, nc, nl) = F(pn, nc, nl) + sorrel * (F(p, nc, nl) - F(pn, nc, nl));
            }
            for (nl = 2; nl <= ny; nl++)
                for (nc = 2; nc <= nx; nc++) dif = dmax(dif, fabs(F(pn, nc, nl) - F(p, nc, nl)));
            if (it > 1 && dif < sortol) { nconv = it; break; }
        }
        break;
    default:
        fprintf(stderr, "Wrong nPpeSolver flag passed to Ppe\n");
        break;
    }
    if (nconv < 1) nconv = msorit;   /* pressure.f:242-246 (after the warning) */
    *nSorConv = nconv;
    free(div); free(pa); free(b); free(aline); free(bline); free(plold); free(pn);
    return 0;
}
#undef PA

/* --------------------------------------------------------------- utility.f */

int wd2o_project(const wd2o_regions *r, double dk, const wd2o_metrics *m,
                 const double *p, double *u, double *v)
{
    int ld = r->ld, ir, jr, i, j;
    const double dFour = 4.0;
    const double *dju = m->dju, *djv = m->djv, *yeu = m->yeu, *xzv = m->xzv, *yzu = m->yzu, *xev = m->xev;
    for (jr = 1; jr <= r->nReg[1]; jr++)
        for (ir = 1; ir <= r->nReg[0]; ir++) {
            if (rtype(r, ir, jr) == RM_BLOCKG) continue;
            int iW = brd(r, ir, jr, WEST), iE = brd(r, ir, jr, EAST);
            int jS = brd(r, ir, jr, SOUTH), jN = brd(r, ir, jr, NORTH);
            for (j = jS + 1; j <= jN; j++)
                for (i = iW + 1; i <= iE - 1; i++) {
                    double pzi = F(p, i + 1, j) - F(p, i, j);
                    double pet = (F(p, i + 1, j + 1) + F(p, i, j + 1) - F(p, i + 1, j - 1) - F(p, i, j - 1)) / dFour;
                    double djk = F(dju, i, j) * dk;
                    F(u, i, j) = F(u, i, j) - djk * (F(yeu, i, j) * pzi - F(yzu, i, j) * pet);
                }
            switch (btp(r, ir, jr, WEST)) {
            case BM_OUTLT1:
                for (j = jS + 1; j <= jN; j++) {
                    double pzi = F(p, iW + 1, j) - F(p, iW, j);
                    double pet = (F(p, iW + 1, j + 1) + F(p, iW, j + 1) - F(p, iW + 1, j - 1) - F(p, iW, j - 1)) / dFour;
                    double djk = F(dju, iW, j) * dk;
                    F(u, iW, j) = F(u, iW, j) - djk * (F(yeu, iW, j) * pzi - F(yzu, iW, j) * pet);
                }
                break;
            case BM_INTERN: case BM_WALL1: case BM_WALL2: case BM_INLET: case BM_OUTLT2: break;
            default: fprintf(stderr, "Wrong nBdTypeW flag in region %d,%d\n", ir, jr);
            }
            switch (btp(r, ir, jr, EAST)) {
            case BM_INTERN: case BM_OUTLT1:
                for (j = jS + 1; j <= jN; j++) {
                    double pzi = F(p, iE + 1, j) - F(p, iE, j);
                    double pet = (F(p, iE + 1, j + 1) + F(p, iE, j + 1) - F(p, iE + 1, j - 1) - F(p, iE, j - 1)) / dFour;
                    double djk = F(dju, iE, j) * dk;
                    F(u, iE, j) = F(u, iE, j) - djk * (F(yeu, iE, j) * pzi - F(yzu, iE, j) * pet);
                }
                break;
            case BM_WALL1: case BM_WALL2: case BM_INLET: case BM_OUTLT2: break;
            default: fprintf(stderr, "Wrong nBdTypeE flag in region %d,%d\n", ir, jr);
            }
        }
    for (jr = 1; jr <= r->nReg[1]; jr++)
        for (ir = 1; ir <= r->nReg[0]; ir++) {
            if (rtype(r, ir, jr) == RM_BLOCKG) continue;
            int iW = brd(r, ir, jr, WEST), iE = brd(r, ir, jr, EAST);
            int jS = brd(r, ir, jr, SOUTH), jN = brd(r, ir, jr, NORTH);
            for (j = jS + 1; j <= jN - 1; j++)
                for (i = iW + 1; i <= iE; i++) {
                    double pzi = (F(p, i + 1, j + 1) + F(p, i + 1, j) - F(p, i - 1, j + 1) - F(p, i - 1, j)) / dFour;
                    double pet = F(p, i, j + 1) - F(p, i, j);
                    double djk = F(djv, i, j) * dk;
                    F(v, i, j) = F(v, i, j) - djk * (-F(xev, i, j) * pzi + F(xzv, i, j) * pet);
                }
            switch (btp(r, ir, jr, SOUTH)) {
            case BM_OUTLT1:
                for (i = iW + 1; i <= iE; i++) {
                    double pzi = (F(p, i + 1, jS + 1) + F(p, i + 1, jS) - F(p, i - 1, jS + 1) - F(p, i - 1, jS)) / dFour;
                    double pet = F(p, i, jS + 1) - F(p, i, jS);
                    double djk = F(djv, i, jS) * dk;
                    F(v, i, jS) = F(v, i, jS) - djk * (-F(xev, i, jS) * pzi + F(xzv, i, jS) * pet);
                }
                break;
            case BM_INTERN: case BM_WALL1: case BM_WALL2: case BM_INLET: case BM_OUTLT2: break;
            default: fprintf(stderr, "Wrong nBdTypeS flag in region %d,%d\n", ir, jr);
            }
            switch (btp(r, ir, jr, NORTH)) {
            case BM_INTERN: case BM_OUTLT1:
                for (i = iW + 1; i <= iE; i++) {
                    double pzi = (F(p, i + 1, jN + 1) + F(p, i + 1, jN) - F(p, i - 1, jN + 1) - F(p, i - 1, jN)) / dFour;
                    double pet = F(p, i, jN + 1) - F(p, i, jN);
                    double djk = F(djv, i, jN) * dk;
                    F(v, i, jN) = F(v, i, jN) - djk * (-F(xev, i, jN) * pzi + F(xzv, i, jN) * pet);
                }
                break;
            case BM_WALL1: case BM_WALL2: case BM_INLET: case BM_OUTLT2: break;
            default: fprintf(stderr, "Wrong nBdTypeN flag in region %d,%d\n", ir, jr);
            }
        }
    return 0;
}

/* ------------------------------------------------------------------ main.f */

int wd2o_cold_start(const wd2o_regions *r, const wd2o_metrics *m, int lCartesGrid,
                    int nPpeSolver, int msorit, double sortol, double sorrel, double dk,
                    double *u, double *v, double *p, int *nSorConv)
{
    int st;
    if ((st = wd2o_vel_bound_cond(r, u, v)) != 0) return st;
    if ((st = wd2o_ppe(r, lCartesGrid, nPpeSolver, msorit, nSorConv, dk, sortol, sorrel, m, u, v, p)) != 0) return st;
    if ((st = wd2o_pres_bound_cond(r, p)) != 0) return st;
    if ((st = wd2o_project(r, dk, m, p, u, v)) != 0) return st;
    return wd2o_vel_bound_cond(r, u, v);
}

int wd2o_step(const wd2o_regions *r, const wd2o_metrics *m, int lCartesGrid,
              int mqiter, double qtol, int nPpeSolver, int msorit, double sortol,
              double sorrel, double dk, double re, double fr,
              const double *d, double *dn,
              double *u, double *v, double *p,
              double *pn, double *un, double *vn, double *us, double *vs,
              int *out_iter, double *out_dif)
{
    int nx = r->nx, ny = r->ny, ld = r->ld, i, j, st, nql, nsor = 0;
    for (j = 0; j <= ny + 1; j++)
        for (i = 0; i <= nx + 1; i++) {
            F(pn, i, j) = F(p, i, j);
            F(un, i, j) = F(u, i, j);
            F(vn, i, j) = F(v, i, j);
            F(dn, i, j) = F(d, i, j);
        }
    for (j = 0; j <= ny + 1; j++)
        for (i = 0; i <= nx + 1; i++) {
            F(us, i, j) = F(u, i, j);
            F(vs, i, j) = F(v, i, j);
        }
    nql = wd2o_naux_momentum(r, mqiter, dk, re, fr, qtol, m, d, dn, un, vn, us, vs);
    if (nql <= ERR_FLAG) return nql;
    if ((st = wd2o_vel_bound_cond(r, us, vs)) != 0) return st;
    if ((st = wd2o_pres_bound_cond(r, p)) != 0) return st;
    if ((st = wd2o_ppe(r, lCartesGrid, nPpeSolver, msorit, &nsor, dk, sortol, sorrel, m, us, vs, p)) != 0) return st;
    if ((st = wd2o_pres_bound_cond(r, p)) != 0) return st;
    if ((st = wd2o_project(r, dk, m, p, us, vs)) != 0) return st;
    if ((st = wd2o_vel_bound_cond(r, us, vs)) != 0) return st;
    if ((st = wd2o_pres_bound_cond(r, p)) != 0) return st;
    for (j = 0; j <= ny + 1; j++)
        for (i = 0; i <= nx + 1; i++) {
            F(u, i, j) = F(us, i, j);
            F(v, i, j) = F(vs, i, j);
        }
    if ((st = wd2o_vel_bound_cond(r, u, v)) != 0) return st;
    if ((st = wd2o_pres_bound_cond(r, p)) != 0) return st;
    out_iter[0] = nql;
    out_iter[1] = nsor;
    out_dif[0] = wd2o_diff_max_norm(nx, ny, ld, pn, p);
    out_dif[1] = wd2o_diff_max_norm(nx, ny, ld, un, u);
    out_dif[2] = wd2o_diff_max_norm(nx, ny, ld, vn, v);
    return 0;
}

/* ------------------------------------------------------------------ thermal */

/* wolfd2.h:24-31 */
enum { RT_NOSRCE = 0, RT_HEATGN = 1, RT_TEMPER = 2 };
enum { BT_INTERN = 0, BT_TEMPER = 1, BT_HTFLUX = 2 };

static inline int trtype(const wd2o_thermal *th, const wd2o_regions *r, int ir, int jr) {
    return th->nTRgType[(ir - 1) + r->mgri * (jr - 1)];
}
static inline int tbtp(const wd2o_thermal *th, const wd2o_regions *r, int ir, int jr, int side) {
    return th->nTemBdTp[(ir - 1) + r->mgri * ((jr - 1) + r->mgrj * (side - 1))];
}

/* bound_cond.f:TempBoundCond 1030-1190 */
int wd2o_temp_bound_cond(const wd2o_regions *r, const wd2o_thermal *th, double *t)
{
    int ld = r->ld, ir, jr, i, j;
    const double dTwo = 2.0;
    for (jr = 1; jr <= r->nReg[1]; jr++)
        for (ir = 1; ir <= r->nReg[0]; ir++) {
            int iW = brd(r, ir, jr, WEST), iE = brd(r, ir, jr, EAST);
            int jS = brd(r, ir, jr, SOUTH), jN = brd(r, ir, jr, NORTH);
            if (trtype(th, r, ir, jr) == RT_TEMPER) {
                const double tv = tab2(r, th->dTRgVal, ir, jr);
                for (j = jS + 1; j <= jN; j++)
                    for (i = iW + 1; i <= iE; i++) F(t, i, j) = tv;
                for (j = jS + 1; j <= jN; j++) F(t, iW + 1, j) = dTwo * bcv(r, ir, jr, WEST, VT) - F(t, iW, j);
                for (j = jS + 1; j <= jN; j++) F(t, iE, j) = dTwo * bcv(r, ir, jr, EAST, VT) - F(t, iE + 1, j);
                for (i = iW + 1; i <= iE; i++) F(t, i, jS + 1) = dTwo * bcv(r, ir, jr, SOUTH, VT) - F(t, i, jS);
                for (i = iW + 1; i <= iE; i++) F(t, i, jN) = dTwo * bcv(r, ir, jr, NORTH, VT) - F(t, i, jN + 1);
                continue;
            }
            switch (tbtp(th, r, ir, jr, WEST)) {
            case BT_INTERN: break;
            case BT_TEMPER:
                for (j = jS + 1; j <= jN; j++) F(t, iW, j) = dTwo * bcv(r, ir, jr, WEST, VT) - F(t, iW + 1, j);
                break;
            case BT_HTFLUX:
                for (j = jS + 1; j <= jN; j++) F(t, iW, j) = bcv(r, ir, jr, WEST, VT) + F(t, iW + 1, j);
                break;
            default:
                fprintf(stderr, "Wrong nTemBdTp W flag in region %d,%d\n", ir, jr);
                return ERR_FLAG;
            }
            switch (tbtp(th, r, ir, jr, EAST)) {
            case BT_INTERN: break;
            case BT_TEMPER:
                for (j = jS + 1; j <= jN; j++) F(t, iE + 1, j) = dTwo * bcv(r, ir, jr, EAST, VT) - F(t, iE, j);
                break;
            case BT_HTFLUX:
                for (j = jS + 1; j <= jN; j++) F(t, iE + 1, j) = bcv(r, ir, jr, EAST, VT) + F(t, iE, j);
                break;
            default:
                fprintf(stderr, "Wrong nTemBdTp E flag in region %d,%d\n", ir, jr);
                return ERR_FLAG;
            }
            switch (tbtp(th, r, ir, jr, SOUTH)) {
            case BT_INTERN: break;
            case BT_TEMPER:
                for (i = iW + 1; i <= iE; i++) F(t, i, jS) = dTwo * bcv(r, ir, jr, SOUTH, VT) - F(t, i, jS + 1);
                break;
            case BT_HTFLUX:
                for (i = iW + 1; i <= iE; i++) F(t, i, jS) = bcv(r, ir, jr, SOUTH, VT) + F(t, i, jS + 1);
                break;
            default:
                fprintf(stderr, "Wrong nTemBdTp S flag in region %d,%d\n", ir, jr);
                return ERR_FLAG;
            }
            switch (tbtp(th, r, ir, jr, NORTH)) {
            case BT_INTERN: break;
            case BT_TEMPER:
                for (i = iW + 1; i <= iE; i++) F(t, i, jN + 1) = dTwo * bcv(r, ir, jr, NORTH, VT) - F(t, i, jN);
                break;
            case BT_HTFLUX:
                for (i = iW + 1; i <= iE; i++) F(t, i, jN + 1) = bcv(r, ir, jr, NORTH, VT) + F(t, i, jN);
                break;
            default:
                fprintf(stderr, "Wrong nTemBdTp N flag in region %d,%d\n", ir, jr);
                return ERR_FLAG;
            }
        }
    return 0;
}

/* thermal.f:ThermEnergy 24-274.  Both splits hand AltTridLU one chain of
 * (nx-1)(ny-1) rows ordered i fastest (ind = (j-2)(nx-1)+i-1); the second
 * split's rows carry the j-direction coefficients in that same i-fastest
 * order (thermal.f:210-245), which this restatement keeps. */
int wd2o_therm_energy(const wd2o_regions *r, const wd2o_thermal *th, double dk, double pe,
                      const wd2o_metrics *m, const double *un, const double *vn,
                      const double *u, const double *v, const double *tn, double *t)
{
    const int nx = r->nx, ny = r->ny, ld = r->ld;
    const size_t nf = (size_t)ld * (ny + 2 > 0 ? (size_t)(ny + 2) : 1);
    const int n = (nx - 1) * (ny - 1);
    const double dZero = 0.0, dOne = 1.0, dTwo = 2.0, dHalf = 0.5;
    const double *rau = m->rau, *rbu = m->rbu, *rbv = m->rbv, *rgv = m->rgv, *djc = m->djc;
    int ir, jr, i, j, st;
    double *cu1 = zalloc(nf), *cun = zalloc(nf), *cv1 = zalloc(nf), *cvn = zalloc(nf), *s = zalloc(nf);
    double *a = zalloc(3 * (size_t)n), *b = zalloc((size_t)n);
    const double pe1 = dOne / pe;
#define A(k, ind) a[3 * ((ind) - 1) + (k) - 1]
#define B(ind) b[(ind) - 1]
    st = wd2o_temp_bound_cond(r, th, t);
    if (st) goto out;
    wd2o_conv_coef(nx, ny, ld, 6, 0, m->xzc, m->xec, m->yzc, m->yec, u, v, cu1, cv1);
    wd2o_conv_coef(nx, ny, ld, 3, 0, m->xzv, m->xeu, m->yzv, m->yeu, un, vn, cun, cvn);
    /* heat generation term */
    for (jr = 1; jr <= r->nReg[1]; jr++)
        for (ir = 1; ir <= r->nReg[0]; ir++) {
            int iW = brd(r, ir, jr, WEST), iE = brd(r, ir, jr, EAST);
            int jS = brd(r, ir, jr, SOUTH), jN = brd(r, ir, jr, NORTH);
            const double sv = trtype(th, r, ir, jr) == RT_HEATGN ? tab2(r, th->dHGSTval, ir, jr) : dZero;
            for (j = jS + 1; j <= jN; j++)
                for (i = iW + 1; i <= iE; i++) F(s, i, j) = sv;
        }
    /* first split step */
    for (j = 2; j <= ny; j++)
        for (i = 2; i <= nx; i++) {
            const int ind = (j - 2) * (nx - 1) + i - 1;
            const double rkj = dk * F(djc, i, j) * dHalf;
            if (F(cu1, i, j) >= dZero) {
                A(1, ind) = rkj * (-F(cu1, i - 1, j) - pe1 * F(rau, i - 1, j));
                A(2, ind) = dOne + rkj * (F(cu1, i, j) + pe1 * (F(rau, i, j) + F(rau, i - 1, j)));
                A(3, ind) = rkj * (-pe1 * F(rau, i, j));
            } else {
                A(1, ind) = rkj * (-pe1 * F(rau, i - 1, j));
                A(2, ind) = dOne + rkj * (-F(cu1, i, j) + pe1 * (F(rau, i, j) + F(rau, i - 1, j)));
                A(3, ind) = rkj * (F(cu1, i + 1, j) - pe1 * F(rau, i, j));
            }
            const double c = -F(cvn, i, j - 1) * F(tn, i, j - 1) - F(cun, i - 1, j) * F(tn, i - 1, j)
                             + (F(cun, i, j) - F(cun, i - 1, j) + F(cvn, i, j) - F(cvn, i, j - 1)) * F(tn, i, j)
                             + F(cun, i, j) * F(tn, i + 1, j) + F(cvn, i, j) * F(tn, i, j + 1);
            const double d = F(rau, i, j) * (F(tn, i + 1, j) - F(tn, i, j))
                             - F(rau, i - 1, j) * (F(tn, i, j) - F(tn, i - 1, j))
                             + F(rbu, i, j) * (F(tn, i + 1, j + 1) + F(tn, i, j + 1) - F(tn, i + 1, j - 1) - F(tn, i, j - 1))
                             - F(rbu, i - 1, j) * (F(tn, i, j + 1) + F(tn, i - 1, j + 1) - F(tn, i, j - 1) - F(tn, i - 1, j - 1))
                             + F(rbv, i, j) * (F(tn, i + 1, j + 1) + F(tn, i + 1, j) - F(tn, i - 1, j + 1) - F(tn, i - 1, j))
                             - F(rbv, i, j - 1) * (F(tn, i + 1, j) + F(tn, i + 1, j - 1) - F(tn, i - 1, j) - F(tn, i - 1, j - 1))
                             + F(rgv, i, j) * (F(tn, i, j + 1) - F(tn, i, j))
                             - F(rgv, i, j - 1) * (F(tn, i, j) - F(tn, i, j - 1));
            B(ind) = dk * F(s, i, j) * dHalf + dTwo * rkj * (-c + pe1 * d);
        }
    wd2o_alt_trid_lu(n, a, b);
    /* second split step: the right-hand side is the first split's solution */
    for (j = 2; j <= ny; j++)
        for (i = 2; i <= nx; i++) {
            const int ind = (j - 2) * (nx - 1) + i - 1;
            const double rkj = dk * F(djc, i, j) * dHalf;
            if (F(cv1, i, j) >= dZero) {
                A(1, ind) = rkj * (-F(cv1, i, j - 1) - pe1 * F(rgv, i, j - 1));
                A(2, ind) = dOne + rkj * (F(cv1, i, j) + pe1 * (F(rgv, i, j) + F(rgv, i, j - 1)));
                A(3, ind) = rkj * (-pe1 * F(rgv, i, j));
            } else {
                A(1, ind) = rkj * (-pe1 * F(rgv, i, j - 1));
                A(2, ind) = dOne + rkj * (-F(cv1, i, j) + pe1 * (F(rgv, i, j) + F(rgv, i, j - 1)));
                A(3, ind) = rkj * (F(cv1, i, j + 1) - pe1 * F(rgv, i, j));
            }
        }
    /* interior of fixed-temperature regions */
    for (jr = 1; jr <= r->nReg[1]; jr++)
        for (ir = 1; ir <= r->nReg[0]; ir++) {
            if (trtype(th, r, ir, jr) != RT_TEMPER) continue;
            int iW = brd(r, ir, jr, WEST), iE = brd(r, ir, jr, EAST);
            int jS = brd(r, ir, jr, SOUTH), jN = brd(r, ir, jr, NORTH);
            for (j = jS + 1; j <= jN; j++)
                for (i = iW + 1; i <= iE; i++) {
                    const int ind = (j - 2) * (nx - 1) + i - 1;
                    A(1, ind) = dZero;
                    A(2, ind) = dOne;
                    A(3, ind) = dZero;
                    B(ind) = dZero;
                }
        }
    wd2o_alt_trid_lu(n, a, b);
    for (j = 2; j <= ny; j++)
        for (i = 2; i <= nx; i++) {
            const int ind = (j - 2) * (nx - 1) + i - 1;
            F(t, i, j) = F(t, i, j) + B(ind);
        }
#undef A
#undef B
out:
    free(cu1); free(cun); free(cv1); free(cvn); free(s); free(a); free(b);
    return st;
}

/* thermal.f:EqState 280-328 */
void wd2o_eq_state(int nx, int ny, int ld, double uref, double densref, double tmax, double tref,
                   double rconst, const double *p, const double *t, double *den)
{
    int i, j;
    const double dZero = 0.0, dOne = 1.0;
    const double pref = densref * rconst * tref;
    for (j = 2; j <= ny; j++)
        for (i = 2; i <= nx; i++) {
            const double c1 = F(p, i, j) * densref * (uref * uref) + pref;
            const double c2 = densref * rconst * (F(t, i, j) * (tmax - tref) + tref);
            F(den, i, j) = c1 / c2 - dOne;
            if (fabs(F(den, i, j)) < 1.e-10) F(den, i, j) = dZero;
        }
}

/* utility.f:Filter 31-247 */
int wd2o_filter(const wd2o_regions *r, const wd2o_thermal *th, int ncomp, double fp, double *qu)
{
    const int nx = r->nx, ny = r->ny, ld = r->ld;
    const double dFour = 4.0;
    int ir, jr, i, j;
    if (ncomp != VU && ncomp != VV && ncomp != VT) {
        fprintf(stderr, "Wrong ncomp flag passed to Filter\n");
        return ERR_FLAG;
    }
    double *qh = zalloc((size_t)ld * (ny + 2));
#define FILT(i, j) ((F(qu, i, j - 1) + F(qu, i - 1, j) + F(qu, i, j + 1) + F(qu, i + 1, j) + fp * F(qu, i, j)) / (fp + dFour))
    for (j = 0; j <= ny + 1; j++)
        for (i = 0; i <= nx + 1; i++) F(qh, i, j) = F(qu, i, j);
    for (jr = 1; jr <= r->nReg[1]; jr++)
        for (ir = 1; ir <= r->nReg[0]; ir++) {
            int iW = brd(r, ir, jr, WEST), iE = brd(r, ir, jr, EAST);
            int jS = brd(r, ir, jr, SOUTH), jN = brd(r, ir, jr, NORTH);
            if (ncomp == VU) {
                if (rtype(r, ir, jr) == RM_BLOCKG) continue;
                for (j = jS + 1; j <= jN; j++)
                    for (i = iW + 1; i <= iE - 1; i++) F(qh, i, j) = FILT(i, j);
                int tw = btp(r, ir, jr, WEST), te = btp(r, ir, jr, EAST);
                if (tw == BM_OUTLT1)
                    for (j = jS + 1; j <= jN; j++) F(qh, iW, j) = FILT(iW, j);
                else if (tw < BM_INTERN || tw > BM_OUTLT2)
                    fprintf(stderr, "Wrong nBdTypeW flag in region %d,%d\n", ir, jr);
                if (te == BM_INTERN || te == BM_OUTLT1)
                    for (j = jS + 1; j <= jN; j++) F(qh, iE, j) = FILT(iE, j);
                else if (te < BM_INTERN || te > BM_OUTLT2)
                    fprintf(stderr, "Wrong nBdTypeE flag in region %d,%d\n", ir, jr);
            } else if (ncomp == VV) {
                if (rtype(r, ir, jr) == RM_BLOCKG) continue;
                for (j = jS + 1; j <= jN - 1; j++)
                    for (i = iW + 1; i <= iE; i++) F(qh, i, j) = FILT(i, j);
                int ts = btp(r, ir, jr, SOUTH), tn = btp(r, ir, jr, NORTH);
                if (ts == BM_OUTLT1)
                    for (i = iW + 1; i <= iE; i++) F(qh, i, jS) = FILT(i, jS);
                else if (ts < BM_INTERN || ts > BM_OUTLT2)
                    fprintf(stderr, "Wrong nBdTypeS flag in region %d,%d\n", ir, jr);
                if (tn == BM_INTERN || tn == BM_OUTLT1)
                    for (i = iW + 1; i <= iE; i++) F(qh, i, jN) = FILT(i, jN);
                else if (tn < BM_INTERN || tn > BM_OUTLT2)
                    fprintf(stderr, "Wrong nBdTypeN flag in region %d,%d\n", ir, jr);
            } else {
                /* utility.f:209 compares the thermal region type with BT_TEMPER
                 * (= RT_HEATGN): heat-generating regions are the ones skipped */
                if (trtype(th, r, ir, jr) == BT_TEMPER) continue;
                for (j = jS + 1; j <= jN; j++)
                    for (i = iW + 1; i <= iE; i++) F(qh, i, j) = FILT(i, j);
            }
        }
#undef FILT
    for (j = 0; j <= ny + 1; j++)
        for (i = 0; i <= nx + 1; i++) F(qu, i, j) = F(qh, i, j);
    free(qh);
    return 0;
}

/* One time step of main.f:690-972 with the energy equation, buoyancy,
 * momentum/energy iterations and filters (wd2o_step_options).  Work arrays
 * (pn, un, vn, tn, us, vs, ts: ld*(ny+2) doubles each) are caller-owned;
 * d is the density (EqState's output), dn its time-level-n copy.
 * out_iter = {nQLiter, nSorConv} of the last iteration; out_dif = the four
 * PrintDiff norms {|p1-pn|, |u1-un|, |v1-vn|, |t1-tn|}. */
int wd2o_step_ex(const wd2o_regions *r, const wd2o_thermal *th, const wd2o_metrics *m,
                 const wd2o_step_options *o, int lCartesGrid, int mqiter, double qtol,
                 int nPpeSolver, int msorit, double sortol, double sorrel, double dk, double re,
                 double fr, double *d, double *dn, double *u, double *v, double *p, double *t,
                 double *pn, double *un, double *vn, double *tn, double *us, double *vs, double *ts,
                 int *out_iter, double *out_dif)
{
    int nx = r->nx, ny = r->ny, ld = r->ld, i, j, l, st, nql = 0, nsor = 0;
    int nmeiter = o->nmeiter;
    for (j = 0; j <= ny + 1; j++)
        for (i = 0; i <= nx + 1; i++) {
            F(pn, i, j) = F(p, i, j);
            F(un, i, j) = F(u, i, j);
            F(vn, i, j) = F(v, i, j);
            F(tn, i, j) = F(t, i, j);
            F(dn, i, j) = F(d, i, j);
        }
    if (o->nthermen != 1 && nmeiter > 0) nmeiter = 1;   /* main.f:735 */
    for (l = 1; l <= nmeiter; l++) {
        double dif1, dif2, dif3, difmax;
        for (j = 0; j <= ny + 1; j++)
            for (i = 0; i <= nx + 1; i++) {
                F(us, i, j) = F(u, i, j);
                F(vs, i, j) = F(v, i, j);
                F(ts, i, j) = F(t, i, j);
            }
        nql = wd2o_naux_momentum(r, mqiter, dk, re, fr, qtol, m, d, dn, un, vn, us, vs);
        if (nql <= ERR_FLAG) return nql;
        if (o->nfiltu == 1 && (st = wd2o_filter(r, th, VU, o->fpu, us)) != 0) return st;
        if (o->nfiltv == 1 && (st = wd2o_filter(r, th, VV, o->fpv, vs)) != 0) return st;
        if ((st = wd2o_vel_bound_cond(r, us, vs)) != 0) return st;
        if ((st = wd2o_pres_bound_cond(r, p)) != 0) return st;
        if ((st = wd2o_ppe(r, lCartesGrid, nPpeSolver, msorit, &nsor, dk, sortol, sorrel, m, us, vs, p)) != 0)
            return st;
        if ((st = wd2o_pres_bound_cond(r, p)) != 0) return st;
        if ((st = wd2o_project(r, dk, m, p, us, vs)) != 0) return st;
        if ((st = wd2o_vel_bound_cond(r, us, vs)) != 0) return st;
        if ((st = wd2o_pres_bound_cond(r, p)) != 0) return st;
        if (o->nthermen == 1 && (st = wd2o_therm_energy(r, th, dk, o->pe, m, un, vn, us, vs, tn, ts)) != 0)
            return st;
        if (o->neqstate == 1)
            wd2o_eq_state(nx, ny, ld, o->uref, o->densref, o->tmax, o->tref, o->rconst, p, ts, d);
        dif1 = wd2o_diff_max_norm(nx, ny, ld, u, us);
        dif2 = wd2o_diff_max_norm(nx, ny, ld, v, vs);
        dif3 = wd2o_diff_max_norm(nx, ny, ld, t, ts);
        for (j = 0; j <= ny + 1; j++)
            for (i = 0; i <= nx + 1; i++) {
                F(u, i, j) = F(us, i, j);
                F(v, i, j) = F(vs, i, j);
                F(t, i, j) = F(ts, i, j);
            }
        /* main.f:873-877, Fortran MAX of the three */
        difmax = dif1;
        if (dif2 > difmax) difmax = dif2;
        if (dif3 > difmax) difmax = dif3;
        if (l > 1 && difmax < o->dmeittol) break;
    }
    if (o->nthermen == 1 && o->nfiltt == 1 && (st = wd2o_filter(r, th, VT, o->fpt, t)) != 0) return st;
    if ((st = wd2o_vel_bound_cond(r, u, v)) != 0) return st;
    if ((st = wd2o_pres_bound_cond(r, p)) != 0) return st;
    if (o->nthermen == 1 && (st = wd2o_temp_bound_cond(r, th, t)) != 0) return st;
    out_iter[0] = nql;
    out_iter[1] = nsor;
    out_dif[0] = wd2o_diff_max_norm(nx, ny, ld, pn, p);
    out_dif[1] = wd2o_diff_max_norm(nx, ny, ld, un, u);
    out_dif[2] = wd2o_diff_max_norm(nx, ny, ld, vn, v);
    out_dif[3] = wd2o_diff_max_norm(nx, ny, ld, tn, t);
    return 0;
}
