/*
 * ns_oracle.c -- CPU FP64 restatement of the NS_3D_Staggered_dp time step.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the parity oracle: tests/, the
 * smoke() check and bench.py's cpu_baseline leg may load it, and only as the
 * checker or the timed CPU baseline.  No product code path calls it.
 *
 * Every routine restates one piece of the reference (paths relative to the
 * reference checkout's NS_3D_Staggered_dp/ folder) and operates on the
 * reference's own storage: Fortran column-major ghosted arrays
 *   u(nx+1,ny+2,nz+2)  v(nx+2,ny+1,nz+2)  w(nx+2,ny+2,nz+1)  p,dp(nx+2,ny+2,nz+2)
 * (big_periodic_3D.f90:14-18,27-28).  Index helpers below are 1-based like
 * the Fortran so that each line can be read against the reference.
 *
 * Boundary-plane sets are passed packed (see nso_plane_layout): the 18
 * velocity planes in vel_bc_staggered's argument order
 * (NS_functions.f90:213-217), then the 6 pressure planes, then one scalar,
 * mean_dp (NS_functions.f90:1577).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <complex.h>
#undef I /* the reference names cell indices I, J, K */
#define CI _Complex_I
#ifdef _OPENMP
#include <omp.h>
#endif

#define NSO_EULER 0
#define NSO_AB2 1
#define NSO_AB2_CN 2
#define NSO_FLAG_GAUGE_CORNER 1 /* dp = dp - dp(1,1,1): big_periodic_3D.f90:562 */
#define NSO_FLAG_CN_CYCLIC 2    /* correct cyclic row 1 instead of mat_CN's (NS_functions.f90:1273) */

typedef struct {
    int nx, ny, nz;
    double dx, dy, dz, nu, dt;
    int scheme;
    int bc[3], pbc[3];
    int dp_flag;
    int flags;
} nso_config;

/* 1-based Fortran index into an (n1,n2,*) array */
#define IX(i, j, k, n1, n2) ((size_t)((i) - 1) + (size_t)(n1) * ((size_t)((j) - 1) + (size_t)(n2) * (size_t)((k) - 1)))

static const double PI = 3.1415926535897932; /* big_periodic_3D.f90:6 */

/* ------------------------------------------------------------------ */
/* packed plane layout                                                 */
/* ------------------------------------------------------------------ */
/* sizes of the 24 planes, in order: u(bx1,bxn,by1,byn,bz1,bzn), v(...), w(...), p(...) */
void nso_plane_layout(int nx, int ny, int nz, long *sizes /*24*/, long *offsets /*25*/)
{
    long d[4][3] = {{nx + 1, ny + 2, nz + 2}, {nx + 2, ny + 1, nz + 2}, {nx + 2, ny + 2, nz + 1}, {nx + 2, ny + 2, nz + 2}};
    int q = 0;
    for (int c = 0; c < 4; c++) {
        long sx = d[c][1] * d[c][2], sy = d[c][0] * d[c][2], sz = d[c][0] * d[c][1];
        sizes[q++] = sx; sizes[q++] = sx;
        sizes[q++] = sy; sizes[q++] = sy;
        sizes[q++] = sz; sizes[q++] = sz;
    }
    long off = 0;
    for (q = 0; q < 24; q++) { offsets[q] = off; off += sizes[q]; }
    offsets[24] = off; /* mean_dp scalar lives here; total length = off + 1 */
}

static const double *plane(const double *planes, int nx, int ny, int nz, int q)
{
    long s[24], o[25];
    nso_plane_layout(nx, ny, nz, s, o);
    return planes + o[q];
}
static double *plane_w(double *planes, int nx, int ny, int nz, int q)
{
    long s[24], o[25];
    nso_plane_layout(nx, ny, nz, s, o);
    return planes + o[q];
}

/* ------------------------------------------------------------------ */
/* FD primitives: FD_functions.f90:222-527                              */
/* ------------------------------------------------------------------ */
/* avg(A,k,bc): bc==1 keeps the length and wraps B(1)=(A(n)+A(n-1))/2
 * (FD_functions.f90:246-301); otherwise length n-1, B(i)=(A(i+1)+A(i))/2
 * (:302-339). */
void nso_avg(const double *A, int n1, int n2, int n3, int k, int bc, double *B)
{
    int m1 = n1, m2 = n2, m3 = n3;
    if (bc != 1) { if (k == 1) m1--; else if (k == 2) m2--; else m3--; }
    for (int c = 1; c <= m3; c++)
        for (int b = 1; b <= m2; b++)
            for (int a = 1; a <= m1; a++) {
                double r;
                if (bc == 1) {
                    if (k == 1) r = (a == 1) ? (A[IX(n1, b, c, n1, n2)] + A[IX(n1 - 1, b, c, n1, n2)]) / 2.0
                                             : (A[IX(a, b, c, n1, n2)] + A[IX(a - 1, b, c, n1, n2)]) / 2.0;
                    else if (k == 2) r = (b == 1) ? (A[IX(a, n2, c, n1, n2)] + A[IX(a, n2 - 1, c, n1, n2)]) / 2.0
                                                  : (A[IX(a, b, c, n1, n2)] + A[IX(a, b - 1, c, n1, n2)]) / 2.0;
                    else r = (c == 1) ? (A[IX(a, b, n3, n1, n2)] + A[IX(a, b, n3 - 1, n1, n2)]) / 2.0
                                      : (A[IX(a, b, c, n1, n2)] + A[IX(a, b, c - 1, n1, n2)]) / 2.0;
                } else {
                    if (k == 1) r = (A[IX(a + 1, b, c, n1, n2)] + A[IX(a, b, c, n1, n2)]) / 2.0;
                    else if (k == 2) r = (A[IX(a, b + 1, c, n1, n2)] + A[IX(a, b, c, n1, n2)]) / 2.0;
                    else r = (A[IX(a, b, c + 1, n1, n2)] + A[IX(a, b, c, n1, n2)]) / 2.0;
                }
                B[IX(a, b, c, m1, m2)] = r;
            }
}

/* diff(A,n,k): forward difference, length n-1 along k (FD_functions.f90:345-410) */
void nso_diff(const double *A, int n1, int n2, int n3, int k, double *B)
{
    int m1 = n1 - (k == 1), m2 = n2 - (k == 2), m3 = n3 - (k == 3);
    for (int c = 1; c <= m3; c++)
        for (int b = 1; b <= m2; b++)
            for (int a = 1; a <= m1; a++) {
                double r;
                if (k == 1) r = A[IX(a + 1, b, c, n1, n2)] - A[IX(a, b, c, n1, n2)];
                else if (k == 2) r = A[IX(a, b + 1, c, n1, n2)] - A[IX(a, b, c, n1, n2)];
                else r = A[IX(a, b, c + 1, n1, n2)] - A[IX(a, b, c, n1, n2)];
                B[IX(a, b, c, m1, m2)] = r;
            }
}

/* diff2(A,k,bc): bc==1 -> length n-1 with B(1)=A(n-1)-2A(1)+A(2) (:437-486);
 * otherwise length n-2, B(i)=A(i)-2A(i+1)+A(i+2) (:487-523). */
void nso_diff2(const double *A, int n1, int n2, int n3, int k, int bc, double *B)
{
    int s = (bc == 1) ? 1 : 2;
    int m1 = n1 - (k == 1) * s, m2 = n2 - (k == 2) * s, m3 = n3 - (k == 3) * s;
    for (int c = 1; c <= m3; c++)
        for (int b = 1; b <= m2; b++)
            for (int a = 1; a <= m1; a++) {
                double r;
                if (bc == 1) {
                    if (k == 1) r = (a == 1) ? A[IX(n1 - 1, b, c, n1, n2)] - 2.0 * A[IX(1, b, c, n1, n2)] + A[IX(2, b, c, n1, n2)]
                                             : A[IX(a - 1, b, c, n1, n2)] - 2.0 * A[IX(a, b, c, n1, n2)] + A[IX(a + 1, b, c, n1, n2)];
                    else if (k == 2) r = (b == 1) ? A[IX(a, n2 - 1, c, n1, n2)] - 2.0 * A[IX(a, 1, c, n1, n2)] + A[IX(a, 2, c, n1, n2)]
                                                  : A[IX(a, b - 1, c, n1, n2)] - 2.0 * A[IX(a, b, c, n1, n2)] + A[IX(a, b + 1, c, n1, n2)];
                    else r = (c == 1) ? A[IX(a, b, n3 - 1, n1, n2)] - 2.0 * A[IX(a, b, 1, n1, n2)] + A[IX(a, b, 2, n1, n2)]
                                      : A[IX(a, b, c - 1, n1, n2)] - 2.0 * A[IX(a, b, c, n1, n2)] + A[IX(a, b, c + 1, n1, n2)];
                } else {
                    if (k == 1) r = A[IX(a, b, c, n1, n2)] - 2.0 * A[IX(a + 1, b, c, n1, n2)] + A[IX(a + 2, b, c, n1, n2)];
                    else if (k == 2) r = A[IX(a, b, c, n1, n2)] - 2.0 * A[IX(a, b + 1, c, n1, n2)] + A[IX(a, b + 2, c, n1, n2)];
                    else r = A[IX(a, b, c, n1, n2)] - 2.0 * A[IX(a, b, c + 1, n1, n2)] + A[IX(a, b, c + 2, n1, n2)];
                }
                B[IX(a, b, c, m1, m2)] = r;
            }
}

/* ------------------------------------------------------------------ */
/* cal_conv: NS_functions.f90:53-158 (pointwise form of the avg/diff   */
/* composition).  Output shapes: conv_x (nx or nx-1, ny, nz) etc.       */
/* ------------------------------------------------------------------ */
void nso_cal_conv(const double *u, const double *v, const double *w, int nx, int ny, int nz,
                  int bcx, int bcy, int bcz, double dx, double dy, double dz,
                  double *cx, double *cy, double *cz)
{
    const int U1 = nx + 1, U2 = ny + 2, V1 = nx + 2, V2 = ny + 1, W1 = nx + 2, W2 = ny + 2;
#define Uv(i, j, k) u[IX(i, j, k, U1, U2)]
#define Vv(i, j, k) v[IX(i, j, k, V1, V2)]
#define Wv(i, j, k) w[IX(i, j, k, W1, W2)]
    /* edge fluxes: E on xy-edges (u face i, v face j), F on xz-edges, G on yz-edges */
#define EF(i, j, k) (((Uv(i, (j) + 1, k) + Uv(i, j, k)) / 2.0) * ((Vv((i) + 1, j, k) + Vv(i, j, k)) / 2.0))
#define FF(i, j, k) (((Uv(i, j, (k) + 1) + Uv(i, j, k)) / 2.0) * ((Wv((i) + 1, j, k) + Wv(i, j, k)) / 2.0))
#define GF(i, j, k) (((Vv(i, j, (k) + 1) + Vv(i, j, k)) / 2.0) * ((Wv(i, (j) + 1, k) + Wv(i, j, k)) / 2.0))
    /* x-component */
    {
        int m1 = (bcx == 1) ? nx : nx - 1, off = (bcx == 1) ? 0 : 1;
#pragma omp parallel for collapse(2) schedule(static)
        for (int c = 1; c <= nz; c++)
            for (int b = 1; b <= ny; b++)
                for (int a = 1; a <= m1; a++) {
                    int fi = a + off, J = b + 1, K = c + 1;
                    double ur = (Uv(fi + 1, J, K) + Uv(fi, J, K)) / 2.0;
                    double ul = (bcx == 1 && fi == 1) ? (Uv(nx + 1, J, K) + Uv(nx, J, K)) / 2.0
                                                      : (Uv(fi, J, K) + Uv(fi - 1, J, K)) / 2.0;
                    double t1 = ur * ur - ul * ul;
                    double t2 = EF(fi, J, K) - EF(fi, J - 1, K);
                    double t3 = FF(fi, J, K) - FF(fi, J, K - 1);
                    double r = t1 / dx;
                    r = r + t2 / dy;
                    r = r + t3 / dz;
                    cx[IX(a, b, c, m1, ny)] = r;
                }
    }
    /* y-component */
    {
        int m2 = (bcy == 1) ? ny : ny - 1, off = (bcy == 1) ? 0 : 1;
#pragma omp parallel for collapse(2) schedule(static)
        for (int c = 1; c <= nz; c++)
            for (int b = 1; b <= m2; b++)
                for (int a = 1; a <= nx; a++) {
                    int I = a + 1, fj = b + off, K = c + 1;
                    double t1 = EF(I, fj, K) - EF(I - 1, fj, K);
                    double vr = (Vv(I, fj + 1, K) + Vv(I, fj, K)) / 2.0;
                    double vl = (bcy == 1 && fj == 1) ? (Vv(I, ny + 1, K) + Vv(I, ny, K)) / 2.0
                                                      : (Vv(I, fj, K) + Vv(I, fj - 1, K)) / 2.0;
                    double t2 = vr * vr - vl * vl;
                    double t3 = GF(I, fj, K) - GF(I, fj, K - 1);
                    double r = t1 / dx;
                    r = r + t2 / dy;
                    r = r + t3 / dz;
                    cy[IX(a, b, c, nx, m2)] = r;
                }
    }
    /* z-component */
    {
        int m3 = (bcz == 1) ? nz : nz - 1, off = (bcz == 1) ? 0 : 1;
#pragma omp parallel for collapse(2) schedule(static)
        for (int c = 1; c <= m3; c++)
            for (int b = 1; b <= ny; b++)
                for (int a = 1; a <= nx; a++) {
                    int I = a + 1, J = b + 1, fk = c + off;
                    double t1 = FF(I, J, fk) - FF(I - 1, J, fk);
                    double t2 = GF(I, J, fk) - GF(I, J - 1, fk);
                    double wr = (Wv(I, J, fk + 1) + Wv(I, J, fk)) / 2.0;
                    double wl = (bcz == 1 && fk == 1) ? (Wv(I, J, nz + 1) + Wv(I, J, nz)) / 2.0
                                                      : (Wv(I, J, fk) + Wv(I, J, fk - 1)) / 2.0;
                    double t3 = wr * wr - wl * wl;
                    double r = t1 / dx;
                    r = r + t2 / dy;
                    r = r + t3 / dz;
                    cz[IX(a, b, c, nx, ny)] = r;
                }
    }
#undef EF
#undef FF
#undef GF
}

/* cal_diff: NS_functions.f90:161-210 */
void nso_cal_diff(const double *u, const double *v, const double *w, int nx, int ny, int nz,
                  int bcx, int bcy, int bcz, double dx2, double dy2, double dz2,
                  double *dfx, double *dfy, double *dfz)
{
    const int U1 = nx + 1, U2 = ny + 2, V1 = nx + 2, V2 = ny + 1, W1 = nx + 2, W2 = ny + 2;
    {
        int m1 = (bcx == 1) ? nx : nx - 1, off = (bcx == 1) ? 0 : 1;
#pragma omp parallel for collapse(2) schedule(static)
        for (int c = 1; c <= nz; c++)
            for (int b = 1; b <= ny; b++)
                for (int a = 1; a <= m1; a++) {
                    int fi = a + off, J = b + 1, K = c + 1;
                    double l = (bcx == 1 && fi == 1) ? Uv(nx, J, K) : Uv(fi - 1, J, K);
                    double t1 = l - 2.0 * Uv(fi, J, K) + Uv(fi + 1, J, K);
                    double t2 = Uv(fi, J - 1, K) - 2.0 * Uv(fi, J, K) + Uv(fi, J + 1, K);
                    double t3 = Uv(fi, J, K - 1) - 2.0 * Uv(fi, J, K) + Uv(fi, J, K + 1);
                    double r = t1 / dx2;
                    r = r + t2 / dy2;
                    r = r + t3 / dz2;
                    dfx[IX(a, b, c, m1, ny)] = r;
                }
    }
    {
        int m2 = (bcy == 1) ? ny : ny - 1, off = (bcy == 1) ? 0 : 1;
#pragma omp parallel for collapse(2) schedule(static)
        for (int c = 1; c <= nz; c++)
            for (int b = 1; b <= m2; b++)
                for (int a = 1; a <= nx; a++) {
                    int I = a + 1, fj = b + off, K = c + 1;
                    double l = (bcy == 1 && fj == 1) ? Vv(I, ny, K) : Vv(I, fj - 1, K);
                    double t1 = Vv(I - 1, fj, K) - 2.0 * Vv(I, fj, K) + Vv(I + 1, fj, K);
                    double t2 = l - 2.0 * Vv(I, fj, K) + Vv(I, fj + 1, K);
                    double t3 = Vv(I, fj, K - 1) - 2.0 * Vv(I, fj, K) + Vv(I, fj, K + 1);
                    double r = t1 / dx2;
                    r = r + t2 / dy2;
                    r = r + t3 / dz2;
                    dfy[IX(a, b, c, nx, m2)] = r;
                }
    }
    {
        int m3 = (bcz == 1) ? nz : nz - 1, off = (bcz == 1) ? 0 : 1;
#pragma omp parallel for collapse(2) schedule(static)
        for (int c = 1; c <= m3; c++)
            for (int b = 1; b <= ny; b++)
                for (int a = 1; a <= nx; a++) {
                    int I = a + 1, J = b + 1, fk = c + off;
                    double l = (bcz == 1 && fk == 1) ? Wv(I, J, nz) : Wv(I, J, fk - 1);
                    double t1 = Wv(I - 1, J, fk) - 2.0 * Wv(I, J, fk) + Wv(I + 1, J, fk);
                    double t2 = Wv(I, J - 1, fk) - 2.0 * Wv(I, J, fk) + Wv(I, J + 1, fk);
                    double t3 = l - 2.0 * Wv(I, J, fk) + Wv(I, J, fk + 1);
                    double r = t1 / dx2;
                    r = r + t2 / dy2;
                    r = r + t3 / dz2;
                    dfz[IX(a, b, c, nx, ny)] = r;
                }
    }
#undef Uv
#undef Vv
#undef Wv
}

/* ------------------------------------------------------------------ */
/* velocity boundary fill: NS_functions.f90:213-310 (x, then y, then z) */
/* ------------------------------------------------------------------ */
/* generic helper: one face pair of one component along axis ax.
 * mode: 0 periodic copy (normal: end<-first; tangential: ghost wraps),
 *       2 Dirichlet (normal: value; tangential: 2b - interior), 4 ghost value */
static void fill_face(double *A, const int n[3], int ax, int normal, int bc, const double *b1, const double *bn)
{
    int o1 = (ax == 0) ? 1 : 0, o2 = (ax == 2) ? 1 : 2; /* the two other axes */
    if (ax == 1) { o1 = 0; o2 = 2; }
    int N = n[ax];
    for (int q = 1; q <= n[o2]; q++)
        for (int p = 1; p <= n[o1]; p++) {
            int idx1[3], idx2[3];
            idx1[o1] = p; idx1[o2] = q; idx2[o1] = p; idx2[o2] = q;
#define AT(ix) A[IX(ix[0], ix[1], ix[2], n[0], n[1])]
            size_t pl = (size_t)(p - 1) + (size_t)n[o1] * (size_t)(q - 1);
            if (bc == 1) {
                if (normal) { /* last <- first */
                    idx1[ax] = N; idx2[ax] = 1; AT(idx1) = AT(idx2);
                } else {
                    idx1[ax] = 1; idx2[ax] = N - 1; AT(idx1) = AT(idx2);
                    idx1[ax] = N; idx2[ax] = 2; AT(idx1) = AT(idx2);
                }
            } else if (bc == 2) {
                if (normal) {
                    idx1[ax] = 1; AT(idx1) = b1[pl];
                    idx1[ax] = N; AT(idx1) = bn[pl];
                } else {
                    idx1[ax] = 1; idx2[ax] = 2; AT(idx1) = 2 * b1[pl] - AT(idx2);
                    idx1[ax] = N; idx2[ax] = N - 1; AT(idx1) = 2 * bn[pl] - AT(idx2);
                }
            } else if (bc == 4) {
                idx1[ax] = 1; AT(idx1) = b1[pl];
                idx1[ax] = N; AT(idx1) = bn[pl];
            }
#undef AT
        }
}

void nso_vel_bc(double *u, double *v, double *w, int nx, int ny, int nz, const double *planes, int bcx, int bcy, int bcz)
{
    int nu_[3] = {nx + 1, ny + 2, nz + 2}, nv_[3] = {nx + 2, ny + 1, nz + 2}, nw_[3] = {nx + 2, ny + 2, nz + 1};
    int bc[3] = {bcx, bcy, bcz};
    for (int ax = 0; ax < 3; ax++) {
        const double *p0 = planes ? plane(planes, nx, ny, nz, 0 + 2 * ax) : NULL, *p1 = planes ? plane(planes, nx, ny, nz, 1 + 2 * ax) : NULL;
        if (bc[ax] != 1 && bc[ax] != 3 && !planes) continue;
        fill_face(u, nu_, ax, ax == 0, bc[ax], p0, p1);
        p0 = planes ? plane(planes, nx, ny, nz, 6 + 2 * ax) : NULL; p1 = planes ? plane(planes, nx, ny, nz, 7 + 2 * ax) : NULL;
        fill_face(v, nv_, ax, ax == 1, bc[ax], p0, p1);
        p0 = planes ? plane(planes, nx, ny, nz, 12 + 2 * ax) : NULL; p1 = planes ? plane(planes, nx, ny, nz, 13 + 2 * ax) : NULL;
        fill_face(w, nw_, ax, ax == 2, bc[ax], p0, p1);
    }
}

/* vel_bc_staggered_CN2: NS_functions.f90:313-410.  bc==1 zeroes the boundary
 * rows; bc 2/4 put the given planes (the dir-th pair of each component of
 * the packed set) into them. */
void nso_vel_bc_cn2(double *u, double *v, double *w, int nx, int ny, int nz, const double *planes, int bc, int dir)
{
    int n_[3][3] = {{nx + 1, ny + 2, nz + 2}, {nx + 2, ny + 1, nz + 2}, {nx + 2, ny + 2, nz + 1}};
    double *A_[3] = {u, v, w};
    int ax = dir - 1;
    if (bc == 3) return;
    for (int c = 0; c < 3; c++) {
        int *n = n_[c];
        double *A = A_[c];
        int o1 = (ax == 0) ? 1 : 0, o2 = (ax == 2) ? 1 : 2;
        const double *b1 = (bc != 1) ? plane(planes, nx, ny, nz, 6 * c + 2 * ax) : NULL;
        const double *bn = (bc != 1) ? plane(planes, nx, ny, nz, 6 * c + 2 * ax + 1) : NULL;
        for (int q = 1; q <= n[o2]; q++)
            for (int p = 1; p <= n[o1]; p++) {
                int id[3];
                id[o1] = p; id[o2] = q;
                size_t pl = (size_t)(p - 1) + (size_t)n[o1] * (size_t)(q - 1);
                if (bc == 1) {
                    /* periodic: the normal component only zeroes its last row */
                    if (c != ax) { id[ax] = 1; A[IX(id[0], id[1], id[2], n[0], n[1])] = 0.0; }
                    id[ax] = n[ax]; A[IX(id[0], id[1], id[2], n[0], n[1])] = 0.0;
                } else {
                    id[ax] = 1; A[IX(id[0], id[1], id[2], n[0], n[1])] = b1[pl];
                    id[ax] = n[ax]; A[IX(id[0], id[1], id[2], n[0], n[1])] = bn[pl];
                }
            }
    }
}

/* get_velpr_bc / get_pr_bc: NS_functions.f90:414-562.  Writes into the
 * packed plane set.  p_prev is used only when !(dp_flag || using_ustar). */
void nso_get_pr_bc(const double *p, int nx, int ny, int nz, int pbcx, int pbcy, int pbcz, double dx, double dy, double dz, double *planes)
{
    int np[3] = {nx + 2, ny + 2, nz + 2}, pbc[3] = {pbcx, pbcy, pbcz};
    double d[3] = {dx, dy, dz};
    for (int ax = 0; ax < 3; ax++) {
        double *b1 = plane_w(planes, nx, ny, nz, 18 + 2 * ax), *bn = plane_w(planes, nx, ny, nz, 19 + 2 * ax);
        int o1 = (ax == 0) ? 1 : 0, o2 = (ax == 2) ? 1 : 2;
        for (int q = 1; q <= np[o2]; q++)
            for (int pp = 1; pp <= np[o1]; pp++) {
                int i1[3], i2[3], j1[3], j2[3];
                i1[o1] = i2[o1] = j1[o1] = j2[o1] = pp;
                i1[o2] = i2[o2] = j1[o2] = j2[o2] = q;
                i1[ax] = 1; i2[ax] = 2; j1[ax] = np[ax]; j2[ax] = np[ax] - 1;
                double a1 = p[IX(i1[0], i1[1], i1[2], np[0], np[1])], a2 = p[IX(i2[0], i2[1], i2[2], np[0], np[1])];
                double e1 = p[IX(j1[0], j1[1], j1[2], np[0], np[1])], e2 = p[IX(j2[0], j2[1], j2[2], np[0], np[1])];
                size_t pl = (size_t)(pp - 1) + (size_t)np[o1] * (size_t)(q - 1);
                if (pbc[ax] == 2) { b1[pl] = (a1 + a2) / 2; bn[pl] = (e1 + e2) / 2; }
                else if (pbc[ax] == 4) { b1[pl] = a1; bn[pl] = e1; }
                else if (pbc[ax] == 3) { b1[pl] = (a2 - a1) / d[ax]; bn[pl] = (e1 - e2) / d[ax]; }
            }
    }
}

void nso_get_velpr_bc(const double *ubc, const double *vbc, const double *wbc, const double *pbcarr,
                      int bcx, int bcy, int bcz, int pbcx, int pbcy, int pbcz, int nx, int ny, int nz,
                      double dx, double dy, double dz, double dt, int dp_flag, int using_ustar,
                      const double *p_prev, double *planes)
{
    int n_[3][3] = {{nx + 1, ny + 2, nz + 2}, {nx + 2, ny + 1, nz + 2}, {nx + 2, ny + 2, nz + 1}};
    const double *src[3] = {ubc, vbc, wbc};
    int bc[3] = {bcx, bcy, bcz};
    double d[3] = {dx, dy, dz};
    double *corr[3] = {NULL, NULL, NULL};
    int need_corr = !(dp_flag || using_ustar) && (bcx != 1 || bcy != 1 || bcz != 1);
    if (need_corr) {
        /* u_bc_corr = u_bc + diff(p_previous,1,1)/dx*dt  (NS_functions.f90:434-444) */
        int P1 = nx + 2, P2 = ny + 2;
        for (int c = 0; c < 3; c++) {
            int *n = n_[c];
            size_t N = (size_t)n[0] * n[1] * n[2];
            corr[c] = (double *)malloc(N * sizeof(double));
            for (int k = 1; k <= n[2]; k++)
                for (int j = 1; j <= n[1]; j++)
                    for (int i = 1; i <= n[0]; i++) {
                        int a[3] = {i, j, k}, b[3] = {i, j, k};
                        b[c]++;
                        double g = p_prev[IX(b[0], b[1], b[2], P1, P2)] - p_prev[IX(a[0], a[1], a[2], P1, P2)];
                        corr[c][IX(i, j, k, n[0], n[1])] = src[c][IX(i, j, k, n[0], n[1])] + g / d[c] * dt;
                    }
            src[c] = corr[c];
        }
    }
    for (int ax = 0; ax < 3; ax++) {
        if (bc[ax] != 2 && bc[ax] != 4) continue;
        for (int c = 0; c < 3; c++) {
            int *n = n_[c];
            double *b1 = plane_w(planes, nx, ny, nz, 6 * c + 2 * ax), *bn = plane_w(planes, nx, ny, nz, 6 * c + 2 * ax + 1);
            int o1 = (ax == 0) ? 1 : 0, o2 = (ax == 2) ? 1 : 2;
            const double *A = src[c];
            for (int q = 1; q <= n[o2]; q++)
                for (int pp = 1; pp <= n[o1]; pp++) {
                    int i1[3], i2[3], j1[3], j2[3];
                    i1[o1] = i2[o1] = j1[o1] = j2[o1] = pp;
                    i1[o2] = i2[o2] = j1[o2] = j2[o2] = q;
                    i1[ax] = 1; i2[ax] = 2; j1[ax] = n[ax]; j2[ax] = n[ax] - 1;
                    size_t pl = (size_t)(pp - 1) + (size_t)n[o1] * (size_t)(q - 1);
                    double a1 = A[IX(i1[0], i1[1], i1[2], n[0], n[1])], a2 = A[IX(i2[0], i2[1], i2[2], n[0], n[1])];
                    double e1 = A[IX(j1[0], j1[1], j1[2], n[0], n[1])], e2 = A[IX(j2[0], j2[1], j2[2], n[0], n[1])];
                    if (bc[ax] == 4 || c == ax) { b1[pl] = a1; bn[pl] = e1; }
                    else { b1[pl] = (a1 + a2) / 2; bn[pl] = (e1 + e2) / 2; }
                }
        }
    }
    for (int c = 0; c < 3; c++) free(corr[c]);
    nso_get_pr_bc(pbcarr, nx, ny, nz, pbcx, pbcy, pbcz, dx, dy, dz, planes);
}

/* ------------------------------------------------------------------ */
/* pressure: pr_bc_staggered_modify_rhs (NS_functions.f90:1345-1400),   */
/* pr_bc_staggered (:1403-1464)                                         */
/* ------------------------------------------------------------------ */
void nso_pr_bc_modify_rhs(double *rhs, int nx, int ny, int nz, const double *planes, int pbcx, int pbcy, int pbcz,
                          double dx, double dy, double dz)
{
    int n[3] = {nx, ny, nz}, np[3] = {nx + 2, ny + 2, nz + 2}, pbc[3] = {pbcx, pbcy, pbcz};
    double d[3] = {dx, dy, dz}, d2[3] = {dx * dx, dy * dy, dz * dz};
    for (int ax = 0; ax < 3; ax++) {
        if (pbc[ax] == 1) continue;
        const double *b1 = plane(planes, nx, ny, nz, 18 + 2 * ax), *bn = plane(planes, nx, ny, nz, 19 + 2 * ax);
        int o1 = (ax == 0) ? 1 : 0, o2 = (ax == 2) ? 1 : 2;
        for (int q = 1; q <= n[o2]; q++)
            for (int pp = 1; pp <= n[o1]; pp++) {
                int i1[3], in[3];
                i1[o1] = in[o1] = pp; i1[o2] = in[o2] = q;
                i1[ax] = 1; in[ax] = n[ax];
                /* planes are full (nxp,...) sized; use their interior (2:np-1) */
                size_t pl = (size_t)pp + (size_t)np[o1] * (size_t)q;
                double *r1 = &rhs[IX(i1[0], i1[1], i1[2], nx, ny)], *rn = &rhs[IX(in[0], in[1], in[2], nx, ny)];
                if (pbc[ax] == 2) { *r1 = *r1 - 2 * b1[pl] / d2[ax]; *rn = *rn - 2 * bn[pl] / d2[ax]; }
                else if (pbc[ax] == 3) { *r1 = *r1 + b1[pl] / d[ax]; *rn = *rn - bn[pl] / d[ax]; }
                else if (pbc[ax] == 4) { *r1 = *r1 - b1[pl] / d2[ax]; *rn = *rn - bn[pl] / d2[ax]; }
            }
    }
}

void nso_pr_bc(double *dp, int nx, int ny, int nz, const double *planes, int pbcx, int pbcy, int pbcz, double dx, double dy, double dz)
{
    int np[3] = {nx + 2, ny + 2, nz + 2}, pbc[3] = {pbcx, pbcy, pbcz};
    double d[3] = {dx, dy, dz};
    for (int ax = 0; ax < 3; ax++) {
        const double *b1 = (pbc[ax] != 1) ? plane(planes, nx, ny, nz, 18 + 2 * ax) : NULL;
        const double *bn = (pbc[ax] != 1) ? plane(planes, nx, ny, nz, 19 + 2 * ax) : NULL;
        int o1 = (ax == 0) ? 1 : 0, o2 = (ax == 2) ? 1 : 2;
        for (int q = 1; q <= np[o2]; q++)
            for (int pp = 1; pp <= np[o1]; pp++) {
                int i1[3], i2[3], j1[3], j2[3];
                i1[o1] = i2[o1] = j1[o1] = j2[o1] = pp;
                i1[o2] = i2[o2] = j1[o2] = j2[o2] = q;
                size_t pl = (size_t)(pp - 1) + (size_t)np[o1] * (size_t)(q - 1);
#define DP(ix) dp[IX(ix[0], ix[1], ix[2], np[0], np[1])]
                if (pbc[ax] == 1) {
                    i1[ax] = 1; i2[ax] = np[ax] - 1; DP(i1) = DP(i2);
                    j1[ax] = np[ax]; j2[ax] = 2; DP(j1) = DP(j2);
                } else if (pbc[ax] == 2) {
                    i1[ax] = 1; i2[ax] = 2; DP(i1) = 2 * b1[pl] - DP(i2);
                    j1[ax] = np[ax]; j2[ax] = np[ax] - 1; DP(j1) = 2 * bn[pl] - DP(j2);
                } else if (pbc[ax] == 3) {
                    i1[ax] = 1; i2[ax] = 2; DP(i1) = -d[ax] * b1[pl] + DP(i2);
                    j1[ax] = np[ax]; j2[ax] = np[ax] - 1; DP(j1) = d[ax] * bn[pl] + DP(j2);
                } else if (pbc[ax] == 4) {
                    i1[ax] = 1; DP(i1) = b1[pl];
                    j1[ax] = np[ax]; DP(j1) = bn[pl];
                }
#undef DP
            }
    }
}

/* ------------------------------------------------------------------ */
/* CN matrices and dense LU (mat_CN NS_functions.f90:1245-1294;        */
/* getrf/getrs as called at big_periodic_3D.f90:181-186,425-483)       */
/* ------------------------------------------------------------------ */
/* A (n+1)^2 and B (n+2)^2, column-major.  cyclic!=0 replaces mat_CN's
 * periodic row-1 corner column n+1 by column n (the wrap stencil). */
void nso_mat_cn(int n, int bc, double dx2, double dt, double nu, int cyclic, double *A, double *B)
{
    int NA = n + 1, NB = n + 2;
    double idx2 = 1.0 / dx2;
    memset(A, 0, sizeof(double) * NA * NA);
    memset(B, 0, sizeof(double) * NB * NB);
#define AM(r, c) A[(size_t)((r) - 1) + (size_t)NA * ((c) - 1)]
#define BM(r, c) B[(size_t)((r) - 1) + (size_t)NB * ((c) - 1)]
    double dg = 1.0 + 0.5 * dt * nu * 2.0 * idx2, of = -(0.5 * dt * nu * 1.0 * idx2);
    for (int i = 2; i <= n; i++) { AM(i, i) = dg; AM(i, i - 1) = of; AM(i, i + 1) = of; }
    for (int i = 2; i <= n + 1; i++) { BM(i, i) = dg; BM(i, i - 1) = of; BM(i, i + 1) = of; }
    if (bc == 1) {
        AM(1, 1) = dg; AM(1, 2) = of; AM(1, cyclic ? n : n + 1) = of;
        AM(n + 1, n + 1) = 1.0; AM(n + 1, 1) = -1.0;
        BM(1, 1) = 1.0; BM(1, n + 1) = -1.0;
        BM(n + 2, n + 2) = 1.0; BM(n + 2, 2) = -1.0;
    } else if (bc == 2) {
        AM(1, 1) = 1.0; AM(n + 1, n + 1) = 1.0;
        BM(1, 1) = 0.5; BM(1, 2) = 0.5;
        BM(n + 2, n + 2) = 0.5; BM(n + 2, n + 1) = 0.5;
    } else if (bc == 4) {
        AM(1, 1) = 1.0; AM(n + 1, n + 1) = 1.0;
        BM(1, 1) = 1.0; BM(n + 2, n + 2) = 1.0;
    }
#undef AM
#undef BM
}

/* LU with partial pivoting, in place (column-major N x N). */
static void lu_factor(double *M, int N, int *piv)
{
    for (int c = 0; c < N; c++) {
        int pr = c;
        double best = fabs(M[c + (size_t)N * c]);
        for (int r = c + 1; r < N; r++) {
            double a = fabs(M[r + (size_t)N * c]);
            if (a > best) { best = a; pr = r; }
        }
        piv[c] = pr;
        if (pr != c)
            for (int k = 0; k < N; k++) { double t = M[c + (size_t)N * k]; M[c + (size_t)N * k] = M[pr + (size_t)N * k]; M[pr + (size_t)N * k] = t; }
        double inv = 1.0 / M[c + (size_t)N * c];
        for (int r = c + 1; r < N; r++) M[r + (size_t)N * c] *= inv;
        for (int k = c + 1; k < N; k++) {
            double m = M[c + (size_t)N * k];
            if (m != 0.0)
                for (int r = c + 1; r < N; r++) M[r + (size_t)N * k] -= M[r + (size_t)N * c] * m;
        }
    }
}
static void lu_solve(const double *M, int N, const int *piv, double *x)
{
    for (int c = 0; c < N; c++) { int p = piv[c]; if (p != c) { double t = x[c]; x[c] = x[p]; x[p] = t; } }
    for (int c = 0; c < N; c++) { double xc = x[c]; for (int r = c + 1; r < N; r++) x[r] -= M[r + (size_t)N * c] * xc; }
    for (int c = N - 1; c >= 0; c--) { x[c] /= M[c + (size_t)N * c]; double xc = x[c]; for (int r = 0; r < c; r++) x[r] -= M[r + (size_t)N * c] * xc; }
}

/* One ADI sweep along `dir` over every line of u, v, w (including ghost
 * lines), as at big_periodic_3D.f90:425-483 / ...interBC_combine.f90:656-701. */
static void adi_sweep(double *u, double *v, double *w, const nso_config *cfg, int dir)
{
    int nx = cfg->nx, ny = cfg->ny, nz = cfg->nz;
    int n_[3][3] = {{nx + 1, ny + 2, nz + 2}, {nx + 2, ny + 1, nz + 2}, {nx + 2, ny + 2, nz + 1}};
    double *A_[3] = {u, v, w};
    int ax = dir - 1;
    int nn = (ax == 0) ? nx : (ax == 1) ? ny : nz;
    double h = (ax == 0) ? cfg->dx : (ax == 1) ? cfg->dy : cfg->dz;
    int NA = nn + 1, NB = nn + 2;
    double *MA = (double *)malloc(sizeof(double) * NA * NA), *MB = (double *)malloc(sizeof(double) * NB * NB);
    int *pa = (int *)malloc(sizeof(int) * NA), *pb = (int *)malloc(sizeof(int) * NB);
    nso_mat_cn(nn, cfg->bc[ax], h * h, cfg->dt, cfg->nu, (cfg->flags & NSO_FLAG_CN_CYCLIC) != 0, MA, MB);
    lu_factor(MA, NA, pa);
    lu_factor(MB, NB, pb);
    for (int c = 0; c < 3; c++) {
        int *n = n_[c];
        double *A = A_[c];
        const double *M = (c == ax) ? MA : MB;
        const int *pv = (c == ax) ? pa : pb;
        int N = n[ax];
        int o1 = (ax == 0) ? 1 : 0, o2 = (ax == 2) ? 1 : 2;
        size_t stride = (ax == 0) ? 1 : (ax == 1) ? (size_t)n[0] : (size_t)n[0] * n[1];
#pragma omp parallel
        {
            double *line = (double *)malloc(sizeof(double) * N);
#pragma omp for collapse(2) schedule(static)
            for (int q = 1; q <= n[o2]; q++)
                for (int p = 1; p <= n[o1]; p++) {
                    int id[3];
                    id[o1] = p; id[o2] = q; id[ax] = 1;
                    size_t base = IX(id[0], id[1], id[2], n[0], n[1]);
                    for (int t = 0; t < N; t++) line[t] = A[base + stride * t];
                    lu_solve(M, N, pv, line);
                    for (int t = 0; t < N; t++) A[base + stride * t] = line[t];
                }
            free(line);
        }
    }
    free(MA); free(MB); free(pa); free(pb);
}

void nso_adi_sweep(double *u, double *v, double *w, const nso_config *cfg, int dir) { adi_sweep(u, v, w, cfg, dir); }

/* ------------------------------------------------------------------ */
/* FFT (hand-written radix-2 + direct DFT fallback) for the periodic    */
/* Poisson solve (big_periodic_3D.f90:215-256,526-548)                  */
/* ------------------------------------------------------------------ */
static int is_pow2(int n) { return n > 0 && (n & (n - 1)) == 0; }

/* in-place complex FFT of length n, sign -1 forward / +1 backward (unscaled) */
static void fft1(double complex *x, int n, int sign, double complex *tmp)
{
    if (is_pow2(n)) {
        for (int i = 1, j = 0; i < n; i++) {
            int bit = n >> 1;
            for (; j & bit; bit >>= 1) j ^= bit;
            j ^= bit;
            if (i < j) { double complex t = x[i]; x[i] = x[j]; x[j] = t; }
        }
        for (int len = 2; len <= n; len <<= 1) {
            int half = len >> 1;
            for (int k = 0; k < half; k++) {
                double ang = sign * 2.0 * PI * (double)k / (double)len;
                double complex wk = cos(ang) + CI * sin(ang);
                for (int s = 0; s < n; s += len) {
                    double complex a = x[s + k], b = x[s + k + half] * wk;
                    x[s + k] = a + b;
                    x[s + k + half] = a - b;
                }
            }
        }
    } else {
        for (int k = 0; k < n; k++) {
            double complex acc = 0;
            for (int j = 0; j < n; j++) {
                long m = ((long)j * k) % n;
                double ang = sign * 2.0 * PI * (double)m / (double)n;
                acc += x[j] * (cos(ang) + CI * sin(ang));
            }
            tmp[k] = acc;
        }
        memcpy(x, tmp, sizeof(double complex) * n);
    }
}

/* 3-D complex FFT over an (n1,n2,n3) column-major array */
static void fft3(double complex *a, int n1, int n2, int n3, int sign)
{
    int nmax = n1 > n2 ? n1 : n2;
    nmax = nmax > n3 ? nmax : n3;
#pragma omp parallel
    {
        double complex *line = (double complex *)malloc(sizeof(double complex) * nmax);
        double complex *tmp = (double complex *)malloc(sizeof(double complex) * nmax);
#pragma omp for schedule(static)
        for (long l = 0; l < (long)n2 * n3; l++) fft1(a + (size_t)l * n1, n1, sign, tmp);
#pragma omp for collapse(2) schedule(static)
        for (int k = 0; k < n3; k++)
            for (int i = 0; i < n1; i++) {
                double complex *b = a + (size_t)i + (size_t)n1 * n2 * k;
                for (int j = 0; j < n2; j++) line[j] = b[(size_t)n1 * j];
                fft1(line, n2, sign, tmp);
                for (int j = 0; j < n2; j++) b[(size_t)n1 * j] = line[j];
            }
#pragma omp for collapse(2) schedule(static)
        for (int j = 0; j < n2; j++)
            for (int i = 0; i < n1; i++) {
                double complex *b = a + (size_t)i + (size_t)n1 * j;
                for (int k = 0; k < n3; k++) line[k] = b[(size_t)n1 * n2 * k];
                fft1(line, n3, sign, tmp);
                for (int k = 0; k < n3; k++) b[(size_t)n1 * n2 * k] = line[k];
            }
        free(line); free(tmp);
    }
}

void nso_fft3(double *re, double *im, int n1, int n2, int n3, int sign)
{
    size_t N = (size_t)n1 * n2 * n3;
    double complex *a = (double complex *)malloc(sizeof(double complex) * N);
    for (size_t q = 0; q < N; q++) a[q] = re[q] + CI * im[q];
    fft3(a, n1, n2, n3, sign);
    for (size_t q = 0; q < N; q++) { re[q] = creal(a[q]); im[q] = cimag(a[q]); }
    free(a);
}

/* Periodic Poisson: forward DFT (unscaled), divide by
 * eig = -4*((sin(pi*(i-1)/nx)/dx)^2 + ...), mode (1,1,1) -> 0 (eig=+inf),
 * backward DFT scaled by 1/(nx*ny*nz).  in: rhs(nx,ny,nz), out: sol(nx,ny,nz). */
void nso_poisson_fft(const double *rhs, double *sol, int nx, int ny, int nz, double dx, double dy, double dz)
{
    size_t N = (size_t)nx * ny * nz;
    double complex *a = (double complex *)malloc(sizeof(double complex) * N);
    double *ex = (double *)malloc(sizeof(double) * nx), *ey = (double *)malloc(sizeof(double) * ny), *ez = (double *)malloc(sizeof(double) * nz);
    for (int i = 1; i <= nx; i++) { double s = sin(PI * (i - 1) / nx) / dx; ex[i - 1] = s * s; }
    for (int i = 1; i <= ny; i++) { double s = sin(PI * (i - 1) / ny) / dy; ey[i - 1] = s * s; }
    for (int i = 1; i <= nz; i++) { double s = sin(PI * (i - 1) / nz) / dz; ez[i - 1] = s * s; }
    for (size_t q = 0; q < N; q++) a[q] = rhs[q];
    fft3(a, nx, ny, nz, -1);
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = 0; k < nz; k++)
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++) {
                size_t q = (size_t)i + (size_t)nx * (j + (size_t)ny * k);
                if (i == 0 && j == 0 && k == 0) { a[q] = 0; continue; }
                double e = ex[i];
                e = e + ey[j];
                e = e + ez[k];
                e = -4.0 * e;
                a[q] = creal(a[q]) / e + CI * (cimag(a[q]) / e);
            }
    fft3(a, nx, ny, nz, +1);
    double sc = 1.0 / ((double)nx * ny * nz);
    for (size_t q = 0; q < N; q++) sol[q] = creal(a[q]) * sc;
    free(a); free(ex); free(ey); free(ez);
}

/* ------------------------------------------------------------------ */
/* DCT Poisson: DCT_poisson_solver NS_functions.f90:1505-1630 with the  */
/* MKL staggered cosine transform (backward = unnormalised DCT-II,      */
/* forward = its exact inverse) and the DST-II built from it by sign    */
/* flip + reversal (:1522-1525).  Eigenvalues: ...interBC:321-351.      */
/* ------------------------------------------------------------------ */
static double dct_cos(int k, int j, int n) /* cos(pi*k*(2j+1)/(2n)) with exact argument reduction */
{
    long m = ((long)k * (2 * j + 1)) % (4L * n);
    return cos(PI * (double)m / (2.0 * n));
}
/* unnormalised DCT-II: X_k = sum_j x_j cos(pi k (2j+1)/(2n)) */
static void dct2(const double *x, double *X, int n)
{
    for (int k = 0; k < n; k++) {
        double s = 0;
        for (int j = 0; j < n; j++) s += x[j] * dct_cos(k, j, n);
        X[k] = s;
    }
}
/* exact inverse: x_j = X_0/n + (2/n) sum_{k>=1} X_k cos(...) */
static void idct2(const double *X, double *x, int n)
{
    for (int j = 0; j < n; j++) {
        double s = 0;
        for (int k = 1; k < n; k++) s += X[k] * dct_cos(k, j, n);
        x[j] = X[0] / n + 2.0 * s / n;
    }
}
/* forward transform of one line for pressure BC code pbc (2: DST-II, 3: DCT-II) */
static void tt_forward(double *f, int n, int pbc, double *t)
{
    if (pbc == 2) {
        for (int j = 0; j < n; j += 2) f[j] = -f[j];
        dct2(f, t, n);
        for (int j = 0; j < n; j++) f[j] = t[n - 1 - j];
    } else {
        dct2(f, t, n);
        memcpy(f, t, sizeof(double) * n);
    }
}
static void tt_backward(double *f, int n, int pbc, double *t)
{
    if (pbc == 2) {
        for (int j = 0; j < n; j++) t[j] = f[n - 1 - j];
        idct2(t, f, n);
        for (int j = 0; j < n; j += 2) f[j] = -f[j];
    } else {
        idct2(f, t, n);
        memcpy(f, t, sizeof(double) * n);
    }
}

void nso_dct1(const double *x, double *X, int n, int pbc, int inverse)
{
    double *f = (double *)malloc(sizeof(double) * n), *t = (double *)malloc(sizeof(double) * n);
    memcpy(f, x, sizeof(double) * n);
    if (inverse) tt_backward(f, n, pbc, t); else tt_forward(f, n, pbc, t);
    memcpy(X, f, sizeof(double) * n);
    free(f); free(t);
}

static double tt_eig(int i /*1-based*/, int n, int pbc, double d2)
{
    double s;
    if (pbc == 2) s = sin(i * PI / 2 / n);
    else s = sin((i - 1) * PI / 2 / n);
    return s * s / d2;
}

/* in-place on rhs(nx,ny,nz).  Supports pbc in {2,3} on every axis. */
void nso_poisson_dct(double *rhs, int nx, int ny, int nz, int pbcx, int pbcy, int pbcz,
                     double dx, double dy, double dz, double mean_dp)
{
    int n[3] = {nx, ny, nz}, pbc[3] = {pbcx, pbcy, pbcz};
    size_t st[3] = {1, (size_t)nx, (size_t)nx * ny};
    for (int pass = 0; pass < 6; pass++) {
        int ax = pass < 3 ? pass : 5 - pass;
        int inv = pass >= 3;
        int o1 = (ax == 0) ? 1 : 0, o2 = (ax == 2) ? 1 : 2;
        if (pass == 3) {
            /* divide by the eigenvalues; the zero mode carries mean_dp (NS_functions.f90:1576-1577) */
            double d2[3] = {dx * dx, dy * dy, dz * dz};
            int zero_mode = 0;
#pragma omp parallel for collapse(2) schedule(static)
            for (int k = 1; k <= nz; k++)
                for (int j = 1; j <= ny; j++)
                    for (int i = 1; i <= nx; i++) {
                        double e = tt_eig(i, nx, pbcx, d2[0]);
                        e = e + tt_eig(j, ny, pbcy, d2[1]);
                        e = e + tt_eig(k, nz, pbcz, d2[2]);
                        e = -4.0 * e;
                        size_t q = IX(i, j, k, nx, ny);
                        if (e == 0.0) rhs[q] = rhs[q] / INFINITY;
                        else rhs[q] = rhs[q] / e;
                    }
            if (tt_eig(1, nx, pbcx, 1) == 0 && tt_eig(1, ny, pbcy, 1) == 0 && tt_eig(1, nz, pbcz, 1) == 0) zero_mode = 1;
            if (zero_mode) rhs[0] = mean_dp * nx * ny * nz;
        }
#pragma omp parallel
        {
            int N = n[ax];
            double *line = (double *)malloc(sizeof(double) * N), *t = (double *)malloc(sizeof(double) * N);
#pragma omp for collapse(2) schedule(static)
            for (int q = 0; q < n[o2]; q++)
                for (int p = 0; p < n[o1]; p++) {
                    size_t base = st[o1] * p + st[o2] * q;
                    for (int s = 0; s < N; s++) line[s] = rhs[base + st[ax] * s];
                    if (inv) tt_backward(line, N, pbc[ax], t); else tt_forward(line, N, pbc[ax], t);
                    for (int s = 0; s < N; s++) rhs[base + st[ax] * s] = line[s];
                }
            free(line); free(t);
        }
    }
}

/* ------------------------------------------------------------------ */
/* divergence RHS (big_periodic_3D.f90:498-501) and projection (:562-567) */
/* ------------------------------------------------------------------ */
void nso_divergence(const double *u, const double *v, const double *w, int nx, int ny, int nz,
                    double dx, double dy, double dz, double dt, double *rhs)
{
    const int U1 = nx + 1, U2 = ny + 2, V1 = nx + 2, V2 = ny + 1, W1 = nx + 2, W2 = ny + 2;
#pragma omp parallel for collapse(2) schedule(static)
    for (int c = 1; c <= nz; c++)
        for (int b = 1; b <= ny; b++)
            for (int a = 1; a <= nx; a++) {
                double r = (u[IX(a + 1, b + 1, c + 1, U1, U2)] - u[IX(a, b + 1, c + 1, U1, U2)]) / dx;
                r = r + (v[IX(a + 1, b + 1, c + 1, V1, V2)] - v[IX(a + 1, b, c + 1, V1, V2)]) / dy;
                r = r + (w[IX(a + 1, b + 1, c + 1, W1, W2)] - w[IX(a + 1, b + 1, c, W1, W2)]) / dz;
                rhs[IX(a, b, c, nx, ny)] = r / dt;
            }
}

/* u = -dt*diff(dp,1,1)/dx + u_star over the full ghosted arrays; p = dp_flag*p + dp */
void nso_project(double *u, double *v, double *w, double *p, const double *us, const double *vs, const double *ws,
                 const double *dp, int nx, int ny, int nz, double dx, double dy, double dz, double dt, int dp_flag)
{
    int P1 = nx + 2, P2 = ny + 2;
    int n_[3][3] = {{nx + 1, ny + 2, nz + 2}, {nx + 2, ny + 1, nz + 2}, {nx + 2, ny + 2, nz + 1}};
    double *o[3] = {u, v, w};
    const double *s[3] = {us, vs, ws};
    double d[3] = {dx, dy, dz};
    for (int c = 0; c < 3; c++) {
        int *n = n_[c];
#pragma omp parallel for collapse(2) schedule(static)
        for (int k = 1; k <= n[2]; k++)
            for (int j = 1; j <= n[1]; j++)
                for (int i = 1; i <= n[0]; i++) {
                    int b[3] = {i, j, k};
                    b[c]++;
                    double g = dp[IX(b[0], b[1], b[2], P1, P2)] - dp[IX(i, j, k, P1, P2)];
                    size_t q = IX(i, j, k, n[0], n[1]);
                    o[c][q] = s[c][q] - dt * g / d[c];
                }
    }
    size_t NP = (size_t)(nx + 2) * (ny + 2) * (nz + 2);
    for (size_t q = 0; q < NP; q++) p[q] = dp_flag * p[q] + dp[q];
}


/* ------------------------------------------------------------------ */
/* parent-run capture of a sub-box's plane set (big_periodic_3D.f90:  */
/* 433-489: post-sweep planes; get_velpr_bc/get_pr_bc semantics)      */
/* ------------------------------------------------------------------ */
static struct { int on, x0, y0, z0, nx, ny, nz; double *dst; } g_cap;

void nso_set_capture(int x0, int y0, int z0, int nx0, int ny0, int nz0, double *dst)
{
    g_cap.on = dst != NULL;
    g_cap.x0 = x0; g_cap.y0 = y0; g_cap.z0 = z0;
    g_cap.nx = nx0; g_cap.ny = ny0; g_cap.nz = nz0;
    g_cap.dst = dst;
}

/* value of parent component array A (dims np) at sub-box 1-based index (i,j,k) */
static double sub_at(const double *A, const int np[3], int i, int j, int k)
{
    return A[IX(i + g_cap.x0 - 1, j + g_cap.y0 - 1, k + g_cap.z0 - 1, np[0], np[1])];
}

static void capture_vel_axis(const double *us, const double *vs, const double *ws, const nso_config *cfg, int ax)
{
    int nx = cfg->nx, ny = cfg->ny, nz = cfg->nz, sx = g_cap.nx, sy = g_cap.ny, sz = g_cap.nz;
    int npar[3][3] = {{nx + 1, ny + 2, nz + 2}, {nx + 2, ny + 1, nz + 2}, {nx + 2, ny + 2, nz + 1}};
    int nsub[3][3] = {{sx + 1, sy + 2, sz + 2}, {sx + 2, sy + 1, sz + 2}, {sx + 2, sy + 2, sz + 1}};
    const double *src[3] = {us, vs, ws};
    for (int c = 0; c < 3; c++) {
        int *n = nsub[c];
        double *b1 = plane_w(g_cap.dst, sx, sy, sz, 6 * c + 2 * ax), *bn = plane_w(g_cap.dst, sx, sy, sz, 6 * c + 2 * ax + 1);
        int o1 = (ax == 0) ? 1 : 0, o2 = (ax == 2) ? 1 : 2;
        for (int q = 1; q <= n[o2]; q++)
            for (int pp = 1; pp <= n[o1]; pp++) {
                int a[4][3];
                for (int t = 0; t < 4; t++) { a[t][o1] = pp; a[t][o2] = q; }
                a[0][ax] = 1; a[1][ax] = 2; a[2][ax] = n[ax]; a[3][ax] = n[ax] - 1;
                double v1 = sub_at(src[c], npar[c], a[0][0], a[0][1], a[0][2]);
                double v2 = sub_at(src[c], npar[c], a[1][0], a[1][1], a[1][2]);
                double e1 = sub_at(src[c], npar[c], a[2][0], a[2][1], a[2][2]);
                double e2 = sub_at(src[c], npar[c], a[3][0], a[3][1], a[3][2]);
                size_t pl = (size_t)(pp - 1) + (size_t)n[o1] * (size_t)(q - 1);
                if (c == ax) { b1[pl] = v1; bn[pl] = e1; }
                else { b1[pl] = (v1 + v2) / 2; bn[pl] = (e1 + e2) / 2; }
            }
    }
}

static void capture_pressure(const double *dp, const nso_config *cfg)
{
    int nx = cfg->nx, ny = cfg->ny, sx = g_cap.nx, sy = g_cap.ny, sz = g_cap.nz;
    int npar[3] = {nx + 2, ny + 2, cfg->nz + 2}, n[3] = {sx + 2, sy + 2, sz + 2};
    double d[3] = {cfg->dx, cfg->dy, cfg->dz};
    for (int ax = 0; ax < 3; ax++) {
        double *b1 = plane_w(g_cap.dst, sx, sy, sz, 18 + 2 * ax), *bn = plane_w(g_cap.dst, sx, sy, sz, 19 + 2 * ax);
        int o1 = (ax == 0) ? 1 : 0, o2 = (ax == 2) ? 1 : 2;
        for (int q = 1; q <= n[o2]; q++)
            for (int pp = 1; pp <= n[o1]; pp++) {
                int a[4][3];
                for (int t = 0; t < 4; t++) { a[t][o1] = pp; a[t][o2] = q; }
                a[0][ax] = 1; a[1][ax] = 2; a[2][ax] = n[ax]; a[3][ax] = n[ax] - 1;
                double v1 = sub_at(dp, npar, a[0][0], a[0][1], a[0][2]), v2 = sub_at(dp, npar, a[1][0], a[1][1], a[1][2]);
                double e1 = sub_at(dp, npar, a[2][0], a[2][1], a[2][2]), e2 = sub_at(dp, npar, a[3][0], a[3][1], a[3][2]);
                size_t pl = (size_t)(pp - 1) + (size_t)n[o1] * (size_t)(q - 1);
                b1[pl] = (v2 - v1) / d[ax];
                bn[pl] = (e1 - e2) / d[ax];
            }
    }
    double s = 0;
    for (int k = 2; k <= sz + 1; k++)
        for (int j = 2; j <= sy + 1; j++)
            for (int i = 2; i <= sx + 1; i++) s += sub_at(dp, npar, i, j, k);
    long s24[24], o25[25];
    nso_plane_layout(sx, sy, sz, s24, o25);
    g_cap.dst[o25[24]] = s / ((double)sx * sy * sz);
}

/* ------------------------------------------------------------------ */
/* one time step: the loop body of big_periodic_3D.f90:326-571 and      */
/* re_simulation_sub_tstep_interBC_combine.f90:508-802                  */
/* ------------------------------------------------------------------ */
/* state arrays: u,v,w,p (in/out), hx,hy,hz = rhs_*_previous (in/out),
 * us,vs,ws,dp (out: u*, dp of this step).  planes: packed set (may be NULL
 * when every code is periodic).  first!=0: Euler-type first step. */
int nso_step(const nso_config *cfg, double *u, double *v, double *w, double *p,
             double *hx, double *hy, double *hz, double *us, double *vs, double *ws, double *dp,
             const double *planes, int first)
{
    int nx = cfg->nx, ny = cfg->ny, nz = cfg->nz;
    int bcx = cfg->bc[0], bcy = cfg->bc[1], bcz = cfg->bc[2];
    double dx = cfg->dx, dy = cfg->dy, dz = cfg->dz, dt = cfg->dt, nu = cfg->nu;
    int all_periodic_p = cfg->pbc[0] == 1 && cfg->pbc[1] == 1 && cfg->pbc[2] == 1;
    int none_periodic_p = cfg->pbc[0] != 1 && cfg->pbc[1] != 1 && cfg->pbc[2] != 1;
    if (!all_periodic_p && !none_periodic_p) return -2;
    int mx = bcx == 1 ? nx : nx - 1, my = bcy == 1 ? ny : ny - 1, mz = bcz == 1 ? nz : nz - 1;
    size_t NU = (size_t)(nx + 1) * (ny + 2) * (nz + 2), NV = (size_t)(nx + 2) * (ny + 1) * (nz + 2), NW = (size_t)(nx + 2) * (ny + 2) * (nz + 1);
    size_t NP = (size_t)(nx + 2) * (ny + 2) * (nz + 2), NI = (size_t)nx * ny * nz;
    double *cx = malloc(sizeof(double) * (size_t)mx * ny * nz), *cy = malloc(sizeof(double) * (size_t)nx * my * nz), *cz = malloc(sizeof(double) * (size_t)nx * ny * mz);
    double *fx = malloc(sizeof(double) * (size_t)mx * ny * nz), *fy = malloc(sizeof(double) * (size_t)nx * my * nz), *fz = malloc(sizeof(double) * (size_t)nx * ny * mz);
    nso_cal_conv(u, v, w, nx, ny, nz, bcx, bcy, bcz, dx, dy, dz, cx, cy, cz);
    nso_cal_diff(u, v, w, nx, ny, nz, bcx, bcy, bcz, dx * dx, dy * dy, dz * dz, fx, fy, fz);

    /* predictor over the full ghosted arrays; rhs is zero outside the ranges */
    double *S[3] = {us, vs, ws}, *H[3] = {hx, hy, hz};
    const double *V0[3] = {u, v, w}, *C[3] = {cx, cy, cz}, *D[3] = {fx, fy, fz};
    int n_[3][3] = {{nx + 1, ny + 2, nz + 2}, {nx + 2, ny + 1, nz + 2}, {nx + 2, ny + 2, nz + 1}};
    int m_[3][3] = {{mx, ny, nz}, {nx, my, nz}, {nx, ny, mz}};
    int bcv[3] = {bcx, bcy, bcz};
    double dd[3] = {dx, dy, dz};
    int P1 = nx + 2, P2 = ny + 2;
    for (int c = 0; c < 3; c++) {
        int *n = n_[c];
        int lo[3], hi[3];
        for (int a = 0; a < 3; a++) {
            if (a == c) { lo[a] = (bcv[a] == 1) ? 1 : 2; hi[a] = (bcv[a] == 1) ? (a == 0 ? nx : a == 1 ? ny : nz) : (a == 0 ? nx : a == 1 ? ny : nz); }
            else { lo[a] = 2; hi[a] = (a == 0 ? nx : a == 1 ? ny : nz) + 1; }
        }
#pragma omp parallel for collapse(2) schedule(static)
        for (int k = 1; k <= n[2]; k++)
            for (int j = 1; j <= n[1]; j++)
                for (int i = 1; i <= n[0]; i++) {
                    size_t q = IX(i, j, k, n[0], n[1]);
                    int in = i >= lo[0] && i <= hi[0] && j >= lo[1] && j <= hi[1] && k >= lo[2] && k <= hi[2];
                    double cv = 0, df = 0;
                    if (in) {
                        size_t r = IX(i - lo[0] + 1, j - lo[1] + 1, k - lo[2] + 1, m_[c][0], m_[c][1]);
                        cv = C[c][r]; df = D[c][r];
                    }
                    double dpd = 0.0;
                    if (cfg->dp_flag) {
                        int b[3] = {i, j, k};
                        b[c]++;
                        dpd = (p[IX(b[0], b[1], b[2], P1, P2)] - p[IX(i, j, k, P1, P2)]) / dd[c];
                    }
                    double dpf = (double)cfg->dp_flag;
                    double rhs, val;
                    if (cfg->scheme == NSO_EULER) {
                        rhs = in ? nu * df - cv : 0.0;
                        val = dt * (1.0 * rhs - dpf * dpd) + V0[c][q];
                        H[c][q] = 0.0;
                    } else if (cfg->scheme == NSO_AB2) {
                        rhs = in ? nu * df - cv : 0.0;
                        if (first) val = dt * (1.0 * rhs - dpf * dpd) + V0[c][q];
                        else val = dt * (1.5 * rhs - 0.5 * H[c][q] - dpf * dpd) + V0[c][q];
                        H[c][q] = rhs;
                    } else {
                        if (first) {
                            rhs = in ? 0.5 * nu * df - cv : 0.0;
                            val = dt * (rhs - dpf * dpd) + V0[c][q];
                        } else {
                            rhs = in ? 0.5 * nu * df - 1.5 * cv : 0.0;
                            val = dt * (rhs - 0.5 * H[c][q] - dpf * dpd) + V0[c][q];
                        }
                        if (in) H[c][q] = -cv;
                    }
                    S[c][q] = val;
                }
    }
    free(cx); free(cy); free(cz); free(fx); free(fy); free(fz);

    if (cfg->scheme == NSO_AB2_CN) {
        for (int dir = 1; dir <= 3; dir++) {
            nso_vel_bc_cn2(us, vs, ws, nx, ny, nz, planes, bcv[dir - 1], dir);
            adi_sweep(us, vs, ws, cfg, dir);
            if (g_cap.on) capture_vel_axis(us, vs, ws, cfg, dir - 1);
        }
    } else {
        nso_vel_bc(us, vs, ws, nx, ny, nz, planes, bcx, bcy, bcz);
        if (g_cap.on)
            for (int ax = 0; ax < 3; ax++) capture_vel_axis(us, vs, ws, cfg, ax);
    }

    double *rhs = malloc(sizeof(double) * NI);
    nso_divergence(us, vs, ws, nx, ny, nz, dx, dy, dz, dt, rhs);
    if (!all_periodic_p) nso_pr_bc_modify_rhs(rhs, nx, ny, nz, planes, cfg->pbc[0], cfg->pbc[1], cfg->pbc[2], dx, dy, dz);
    double *sol = malloc(sizeof(double) * NI);
    if (all_periodic_p) {
        nso_poisson_fft(rhs, sol, nx, ny, nz, dx, dy, dz);
    } else {
        long s24[24], o25[25];
        nso_plane_layout(nx, ny, nz, s24, o25);
        double mean_dp = planes[o25[24]];
        memcpy(sol, rhs, sizeof(double) * NI);
        nso_poisson_dct(sol, nx, ny, nz, cfg->pbc[0], cfg->pbc[1], cfg->pbc[2], dx, dy, dz, mean_dp);
    }
    for (int k = 1; k <= nz; k++)
        for (int j = 1; j <= ny; j++)
            for (int i = 1; i <= nx; i++) dp[IX(i + 1, j + 1, k + 1, P1, P2)] = sol[IX(i, j, k, nx, ny)];
    nso_pr_bc(dp, nx, ny, nz, planes, cfg->pbc[0], cfg->pbc[1], cfg->pbc[2], dx, dy, dz);
    if (cfg->flags & NSO_FLAG_GAUGE_CORNER) {
        double c0 = dp[0];
        for (size_t q = 0; q < NP; q++) dp[q] = dp[q] - c0;
    }
    if (g_cap.on) capture_pressure(dp, cfg);
    nso_project(u, v, w, p, us, vs, ws, dp, nx, ny, nz, dx, dy, dz, dt, cfg->dp_flag);
    free(rhs); free(sol);
    (void)NU; (void)NV; (void)NW;
    return 0;
}

/* max |div u| over interior cells (big_periodic_3D.f90:576-579) */
double nso_max_div(const double *u, const double *v, const double *w, int nx, int ny, int nz, double dx, double dy, double dz)
{
    double *r = malloc(sizeof(double) * (size_t)nx * ny * nz), m = 0;
    nso_divergence(u, v, w, nx, ny, nz, dx, dy, dz, 1.0, r);
    for (size_t q = 0; q < (size_t)nx * ny * nz; q++) if (fabs(r[q]) > m) m = fabs(r[q]);
    free(r);
    return m;
}

/* linear time interpolation of a packed plane set: out = (1-th)*a + th*b */
void nso_bc_interp_linear(const double *a, const double *b, double th, double *out, long n)
{
    for (long q = 0; q < n; q++) out[q] = (1.0 - th) * a[q] + th * b[q];
}

int nso_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
