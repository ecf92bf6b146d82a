// Discrete adjoint of one time step (Euler / AB2 / AB2-CN predictor, velocity
// fill or, for AB2-CN, the boundary rows and ADI sweeps, divergence +
// boundary RHS terms, Poisson solve, pressure fill, projection), FP64, for
// gradients of a function of the final state with respect to the initial
// state and to every step's boundary planes.
//
// Forward step (as ns_step runs it without lazy projection):
//   S1 u* = u + dt(c1 rhs(u) - c2 H - dp_flag grad p),  H' = rhs(u) on the
//      computed faces, rhs = nu diff(u) - conv(u)       (big_periodic_3D.f90:350-396)
//   S2 ghost fill of u* from the planes, x then y then z   (NS_functions.f90:213-310);
//      AB2-CN instead, per axis x, y, z: line ends of u* := planes
//      (NS_functions.f90:313-410), then u* := A^{-1} u* along that axis
//      (big_periodic_3D.f90:425-483); the adjoint solves with A^T
//   S3 r = div(u*)/dt + boundary terms; dp = P r + mean_dp  (…interBC:771, NS_functions.f90:1345-1400)
//   S4 ghost fill of dp, x then y then z                  (NS_functions.f90:1403-1464)
//   S5 u = u* - dt grad dp (all faces), p = dp_flag p + dp (big_periodic_3D.f90:564-567)
// The adjoint runs S5..S1 transposed.  P is symmetric (an orthogonal
// transform, a diagonal, its inverse) apart from the zero mode, which is set
// to mean_dp: the adjoint solve is the same solve with the zero mode dropped
// and d mean_dp = sum of the adjoint dp.  conv(u) is written through its cell
// and edge fluxes (u^2 at x-cells, u_avg*v_avg at x-y edges, ...): each flux
// adjoint is formed from the output adjoints around it and spread back to the
// faces it averages, so every adjoint value is gathered (no atomics, the
// result does not depend on thread order).
//
// The adjoint arrays are dense (the reference layout: row pitch = row length),
// and these kernels index them so; every launcher takes a dense Grid.
#include <cassert>

#include "ns_internal.h"

namespace ns {

namespace {

constexpr int AT = 256;

inline unsigned int ablocks(size_t n) { return (unsigned int)((n + AT - 1) / AT); }

// 1-based read of a ghosted array, 0 outside it
__device__ __forceinline__ double rd(const double *a, const int n[3], int i, int j, int k) {
    if (i < 1 || i > n[0] || j < 1 || j > n[1] || k < 1 || k > n[2]) return 0.0;
    return a[GIX(i, j, k, n[0], n[1])];
}

// the same without the range test, for points whose every access is in range
template <bool SAFE>
__device__ __forceinline__ double rdt(const double *a, const int n[3], int i, int j, int k) {
    if (SAFE) return rd(a, n, i, j, k);
    return a[(size_t)(i - 1) + (size_t)n[0] * ((size_t)(j - 1) + (size_t)n[1] * (size_t)(k - 1))];
}

__device__ __forceinline__ void dims_of(const Grid &g, int c, int n[3]) {
    n[0] = g.nx + 2; n[1] = g.ny + 2; n[2] = g.nz + 2;
    if (c < 3) n[c] -= 1;
}

// 1-based (i, j, k) of flat index q (arrays hold fewer than 2^32 elements,
// checked at context creation: 32-bit division through a double reciprocal)
__device__ __forceinline__ void unflat(size_t q, const int n[3], int &i, int &j, int &k) {
    const unsigned qq = (unsigned)q, n0 = (unsigned)n[0], n1 = (unsigned)n[1];
    const unsigned r = udiv_rd(qq, n0, 1.0 / n[0]);
    const unsigned kk = udiv_rd(r, n1, 1.0 / n[1]);
    i = (int)(qq - r * n0) + 1;
    j = (int)(r - kk * n1) + 1;
    k = (int)kk + 1;
}

// computed faces of component c (the reference's x_range0/x_range1 etc.)
__device__ __forceinline__ bool computed(const Grid &g, const int per[3], int c, int i, int j, int k) {
    const int id[3] = {i, j, k}, n[3] = {g.nx, g.ny, g.nz};
    for (int a = 0; a < 3; a++) {
        const int lo = (a == c) ? (per[a] ? 1 : 2) : 2;
        const int hi = (a == c) ? n[a] : n[a] + 1;
        if (id[a] < lo || id[a] > hi) return false;
    }
    return true;
}

// ---- S5: projection --------------------------------------------------------
struct AdjProj {
    Grid g;
    double dt, rd3[3];  // dt and 1/dx, 1/dy, 1/dz
    int dp_flag;
    const double *ub[3];
    double *usb[3];
    double *pb;   // in: adjoint p; out: dp_flag * it
    double *dpb;  // out
};

// u* adjoint = u adjoint (all faces)
__global__ void __launch_bounds__(AT) k_adj_proj_vel(AdjProj P, int c) {
    const size_t n = (c == 0) ? P.g.nu() : (c == 1) ? P.g.nv() : P.g.nw();
    const size_t q = (size_t)blockIdx.x * AT + threadIdx.x;
    if (q < n) P.usb[c][q] = P.ub[c][q];
}

// dp adjoint = p adjoint + grad^T(u adjoint) with u(m) = u*(m) - dt (dp(m+1) - dp(m))/dx
__global__ void __launch_bounds__(AT) k_adj_proj_dp(AdjProj P) {
    const Grid g = P.g;
    const int np_[3] = {g.nx + 2, g.ny + 2, g.nz + 2};
    const size_t q = (size_t)blockIdx.x * AT + threadIdx.x;
    if (q >= g.np()) return;
    int i, j, k;
    unflat(q, np_, i, j, k);
    double s = 0.0;
    for (int c = 0; c < 3; c++) {
        int n[3];
        dims_of(g, c, n);
        int m1[3] = {i, j, k};
        m1[c] -= 1;
        const double hi = rd(P.ub[c], n, i, j, k), lo = rd(P.ub[c], n, m1[0], m1[1], m1[2]);
        s += P.dt * P.rd3[c] * (hi - lo);
    }
    const double pb = P.pb[q];
    P.dpb[q] = pb + s;
    P.pb[q] = P.dp_flag ? pb : 0.0;
}

// ---- S4 / S2: ghost fills, transposed (one axis, one component) ------------
struct AdjFill {
    Grid g;
    int ax, bc, comp;   // comp 0..2 velocity, 3 pressure
    double d;           // spacing (pressure code 3)
    PlaneLayout L;
    double *a;          // adjoint array (in place)
    double *av[3];      // the three velocity adjoints (velocity fills)
    double *pbar;       // plane-set adjoint (accumulated)
};

// forward rules (velocity, NS_functions.f90:235-308; pressure :1421-1462):
//   bc 1 normal: a(N) = a(1);  other: a(1) = a(N-1), a(N) = a(2)
//   bc 2 velocity normal: a(1) = b1, a(N) = bn;  tangential: a(1) = 2 b1 - a(2), a(N) = 2 bn - a(N-1)
//   bc 2 pressure: a(1) = 2 b1 - a(2), a(N) = 2 bn - a(N-1)
//   bc 3 pressure: a(1) = -d b1 + a(2), a(N) = d bn + a(N-1)
//   bc 4: a(1) = b1, a(N) = bn
// transposed in reverse order of the assignments (a(N) first)
__global__ void __launch_bounds__(AT) k_adj_fill(const AdjFill P) {
    const int comp = (P.comp < 3) ? (int)blockIdx.y : 3;  // velocity: one component per blockIdx.y
    double *const a = (comp < 3) ? P.av[comp] : P.a;
    int n[3];
    dims_of(P.g, comp, n);
    const int ax = P.ax;
    const int o1 = (ax == 0) ? 1 : 0, o2 = (ax == 2) ? 1 : 2;
    const size_t t = (size_t)blockIdx.x * AT + threadIdx.x;
    if (t >= (size_t)n[o1] * n[o2]) return;
    int id[3];
    id[o1] = (int)(t % n[o1]) + 1;
    id[o2] = (int)(t / n[o1]) + 1;
    const int N = n[ax];
    auto at = [&](int pos) -> double & {
        int q[3] = {id[0], id[1], id[2]};
        q[ax] = pos;
        return a[GIX(q[0], q[1], q[2], n[0], n[1])];
    };
    const bool pres = comp == 3;
    const long o1p = pres ? P.L.off[18 + 2 * ax] : P.L.off[6 * comp + 2 * ax];
    const long onp = o1p + (pres ? (P.L.off[19 + 2 * ax] - P.L.off[18 + 2 * ax])
                                 : (P.L.off[6 * comp + 2 * ax + 1] - P.L.off[6 * comp + 2 * ax]));
    double *b1 = P.pbar ? P.pbar + o1p : nullptr, *bn = P.pbar ? P.pbar + onp : nullptr;
    const size_t pl = t;
    const bool normal = (comp == ax);
    if (P.bc == 1) {
        if (!pres && normal) {
            at(1) += at(N);
            at(N) = 0.0;
        } else {
            at(2) += at(N);
            at(N) = 0.0;
            at(N - 1) += at(1);
            at(1) = 0.0;
        }
    } else if (P.bc == 2 && !pres && normal) {
        bn[pl] += at(N);
        at(N) = 0.0;
        b1[pl] += at(1);
        at(1) = 0.0;
    } else if (P.bc == 2) {
        bn[pl] += 2.0 * at(N);
        at(N - 1) -= at(N);
        at(N) = 0.0;
        b1[pl] += 2.0 * at(1);
        at(2) -= at(1);
        at(1) = 0.0;
    } else if (P.bc == 3) {
        bn[pl] += P.d * at(N);
        at(N - 1) += at(N);
        at(N) = 0.0;
        b1[pl] += -P.d * at(1);
        at(2) += at(1);
        at(1) = 0.0;
    } else if (P.bc == 4) {
        bn[pl] += at(N);
        at(N) = 0.0;
        b1[pl] += at(1);
        at(1) = 0.0;
    }
}

// ---- S3: interior gather, boundary RHS terms and divergence, transposed -----
__global__ void __launch_bounds__(AT) k_adj_take_interior(Grid g, const double *dpb, double *r) {
    const size_t q = (size_t)blockIdx.x * AT + threadIdx.x;
    if (q >= g.ni()) return;
    const int a = (int)(q % g.nx) + 1;
    const size_t s = q / g.nx;
    const int b = (int)(s % g.ny) + 1, c = (int)(s / g.ny) + 1;
    r[q] = dpb[GIX(a + 1, b + 1, c + 1, g.nx + 2, g.ny + 2)];
}

// pr_bc_staggered_modify_rhs transposed (NS_functions.f90:1362-1398): one axis
struct AdjRhs {
    Grid g;
    int pbc[3];
    double d[3];
    PlaneLayout L;
    const double *rb;  // adjoint RHS (interior)
    double *pbar;
};
// blockIdx.y = axis (the three axes write disjoint planes)
__global__ void __launch_bounds__(AT) k_adj_modify_rhs(AdjRhs P) {
    const int ax = blockIdx.y;
    const int pbc = P.pbc[ax];
    if (pbc == 1) return;
    const int n[3] = {P.g.nx, P.g.ny, P.g.nz}, np_[3] = {P.g.nx + 2, P.g.ny + 2, P.g.nz + 2};
    const int o1 = (ax == 0) ? 1 : 0, o2 = (ax == 2) ? 1 : 2;
    const size_t t = (size_t)blockIdx.x * AT + threadIdx.x;
    if (t >= (size_t)n[o1] * n[o2]) return;
    const int pp = (int)(t % n[o1]) + 1, qq = (int)(t / n[o1]) + 1;
    int i1[3], iN[3];
    i1[o1] = iN[o1] = pp;
    i1[o2] = iN[o2] = qq;
    i1[ax] = 1;
    iN[ax] = n[ax];
    const size_t pl = (size_t)pp + (size_t)np_[o1] * (size_t)qq;
    const double r1 = P.rb[GIX(i1[0], i1[1], i1[2], P.g.nx, P.g.ny)], rn = P.rb[GIX(iN[0], iN[1], iN[2], P.g.nx, P.g.ny)];
    double *b1 = P.pbar + P.L.off[18 + 2 * ax], *bn = P.pbar + P.L.off[19 + 2 * ax];
    const double d = P.d[ax], d2 = d * d;
    if (pbc == 2) { b1[pl] += -2.0 * r1 / d2; bn[pl] += -2.0 * rn / d2; }
    else if (pbc == 3) { b1[pl] += r1 / d; bn[pl] += -rn / d; }
    else if (pbc == 4) { b1[pl] += -r1 / d2; bn[pl] += -rn / d2; }
}

// u* adjoint += div^T(r adjoint): r(a,b,c) = sum_c' (s(face+1) - s(face))/h / dt
// blockIdx.y = component c (uniform: the parameter selections below are scalar)
__global__ void __launch_bounds__(AT) k_adj_div(Grid g, double3 rh_dt3, const double *rb, double *usb0, double *usb1,
                                              double *usb2) {
    const int c = blockIdx.y;
    const double rh_dt = (c == 0) ? rh_dt3.x : (c == 1) ? rh_dt3.y : rh_dt3.z;
    double *usb = (c == 0) ? usb0 : (c == 1) ? usb1 : usb2;
    int n[3];
    dims_of(g, c, n);
    const size_t tot = (size_t)n[0] * n[1] * n[2];
    const size_t q = (size_t)blockIdx.x * AT + threadIdx.x;
    if (q >= tot) return;
    int i, j, k;
    unflat(q, n, i, j, k);
    // interior cell (0-based) left of the face along c is (face index - 2), right is (face - 1); others: index - 2
    int cl[3] = {i - 2, j - 2, k - 2}, cr[3] = {i - 2, j - 2, k - 2};
    cr[c] += 1;
    const int m[3] = {g.nx, g.ny, g.nz};
    auto R = [&](const int *a) -> double {
        for (int x = 0; x < 3; x++)
            if (a[x] < 0 || a[x] >= m[x]) return 0.0;
        return rb[(size_t)a[0] + (size_t)g.nx * ((size_t)a[1] + (size_t)g.ny * a[2])];
    };
    usb[q] += (R(cl) - R(cr)) * rh_dt;
}

// ---- S1: predictor ---------------------------------------------------------
struct AdjPred {
    Grid g;
    int per[3];
    double dt, nu, dpf, rd3[3], rd2[3];  // 1/h and 1/h^2
    // u* = u + dt*(a_df*nu*diff - a_cv*conv - c_hist*H - dp_flag*grad p),
    // H' = h_df*nu*diff - h_cv*conv on the computed faces
    double a_df, a_cv, c_hist, h_df, h_cv;
    int split;                           // gcb is an array of its own (a_df != a_cv or h_df != h_cv)
    const double *u[3];                  // forward state at the step start
    const double *usb[3];                // adjoint of u* (after the fill transpose)
    double *hb[3];                       // in: adjoint of H'; out: adjoint of H
    double *gb[3];                       // out: adjoint of nu*diff on the computed faces (0 elsewhere)
    double *gcb[3];                      // out: adjoint of -conv (gb itself unless split)
    double *ub[3];                       // out: adjoint of u
    double *pb;                          // adjoint of p (accumulated)
};

// blockIdx.y = component c
__global__ void __launch_bounds__(AT) k_adj_pred_prep(AdjPred P) {
    const int c = blockIdx.y;
    int n[3];
    dims_of(P.g, c, n);
    const size_t tot = (size_t)n[0] * n[1] * n[2];
    const size_t q = (size_t)blockIdx.x * AT + threadIdx.x;
    if (q >= tot) return;
    int i, j, k;
    unflat(q, n, i, j, k);
    const bool on = computed(P.g, P.per, c, i, j, k);
    const double us = on ? P.usb[c][q] : 0.0;
    const double hb = P.hb[c][q];
    P.gb[c][q] = on ? (P.dt * P.a_df * us + P.h_df * hb) : 0.0;
    if (P.split) P.gcb[c][q] = on ? (P.dt * P.a_cv * us + P.h_cv * hb) : 0.0;
    P.hb[c][q] = on ? -P.dt * P.c_hist * us : hb;
    P.ub[c][q] = us;  // u* = u + ...
}

// p adjoint += (grad p)^T: u*(q) -= dt dp_flag (p(q+e) - p(q))/h on computed faces
__global__ void __launch_bounds__(AT) k_adj_pred_p(AdjPred P) {
    const Grid g = P.g;
    const int np_[3] = {g.nx + 2, g.ny + 2, g.nz + 2};
    const size_t q = (size_t)blockIdx.x * AT + threadIdx.x;
    if (q >= g.np()) return;
    int i, j, k;
    unflat(q, np_, i, j, k);
    double s = 0.0;
    for (int c = 0; c < 3; c++) {
        int n[3];
        dims_of(g, c, n);
        int m1[3] = {i, j, k};
        m1[c] -= 1;
        // face m1 reads p(m1+e) = p(i,j,k) with weight -1; face (i,j,k) reads p(i,j,k) with +1
        double a = 0.0, b = 0.0;
        if (m1[c] >= 1 && computed(g, P.per, c, m1[0], m1[1], m1[2])) a = rd(P.usb[c], n, m1[0], m1[1], m1[2]);
        if (i <= n[0] && j <= n[1] && k <= n[2] && computed(g, P.per, c, i, j, k)) b = rd(P.usb[c], n, i, j, k);
        s += P.rd3[c] * (b - a);
    }
    P.pb[q] += P.dt * P.dpf * s;
}

// u adjoint += nu diff^T(g) - conv^T(g; u): gathered at each face of component c
// (range-checked reads; the ghost faces use this form, the others convdiff_inner).
template <bool SAFE>
__device__ __forceinline__ double convdiff_at(const AdjPred &P, int c, int i, int j, int k, const int nu_[3],
                                              const int nv_[3], const int nw_[3]) {
    const Grid &g = P.g;
    const int *nc = (c == 0) ? nu_ : (c == 1) ? nv_ : nw_;
    const double *U = P.u[0], *V = P.u[1], *W = P.u[2];
    const double *GU = P.gcb[0], *GV = P.gcb[1], *GW = P.gcb[2];  // convective part
    const double rx = P.rd3[0], ry = P.rd3[1], rz = P.rd3[2];
    const int nn[3] = {g.nx, g.ny, g.nz};
    const double *Gd = P.gb[c];   // diffusive part
    const double *Gc = P.gcb[c];  // convective part
    const double *Xc = P.u[c];
    const int id[3] = {i, j, k};
    // diffusion, diff2 along each axis (own axis wraps output 1's left neighbour to n when periodic)
    double dif = 0.0;
    const double gc0 = rdt<SAFE>(Gd, nc, i, j, k);
#pragma unroll
    for (int a = 0; a < 3; a++) {
        int m[3] = {i, j, k}, p1[3] = {i, j, k};
        m[a] -= 1;
        p1[a] += 1;
        double s = rdt<SAFE>(Gd, nc, m[0], m[1], m[2]) - 2.0 * gc0 + rdt<SAFE>(Gd, nc, p1[0], p1[1], p1[2]);
        if (a == c && P.per[c] && id[c] == nn[c]) {
            int o[3] = {i, j, k};
            o[c] = 1;
            s += rdt<SAFE>(Gd, nc, o[0], o[1], o[2]);
        }
        dif += s * P.rd2[a];
    }
    // convection: own-axis squares at the cells m = id-1 and m = id along c
    double conv = 0.0;
    {
        const double rh = P.rd3[c];
        auto cellbar = [&](int m) -> double {  // 2 a_m times the adjoint of a_m^2 (cell between faces m, m+1)
            if (m < 1 || m > nn[c]) return 0.0;
            int f0[3] = {i, j, k}, f1[3] = {i, j, k};
            f0[c] = m;
            f1[c] = m + 1;
            double X = (rdt<SAFE>(Gc, nc, f0[0], f0[1], f0[2]) - rdt<SAFE>(Gc, nc, f1[0], f1[1], f1[2])) * rh;
            if (P.per[c] && m == nn[c]) {  // a_n is also output 1's left average
                int o[3] = {i, j, k};
                o[c] = 1;
                X -= rdt<SAFE>(Gc, nc, o[0], o[1], o[2]) * rh;
            }
            const double am = 0.5 * (rdt<SAFE>(Xc, nc, f1[0], f1[1], f1[2]) + rdt<SAFE>(Xc, nc, f0[0], f0[1], f0[2]));
            return 2.0 * am * X;
        };
        conv += 0.5 * (cellbar(id[c]) + cellbar(id[c] - 1));
    }
    // edge fluxes: EF (u,v) in x-y, FF (u,w) in x-z, GF (v,w) in y-z
    // EF(i,j,k) = A*B, A = (u(i,j+1)+u(i,j))/2, B = (v(i+1,j)+v(i,j))/2
    auto EFbar = [&](int a, int b, int kk) -> double {
        return (rdt<SAFE>(GU, nu_, a, b, kk) - rdt<SAFE>(GU, nu_, a, b + 1, kk)) * ry +
               (rdt<SAFE>(GV, nv_, a, b, kk) - rdt<SAFE>(GV, nv_, a + 1, b, kk)) * rx;
    };
    auto FFbar = [&](int a, int b, int kk) -> double {
        return (rdt<SAFE>(GU, nu_, a, b, kk) - rdt<SAFE>(GU, nu_, a, b, kk + 1)) * rz +
               (rdt<SAFE>(GW, nw_, a, b, kk) - rdt<SAFE>(GW, nw_, a + 1, b, kk)) * rx;
    };
    auto GFbar = [&](int a, int b, int kk) -> double {
        return (rdt<SAFE>(GV, nv_, a, b, kk) - rdt<SAFE>(GV, nv_, a, b, kk + 1)) * rz +
               (rdt<SAFE>(GW, nw_, a, b, kk) - rdt<SAFE>(GW, nw_, a, b + 1, kk)) * ry;
    };
    auto avg = [&](const double *A, const int *n, int a0, int b0, int c0, int a1, int b1, int c1) {
        return 0.5 * (rdt<SAFE>(A, n, a1, b1, c1) + rdt<SAFE>(A, n, a0, b0, c0));
    };
    if (c == 0) {
        // u(i,j) in A(i,j) and A(i,j-1) of EF; in A(i,j,k), A(i,j,k-1) of FF
        conv += 0.5 * (EFbar(i, j, k) * avg(V, nv_, i, j, k, i + 1, j, k) + EFbar(i, j - 1, k) * avg(V, nv_, i, j - 1, k, i + 1, j - 1, k));
        conv += 0.5 * (FFbar(i, j, k) * avg(W, nw_, i, j, k, i + 1, j, k) + FFbar(i, j, k - 1) * avg(W, nw_, i, j, k - 1, i + 1, j, k - 1));
    } else if (c == 1) {
        // v(i,j) in B(i,j) and B(i-1,j) of EF; in A(i,j,k), A(i,j,k-1) of GF
        conv += 0.5 * (EFbar(i, j, k) * avg(U, nu_, i, j, k, i, j + 1, k) + EFbar(i - 1, j, k) * avg(U, nu_, i - 1, j, k, i - 1, j + 1, k));
        conv += 0.5 * (GFbar(i, j, k) * avg(W, nw_, i, j, k, i, j + 1, k) + GFbar(i, j, k - 1) * avg(W, nw_, i, j, k - 1, i, j + 1, k - 1));
    } else {
        // w(i,j,k) in B(i,j,k), B(i-1,j,k) of FF; in B(i,j,k), B(i,j-1,k) of GF
        conv += 0.5 * (FFbar(i, j, k) * avg(U, nu_, i, j, k, i, j, k + 1) + FFbar(i - 1, j, k) * avg(U, nu_, i - 1, j, k, i - 1, j, k + 1));
        conv += 0.5 * (GFbar(i, j, k) * avg(V, nv_, i, j, k, i, j, k + 1) + GFbar(i, j - 1, k) * avg(V, nv_, i, j - 1, k, i, j - 1, k + 1));
    }
    return P.nu * dif - conv;
}

// The same sum along z columns: a block takes a 64 x 4 tile of (i, j)
// columns and each thread walks z planes [k0, k1] of its column.  The values a
// face reads at k - 1 and k + 1 of its own or a neighbouring column (the z
// stencil, the z edge fluxes) stay in registers from plane to plane, and the
// in-plane neighbours of consecutive planes are read by the same block close
// together in time, so each array streams about once.  Non-ghost faces use the
// sums below with one linear index per array (every access inside the arrays
// there; the expression order is that of convdiff_at); ghost faces use
// convdiff_at.
constexpr int CD_TX = 64, CD_TY = 4;

// the ghost faces of component c (i, j or k at either end of its array):
// the two z end planes, then the two y end rows of the other planes, then the
// two x end columns of the rest
__global__ void __launch_bounds__(AT) k_adj_convdiff_surface(AdjPred P) {
    const int c = blockIdx.y;
    int nu_[3], nv_[3], nw_[3];
    dims_of(P.g, 0, nu_);
    dims_of(P.g, 1, nv_);
    dims_of(P.g, 2, nw_);
    const int *nc = (c == 0) ? nu_ : (c == 1) ? nv_ : nw_;
    const long n0 = nc[0], n1 = nc[1], n2 = nc[2];
    const long sz = 2 * n0 * n1, sy = 2 * n0 * (n2 - 2), sx = 2 * (n1 - 2) * (n2 - 2);
    long t = (long)blockIdx.x * AT + threadIdx.x;
    int i, j, k;
    if (t < sz) {
        k = (t < n0 * n1) ? 1 : (int)n2;
        t %= n0 * n1;
        i = (int)(t % n0) + 1;
        j = (int)(t / n0) + 1;
    } else if ((t -= sz) < sy) {
        j = (t < n0 * (n2 - 2)) ? 1 : (int)n1;
        t %= n0 * (n2 - 2);
        i = (int)(t % n0) + 1;
        k = (int)(t / n0) + 2;
    } else if ((t -= sy) < sx) {
        i = (t < (n1 - 2) * (n2 - 2)) ? 1 : (int)n0;
        t %= (n1 - 2) * (n2 - 2);
        j = (int)(t % (n1 - 2)) + 2;
        k = (int)(t / (n1 - 2)) + 2;
    } else {
        return;
    }
    P.ub[c][(size_t)(i - 1) + (size_t)n0 * ((size_t)(j - 1) + (size_t)n1 * (size_t)(k - 1))] +=
        convdiff_at<true>(P, c, i, j, k, nu_, nv_, nw_);
}

template <int C>
__global__ void __launch_bounds__(CD_TX * CD_TY) k_adj_convdiff_z(AdjPred P, int ntx, int nty, int nzc, int cz) {
    const Grid g = P.g;
    int nu_[3], nv_[3], nw_[3];
    dims_of(g, 0, nu_);
    dims_of(g, 1, nv_);
    dims_of(g, 2, nw_);
    const int *nc = (C == 0) ? nu_ : (C == 1) ? nv_ : nw_;
    const int nb = ntx * nty * nzc;
    const int b = xcd_slab_block(blockIdx.x, gridDim.x);
    if (b >= nb) return;
    // tiles and chunks cover the non-ghost faces i, j, k in [2, n - 1] only: the
    // ghost faces (the box's surface) are k_adj_convdiff_surface's
    const int tx = b % ntx, ty = (b / ntx) % nty, kz = b / (ntx * nty);
    const int i = tx * CD_TX + (int)(threadIdx.x & 63) + 2, j = ty * CD_TY + (int)(threadIdx.x >> 6) + 2;
    const int k0 = kz * cz + 2, k1 = (k0 + cz - 1 < nc[2] - 1) ? k0 + cz - 1 : nc[2] - 1;
    if (i > nc[0] - 1 || j > nc[1] - 1) return;
    const long sc1 = nc[0], sc2 = (long)nc[0] * nc[1];
    double *out = P.ub[C] + (i - 1) + sc1 * (j - 1);
    const double *U = P.u[0], *V = P.u[1], *W = P.u[2];
    const double *GU = P.gcb[0], *GV = P.gcb[1], *GW = P.gcb[2];  // convective part
    const double *GD = P.gb[C];                                    // diffusive part, component C
    const double rx = P.rd3[0], ry = P.rd3[1], rz = P.rd3[2];
    const long su1 = nu_[0], su2 = (long)nu_[0] * nu_[1];
    const long sv1 = nv_[0], sv2 = (long)nv_[0] * nv_[1];
    const long sw1 = nw_[0], sw2 = (long)nw_[0] * nw_[1];
    // plane-1 offsets of the column (i, j) in each array
    const long cu = (i - 1) + su1 * (j - 1), cv = (i - 1) + sv1 * (j - 1), cw = (i - 1) + sw1 * (j - 1);
    // value at plane kk of a column, 0 outside the array's planes
    auto zk = [](const double *A, long col, long s2, int n2, int kk) {
        return (kk >= 1 && kk <= n2) ? A[col + s2 * (kk - 1)] : 0.0;
    };
    const int nzC = nc[2];
    if constexpr (C == 0) {
        // windows: GD, GU at (i, j) planes k-1..k+1; GW, W at (i, j), (i+1, j) planes k-1, k
        double gdm = zk(GD, cu, su2, nzC, k0 - 1), gdc = zk(GD, cu, su2, nzC, k0), gdp = zk(GD, cu, su2, nzC, k0 + 1);
        double gum = zk(GU, cu, su2, nzC, k0 - 1), guc = zk(GU, cu, su2, nzC, k0), gup = zk(GU, cu, su2, nzC, k0 + 1);
        double gw0m = zk(GW, cw, sw2, nw_[2], k0 - 1), gw0c = zk(GW, cw, sw2, nw_[2], k0);
        double gw1m = zk(GW, cw + 1, sw2, nw_[2], k0 - 1), gw1c = zk(GW, cw + 1, sw2, nw_[2], k0);
        double w0m = zk(W, cw, sw2, nw_[2], k0 - 1), w0c = zk(W, cw, sw2, nw_[2], k0);
        double w1m = zk(W, cw + 1, sw2, nw_[2], k0 - 1), w1c = zk(W, cw + 1, sw2, nw_[2], k0);
        for (int k = k0; k <= k1; k++) {
            const long qu = cu + su2 * (k - 1), qv = cv + sv2 * (k - 1);
            {
                double dif = (GD[qu - 1] - 2.0 * gdc + GD[qu + 1]) * P.rd2[0] +
                             (GD[qu - su1] - 2.0 * gdc + GD[qu + su1]) * P.rd2[1] + (gdm - 2.0 * gdc + gdp) * P.rd2[2];
                double X1 = (guc - GU[qu + 1]) * rx;
                if (P.per[0] && i == g.nx) {
                    dif += GD[qu - (i - 1)] * P.rd2[0];
                    X1 -= GU[qu - (i - 1)] * rx;
                }
                const double X0 = (GU[qu - 1] - guc) * rx;
                double conv = 0.5 * (2.0 * (0.5 * (U[qu + 1] + U[qu])) * X1 + 2.0 * (0.5 * (U[qu] + U[qu - 1])) * X0);
                const double ef1 = (guc - GU[qu + su1]) * ry + (GV[qv] - GV[qv + 1]) * rx;
                const double ef0 = (GU[qu - su1] - guc) * ry + (GV[qv - sv1] - GV[qv - sv1 + 1]) * rx;
                conv += 0.5 * (ef1 * (0.5 * (V[qv + 1] + V[qv])) + ef0 * (0.5 * (V[qv - sv1 + 1] + V[qv - sv1])));
                const double ff1 = (guc - gup) * rz + (gw0c - gw1c) * rx;
                const double ff0 = (gum - guc) * rz + (gw0m - gw1m) * rx;
                conv += 0.5 * (ff1 * (0.5 * (w1c + w0c)) + ff0 * (0.5 * (w1m + w0m)));
                out[sc2 * (k - 1)] += P.nu * dif - conv;
            }
            gdm = gdc; gdc = gdp; gdp = zk(GD, cu, su2, nzC, k + 2);
            gum = guc; guc = gup; gup = zk(GU, cu, su2, nzC, k + 2);
            gw0m = gw0c; gw0c = zk(GW, cw, sw2, nw_[2], k + 1);
            gw1m = gw1c; gw1c = zk(GW, cw + 1, sw2, nw_[2], k + 1);
            w0m = w0c; w0c = zk(W, cw, sw2, nw_[2], k + 1);
            w1m = w1c; w1c = zk(W, cw + 1, sw2, nw_[2], k + 1);
        }
    } else if constexpr (C == 1) {
        // windows: GD, GV at (i, j) planes k-1..k+1; GW, W at (i, j), (i, j+1) planes k-1, k
        double gdm = zk(GD, cv, sv2, nzC, k0 - 1), gdc = zk(GD, cv, sv2, nzC, k0), gdp = zk(GD, cv, sv2, nzC, k0 + 1);
        double gvm = zk(GV, cv, sv2, nzC, k0 - 1), gvc = zk(GV, cv, sv2, nzC, k0), gvp = zk(GV, cv, sv2, nzC, k0 + 1);
        double gw0m = zk(GW, cw, sw2, nw_[2], k0 - 1), gw0c = zk(GW, cw, sw2, nw_[2], k0);
        double gw1m = zk(GW, cw + sw1, sw2, nw_[2], k0 - 1), gw1c = zk(GW, cw + sw1, sw2, nw_[2], k0);
        double w0m = zk(W, cw, sw2, nw_[2], k0 - 1), w0c = zk(W, cw, sw2, nw_[2], k0);
        double w1m = zk(W, cw + sw1, sw2, nw_[2], k0 - 1), w1c = zk(W, cw + sw1, sw2, nw_[2], k0);
        for (int k = k0; k <= k1; k++) {
            const long qu = cu + su2 * (k - 1), qv = cv + sv2 * (k - 1);
            {
                double dif = (GD[qv - 1] - 2.0 * gdc + GD[qv + 1]) * P.rd2[0] +
                             (GD[qv - sv1] - 2.0 * gdc + GD[qv + sv1]) * P.rd2[1] + (gdm - 2.0 * gdc + gdp) * P.rd2[2];
                double X1 = (gvc - GV[qv + sv1]) * ry;
                if (P.per[1] && j == g.ny) {
                    dif += GD[qv - (long)(j - 1) * sv1] * P.rd2[1];
                    X1 -= GV[qv - (long)(j - 1) * sv1] * ry;
                }
                const double X0 = (GV[qv - sv1] - gvc) * ry;
                double conv =
                    0.5 * (2.0 * (0.5 * (V[qv + sv1] + V[qv])) * X1 + 2.0 * (0.5 * (V[qv] + V[qv - sv1])) * X0);
                const double ef1 = (GU[qu] - GU[qu + su1]) * ry + (gvc - GV[qv + 1]) * rx;
                const double ef0 = (GU[qu - 1] - GU[qu - 1 + su1]) * ry + (GV[qv - 1] - gvc) * rx;
                conv += 0.5 * (ef1 * (0.5 * (U[qu + su1] + U[qu])) + ef0 * (0.5 * (U[qu - 1 + su1] + U[qu - 1])));
                const double gf1 = (gvc - gvp) * rz + (gw0c - gw1c) * ry;
                const double gf0 = (gvm - gvc) * rz + (gw0m - gw1m) * ry;
                conv += 0.5 * (gf1 * (0.5 * (w1c + w0c)) + gf0 * (0.5 * (w1m + w0m)));
                out[sc2 * (k - 1)] += P.nu * dif - conv;
            }
            gdm = gdc; gdc = gdp; gdp = zk(GD, cv, sv2, nzC, k + 2);
            gvm = gvc; gvc = gvp; gvp = zk(GV, cv, sv2, nzC, k + 2);
            gw0m = gw0c; gw0c = zk(GW, cw, sw2, nw_[2], k + 1);
            gw1m = gw1c; gw1c = zk(GW, cw + sw1, sw2, nw_[2], k + 1);
            w0m = w0c; w0c = zk(W, cw, sw2, nw_[2], k + 1);
            w1m = w1c; w1c = zk(W, cw + sw1, sw2, nw_[2], k + 1);
        }
    } else {
        // windows: GD, GW, W at (i, j) planes k-1..k+1; GU, U at (i, j), (i-1, j)
        // and GV, V at (i, j), (i, j-1), planes k, k+1
        double gdm = zk(GD, cw, sw2, nzC, k0 - 1), gdc = zk(GD, cw, sw2, nzC, k0), gdp = zk(GD, cw, sw2, nzC, k0 + 1);
        double gwm = zk(GW, cw, sw2, nzC, k0 - 1), gwc = zk(GW, cw, sw2, nzC, k0), gwp = zk(GW, cw, sw2, nzC, k0 + 1);
        double wm = zk(W, cw, sw2, nzC, k0 - 1), wc = zk(W, cw, sw2, nzC, k0), wp = zk(W, cw, sw2, nzC, k0 + 1);
        double gu0c = zk(GU, cu, su2, nu_[2], k0), gu0p = zk(GU, cu, su2, nu_[2], k0 + 1);
        double gu1c = zk(GU, cu - 1, su2, nu_[2], k0), gu1p = zk(GU, cu - 1, su2, nu_[2], k0 + 1);
        double u0c = zk(U, cu, su2, nu_[2], k0), u0p = zk(U, cu, su2, nu_[2], k0 + 1);
        double u1c = zk(U, cu - 1, su2, nu_[2], k0), u1p = zk(U, cu - 1, su2, nu_[2], k0 + 1);
        double gv0c = zk(GV, cv, sv2, nv_[2], k0), gv0p = zk(GV, cv, sv2, nv_[2], k0 + 1);
        double gv1c = zk(GV, cv - sv1, sv2, nv_[2], k0), gv1p = zk(GV, cv - sv1, sv2, nv_[2], k0 + 1);
        double v0c = zk(V, cv, sv2, nv_[2], k0), v0p = zk(V, cv, sv2, nv_[2], k0 + 1);
        double v1c = zk(V, cv - sv1, sv2, nv_[2], k0), v1p = zk(V, cv - sv1, sv2, nv_[2], k0 + 1);
        for (int k = k0; k <= k1; k++) {
            const long qw = cw + sw2 * (k - 1);
            {
                double dif = (GD[qw - 1] - 2.0 * gdc + GD[qw + 1]) * P.rd2[0] +
                             (GD[qw - sw1] - 2.0 * gdc + GD[qw + sw1]) * P.rd2[1] + (gdm - 2.0 * gdc + gdp) * P.rd2[2];
                double X1 = (gwc - gwp) * rz;
                if (P.per[2] && k == g.nz) {
                    dif += GD[qw - (long)(k - 1) * sw2] * P.rd2[2];
                    X1 -= GW[qw - (long)(k - 1) * sw2] * rz;
                }
                const double X0 = (gwm - gwc) * rz;
                double conv = 0.5 * (2.0 * (0.5 * (wp + wc)) * X1 + 2.0 * (0.5 * (wc + wm)) * X0);
                const double ff1 = (gu0c - gu0p) * rz + (gwc - GW[qw + 1]) * rx;
                const double ff0 = (gu1c - gu1p) * rz + (GW[qw - 1] - gwc) * rx;
                conv += 0.5 * (ff1 * (0.5 * (u0p + u0c)) + ff0 * (0.5 * (u1p + u1c)));
                const double gf1 = (gv0c - gv0p) * rz + (gwc - GW[qw + sw1]) * ry;
                const double gf0 = (gv1c - gv1p) * rz + (GW[qw - sw1] - gwc) * ry;
                conv += 0.5 * (gf1 * (0.5 * (v0p + v0c)) + gf0 * (0.5 * (v1p + v1c)));
                out[sc2 * (k - 1)] += P.nu * dif - conv;
            }
            gdm = gdc; gdc = gdp; gdp = zk(GD, cw, sw2, nzC, k + 2);
            gwm = gwc; gwc = gwp; gwp = zk(GW, cw, sw2, nzC, k + 2);
            wm = wc; wc = wp; wp = zk(W, cw, sw2, nzC, k + 2);
            gu0c = gu0p; gu0p = zk(GU, cu, su2, nu_[2], k + 2);
            gu1c = gu1p; gu1p = zk(GU, cu - 1, su2, nu_[2], k + 2);
            u0c = u0p; u0p = zk(U, cu, su2, nu_[2], k + 2);
            u1c = u1p; u1p = zk(U, cu - 1, su2, nu_[2], k + 2);
            gv0c = gv0p; gv0p = zk(GV, cv, sv2, nv_[2], k + 2);
            gv1c = gv1p; gv1p = zk(GV, cv - sv1, sv2, nv_[2], k + 2);
            v0c = v0p; v0p = zk(V, cv, sv2, nv_[2], k + 2);
            v1c = v1p; v1p = zk(V, cv - sv1, sv2, nv_[2], k + 2);
        }
    }
}

__global__ void __launch_bounds__(AT) k_block_sum(const double *a, size_t n, double *partial) {
    double s = 0.0;
    for (size_t q = (size_t)blockIdx.x * AT + threadIdx.x; q < n; q += (size_t)gridDim.x * AT) s += a[q];
    __shared__ double red[AT];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int h = AT / 2; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

__global__ void k_add_at(double *a, size_t q, double v) {
    if (threadIdx.x == 0 && blockIdx.x == 0) a[q] += v;
}

// a[q] += scale * (sum of the nb block partials), summed in a fixed order
__global__ void __launch_bounds__(AT) k_add_partials(double *a, size_t q, const double *partial, int nb,
                                                     double scale) {
    double s = 0.0;
    for (int i = threadIdx.x; i < nb; i += AT) s += partial[i];
    __shared__ double red[AT];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int h = AT / 2; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) a[q] += scale * red[0];
}

}  // namespace

int launch_block_sums(const double *a, size_t n, double *partial, int max_blocks, hipStream_t st) {
    size_t nb = (n + AT - 1) / AT;
    if (nb > (size_t)max_blocks) nb = max_blocks;
    if (nb < 1) nb = 1;
    hipLaunchKernelGGL(k_block_sum, dim3((unsigned)nb), dim3(AT), 0, st, a, n, partial);
    return (int)nb;
}

void launch_add_partials(double *a, size_t q, const double *partial, int nb, double scale, hipStream_t st) {
    hipLaunchKernelGGL(k_add_partials, dim3(1), dim3(AT), 0, st, a, q, partial, nb, scale);
}

void launch_add_at(double *a, size_t q, double v, hipStream_t st) {
    hipLaunchKernelGGL(k_add_at, dim3(1), dim3(64), 0, st, a, q, v);
}

void launch_adj_project(const Grid &g, double dt, const double h[3], int dp_flag, const double *const ub[3],
                        double *const usb[3], double *pb, double *dpb, hipStream_t st) {
    assert(g.dense());
    AdjProj P;
    P.g = g;
    P.dt = dt;
    for (int a = 0; a < 3; a++) {
        P.rd3[a] = 1.0 / h[a];
        P.ub[a] = ub[a];
        P.usb[a] = usb[a];
    }
    P.dp_flag = dp_flag;
    P.pb = pb;
    P.dpb = dpb;
    const size_t n[3] = {g.nu(), g.nv(), g.nw()};
    for (int c = 0; c < 3; c++)
        if (usb[c] != ub[c]) hipLaunchKernelGGL(k_adj_proj_vel, dim3(ablocks(n[c])), dim3(AT), 0, st, P, c);
    hipLaunchKernelGGL(k_adj_proj_dp, dim3(ablocks(g.np())), dim3(AT), 0, st, P);
}

void launch_adj_fill(const Grid &g, int comp, const int bc[3], const double h[3], double *const a[3], double *pbar,
                     hipStream_t st) {
    assert(g.dense());
    AdjFill P;
    P.g = g;
    P.comp = comp;
    P.L = plane_layout(g.nx, g.ny, g.nz);
    P.a = a[0];
    for (int c = 0; c < 3; c++) P.av[c] = a[c];
    P.pbar = pbar;
    const int n[3] = {g.nx + 2, g.ny + 2, g.nz + 2};  // the largest plane of any component
    for (int ax = 2; ax >= 0; ax--) {  // forward order x, y, z
        P.ax = ax;
        P.bc = bc[ax];
        P.d = h[ax];
        if (bc[ax] < 1 || bc[ax] > 4) continue;
        const int o1 = (ax == 0) ? 1 : 0, o2 = (ax == 2) ? 1 : 2;
        hipLaunchKernelGGL(k_adj_fill, dim3(ablocks((size_t)n[o1] * n[o2]), comp < 3 ? 3 : 1), dim3(AT), 0, st, P);
    }
}

void launch_adj_take_interior(const Grid &g, const double *dpb, double *r, hipStream_t st) {
    assert(g.dense());
    hipLaunchKernelGGL(k_adj_take_interior, dim3(ablocks(g.ni())), dim3(AT), 0, st, g, dpb, r);
}

void launch_adj_modify_rhs(const Grid &g, const int pbc[3], const double h[3], const double *rb, double *pbar,
                           hipStream_t st) {
    assert(g.dense());
    AdjRhs R;
    R.g = g;
    R.L = plane_layout(g.nx, g.ny, g.nz);
    R.rb = rb;
    R.pbar = pbar;
    if (!pbar || (pbc[0] == 1 && pbc[1] == 1 && pbc[2] == 1)) return;
    for (int ax = 0; ax < 3; ax++) {
        R.pbc[ax] = pbc[ax];
        R.d[ax] = h[ax];
    }
    const size_t mx = std::max((size_t)g.ny * g.nz, std::max((size_t)g.nx * g.nz, (size_t)g.nx * g.ny));
    hipLaunchKernelGGL(k_adj_modify_rhs, dim3(ablocks(mx), 3), dim3(AT), 0, st, R);
}

void launch_adj_div(const Grid &g, const double h[3], double dt, const double *rb, double *const usb[3],
                    hipStream_t st) {
    assert(g.dense());
    const size_t mx = std::max(g.nu(), std::max(g.nv(), g.nw()));
    const double3 r = make_double3(1.0 / (h[0] * dt), 1.0 / (h[1] * dt), 1.0 / (h[2] * dt));
    hipLaunchKernelGGL(k_adj_div, dim3(ablocks(mx), 3), dim3(AT), 0, st, g, r, rb, usb[0], usb[1], usb[2]);
}

void launch_adj_rhs_div(const Grid &g, const int pbc[3], const double h[3], double dt, const double *rb,
                        double *const usb[3], double *pbar, hipStream_t st) {
    assert(g.dense());
    launch_adj_modify_rhs(g, pbc, h, rb, pbar, st);
    launch_adj_div(g, h, dt, rb, usb, st);
}

namespace {

// transpose of vel_bc_staggered_CN2 along axis ax (NS_functions.f90:313-410):
// on walls the forward sets both line ends of every component to its plane
// values, so their adjoints move to the planes and the ends' adjoints become
// zero; on periodic lines it sets the ghost rows to zero, so their adjoints
// become zero (one thread per line and component)
__global__ void __launch_bounds__(AT) k_adj_cn2(Grid g, int ax, int bc, PlaneLayout L, double *a0, double *a1,
                                              double *a2, double *pbar) {
    const int c = blockIdx.y;
    int n[3];
    dims_of(g, c, n);
    const int o1 = (ax == 0) ? 1 : 0, o2 = (ax == 2) ? 1 : 2;
    const size_t t = (size_t)blockIdx.x * AT + threadIdx.x;
    if (t >= (size_t)n[o1] * n[o2]) return;
    int id[3];
    id[o1] = (int)(t % n[o1]) + 1;
    id[o2] = (int)(t / n[o1]) + 1;
    double *A = (c == 0) ? a0 : (c == 1) ? a1 : a2;
    id[ax] = 1;
    const size_t q1 = GIX(id[0], id[1], id[2], n[0], n[1]);
    id[ax] = n[ax];
    const size_t qn = GIX(id[0], id[1], id[2], n[0], n[1]);
    if (bc == 1) {
        // the forward zeroes the ghost rows (row N; row 1 too for the tangential components)
        if (c != ax) A[q1] = 0.0;
        A[qn] = 0.0;
        return;
    }
    pbar[L.off[6 * c + 2 * ax] + t] += A[q1];
    pbar[L.off[6 * c + 2 * ax + 1] + t] += A[qn];
    A[q1] = 0.0;
    A[qn] = 0.0;
}

}  // namespace

void launch_adj_cn2(const Grid &g, int ax, int bc, double *const a[3], double *pbar, hipStream_t st) {
    assert(g.dense());
    if (bc != 1 && bc != 2 && bc != 4) return;  // code 3: the forward leaves the lines as they are
    size_t mx = 0;
    for (int c = 0; c < 3; c++) {
        int n[3];
        comp_dims(g, c, n);
        const int o1 = (ax == 0) ? 1 : 0, o2 = (ax == 2) ? 1 : 2;
        const size_t m = (size_t)n[o1] * n[o2];
        mx = m > mx ? m : mx;
    }
    hipLaunchKernelGGL(k_adj_cn2, dim3(ablocks(mx), 3), dim3(AT), 0, st, g, ax, bc, plane_layout(g.nx, g.ny, g.nz),
                       a[0], a[1], a[2], pbar);
}

void launch_adj_predict(const Grid &g, const int per[3], double dt, double nu, int dp_flag, const double h[3],
                        int scheme, int first, const double *const u[3], const double *const usb[3],
                        double *const hb[3], double *const gb[3], double *const gcb[3], double *const ub[3],
                        double *pb, hipStream_t st) {
    assert(g.dense());
    AdjPred P;
    P.g = g;
    for (int a = 0; a < 3; a++) {
        P.per[a] = per[a];
        P.rd3[a] = 1.0 / h[a];
        P.rd2[a] = 1.0 / (h[a] * h[a]);
        P.u[a] = u[a];
        P.usb[a] = usb[a];
        P.hb[a] = hb[a];
        P.gb[a] = gb[a];
        P.gcb[a] = gcb[a];
        P.ub[a] = ub[a];
    }
    P.dt = dt;
    P.nu = nu;
    P.dpf = (double)dp_flag;
    P.pb = pb;
    // weights of the predictor (big_periodic_3D.f90:350-396): Euler
    // u* = u + dt*rhs, H' = 0; AB2 u* = u + dt*(1.5 rhs - 0.5 H), H' = rhs
    // (first step: Euler weights, H' = rhs); AB2-CN u* = u + dt*(nu/2 diff
    // - 1.5 conv - 0.5 H), H' = -conv (first step: - conv)
    if (scheme == NS_AB2_CN) {
        P.a_df = 0.5;
        P.a_cv = first ? 1.0 : 1.5;
        P.c_hist = first ? 0.0 : 0.5;
        P.h_df = 0.0;
        P.h_cv = 1.0;
    } else {
        const bool ab2 = (scheme == NS_AB2);
        P.a_df = P.a_cv = (ab2 && !first) ? 1.5 : 1.0;
        P.c_hist = (ab2 && !first) ? 0.5 : 0.0;
        P.h_df = P.h_cv = ab2 ? 1.0 : 0.0;
    }
    P.split = (gcb[0] != gb[0]) ? 1 : 0;
    const size_t tot[3] = {g.nu(), g.nv(), g.nw()};
    hipLaunchKernelGGL(k_adj_pred_prep, dim3(ablocks(std::max(tot[0], std::max(tot[1], tot[2]))), 3), dim3(AT), 0, st,
                       P);
    if (dp_flag) hipLaunchKernelGGL(k_adj_pred_p, dim3(ablocks(g.np())), dim3(AT), 0, st, P);
    // the box's surface faces, then the z columns of the others: chunks of cz
    // planes, about one round of 4 resident tiles per CU
    long surf = 0;
    for (int c = 0; c < 3; c++) {
        int n[3];
        comp_dims(g, c, n);
        const long m = 2L * n[0] * n[1] + 2L * n[0] * (n[2] - 2) + 2L * (n[1] - 2) * (n[2] - 2);
        surf = m > surf ? m : surf;
    }
    hipLaunchKernelGGL(k_adj_convdiff_surface, dim3(ablocks((size_t)surf), 3), dim3(AT), 0, st, P);
    for (int c = 0; c < 3; c++) {
        int n[3];
        comp_dims(g, c, n);
        const int ntx = (n[0] - 2 + CD_TX - 1) / CD_TX, nty = (n[1] - 2 + CD_TY - 1) / CD_TY;
        if (ntx < 1 || nty < 1 || n[2] < 3) continue;
        int cz = (int)(((long)(n[2] - 2) * ntx * nty + 1023) / 1024);
        cz = cz < 4 ? 4 : cz > 32 ? 32 : cz;
        const int nzc = (n[2] - 2 + cz - 1) / cz;
        const dim3 grid(round8((unsigned)(ntx * nty * nzc))), blk(CD_TX * CD_TY);
        if (c == 0) hipLaunchKernelGGL(k_adj_convdiff_z<0>, grid, blk, 0, st, P, ntx, nty, nzc, cz);
        else if (c == 1) hipLaunchKernelGGL(k_adj_convdiff_z<1>, grid, blk, 0, st, P, ntx, nty, nzc, cz);
        else hipLaunchKernelGGL(k_adj_convdiff_z<2>, grid, blk, 0, st, P, ntx, nty, nzc, cz);
    }
}

}  // namespace ns
