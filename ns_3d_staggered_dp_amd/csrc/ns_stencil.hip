// Staggered-grid stencils, boundary fills, divergence and projection for
// gfx950, FP64.  Formulas and operation order follow the reference so that
// each value rounds the same way (paths relative to NS_3D_Staggered_dp/):
//   cal_conv  NS_functions.f90:53-158   (avg/diff composition, FD_functions.f90:222-410)
//   cal_diff  NS_functions.f90:161-210  (diff2, FD_functions.f90:413-527)
//   predictor big_periodic_3D.f90:350-424, ...interBC_combine.f90:543-654
//   vel_bc_staggered(_CN2) NS_functions.f90:213-410, pr_bc_staggered :1403-1464
//   get_velpr_bc/get_pr_bc :414-562, pr_bc_staggered_modify_rhs :1345-1400
//   divergence big_periodic_3D.f90:498-501, projection :562-567
#include "ns_internal.h"

namespace ns {

namespace {

struct FieldsC {
    const double *u, *v, *w;
    int nx, ny, nz;
    int pu, pc;  // row pitches of u and of v, w
};

#define UU(i, j, k) f.u[GIX(i, j, k, f.pu, f.ny + 2)]
#define VV(i, j, k) f.v[GIX(i, j, k, f.pc, f.ny + 1)]
#define WW(i, j, k) f.w[GIX(i, j, k, f.pc, f.ny + 2)]

// edge fluxes shared by two components: E = avg_y(u)*avg_x(v) on xy-edges,
// F = avg_z(u)*avg_x(w) on xz-edges, G = avg_z(v)*avg_y(w) on yz-edges
__device__ __forceinline__ double fluxE(const FieldsC &f, int i, int j, int k) {
    return ((UU(i, j + 1, k) + UU(i, j, k)) / 2.0) * ((VV(i + 1, j, k) + VV(i, j, k)) / 2.0);
}
__device__ __forceinline__ double fluxF(const FieldsC &f, int i, int j, int k) {
    return ((UU(i, j, k + 1) + UU(i, j, k)) / 2.0) * ((WW(i + 1, j, k) + WW(i, j, k)) / 2.0);
}
__device__ __forceinline__ double fluxG(const FieldsC &f, int i, int j, int k) {
    return ((VV(i, j, k + 1) + VV(i, j, k)) / 2.0) * ((WW(i, j + 1, k) + WW(i, j, k)) / 2.0);
}

// conv and diff of one component at one in-range point (1-based ghosted indices)
__device__ __forceinline__ void conv_diff_u(const FieldsC &f, int per, int fi, int J, int K, double dx, double dy,
                                            double dz, double dx2, double dy2, double dz2, double &conv, double &dif) {
    int fl = (per && fi == 1) ? f.nx : fi - 1;
    double c0 = UU(fi, J, K);
    double ur = (UU(fi + 1, J, K) + c0) / 2.0;
    // avg(bc=1) wraps with the array's last two entries: (u(nx+1)+u(nx))/2 (FD_functions.f90:252)
    double ul = (per && fi == 1) ? (UU(f.nx + 1, J, K) + UU(f.nx, J, K)) / 2.0 : (c0 + UU(fl, J, K)) / 2.0;
    double t1 = ur * ur - ul * ul;
    double t2 = fluxE(f, fi, J, K) - fluxE(f, fi, J - 1, K);
    double t3 = fluxF(f, fi, J, K) - fluxF(f, fi, J, K - 1);
    double r = t1 / dx;
    r = r + t2 / dy;
    r = r + t3 / dz;
    conv = r;
    double d1 = UU(fl, J, K) - 2.0 * c0 + UU(fi + 1, J, K);
    double d2 = UU(fi, J - 1, K) - 2.0 * c0 + UU(fi, J + 1, K);
    double d3 = UU(fi, J, K - 1) - 2.0 * c0 + UU(fi, J, K + 1);
    double s = d1 / dx2;
    s = s + d2 / dy2;
    s = s + d3 / dz2;
    dif = s;
}
__device__ __forceinline__ void conv_diff_v(const FieldsC &f, int per, int I, int fj, int K, double dx, double dy,
                                            double dz, double dx2, double dy2, double dz2, double &conv, double &dif) {
    int fl = (per && fj == 1) ? f.ny : fj - 1;
    double c0 = VV(I, fj, K);
    double t1 = fluxE(f, I, fj, K) - fluxE(f, I - 1, fj, K);
    double vr = (VV(I, fj + 1, K) + c0) / 2.0;
    double vl = (per && fj == 1) ? (VV(I, f.ny + 1, K) + VV(I, f.ny, K)) / 2.0 : (c0 + VV(I, fl, K)) / 2.0;
    double t2 = vr * vr - vl * vl;
    double t3 = fluxG(f, I, fj, K) - fluxG(f, I, fj, K - 1);
    double r = t1 / dx;
    r = r + t2 / dy;
    r = r + t3 / dz;
    conv = r;
    double d1 = VV(I - 1, fj, K) - 2.0 * c0 + VV(I + 1, fj, K);
    double d2 = VV(I, fl, K) - 2.0 * c0 + VV(I, fj + 1, K);
    double d3 = VV(I, fj, K - 1) - 2.0 * c0 + VV(I, fj, K + 1);
    double s = d1 / dx2;
    s = s + d2 / dy2;
    s = s + d3 / dz2;
    dif = s;
}
__device__ __forceinline__ void conv_diff_w(const FieldsC &f, int per, int I, int J, int fk, double dx, double dy,
                                            double dz, double dx2, double dy2, double dz2, double &conv, double &dif) {
    int fl = (per && fk == 1) ? f.nz : fk - 1;
    double c0 = WW(I, J, fk);
    double t1 = fluxF(f, I, J, fk) - fluxF(f, I - 1, J, fk);
    double t2 = fluxG(f, I, J, fk) - fluxG(f, I, J - 1, fk);
    double wr = (WW(I, J, fk + 1) + c0) / 2.0;
    double wl = (per && fk == 1) ? (WW(I, J, f.nz + 1) + WW(I, J, f.nz)) / 2.0 : (c0 + WW(I, J, fl)) / 2.0;
    double t3 = wr * wr - wl * wl;
    double r = t1 / dx;
    r = r + t2 / dy;
    r = r + t3 / dz;
    conv = r;
    double d1 = WW(I - 1, J, fk) - 2.0 * c0 + WW(I + 1, J, fk);
    double d2 = WW(I, J - 1, fk) - 2.0 * c0 + WW(I, J + 1, fk);
    double d3 = WW(I, J, fl) - 2.0 * c0 + WW(I, J, fk + 1);
    double s = d1 / dx2;
    s = s + d2 / dy2;
    s = s + d3 / dz2;
    dif = s;
}

__device__ __forceinline__ void conv_diff(const FieldsC &f, int c, const int bc[3], int i, int j, int k, double dx,
                                          double dy, double dz, double dx2, double dy2, double dz2, double &conv,
                                          double &dif) {
    if (c == 0) conv_diff_u(f, bc[0] == 1, i, j, k, dx, dy, dz, dx2, dy2, dz2, conv, dif);
    else if (c == 1) conv_diff_v(f, bc[1] == 1, i, j, k, dx, dy, dz, dx2, dy2, dz2, conv, dif);
    else conv_diff_w(f, bc[2] == 1, i, j, k, dx, dy, dz, dx2, dy2, dz2, conv, dif);
}

struct CDParams {
    int nx, ny, nz;
    int pu, pc;
    double h0, h1, h2;
    int bc[3];
    int which;  // 0 conv, 1 diff
    const double *u, *v, *w;
    double *o[3];
};

// compact outputs conv_x(m1,ny,nz) etc. (cal_conv/cal_diff mirrors)
__global__ void __launch_bounds__(256) k_conv_diff(CDParams P) {
    const int c = blockIdx.y;
    int m[3] = {P.nx, P.ny, P.nz};
    if (P.bc[c] != 1) m[c] -= 1;
    size_t tot = (size_t)m[0] * m[1] * m[2];
    size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= tot) return;
    int a = (int)(q % m[0]) + 1;
    size_t r = q / m[0];
    int b = (int)(r % m[1]) + 1;
    int cc = (int)(r / m[1]) + 1;
    int id[3] = {a + 1, b + 1, cc + 1};
    id[c] = (P.bc[c] == 1) ? id[c] - 1 : id[c];  // own axis: periodic 1..n, walls 2..n
    FieldsC f{P.u, P.v, P.w, P.nx, P.ny, P.nz, P.pu, P.pc};
    double cv, df;
    // conv mirror gets (dx,dy,dz); diff mirror gets (dx2,dy2,dz2) as cal_diff does
    conv_diff(f, c, P.bc, id[0], id[1], id[2], P.h0, P.h1, P.h2, P.h0, P.h1, P.h2, cv, df);
    P.o[c][q] = (P.which == 0) ? cv : df;
}

struct BCParams {
    int nx, ny, nz;
    int pu, pc;    // row pitches of u and of v, w, p
    int ax;        // axis of a one-axis launch; first filled axis of a whole fill
    int bc;
    int pbc_mode;  // 0 velocity fill, 1 CN2 rows, 2 pressure fill (one axis); 3 velocity, 4 pressure (whole fill)
    double d;
    int bcs[3];    // whole fill: codes per axis
    double ds[3];  // whole pressure fill: spacings (code 3 rows)
    PlaneLayout L;
    const double *planes;
    double *a[3];
};

// ---- a whole ghost fill (x, y, z passes) in one launch ----
// vel_bc_staggered (NS_functions.f90:213-310) and pr_bc_staggered
// (:1403-1464) assign the end entries of every line along x, then y, then z,
// each from values the earlier passes left.  The final value of an entry is
// the last pass that assigns it, applied to what the earlier passes produced;
// following that chain back ends at entries no pass assigns (the substituted
// indices are source rows 2, n-1 or the periodic normal face 1, which no pass
// assigns).  So one thread per assigned entry computes its final value from
// unassigned entries only (no thread reads an entry another writes), with the
// operations and their order of the three passes: bitwise their result.

// is entry m of an axis line (length N) assigned by that axis' pass?
__device__ __forceinline__ bool fill_assigned(bool pres, bool normal, int bc, int m, int N) {
    if (m != 1 && m != N) return false;
    if (pres) return bc >= 1 && bc <= 4;
    if (bc == 1) return normal ? m == N : true;
    return bc == 2 || bc == 4;  // velocity code 3: the pass leaves the line
}

// the pass' value for the low (m = 1) or high (m = N) end; src(m'): the line's
// value at m' before the pass, b: the line's plane entry (codes 2-4)
template <class Src>
__device__ __forceinline__ double fill_rule(bool pres, bool normal, int bc, bool lo, int N, double d, const double *b,
                                            Src src) {
    if (!pres) {
        if (bc == 1) return normal ? src(1) : (lo ? src(N - 1) : src(2));
        if (bc == 2) return normal ? *b : (lo ? 2 * *b - src(2) : 2 * *b - src(N - 1));
        return *b;
    }
    if (bc == 1) return lo ? src(N - 1) : src(2);
    if (bc == 2) return lo ? 2 * *b - src(2) : 2 * *b - src(N - 1);
    if (bc == 3) return lo ? -d * *b + src(2) : d * *b + src(N - 1);
    return *b;
}

// entry (i, j, k) of the box surface enumerated by t: the two z end planes,
// the two y end rows of the other planes, the two x end columns of the rest
struct Entry {
    int i, j, k;
    bool ok;
};
__device__ __forceinline__ Entry surface_entry(size_t t, size_t n0, size_t n1, size_t n2) {
    const size_t sz = 2 * n0 * n1, sy = 2 * n0 * (n2 - 2), sx = 2 * (n1 - 2) * (n2 - 2);
    if (t < sz) {
        const size_t r = t % (n0 * n1);
        return {(int)(r % n0) + 1, (int)(r / n0) + 1, (t < n0 * n1) ? 1 : (int)n2, true};
    }
    if ((t -= sz) < sy) {
        const size_t r = t % (n0 * (n2 - 2));
        return {(int)(r % n0) + 1, (t < n0 * (n2 - 2)) ? 1 : (int)n1, (int)(r / n0) + 2, true};
    }
    if ((t -= sy) < sx) {
        const size_t r = t % ((n1 - 2) * (n2 - 2));
        return {(t < (n1 - 2) * (n2 - 2)) ? 1 : (int)n0, (int)(r % (n1 - 2)) + 2, (int)(r / (n1 - 2)) + 2, true};
    }
    return {0, 0, 0, false};
}

__host__ __device__ inline size_t surface_size(const int n[3]) {
    return 2 * (size_t)n[0] * n[1] + 2 * (size_t)n[0] * (n[2] - 2) + 2 * (size_t)(n[1] - 2) * (n[2] - 2);
}

// component c (velocity) or dp (pressure); passes along axes >= P.ax
__device__ __forceinline__ void whole_fill_body(const BCParams &P, const size_t t, const int c) {
    const bool pres = (P.pbc_mode == 4);
    // every array and parameter index below is a constant (a run-time index
    // would copy the parameters to scratch memory)
    const int n0 = P.nx + 2 - (!pres && c == 0), n1 = P.ny + 2 - (!pres && c == 1), n2 = P.nz + 2 - (!pres && c == 2);
    const int px = (!pres && c == 0) ? P.pu : P.pc;
    const Entry e = surface_entry(t, n0, n1, n2);
    if (!e.ok) return;
    const int i = e.i, j = e.j, k = e.k;
    double *A = pres ? P.a[0] : (c == 0) ? P.a[0] : (c == 1) ? P.a[1] : P.a[2];
    auto asg = [&](int ax, int m) {
        const int na = (ax == 0) ? n0 : (ax == 1) ? n1 : n2;
        const int bc = (ax == 0) ? P.bcs[0] : (ax == 1) ? P.bcs[1] : P.bcs[2];
        return ax >= P.ax && fill_assigned(pres, c == ax, bc, m, na);
    };
    if (!asg(0, i) && !asg(1, j) && !asg(2, k)) return;
    // the plane entry of an axis' low or high end for the line through (o1, o2)
    auto plane = [&](int ax, bool lo, size_t pl) -> const double * {
        if (((ax == 0) ? P.bcs[0] : (ax == 1) ? P.bcs[1] : P.bcs[2]) == 1) return nullptr;
        const int e = 2 * ax;
        const long o = pres ? (lo ? P.L.off[18 + e] : P.L.off[19 + e])
                       : (c == 0) ? (lo ? P.L.off[e] : P.L.off[e + 1])
                       : (c == 1) ? (lo ? P.L.off[6 + e] : P.L.off[7 + e])
                                  : (lo ? P.L.off[12 + e] : P.L.off[13 + e]);
        return P.planes + o + pl;
    };
    auto vx = [&](int i_, int j_, int k_) -> double {
        if (asg(0, i_))
            return fill_rule(pres, c == 0, P.bcs[0], i_ == 1, n0, P.ds[0],
                             plane(0, i_ == 1, (size_t)(j_ - 1) + (size_t)(k_ - 1) * n1),
                             [&](int m) { return A[GIX(m, j_, k_, px, n1)]; });
        return A[GIX(i_, j_, k_, px, n1)];
    };
    auto vy = [&](int i_, int j_, int k_) -> double {
        if (asg(1, j_))
            return fill_rule(pres, c == 1, P.bcs[1], j_ == 1, n1, P.ds[1],
                             plane(1, j_ == 1, (size_t)(i_ - 1) + (size_t)(k_ - 1) * n0),
                             [&](int m) { return vx(i_, m, k_); });
        return vx(i_, j_, k_);
    };
    double v;
    if (asg(2, k))
        v = fill_rule(pres, c == 2, P.bcs[2], k == 1, n2, P.ds[2],
                      plane(2, k == 1, (size_t)(i - 1) + (size_t)(j - 1) * n0), [&](int m) { return vy(i, j, m); });
    else
        v = vy(i, j, k);
    A[GIX(i, j, k, px, n1)] = v;
    // pressure fill with an accumulation target (the lazily projected
    // dp_flag = 1 step): p = 1*p + dp at the ghost entry, as the projection
    // would form it
    if (pres && P.a[1]) P.a[1][GIX(i, j, k, px, n1)] = P.a[1][GIX(i, j, k, px, n1)] + v;
}

__global__ void __launch_bounds__(256) k_whole_fill(BCParams P) {
    whole_fill_body(P, (size_t)blockIdx.x * blockDim.x + threadIdx.x, blockIdx.y);
}

__global__ void __launch_bounds__(256) k_whole_fill_batch(const BCParams *__restrict__ Ps,
                                                          const int *__restrict__ prefix, int nbox) {
    int blk;
    const BCParams &P = batch_params(Ps, prefix, nbox, blk);
    whole_fill_body(P, (size_t)blk * blockDim.x + threadIdx.x, blockIdx.y);
}

__device__ __forceinline__ void other_axes(int ax, int &o1, int &o2) {
    o1 = (ax == 0) ? 1 : 0;
    o2 = (ax == 2) ? 1 : 2;
}

// boundary rows before an ADI sweep (vel_bc_staggered_CN2, NS_functions.f90:313-410)
__device__ __forceinline__ void cn2_body(const BCParams &P, const size_t t) {
    const Grid g(P.nx, P.ny, P.nz, P.pu, P.pc);
    int o1, o2;
    other_axes(P.ax, o1, o2);
    for (int c = 0; c < 3; c++) {
        int n[3];
        comp_dims(g, c, n);
        const int px = comp_pitch(g, c);
        size_t tot = (size_t)n[o1] * n[o2];
        if (t >= tot) continue;
        int id[3];
        id[o1] = (int)(t % n[o1]) + 1;
        id[o2] = (int)(t / n[o1]) + 1;
        double *A = P.a[c];
        const int N = n[P.ax];
        if (P.bc == 1) {
            if (c != P.ax) { id[P.ax] = 1; A[GIX(id[0], id[1], id[2], px, n[1])] = 0.0; }
            id[P.ax] = N; A[GIX(id[0], id[1], id[2], px, n[1])] = 0.0;
        } else if (P.bc == 2 || P.bc == 4) {
            const double *b1 = P.planes + P.L.off[6 * c + 2 * P.ax];
            const double *bn = P.planes + P.L.off[6 * c + 2 * P.ax + 1];
            id[P.ax] = 1; A[GIX(id[0], id[1], id[2], px, n[1])] = b1[t];
            id[P.ax] = N; A[GIX(id[0], id[1], id[2], px, n[1])] = bn[t];
        }
    }
}

__global__ void __launch_bounds__(256) k_cn2_axis(BCParams P) {
    cn2_body(P, (size_t)blockIdx.x * blockDim.x + threadIdx.x);
}

__global__ void __launch_bounds__(256) k_cn2_axis_batch(const BCParams *__restrict__ Ps, const int *__restrict__ prefix,
                                                        int nbox) {
    int blk;
    const BCParams &P = batch_params(Ps, prefix, nbox, blk);
    cn2_body(P, (size_t)blk * blockDim.x + threadIdx.x);
}

// Poisson RHS boundary terms along one axis (NS_functions.f90:1362-1398)
__global__ void __launch_bounds__(256) k_modify_rhs_axis(BCParams P) {
    int n[3] = {P.nx, P.ny, P.nz}, np_[3] = {P.nx + 2, P.ny + 2, P.nz + 2};
    int o1, o2;
    other_axes(P.ax, o1, o2);
    size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)n[o1] * n[o2]) return;
    int pp = (int)(t % n[o1]) + 1, qq = (int)(t / n[o1]) + 1;
    int i1[3], iN[3];
    i1[o1] = iN[o1] = pp;
    i1[o2] = iN[o2] = qq;
    i1[P.ax] = 1;
    iN[P.ax] = n[P.ax];
    size_t pl = (size_t)pp + (size_t)np_[o1] * (size_t)qq;  // interior (2:np-1) of the full plane
    const double *b1 = P.planes + P.L.off[18 + 2 * P.ax], *bn = P.planes + P.L.off[19 + 2 * P.ax];
    double *r1 = &P.a[0][GIX(i1[0], i1[1], i1[2], P.nx, P.ny)];
    double *rn = &P.a[0][GIX(iN[0], iN[1], iN[2], P.nx, P.ny)];
    const double d = P.d, d2 = P.d * P.d;
    if (P.bc == 2) { *r1 = *r1 - 2 * b1[pl] / d2; *rn = *rn - 2 * bn[pl] / d2; }
    else if (P.bc == 3) { *r1 = *r1 + b1[pl] / d; *rn = *rn - bn[pl] / d; }
    else if (P.bc == 4) { *r1 = *r1 - b1[pl] / d2; *rn = *rn - bn[pl] / d2; }
}

__global__ void __launch_bounds__(256) k_divergence(Grid g, double dx, double dy, double dz, double dt,
                                                    const double *__restrict__ us, const double *__restrict__ vs,
                                                    const double *__restrict__ ws, double *__restrict__ rhs) {
    size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= g.ni()) return;
    int a = (int)(q % g.nx) + 1;
    size_t r = q / g.nx;
    int b = (int)(r % g.ny) + 1, c = (int)(r / g.ny) + 1;
    const int U1 = g.pu, U2 = g.ny + 2, V1 = g.pc, V2 = g.ny + 1, W1 = g.pc, W2 = g.ny + 2;
    double s = (us[GIX(a + 1, b + 1, c + 1, U1, U2)] - us[GIX(a, b + 1, c + 1, U1, U2)]) / dx;
    s = s + (vs[GIX(a + 1, b + 1, c + 1, V1, V2)] - vs[GIX(a + 1, b, c + 1, V1, V2)]) / dy;
    s = s + (ws[GIX(a + 1, b + 1, c + 1, W1, W2)] - ws[GIX(a + 1, b + 1, c, W1, W2)]) / dz;
    rhs[q] = s / dt;
}

struct ProjParams {
    Grid g;
    double d[3], dt;
    int dp_flag;
    const double *s[3];
    const double *dp, *gauge;
    double *o[4];
    double rp[4], r1[4];  // 1/row pitch and 1/rows per plane of each component's arrays
    int org[4];           // offset of each component's array origin from its line-aligned start
};

// storage slot t of component c (c = 3: p) of the projection
// (big_periodic_3D.f90:562-567): u = u* - dt*(dp(i+1)-dp(i))/dx, p = dp_flag*p + dp.
// Slot t is element t - org[c] of the arrays, so the 64 slots of a wavefront
// are one or more whole 128-byte lines; slots before the origin and the spare
// elements at the row ends are skipped.
__device__ __forceinline__ void project_body(const ProjParams &P, const size_t t, const int c) {
    int n[3];
    comp_dims(P.g, c, n);
    const int px = comp_pitch(P.g, c);
    const long qs = (long)t - P.org[c];
    if (qs < 0 || (size_t)qs >= (size_t)px * n[1] * n[2]) return;
    // arrays hold fewer than 2^32 elements (checked at context creation)
    const unsigned qq = (unsigned)qs;
    const unsigned r = udiv_rd(qq, (unsigned)px, P.rp[c]);
    const int i = (int)(qq - r * (unsigned)px) + 1;
    if (i > n[0]) return;
    const size_t q = qq;
    const double gc = P.gauge ? *P.gauge : 0.0;
    if (c == 3) {
        double dpv = P.gauge ? P.dp[q] - gc : P.dp[q];
        P.o[3][q] = (double)P.dp_flag * P.o[3][q] + dpv;
        return;
    }
    const unsigned kk = udiv_rd(r, (unsigned)n[1], P.r1[c]);
    const int j = (int)(r - kk * (unsigned)n[1]) + 1, k = (int)kk + 1;
    int b[3] = {i, j, k};
    b[c]++;
    const int P1 = P.g.pc, P2 = P.g.ny + 2;
    double hi = P.dp[GIX(b[0], b[1], b[2], P1, P2)], lo = P.dp[GIX(i, j, k, P1, P2)];
    if (P.gauge) { hi = hi - gc; lo = lo - gc; }
    double gr = hi - lo;
    P.o[c][q] = P.s[c][q] - P.dt * gr / P.d[c];
}

// a block projects PROJ_PER * 256 consecutive elements of one component,
// 256 apart per thread (coalesced, independent loads in flight)
constexpr int PROJ_PER = 4;

__global__ void __launch_bounds__(256) k_project(ProjParams P) {
    const size_t q0 = (size_t)xcd_slab_block(blockIdx.x, gridDim.x) * (256 * PROJ_PER) + threadIdx.x;
#pragma unroll
    for (int m = 0; m < PROJ_PER; m++) project_body(P, q0 + (size_t)m * 256, blockIdx.y);
}

__global__ void __launch_bounds__(256) k_project_batch(const ProjParams *__restrict__ Ps, const int *__restrict__ prefix,
                                                       int nbox) {
    int blk;
    const ProjParams &P = batch_params(Ps, prefix, nbox, blk);
    const size_t q0 = (size_t)blk * (256 * PROJ_PER) + threadIdx.x;
#pragma unroll
    for (int m = 0; m < PROJ_PER; m++) project_body(P, q0 + (size_t)m * 256, blockIdx.y);
}

struct GetBCParams {
    int nx, ny, nz;
    int pu, pc;  // row pitches of u and of v, w, p
    int ax, bc, corr;
    double d[3], dt;
    PlaneLayout L;
    const double *src[3];
    const double *pprev;
    double *planes;
};

// velocity planes from a (parent) field (get_velpr_bc, NS_functions.f90:446-503)
__global__ void __launch_bounds__(256) k_get_vel_bc_axis(GetBCParams P) {
    const Grid g(P.nx, P.ny, P.nz, P.pu, P.pc);
    int o1, o2;
    other_axes(P.ax, o1, o2);
    size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int P1 = P.pc, P2 = P.ny + 2;
    for (int c = 0; c < 3; c++) {
        int n[3];
        comp_dims(g, c, n);
        const int px = comp_pitch(g, c);
        if (t >= (size_t)n[o1] * n[o2]) continue;
        int id[3];
        id[o1] = (int)(t % n[o1]) + 1;
        id[o2] = (int)(t / n[o1]) + 1;
        const int N = n[P.ax];
        auto val = [&](int pos) -> double {
            int a[3] = {id[0], id[1], id[2]};
            a[P.ax] = pos;
            double x = P.src[c][GIX(a[0], a[1], a[2], px, n[1])];
            if (P.corr) {
                int b[3] = {a[0], a[1], a[2]};
                b[c]++;
                double gr = P.pprev[GIX(b[0], b[1], b[2], P1, P2)] - P.pprev[GIX(a[0], a[1], a[2], P1, P2)];
                x = x + gr / P.d[c] * P.dt;
            }
            return x;
        };
        double *b1 = P.planes + P.L.off[6 * c + 2 * P.ax], *bn = P.planes + P.L.off[6 * c + 2 * P.ax + 1];
        if (P.bc == 4 || c == P.ax) { b1[t] = val(1); bn[t] = val(N); }
        else { b1[t] = (val(1) + val(2)) / 2; bn[t] = (val(N) + val(N - 1)) / 2; }
    }
}

// pressure planes (get_pr_bc, NS_functions.f90:518-560)
__global__ void __launch_bounds__(256) k_get_pr_bc_axis(GetBCParams P) {
    int n[3] = {P.nx + 2, P.ny + 2, P.nz + 2};
    int o1, o2;
    other_axes(P.ax, o1, o2);
    size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)n[o1] * n[o2]) return;
    int id[3];
    id[o1] = (int)(t % n[o1]) + 1;
    id[o2] = (int)(t / n[o1]) + 1;
    const double *p = P.src[0];
    const int N = n[P.ax];
    auto val = [&](int pos) -> double {
        int a[3] = {id[0], id[1], id[2]};
        a[P.ax] = pos;
        return p[GIX(a[0], a[1], a[2], P.pc, n[1])];
    };
    double *b1 = P.planes + P.L.off[18 + 2 * P.ax], *bn = P.planes + P.L.off[19 + 2 * P.ax];
    double d = P.d[P.ax];
    if (P.bc == 2) { b1[t] = (val(1) + val(2)) / 2; bn[t] = (val(N) + val(N - 1)) / 2; }
    else if (P.bc == 4) { b1[t] = val(1); bn[t] = val(N); }
    else if (P.bc == 3) { b1[t] = (val(2) - val(1)) / d; bn[t] = (val(N) - val(N - 1)) / d; }
}

__global__ void __launch_bounds__(256) k_axpby(const double *__restrict__ a, const double *__restrict__ b, double th,
                                               double *__restrict__ out, long n) {
    long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) out[q] = (1.0 - th) * a[q] + th * b[q];
}

// out = sum_i w[i] * a[i*stride + q], i = 0..m-1 (m <= NS_WSUM_MAX), summed in index order
struct WSumParams {
    const double *a;
    long stride, n;
    int m;
    double w[NS_WSUM_MAX];
    double *out;
};
__global__ void __launch_bounds__(256) k_wsum(WSumParams P) {
    long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= P.n) return;
    double s = P.w[0] * P.a[q];
    for (int i = 1; i < P.m; i++) s = s + P.w[i] * P.a[(long)i * P.stride + q];
    P.out[q] = s;
}

__global__ void __launch_bounds__(256) k_copy_interior(Grid g, const double *__restrict__ sol, double *__restrict__ dp) {
    size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= g.ni()) return;
    int a = (int)(q % g.nx) + 1;
    size_t r = q / g.nx;
    int b = (int)(r % g.ny) + 1, c = (int)(r / g.ny) + 1;
    dp[GIX(a + 1, b + 1, c + 1, g.pc, g.ny + 2)] = sol[q];
}

// dst row r = src row r [- *gauge], n0 entries per row, rows ps (src) and pd
// (dst) elements apart: an array between the reference and a context's layout
__global__ void __launch_bounds__(256) k_relayout(const double *__restrict__ src, int ps, double *__restrict__ dst,
                                                  int pd, int n0, unsigned tot, double rn0, const double *gauge) {
    const double gc = gauge ? *gauge : 0.0;
    for (unsigned e = blockIdx.x * 256u + threadIdx.x; e < tot; e += gridDim.x * 256u) {
        const unsigned r = udiv_rd(e, (unsigned)n0, rn0);
        const unsigned i = e - r * (unsigned)n0;
        double x = src[(size_t)r * ps + i];
        if (gauge) x = x - gc;
        dst[(size_t)r * pd + i] = x;
    }
}

__global__ void __launch_bounds__(256) k_sub_gauge(const double *__restrict__ src, const double *gauge,
                                                   double *__restrict__ dst, size_t n) {
    size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) dst[q] = src[q] - *gauge;
}

// dst(e0,e1,e2) = src(o0+1.., o1+1.., o2+1..) [- *gauge]: a box of a ghosted
// array of row pitch px and n1 rows per plane, 0-based origin o
__global__ void __launch_bounds__(256) k_extract_box(const double *__restrict__ src, int px, int n1, int o0, int o1,
                                                     int o2, int e0, int e1, int e2, const double *gauge,
                                                     double *__restrict__ dst) {
    size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t tot = (size_t)e0 * e1 * e2;
    if (q >= tot) return;
    const int i = (int)(q % e0);
    const size_t r = q / e0;
    const int j = (int)(r % e1), k = (int)(r / e1);
    double x = src[GIX(o0 + i + 1, o1 + j + 1, o2 + k + 1, px, n1)];
    if (gauge) x = x - *gauge;
    dst[q] = x;
}

__global__ void __launch_bounds__(256) k_max_div(Grid g, double dx, double dy, double dz, const double *__restrict__ u,
                                                 const double *__restrict__ v, const double *__restrict__ w,
                                                 double *partial) {
    __shared__ double red[256];
    double m = 0.0;
    const int U1 = g.pu, U2 = g.ny + 2, V1 = g.pc, V2 = g.ny + 1, W1 = g.pc, W2 = g.ny + 2;
    for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < g.ni(); q += (size_t)gridDim.x * blockDim.x) {
        int a = (int)(q % g.nx) + 1;
        size_t r = q / g.nx;
        int b = (int)(r % g.ny) + 1, c = (int)(r / g.ny) + 1;
        double s = (u[GIX(a + 1, b + 1, c + 1, U1, U2)] - u[GIX(a, b + 1, c + 1, U1, U2)]) / dx;
        s = s + (v[GIX(a + 1, b + 1, c + 1, V1, V2)] - v[GIX(a + 1, b, c + 1, V1, V2)]) / dy;
        s = s + (w[GIX(a + 1, b + 1, c + 1, W1, W2)] - w[GIX(a + 1, b + 1, c, W1, W2)]) / dz;
        m = fmax(m, fabs(s));
    }
    red[threadIdx.x] = m;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

__global__ void __launch_bounds__(256) k_count_nonfinite(const double *__restrict__ a, size_t n, unsigned int *count) {
    unsigned int local = 0;
    for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x)
        if (!isfinite(a[q])) local++;
    if (local) atomicAdd(count, local);
}


// ---- parent-run capture of a sub-box's boundary planes ----
// (big_periodic_3D.f90:433-489 stores the post-sweep planes of the sub-box;
// get_velpr_bc/get_pr_bc, NS_functions.f90:446-560, define the plane values)
struct CapParams {
    Grid gp, gs;      // parent and sub-box grids
    int o[3];         // 0-based offset of the sub-box origin in the parent arrays
    int ax;           // axis of the planes written by this launch
    int what;         // 0 velocity planes (wall code 2), 1 pressure-gradient planes (code 3)
    double d[3];
    const double *gauge;  // parent corner gauge value (dp(1,1,1)) or null
    PlaneLayout L;        // layout of the sub-box plane set
    const double *src[3];
    double *planes;
};

__device__ __forceinline__ int sel3(const int a[3], int k) { return k == 0 ? a[0] : (k == 1 ? a[1] : a[2]); }

__global__ void __launch_bounds__(256) k_capture_axis(CapParams P) {
    int o1, o2;
    other_axes(P.ax, o1, o2);
    size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int nc = (P.what == 0) ? 3 : 1;
    for (int c = 0; c < nc; c++) {
        const int cc = (P.what == 0) ? c : 3;
        int ns_[3], np_[3];
        comp_dims(P.gs, cc, ns_);
        comp_dims(P.gp, cc, np_);
        const int ppx = comp_pitch(P.gp, cc);
        // every small array is read through selects (a run-time index would
        // place it in scratch memory)
        const int n1 = sel3(ns_, o1), n2 = sel3(ns_, o2);
        if (t >= (size_t)n1 * n2) continue;
        const int p1 = (int)(t % n1), p2 = (int)(t / n1);  // 0-based along o1, o2
        const int N = sel3(ns_, P.ax);
        const double *A = P.src[c];
        const double gc = P.gauge ? *P.gauge : 0.0;
        auto val = [&](int pos) -> double {  // pos is 1-based along ax in the sub-box
            const int ax = P.ax, q = pos - 1;
            const int a0 = (ax == 0 ? q : (o1 == 0 ? p1 : p2)) + P.o[0];
            const int a1 = (ax == 1 ? q : (o1 == 1 ? p1 : p2)) + P.o[1];
            const int a2 = (ax == 2 ? q : (o1 == 2 ? p1 : p2)) + P.o[2];
            double x = A[(size_t)a0 + (size_t)ppx * ((size_t)a1 + (size_t)np_[1] * (size_t)a2)];
            return P.gauge ? x - gc : x;
        };
        if (P.what == 0) {
            double *b1 = P.planes + P.L.off[6 * c + 2 * P.ax], *bn = P.planes + P.L.off[6 * c + 2 * P.ax + 1];
            if (c == P.ax) { b1[t] = val(1); bn[t] = val(N); }
            else { b1[t] = (val(1) + val(2)) / 2; bn[t] = (val(N) + val(N - 1)) / 2; }
        } else {
            double *b1 = P.planes + P.L.off[18 + 2 * P.ax], *bn = P.planes + P.L.off[19 + 2 * P.ax];
            double dd = P.d[P.ax];
            b1[t] = (val(2) - val(1)) / dd;
            bn[t] = (val(N) - val(N - 1)) / dd;
        }
    }
}

// deterministic two-pass mean of the gauged dp over the sub-box interior
__global__ void __launch_bounds__(256) k_subbox_partial(CapParams P, double *partial) {
    __shared__ double red[256];
    const size_t n = P.gs.ni();
    const double gc = P.gauge ? *P.gauge : 0.0;
    double s = 0.0;
    const int P1 = P.gp.pc, P2 = P.gp.ny + 2;
    for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x) {
        int a = (int)(q % P.gs.nx);
        size_t r = q / P.gs.nx;
        int b = (int)(r % P.gs.ny), c = (int)(r / P.gs.ny);
        double x = P.src[0][(size_t)(a + 1 + P.o[0]) + (size_t)P1 * ((size_t)(b + 1 + P.o[1]) + (size_t)P2 * (size_t)(c + 1 + P.o[2]))];
        s += P.gauge ? x - gc : x;
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if ((int)threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

__global__ void k_subbox_mean(const double *partial, int np, double inv_n, double *out) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        double s = 0.0;
        for (int i = 0; i < np; i++) s += partial[i];
        *out = s * inv_n;
    }
}

inline unsigned int nblocks(size_t n, int bs = 256) { return (unsigned int)((n + bs - 1) / bs); }

}  // namespace

PlaneLayout plane_layout(int nx, int ny, int nz) {
    PlaneLayout L;
    long d[4][3] = {{nx + 1, ny + 2, nz + 2}, {nx + 2, ny + 1, nz + 2}, {nx + 2, ny + 2, nz + 1}, {nx + 2, ny + 2, nz + 2}};
    int q = 0;
    for (int c = 0; c < 4; c++) {
        long sx = d[c][1] * d[c][2], sy = d[c][0] * d[c][2], sz = d[c][0] * d[c][1];
        L.size[q++] = sx; L.size[q++] = sx;
        L.size[q++] = sy; L.size[q++] = sy;
        L.size[q++] = sz; L.size[q++] = sz;
    }
    long off = 0;
    for (q = 0; q < 24; q++) { L.off[q] = off; off += L.size[q]; }
    L.off[24] = off;
    L.total = off + 1;
    return L;
}

void launch_conv_diff(const Grid &g, const int bc[3], double h0, double h1, double h2, int which_diff, const double *u,
                      const double *v, const double *w, double *ox, double *oy, double *oz, hipStream_t st) {
    CDParams P;
    P.nx = g.nx; P.ny = g.ny; P.nz = g.nz;
    P.pu = g.pu; P.pc = g.pc;
    P.h0 = h0; P.h1 = h1; P.h2 = h2;
    for (int a = 0; a < 3; a++) P.bc[a] = bc[a];
    P.which = which_diff;
    P.u = u; P.v = v; P.w = w;
    P.o[0] = ox; P.o[1] = oy; P.o[2] = oz;
    hipLaunchKernelGGL(k_conv_diff, dim3(nblocks(g.ni()), 3), dim3(256), 0, st, P);
}

static size_t max_plane(const Grid &g, int ax) {
    size_t m = 0;
    for (int c = 0; c < 3; c++) {
        int n[3];
        comp_dims(g, c, n);
        int o1, o2;
        o1 = (ax == 0) ? 1 : 0;
        o2 = (ax == 2) ? 1 : 2;
        size_t s = (size_t)n[o1] * n[o2];
        m = s > m ? s : m;
    }
    return m;
}

namespace {

// All-periodic ghost fill of up to three arrays in one pass.  The x, y, z
// fills copy interior values and each later axis copies the earlier axes'
// ghost columns, so every ghost entry ends up as the interior value at its
// index wrapped along every axis (pure copies: bitwise the sequential fills).
// Wrap along axis b of length N: the normal axis of a velocity component maps
// N -> 1 (NS_functions.f90:236), the other axes and pressure map 1 -> N-1,
// N -> 2 (:237-240, :1420-1422; the y and z blocks alike).  Sources are never
// ghosts, so no value is read after another thread writes it.
struct PerFill {
    Grid g;
    int comp[3];  // 0..2 velocity component, 3 pressure
    double *a[3];
    int first_axis;  // faces of the axes first_axis..2 are written
};

__global__ void __launch_bounds__(256) k_periodic_fill(const PerFill P) {
    const int k = blockIdx.y, face = blockIdx.z, ax = face >> 1;
    if (ax < P.first_axis) return;
    const int c = P.comp[k];
    int n[3];
    comp_dims(P.g, c < 3 ? c : 3, n);
    const int o1 = (ax == 0) ? 1 : 0, o2 = (ax == 2) ? 1 : 2;
    const size_t t = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (t >= (size_t)n[o1] * n[o2]) return;
    int id[3];
    id[o1] = (int)(t % n[o1]) + 1;
    id[o2] = (int)(t / n[o1]) + 1;
    id[ax] = (face & 1) ? n[ax] : 1;
    int src[3];
#pragma unroll
    for (int b = 0; b < 3; b++) {
        const int N = n[b], i = id[b];
        if (b == c) src[b] = (i == N) ? 1 : i;
        else src[b] = (i == 1) ? N - 1 : (i == N) ? 2 : i;
    }
    if (src[0] == id[0] && src[1] == id[1] && src[2] == id[2]) return;
    double *A = P.a[k];
    const int px = comp_pitch(P.g, c < 3 ? c : 3);
    A[GIX(id[0], id[1], id[2], px, n[1])] = A[GIX(src[0], src[1], src[2], px, n[1])];
}

void launch_periodic_fill(const Grid &g, int na, const int comp[3], double *const a[3], int first_axis,
                          hipStream_t st) {
    PerFill P;
    P.g = g;
    for (int k = 0; k < 3; k++) {
        P.comp[k] = k < na ? comp[k] : 3;
        P.a[k] = k < na ? a[k] : nullptr;
    }
    P.first_axis = first_axis;
    size_t mx = 0;
    for (int k = 0; k < na; k++) {
        int n[3];
        comp_dims(g, comp[k], n);
        const size_t f[3] = {(size_t)n[1] * n[2], (size_t)n[0] * n[2], (size_t)n[0] * n[1]};
        for (int b = 0; b < 3; b++) mx = f[b] > mx ? f[b] : mx;
    }
    hipLaunchKernelGGL(k_periodic_fill, dim3(nblocks(mx), na, 6), dim3(256), 0, st, P);
}

}  // namespace

int make_fill_params(const Grid &g, bool pres, const int bc[3], const double *planes, const double d[3], double *a0,
                     double *a1, double *a2, int first_axis, void *out) {
    BCParams &P = *(BCParams *)out;
    P = BCParams{};
    P.nx = g.nx; P.ny = g.ny; P.nz = g.nz;
    P.pu = g.pu; P.pc = g.pc;
    P.L = plane_layout(g.nx, g.ny, g.nz);
    P.planes = planes;
    P.a[0] = a0; P.a[1] = a1; P.a[2] = a2;
    P.pbc_mode = pres ? 4 : 3;
    P.ax = first_axis;
    bool any = false;
    for (int a = 0; a < 3; a++) {
        P.bcs[a] = bc[a];
        P.ds[a] = d[a];
        if (a >= first_axis && (pres || bc[a] != 3)) any = true;
    }
    if (!any) return 0;
    size_t mx = 0;
    for (int c = 0; c < (pres ? 1 : 3); c++) {
        int n[3];
        comp_dims(g, pres ? 3 : c, n);
        const size_t m = surface_size(n);
        mx = m > mx ? m : mx;
    }
    return (int)nblocks(mx);
}

void launch_whole_fill_batch(const void *params, const int *prefix, int nbox, int nblk, int ncomp, hipStream_t st) {
    hipLaunchKernelGGL(k_whole_fill_batch, dim3(nblk, ncomp), dim3(256), 0, st, (const BCParams *)params, prefix,
                       nbox);
}

void launch_vel_bc(const Grid &g, const int bc[3], const double *planes, double *u, double *v, double *w,
                   hipStream_t st, int first_axis) {
    if (bc[0] == 1 && bc[1] == 1 && bc[2] == 1) {
        const int comp[3] = {0, 1, 2};
        double *const a[3] = {u, v, w};
        launch_periodic_fill(g, 3, comp, a, first_axis, st);
        return;
    }
    BCParams P;
    const double d[3] = {0.0, 0.0, 0.0};
    const int nb = make_fill_params(g, false, bc, planes, d, u, v, w, first_axis, &P);
    if (nb) hipLaunchKernelGGL(k_whole_fill, dim3(nb, 3), dim3(256), 0, st, P);
}

int make_cn2_params(const Grid &g, int bc, int ax, const double *planes, double *u, double *v, double *w, void *out) {
    BCParams &P = *(BCParams *)out;
    P.nx = g.nx; P.ny = g.ny; P.nz = g.nz;
    P.pu = g.pu; P.pc = g.pc;
    P.L = plane_layout(g.nx, g.ny, g.nz);
    P.planes = planes;
    P.a[0] = u; P.a[1] = v; P.a[2] = w;
    P.ax = ax;
    P.bc = bc;
    P.pbc_mode = 1;
    P.d = 0;
    return bc == 3 ? 0 : (int)nblocks(max_plane(g, ax));
}

void launch_vel_bc_cn2(const Grid &g, int bc, int dir, const double *planes, double *u, double *v, double *w,
                       hipStream_t st) {
    BCParams P;
    const int nb = make_cn2_params(g, bc, dir - 1, planes, u, v, w, &P);
    if (nb) hipLaunchKernelGGL(k_cn2_axis, dim3(nb), dim3(256), 0, st, P);
}

void launch_cn2_batch(const void *params, const int *prefix, int nbox, int nblk, hipStream_t st) {
    hipLaunchKernelGGL(k_cn2_axis_batch, dim3(nblk), dim3(256), 0, st, (const BCParams *)params, prefix, nbox);
}

void launch_pr_bc(const Grid &g, const int pbc[3], const double *planes, double dx, double dy, double dz, double *dp,
                  hipStream_t st, int first_axis, double *p_acc) {
    if (pbc[0] == 1 && pbc[1] == 1 && pbc[2] == 1) {
        const int comp[3] = {3, 3, 3};
        double *const a[3] = {dp, nullptr, nullptr};
        launch_periodic_fill(g, 1, comp, a, first_axis, st);
        return;
    }
    BCParams P;
    const double d[3] = {dx, dy, dz};
    const int nb = make_fill_params(g, true, pbc, planes, d, dp, p_acc, nullptr, first_axis, &P);
    if (nb) hipLaunchKernelGGL(k_whole_fill, dim3(nb, 1), dim3(256), 0, st, P);
}

void launch_modify_rhs(const Grid &g, const int pbc[3], const double *planes, double dx, double dy, double dz,
                       double *rhs, hipStream_t st) {
    BCParams P;
    P.nx = g.nx; P.ny = g.ny; P.nz = g.nz;
    P.pu = g.pu; P.pc = g.pc;
    P.L = plane_layout(g.nx, g.ny, g.nz);
    P.planes = planes;
    P.a[0] = rhs; P.a[1] = nullptr; P.a[2] = nullptr;
    double d[3] = {dx, dy, dz};
    int n[3] = {g.nx, g.ny, g.nz};
    for (int ax = 0; ax < 3; ax++) {
        if (pbc[ax] == 1) continue;
        P.ax = ax;
        P.bc = pbc[ax];
        P.d = d[ax];
        int o1 = (ax == 0) ? 1 : 0, o2 = (ax == 2) ? 1 : 2;
        hipLaunchKernelGGL(k_modify_rhs_axis, dim3(nblocks((size_t)n[o1] * n[o2])), dim3(256), 0, st, P);
    }
}

void launch_divergence(const Grid &g, double dx, double dy, double dz, double dt, const double *us, const double *vs,
                       const double *ws, double *rhs, hipStream_t st) {
    hipLaunchKernelGGL(k_divergence, dim3(nblocks(g.ni())), dim3(256), 0, st, g, dx, dy, dz, dt, us, vs, ws, rhs);
}

void launch_project(const Grid &g, double dx, double dy, double dz, double dt, int dp_flag, const double *us,
                    const double *vs, const double *ws, const double *dp, const double *gauge, double *u, double *v,
                    double *w, double *p, hipStream_t st) {
    ProjParams P;
    const double d[3] = {dx, dy, dz};
    const int nb = make_project_params(g, d, dt, dp_flag, us, vs, ws, dp, gauge, u, v, w, p, &P);
    // p null: the velocity only (the pressure is already formed)
    hipLaunchKernelGGL(k_project, dim3(nb, p ? 4 : 3), dim3(256), 0, st, P);
}

void launch_get_velpr_bc(const Grid &g, const int bc[3], const int pbc[3], double dx, double dy, double dz, double dt,
                         int corr, const double *ubc, const double *vbc, const double *wbc, const double *pbc_arr,
                         const double *pprev, double *planes, hipStream_t st) {
    GetBCParams P;
    P.nx = g.nx; P.ny = g.ny; P.nz = g.nz;
    P.pu = g.pu; P.pc = g.pc;
    P.d[0] = dx; P.d[1] = dy; P.d[2] = dz;
    P.dt = dt;
    P.corr = corr && (bc[0] != 1 || bc[1] != 1 || bc[2] != 1);
    P.L = plane_layout(g.nx, g.ny, g.nz);
    P.src[0] = ubc; P.src[1] = vbc; P.src[2] = wbc;
    P.pprev = pprev;
    P.planes = planes;
    for (int ax = 0; ax < 3; ax++) {
        if (bc[ax] != 2 && bc[ax] != 4) continue;
        P.ax = ax;
        P.bc = bc[ax];
        hipLaunchKernelGGL(k_get_vel_bc_axis, dim3(nblocks(max_plane(g, ax))), dim3(256), 0, st, P);
    }
    launch_get_pr_bc(g, pbc, dx, dy, dz, pbc_arr, planes, st);
}

void launch_get_pr_bc(const Grid &g, const int pbc[3], double dx, double dy, double dz, const double *p,
                      double *planes, hipStream_t st) {
    GetBCParams P;
    P.nx = g.nx; P.ny = g.ny; P.nz = g.nz;
    P.pu = g.pu; P.pc = g.pc;
    P.d[0] = dx; P.d[1] = dy; P.d[2] = dz;
    P.dt = 0;
    P.corr = 0;
    P.L = plane_layout(g.nx, g.ny, g.nz);
    P.src[0] = p; P.src[1] = nullptr; P.src[2] = nullptr;
    P.pprev = nullptr;
    P.planes = planes;
    int n[3] = {g.nx + 2, g.ny + 2, g.nz + 2};
    for (int ax = 0; ax < 3; ax++) {
        if (pbc[ax] == 1) continue;
        P.ax = ax;
        P.bc = pbc[ax];
        int o1 = (ax == 0) ? 1 : 0, o2 = (ax == 2) ? 1 : 2;
        hipLaunchKernelGGL(k_get_pr_bc_axis, dim3(nblocks((size_t)n[o1] * n[o2])), dim3(256), 0, st, P);
    }
}

void launch_wsum(const double *a, long stride, const double *w, int m, double *out, long n, hipStream_t st) {
    WSumParams P;
    P.a = a;
    P.stride = stride;
    P.n = n;
    P.m = m;
    for (int i = 0; i < NS_WSUM_MAX; i++) P.w[i] = i < m ? w[i] : 0.0;
    P.out = out;
    hipLaunchKernelGGL(k_wsum, dim3(nblocks(n)), dim3(256), 0, st, P);
}

void launch_axpby(const double *a, const double *b, double th, double *out, long n, hipStream_t st) {
    hipLaunchKernelGGL(k_axpby, dim3(nblocks(n)), dim3(256), 0, st, a, b, th, out, n);
}

void launch_copy_interior(const Grid &g, const double *sol, double *dp, hipStream_t st) {
    hipLaunchKernelGGL(k_copy_interior, dim3(nblocks(g.ni())), dim3(256), 0, st, g, sol, dp);
}

void launch_extract_box(const double *src, const int n[3], const int o[3], const int e[3], const double *gauge,
                        double *dst, hipStream_t st) {
    const size_t tot = (size_t)e[0] * e[1] * e[2];
    hipLaunchKernelGGL(k_extract_box, dim3(nblocks(tot)), dim3(256), 0, st, src, n[0], n[1], o[0], o[1], o[2], e[0],
                       e[1], e[2], gauge, dst);
}

void launch_relayout(const double *src, int ps, double *dst, int pd, int n0, size_t rows, const double *gauge,
                     hipStream_t st) {
    const size_t tot = (size_t)n0 * rows;
    size_t nb = nblocks(tot);
    if (nb > 8192) nb = 8192;
    if (nb < 1) nb = 1;
    hipLaunchKernelGGL(k_relayout, dim3((unsigned)nb), dim3(256), 0, st, src, ps, dst, pd, n0, (unsigned)tot,
                       1.0 / n0, gauge);
}

void launch_sub_gauge(const double *src, const double *gauge, double *dst, size_t n, hipStream_t st) {
    hipLaunchKernelGGL(k_sub_gauge, dim3(nblocks(n)), dim3(256), 0, st, src, gauge, dst, n);
}

int max_div_partials(const Grid &g) {
    size_t nb = nblocks(g.ni());
    return (int)(nb < 2048 ? nb : 2048);
}

int launch_max_div(const Grid &g, double dx, double dy, double dz, const double *u, const double *v, const double *w,
                   double *partial, hipStream_t st) {
    int nb = max_div_partials(g);
    hipLaunchKernelGGL(k_max_div, dim3(nb), dim3(256), 0, st, g, dx, dy, dz, u, v, w, partial);
    return nb;
}

void launch_count_nonfinite(const double *a, size_t n, unsigned int *count, hipStream_t st) {
    size_t nb = nblocks(n);
    if (nb > 4096) nb = 4096;
    hipLaunchKernelGGL(k_count_nonfinite, dim3((unsigned)nb), dim3(256), 0, st, a, n, count);
}


void launch_capture(const Grid &gp, const Grid &gs, const int o[3], int ax, int what, const double d[3],
                    const double *gauge, const double *s0, const double *s1, const double *s2, double *planes,
                    hipStream_t st) {
    CapParams P;
    P.gp = gp; P.gs = gs;
    for (int a = 0; a < 3; a++) { P.o[a] = o[a]; P.d[a] = d[a]; }
    P.ax = ax; P.what = what;
    P.gauge = gauge;
    P.L = plane_layout(gs.nx, gs.ny, gs.nz);
    P.src[0] = s0; P.src[1] = s1; P.src[2] = s2;
    P.planes = planes;
    size_t m = what == 0 ? max_plane(gs, ax) : 0;
    if (what == 1) {
        int n[3] = {gs.nx + 2, gs.ny + 2, gs.nz + 2};
        int o1 = (ax == 0) ? 1 : 0, o2 = (ax == 2) ? 1 : 2;
        m = (size_t)n[o1] * n[o2];
    }
    hipLaunchKernelGGL(k_capture_axis, dim3(nblocks(m)), dim3(256), 0, st, P);
}

void launch_subbox_mean(const Grid &gp, const Grid &gs, const int o[3], const double *gauge, const double *dp,
                        double *partial, double *out, hipStream_t st) {
    CapParams P{};
    P.gp = gp; P.gs = gs;
    for (int a = 0; a < 3; a++) P.o[a] = o[a];
    P.gauge = gauge;
    P.src[0] = dp;
    int nb = (int)nblocks(gs.ni());
    if (nb > 1024) nb = 1024;
    hipLaunchKernelGGL(k_subbox_partial, dim3(nb), dim3(256), 0, st, P, partial);
    hipLaunchKernelGGL(k_subbox_mean, dim3(1), dim3(64), 0, st, partial, nb, 1.0 / (double)gs.ni(), out);
}

// ---- per-box parameters and batched launches (ns_batch) ----
size_t bc_params_size() { return sizeof(BCParams); }
size_t proj_params_size() { return sizeof(ProjParams); }

int make_project_params(const Grid &g, const double d[3], double dt, int dp_flag, const double *us, const double *vs,
                        const double *ws, const double *dp, const double *gauge, double *u, double *v, double *w,
                        double *p, void *out) {
    ProjParams &P = *(ProjParams *)out;
    P.g = g;
    P.d[0] = d[0]; P.d[1] = d[1]; P.d[2] = d[2];
    P.dt = dt;
    P.dp_flag = dp_flag;
    P.s[0] = us; P.s[1] = vs; P.s[2] = ws;
    P.dp = dp;
    P.gauge = gauge;
    P.o[0] = u; P.o[1] = v; P.o[2] = w; P.o[3] = p;
    for (int c = 0; c < 4; c++) {
        int n[3];
        comp_dims(g, c, n);
        P.rp[c] = 1.0 / comp_pitch(g, c);
        P.r1[c] = 1.0 / n[1];
        P.org[c] = field_origin(g, c);
    }
    // storage slots of the largest array, its origin offset included
    return (int)round8(nblocks(g.np() + (size_t)field_origin(g, 3), 256 * PROJ_PER));
}

void launch_project_batch(const void *params, const int *prefix, int nbox, int nblk, hipStream_t st) {
    hipLaunchKernelGGL(k_project_batch, dim3(nblk, 4), dim3(256), 0, st, (const ProjParams *)params, prefix, nbox);
}

}  // namespace ns
