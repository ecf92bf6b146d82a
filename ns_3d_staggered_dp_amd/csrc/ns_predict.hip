// Fused predictor for gfx950, FP64: cal_conv + cal_diff + Euler/AB2/AB2-CN
// update (+ pressure gradient) for all three velocity components in one pass
// (NS_functions.f90:53-210, big_periodic_3D.f90:343-424,
// re_simulation_sub_tstep_interBC_combine.f90:543-654).
//
// Layout of the work: a block owns a 64 x 4 column of cells and streams
// through z ("2.5-D" blocking).  Each component keeps a ring of three z-slices
// of its halo'd 66 x 6 tile in LDS, so every input value is fetched from
// memory about 1.5 times (the x/y halo, mostly served by L2) instead of once
// per stencil use; the next slices are prefetched into registers while the
// current one is computed.  Only
// the points of the reference's assignment ranges are computed (x_range0 etc.,
// big_periodic_3D.f90:126-146); the ghost entries of u* are rewritten by the
// boundary fill or discarded by the ADI sweeps in every scheme, so they are
// left untouched, and so are the ghost entries of the AB2 history.
//
// LAZY: the input velocity is the previous step's u* and dp, and the
// projection u = u* - dt*(dp(i+1)-dp(i))/dx (big_periodic_3D.f90:562-567,
// with the corner gauge) is evaluated while the tiles are loaded, exactly as
// the projection kernel computes it.  This removes one full read/write pass
// of u, v, w per step.
//
// Operation order inside each formula is the reference's (see ns_stencil.hip).
// Divisions by the (per-launch constant) grid spacings are evaluated as
// q = a*r, q += (a - q*d)*r with r = 1/d rounded on the host (Markstein's
// correction, two FMAs): the quotient is the correctly rounded a/d except in
// rare last-bit cases, at a third of the cost of the full IEEE division
// sequence.  Results therefore agree with the per-point kernel to the last
// bit almost everywhere and to within one rounding elsewhere.
#include "ns_internal.h"

namespace ns {

namespace {

constexpr int TX = 64, TY = 4;
constexpr int HX = TX + 2, HY = TY + 2;

struct TileParams {
    int nx, ny, nz;
    double dx, dy, dz, nu, dt;
    double rd[3], rd2[3];  // 1/dx, 1/dy, 1/dz and 1/dx^2, ... rounded on the host
    int scheme, first, dp_flag;
    int per[3];          // own-axis periodicity of u, v, w (bc == 1)
    int zc;              // z cells per block
    int tiles_x, tiles_y, tiles_z, ntiles;
    const double *src[3];  // u, v, w (or u*, v*, w* when lazy)
    const double *dp, *gauge;  // lazy projection inputs
    const double *p;
    double *h[3];
    double *out[3];
};

// a/d with r = RN(1/d): one product and one FMA correction step
__device__ __forceinline__ double qdiv(double a, double d, double r) {
    const double q = a * r;
    const double e = __builtin_fma(-q, d, a);
    return __builtin_fma(e, r, q);
}

__device__ __forceinline__ void cdims(const TileParams &P, int c, int n[3]) {
    n[0] = P.nx + 2; n[1] = P.ny + 2; n[2] = P.nz + 2;
    n[c] -= 1;
}

// global 1-based index of interior cell a (0-based) along axis ax for component c
__device__ __forceinline__ int gidx(int c, int ax, int a) { return a + ((ax == c) ? 1 : 2); }

template <bool LAZY>
__device__ __forceinline__ double load_val(const TileParams &P, int c, int gi, int gj, int gk) {
    int n[3];
    cdims(P, c, n);
    // periodic wrap of the own-axis left neighbour of face 1 (index 0 -> n)
    if (c == 0 && gi == 0 && P.per[0]) gi = P.nx;
    if (c == 1 && gj == 0 && P.per[1]) gj = P.ny;
    if (c == 2 && gk == 0 && P.per[2]) gk = P.nz;
    gi = gi < 1 ? 1 : (gi > n[0] ? n[0] : gi);
    gj = gj < 1 ? 1 : (gj > n[1] ? n[1] : gj);
    gk = gk < 1 ? 1 : (gk > n[2] ? n[2] : gk);
    size_t q = FIX(gi, gj, gk, n[0], n[1]);
    double x = P.src[c][q];
    if (LAZY) {
        const int P1 = P.nx + 2, P2 = P.ny + 2;
        int b[3] = {gi, gj, gk};
        b[c]++;
        double hi = P.dp[FIX(b[0], b[1], b[2], P1, P2)], lo = P.dp[FIX(gi, gj, gk, P1, P2)];
        if (P.gauge) {
            const double g = *P.gauge;
            hi = hi - g;
            lo = lo - g;
        }
        double gr = hi - lo;
        const double d = (c == 0) ? P.dx : (c == 1) ? P.dy : P.dz;
        x = x - qdiv(P.dt * gr, d, P.rd[c]);
    }
    return x;
}

template <bool LAZY>
__global__ void __launch_bounds__(TX *TY) k_predict_tiled(TileParams P) {
    __shared__ double S[3][3][HY][HX];  // [component][slot][y][x]
    // XCD-aware tile order: consecutive tiles of one XCD are x/y neighbours
    const int nb = P.ntiles;
    int b = blockIdx.x;
    const int per_xcd = (nb + 7) / 8;
    int t = (b % 8) * per_xcd + b / 8;
    if (t >= nb) return;  // only when nb is not a multiple of 8: padding blocks exit at once
    const int tx_ = t % P.tiles_x;
    const int ty_ = (t / P.tiles_x) % P.tiles_y;
    const int tz_ = t / (P.tiles_x * P.tiles_y);
    const int a0 = tx_ * TX, b0 = ty_ * TY, z0 = tz_ * P.zc;
    const int z1 = min(z0 + P.zc, P.nz);
    const int tx = threadIdx.x % TX, ty = threadIdx.x / TX;
    // tile origins (global index of local 0) per component
    int ox[3], oy[3];
    for (int c = 0; c < 3; c++) {
        ox[c] = gidx(c, 0, a0) - 1;
        oy[c] = gidx(c, 1, b0) - 1;
    }
    // Per-thread load descriptors: thread e loads halo elements e and e+TX*TY
    // (HX x HY per slice) of each component.  Their x/y offsets (with the
    // periodic wrap of the own-axis index 0 and clamping) are fixed for the
    // whole column, so each slice load is one add and one load.
    constexpr int NE = 2;
    unsigned int offc[3][NE], offp[3][NE];  // < 2^32 elements per array (up to ~1600^3)
    bool valid[NE];
    int lxy[NE][2];
    const int P1 = P.nx + 2, P2 = P.ny + 2;
    const size_t pp = (size_t)P1 * P2;
    size_t ps[3];
    int nz3[3];
    for (int s_ = 0; s_ < NE; s_++) {
        const int e = threadIdx.x + s_ * TX * TY;
        valid[s_] = e < HX * HY;
        const int la = e % HX, lb = e / HX;
        lxy[s_][0] = la;
        lxy[s_][1] = lb;
        for (int c = 0; c < 3; c++) {
            int n[3];
            cdims(P, c, n);
            int gi = ox[c] + la, gj = oy[c] + lb;
            if (c == 0 && gi == 0 && P.per[0]) gi = P.nx;
            if (c == 1 && gj == 0 && P.per[1]) gj = P.ny;
            gi = gi < 1 ? 1 : (gi > n[0] ? n[0] : gi);
            gj = gj < 1 ? 1 : (gj > n[1] ? n[1] : gj);
            offc[c][s_] = (unsigned int)((gi - 1) + n[0] * (gj - 1));
            offp[c][s_] = (unsigned int)((gi - 1) + P1 * (gj - 1));
            if (s_ == 0) {
                ps[c] = (size_t)n[0] * n[1];
                nz3[c] = n[2];
            }
        }
    }
    const double gauge = (LAZY && P.gauge) ? *P.gauge : 0.0;
    const double dgrad[3] = {P.dx, P.dy, P.dz};
    // own-axis step of the dp neighbour used by the lazy projection of component c
    const size_t dstep[3] = {1, (size_t)P1, pp};
    auto keff = [&](int c, int gk) {
        if (c == 2 && gk == 0 && P.per[2]) gk = P.nz;
        return gk < 1 ? 1 : (gk > nz3[c] ? nz3[c] : gk);
    };
    struct Regs {
        double v[3][NE], lo[3][NE], hi[3][NE];
    };
    auto fetch = [&](Regs &R, int c, int gk) {
        const int k = keff(c, gk);
        const double *src = P.src[c] + (size_t)(k - 1) * ps[c];
        const double *dpk = LAZY ? P.dp + (size_t)(k - 1) * pp : nullptr;
#pragma unroll
        for (int s_ = 0; s_ < NE; s_++) {
            if (!valid[s_]) continue;
            R.v[c][s_] = src[offc[c][s_]];
            if (LAZY) {
                R.lo[c][s_] = dpk[offp[c][s_]];
                R.hi[c][s_] = dpk[offp[c][s_] + dstep[c]];
            }
        }
    };
    auto store = [&](const Regs &R, int c, int gk) {
        double(*dst)[HX] = S[c][gk % 3];
#pragma unroll
        for (int s_ = 0; s_ < NE; s_++) {
            if (!valid[s_]) continue;
            double x = R.v[c][s_];
            if (LAZY) {
                // the projection of big_periodic_3D.f90:562-567, gauge applied first
                double hi = R.hi[c][s_], lo = R.lo[c][s_];
                if (P.gauge) {
                    hi = hi - gauge;
                    lo = lo - gauge;
                }
                double gr = hi - lo;
                x = x - qdiv(P.dt * gr, dgrad[c], P.rd[c]);
            }
            dst[lxy[s_][1]][lxy[s_][0]] = x;
        }
    };
    Regs R;
    // prologue: u, v need slices K-1..K+1 with K = az+2; w needs fk-1..fk+1 with fk = az+1
    for (int q = 0; q < 3; q++) {
        fetch(R, 0, z0 + 1 + q); fetch(R, 1, z0 + 1 + q); fetch(R, 2, z0 + q);
        store(R, 0, z0 + 1 + q); store(R, 1, z0 + 1 + q); store(R, 2, z0 + q);
    }
    __syncthreads();
    const double dx = P.dx, dy = P.dy, dz = P.dz, dx2 = dx * dx, dy2 = dy * dy, dz2 = dz * dz;
    const double rx = P.rd[0], ry = P.rd[1], rz = P.rd[2], rx2 = P.rd2[0], ry2 = P.rd2[1], rz2 = P.rd2[2];
    const double dt = P.dt, nu = P.nu, dpf = (double)P.dp_flag;
    const int a = a0 + tx, bb = b0 + ty;
    for (int az = z0; az < z1; az++) {
        const bool more = az + 1 < z1;
        // prefetch the slices of the next iteration; their latency overlaps the compute below
        if (more) {
            fetch(R, 0, az + 4);
            fetch(R, 1, az + 4);
            fetch(R, 2, az + 3);
        }
        if (a < P.nx && bb < P.ny) {
#define LU(gi, gj, gk) S[0][(gk) % 3][(gj) - oy[0]][(gi) - ox[0]]
#define LV(gi, gj, gk) S[1][(gk) % 3][(gj) - oy[1]][(gi) - ox[1]]
#define LW(gi, gj, gk) S[2][(gk) % 3][(gj) - oy[2]][(gi) - ox[2]]
#define EF(i, j, k) (((LU(i, (j) + 1, k) + LU(i, j, k)) / 2.0) * ((LV((i) + 1, j, k) + LV(i, j, k)) / 2.0))
#define FF(i, j, k) (((LU(i, j, (k) + 1) + LU(i, j, k)) / 2.0) * ((LW((i) + 1, j, k) + LW(i, j, k)) / 2.0))
#define GF(i, j, k) (((LV(i, j, (k) + 1) + LV(i, j, k)) / 2.0) * ((LW(i, (j) + 1, k) + LW(i, j, k)) / 2.0))
            for (int c = 0; c < 3; c++) {
                // own-axis in-range: periodic 0..n-1, walls 1..n-1 (faces 2..n)
                const int own = (c == 0) ? a : (c == 1) ? bb : az;
                if (!P.per[c] && own == 0) continue;
                const int gi = gidx(c, 0, a), gj = gidx(c, 1, bb), gk = gidx(c, 2, az);
                double cv, df, c0;
                if (c == 0) {
                    const int fi = gi, J = gj, K = gk;
                    c0 = LU(fi, J, K);
                    const int fl = fi - 1;  // tile local index 0 holds the periodic wrap when fi == 1
                    double ur = (LU(fi + 1, J, K) + c0) / 2.0;
                    // avg(bc=1) wraps with u(nx+1) and u(nx) (FD_functions.f90:252)
                    double ul = (P.per[0] && fi == 1)
                                    ? (load_val<LAZY>(P, 0, P.nx + 1, J, K) + LU(fl, J, K)) / 2.0
                                    : (c0 + LU(fl, J, K)) / 2.0;
                    double t1 = ur * ur - ul * ul;
                    double t2 = EF(fi, J, K) - EF(fi, J - 1, K);
                    double t3 = FF(fi, J, K) - FF(fi, J, K - 1);
                    double r = qdiv(t1, dx, rx);
                    r = r + qdiv(t2, dy, ry);
                    r = r + qdiv(t3, dz, rz);
                    cv = r;
                    double d1 = LU(fl, J, K) - 2.0 * c0 + LU(fi + 1, J, K);
                    double d2 = LU(fi, J - 1, K) - 2.0 * c0 + LU(fi, J + 1, K);
                    double d3 = LU(fi, J, K - 1) - 2.0 * c0 + LU(fi, J, K + 1);
                    double s = qdiv(d1, dx2, rx2);
                    s = s + qdiv(d2, dy2, ry2);
                    s = s + qdiv(d3, dz2, rz2);
                    df = s;
                } else if (c == 1) {
                    const int I = gi, fj = gj, K = gk;
                    c0 = LV(I, fj, K);
                    const int fl = fj - 1;
                    double t1 = EF(I, fj, K) - EF(I - 1, fj, K);
                    double vr = (LV(I, fj + 1, K) + c0) / 2.0;
                    double vl = (P.per[1] && fj == 1)
                                    ? (load_val<LAZY>(P, 1, I, P.ny + 1, K) + LV(I, fl, K)) / 2.0
                                    : (c0 + LV(I, fl, K)) / 2.0;
                    double t2 = vr * vr - vl * vl;
                    double t3 = GF(I, fj, K) - GF(I, fj, K - 1);
                    double r = qdiv(t1, dx, rx);
                    r = r + qdiv(t2, dy, ry);
                    r = r + qdiv(t3, dz, rz);
                    cv = r;
                    double d1 = LV(I - 1, fj, K) - 2.0 * c0 + LV(I + 1, fj, K);
                    double d2 = LV(I, fl, K) - 2.0 * c0 + LV(I, fj + 1, K);
                    double d3 = LV(I, fj, K - 1) - 2.0 * c0 + LV(I, fj, K + 1);
                    double s = qdiv(d1, dx2, rx2);
                    s = s + qdiv(d2, dy2, ry2);
                    s = s + qdiv(d3, dz2, rz2);
                    df = s;
                } else {
                    const int I = gi, J = gj, fk = gk;
                    c0 = LW(I, J, fk);
                    const int fl = fk - 1;
                    double t1 = FF(I, J, fk) - FF(I - 1, J, fk);
                    double t2 = GF(I, J, fk) - GF(I, J - 1, fk);
                    double wr = (LW(I, J, fk + 1) + c0) / 2.0;
                    double wl = (P.per[2] && fk == 1)
                                    ? (load_val<LAZY>(P, 2, I, J, P.nz + 1) + LW(I, J, fl)) / 2.0
                                    : (c0 + LW(I, J, fl)) / 2.0;
                    double t3 = wr * wr - wl * wl;
                    double r = qdiv(t1, dx, rx);
                    r = r + qdiv(t2, dy, ry);
                    r = r + qdiv(t3, dz, rz);
                    cv = r;
                    double d1 = LW(I - 1, J, fk) - 2.0 * c0 + LW(I + 1, J, fk);
                    double d2 = LW(I, J - 1, fk) - 2.0 * c0 + LW(I, J + 1, fk);
                    double d3 = LW(I, J, fl) - 2.0 * c0 + LW(I, J, fk + 1);
                    double s = qdiv(d1, dx2, rx2);
                    s = s + qdiv(d2, dy2, ry2);
                    s = s + qdiv(d3, dz2, rz2);
                    df = s;
                }
                int n[3];
                cdims(P, c, n);
                const size_t q = FIX(gi, gj, gk, n[0], n[1]);
                double dpd = 0.0;
                if (P.dp_flag) {
                    int bq[3] = {gi, gj, gk};
                    bq[c]++;
                    const double dd = (c == 0) ? dx : (c == 1) ? dy : dz;
                    dpd = qdiv(P.p[FIX(bq[0], bq[1], bq[2], P.nx + 2, P.ny + 2)] - P.p[FIX(gi, gj, gk, P.nx + 2, P.ny + 2)],
                               dd, P.rd[c]);
                }
                double *H = P.h[c];
                double rhs, val;
                if (P.scheme == NS_EULER) {
                    rhs = nu * df - cv;
                    val = dt * (1.0 * rhs - dpf * dpd) + c0;
                    H[q] = 0.0;
                } else if (P.scheme == NS_AB2) {
                    rhs = nu * df - cv;
                    if (P.first) val = dt * (1.0 * rhs - dpf * dpd) + c0;
                    else val = dt * (1.5 * rhs - 0.5 * H[q] - dpf * dpd) + c0;
                    H[q] = rhs;
                } else {
                    if (P.first) {
                        rhs = 0.5 * nu * df - cv;
                        val = dt * (rhs - dpf * dpd) + c0;
                    } else {
                        rhs = 0.5 * nu * df - 1.5 * cv;
                        val = dt * (rhs - 0.5 * H[q] - dpf * dpd) + c0;
                    }
                    H[q] = -cv;
                }
                P.out[c][q] = val;
            }
#undef EF
#undef FF
#undef GF
#undef LU
#undef LV
#undef LW
        }
        __syncthreads();
        if (more) {
            store(R, 0, az + 4);
            store(R, 1, az + 4);
            store(R, 2, az + 3);
        }
        __syncthreads();
    }
}

}  // namespace

void launch_predict_tiled(const ns_config &c, int first, bool lazy, const double *u, const double *v, const double *w,
                          const double *dp, const double *gauge, const double *p, double *hx, double *hy, double *hz,
                          double *us, double *vs, double *ws, hipStream_t st) {
    TileParams P;
    P.nx = c.nx; P.ny = c.ny; P.nz = c.nz;
    P.dx = c.dx; P.dy = c.dy; P.dz = c.dz; P.nu = c.nu; P.dt = c.dt;
    const double d[3] = {c.dx, c.dy, c.dz};
    for (int a = 0; a < 3; a++) {
        P.rd[a] = 1.0 / d[a];
        P.rd2[a] = 1.0 / (d[a] * d[a]);
    }
    P.scheme = c.scheme; P.first = first; P.dp_flag = c.dp_flag;
    for (int a = 0; a < 3; a++) P.per[a] = (c.bc[a] == 1);
    P.zc = 16;
    P.tiles_x = (c.nx + TX - 1) / TX;
    P.tiles_y = (c.ny + TY - 1) / TY;
    P.tiles_z = (c.nz + P.zc - 1) / P.zc;
    P.ntiles = P.tiles_x * P.tiles_y * P.tiles_z;
    P.src[0] = u; P.src[1] = v; P.src[2] = w;
    P.dp = dp; P.gauge = gauge;
    P.p = p;
    P.h[0] = hx; P.h[1] = hy; P.h[2] = hz;
    P.out[0] = us; P.out[1] = vs; P.out[2] = ws;
    const int nblk = ((P.ntiles + 7) / 8) * 8;
    if (lazy)
        hipLaunchKernelGGL(k_predict_tiled<true>, dim3(nblk), dim3(TX * TY), 0, st, P);
    else
        hipLaunchKernelGGL(k_predict_tiled<false>, dim3(nblk), dim3(TX * TY), 0, st, P);
}

}  // namespace ns
