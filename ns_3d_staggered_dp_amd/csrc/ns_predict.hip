// Fused predictor for gfx950, FP64: cal_conv + cal_diff + Euler/AB2/AB2-CN
// update (+ pressure gradient) for all three velocity components in one pass
// (NS_functions.f90:53-210, big_periodic_3D.f90:343-424,
// re_simulation_sub_tstep_interBC_combine.f90:543-654).
//
// Layout of the work: a block owns a 64 x 4 column of cells and streams
// through z ("2.5-D" blocking).  Each component keeps a ring of four z-slices
// of its halo'd 66 x 6 tile in LDS, so every input value is fetched from
// memory about 1.5 times (the x/y halo, mostly served by L2) instead of once
// per stencil use.  Each z step stores the slices loaded into registers one
// step earlier, issues the loads of the next step (and the AB2 history of
// the next cell), passes one barrier and computes; four slots make that
// single barrier sufficient.  Every load and LDS store is unconditional
// straight-line code (clamped addresses), so the compiler waits for exactly
// the loads a store needs and the next step's loads stay in flight across
// the barrier and the compute.  Only the points of the reference's
// assignment ranges are computed (x_range0 etc., big_periodic_3D.f90:126-146);
// the ghost entries of u* are rewritten by the boundary fill or discarded by
// the ADI sweeps in every scheme, so they are left untouched, and so are the
// ghost entries of the AB2 history.
//
// LAZY: the input velocity is the previous step's u* and dp, and the
// projection u = u* - dt*(dp(i+1)-dp(i))/dx (big_periodic_3D.f90:562-567,
// with the corner gauge) is applied when the slices are stored to LDS, with
// dp taken from a three-slice LDS ring of its own (67 x 7 tile), so dp is read
// from memory once.  This removes one full read/write pass of u, v, w per
// step.  The periodic wraps use dp(n+1) == dp(1) and dp(n+2) == dp(2), which
// the periodic pressure fill makes exact copies; the context only runs lazily
// when every periodic velocity axis has a periodic pressure axis.
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

constexpr int NVS = 4;       // velocity ring slots: one barrier per z step
constexpr int NDS = 3;       // dp ring slots
constexpr int NT = 256;      // threads per block, one per cell of the tile
constexpr int NE = 2;        // load slots per thread and array (every tile has <= 2*NT elements)

// Tile shapes: TXV x (NT / TXV) cells per block; 64 x 4, or 32 x 8 for boxes
// up to 128 wide and where a box's width leaves fewer idle lanes that way (a
// 70-wide box: two 64-wide tiles use 55% of their lanes, three 32-wide ones 73%)
template <int TXV>
struct Shape {
    static constexpr int TX = TXV, TY = NT / TXV;
    static constexpr int HX = TX + 2, HY = TY + 2;
    static constexpr int UW = HX + 1;   // u rows: local x = HX holds u(nx+1) (periodic wrap of the x average)
    static constexpr int VH = HY + 1;   // v tile: local row HY holds v(ny+1) (periodic wrap of the y average)
    static constexpr int DXT = HX + 1, DYT = HY + 1;  // dp tile, covering the lo/hi points of all three components
    static_assert(HY * UW <= NE * NT && VH * HX <= NE * NT && HY * HX <= NE * NT && DYT * DXT <= NE * NT,
                  "tile larger than the per-thread load slots");
};

struct TileParams {
    int nx, ny, nz;
    int pu, pc;          // row pitches of the u-shaped arrays and of v, w, p
    double dx, dy, dz, nu, dt;
    double rd[3], rd2[3];  // 1/dx, 1/dy, 1/dz and 1/dx^2, ... rounded on the host
    int scheme, first, dp_flag;
    int per[3];          // own-axis periodicity of u, v, w (bc == 1)
    int wrap;            // every axis periodic: ghost reads take the interior image
    int tx;              // tile width (64 or 32; the tile is tx x 256/tx cells)
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
__device__ __forceinline__ int cpitch(const TileParams &P, int c) { return c == 0 ? P.pu : P.pc; }

// global 1-based index of interior cell a (0-based) along axis ax for component c
__device__ __forceinline__ int gidx(int c, int ax, int a) { return a + ((ax == c) ? 1 : 2); }

__device__ __forceinline__ int clampi(int x, int lo, int hi) { return x < lo ? lo : (x > hi ? hi : x); }

// In an all-periodic box every ghost entry of u, v, w is a copy of the
// interior value at its index wrapped along every axis (vel_bc_staggered with
// bc 1, NS_functions.f90:213-310): along a component's own axis (n entries)
// entry n is entry 1, along the others entry 1 is n-1 and entry n is 2.
// Reading through this map makes the ghosts' contents irrelevant.
__device__ __forceinline__ int pwrap(int i, int n, bool own) {
    if (own) return (i == n) ? 1 : i;
    return (i == 1) ? n - 1 : ((i == n) ? 2 : i);
}

// one velocity value with the lazy projection applied, read straight from
// memory (used for the periodic z wrap of w, once per column)
template <bool LAZY>
__device__ __forceinline__ double load_val(const TileParams &P, int c, int gi, int gj, int gk) {
    int n[3];
    cdims(P, c, n);
    gi = clampi(gi, 1, n[0]);
    gj = clampi(gj, 1, n[1]);
    gk = clampi(gk, 1, n[2]);
    int vi = gi, vj = gj, vk = gk;
    if (P.wrap) {
        vi = pwrap(vi, n[0], c == 0);
        vj = pwrap(vj, n[1], c == 1);
        vk = pwrap(vk, n[2], c == 2);
    }
    double x = P.src[c][GIX(vi, vj, vk, cpitch(P, c), n[1])];
    if (LAZY) {
        const int P1 = P.pc, P2 = P.ny + 2;
        int b[3] = {gi, gj, gk};
        b[c]++;
        double hi = P.dp[GIX(b[0], b[1], b[2], P1, P2)], lo = P.dp[GIX(gi, gj, gk, P1, P2)];
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

// LDS tiles of one block: rings of z-slices of the halo'd tiles
template <bool LAZY, int TXV>
struct Tiles {
    using S = Shape<TXV>;
    double u[NVS][S::HY][S::UW];
    double v[NVS][S::VH][S::HX];
    double w[NVS][S::HY][S::HX];
    double d[LAZY ? NDS : 1][S::DYT][S::DXT];
    double spare[64];  // target of the slice stores of lanes past a tile's end
};

// one block's column of tiles; b is the block's index within its box
template <bool LAZY, int TXV>
__device__ __forceinline__ void predict_tile(const TileParams P, const int b) {
    using S = Shape<TXV>;
    constexpr int TX = S::TX, TY = S::TY, HX = S::HX, HY = S::HY, UW = S::UW, VH = S::VH, DXT = S::DXT,
                  DYT = S::DYT;
    __shared__ Tiles<LAZY, TXV> T;
    // XCD-aware tile order: consecutive tiles of one XCD are x/y neighbours
    const int nb = P.ntiles;
    const int per_xcd = (nb + 7) / 8;
    const int t = (b % 8) * per_xcd + b / 8;
    if (t >= nb) return;  // only when nb is not a multiple of 8: padding blocks exit at once
    const int tx_ = t % P.tiles_x;
    const int ty_ = (t / P.tiles_x) % P.tiles_y;
    const int tz_ = t / (P.tiles_x * P.tiles_y);
    const int a0 = tx_ * TX, b0 = ty_ * TY, z0 = tz_ * P.zc;
    const int z1 = min(z0 + P.zc, P.nz);
    const int tx = threadIdx.x % TX, ty = threadIdx.x / TX;
    const int P1 = P.nx + 2, P2 = P.ny + 2;
    const size_t pp = (size_t)P.pc * P2;
    // tile origins (global index of local 0) per component; the dp tile starts at (a0, b0)
    const int ox[3] = {a0, a0 + 1, a0 + 1}, oy[3] = {b0 + 1, b0, b0 + 1};
    int nn[3][3];
    for (int c = 0; c < 3; c++) cdims(P, c, nn[c]);

    // Per-thread load descriptors (element e = threadIdx.x + s*NT of each tile).
    // x/y offsets, with the periodic wrap of the own-axis index 0 and clamping,
    // are fixed for the whole column; each slice load is one add and one load.
    // Clamped points only feed outputs outside the computed ranges.
    unsigned int off[3][NE];   // element offset within a z-slice of u, v, w
    short lds[3][NE];          // LDS position within the slice (-1: no element)
    unsigned int offd[NE];     // dp tile
    short ldsd[NE];
    short plo[3][NE], phi[3][NE];  // dp tile positions of the lazy projection (phi < 0: next dp slice, w)
    // Lanes past a tile's end load the tile's last element (the same address
    // as a lane that needs it, so no extra traffic) and store into T.spare:
    // every load and LDS store is then unconditional, straight-line code,
    // which lets the compiler count the outstanding loads exactly instead of
    // draining them all (a branch around a load makes it wait for every
    // earlier load, including the prefetch for two steps ahead).
    for (int s_ = 0; s_ < NE; s_++) {
        const int e0 = threadIdx.x + s_ * NT;
        // u: HY rows of UW; column HX is u(nx+1)
        {
            const bool ok = e0 < HY * UW;
            const int e = ok ? e0 : HY * UW - 1;
            const int la = e % UW, lb = e / UW;
            int gi = (la == HX) ? P.nx + 1 : ox[0] + la, gj = oy[0] + lb;
            if (gi == 0 && P.per[0]) gi = P.nx;
            if (P.wrap) {
                gi = pwrap(gi, nn[0][0], true);
                gj = pwrap(gj, nn[0][1], false);
            }
            gi = clampi(gi, 1, nn[0][0]);
            gj = clampi(gj, 1, nn[0][1]);
            off[0][s_] = (unsigned int)((gi - 1) + P.pu * (gj - 1));
            lds[0][s_] = ok ? (short)(lb * UW + la) : (short)-1;
            // dp(nx+1) == dp(1) and dp(nx+2) == dp(2) after the periodic pressure fill
            const int dx0 = (la == HX) ? 1 : la;
            plo[0][s_] = (short)((lb + 1) * DXT + dx0);
            phi[0][s_] = (short)((lb + 1) * DXT + dx0 + 1);
        }
        // v: VH rows of HX; row HY is v(ny+1)
        {
            const bool ok = e0 < VH * HX;
            const int e = ok ? e0 : VH * HX - 1;
            const int la = e % HX, lb = e / HX;
            int gi = ox[1] + la, gj = (lb == HY) ? P.ny + 1 : oy[1] + lb;
            if (gj == 0 && P.per[1]) gj = P.ny;
            if (P.wrap) {
                gi = pwrap(gi, nn[1][0], false);
                gj = pwrap(gj, nn[1][1], true);
            }
            gi = clampi(gi, 1, nn[1][0]);
            gj = clampi(gj, 1, nn[1][1]);
            off[1][s_] = (unsigned int)((gi - 1) + P.pc * (gj - 1));
            lds[1][s_] = ok ? (short)(lb * HX + la) : (short)-1;
            const int dy0 = (lb == HY) ? 1 : lb;
            plo[1][s_] = (short)(dy0 * DXT + la + 1);
            phi[1][s_] = (short)((dy0 + 1) * DXT + la + 1);
        }
        // w: HY rows of HX
        {
            const bool ok = e0 < HY * HX;
            const int e = ok ? e0 : HY * HX - 1;
            const int la = e % HX, lb = e / HX;
            int gi = ox[2] + la, gj = oy[2] + lb;
            if (P.wrap) {
                gi = pwrap(gi, nn[2][0], false);
                gj = pwrap(gj, nn[2][1], false);
            }
            gi = clampi(gi, 1, nn[2][0]);
            gj = clampi(gj, 1, nn[2][1]);
            off[2][s_] = (unsigned int)((gi - 1) + P.pc * (gj - 1));
            lds[2][s_] = ok ? (short)(lb * HX + la) : (short)-1;
            plo[2][s_] = (short)((lb + 1) * DXT + la + 1);
            phi[2][s_] = plo[2][s_];
        }
        // dp: DYT rows of DXT starting at (a0, b0)
        {
            const bool ok = LAZY && e0 < DYT * DXT;
            const int e = (e0 < DYT * DXT) ? e0 : DYT * DXT - 1;
            const int la = e % DXT, lb = e / DXT;
            int gi = a0 + la, gj = b0 + lb;
            if (gi == 0 && P.per[0]) gi = P.nx;
            if (gj == 0 && P.per[1]) gj = P.ny;
            gi = clampi(gi, 1, P1);
            gj = clampi(gj, 1, P2);
            offd[s_] = (unsigned int)((gi - 1) + P.pc * (gj - 1));
            ldsd[s_] = ok ? (short)e : (short)-1;
        }
    }
    const size_t ps[3] = {(size_t)P.pu * nn[0][1], (size_t)P.pc * nn[1][1], (size_t)P.pc * nn[2][1]};
    const double gauge = (LAZY && P.gauge) ? *P.gauge : 0.0;
    // z index of a velocity slice (0 -> nz for periodic w) and of a dp slice
    auto kvel = [&](int c, int s) {
        if (c == 2 && s == 0 && P.per[2]) s = P.nz;
        if (P.wrap) s = pwrap(s, nn[c][2], c == 2);
        return clampi(s, 1, nn[c][2]);
    };
    auto kdp = [&](int s) {
        if (s == 0 && P.per[2]) s = P.nz;  // dp(nz+1) == dp(1) after the periodic fill
        return clampi(s, 1, P.nz + 2);
    };
    struct VRegs {
        double v[3][NE];
    };
    struct DRegs {
        double d[NE];
    };
    // vel slices: u, v at su, w at sw
    double *const spare = &T.spare[threadIdx.x & 63];
    auto fetch_vel = [&](VRegs &R, int su, int sw) {
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const double *src = P.src[c] + (size_t)(kvel(c, c == 2 ? sw : su) - 1) * ps[c];
#pragma unroll
            for (int s_ = 0; s_ < NE; s_++) R.v[c][s_] = src[off[c][s_]];
        }
    };
    auto fetch_dp = [&](DRegs &D, int s) {
        const double *src = P.dp + (size_t)(kdp(s) - 1) * pp;
#pragma unroll
        for (int s_ = 0; s_ < NE; s_++) D.d[s_] = src[offd[s_]];
    };
    auto store_dp = [&](const DRegs &D, int s) {
        double *dst = &T.d[s % NDS][0][0];
#pragma unroll
        for (int s_ = 0; s_ < NE; s_++) {
            double *q = (ldsd[s_] >= 0) ? dst + ldsd[s_] : spare;
            *q = D.d[s_] - gauge;
        }
    };
    // projection u = u* - dt*(dp(+1) - dp)/d (big_periodic_3D.f90:562-567), with the
    // dp slices (gauge already applied) taken from the LDS ring
    auto store_vel = [&](const VRegs &R, int su, int sw) {
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const int s = (c == 2) ? sw : su;
            double *dst = (c == 0) ? &T.u[s % NVS][0][0] : (c == 1) ? &T.v[s % NVS][0][0] : &T.w[s % NVS][0][0];
            const double d = (c == 0) ? P.dx : (c == 1) ? P.dy : P.dz;
#pragma unroll
            for (int s_ = 0; s_ < NE; s_++) {
                double x = R.v[c][s_];
                if (LAZY) {
                    const double *dl = &T.d[s % NDS][0][0];
                    const double *dh = (c == 2) ? &T.d[(s + 1) % NDS][0][0] : dl;
                    const double gr = dh[phi[c][s_]] - dl[plo[c][s_]];
                    x = x - qdiv(P.dt * gr, d, P.rd[c]);
                }
                double *q = (lds[c][s_] >= 0) ? dst + lds[c][s_] : spare;
                *q = x;
            }
        }
    };

    // output cell of this thread and its first-slice offsets
    const int a = a0 + tx, bb = b0 + ty;
    const bool act = a < P.nx && bb < P.ny;
    size_t q0[3];
    bool on[3];
    // lanes past the box read a valid cell's history (discarded)
    const int ac = min(a, P.nx - 1), bc = min(bb, P.ny - 1);
#pragma unroll
    for (int c = 0; c < 3; c++) {
        const int own = (c == 0) ? a : bb;
        on[c] = act && (c == 2 || P.per[c] || own != 0);
        q0[c] = GIX(gidx(c, 0, ac), gidx(c, 1, bc), gidx(c, 2, 0), cpitch(P, c), nn[c][1]);
    }
    // the history of cells this thread does not compute is loaded and never
    // used (nor is it used by the first / Euler steps), so the load needs no
    // condition: a select on the loaded value would wait for the load at once
    auto fetch_h = [&](double hr[3], int az) {
        const size_t k = (size_t)min(az, P.nz - 1);  // past the column end: a valid cell, not used
#pragma unroll
        for (int c = 0; c < 3; c++) hr[c] = P.h[c][q0[c] + k * ps[c]];
    };

    // prologue: dp slices z0..z0+2, then u, v slices z0+1, z0+2 and w slices z0, z0+1
    // into LDS, then the registers of the first step's store phase
    VRegs R;
    DRegs D;
    if (LAZY) {
        for (int s = z0; s < z0 + 3; s++) {
            fetch_dp(D, s);
            store_dp(D, s);
        }
        __syncthreads();
    }
    for (int q = 0; q < 2; q++) {
        fetch_vel(R, z0 + 1 + q, z0 + q);
        store_vel(R, z0 + 1 + q, z0 + q);
    }
    if (LAZY) {
        fetch_dp(D, z0 + 3);
        __syncthreads();  // dp slot (z0+3) % 3 was read above
        store_dp(D, z0 + 3);
        fetch_dp(D, z0 + 4);
    }
    fetch_vel(R, z0 + 3, z0 + 2);
    double hreg[3];
    fetch_h(hreg, z0);
    if (LAZY) __syncthreads();  // dp slice z0+3 visible to the first store phase

    const double dx = P.dx, dy = P.dy, dz = P.dz, dx2 = dx * dx, dy2 = dy * dy, dz2 = dz * dz;
    const double rx = P.rd[0], ry = P.rd[1], rz = P.rd[2], rx2 = P.rd2[0], ry2 = P.rd2[1], rz2 = P.rd2[2];
    const double dt = P.dt, nu = P.nu, dpf = (double)P.dp_flag;
    // one z step: store the slices loaded one step earlier (u, v at az+3,
    // w at az+2; dp at az+4), load the next step's into the same registers
    // (they overlap the barrier and the compute; past the end of the chunk
    // they read clamped, valid addresses and are not used) and compute
    auto zstep = [&](const int az) {
        if (LAZY) store_dp(D, az + 4);
        store_vel(R, az + 3, az + 2);
        fetch_vel(R, az + 4, az + 3);
        if (LAZY) fetch_dp(D, az + 5);
        __syncthreads();
        if (act) {
#define LU(gi, gj, gk) T.u[(gk) % NVS][(gj) - oy[0]][(gi) - ox[0]]
#define LV(gi, gj, gk) T.v[(gk) % NVS][(gj) - oy[1]][(gi) - ox[1]]
#define LW(gi, gj, gk) T.w[(gk) % NVS][(gj) - oy[2]][(gi) - ox[2]]
#define EF(i, j, k) (((LU(i, (j) + 1, k) + LU(i, j, k)) / 2.0) * ((LV((i) + 1, j, k) + LV(i, j, k)) / 2.0))
#define FF(i, j, k) (((LU(i, j, (k) + 1) + LU(i, j, k)) / 2.0) * ((LW((i) + 1, j, k) + LW(i, j, k)) / 2.0))
#define GF(i, j, k) (((LV(i, j, (k) + 1) + LV(i, j, k)) / 2.0) * ((LW(i, (j) + 1, k) + LW(i, j, k)) / 2.0))
#pragma unroll
            for (int c = 0; c < 3; c++) {
                // own-axis in-range: periodic 0..n-1, walls 1..n-1 (faces 2..n)
                if (!on[c] || (c == 2 && !P.per[2] && az == 0)) continue;
                const int gi = gidx(c, 0, a), gj = gidx(c, 1, bb), gk = gidx(c, 2, az);
                double cv, df, c0;
                if (c == 0) {
                    const int fi = gi, J = gj, K = gk;
                    c0 = LU(fi, J, K);
                    const int fl = fi - 1;  // tile local index 0 holds the periodic wrap when fi == 1
                    double ur = (LU(fi + 1, J, K) + c0) / 2.0;
                    // avg(bc=1) wraps with u(nx+1) and u(nx) (FD_functions.f90:252)
                    double ul = (P.per[0] && fi == 1) ? (T.u[K % NVS][J - oy[0]][HX] + LU(fl, J, K)) / 2.0
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
                    double vl = (P.per[1] && fj == 1) ? (T.v[K % NVS][HY][I - ox[1]] + LV(I, fl, K)) / 2.0
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
                    double wl = (P.per[2] && fk == 1) ? (load_val<LAZY>(P, 2, I, J, P.nz + 1) + LW(I, J, fl)) / 2.0
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
                const size_t q = q0[c] + (size_t)az * ps[c];
                double dpd = 0.0;
                if (P.dp_flag) {
                    int bq[3] = {gi, gj, gk};
                    bq[c]++;
                    const double dd = (c == 0) ? dx : (c == 1) ? dy : dz;
                    dpd = qdiv(P.p[GIX(bq[0], bq[1], bq[2], P.pc, P2)] - P.p[GIX(gi, gj, gk, P.pc, P2)], dd, P.rd[c]);
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
                    else val = dt * (1.5 * rhs - 0.5 * hreg[c] - dpf * dpd) + c0;
                    H[q] = rhs;
                } else {
                    if (P.first) {
                        rhs = 0.5 * nu * df - cv;
                        val = dt * (rhs - dpf * dpd) + c0;
                    } else {
                        rhs = 0.5 * nu * df - 1.5 * cv;
                        val = dt * (rhs - 0.5 * hreg[c] - dpf * dpd) + c0;
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
        // the next cell's AB2 history, loaded after its last use here: it
        // lands during the next step's store phase and conv/diff compute.
        // The empty asm (a compiler-only memory barrier, no instruction)
        // keeps the compiler from sinking the load to its use in the next
        // step, where it would wait for every outstanding load.
        fetch_h(hreg, az + 1);
        asm volatile("" ::: "memory");
    };
    for (int az = z0; az < z1; az++) zstep(az);
}

template <bool LAZY, int TXV>
__global__ void __launch_bounds__(NT, 3) k_predict_tiled(TileParams P) {
    predict_tile<LAZY, TXV>(P, blockIdx.x);
}

// many boxes in one launch: prefix[i] is the first block of box i
template <bool LAZY, int TXV>
__global__ void __launch_bounds__(NT, 3) k_predict_tiled_batch(const TileParams *__restrict__ Ps,
                                                               const int *__restrict__ prefix, int nbox) {
    int blk;
    const TileParams &P = batch_params(Ps, prefix, nbox, blk);
    predict_tile<LAZY, TXV>(P, blk);
}

}  // namespace

namespace {

// blocks of the predictor resident on the whole device at once (cached per device)
int resident_blocks(bool lazy, int tx) {
    static int cache[2][2][64] = {};
    int dev = 0;
    (void)hipGetDevice(&dev);
    int &slots = cache[lazy ? 1 : 0][tx == 32 ? 1 : 0][dev & 63];
    if (!slots) {
        int per_cu = 0, cus = 0;
        if (lazy && tx == 32) (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_predict_tiled<true, 32>, NT, 0);
        else if (lazy) (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_predict_tiled<true, 64>, NT, 0);
        else if (tx == 32) (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_predict_tiled<false, 32>, NT, 0);
        else (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_predict_tiled<false, 64>, NT, 0);
        (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
        slots = (per_cu > 0 ? per_cu : 1) * (cus > 0 ? cus : 1);
    }
    return slots;
}

// z cells per block: the column split into tz chunks so that the tiles fill
// whole rounds of resident blocks (a partial last round leaves most of the
// device idle); cost = rounds x (chunk + the ~3 slices of the prologue)
int choose_zc(int nz, long txy, int slots) {
    int best_tz = 1;
    long best = -1;
    for (int tz = 1; tz <= nz; tz++) {
        const int zc = (nz + tz - 1) / tz;
        if (zc < 4 && tz > 1) break;
        const long nt = txy * tz;
        const long cost = ((nt + slots - 1) / slots) * (zc + 3);
        if (best < 0 || cost < best) {
            best = cost;
            best_tz = tz;
        }
    }
    return (nz + best_tz - 1) / best_tz;
}

}  // namespace

size_t predict_params_size() { return sizeof(TileParams); }

int predict_shape(const ns_config &c, int tile) {
    if (tile == 64 || tile == 32) return tile;
    // boxes up to 128 wide: 32 x 8 (a smaller halo, 34 x 10 around 256 cells
    // against 66 x 6, and fewer idle lanes on such widths); wider ones the
    // shape with fewer idle lanes, 64 x 4 on ties (its longer rows stream
    // better at 256^3: 0.421 against 0.432 ms)
    const long a64 = (long)((c.nx + 63) / 64) * 64 * ((c.ny + 3) / 4) * 4;
    const long a32 = (long)((c.nx + 31) / 32) * 32 * ((c.ny + 7) / 8) * 8;
    return (c.nx <= 128 || a32 < a64) ? 32 : 64;
}

long predict_tile_columns(const ns_config &c, int tile) {
    const int tx = predict_shape(c, tile), ty = NT / tx;
    return (long)((c.nx + tx - 1) / tx) * ((c.ny + ty - 1) / ty) * c.nz;
}

int predict_resident_blocks(bool lazy, int tile) { return resident_blocks(lazy, tile == 32 ? 32 : 64); }

int make_predict_params(const ns_config &c, const Grid &g, int first, bool lazy, const double *u, const double *v, const double *w,
                        const double *dp, const double *gauge, const double *p, double *hx, double *hy, double *hz,
                        double *us, double *vs, double *ws, void *out, int zc_hint, bool wrap, int tile) {
    TileParams &P = *(TileParams *)out;
    P.nx = c.nx; P.ny = c.ny; P.nz = c.nz;
    P.pu = g.pu; P.pc = g.pc;
    P.dx = c.dx; P.dy = c.dy; P.dz = c.dz; P.nu = c.nu; P.dt = c.dt;
    const double d[3] = {c.dx, c.dy, c.dz};
    for (int a = 0; a < 3; a++) {
        P.rd[a] = 1.0 / d[a];
        P.rd2[a] = 1.0 / (d[a] * d[a]);
    }
    P.scheme = c.scheme; P.first = first; P.dp_flag = c.dp_flag;
    for (int a = 0; a < 3; a++) P.per[a] = (c.bc[a] == 1);
    P.wrap = (wrap && P.per[0] && P.per[1] && P.per[2]) ? 1 : 0;
    P.tx = predict_shape(c, tile);
    const int ty = NT / P.tx;
    P.tiles_x = (c.nx + P.tx - 1) / P.tx;
    P.tiles_y = (c.ny + ty - 1) / ty;
    if (zc_hint > 0) {
        // equal chunks of about zc_hint cells
        const int tz = (c.nz + zc_hint - 1) / zc_hint;
        P.zc = (c.nz + tz - 1) / tz;
    } else {
        // a box alone: chunks that fill whole rounds of the resident blocks
        P.zc = choose_zc(c.nz, (long)P.tiles_x * P.tiles_y, resident_blocks(lazy, P.tx));
    }
    P.tiles_z = (c.nz + P.zc - 1) / P.zc;
    P.ntiles = P.tiles_x * P.tiles_y * P.tiles_z;
    P.src[0] = u; P.src[1] = v; P.src[2] = w;
    P.dp = dp; P.gauge = gauge;
    P.p = p;
    P.h[0] = hx; P.h[1] = hy; P.h[2] = hz;
    P.out[0] = us; P.out[1] = vs; P.out[2] = ws;
    return ((P.ntiles + 7) / 8) * 8;
}

void launch_predict_batch(bool lazy, int tile, const void *params, const int *prefix, int nbox, int nblocks,
                          hipStream_t st) {
    const TileParams *Ps = (const TileParams *)params;
    if (lazy && tile == 32)
        hipLaunchKernelGGL((k_predict_tiled_batch<true, 32>), dim3(nblocks), dim3(NT), 0, st, Ps, prefix, nbox);
    else if (lazy)
        hipLaunchKernelGGL((k_predict_tiled_batch<true, 64>), dim3(nblocks), dim3(NT), 0, st, Ps, prefix, nbox);
    else if (tile == 32)
        hipLaunchKernelGGL((k_predict_tiled_batch<false, 32>), dim3(nblocks), dim3(NT), 0, st, Ps, prefix, nbox);
    else
        hipLaunchKernelGGL((k_predict_tiled_batch<false, 64>), dim3(nblocks), dim3(NT), 0, st, Ps, prefix, nbox);
}

void launch_predict_tiled(const ns_config &c, const Grid &g, int first, bool lazy, const double *u, const double *v, const double *w,
                          const double *dp, const double *gauge, const double *p, double *hx, double *hy, double *hz,
                          double *us, double *vs, double *ws, hipStream_t st, bool wrap) {
    TileParams P;
    const int nblk = make_predict_params(c, g, first, lazy, u, v, w, dp, gauge, p, hx, hy, hz, us, vs, ws, &P, 0, wrap);
    if (lazy && P.tx == 32) hipLaunchKernelGGL((k_predict_tiled<true, 32>), dim3(nblk), dim3(NT), 0, st, P);
    else if (lazy) hipLaunchKernelGGL((k_predict_tiled<true, 64>), dim3(nblk), dim3(NT), 0, st, P);
    else if (P.tx == 32) hipLaunchKernelGGL((k_predict_tiled<false, 32>), dim3(nblk), dim3(NT), 0, st, P);
    else hipLaunchKernelGGL((k_predict_tiled<false, 64>), dim3(nblk), dim3(NT), 0, st, P);
}

}  // namespace ns
