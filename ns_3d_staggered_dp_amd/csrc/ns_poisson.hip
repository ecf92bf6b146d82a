// Pressure Poisson solves for gfx950, FP64.
//
// Periodic (big_periodic_3D.f90:215-256,526-548): r2c forward transform of
// the divergence (unscaled), division by
//   eig = -4*((sin(pi*kx/nx)/dx)^2 + (sin(pi*ky/ny)/dy)^2 + (sin(pi*kz/nz)/dz)^2),
// mode (0,0,0) -> 0, c2r backward transform scaled by 1/(nx*ny*nz).
// Neumann/Dirichlet walls (DCT_poisson_solver, NS_functions.f90:1505-1630):
// per-axis DCT-II (pbc=3) or DST-II (pbc=2, sign flip + DCT-II + reversal,
// :1522-1525), division by the eigenvalues of ...interBC_combine.f90:321-351,
// zero mode := mean_dp*nx*ny*nz (:1577), exact inverse transforms.
//
// Structure: five line passes over the volume (x forward fused with the
// divergence and the boundary RHS terms; y forward; z forward + modal
// division + z inverse; y inverse; x inverse writing dp).  Each pass stages
// L lines in LDS and runs a mixed-radix Stockham FFT (radices 2,3,4,5,7,8).
// Real lines are transformed two at a time as one complex line.  The DCT is
// Makhoul's (IEEE TASSP 28(1), 1980): even/odd reorder, complex FFT of length
// n, quarter-wave twiddle.
#include "ns_internal.h"

namespace ns {

namespace {

__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// multiply by sign*i
__device__ __forceinline__ double2 cmuli(double2 a, int sign) {
    return sign < 0 ? make_double2(a.y, -a.x) : make_double2(-a.y, a.x);
}

// small DFTs: X[k] = sum_r v[r] exp(sign*2*pi*i*r*k/R)
template <int R> struct SmallDFT;

template <> struct SmallDFT<2> {
    __device__ static void run(double2 *v, int) {
        double2 a = v[0], b = v[1];
        v[0] = cadd(a, b);
        v[1] = csub(a, b);
    }
};
template <> struct SmallDFT<4> {
    __device__ static void run(double2 *v, int s) {
        double2 t0 = cadd(v[0], v[2]), t1 = csub(v[0], v[2]);
        double2 t2 = cadd(v[1], v[3]), t3 = cmuli(csub(v[1], v[3]), s);
        v[0] = cadd(t0, t2);
        v[2] = csub(t0, t2);
        v[1] = cadd(t1, t3);
        v[3] = csub(t1, t3);
    }
};
template <> struct SmallDFT<8> {
    __device__ static void run(double2 *v, int s) {
        const double h = 0.70710678118654752440;
        double2 e[4] = {v[0], v[2], v[4], v[6]}, o[4] = {v[1], v[3], v[5], v[7]};
        SmallDFT<4>::run(e, s);
        SmallDFT<4>::run(o, s);
        // o[k] *= W8^k, W8 = exp(s*2*pi*i/8)
        double2 w1 = make_double2(h, s * h), w3 = make_double2(-h, s * h);
        o[1] = cmul(o[1], w1);
        o[2] = cmuli(o[2], s);
        o[3] = cmul(o[3], w3);
#pragma unroll
        for (int k = 0; k < 4; k++) {
            v[k] = cadd(e[k], o[k]);
            v[k + 4] = csub(e[k], o[k]);
        }
    }
};
// odd radices: X[k] = v0 + sum_r a_r cos(2 pi r k/R) + s i sum_r b_r sin(2 pi r k/R)
// over the symmetric pairs a_r = v_r + v_{R-r}, b_r = v_r - v_{R-r}, r = 1..(R-1)/2
template <int R> struct OddTab;
template <> struct OddTab<3> {
    static constexpr double c[2] = {1.0, -0.5};
    static constexpr double s[2] = {0.0, 0.86602540378443864676};
};
template <> struct OddTab<5> {
    static constexpr double c[3] = {1.0, 0.30901699437494742410, -0.80901699437494742410};
    static constexpr double s[3] = {0.0, 0.95105651629515357212, 0.58778525229247312917};
};
template <> struct OddTab<7> {
    static constexpr double c[4] = {1.0, 0.62348980185873353053, -0.22252093395631440429, -0.90096886790241912624};
    static constexpr double s[4] = {0.0, 0.78183148246802980871, 0.97492791218182360702, 0.43388373911755812048};
};
template <int R> struct SmallDFTOdd {
    __device__ __forceinline__ static void run(double2 *v, int sg) {
        constexpr int H = (R - 1) / 2;
        double2 a[H + 1], b[H + 1];
        double2 x0 = v[0];
#pragma unroll
        for (int r = 1; r <= H; r++) {
            a[r] = cadd(v[r], v[R - r]);
            b[r] = csub(v[r], v[R - r]);
            x0 = cadd(x0, a[r]);
        }
#pragma unroll
        for (int k = 1; k <= H; k++) {
            double2 cr = v[0], si = make_double2(0.0, 0.0);
#pragma unroll
            for (int r = 1; r <= H; r++) {
                const int m = (r * k) % R;  // cos(2 pi m/R) and sin(2 pi m/R) from the half table
                const double cm = OddTab<R>::c[m <= H ? m : R - m];
                const double sm = (m <= H) ? OddTab<R>::s[m] : -OddTab<R>::s[R - m];
                cr = make_double2(cr.x + a[r].x * cm, cr.y + a[r].y * cm);
                si = make_double2(si.x + b[r].x * sm, si.y + b[r].y * sm);
            }
            // sign*i*si
            const double2 t = cmuli(si, sg);
            v[k] = cadd(cr, t);
            v[R - k] = csub(cr, t);
        }
        v[0] = x0;
    }
};
template <> struct SmallDFT<3> { __device__ static void run(double2 *v, int s) { SmallDFTOdd<3>::run(v, s); } };
template <> struct SmallDFT<5> { __device__ static void run(double2 *v, int s) { SmallDFTOdd<5>::run(v, s); } };
template <> struct SmallDFT<7> { __device__ static void run(double2 *v, int s) { SmallDFTOdd<7>::run(v, s); } };

// LDS layout of L lines of length N: element (line, pos) at line*SL + pos*SP.
// Line-major (SL = padded length, SP = 1) suits lines that are contiguous in
// memory; line-interleaved (SL = 1, SP = L) suits strided lines, whose global
// loads arrive line-fastest.  Threads walk butterflies in the same fastest
// index so LDS accesses of a wavefront are unit-stride in either layout.
struct Lay {
    int SL, SP;
};

// one Stockham stage over L lines of length N
template <int R>
__device__ __forceinline__ void stage(const double2 *__restrict__ A, double2 *__restrict__ B, int N, Lay ly, int L,
                                      int Ns, const double2 *__restrict__ tw, int sign) {
    const int NR = N / R;
    const int nb = L * NR;
    const int tstep = N / (Ns * R);
    const bool line_fast = (ly.SL == 1);
    for (int t = threadIdx.x; t < nb; t += blockDim.x) {
        int line, j;
        if (line_fast) { line = t % L; j = t / L; }
        else { line = t / NR; j = t - line * NR; }
        const double2 *in = A + line * ly.SL;
        double2 v[R];
#pragma unroll
        for (int r = 0; r < R; r++) v[r] = in[(j + r * NR) * ly.SP];
        int jm = j % Ns;
        if (jm) {
#pragma unroll
            for (int r = 1; r < R; r++) {
                double2 w = tw[jm * r * tstep];
                if (sign > 0) w.y = -w.y;
                v[r] = cmul(v[r], w);
            }
        }
        SmallDFT<R>::run(v, sign);
        int idxD = (j - jm) * R + jm;
        double2 *out = B + line * ly.SL;
#pragma unroll
        for (int r = 0; r < R; r++) out[(idxD + r * Ns) * ly.SP] = v[r];
    }
}

// full transform; data starts in A, result pointer returned (A or B)
__device__ double2 *fft_lds(double2 *A, double2 *B, const FFTPlan &pl, Lay ly, int L, int sign) {
    int Ns = 1;
    for (int s = 0; s < pl.nst; s++) {
        switch (pl.radix[s]) {
            case 8: stage<8>(A, B, pl.n, ly, L, Ns, pl.tw, sign); break;
            case 4: stage<4>(A, B, pl.n, ly, L, Ns, pl.tw, sign); break;
            case 2: stage<2>(A, B, pl.n, ly, L, Ns, pl.tw, sign); break;
            case 3: stage<3>(A, B, pl.n, ly, L, Ns, pl.tw, sign); break;
            case 5: stage<5>(A, B, pl.n, ly, L, Ns, pl.tw, sign); break;
            case 7: stage<7>(A, B, pl.n, ly, L, Ns, pl.tw, sign); break;
        }
        __syncthreads();
        Ns *= pl.radix[s];
        double2 *t = A;
        A = B;
        B = t;
    }
    return A;
}

struct DivSrc {
    // divergence source: either rhs (interior array) or the staggered u*, v*, w*
    const double *rhs;
    const double *us, *vs, *ws;
    double dx, dy, dz, dt;
    // boundary RHS terms (pr_bc_staggered_modify_rhs), DCT family only
    const double *planes;
    PlaneLayout L;
    int pbc[3];
};

// divergence RHS at interior cell (a,b,c), 0-based (big_periodic_3D.f90:498-501)
__device__ __forceinline__ double div_at(const Grid &g, const DivSrc &S, int a, int b, int c) {
    double s;
    if (S.rhs) {
        s = S.rhs[(size_t)a + (size_t)g.nx * ((size_t)b + (size_t)g.ny * c)];
    } else {
        const int U1 = g.nx + 1, U2 = g.ny + 2, V1 = g.nx + 2, V2 = g.ny + 1, W1 = g.nx + 2, W2 = g.ny + 2;
        const int A = a + 1, B = b + 1, C = c + 1;
        s = (S.us[FIX(A + 1, B + 1, C + 1, U1, U2)] - S.us[FIX(A, B + 1, C + 1, U1, U2)]) / S.dx;
        s = s + (S.vs[FIX(A + 1, B + 1, C + 1, V1, V2)] - S.vs[FIX(A + 1, B, C + 1, V1, V2)]) / S.dy;
        s = s + (S.ws[FIX(A + 1, B + 1, C + 1, W1, W2)] - S.ws[FIX(A + 1, B + 1, C, W1, W2)]) / S.dz;
        s = s / S.dt;
    }
    if (S.planes) {
        // x, then y, then z terms, in the reference's order (NS_functions.f90:1362-1398)
        const int id[3] = {a, b, c}, n[3] = {g.nx, g.ny, g.nz}, np_[3] = {g.nx + 2, g.ny + 2, g.nz + 2};
        const double d[3] = {S.dx, S.dy, S.dz};
#pragma unroll
        for (int ax = 0; ax < 3; ax++) {
            int pbc = S.pbc[ax];
            if (pbc == 1) continue;
            int o1 = (ax == 0) ? 1 : 0, o2 = (ax == 2) ? 1 : 2;
            size_t pl = (size_t)(id[o1] + 1) + (size_t)np_[o1] * (size_t)(id[o2] + 1);
            double dd = d[ax], d2 = d[ax] * d[ax];
            if (id[ax] == 0) {
                double b1 = S.planes[S.L.off[18 + 2 * ax] + pl];
                if (pbc == 2) s = s - 2 * b1 / d2;
                else if (pbc == 3) s = s + b1 / dd;
                else if (pbc == 4) s = s - b1 / d2;
            }
            if (id[ax] == n[ax] - 1) {
                double bn = S.planes[S.L.off[19 + 2 * ax] + pl];
                if (pbc == 2) s = s - 2 * bn / d2;
                else if (pbc == 3) s = s - bn / dd;
                else if (pbc == 4) s = s - bn / d2;
            }
        }
    }
    return s;
}

struct PassParams {
    Grid g;
    FFTPlan pl;  // plan of this pass's axis
    int L;       // complex lines per block
    int LS;      // LDS line stride (complex)
    int nxh;     // nx/2+1 (periodic)
    int nxhp;    // spectrum row stride: nxh rounded up to a multiple of 8
    DivSrc src;
    double2 *spec;                          // periodic spectrum [z][y][kx]
    double *work;                           // DCT work array [z][y][x]
    const double *ex, *ey, *ez;             // eigenvalue parts
    const double2 *wq;                      // Makhoul twiddles of this axis
    int pbc;                                // this axis' pressure code (DCT)
    const double *mean_dp_dev;              // DCT zero mode source (device) or null
    double mean_dp;                         // host value if mean_dp_dev is null
    double scale;                           // backward scale (periodic)
    double *dp_out;                         // ghosted dp (interior written) or null
    double *sol_out;                        // interior (nx,ny,nz) or null
};

// ---------------- periodic passes ----------------
//
// In-place Stockham: each stage reads its butterflies into registers, waits
// at a barrier, then writes them back, so a block needs one line buffer
// (L*N <= FFT_ELEMS complex) instead of two, and twice as many blocks fit on
// a CU.  L is a power of two and threads walk the lines fastest; line-major
// buffers (x passes) use a row stride of 2 (mod 16) complex so that the L
// lines of one butterfly index fall in distinct LDS banks.  Twiddles are
// staged in LDS.

constexpr int FFT_NT = 256;      // threads per block
constexpr int FFT_ELEMS = 2048;  // complex elements per block (L*N)

// x / d for 0 <= x < 2^24 with r = 1/d in single precision (one correction)
struct IDiv {
    int d;
    float r;
};
__device__ __forceinline__ int idiv(int x, IDiv D) {
    int q = (int)((float)x * D.r);
    const int rem = x - q * D.d;
    if (rem < 0) q--;
    else if (rem >= D.d) q++;
    return q;
}

template <int R>
__device__ __forceinline__ void stage_ip(double2 *A, int N, Lay ly, int lsh, int Ns, int tstep,
                                         const double2 *__restrict__ tw, int sign) {
    constexpr int MB = (FFT_ELEMS / R + FFT_NT - 1) / FFT_NT;
    const int NR = N / R;
    const int nb = NR << lsh;
    const int lm = (1 << lsh) - 1;
    const bool p2 = (Ns & (Ns - 1)) == 0;
    double2 v[MB][R];
#pragma unroll
    for (int m = 0; m < MB; m++) {
        const int t = threadIdx.x + m * FFT_NT;
        if (t < nb) {
            const int line = t & lm, j = t >> lsh;
            const double2 *in = A + line * ly.SL;
#pragma unroll
            for (int r = 0; r < R; r++) v[m][r] = in[(j + r * NR) * ly.SP];
            const int jm = p2 ? (j & (Ns - 1)) : (j % Ns);
            if (jm) {
#pragma unroll
                for (int r = 1; r < R; r++) {
                    double2 w = tw[jm * r * tstep];
                    if (sign > 0) w.y = -w.y;
                    v[m][r] = cmul(v[m][r], w);
                }
            }
            SmallDFT<R>::run(v[m], sign);
        }
    }
    __syncthreads();
#pragma unroll
    for (int m = 0; m < MB; m++) {
        const int t = threadIdx.x + m * FFT_NT;
        if (t < nb) {
            const int line = t & lm, j = t >> lsh;
            const int jm = p2 ? (j & (Ns - 1)) : (j % Ns);
            const int idxD = (j - jm) * R + jm;
            double2 *out = A + line * ly.SL;
#pragma unroll
            for (int r = 0; r < R; r++) out[(idxD + r * Ns) * ly.SP] = v[m][r];
        }
    }
    __syncthreads();
}

// full in-place transform of the 2^lsh lines in A (callers synchronise before).
// ODD: the plan may contain radices 3, 5, 7; power-of-two lengths use the
// variant without them, whose register footprint is a third.
template <bool ODD>
__device__ void fft_ip(double2 *A, const FFTPlan &pl, Lay ly, int lsh, const double2 *tw, int sign) {
    int Ns = 1;
    for (int s = 0; s < pl.nst; s++) {
        const int R = pl.radix[s];
        const int tstep = pl.n / (Ns * R);
        switch (R) {
            case 8: stage_ip<8>(A, pl.n, ly, lsh, Ns, tstep, tw, sign); break;
            case 4: stage_ip<4>(A, pl.n, ly, lsh, Ns, tstep, tw, sign); break;
            case 2: stage_ip<2>(A, pl.n, ly, lsh, Ns, tstep, tw, sign); break;
            case 3: if (ODD) stage_ip<3>(A, pl.n, ly, lsh, Ns, tstep, tw, sign); break;
            case 5: if (ODD) stage_ip<5>(A, pl.n, ly, lsh, Ns, tstep, tw, sign); break;
            case 7: if (ODD) stage_ip<7>(A, pl.n, ly, lsh, Ns, tstep, tw, sign); break;
        }
        Ns *= R;
    }
}

struct FFTPass {
    Grid g;
    FFTPlan pl;   // plan of this pass's axis
    int lsh;      // log2 of the complex lines per block
    int LS;       // LDS row stride (complex) of line-major buffers
    int nxh;      // nx/2+1
    int nxhp;     // spectrum row stride: nxh rounded up to a multiple of 8
    IDiv dN, dNxh, dNy;
    // divergence source: rhs (interior array) or the staggered u*, v*, w*
    const double *rhs;
    const double *us, *vs, *ws;
    double dx, dy, dz, dt, rdx, rdy, rdz, rdt;  // spacings and their reciprocals
    double2 *spec;                // half spectrum [z][y][kx]
    const double *ex, *ey, *ez;   // eigenvalue parts
    double scale;                 // backward scale 1/(nx*ny*nz)
    double *dp_out;               // ghosted dp (interior written) or null
    int dp_ghost_x;               // also write dp's periodic x ghosts
    double *sol_out;              // interior (nx,ny,nz) or null
};

// a/d with r = RN(1/d): one product and one FMA correction step (the
// correctly rounded quotient except in rare last-bit cases)
__device__ __forceinline__ double qdiv(double a, double d, double r) {
    const double q = a * r;
    const double e = __builtin_fma(-q, d, a);
    return __builtin_fma(e, r, q);
}

// divergence RHS at interior cell (a,b,c), 0-based (big_periodic_3D.f90:498-501)
__device__ __forceinline__ double div_rhs(const FFTPass &P, int a, int b, int c) {
    const Grid &g = P.g;
    if (P.rhs) return P.rhs[(size_t)a + (size_t)g.nx * ((size_t)b + (size_t)g.ny * c)];
    const int U1 = g.nx + 1, U2 = g.ny + 2, V1 = g.nx + 2, V2 = g.ny + 1, W1 = g.nx + 2, W2 = g.ny + 2;
    const int A = a + 1, B = b + 1, C = c + 1;
    double s = qdiv(P.us[FIX(A + 1, B + 1, C + 1, U1, U2)] - P.us[FIX(A, B + 1, C + 1, U1, U2)], P.dx, P.rdx);
    s = s + qdiv(P.vs[FIX(A + 1, B + 1, C + 1, V1, V2)] - P.vs[FIX(A + 1, B, C + 1, V1, V2)], P.dy, P.rdy);
    s = s + qdiv(P.ws[FIX(A + 1, B + 1, C + 1, W1, W2)] - P.ws[FIX(A + 1, B + 1, C, W1, W2)], P.dz, P.rdz);
    return qdiv(s, P.dt, P.rdt);
}

template <bool ODD>
__global__ void __launch_bounds__(FFT_NT) k_fft_x_fwd(FFTPass P) {
    extern __shared__ double2 smem[];
    const Grid g = P.g;
    const int N = g.nx, L = 1 << P.lsh, LS = P.LS, nxh = P.nxh;
    double2 *A = smem, *tws = smem + (size_t)L * LS;
    for (int f = threadIdx.x; f < N; f += FFT_NT) tws[f] = P.pl.tw[f];
    const int RL = g.ny * g.nz;
    const int rl0 = blockIdx.x * 2 * L;
    const int cnt = 2 * L * N;
#pragma unroll 4
    for (int f = threadIdx.x; f < cnt; f += FFT_NT) {
        const int rline = idiv(f, P.dN), x = f - rline * N;
        const int rl = rl0 + rline;
        double val = 0.0;
        if (rl < RL) {
            const int k = idiv(rl, P.dNy);
            val = div_rhs(P, x, rl - k * g.ny, k);
        }
        ((double *)&A[(rline >> 1) * LS + x])[rline & 1] = val;
    }
    __syncthreads();
    fft_ip<ODD>(A, P.pl, Lay{LS, 1}, P.lsh, tws, -1);
    // split the paired transform into the two half spectra
    const int cnto = 2 * L * nxh;
    for (int f = threadIdx.x; f < cnto; f += FFT_NT) {
        const int rline = idiv(f, P.dNxh), kx = f - rline * nxh;
        const int rl = rl0 + rline;
        if (rl >= RL) continue;
        const double2 *z = A + (rline >> 1) * LS;
        const double2 a = z[kx], b = z[kx == 0 ? 0 : N - kx];
        double2 X;
        if ((rline & 1) == 0) X = make_double2((a.x + b.x) * 0.5, (a.y - b.y) * 0.5);
        else X = make_double2((a.y + b.y) * 0.5, (b.x - a.x) * 0.5);
        P.spec[(size_t)rl * P.nxhp + kx] = X;
    }
}

// strided complex lines: along y (stride nxhp) for fixed (kx,z), or along z
// (stride nxhp*ny) for fixed (kx,y); L consecutive kx per block, held
// line-interleaved in LDS (element (l,pos) at pos*L + l).  Spectrum rows are
// padded to nxhp (a multiple of 8 complex = 128 B), so a group of 8
// wavenumbers is one aligned cache line.
template <int AXIS>
__device__ __forceinline__ size_t spec_index(const Grid &g, int nxhp, int kx, int pos, int other) {
    if (AXIS == 1) return ((size_t)other * g.ny + pos) * nxhp + kx;
    return ((size_t)pos * g.ny + other) * nxhp + kx;
}

template <int AXIS, int SIGN, int SOLVE, bool ODD>
__global__ void __launch_bounds__(FFT_NT) k_fft_yz(FFTPass P) {
    extern __shared__ double2 smem[];
    const Grid g = P.g;
    const int N = (AXIS == 1) ? g.ny : g.nz, lsh = P.lsh, L = 1 << lsh, nxh = P.nxh, nxhp = P.nxhp;
    double2 *A = smem, *tws = smem + (size_t)L * N;
    for (int f = threadIdx.x; f < N; f += FFT_NT) tws[f] = P.pl.tw[f];
    const int ngrp = (nxh + L - 1) >> lsh;
    const int kx0 = (blockIdx.x % ngrp) << lsh;
    const int other = blockIdx.x / ngrp;
    const int cnt = L * N;
#pragma unroll 4
    for (int f = threadIdx.x; f < cnt; f += FFT_NT) {
        const int l = f & (L - 1), pos = f >> lsh;
        const int kx = kx0 + l;
        A[f] = (kx < nxh) ? P.spec[spec_index<AXIS>(g, nxhp, kx, pos, other)] : make_double2(0.0, 0.0);
    }
    __syncthreads();
    const Lay ly{1, L};
    fft_ip<ODD>(A, P.pl, ly, lsh, tws, SIGN);
    if (SOLVE) {
        // modal division (z pass): eig = -4*((ex + ey) + ez); mode (0,0,0) -> 0
        for (int f = threadIdx.x; f < cnt; f += FFT_NT) {
            const int l = f & (L - 1), kz = f >> lsh;
            const int kx = kx0 + l;
            if (kx >= nxh) continue;
            double2 *z = A + f;
            double e = P.ex[kx];
            e = e + P.ey[other];
            e = e + P.ez[kz];
            e = -4.0 * e;
            if (kx == 0 && other == 0 && kz == 0) *z = make_double2(0.0, 0.0);
            else *z = make_double2(z->x / e, z->y / e);
        }
        __syncthreads();
        fft_ip<ODD>(A, P.pl, ly, lsh, tws, +1);
    }
    for (int f = threadIdx.x; f < cnt; f += FFT_NT) {
        const int l = f & (L - 1), pos = f >> lsh;
        const int kx = kx0 + l;
        if (kx < nxh) P.spec[spec_index<AXIS>(g, nxhp, kx, pos, other)] = A[f];
    }
}

template <bool ODD>
__global__ void __launch_bounds__(FFT_NT) k_fft_x_inv(FFTPass P) {
    extern __shared__ double2 smem[];
    const Grid g = P.g;
    const int N = g.nx, L = 1 << P.lsh, LS = P.LS, nxh = P.nxh;
    double2 *A = smem, *tws = smem + (size_t)L * LS;
    for (int f = threadIdx.x; f < N; f += FFT_NT) tws[f] = P.pl.tw[f];
    const int RL = g.ny * g.nz;
    const int rl0 = blockIdx.x * 2 * L;
    // Z[m] = Xa[m] + i Xb[m] of the paired lines, Hermitian extension for m >= nxh;
    // c2r ignores the imaginary parts of DC and Nyquist
    const int cnt = L * N;
#pragma unroll 4
    for (int f = threadIdx.x; f < cnt; f += FFT_NT) {
        const int l = idiv(f, P.dN), m = f - l * N;
        const int ra = rl0 + 2 * l;
        const bool cj = m >= nxh;
        const int km = cj ? N - m : m;
        double2 a = (ra < RL) ? P.spec[(size_t)ra * P.nxhp + km] : make_double2(0.0, 0.0);
        double2 b = (ra + 1 < RL) ? P.spec[(size_t)(ra + 1) * P.nxhp + km] : make_double2(0.0, 0.0);
        if (km == 0 || 2 * km == N) {
            a.y = 0.0;
            b.y = 0.0;
        }
        if (cj) {
            a.y = -a.y;
            b.y = -b.y;
        }
        A[l * LS + m] = make_double2(a.x - b.y, a.y + b.x);
    }
    __syncthreads();
    fft_ip<ODD>(A, P.pl, Lay{LS, 1}, P.lsh, tws, +1);
    const double sc = P.scale;
    const int cnto = 2 * L * N;
    for (int f = threadIdx.x; f < cnto; f += FFT_NT) {
        const int rline = idiv(f, P.dN), x = f - rline * N;
        const int rl = rl0 + rline;
        if (rl >= RL) continue;
        const double2 z = A[(rline >> 1) * LS + x];
        const double val = ((rline & 1) ? z.y : z.x) * sc;
        const int k = idiv(rl, P.dNy), j = rl - k * g.ny;
        if (P.dp_out) {
            const size_t q = FIX(x + 2, j + 2, k + 2, g.nx + 2, g.ny + 2);
            P.dp_out[q] = val;
            // x part of pr_bc_staggered (pbc 1): dp(1) = dp(nx+1), dp(nx+2) = dp(2)
            if (P.dp_ghost_x) {
                if (x == 0) P.dp_out[q + g.nx] = val;
                if (x == g.nx - 1) P.dp_out[q - g.nx] = val;
            }
        }
        if (P.sol_out) P.sol_out[(size_t)x + (size_t)g.nx * rl] = val;
    }
}

// ---------------- DCT family passes ----------------
// Makhoul reorder: v[m] = x[src(m)]
__device__ __forceinline__ int mk_src(int m, int N) {
    int h = (N + 1) / 2;
    return (m < h) ? 2 * m : 2 * (N - 1 - m) + 1;
}

// forward DCT-II (pbc 3) or DST-II (pbc 2) of the pair of real lines held as
// Z (after the forward FFT of the reordered data); writes X_a, X_b into out
__device__ __forceinline__ void makhoul_post(const double2 *Z, int N, int k, const double2 *wq, double &xa, double &xb) {
    double2 a = Z[k], b = Z[(N - k) % N];
    // Va = (Z[k] + conj(Z[N-k]))/2, Vb = -i (Z[k] - conj(Z[N-k]))/2
    double2 va = make_double2((a.x + b.x) * 0.5, (a.y - b.y) * 0.5);
    double2 vb = make_double2((a.y + b.y) * 0.5, (b.x - a.x) * 0.5);
    double c = wq[k].x, s = wq[k].y;  // exp(-i pi k/(2N)) = c - i s
    xa = c * va.x + s * va.y;
    xb = c * vb.x + s * vb.y;
}

// inverse: given X (length N, real) returns V[k] = exp(i pi k/(2N)) (X[k] - i X[N-k])
__device__ __forceinline__ double2 makhoul_pre_inv(double xk, double xnk, int k, const double2 *wq) {
    double c = wq[k].x, s = wq[k].y;
    return make_double2(c * xk + s * xnk, s * xk - c * xnk);
}

// generic real-line gather/scatter for the DCT passes.
// AXIS 0: lines along x, pairs of consecutive flattened (y,z) lines.
// AXIS 1/2: lines along y/z, pairs of consecutive x at fixed z (AXIS 1) or y (AXIS 2).
struct RealLines {
    __device__ static void locate(const Grid &g, int AXIS, int blk, int L, int rline, long &base, long &stride,
                                  bool &valid) {
        if (AXIS == 0) {
            long rl = (long)blk * 2 * L + rline;
            valid = rl < (long)g.ny * g.nz;
            base = rl * g.nx;
            stride = 1;
        } else {
            int nxp = (g.nx + 1) / 2;               // pairs per row
            int ngrp = (nxp + L - 1) / L;
            int other = blk / ngrp;
            int x = (blk % ngrp) * 2 * L + rline;
            valid = x < g.nx;
            if (AXIS == 1) { base = (long)other * g.nx * g.ny + x; stride = g.nx; }
            else { base = (long)other * g.nx + x; stride = (long)g.nx * g.ny; }
        }
    }
};

template <int AXIS>
__device__ __forceinline__ int dct_n(const Grid &g) { return AXIS == 0 ? g.nx : AXIS == 1 ? g.ny : g.nz; }

// forward DCT along AXIS; AXIS 0 reads the divergence (+ boundary RHS terms)
template <int AXIS, int SOLVE>
__device__ __forceinline__ void dct_fwd_body(const PassParams &P, const int blk) {
    extern __shared__ double2 smem[];
    const Grid g = P.g;
    const int N = dct_n<AXIS>(g), L = P.L, LS = P.LS;
    double2 *A = smem, *B = smem + (size_t)L * LS;
    const bool dst = (P.pbc == 2);
    // load 2L real lines, reorder into v, pack pairs as complex
    for (int f = threadIdx.x; f < 2 * L * N; f += blockDim.x) {
        int rline, pos;
        if (AXIS == 0) { rline = f / N; pos = f - rline * N; }
        else { rline = f % (2 * L); pos = f / (2 * L); }
        long base, stride;
        bool valid;
        RealLines::locate(g, AXIS, blk, L, rline, base, stride, valid);
        double val = 0.0;
        if (valid) {
            if (AXIS == 0) {
                long rl = base / g.nx;
                val = div_at(g, P.src, pos, (int)(rl % g.ny), (int)(rl / g.ny));
            } else {
                val = P.work[base + stride * pos];
            }
            if (dst && (pos & 1) == 0) val = -val;
        }
        // x[pos] goes to v[m] with mk_src(m) == pos
        int m = (pos & 1) ? (N - 1 - (pos >> 1)) : (pos >> 1);
        double *slot = (double *)&A[(rline >> 1) * LS + m];
        slot[rline & 1] = val;
    }
    __syncthreads();
    double2 *Z = fft_lds(A, B, P.pl, Lay{LS, 1}, L, -1);
    double2 *O = (Z == A) ? B : A;
    // post-twiddle into real coefficients, stored as (xa, xb) pairs in O
    for (int f = threadIdx.x; f < L * N; f += blockDim.x) {
        int l = f / N, k = f - l * N;
        double xa, xb;
        makhoul_post(Z + l * LS, N, k, P.wq, xa, xb);
        int kk = dst ? (N - 1 - k) : k;  // DST-II reverses the output order
        O[l * LS + kk] = make_double2(xa, xb);
    }
    __syncthreads();
    if (SOLVE) {
        // modal division; the zero mode carries mean_dp (NS_functions.f90:1576-1577)
        for (int f = threadIdx.x; f < L * N; f += blockDim.x) {
            int l = f / N, kz = f - l * N;
            for (int h = 0; h < 2; h++) {
                int rline = 2 * l + h;
                long base, stride;
                bool valid;
                RealLines::locate(g, AXIS, blk, L, rline, base, stride, valid);
                if (!valid) continue;
                int xi = (int)(base % g.nx), yj = (int)((base / g.nx) % g.ny);
                double e = P.ex[xi];
                e = e + P.ey[yj];
                e = e + P.ez[kz];
                e = -4.0 * e;
                double *slot = (double *)&O[l * LS + kz];
                if (e == 0.0) slot[h] = P.mean_dp_dev ? (*P.mean_dp_dev) * g.nx * g.ny * g.nz : P.mean_dp * g.nx * g.ny * g.nz;
                else slot[h] = slot[h] / e;
            }
        }
        __syncthreads();
        // inverse along z: (DST: reverse input) -> V -> inverse FFT -> un-reorder (-> DST sign flip)
        for (int f = threadIdx.x; f < L * N; f += blockDim.x) {
            int l = f / N, k = f - l * N;
            const double2 *X = O + l * LS;
            int k0 = dst ? (N - 1 - k) : k;
            double2 xk = X[k0];
            double2 xnk = make_double2(0.0, 0.0);
            if (k > 0) xnk = X[dst ? (k - 1) : (N - k)];
            double2 va = makhoul_pre_inv(xk.x, xnk.x, k, P.wq);
            double2 vb = makhoul_pre_inv(xk.y, xnk.y, k, P.wq);
            Z[l * LS + k] = make_double2(va.x - vb.y, va.y + vb.x);
        }
        __syncthreads();
        double2 *R = fft_lds(Z, O, P.pl, Lay{LS, 1}, L, +1);
        const double inv = 1.0 / N;
        for (int f = threadIdx.x; f < 2 * L * N; f += blockDim.x) {
            int rline = f % (2 * L), pos = f / (2 * L);
            long base, stride;
            bool valid;
            RealLines::locate(g, AXIS, blk, L, rline, base, stride, valid);
            if (!valid) continue;
            int m = (pos & 1) ? (N - 1 - (pos >> 1)) : (pos >> 1);
            double2 z = R[(rline >> 1) * LS + m];
            double val = ((rline & 1) ? z.y : z.x) * inv;
            if (dst && (pos & 1) == 0) val = -val;
            P.work[base + stride * pos] = val;
        }
        return;
    }
    for (int f = threadIdx.x; f < 2 * L * N; f += blockDim.x) {
        int rline, pos;
        if (AXIS == 0) { rline = f / N; pos = f - rline * N; }
        else { rline = f % (2 * L); pos = f / (2 * L); }
        long base, stride;
        bool valid;
        RealLines::locate(g, AXIS, blk, L, rline, base, stride, valid);
        if (!valid) continue;
        double2 z = O[(rline >> 1) * LS + pos];
        P.work[base + stride * pos] = (rline & 1) ? z.y : z.x;
    }
}

template <int AXIS>
__device__ __forceinline__ void dct_inv_body(const PassParams &P, const int blk) {
    extern __shared__ double2 smem[];
    const Grid g = P.g;
    const int N = dct_n<AXIS>(g), L = P.L, LS = P.LS;
    double2 *A = smem, *B = smem + (size_t)L * LS;
    const bool dst = (P.pbc == 2);
    for (int f = threadIdx.x; f < 2 * L * N; f += blockDim.x) {
        int rline, pos;
        if (AXIS == 0) { rline = f / N; pos = f - rline * N; }
        else { rline = f % (2 * L); pos = f / (2 * L); }
        long base, stride;
        bool valid;
        RealLines::locate(g, AXIS, blk, L, rline, base, stride, valid);
        double val = valid ? P.work[base + stride * pos] : 0.0;
        double *slot = (double *)&B[(rline >> 1) * LS + pos];
        slot[rline & 1] = val;
    }
    __syncthreads();
    for (int f = threadIdx.x; f < L * N; f += blockDim.x) {
        int l = f / N, k = f - l * N;
        const double2 *X = B + l * LS;
        int k0 = dst ? (N - 1 - k) : k;
        double2 xk = X[k0];
        double2 xnk = make_double2(0.0, 0.0);
        if (k > 0) xnk = X[dst ? (k - 1) : (N - k)];
        double2 va = makhoul_pre_inv(xk.x, xnk.x, k, P.wq);
        double2 vb = makhoul_pre_inv(xk.y, xnk.y, k, P.wq);
        A[l * LS + k] = make_double2(va.x - vb.y, va.y + vb.x);
    }
    __syncthreads();
    double2 *R = fft_lds(A, B, P.pl, Lay{LS, 1}, L, +1);
    const double inv = 1.0 / N;
    for (int f = threadIdx.x; f < 2 * L * N; f += blockDim.x) {
        int rline, pos;
        if (AXIS == 0) { rline = f / N; pos = f - rline * N; }
        else { rline = f % (2 * L); pos = f / (2 * L); }
        long base, stride;
        bool valid;
        RealLines::locate(g, AXIS, blk, L, rline, base, stride, valid);
        if (!valid) continue;
        int m = (pos & 1) ? (N - 1 - (pos >> 1)) : (pos >> 1);
        double2 z = R[(rline >> 1) * LS + m];
        double val = ((rline & 1) ? z.y : z.x) * inv;
        if (dst && (pos & 1) == 0) val = -val;
        if (AXIS == 0) {
            long rl = base / g.nx;
            int j = (int)(rl % g.ny), k = (int)(rl / g.ny);
            if (P.dp_out) P.dp_out[FIX(pos + 2, j + 2, k + 2, g.nx + 2, g.ny + 2)] = val;
            if (P.sol_out) P.sol_out[base + pos] = val;
        } else {
            P.work[base + stride * pos] = val;
        }
    }
}

template <int AXIS, int SOLVE>
__global__ void __launch_bounds__(256) k_dct_fwd(PassParams P) {
    dct_fwd_body<AXIS, SOLVE>(P, blockIdx.x);
}
template <int AXIS>
__global__ void __launch_bounds__(256) k_dct_inv(PassParams P) {
    dct_inv_body<AXIS>(P, blockIdx.x);
}
template <int AXIS, int SOLVE>
__global__ void __launch_bounds__(256) k_dct_fwd_batch(const PassParams *__restrict__ Ps, const int *__restrict__ prefix,
                                                       int nbox) {
    const int box = batch_box(prefix, nbox, blockIdx.x);
    dct_fwd_body<AXIS, SOLVE>(Ps[box], blockIdx.x - prefix[box]);
}
template <int AXIS>
__global__ void __launch_bounds__(256) k_dct_inv_batch(const PassParams *__restrict__ Ps, const int *__restrict__ prefix,
                                                       int nbox) {
    const int box = batch_box(prefix, nbox, blockIdx.x);
    dct_inv_body<AXIS>(Ps[box], blockIdx.x - prefix[box]);
}

// lines per block of the DCT passes: two LDS line buffers of ~16 KiB each
// (4 resident blocks per CU)
inline int lines_per_block(int LS) {
    int L = 1024 / LS;
    return L < 1 ? 1 : L;
}
inline size_t lds_bytes(int L, int LS) { return (size_t)2 * L * LS * sizeof(double2); }

}  // namespace

int spectrum_row(int nx) { return ((nx / 2 + 1) + 7) / 8 * 8; }

namespace {

// launch with dynamic LDS, raising the kernel's limit above 64 KiB when a
// long line needs it
template <typename K>
void launch_lds(K kernel, unsigned nb, size_t lds, hipStream_t st, const FFTPass &P) {
    if (lds > 65536) (void)hipFuncSetAttribute((const void *)kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    hipLaunchKernelGGL(kernel, dim3(nb), dim3(FFT_NT), lds, st, P);
}

// log2 of the lines per block: the largest power of two L with L*n <= elems
int fft_lsh(int n, int elems = FFT_ELEMS) {
    int s = 0;
    while ((2 << s) * n <= elems) s++;
    return s;
}
// x passes: contiguous lines need no grouping for coalescing, so they use
// fewer elements per block (more resident blocks per CU; measured at 256^3)
constexpr int FFT_XF_ELEMS = FFT_ELEMS / 4, FFT_XI_ELEMS = FFT_ELEMS / 2;
// line-major row stride: n rounded up to 2 (mod 16) complex
int fft_row(int n) { return n + ((2 - n) % 16 + 16) % 16; }

IDiv make_idiv(int d) { return IDiv{d, 1.0f / (float)d}; }

}  // namespace

void poisson_fft_solve(const Grid &g, const PoissonPlans &pl, double dx, double dy, double dz, double dt,
                       const double *us, const double *vs, const double *ws, const double *rhs_in, double2 *spec,
                       double *dp_out, double *sol_out, hipStream_t st, hipEvent_t *ev, bool dp_ghost_x) {
    FFTPass P{};
    P.g = g;
    P.nxh = g.nx / 2 + 1;
    P.nxhp = spectrum_row(g.nx);
    P.dN = make_idiv(g.nx);
    P.dNxh = make_idiv(P.nxh);
    P.dNy = make_idiv(g.ny);
    P.rhs = rhs_in;
    P.us = us; P.vs = vs; P.ws = ws;
    P.dx = dx; P.dy = dy; P.dz = dz; P.dt = dt;
    P.rdx = 1.0 / dx; P.rdy = 1.0 / dy; P.rdz = 1.0 / dz; P.rdt = 1.0 / dt;
    P.spec = spec;
    P.ex = pl.ex; P.ey = pl.ey; P.ez = pl.ez;
    P.scale = 1.0 / ((double)g.nx * g.ny * g.nz);
    P.dp_out = dp_out;
    P.dp_ghost_x = dp_ghost_x ? 1 : 0;
    P.sol_out = sol_out;
    const long RL = (long)g.ny * g.nz;
    auto lds_x = [&]() { return ((size_t)(1 << P.lsh) * P.LS + g.nx) * sizeof(double2); };
    auto lds_yz = [&](int n) { return ((size_t)(1 << P.lsh) * n + n) * sizeof(double2); };
    auto nb_yz = [&](int others) { return (unsigned)((((P.nxh + (1 << P.lsh) - 1) >> P.lsh)) * others); };
    // x forward (fused with the divergence)
    P.pl = pl.px;
    P.lsh = fft_lsh(g.nx, FFT_XF_ELEMS);
    P.LS = fft_row(g.nx);
    const unsigned nbx = (unsigned)((RL + (2 << P.lsh) - 1) / (2 << P.lsh));
    auto pow2 = [](int n) { return (n & (n - 1)) == 0; };
    if (pow2(g.nx)) launch_lds(k_fft_x_fwd<false>, nbx, lds_x(), st, P);
    else launch_lds(k_fft_x_fwd<true>, nbx, lds_x(), st, P);
    if (ev) (void)hipEventRecord(ev[0], st);
    // y forward
    P.pl = pl.py;
    P.lsh = fft_lsh(g.ny);
    if (pow2(g.ny)) launch_lds(k_fft_yz<1, -1, 0, false>, nb_yz(g.nz), lds_yz(g.ny), st, P);
    else launch_lds(k_fft_yz<1, -1, 0, true>, nb_yz(g.nz), lds_yz(g.ny), st, P);
    if (ev) (void)hipEventRecord(ev[1], st);
    // z forward + modal division + z inverse
    P.pl = pl.pz;
    P.lsh = fft_lsh(g.nz);
    if (pow2(g.nz)) launch_lds(k_fft_yz<2, -1, 1, false>, nb_yz(g.ny), lds_yz(g.nz), st, P);
    else launch_lds(k_fft_yz<2, -1, 1, true>, nb_yz(g.ny), lds_yz(g.nz), st, P);
    if (ev) (void)hipEventRecord(ev[2], st);
    // y inverse
    P.pl = pl.py;
    P.lsh = fft_lsh(g.ny);
    if (pow2(g.ny)) launch_lds(k_fft_yz<1, 1, 0, false>, nb_yz(g.nz), lds_yz(g.ny), st, P);
    else launch_lds(k_fft_yz<1, 1, 0, true>, nb_yz(g.nz), lds_yz(g.ny), st, P);
    if (ev) (void)hipEventRecord(ev[3], st);
    // x inverse (c2r) writing dp
    P.pl = pl.px;
    P.lsh = fft_lsh(g.nx, FFT_XI_ELEMS);
    P.LS = fft_row(g.nx);
    const unsigned nbxi = (unsigned)((RL + (2 << P.lsh) - 1) / (2 << P.lsh));
    if (pow2(g.nx)) launch_lds(k_fft_x_inv<false>, nbxi, lds_x(), st, P);
    else launch_lds(k_fft_x_inv<true>, nbxi, lds_x(), st, P);
    if (ev) (void)hipEventRecord(ev[4], st);
}

// parameters of DCT pass `pass` (0 x forward + divergence, 1 y forward, 2 z
// forward + modal solve + z inverse, 3 y inverse, 4 x inverse); returns the
// block count and the dynamic LDS size
int make_dct_pass_params(const Grid &g, const PoissonPlans &pl, double dx, double dy, double dz, double dt,
                         const double *us, const double *vs, const double *ws, const double *rhs_in,
                         const double *planes, const double *mean_dp_dev, double mean_dp_host, double *work,
                         double *dp_out, double *sol_out, int pass, void *out, size_t *lds) {
    PassParams &P = *(PassParams *)out;
    P = PassParams{};
    P.g = g;
    P.src.rhs = rhs_in;
    P.src.us = us; P.src.vs = vs; P.src.ws = ws;
    P.src.dx = dx; P.src.dy = dy; P.src.dz = dz; P.src.dt = dt;
    P.src.planes = planes;
    P.src.L = plane_layout(g.nx, g.ny, g.nz);
    for (int a = 0; a < 3; a++) P.src.pbc[a] = pl.pbc[a];
    P.work = work;
    P.ex = pl.ex; P.ey = pl.ey; P.ez = pl.ez;
    P.mean_dp_dev = mean_dp_dev;
    P.mean_dp = mean_dp_host;
    P.dp_out = dp_out;
    P.sol_out = sol_out;
    const long RL = (long)g.ny * g.nz;
    const int ax = (pass == 0 || pass == 4) ? 0 : (pass == 2) ? 2 : 1;
    const FFTPlan plans[3] = {pl.px, pl.py, pl.pz};
    const double2 *wq[3] = {pl.wq_x, pl.wq_y, pl.wq_z};
    const int n[3] = {g.nx, g.ny, g.nz};
    P.pl = plans[ax];
    P.wq = wq[ax];
    P.pbc = pl.pbc[ax];
    P.LS = n[ax] + 1;
    P.L = lines_per_block(P.LS);
    *lds = lds_bytes(P.L, P.LS);
    if (ax == 0) return (int)((((RL + 1) / 2) + P.L - 1) / P.L);
    const int others = (ax == 1) ? g.nz : g.ny;
    return (int)(((((g.nx + 1) / 2) + P.L - 1) / P.L) * others);
}

size_t pass_params_size() { return sizeof(PassParams); }

void launch_dct_pass(int pass, const void *params, int nb, size_t lds, hipStream_t st) {
    const PassParams &P = *(const PassParams *)params;
    switch (pass) {
        case 0: hipLaunchKernelGGL((k_dct_fwd<0, 0>), dim3(nb), dim3(256), lds, st, P); break;
        case 1: hipLaunchKernelGGL((k_dct_fwd<1, 0>), dim3(nb), dim3(256), lds, st, P); break;
        case 2: hipLaunchKernelGGL((k_dct_fwd<2, 1>), dim3(nb), dim3(256), lds, st, P); break;
        case 3: hipLaunchKernelGGL((k_dct_inv<1>), dim3(nb), dim3(256), lds, st, P); break;
        case 4: hipLaunchKernelGGL((k_dct_inv<0>), dim3(nb), dim3(256), lds, st, P); break;
    }
}

void launch_dct_batch(int pass, const void *params, const int *prefix, int nbox, int nb, size_t lds, hipStream_t st) {
    const PassParams *Ps = (const PassParams *)params;
    switch (pass) {
        case 0: hipLaunchKernelGGL((k_dct_fwd_batch<0, 0>), dim3(nb), dim3(256), lds, st, Ps, prefix, nbox); break;
        case 1: hipLaunchKernelGGL((k_dct_fwd_batch<1, 0>), dim3(nb), dim3(256), lds, st, Ps, prefix, nbox); break;
        case 2: hipLaunchKernelGGL((k_dct_fwd_batch<2, 1>), dim3(nb), dim3(256), lds, st, Ps, prefix, nbox); break;
        case 3: hipLaunchKernelGGL((k_dct_inv_batch<1>), dim3(nb), dim3(256), lds, st, Ps, prefix, nbox); break;
        case 4: hipLaunchKernelGGL((k_dct_inv_batch<0>), dim3(nb), dim3(256), lds, st, Ps, prefix, nbox); break;
    }
}

void poisson_dct_solve(const Grid &g, const PoissonPlans &pl, double dx, double dy, double dz, double dt,
                       const double *us, const double *vs, const double *ws, const double *rhs_in,
                       const double *planes, const double *mean_dp_dev, double mean_dp_host, double *work,
                       double *dp_out, double *sol_out, hipStream_t st, hipEvent_t *ev) {
    for (int pass = 0; pass < 5; pass++) {
        PassParams P;
        size_t lds = 0;
        const int nb = make_dct_pass_params(g, pl, dx, dy, dz, dt, us, vs, ws, rhs_in, planes, mean_dp_dev,
                                            mean_dp_host, work, dp_out, sol_out, pass, &P, &lds);
        launch_dct_pass(pass, &P, nb, lds, st);
        if (ev) (void)hipEventRecord(ev[pass], st);
    }
}

}  // namespace ns
