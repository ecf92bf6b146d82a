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
// complex product with one fused multiply-add per part (the transforms'
// twiddles; measured 4-5 us less over the 256^3 y/z passes than separate
// products)
__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
    return make_double2(__builtin_fma(a.x, b.x, -(a.y * b.y)), __builtin_fma(a.x, b.y, a.y * b.x));
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
// v *= W16^m = exp(s*2*pi*i*m/16), m a compile-time constant after unrolling
__device__ __forceinline__ double2 w16mul(double2 a, int m, int s) {
    const double c1 = 0.92387953251128675613, s1 = 0.38268343236508977173, h = 0.70710678118654752440;
    switch (m) {
        case 0: return a;
        case 1: return cmul(a, make_double2(c1, s * s1));
        case 2: return cmul(a, make_double2(h, s * h));
        case 3: return cmul(a, make_double2(s1, s * c1));
        case 4: return cmuli(a, s);
        case 6: return cmul(a, make_double2(-h, s * h));
        default: return cmul(a, make_double2(-c1, -s * s1));  // m = 9
    }
}
// 16 = 4 x 4: inner DFTs over r2 (r = r1 + 4*r2), twiddles W16^(r1*k2),
// outer DFTs over r1 give X[k2 + 4*k1]
template <> struct SmallDFT<16> {
    __device__ static void run(double2 *v, int s) {
        double2 t[4][4];
#pragma unroll
        for (int r1 = 0; r1 < 4; r1++) {
#pragma unroll
            for (int r2 = 0; r2 < 4; r2++) t[r1][r2] = v[r1 + 4 * r2];
            SmallDFT<4>::run(t[r1], s);
#pragma unroll
            for (int k2 = 1; k2 < 4; k2++) t[r1][k2] = w16mul(t[r1][k2], r1 * k2, s);
        }
#pragma unroll
        for (int k2 = 0; k2 < 4; k2++) {
            double2 u[4] = {t[0][k2], t[1][k2], t[2][k2], t[3][k2]};
            SmallDFT<4>::run(u, s);
#pragma unroll
            for (int k1 = 0; k1 < 4; k1++) v[k2 + 4 * k1] = u[k1];
        }
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
constexpr int FFT_ELEMS = 2048;  // complex elements per block (L*N), at most
// x passes: contiguous lines need no grouping for coalescing, so they use
// fewer elements per block (more resident blocks per CU; measured at 256^3)
constexpr int FFT_XF_ELEMS = FFT_ELEMS / 4, FFT_XI_ELEMS = FFT_ELEMS / 2;
// DCT passes: 1024 elements per block
constexpr int DCT_ELEMS = 1024;
// Stockham periodic z pass (two transforms per visit; measured at 256^3 before
// the register-resident kernels took the 64-, 256- and 512-long lines)
constexpr int FFT_ZS_ELEMS = 1024;

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

// EL: the block's element cap (L*N <= EL), which sizes the register sets
template <int R, int EL>
__device__ __forceinline__ void stage_ip(double2 *A, int N, Lay ly, int lsh, int Ns, int tstep,
                                         const double2 *__restrict__ tw, int sign) {
    constexpr int MB = (EL / R + FFT_NT - 1) / FFT_NT;
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
template <bool ODD, int EL = FFT_ELEMS>
__device__ __forceinline__ void fft_ip(double2 *A, const FFTPlan &pl, Lay ly, int lsh, const double2 *tw, int sign) {
    int Ns = 1;
    for (int s = 0; s < pl.nst; s++) {
        const int R = pl.radix[s];
        const int tstep = pl.n / (Ns * R);
        switch (R) {
            case 8: stage_ip<8, EL>(A, pl.n, ly, lsh, Ns, tstep, tw, sign); break;
            case 4: stage_ip<4, EL>(A, pl.n, ly, lsh, Ns, tstep, tw, sign); break;
            case 2: stage_ip<2, EL>(A, pl.n, ly, lsh, Ns, tstep, tw, sign); break;
            case 3: if (ODD) stage_ip<3, EL>(A, pl.n, ly, lsh, Ns, tstep, tw, sign); break;
            case 5: if (ODD) stage_ip<5, EL>(A, pl.n, ly, lsh, Ns, tstep, tw, sign); break;
            case 7: if (ODD) stage_ip<7, EL>(A, pl.n, ly, lsh, Ns, tstep, tw, sign); break;
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
    int wrap;                     // all-periodic velocity: read u(nx+1), v(ny+1), w(nz+1) as u(1), v(1), w(1)
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
    const int U1 = g.pu, U2 = g.ny + 2, V1 = g.pc, V2 = g.ny + 1, W1 = g.pc, W2 = g.ny + 2;
    const int A = a + 1, B = b + 1, C = c + 1;
    // the high face of each component's own axis: its periodic copy of face 1
    // when every velocity axis is periodic (the only ghosts read here)
    const int iu = (P.wrap && A == g.nx) ? 1 : A + 1;
    const int jv = (P.wrap && B == g.ny) ? 1 : B + 1;
    const int kw = (P.wrap && C == g.nz) ? 1 : C + 1;
    double s = qdiv(P.us[GIX(iu, B + 1, C + 1, U1, U2)] - P.us[GIX(A, B + 1, C + 1, U1, U2)], P.dx, P.rdx);
    s = s + qdiv(P.vs[GIX(A + 1, jv, C + 1, V1, V2)] - P.vs[GIX(A + 1, B, C + 1, V1, V2)], P.dy, P.rdy);
    s = s + qdiv(P.ws[GIX(A + 1, B + 1, kw, W1, W2)] - P.ws[GIX(A + 1, B + 1, C, W1, W2)], P.dz, P.rdz);
    return qdiv(s, P.dt, P.rdt);
}

template <bool ODD, int EL>
__global__ void __launch_bounds__(FFT_NT) k_fft_x_fwd(FFTPass P) {
    extern __shared__ double2 smem[];
    const Grid g = P.g;
    const int N = g.nx, L = 1 << P.lsh, LS = P.LS, nxh = P.nxh;
    double2 *A = smem, *tws = smem + (size_t)L * LS;
    for (int f = threadIdx.x; f < N; f += FFT_NT) tws[f] = P.pl.tw[f];
    const int RL = g.ny * g.nz;
    const int rl0 = blockIdx.x * 2 * L;
    const int cnt = 2 * L * N;
    {
        // every load of the thread issued before any is used; rows past the
        // grid are clamped to its last row (their values are not used)
        double v[2 * EL / FFT_NT];
#pragma unroll
        for (int u = 0; u < 2 * EL / FFT_NT; u++) {
            const int f = threadIdx.x + u * FFT_NT;
            const int rline = idiv(f, P.dN), x = f - rline * N;
            const int rl = min(rl0 + rline, RL - 1), k = idiv(rl, P.dNy);
            v[u] = div_rhs(P, x, rl - k * g.ny, k);
        }
#pragma unroll
        for (int u = 0; u < 2 * EL / FFT_NT; u++) {
            const int f = threadIdx.x + u * FFT_NT;
            if (f < cnt) {
                const int rline = idiv(f, P.dN), x = f - rline * N;
                ((double *)&A[(rline >> 1) * LS + x])[rline & 1] = (rl0 + rline < RL) ? v[u] : 0.0;
            }
        }
    }
    __syncthreads();
    fft_ip<ODD, EL>(A, P.pl, Lay{LS, 1}, P.lsh, tws, -1);
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

// x forward with the divergence for rows that are a multiple of the block
// size long (nx = 256, 512) in groups of 2L rows of one z plane.  Each u and
// v value the group needs is loaded once, w for both of its planes: u(i+1)
// is read from the u rows staged in LDS, v(j+1) is the same thread's value
// of the next row (N is a multiple of the block size, so a thread's elements
// of consecutive rows share their x), and the group's last row takes v of
// the row after it.  (The generic kernel above loads six values per cell.)
template <bool ODD, int EL, int RPR>
__global__ void __launch_bounds__(FFT_NT) k_fft_x_fwd_rs(FFTPass P) {
    extern __shared__ double2 smem[];
    const Grid g = P.g;
    const int N = g.nx, L = 1 << P.lsh, LS = P.LS, nxh = P.nxh, R2 = 2 * L;
    double2 *A = smem, *tws = smem + (size_t)L * LS;
    double *Us = (double *)(tws + N);  // R2 rows of N+1 (u faces 1..nx+1)
    for (int f = threadIdx.x; f < N; f += FFT_NT) tws[f] = P.pl.tw[f];
    constexpr int PER = 2 * EL / FFT_NT;
    // RPR = N / FFT_NT elements of one row per thread; PER = R2 * RPR
    const int rl0 = blockIdx.x * R2;
    const int c = rl0 / g.ny, y0 = rl0 - c * g.ny;
    const size_t U1 = g.pu, U2 = g.ny + 2, V1 = g.pc, V2 = g.ny + 1, W1 = g.pc, W2 = g.ny + 2;
    const unsigned u1 = g.pu, v1 = g.pc;
    // high faces of the group: w's next plane and v's next row, face 1 again
    // at the end of the box when every velocity axis is periodic
    const size_t kw = (P.wrap && c == g.nz - 1) ? 0 : c + 1;
    const size_t yh = (P.wrap && y0 + R2 == g.ny) ? 0 : y0 + R2;
    const double *ub = P.us + U1 * ((y0 + 1) + U2 * (size_t)(c + 1));
    const double *vb = P.vs + 1 + V1 * (y0 + V2 * (size_t)(c + 1));
    const double *vhb = P.vs + 1 + V1 * (yh + V2 * (size_t)(c + 1));
    const double *wb0 = P.ws + 1 + W1 * ((y0 + 1) + W2 * (size_t)c);
    const double *wb1 = P.ws + 1 + W1 * ((y0 + 1) + W2 * kw);
    double cu[PER], cv[PER], w0[PER], w1[PER], vh[PER];
    // element u of the thread: row u / RPR, x threadIdx.x + (u % RPR) * FFT_NT
#pragma unroll
    for (int u = 0; u < PER; u++) {
        const int r = u / RPR, a = threadIdx.x + (u % RPR) * FFT_NT;
        cu[u] = ub[a + u1 * r];
        cv[u] = vb[a + v1 * r];
        w0[u] = wb0[a + v1 * r];
        w1[u] = wb1[a + v1 * r];
        if (u < RPR) vh[u] = vhb[a];
    }
#pragma unroll
    for (int u = 0; u < PER; u++) {
        const int r = u / RPR, a = threadIdx.x + (u % RPR) * FFT_NT;
        Us[r * (N + 1) + a] = cu[u];
    }
    // the rows' high ghost faces u(nx+1): one load of the first R2 lanes
    if ((int)threadIdx.x < R2) Us[threadIdx.x * (N + 1) + N] = ub[N + u1 * threadIdx.x];
    __syncthreads();
    // the divergence as in div_rhs: ((du/dx + dv/dy) + dw/dz) / dt
#pragma unroll
    for (int u = 0; u < PER; u++) {
        const int r = u / RPR, a = threadIdx.x + (u % RPR) * FFT_NT;
        const double du = Us[r * (N + 1) + ((P.wrap && a == N - 1) ? 0 : a + 1)] - cu[u];
        const double dv = ((u + RPR < PER) ? cv[(u + RPR) % PER] : vh[u % RPR]) - cv[u];
        double s = qdiv(du, P.dx, P.rdx);
        s = s + qdiv(dv, P.dy, P.rdy);
        s = s + qdiv(w1[u] - w0[u], P.dz, P.rdz);
        ((double *)&A[(r >> 1) * LS + a])[r & 1] = qdiv(s, P.dt, P.rdt);
    }
    __syncthreads();
    fft_ip<ODD, EL>(A, P.pl, Lay{LS, 1}, P.lsh, tws, -1);
    // split the paired transform into the two half spectra
    const int cnto = R2 * nxh;
    for (int f = threadIdx.x; f < cnto; f += FFT_NT) {
        const int rline = idiv(f, P.dNxh), kx = f - rline * nxh;
        const double2 *z = A + (rline >> 1) * LS;
        const double2 x0 = z[kx], x1 = z[kx == 0 ? 0 : N - kx];
        double2 X;
        if ((rline & 1) == 0) X = make_double2((x0.x + x1.x) * 0.5, (x0.y - x1.y) * 0.5);
        else X = make_double2((x0.y + x1.y) * 0.5, (x1.x - x0.x) * 0.5);
        P.spec[(size_t)(rl0 + rline) * P.nxhp + kx] = X;
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

template <int AXIS, int SIGN, int SOLVE, bool ODD, int EL>
__global__ void __launch_bounds__(FFT_NT) k_fft_yz(FFTPass P) {
    extern __shared__ double2 smem[];
    const Grid g = P.g;
    const int N = (AXIS == 1) ? g.ny : g.nz, lsh = P.lsh, L = 1 << lsh, nxh = P.nxh, nxhp = P.nxhp;
    double2 *A = smem, *tws = smem + (size_t)L * N;
    for (int f = threadIdx.x; f < N; f += FFT_NT) tws[f] = P.pl.tw[f];
    // blocks of fewer than 8 wavenumbers share a 128-byte line of a spectrum
    // row with G-1 others: those run on one XCD (physical blocks 8 apart), so
    // the line is fetched into that XCD's L2 once
    const int G = (L < 8) ? (8 >> lsh) : 1;
    const int ngrp = ((((nxh + L - 1) >> lsh) + G - 1) / G) * G;
    int blk = (int)blockIdx.x;
    if (G > 1) {
        const int sl = blk >> 3, xcd = blk & 7;
        blk = ((sl / G) * 8 + xcd) * G + (sl % G);
    }
    const int other = blk / ngrp;
    if (other >= ((AXIS == 1) ? g.nz : g.ny)) return;  // padding of the grid (uniform per block)
    const int kx0 = (blk - other * ngrp) << lsh;
    const int cnt = L * N;
    {
        // every load of the thread issued before any is used; indices past the
        // block's lines are clamped to valid addresses (their values are not used)
        double2 v[EL / FFT_NT];
#pragma unroll
        for (int u = 0; u < EL / FFT_NT; u++) {
            const int f = threadIdx.x + u * FFT_NT;
            v[u] = P.spec[spec_index<AXIS>(g, nxhp, min(kx0 + (f & (L - 1)), nxh - 1), min(f >> lsh, N - 1), other)];
        }
#pragma unroll
        for (int u = 0; u < EL / FFT_NT; u++) {
            const int f = threadIdx.x + u * FFT_NT;
            if (f < cnt) A[f] = (kx0 + (f & (L - 1)) < nxh) ? v[u] : make_double2(0.0, 0.0);
        }
    }
    __syncthreads();
    const Lay ly{1, L};
    fft_ip<ODD, EL>(A, P.pl, ly, lsh, tws, SIGN);
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
        fft_ip<ODD, EL>(A, P.pl, ly, lsh, tws, +1);
    }
    for (int f = threadIdx.x; f < cnt; f += FFT_NT) {
        const int l = f & (L - 1), pos = f >> lsh;
        const int kx = kx0 + l;
        if (kx < nxh) P.spec[spec_index<AXIS>(g, nxhp, kx, pos, other)] = A[f];
    }
}

// y/z lines of square length N = R*R as a two-step (four-step without the
// transposes) transform held in registers: with n = n1 + R*n2 and
// k = k2 + R*k1,
//   X[k2 + R*k1] = sum_n1 W_R^(n1*k1) * W_N^(n1*k2) * sum_n2 x[n1 + R*n2] W_R^(n2*k2).
// Thread (l, q) of a block of L lines (l fastest, so a wavefront's loads
// cover L consecutive wavenumbers) loads x[q + R*n2] straight from memory
// (q = n1), runs the inner R-point DFT, twiddles, and passes its column
// through LDS once; the thread then holding k2 = q runs the outer DFT and
// owns X[q + R*k1].  The inverse mirrors the steps, so a z visit with the
// modal division (MODE 2) ends with each thread storing exactly the
// positions it loaded, through two LDS exchanges and two barriers (the
// Stockham passes above make a round trip through LDS per radix stage).
// MODE 0: forward; 1: inverse; 2: forward, division, inverse.
template <int AXIS, int MODE, int R, int L>
__global__ void __launch_bounds__(R * L) k_fft_sq(FFTPass P) {
    // LDS: element (k2, n1, line) at k2*RS + n1*L + line; rows padded for
    // L < 8 so the 8 lanes of one 128-byte LDS cycle fall in distinct banks
    constexpr int RS = R * L + (L < 8 ? L : 0);
    __shared__ double2 S[R * RS];
    const Grid g = P.g;
    const int nxh = P.nxh, nxhp = P.nxhp;
    constexpr int G = (L < 8) ? 8 / L : 1;
    const int ngrp = ((((nxh + L - 1) / L) + G - 1) / G) * G;
    int blk = (int)blockIdx.x;
    if (G > 1) {
        // blocks sharing a 128-byte line of a spectrum row run on one XCD
        const int sl = blk >> 3, xcd = blk & 7;
        blk = ((sl / G) * 8 + xcd) * G + (sl % G);
    }
    const int other = blk / ngrp;
    if (other >= ((AXIS == 1) ? g.nz : g.ny)) return;  // padding of the grid (uniform per block)
    const int l = threadIdx.x & (L - 1), q = threadIdx.x / L;
    const int kx = (blk - other * ngrp) * L + l;
    const bool act = kx < nxh;
    const int kxc = act ? kx : nxh - 1;  // idle lanes read a valid address
    // 32-bit element offsets: the spectrum of a box with a 64- or 256-long
    // y or z axis has fewer than 2^31 elements (every axis is at most 2048)
    const unsigned sp = (AXIS == 1) ? (unsigned)nxhp : (unsigned)g.ny * (unsigned)nxhp;  // line stride
    const unsigned step = sp * R;                                                         // n2 stride
    const unsigned b0 = (AXIS == 1) ? (unsigned)other * (unsigned)g.ny * (unsigned)nxhp : (unsigned)other * (unsigned)nxhp;
    const unsigned bl = b0 + (unsigned)q * sp + (unsigned)kxc;  // element (kxc, pos = q)
    const double2 *__restrict__ tw = P.pl.tw;
    // z visit: the eigenvalue parts staged in LDS before the data loads are
    // issued (visible after the first barrier)
    __shared__ double Ez[MODE == 2 ? R * R : 1];
    double exy = 0.0;
    if (MODE == 2) {
        for (int f = threadIdx.x; f < R * R; f += R * L) Ez[f] = P.ez[f];
        exy = P.ex[kxc] + P.ey[other];
    }
    double2 v[R];
    if (MODE != 1) {
#pragma unroll
        for (int n2 = 0; n2 < R; n2++) v[n2] = P.spec[bl + n2 * step];
        SmallDFT<R>::run(v, -1);
#pragma unroll
        for (int k2 = 1; k2 < R; k2++) v[k2] = cmul(v[k2], tw[q * k2]);
#pragma unroll
        for (int k2 = 0; k2 < R; k2++) S[k2 * RS + q * L + l] = v[k2];
        __syncthreads();
#pragma unroll
        for (int n1 = 0; n1 < R; n1++) v[n1] = S[q * RS + n1 * L + l];
        SmallDFT<R>::run(v, -1);
        // v[k1] = X[q + R*k1]
        if (MODE == 0) {
            if (act) {
                // the store addresses are formed here, not kept live from the loads
                unsigned bs = bl;
                asm volatile("" : "+v"(bs));
#pragma unroll
                for (int k1 = 0; k1 < R; k1++) P.spec[bs + k1 * step] = v[k1];
            }
            return;
        }
        // modal division (z pass): eig = -4*((ex + ey) + ez); mode (0,0,0) -> 0
#pragma unroll
        for (int k1 = 0; k1 < R; k1++) {
            const int kz = q + R * k1;
            double e = exy + Ez[kz];
            e = -4.0 * e;
            // branch-free: mode (0,0,0) has e = 0 and is replaced by zero
            const double r = 1.0 / e;
            const bool zero = (kx | other | kz) == 0;
            v[k1] = make_double2(zero ? 0.0 : v[k1].x * r, zero ? 0.0 : v[k1].y * r);
        }
    } else {
#pragma unroll
        for (int k1 = 0; k1 < R; k1++) v[k1] = P.spec[bl + k1 * step];
    }
    // inverse: the thread holds X[q + R*k1]; inner DFT over k1, twiddle
    // conj(W_N^(n1*q)), exchange, outer DFT over k2 gives x[q + R*n2].  The
    // LDS slots written here are the ones this thread read above, so no
    // barrier is needed before the writes.
    SmallDFT<R>::run(v, +1);
#pragma unroll
    for (int n1 = 1; n1 < R; n1++) {
        const double2 w = tw[n1 * q];
        v[n1] = cmul(v[n1], make_double2(w.x, -w.y));
    }
#pragma unroll
    for (int n1 = 0; n1 < R; n1++) S[q * RS + n1 * L + l] = v[n1];
    __syncthreads();
#pragma unroll
    for (int k2 = 0; k2 < R; k2++) v[k2] = S[k2 * RS + q * L + l];
    SmallDFT<R>::run(v, +1);
    if (act) {
        // the store addresses are formed here, not kept live from the loads
        unsigned bs = bl;
        asm volatile("" : "+v"(bs));
#pragma unroll
        for (int n2 = 0; n2 < R; n2++) P.spec[bs + n2 * step] = v[n2];
    }
}

// y/z lines of 512 = 8^3 points as a three-step transform held in
// registers (used for the z visit; the Stockham y passes were faster at
// 512^3): 64 threads per line, thread t holding the 8 points t + 64*c.
// With t = a + 8*m1 (n = a + 8*m1 + 64*m2), the 64-point transforms of the
// 8 decimated sequences x[a + 8*m] are two 8-point steps with twiddles
// W_64^(m1*kk2), then twiddles W_512^(a*kk) and an 8-point step over a
// give X[kk2 + 8*kk1 + 64*k1], the thread t' = kk2 + 8*kk1 holding k1 = 0..7.
// Two LDS exchanges per direction; the inverse mirrors the steps, so the z
// visit (MODE 2) ends with each thread storing the positions it loaded.
// MODE 0: forward; 1: inverse; 2: forward, division, inverse.
template <int AXIS, int MODE, int L>
__global__ void __launch_bounds__(64 * L) k_fft_cube(FFTPass P) {
    // element (i, j, k, line) at ((k*8 + j)*8 + i)*L + line
    __shared__ double2 S[512 * L];
    auto E = [](int i, int j, int k, int l) { return ((k * 8 + j) * 8 + i) * L + l; };
    const Grid g = P.g;
    const int nxh = P.nxh, nxhp = P.nxhp;
    constexpr int G = (L < 8) ? 8 / L : 1;
    const int ngrp = ((((nxh + L - 1) / L) + G - 1) / G) * G;
    int blk = (int)blockIdx.x;
    if (G > 1) {
        // blocks sharing a 128-byte line of a spectrum row run on one XCD
        const int sl = blk >> 3, xcd = blk & 7;
        blk = ((sl / G) * 8 + xcd) * G + (sl % G);
    }
    const int other = blk / ngrp;
    if (other >= ((AXIS == 1) ? g.nz : g.ny)) return;  // padding of the grid (uniform per block)
    const int l = threadIdx.x % L, t = threadIdx.x / L;
    const int lo = t & 7, hi = t >> 3;
    const int kx = (blk - other * ngrp) * L + l;
    const bool act = kx < nxh;
    const int kxc = act ? kx : nxh - 1;  // idle lanes read a valid address
    // 32-bit element offsets: the spectrum of a box with a 512-long y or z
    // axis has fewer than 2^31 elements (every axis is at most 2048)
    const unsigned sp = (AXIS == 1) ? (unsigned)nxhp : (unsigned)g.ny * (unsigned)nxhp;
    const unsigned step = sp * 64;
    const unsigned b0 = (AXIS == 1) ? (unsigned)other * (unsigned)g.ny * (unsigned)nxhp : (unsigned)other * (unsigned)nxhp;
    const unsigned bl = b0 + (unsigned)t * sp + (unsigned)kxc;  // element (kxc, position t)
    const double2 *__restrict__ tw = P.pl.tw;
    __shared__ double Ez[MODE == 2 ? 512 : 1];
    double exy = 0.0;
    if (MODE == 2) {
        for (int f = threadIdx.x; f < 512; f += 64 * L) Ez[f] = P.ez[f];
        exy = P.ex[kxc] + P.ey[other];
    }
    double2 v[8];
    if (MODE != 1) {
#pragma unroll
        for (int c = 0; c < 8; c++) v[c] = P.spec[bl + c * step];
        // thread (a, m1) = (lo, hi): DFT over m2, twiddle W_64^(m1*kk2)
        SmallDFT<8>::run(v, -1);
#pragma unroll
        for (int k = 1; k < 8; k++) v[k] = cmul(v[k], tw[8 * hi * k]);
#pragma unroll
        for (int k = 0; k < 8; k++) S[E(lo, hi, k, l)] = v[k];
        __syncthreads();
        // thread (a, kk2) = (lo, hi): DFT over m1, twiddle W_512^(a*(kk2 + 8*kk1))
#pragma unroll
        for (int m = 0; m < 8; m++) v[m] = S[E(lo, m, hi, l)];
        SmallDFT<8>::run(v, -1);
#pragma unroll
        for (int k = 0; k < 8; k++) v[k] = cmul(v[k], tw[lo * (hi + 8 * k)]);
        __syncthreads();
#pragma unroll
        for (int k = 0; k < 8; k++) S[E(lo, hi, k, l)] = v[k];
        __syncthreads();
        // thread (kk2, kk1) = (lo, hi): DFT over a gives X[t + 64*k1]
#pragma unroll
        for (int a = 0; a < 8; a++) v[a] = S[E(a, lo, hi, l)];
        SmallDFT<8>::run(v, -1);
        if (MODE == 0) {
            if (act) {
                // the store addresses are formed here, not kept live from the loads
                unsigned bs = bl;
                asm volatile("" : "+v"(bs));
#pragma unroll
                for (int k = 0; k < 8; k++) P.spec[bs + k * step] = v[k];
            }
            return;
        }
        // modal division (z pass): eig = -4*((ex + ey) + ez); mode (0,0,0) -> 0
#pragma unroll
        for (int k = 0; k < 8; k++) {
            const int kz = t + 64 * k;
            double e = exy + Ez[kz];
            e = -4.0 * e;
            // branch-free: mode (0,0,0) has e = 0 and is replaced by zero
            const double r = 1.0 / e;
            const bool zero = (kx | other | kz) == 0;
            v[k] = make_double2(zero ? 0.0 : v[k].x * r, zero ? 0.0 : v[k].y * r);
        }
    } else {
#pragma unroll
        for (int k = 0; k < 8; k++) v[k] = P.spec[bl + k * step];
    }
    // inverse, thread (kk2, kk1) = (lo, hi) holding X[t + 64*k1]: DFT over
    // k1, conjugate twiddle; the LDS slots written first are the ones this
    // thread read last (MODE 2), so no barrier is needed before them
    SmallDFT<8>::run(v, +1);
#pragma unroll
    for (int a = 1; a < 8; a++) {
        const double2 w = tw[a * (lo + 8 * hi)];
        v[a] = cmul(v[a], make_double2(w.x, -w.y));
    }
#pragma unroll
    for (int a = 0; a < 8; a++) S[E(a, lo, hi, l)] = v[a];
    __syncthreads();
    // thread (a, kk2) = (lo, hi): DFT over kk1, conjugate twiddle W_64^(m1*kk2)
#pragma unroll
    for (int k = 0; k < 8; k++) v[k] = S[E(lo, hi, k, l)];
    SmallDFT<8>::run(v, +1);
#pragma unroll
    for (int m = 1; m < 8; m++) {
        const double2 w = tw[8 * m * hi];
        v[m] = cmul(v[m], make_double2(w.x, -w.y));
    }
    __syncthreads();
#pragma unroll
    for (int m = 0; m < 8; m++) S[E(lo, m, hi, l)] = v[m];
    __syncthreads();
    // thread (a, m1) = (lo, hi): DFT over kk2 gives x[t + 64*m2]
#pragma unroll
    for (int k = 0; k < 8; k++) v[k] = S[E(lo, hi, k, l)];
    SmallDFT<8>::run(v, +1);
    if (act) {
        // the store addresses are formed here, not kept live from the loads
        unsigned bs = bl;
        asm volatile("" : "+v"(bs));
#pragma unroll
        for (int c = 0; c < 8; c++) P.spec[bs + c * step] = v[c];
    }
}

// HALF: a power-of-two N with L*N == EL, so the coefficients 0..N/2-1 of the
// block's lines fill EL/(2*FFT_NT) load slots exactly and the Nyquist
// coefficient of each line is one more load of the first L lanes (the
// generic path clamps and issues EL/FFT_NT slots of loads for L*(N/2+1)
// coefficients)
template <bool ODD, int EL, bool HALF = false>
__global__ void __launch_bounds__(FFT_NT) k_fft_x_inv(FFTPass P) {
    extern __shared__ double2 smem[];
    const Grid g = P.g;
    const int N = g.nx, L = 1 << P.lsh, LS = P.LS, nxh = P.nxh;
    double2 *A = smem, *tws = smem + (size_t)L * LS;
    for (int f = threadIdx.x; f < N; f += FFT_NT) tws[f] = P.pl.tw[f];
    const int RL = g.ny * g.nz;
    const int rl0 = blockIdx.x * 2 * L;
    // Z[m] = Xa[m] + i Xb[m] of the paired lines, Hermitian extension for m >= nxh;
    // c2r ignores the imaginary parts of DC and Nyquist.  Each stored
    // coefficient k is loaded once and gives Z[k] and Z[N-k].
    // HALF: coefficients 0..N/2-1 here (f = l*(N/2) + k), Nyquist below
    const int kn = HALF ? N / 2 : nxh;
    const int cnt = L * kn;
    constexpr int NS = HALF ? EL / (2 * FFT_NT) : EL / FFT_NT;
    // every load of the thread issued before any is used; rows past the grid
    // are clamped to its last row (their values are not used)
    double2 va[NS], vb[NS];
#pragma unroll
    for (int u = 0; u < NS; u++) {
        const int f = min((int)threadIdx.x + u * FFT_NT, cnt - 1);
        const int l = HALF ? f >> __builtin_ctz(kn) : idiv(f, P.dNxh), k = f - l * kn;
        const int ra = rl0 + 2 * l;
        va[u] = P.spec[(size_t)min(ra, RL - 1) * P.nxhp + k];
        vb[u] = P.spec[(size_t)min(ra + 1, RL - 1) * P.nxhp + k];
    }
    if (HALF && (int)threadIdx.x < L) {
        // Nyquist coefficient of line threadIdx.x: imaginary parts ignored
        const int ra = rl0 + 2 * threadIdx.x;
        const double an = (ra < RL) ? P.spec[(size_t)ra * P.nxhp + N / 2].x : 0.0;
        const double bn = (ra + 1 < RL) ? P.spec[(size_t)(ra + 1) * P.nxhp + N / 2].x : 0.0;
        A[threadIdx.x * LS + N / 2] = make_double2(an, bn);
    }
#pragma unroll
    for (int u = 0; u < NS; u++) {
        const int f = threadIdx.x + u * FFT_NT;
        if (f >= cnt) break;
        const int l = HALF ? f >> __builtin_ctz(kn) : idiv(f, P.dNxh), k = f - l * kn;
        const int ra = rl0 + 2 * l;
        double2 a = (ra < RL) ? va[u] : make_double2(0.0, 0.0);
        double2 b = (ra + 1 < RL) ? vb[u] : make_double2(0.0, 0.0);
        if (k == 0 || 2 * k == N) {
            a.y = 0.0;
            b.y = 0.0;
        }
        A[l * LS + k] = make_double2(a.x - b.y, a.y + b.x);
        // the conjugate-symmetric partner N-k (k = 1 .. ceil(N/2)-1)
        if (k > 0 && 2 * k < N) A[l * LS + N - k] = make_double2(a.x + b.y, b.x - a.y);
    }
    __syncthreads();
    fft_ip<ODD, EL>(A, P.pl, Lay{LS, 1}, P.lsh, tws, +1);
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
            const size_t q = GIX(x + 2, j + 2, k + 2, g.pc, g.ny + 2);
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
//
// Same in-place machinery as the periodic passes.  A block holds L = 2^lsh
// complex lines; complex line l is the pair of real lines (2l, 2l+1) as real
// and imaginary parts.  x lines are line-major (row stride LS); y and z lines
// are 2L consecutive x values at one z (y pass) or one y (z pass), held
// line-interleaved, so a row of 2L doubles in memory is a row of 2L doubles
// in LDS.  Makhoul's reorder puts x[pos] at m(pos) = pos/2 (even pos) or
// N-1-pos/2 (odd pos).  DST-II lines (pbc 2) flip the sign of the even
// positions and reverse the coefficient order (NS_functions.f90:1522-1525).

struct PassParams {
    Grid g;
    FFTPlan pl;  // plan of this pass's axis
    int lsh;     // log2 of the complex lines per block
    int LS;      // row stride (complex) of line-major buffers (x passes)
    int ngrp;    // y/z passes: groups of 2L consecutive x per row
    int odd;     // the plan has radices 3, 5 or 7
    int big;     // line longer than DCT_ELEMS: full-size register sets
    int nb, nb8; // blocks of the pass, and that count rounded up to a multiple of 8 (launched)
    IDiv dN, dNy, dGrp;
    // x forward: divergence of u*, v*, w* (or the interior array rhs) ...
    const double *rhs;
    const double *us, *vs, *ws;
    double dx, dy, dz, dt, rdx, rdy, rdz, rdt;
    // ... minus the boundary terms of pr_bc_staggered_modify_rhs (planes
    // 18..23 of the plane set: low and high face of x, y, z), when planes
    // is not null
    const double *planes;
    long boff[6];
    int pbcs[3];
    double *work;                // DCT work array [z][y][x]
    const double *ex, *ey, *ez;  // eigenvalue parts
    const double2 *wq;           // exp(-i pi k/(2N)) of this axis
    int pbc;                     // this axis' pressure code (2 DST-II, 3 DCT-II)
    const double *mean_dp_dev;   // zero-mode source on the device, or null
    double mean_dp;              // zero-mode source from the host
    double *dp_out;              // ghosted dp (interior written) or null
    double *sol_out;             // interior (nx,ny,nz) or null
    double *p_acc;               // ghosted p: p += dp at the interior (lazy dp_flag = 1), or null
};

// divergence RHS at interior cell (a,b,c), 0-based (big_periodic_3D.f90:498-501),
// then the boundary terms of the non-periodic axes in x, y, z order
// (NS_functions.f90:1362-1398)
__device__ __forceinline__ double wall_rhs(const PassParams &P, int a, int b, int c) {
    const Grid &g = P.g;
    double s;
    if (P.rhs) {
        s = P.rhs[(size_t)a + (size_t)g.nx * ((size_t)b + (size_t)g.ny * c)];
    } else {
        const int U1 = g.pu, U2 = g.ny + 2, V1 = g.pc, V2 = g.ny + 1, W1 = g.pc, W2 = g.ny + 2;
        const int A = a + 1, B = b + 1, C = c + 1;
        s = qdiv(P.us[GIX(A + 1, B + 1, C + 1, U1, U2)] - P.us[GIX(A, B + 1, C + 1, U1, U2)], P.dx, P.rdx);
        s = s + qdiv(P.vs[GIX(A + 1, B + 1, C + 1, V1, V2)] - P.vs[GIX(A + 1, B, C + 1, V1, V2)], P.dy, P.rdy);
        s = s + qdiv(P.ws[GIX(A + 1, B + 1, C + 1, W1, W2)] - P.ws[GIX(A + 1, B + 1, C, W1, W2)], P.dz, P.rdz);
        s = qdiv(s, P.dt, P.rdt);
    }
    if (P.planes) {
        const int id[3] = {a, b, c}, n[3] = {g.nx, g.ny, g.nz}, np_[3] = {g.nx + 2, g.ny + 2, g.nz + 2};
        const double d[3] = {P.dx, P.dy, P.dz};
#pragma unroll
        for (int ax = 0; ax < 3; ax++) {
            const int pbc = P.pbcs[ax];
            if (pbc == 1) continue;
            const bool lo = id[ax] == 0, hi = id[ax] == n[ax] - 1;
            if (!lo && !hi) continue;
            const int o1 = (ax == 0) ? 1 : 0, o2 = (ax == 2) ? 1 : 2;
            const size_t pl = (size_t)(id[o1] + 1) + (size_t)np_[o1] * (size_t)(id[o2] + 1);
            const double dd = d[ax], d2 = d[ax] * d[ax];
            if (lo) {
                const double b1 = P.planes[P.boff[2 * ax] + pl];
                if (pbc == 2) s = s - 2 * b1 / d2;
                else if (pbc == 3) s = s + b1 / dd;
                else if (pbc == 4) s = s - b1 / d2;
            }
            if (hi) {
                const double bn = P.planes[P.boff[2 * ax + 1] + pl];
                if (pbc == 2) s = s - 2 * bn / d2;
                else if (pbc == 3) s = s - bn / dd;
                else if (pbc == 4) s = s - bn / d2;
            }
        }
    }
    return s;
}

// Makhoul reorder: x[pos] sits at v[mk_pos(pos)]
__device__ __forceinline__ int mk_pos(int pos, int N) { return (pos & 1) ? (N - 1 - (pos >> 1)) : (pos >> 1); }

// DCT-II coefficient k of both real lines of a pair from their paired
// transform (a = Z[k], b = Z[(N-k) mod N]):
// Va = (a + conj b)/2, Vb = -i (a - conj b)/2, X = Re(exp(-i pi k/(2N)) V)
__device__ __forceinline__ void mk_post(double2 a, double2 b, double2 q, double &xa, double &xb) {
    const double2 va = make_double2((a.x + b.x) * 0.5, (a.y - b.y) * 0.5);
    const double2 vb = make_double2((a.y + b.y) * 0.5, (b.x - a.x) * 0.5);
    xa = q.x * va.x + q.y * va.y;  // q = (c, s): exp(-i pi k/(2N)) = c - i s
    xb = q.x * vb.x + q.y * vb.y;
}

// input of the inverse FFT at k for a pair: Va + i Vb with
// V[k] = exp(i pi k/(2N)) (X[k] - i X[N-k]), X[N] = 0
__device__ __forceinline__ double2 mk_pre(double2 xk, double2 xnk, double2 q) {
    const double2 va = make_double2(q.x * xk.x + q.y * xnk.x, q.y * xk.x - q.x * xnk.x);
    const double2 vb = make_double2(q.x * xk.y + q.y * xnk.y, q.y * xk.y - q.x * xnk.y);
    return make_double2(va.x - vb.y, va.y + vb.x);
}

// in place over the L lines of X (coefficients in DCT order): the inverse
// FFT input at k and N-k from X[k] and X[N-k], one thread per pair (k, N-k)
__device__ __forceinline__ void mk_pre_pairs(double2 *A, Lay ly, int lsh, int N, const double2 *__restrict__ wq) {
    const int cnt = (N / 2 + 1) << lsh, lm = (1 << lsh) - 1;
    for (int t = threadIdx.x; t < cnt; t += FFT_NT) {
        const int l = t & lm, h = t >> lsh, h2 = N - h;
        double2 *X = A + l * ly.SL;
        const double2 xh = X[h * ly.SP];
        const double2 xh2 = (h > 0) ? X[h2 * ly.SP] : make_double2(0.0, 0.0);
        const double2 zh = mk_pre(xh, xh2, wq[h]);
        if (h > 0 && h2 != h) X[h2 * ly.SP] = mk_pre(xh2, xh, wq[h2]);
        X[h * ly.SP] = zh;
    }
}

// y/z pass block: first x of its 2L real lines, row base, stride along the line
struct YZGroup {
    int x0;
    size_t base, stride;
};
template <int AXIS>
__device__ __forceinline__ YZGroup yz_group(const PassParams &P, int blk) {
    const Grid &g = P.g;
    const int other = idiv(blk, P.dGrp), grp = blk - other * P.ngrp;
    YZGroup r;
    r.x0 = grp << (P.lsh + 1);
    if (AXIS == 1) {
        r.base = (size_t)other * g.nx * g.ny;
        r.stride = (size_t)g.nx;
    } else {
        r.base = (size_t)other * g.nx;
        r.stride = (size_t)g.nx * g.ny;
    }
    return r;
}

template <int AXIS>
__device__ __forceinline__ int dct_n(const Grid &g) { return AXIS == 0 ? g.nx : AXIS == 1 ? g.ny : g.nz; }

// a y/z block's 2L real lines of work (element f = threadIdx.x + u*FFT_NT of the
// block: line f & (2L-1), position f >> (lsh+1)), every load of the thread
// issued before any is used; indices past the lines are clamped to valid
// addresses (their values are not used)
template <int EL>
__device__ __forceinline__ void yz_load(const PassParams &P, const YZGroup &G, int lsh, int N,
                                        double (&v)[2 * EL / FFT_NT]) {
    const int rm = (2 << lsh) - 1, xl = P.g.nx - 1;
#pragma unroll
    for (int u = 0; u < 2 * EL / FFT_NT; u++) {
        const int f = threadIdx.x + u * FFT_NT;
        const int pos = min(f >> (lsh + 1), N - 1), x = min(G.x0 + (f & rm), xl);
        v[u] = P.work[G.base + G.stride * pos + x];
    }
}

// forward DCT-II / DST-II along x (fused with the RHS) or y, coefficients to work
template <int AXIS, bool ODD, int EL>
__device__ __forceinline__ void dct_fwd_body(const PassParams &P, const int blk) {
    extern __shared__ double2 smem[];
    const Grid &g = P.g;
    const int N = dct_n<AXIS>(g), lsh = P.lsh, L = 1 << lsh;
    const bool dst = (P.pbc == 2);
    const Lay ly = (AXIS == 0) ? Lay{P.LS, 1} : Lay{1, L};
    double2 *A = smem, *tws = smem + (size_t)L * (AXIS == 0 ? P.LS : N);
    double *Ad = (double *)A;
    for (int f = threadIdx.x; f < N; f += FFT_NT) tws[f] = P.pl.tw[f];
    const int cnt = 2 * L * N;
    if (AXIS == 0) {
        const int RL = g.ny * g.nz, rl0 = blk * 2 * L;
        // every load of the thread issued before any is used; rows past the
        // grid are clamped to its last row (their values are not used)
        double v[2 * EL / FFT_NT];
#pragma unroll
        for (int u = 0; u < 2 * EL / FFT_NT; u++) {
            const int f = threadIdx.x + u * FFT_NT;
            const int rline = idiv(f, P.dN), x = f - rline * N;
            const int rl = min(rl0 + rline, RL - 1), k = idiv(rl, P.dNy);
            v[u] = wall_rhs(P, x, rl - k * g.ny, k);
        }
#pragma unroll
        for (int u = 0; u < 2 * EL / FFT_NT; u++) {
            const int f = threadIdx.x + u * FFT_NT;
            if (f < cnt) {
                const int rline = idiv(f, P.dN), x = f - rline * N;
                double val = (rl0 + rline < RL) ? v[u] : 0.0;
                if (dst && (x & 1) == 0) val = -val;
                Ad[2 * ((rline >> 1) * P.LS + mk_pos(x, N)) + (rline & 1)] = val;
            }
        }
    } else {
        const YZGroup G = yz_group<AXIS>(P, blk);
        double v[2 * EL / FFT_NT];
        yz_load<EL>(P, G, lsh, N, v);
#pragma unroll
        for (int u = 0; u < 2 * EL / FFT_NT; u++) {
            const int f = threadIdx.x + u * FFT_NT;
            if (f < cnt) {
                const int rline = f & (2 * L - 1), pos = f >> (lsh + 1);
                double val = (G.x0 + rline < g.nx) ? v[u] : 0.0;
                if (dst && (pos & 1) == 0) val = -val;
                Ad[2 * L * mk_pos(pos, N) + rline] = val;
            }
        }
    }
    __syncthreads();
    fft_ip<ODD, EL>(A, P.pl, ly, lsh, tws, -1);
    const int cnto = L * N;
    if (AXIS == 0) {
        const int RL = g.ny * g.nz, rl0 = blk * 2 * L;
        for (int f = threadIdx.x; f < cnto; f += FFT_NT) {
            const int l = idiv(f, P.dN), k = f - l * N;
            const int ra = rl0 + 2 * l;
            if (ra >= RL) continue;
            const double2 *Z = A + l * P.LS;
            double xa, xb;
            mk_post(Z[k], Z[k == 0 ? 0 : N - k], P.wq[k], xa, xb);
            const int kk = dst ? (N - 1 - k) : k;
            P.work[(size_t)ra * N + kk] = xa;
            if (ra + 1 < RL) P.work[(size_t)(ra + 1) * N + kk] = xb;
        }
    } else {
        const YZGroup G = yz_group<AXIS>(P, blk);
        for (int f = threadIdx.x; f < cnto; f += FFT_NT) {
            const int l = f & (L - 1), k = f >> lsh;
            const int x = G.x0 + 2 * l;
            if (x >= g.nx) continue;
            double xa, xb;
            mk_post(A[k * L + l], A[(k == 0 ? 0 : N - k) * L + l], P.wq[k], xa, xb);
            const int kk = dst ? (N - 1 - k) : k;
            double *o = P.work + G.base + G.stride * kk + x;
            o[0] = xa;
            if (x + 1 < g.nx) o[1] = xb;
        }
    }
}

// z pass: forward transform, modal division (zero mode := mean_dp*nx*ny*nz,
// NS_functions.f90:1576-1577), inverse transform, back to work
template <bool ODD, int EL>
__device__ __forceinline__ void dct_solve_body(const PassParams &P, const int blk) {
    extern __shared__ double2 smem[];
    const Grid &g = P.g;
    const int N = g.nz, lsh = P.lsh, L = 1 << lsh;
    const bool dst = (P.pbc == 2);
    const Lay ly{1, L};
    double2 *A = smem, *tws = smem + (size_t)L * N;
    double *Ad = (double *)A;
    for (int f = threadIdx.x; f < N; f += FFT_NT) tws[f] = P.pl.tw[f];
    const YZGroup G = yz_group<2>(P, blk);
    const int yj = idiv(blk, P.dGrp);
    const int cnt = 2 * L * N, rm = 2 * L - 1;
    {
        double v[2 * EL / FFT_NT];
        yz_load<EL>(P, G, lsh, N, v);
#pragma unroll
        for (int u = 0; u < 2 * EL / FFT_NT; u++) {
            const int f = threadIdx.x + u * FFT_NT;
            if (f < cnt) {
                const int rline = f & rm, pos = f >> (lsh + 1);
                double val = (G.x0 + rline < g.nx) ? v[u] : 0.0;
                if (dst && (pos & 1) == 0) val = -val;
                Ad[2 * L * mk_pos(pos, N) + rline] = val;
            }
        }
    }
    __syncthreads();
    fft_ip<ODD, EL>(A, P.pl, ly, lsh, tws, -1);
    // per pair (k, N-k): coefficients, division by the eigenvalue at the
    // stored position (reversed for DST-II), inverse pre-twiddle, in place
    const double zero_mode = (P.mean_dp_dev ? *P.mean_dp_dev : P.mean_dp) * g.nx * g.ny * g.nz;
    const int cntp = (N / 2 + 1) << lsh;
    for (int t = threadIdx.x; t < cntp; t += FFT_NT) {
        const int l = t & (L - 1), h = t >> lsh, h2 = N - h;
        const int xa_ = G.x0 + 2 * l, xb_ = xa_ + 1;
        // eigenvalue parts of both real lines (lines past nx hold zeros)
        const double ea = (xa_ < g.nx) ? P.ex[xa_] + P.ey[yj] : 1.0;
        const double eb = (xb_ < g.nx) ? P.ex[xb_] + P.ey[yj] : 1.0;
        const double2 zh = A[h * L + l];
        const double2 zh2 = A[(h == 0 ? 0 : h2) * L + l];
        double2 X[2];
        int kz[2];
        const int np = (h > 0 && h2 != h) ? 2 : 1;
        mk_post(zh, zh2, P.wq[h], X[0].x, X[0].y);
        kz[0] = dst ? (N - 1 - h) : h;
        if (np == 2) {
            mk_post(zh2, zh, P.wq[h2], X[1].x, X[1].y);
            kz[1] = dst ? (N - 1 - h2) : h2;
        }
#pragma unroll
        for (int i = 0; i < 2; i++) {
            if (i >= np) break;
            const double ez = P.ez[kz[i]];
            const double e0 = -4.0 * (ea + ez), e1 = -4.0 * (eb + ez);
            X[i].x = (e0 == 0.0) ? zero_mode : X[i].x / e0;
            X[i].y = (e1 == 0.0) ? zero_mode : X[i].y / e1;
        }
        if (np == 2) {
            A[h2 * L + l] = mk_pre(X[1], X[0], P.wq[h2]);
            A[h * L + l] = mk_pre(X[0], X[1], P.wq[h]);
        } else {
            // h == 0 (X[N] = 0) or h == N/2 (X[N-h] = X[h])
            A[h * L + l] = mk_pre(X[0], h == 0 ? make_double2(0.0, 0.0) : X[0], P.wq[h]);
        }
    }
    __syncthreads();
    fft_ip<ODD, EL>(A, P.pl, ly, lsh, tws, +1);
    const double inv = 1.0 / N;
    for (int f = threadIdx.x; f < cnt; f += FFT_NT) {
        const int rline = f & rm, pos = f >> (lsh + 1);
        const int x = G.x0 + rline;
        if (x >= g.nx) continue;
        double val = Ad[2 * L * mk_pos(pos, N) + rline] * inv;
        if (dst && (pos & 1) == 0) val = -val;
        P.work[G.base + G.stride * pos + x] = val;
    }
}

// inverse DCT-II / DST-II along y (work -> work) or x (work -> dp, sol)
template <int AXIS, bool ODD, int EL>
__device__ __forceinline__ void dct_inv_body(const PassParams &P, const int blk) {
    extern __shared__ double2 smem[];
    const Grid &g = P.g;
    const int N = dct_n<AXIS>(g), lsh = P.lsh, L = 1 << lsh;
    const bool dst = (P.pbc == 2);
    const Lay ly = (AXIS == 0) ? Lay{P.LS, 1} : Lay{1, L};
    double2 *A = smem, *tws = smem + (size_t)L * (AXIS == 0 ? P.LS : N);
    double *Ad = (double *)A;
    for (int f = threadIdx.x; f < N; f += FFT_NT) tws[f] = P.pl.tw[f];
    const int cnt = 2 * L * N, rm = 2 * L - 1;
    // coefficients in DCT order: a DST-II line is stored reversed
    double v[2 * EL / FFT_NT];
    if (AXIS == 0) {
        const int RL = g.ny * g.nz, rl0 = blk * 2 * L;
        // the block's lines are one contiguous span of work
        const size_t s0 = (size_t)rl0 * N, last = (size_t)RL * N - 1;
#pragma unroll
        for (int u = 0; u < 2 * EL / FFT_NT; u++) {
            const size_t q = s0 + threadIdx.x + u * FFT_NT;
            v[u] = P.work[q < last ? q : last];
        }
#pragma unroll
        for (int u = 0; u < 2 * EL / FFT_NT; u++) {
            const int f = threadIdx.x + u * FFT_NT;
            if (f < cnt) {
                const int rline = idiv(f, P.dN), pos = f - rline * N;
                const double val = (rl0 + rline < RL) ? v[u] : 0.0;
                Ad[2 * ((rline >> 1) * P.LS + (dst ? N - 1 - pos : pos)) + (rline & 1)] = val;
            }
        }
    } else {
        const YZGroup G = yz_group<AXIS>(P, blk);
        yz_load<EL>(P, G, lsh, N, v);
#pragma unroll
        for (int u = 0; u < 2 * EL / FFT_NT; u++) {
            const int f = threadIdx.x + u * FFT_NT;
            if (f < cnt) {
                const int rline = f & rm, pos = f >> (lsh + 1);
                const double val = (G.x0 + rline < g.nx) ? v[u] : 0.0;
                Ad[2 * L * (dst ? N - 1 - pos : pos) + rline] = val;
            }
        }
    }
    __syncthreads();
    mk_pre_pairs(A, ly, lsh, N, P.wq);
    __syncthreads();
    fft_ip<ODD, EL>(A, P.pl, ly, lsh, tws, +1);
    const double inv = 1.0 / N;
    if (AXIS == 0) {
        const int RL = g.ny * g.nz, rl0 = blk * 2 * L;
        for (int f = threadIdx.x; f < cnt; f += FFT_NT) {
            const int rline = idiv(f, P.dN), pos = f - rline * N;
            const int rl = rl0 + rline;
            if (rl >= RL) continue;
            double val = Ad[2 * ((rline >> 1) * P.LS + mk_pos(pos, N)) + (rline & 1)] * inv;
            if (dst && (pos & 1) == 0) val = -val;
            const int k = idiv(rl, P.dNy), j = rl - k * g.ny;
            const size_t q = GIX(pos + 2, j + 2, k + 2, g.pc, g.ny + 2);
            if (P.dp_out) P.dp_out[q] = val;
            // p = 1*p + dp, the projection's pressure update, where dp is formed
            if (P.p_acc) P.p_acc[q] = P.p_acc[q] + val;
            if (P.sol_out) P.sol_out[(size_t)rl * N + pos] = val;
        }
    } else {
        const YZGroup G = yz_group<AXIS>(P, blk);
        for (int f = threadIdx.x; f < cnt; f += FFT_NT) {
            const int rline = f & rm, pos = f >> (lsh + 1);
            const int x = G.x0 + rline;
            if (x >= g.nx) continue;
            double val = Ad[2 * L * mk_pos(pos, N) + rline] * inv;
            if (dst && (pos & 1) == 0) val = -val;
            P.work[G.base + G.stride * pos + x] = val;
        }
    }
}

// pass 0 x forward + RHS, 1 y forward, 2 z solve, 3 y inverse, 4 x inverse
template <int PASS, bool ODD, int EL>
__device__ __forceinline__ void dct_pass_body(const PassParams &P, const int b) {
    // each XCD takes a contiguous range of the pass's blocks, so the blocks
    // that share a cache line at their row ends share that XCD's L2
    const int blk = xcd_slab_block(b, P.nb8);
    if (blk >= P.nb) return;  // padding (uniform per block)
    if (PASS == 0) dct_fwd_body<0, ODD, EL>(P, blk);
    else if (PASS == 1) dct_fwd_body<1, ODD, EL>(P, blk);
    else if (PASS == 2) dct_solve_body<ODD, EL>(P, blk);
    else if (PASS == 3) dct_inv_body<1, ODD, EL>(P, blk);
    else dct_inv_body<0, ODD, EL>(P, blk);
}

template <int PASS, bool ODD, int EL>
__global__ void __launch_bounds__(FFT_NT) k_dct(PassParams P) {
    dct_pass_body<PASS, ODD, EL>(P, blockIdx.x);
}
template <int PASS, bool ODD, int EL>
__global__ void __launch_bounds__(FFT_NT) k_dct_batch(const PassParams *__restrict__ Ps, const int *__restrict__ prefix,
                                                      int nbox) {
    int blk;
    const PassParams &P = batch_params(Ps, prefix, nbox, blk);
    dct_pass_body<PASS, ODD, EL>(P, blk);
}

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
// line-major row stride: n rounded up to 2 (mod 16) complex
int fft_row(int n) { return n + ((2 - n) % 16 + 16) % 16; }

IDiv make_idiv(int d) { return IDiv{d, 1.0f / (float)d}; }

// square line lengths with a register-resident two-step kernel: 64 and 256
bool sq_ok(int n) { return n == 64 || n == 256; }
template <int AXIS, int MODE, int R, int L>
void launch_sq_l(int others, hipStream_t st, const FFTPass &P) {
    constexpr int G = (L < 8) ? 8 / L : 1;
    const long ngrp = (((P.nxh + L - 1) / L + G - 1) / G) * G;
    const long nb = ngrp * others, q = 8L * G;
    hipLaunchKernelGGL((k_fft_sq<AXIS, MODE, R, L>), dim3((unsigned)((nb + q - 1) / q * q)), dim3(R * L), 0, st, P);
}
// 8 wavenumbers per block: each row segment one whole 128-byte line (the
// 256^3 z visit measured 55.0 us, against 58.3 us with 4 per block)
template <int AXIS, int MODE>
void launch_sq(int n, int others, hipStream_t st, const FFTPass &P) {
    if (n == 64) launch_sq_l<AXIS, MODE, 8, 8>(others, st, P);
    else if (n == 256) launch_sq_l<AXIS, MODE, 16, 8>(others, st, P);
    else if constexpr (MODE == 2) {
        // 512: the three-step kernel, 8 wavenumbers (512 threads) per block
        const long nb = (long)((P.nxh + 7) / 8) * others;
        hipLaunchKernelGGL((k_fft_cube<AXIS, MODE, 8>), dim3((unsigned)((nb + 7) / 8 * 8)), dim3(512), 0, st, P);
    }
}

}  // namespace

void poisson_fft_solve(const Grid &g, const PoissonPlans &pl, double dx, double dy, double dz, double dt,
                       const double *us, const double *vs, const double *ws, const double *rhs_in, double2 *spec,
                       double *dp_out, double *sol_out, hipStream_t st, hipEvent_t *ev, bool dp_ghost_x,
                       bool wrap_vel) {
    FFTPass P{};
    P.wrap = wrap_vel ? 1 : 0;
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
    // blocks of a y/z pass: wavenumber groups per row (padded to whole 128-byte
    // lines when a group is narrower) times rows, a multiple of 8 lines' blocks
    auto nb_yz = [&](int others) {
        const int L = 1 << P.lsh, G = (L < 8) ? 8 / L : 1;
        const long ngrp = (((P.nxh + L - 1) / L + G - 1) / G) * G;
        const long nb = ngrp * others, q = 8L * G;
        return (unsigned)((nb + q - 1) / q * q);
    };
    // the register-resident y/z kernels index the spectrum with 32-bit element
    // offsets (true for every box whose y or z axis they take; checked here)
    const bool spec32 = (size_t)g.nz * g.ny * P.nxhp < ((size_t)1 << 31);
    // x forward (fused with the divergence)
    P.pl = pl.px;
    // lines longer than the x-pass cap use the full-size register sets
    const bool xf_small = g.nx <= FFT_XF_ELEMS, xi_small = g.nx <= FFT_XI_ELEMS;
    P.lsh = fft_lsh(g.nx, xf_small ? FFT_XF_ELEMS : FFT_ELEMS);
    P.LS = fft_row(g.nx);
    const unsigned nbx = (unsigned)((RL + (2 << P.lsh) - 1) / (2 << P.lsh));
    auto pow2 = [](int n) { return (n & (n - 1)) == 0; };
    if (xf_small && !rhs_in && g.nx % FFT_NT == 0 && g.ny % (2 << P.lsh) == 0 &&
        (2 << P.lsh) * (g.nx / FFT_NT) == 2 * FFT_XF_ELEMS / FFT_NT) {
        // rows of 256 or 512: 2L rows fill the block's elements exactly
        const size_t lds = lds_x() + (size_t)(2 << P.lsh) * (g.nx + 1) * sizeof(double);
        if (g.nx == FFT_NT) launch_lds(k_fft_x_fwd_rs<false, FFT_XF_ELEMS, 1>, nbx, lds, st, P);
        else launch_lds(k_fft_x_fwd_rs<false, FFT_XF_ELEMS, 2>, nbx, lds, st, P);
    } else if (xf_small) {
        if (pow2(g.nx)) launch_lds(k_fft_x_fwd<false, FFT_XF_ELEMS>, nbx, lds_x(), st, P);
        else launch_lds(k_fft_x_fwd<true, FFT_XF_ELEMS>, nbx, lds_x(), st, P);
    } else {
        if (pow2(g.nx)) launch_lds(k_fft_x_fwd<false, FFT_ELEMS>, nbx, lds_x(), st, P);
        else launch_lds(k_fft_x_fwd<true, FFT_ELEMS>, nbx, lds_x(), st, P);
    }
    if (ev) (void)hipEventRecord(ev[0], st);
    // y forward
    P.pl = pl.py;
    P.lsh = fft_lsh(g.ny);
    if (sq_ok(g.ny) && spec32) launch_sq<1, 0>(g.ny, g.nz, st, P);
    else if (pow2(g.ny)) launch_lds(k_fft_yz<1, -1, 0, false, FFT_ELEMS>, nb_yz(g.nz), lds_yz(g.ny), st, P);
    else launch_lds(k_fft_yz<1, -1, 0, true, FFT_ELEMS>, nb_yz(g.nz), lds_yz(g.ny), st, P);
    if (ev) (void)hipEventRecord(ev[1], st);
    // z forward + modal division + z inverse
    P.pl = pl.pz;
    const bool zs_small = g.nz <= FFT_ZS_ELEMS;
    P.lsh = fft_lsh(g.nz, zs_small ? FFT_ZS_ELEMS : FFT_ELEMS);
    // z visits of 512-long lines too (measured at 512^3: 875.6 -> 560.0 us;
    // the y passes of 512 keep the Stockham kernel, 427/423 against 439/470 us)
    if ((sq_ok(g.nz) || g.nz == 512) && spec32) launch_sq<2, 2>(g.nz, g.ny, st, P);
    else if (zs_small) {
        if (pow2(g.nz)) launch_lds(k_fft_yz<2, -1, 1, false, FFT_ZS_ELEMS>, nb_yz(g.ny), lds_yz(g.nz), st, P);
        else launch_lds(k_fft_yz<2, -1, 1, true, FFT_ZS_ELEMS>, nb_yz(g.ny), lds_yz(g.nz), st, P);
    } else {
        if (pow2(g.nz)) launch_lds(k_fft_yz<2, -1, 1, false, FFT_ELEMS>, nb_yz(g.ny), lds_yz(g.nz), st, P);
        else launch_lds(k_fft_yz<2, -1, 1, true, FFT_ELEMS>, nb_yz(g.ny), lds_yz(g.nz), st, P);
    }
    if (ev) (void)hipEventRecord(ev[2], st);
    // y inverse
    P.pl = pl.py;
    P.lsh = fft_lsh(g.ny);
    if (sq_ok(g.ny) && spec32) launch_sq<1, 1>(g.ny, g.nz, st, P);
    else if (pow2(g.ny)) launch_lds(k_fft_yz<1, 1, 0, false, FFT_ELEMS>, nb_yz(g.nz), lds_yz(g.ny), st, P);
    else launch_lds(k_fft_yz<1, 1, 0, true, FFT_ELEMS>, nb_yz(g.nz), lds_yz(g.ny), st, P);
    if (ev) (void)hipEventRecord(ev[3], st);
    // x inverse (c2r) writing dp
    P.pl = pl.px;
    P.lsh = fft_lsh(g.nx, xi_small ? FFT_XI_ELEMS : FFT_ELEMS);
    P.LS = fft_row(g.nx);
    const unsigned nbxi = (unsigned)((RL + (2 << P.lsh) - 1) / (2 << P.lsh));
    if (xi_small) {
        // (the Nyquist loads take one lane per line: at most FFT_NT lines, nx >= 4)
        if (pow2(g.nx) && (g.nx << P.lsh) == FFT_XI_ELEMS && (1 << P.lsh) <= FFT_NT)
            launch_lds(k_fft_x_inv<false, FFT_XI_ELEMS, true>, nbxi, lds_x(), st, P);
        else if (pow2(g.nx)) launch_lds(k_fft_x_inv<false, FFT_XI_ELEMS>, nbxi, lds_x(), st, P);
        else launch_lds(k_fft_x_inv<true, FFT_XI_ELEMS>, nbxi, lds_x(), st, P);
    } else {
        if (pow2(g.nx)) launch_lds(k_fft_x_inv<false, FFT_ELEMS>, nbxi, lds_x(), st, P);
        else launch_lds(k_fft_x_inv<true, FFT_ELEMS>, nbxi, lds_x(), st, P);
    }
    if (ev) (void)hipEventRecord(ev[4], st);
}

namespace {

// log2 of the lines per block of a y/z pass over rows of nx values: the
// largest L = 2^s with L*N <= el that leaves the fewest idle lines past nx,
// with 2L >= 16 (a 128-byte row segment) where that fits
int dct_yz_lsh(int nx, int N, int el) {
    const int smax = fft_lsh(N, el);
    const int smin = smax < 3 ? smax : 3;
    int best = smax;
    long bw = -1;
    for (int s = smax; s >= smin; s--) {
        const long w2 = 2L << s;
        const long w = (nx + w2 - 1) / w2 * w2;
        if (bw < 0 || w < bw) {
            bw = w;
            best = s;
        }
    }
    return best;
}

bool plan_has_odd(const FFTPlan &pl) {
    for (int s = 0; s < pl.nst; s++)
        if (pl.radix[s] & 1) return true;
    return false;
}

template <int PASS, bool ODD, int EL>
void dct_launch(const PassParams &P, int nb, size_t lds, hipStream_t st) {
    if (lds > 65536)
        (void)hipFuncSetAttribute((const void *)k_dct<PASS, ODD, EL>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)lds);
    hipLaunchKernelGGL((k_dct<PASS, ODD, EL>), dim3(nb), dim3(FFT_NT), lds, st, P);
}
template <int PASS, bool ODD, int EL>
void dct_launch_batch(const PassParams *Ps, const int *prefix, int nbox, int nb, size_t lds, hipStream_t st) {
    if (lds > 65536)
        (void)hipFuncSetAttribute((const void *)k_dct_batch<PASS, ODD, EL>,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    hipLaunchKernelGGL((k_dct_batch<PASS, ODD, EL>), dim3(nb), dim3(FFT_NT), lds, st, Ps, prefix, nbox);
}
// variant index: odd radices, long lines
template <int PASS>
void dct_launch_v(int v, const PassParams &P, int nb, size_t lds, hipStream_t st) {
    switch (v) {
        case 0: dct_launch<PASS, false, DCT_ELEMS>(P, nb, lds, st); break;
        case 1: dct_launch<PASS, false, FFT_ELEMS>(P, nb, lds, st); break;
        case 2: dct_launch<PASS, true, DCT_ELEMS>(P, nb, lds, st); break;
        default: dct_launch<PASS, true, FFT_ELEMS>(P, nb, lds, st); break;
    }
}
template <int PASS>
void dct_launch_batch_v(int v, const PassParams *Ps, const int *prefix, int nbox, int nb, size_t lds, hipStream_t st) {
    switch (v) {
        case 0: dct_launch_batch<PASS, false, DCT_ELEMS>(Ps, prefix, nbox, nb, lds, st); break;
        case 1: dct_launch_batch<PASS, false, FFT_ELEMS>(Ps, prefix, nbox, nb, lds, st); break;
        case 2: dct_launch_batch<PASS, true, DCT_ELEMS>(Ps, prefix, nbox, nb, lds, st); break;
        default: dct_launch_batch<PASS, true, FFT_ELEMS>(Ps, prefix, nbox, nb, lds, st); break;
    }
}

}  // namespace

// parameters of DCT pass `pass` (0 x forward + divergence, 1 y forward, 2 z
// forward + modal solve + z inverse, 3 y inverse, 4 x inverse); returns the
// block count and the dynamic LDS size
int make_dct_pass_params(const Grid &g, const PoissonPlans &pl, double dx, double dy, double dz, double dt,
                         const double *us, const double *vs, const double *ws, const double *rhs_in,
                         const double *planes, const double *mean_dp_dev, double mean_dp_host, double *work,
                         double *dp_out, double *sol_out, int pass, void *out, size_t *lds, int lsh_drop,
                         double *p_acc) {
    PassParams &P = *(PassParams *)out;
    P = PassParams{};
    P.p_acc = (pass == 4) ? p_acc : nullptr;
    P.g = g;
    P.rhs = rhs_in;
    P.us = us; P.vs = vs; P.ws = ws;
    P.dx = dx; P.dy = dy; P.dz = dz; P.dt = dt;
    P.rdx = 1.0 / dx; P.rdy = 1.0 / dy; P.rdz = 1.0 / dz; P.rdt = 1.0 / dt;
    P.planes = planes;
    if (planes) {
        const PlaneLayout L = plane_layout(g.nx, g.ny, g.nz);
        for (int i = 0; i < 6; i++) P.boff[i] = L.off[18 + i];
    }
    for (int a = 0; a < 3; a++) P.pbcs[a] = pl.pbc[a];
    P.work = work;
    P.ex = pl.ex; P.ey = pl.ey; P.ez = pl.ez;
    P.mean_dp_dev = mean_dp_dev;
    P.mean_dp = mean_dp_host;
    P.dp_out = dp_out;
    P.sol_out = sol_out;
    const int ax = (pass == 0 || pass == 4) ? 0 : (pass == 2) ? 2 : 1;
    const FFTPlan plans[3] = {pl.px, pl.py, pl.pz};
    const double2 *wq[3] = {pl.wq_x, pl.wq_y, pl.wq_z};
    const int n[3] = {g.nx, g.ny, g.nz};
    P.pl = plans[ax];
    P.wq = wq[ax];
    P.pbc = pl.pbc[ax];
    P.odd = plan_has_odd(P.pl) ? 1 : 0;
    P.big = n[ax] > DCT_ELEMS ? 1 : 0;
    const int el = P.big ? FFT_ELEMS : DCT_ELEMS;
    P.dN = make_idiv(g.nx);
    P.dNy = make_idiv(g.ny);
    if (ax == 0) {
        P.lsh = fft_lsh(g.nx, el) - lsh_drop;
        if (P.lsh < 0) P.lsh = 0;
        P.LS = fft_row(g.nx);
        *lds = ((size_t)(1 << P.lsh) * P.LS + g.nx) * sizeof(double2);
        const long RL = (long)g.ny * g.nz, per = 2L << P.lsh;
        P.nb = (int)((RL + per - 1) / per);
        P.nb8 = (int)round8((unsigned)P.nb);
        return P.nb8;
    }
    const int N = n[ax];
    P.lsh = dct_yz_lsh(g.nx, N, el) - lsh_drop;
    if (P.lsh < 0) P.lsh = 0;
    P.ngrp = (int)((g.nx + (2L << P.lsh) - 1) / (2L << P.lsh));
    P.dGrp = make_idiv(P.ngrp);
    *lds = ((size_t)(1 << P.lsh) * N + N) * sizeof(double2);
    P.nb = P.ngrp * (ax == 1 ? g.nz : g.ny);
    P.nb8 = (int)round8((unsigned)P.nb);
    return P.nb8;
}

size_t pass_params_size() { return sizeof(PassParams); }

int dct_params_variant(const void *params) {
    const PassParams &P = *(const PassParams *)params;
    return (P.odd ? 2 : 0) | (P.big ? 1 : 0);
}

void launch_dct_pass(int pass, const void *params, int nb, size_t lds, hipStream_t st) {
    const PassParams &P = *(const PassParams *)params;
    const int v = dct_params_variant(params);
    switch (pass) {
        case 0: dct_launch_v<0>(v, P, nb, lds, st); break;
        case 1: dct_launch_v<1>(v, P, nb, lds, st); break;
        case 2: dct_launch_v<2>(v, P, nb, lds, st); break;
        case 3: dct_launch_v<3>(v, P, nb, lds, st); break;
        default: dct_launch_v<4>(v, P, nb, lds, st); break;
    }
}

void launch_dct_batch(int pass, int variant, const void *params, const int *prefix, int nbox, int nb, size_t lds,
                      hipStream_t st) {
    const PassParams *Ps = (const PassParams *)params;
    switch (pass) {
        case 0: dct_launch_batch_v<0>(variant, Ps, prefix, nbox, nb, lds, st); break;
        case 1: dct_launch_batch_v<1>(variant, Ps, prefix, nbox, nb, lds, st); break;
        case 2: dct_launch_batch_v<2>(variant, Ps, prefix, nbox, nb, lds, st); break;
        case 3: dct_launch_batch_v<3>(variant, Ps, prefix, nbox, nb, lds, st); break;
        default: dct_launch_batch_v<4>(variant, Ps, prefix, nbox, nb, lds, st); break;
    }
}

void poisson_dct_solve(const Grid &g, const PoissonPlans &pl, double dx, double dy, double dz, double dt,
                       const double *us, const double *vs, const double *ws, const double *rhs_in,
                       const double *planes, const double *mean_dp_dev, double mean_dp_host, double *work,
                       double *dp_out, double *sol_out, hipStream_t st, hipEvent_t *ev, double *p_acc) {
    for (int pass = 0; pass < 5; pass++) {
        PassParams P;
        size_t lds = 0;
        const int nb = make_dct_pass_params(g, pl, dx, dy, dz, dt, us, vs, ws, rhs_in, planes, mean_dp_dev,
                                            mean_dp_host, work, dp_out, sol_out, pass, &P, &lds, 0, p_acc);
        launch_dct_pass(pass, &P, nb, lds, st);
        if (ev) (void)hipEventRecord(ev[pass], st);
    }
}

}  // namespace ns
