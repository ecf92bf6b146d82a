// C ABI of libns_hip: context lifecycle, host-side plans (FFT radices and
// twiddles, Poisson eigenvalues, CN line systems), the fused time step and
// the per-subroutine mirrors declared in include/ns_hip.h.
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <deque>
#include <utility>
#include <vector>

#include "ns_internal.h"

using namespace ns;

struct ns_ctx {
    ns_config cfg;
    Grid g;   // the box in the layout of the resident state (dev_grid: line-aligned rows)
    Grid gr;  // the box in the reference layout (caller arrays, mirrors, adjoint buffers)
    PlaneLayout L;
    hipStream_t st = nullptr;
    // AB2-CN: two more streams, so that the three velocity components take
    // their x, y, z sweeps side by side (fork and join events on st)
    hipStream_t side[2] = {nullptr, nullptr};
    hipEvent_t fork = nullptr, join[2] = {nullptr, nullptr};
    int side_ok = 0;
    int dev = 0;  // device current at ns_create (owns the stream and every buffer)
    // resident state: the reference's ghosted arrays, with x rows padded to
    // whole 128-byte lines (layout g)
    double *u = nullptr, *v = nullptr, *w = nullptr, *p = nullptr;
    double *hx = nullptr, *hy = nullptr, *hz = nullptr;
    double *us = nullptr, *vs = nullptr, *ws = nullptr, *dp = nullptr;
    double *us2 = nullptr, *vs2 = nullptr, *ws2 = nullptr;  // second u* buffer (lazy projection)
    // lazy projection (dp_flag == 0): the state is (u*, dp) and u = u* - dt*grad(dp)
    // is formed inside the next predictor; u, v, w, p are materialised on demand
    int lazy = 0, stale = 0;
    // all-periodic lazy AB2/Euler steps leave the ghost entries of u*, v*, w*
    // unfilled (star_fill = 1) and read them through their periodic images;
    // ghost_user: the velocity's ghosts are the caller's (just uploaded), so
    // the next predictor reads them as they are
    int star_fill = 0, ghost_user = 0;
    // lazy with dp_flag = 1 (DCT solve, no gauge): p is kept current by the
    // steps themselves, p += dp where dp is formed (x inverse pass, pressure
    // fill), so only u, v, w wait for the projection
    int pacc = 0;
    double *planes = nullptr;   // current plane set
    double *snaps = nullptr;    // snapshot history for ns_resim_advance
    size_t snaps_len = 0;
    double2 *spec = nullptr;    // periodic spectrum [z][y][kx]
    double *work = nullptr;     // DCT work / interior scratch
    double *partial = nullptr;  // max-div partials
    unsigned int *counter = nullptr;
    PoissonPlans pp{};
    LineSys cn[3][2]{};
    LineSys cnT[3][2]{};  // transposed CN systems (adjoint of the ADI sweeps, walls)
    int cnT_ok = 0;       // cnT matches the current dt
    int first = 1;
    // parent-run capture of a sub-box's boundary plane sets
    int cap_on = 0, cap_every = 1, cap_max = 0, cap_count = 0, cap_step = 0;
    Grid cap_g = Grid(0, 0, 0, 0, 0);
    int cap_o[3] = {0, 0, 0};
    long cap_len = 0;
    double *cap_dst = nullptr;
    double *cap_partial = nullptr;
    int poisson_ok = 0;  // 0: mixed periodic / non-periodic codes (no transform solver)
    // two steady steps captured as one graph, per parity of the lazy u* buffers
    hipGraphExec_t graph[2] = {nullptr, nullptr};
    double *us_base = nullptr;  // u* buffer of parity 0
    int graph_off = 0;          // capture failed or disabled
    double *mpartial = nullptr; // error-metric block partials
    double *hsave[3] = {nullptr, nullptr, nullptr};  // saved AB2 history (sub-step phase)
    // discrete adjoint: adjoint state (u, v, w, p, H) and work arrays
    double *ab[3] = {nullptr, nullptr, nullptr}, *apb = nullptr, *ahb[3] = {nullptr, nullptr, nullptr};
    double *ausb[3] = {nullptr, nullptr, nullptr}, *adpb = nullptr, *agb[3] = {nullptr, nullptr, nullptr};
    double *ar = nullptr, *arsol = nullptr, *apbar = nullptr;
    // adjoint of the convective part of the right-hand side (AB2-CN: the
    // diffusive and convective terms enter u* and H with different weights;
    // explicit schemes alias agb)
    double *agcb[3] = {nullptr, nullptr, nullptr};
    std::vector<void *> allocs;
    // device copies of host arguments of the mirror entry points, one buffer
    // per argument position, grown on demand and kept until ns_destroy, so a
    // host that calls the mirrors every step allocates only on the first call
    std::vector<std::pair<double *, size_t>> stage;
    // one dense array of the largest component: host transfers of the
    // resident state pass through it (relayout on the device)
    double *xfer = nullptr;
};

namespace {

const double PI = 3.1415926535897932;  // big_periodic_3D.f90:6

#define HIPCHK(x)                          \
    do {                                   \
        if ((x) != hipSuccess) {           \
            (void)hipGetLastError();       \
            return NS_ERR_HIP;             \
        }                                  \
    } while (0)

template <class T>
int dalloc(ns_ctx *c, T **p, size_t n) {
    void *q = nullptr;
    if (n == 0) n = 1;
    if (hipMalloc(&q, n * sizeof(T)) != hipSuccess) {
        (void)hipGetLastError();
        return NS_ERR_NOMEM;
    }
    c->allocs.push_back(q);
    *p = (T *)q;
    return NS_OK;
}

template <class T>
int dupload(ns_ctx *c, T **p, const std::vector<T> &h) {
    int rc = dalloc(c, p, h.size());
    if (rc) return rc;
    HIPCHK(hipMemcpy(*p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
    return NS_OK;
}

// a component's ghosted array in layout g: the whole allocation is zeroed
// (spare elements included) and *p is set to the array origin
int dalloc_field(ns_ctx *c, double **p, const Grid &g, int comp) {
    double *raw = nullptr;
    const size_t n = field_alloc(g, comp);
    int rc = dalloc(c, &raw, n);
    if (rc) return rc;
    HIPCHK(hipMemsetAsync(raw, 0, n * sizeof(double), c->st));
    *p = raw + field_origin(g, comp);
    return NS_OK;
}

bool factor_radices(int n, FFTPlan &pl) {
    pl.n = n;
    pl.nst = 0;
    int m = n;
    while (m % 8 == 0 && m > 8) { pl.radix[pl.nst++] = 8; m /= 8; }
    // remaining power of two part
    if (m % 8 == 0) { pl.radix[pl.nst++] = 8; m /= 8; }
    if (m % 4 == 0) { pl.radix[pl.nst++] = 4; m /= 4; }
    if (m % 2 == 0) { pl.radix[pl.nst++] = 2; m /= 2; }
    const int odd[3] = {3, 5, 7};
    for (int q = 0; q < 3; q++)
        while (m % odd[q] == 0) {
            if (pl.nst >= 10) return false;
            pl.radix[pl.nst++] = odd[q];
            m /= odd[q];
        }
    return m == 1 && pl.nst > 0 && pl.nst <= 10;
}

// lengths up to 2048: a transform pass holds its lines (at least one) in
// one block's LDS and registers
int make_fft_plan(ns_ctx *c, int n, FFTPlan &pl) {
    if (n < 2 || n > 2048 || !factor_radices(n, pl)) return NS_ERR_UNSUPPORTED;
    std::vector<double2> tw(n);
    const long double tp = 2.0L * acosl(-1.0L);
    for (int m = 0; m < n; m++) {
        long double a = tp * (long double)m / (long double)n;
        tw[m] = make_double2((double)cosl(a), (double)-sinl(a));
    }
    double2 *d;
    int rc = dupload(c, &d, tw);
    pl.tw = d;
    return rc;
}

int make_quarter_wave(ns_ctx *c, int n, const double2 **out) {
    std::vector<double2> q(n);
    const long double pi = acosl(-1.0L);
    for (int k = 0; k < n; k++) {
        long double a = pi * (long double)k / (2.0L * n);
        q[k] = make_double2((double)cosl(a), (double)sinl(a));
    }
    double2 *d;
    int rc = dupload(c, &d, q);
    *out = d;
    return rc;
}

// Poisson eigenvalue parts, evaluated like the reference:
// FFT: (sin(pi*(i-1)/nx)/dx)**2 (big_periodic_3D.f90:242-252)
// DCT: (sin((i-1)*pi/2/nx))**2/dx2 for pbc 3, (sin(i*pi/2/nx))**2/dx2 for pbc 2 (...interBC:321-351)
int make_eigs(ns_ctx *c) {
    const ns_config &f = c->cfg;
    const int n[3] = {f.nx, f.ny, f.nz};
    const double d[3] = {f.dx, f.dy, f.dz};
    const double *outp[3];
    for (int ax = 0; ax < 3; ax++) {
        std::vector<double> e;
        if (!c->pp.dct) {
            int len = (ax == 0) ? n[0] / 2 + 1 : n[ax];
            e.resize(len);
            for (int i = 1; i <= len; i++) {
                double s = sin(PI * (i - 1) / n[ax]) / d[ax];
                e[i - 1] = s * s;
            }
        } else {
            e.resize(n[ax]);
            double d2 = d[ax] * d[ax];
            for (int i = 1; i <= n[ax]; i++) {
                double s = (f.pbc[ax] == 2) ? sin(i * PI / 2 / n[ax]) : sin((i - 1) * PI / 2 / n[ax]);
                e[i - 1] = s * s / d2;
            }
        }
        double *dd;
        int rc = dupload(c, &dd, e);
        if (rc) return rc;
        outp[ax] = dd;
    }
    c->pp.ex = outp[0];
    c->pp.ey = outp[1];
    c->pp.ez = outp[2];
    return NS_OK;
}

// Thomas factorisation of a tridiagonal (lo_i: row i col i-1, d_i, up_i: row i col i+1)
void thomas_factor(const std::vector<double> &lo, const std::vector<double> &d, const std::vector<double> &up,
                   std::vector<double> &inv, std::vector<double> &cp) {
    size_t m = d.size();
    inv.assign(m, 0.0);
    cp.assign(m, 0.0);
    inv[0] = 1.0 / d[0];
    cp[0] = up[0] * inv[0];
    for (size_t i = 1; i < m; i++) {
        double den = d[i] - lo[i] * cp[i - 1];
        inv[i] = 1.0 / den;
        cp[i] = up[i] * inv[i];
    }
}
std::vector<double> thomas_solve(const std::vector<double> &lo, const std::vector<double> &inv,
                                 const std::vector<double> &cp, std::vector<double> r) {
    size_t m = r.size();
    r[0] = r[0] * inv[0];
    for (size_t i = 1; i < m; i++) r[i] = (r[i] - lo[i] * r[i - 1]) * inv[i];
    for (size_t i = m - 1; i-- > 0;) r[i] = r[i] - cp[i] * r[i + 1];
    return r;
}

// Partitioned solve of a wall line (N rows) in K blocks of at most R rows
// (PART_ROWS, or PART_ROWS_LONG when PART_MAX_BLOCKS blocks of PART_ROWS rows
// do not reach N), block k = rows [kN/K, (k+1)N/K).  Each block is solved on
// its own (local Thomas factors loL, invL, cpL, restarted at the block's first
// row); the couplings to the neighbouring blocks' end values enter through
// the spikes V (block k's solve of lo_{s_k} e_first) and W (of up_{e_k}
// e_last): x = y - x_{s_k - 1} V - x_{e_k + 1} W.  The 2K end values solve
// (I + M) z = ends, M holding the spikes' end entries; G = (I + M)^{-1}.
// Every table depends only on the matrix, which is the same for every line.
// cyclic: the system also couples row 0 to row N-1 (alpha) and row N-1 to
// row 0 (beta), so block 0's left neighbour is block K-1 and block K-1's
// right neighbour block 0 (the same spikes, with the corners as couplings).
// Returns K (0: the line is too long for PART_MAX_BLOCKS blocks), sets R and fills
// tab = loL | invL | cpL | V | W (N each) | G (2K x 2K, row-major).
// (The fewest blocks that fit the rows: more, shorter blocks fill the device
// better for a single small box but cost every lane a longer coupling sum,
// which slowed batches of many small boxes by 14%.)
int part_tables(const std::vector<double> &lo, const std::vector<double> &d, const std::vector<double> &up,
                std::vector<double> &tab, int &R, bool cyclic = false, double alpha = 0.0, double beta = 0.0) {
    const int N = (int)d.size();
    R = (N <= PART_MAX_BLOCKS * PART_ROWS) ? PART_ROWS : PART_ROWS_LONG;
    const int K = (N + R - 1) / R;
    if (K > PART_MAX_BLOCKS || N < 2 * K) {
        R = 0;
        return 0;
    }
    const int K2 = 2 * K;
    tab.assign((size_t)5 * N + (size_t)K2 * K2, 0.0);
    double *loL = tab.data(), *invL = loL + N, *cpL = invL + N, *V = cpL + N, *W = V + N, *G = W + N;
    auto start = [&](int k) { return (k * N) / K; };
    for (int k = 0; k < K; k++)
        for (int i = start(k); i < start(k + 1); i++) {
            const bool first = (i == start(k));
            const double den = first ? d[i] : d[i] - lo[i] * cpL[i - 1];
            invL[i] = 1.0 / den;
            cpL[i] = up[i] * invL[i];
            loL[i] = first ? 0.0 : lo[i];
        }
    // local solve of block k on rhs e (rows of the block)
    auto lsolve = [&](int k, std::vector<double> r) {
        const int a = start(k), L = start(k + 1) - a;
        double dd = 0.0;
        for (int j = 0; j < L; j++) r[j] = dd = (r[j] - loL[a + j] * dd) * invL[a + j];
        for (int j = L - 2; j >= 0; j--) r[j] = r[j] - cpL[a + j] * r[j + 1];
        return r;
    };
    std::vector<double> M((size_t)K2 * K2, 0.0);
    for (int k = 0; k < K; k++) {
        const int a = start(k), b = start(k + 1), L = b - a;
        const int kl = (k > 0) ? k - 1 : K - 1, kr = (k < K - 1) ? k + 1 : 0;
        if (k > 0 || cyclic) {
            std::vector<double> e(L, 0.0);
            e[0] = (k > 0) ? lo[a] : alpha;
            e = lsolve(k, e);
            for (int j = 0; j < L; j++) V[a + j] = e[j];
            M[(size_t)(2 * k) * K2 + 2 * kl + 1] += e[0];
            M[(size_t)(2 * k + 1) * K2 + 2 * kl + 1] += e[L - 1];
        }
        if (k < K - 1 || cyclic) {
            std::vector<double> e(L, 0.0);
            e[L - 1] = (k < K - 1) ? up[b - 1] : beta;
            e = lsolve(k, e);
            for (int j = 0; j < L; j++) W[a + j] = e[j];
            M[(size_t)(2 * k) * K2 + 2 * kr] += e[0];
            M[(size_t)(2 * k + 1) * K2 + 2 * kr] += e[L - 1];
        }
    }
    // G = (I + M)^{-1} by Gauss-Jordan with partial pivoting
    std::vector<double> A((size_t)K2 * K2);
    for (int i = 0; i < K2; i++)
        for (int j = 0; j < K2; j++) {
            A[(size_t)i * K2 + j] = M[(size_t)i * K2 + j] + (i == j ? 1.0 : 0.0);
            G[(size_t)i * K2 + j] = (i == j) ? 1.0 : 0.0;
        }
    for (int c = 0; c < K2; c++) {
        int pr = c;
        for (int r = c + 1; r < K2; r++)
            if (fabs(A[(size_t)r * K2 + c]) > fabs(A[(size_t)pr * K2 + c])) pr = r;
        if (A[(size_t)pr * K2 + c] == 0.0) return 0;
        if (pr != c)
            for (int j = 0; j < K2; j++) {
                std::swap(A[(size_t)pr * K2 + j], A[(size_t)c * K2 + j]);
                std::swap(G[(size_t)pr * K2 + j], G[(size_t)c * K2 + j]);
            }
        const double piv = 1.0 / A[(size_t)c * K2 + c];
        for (int j = 0; j < K2; j++) {
            A[(size_t)c * K2 + j] *= piv;
            G[(size_t)c * K2 + j] *= piv;
        }
        for (int r = 0; r < K2; r++) {
            if (r == c) continue;
            const double f = A[(size_t)r * K2 + c];
            if (f == 0.0) continue;
            for (int j = 0; j < K2; j++) {
                A[(size_t)r * K2 + j] -= f * A[(size_t)c * K2 + j];
                G[(size_t)r * K2 + j] -= f * G[(size_t)c * K2 + j];
            }
        }
    }
    return K;
}

// device copy of a coefficient table: into the table the system held before
// (a rebuild for a new dt: the line lengths of a context never change, and the
// caller has drained the stream) or into a new one
template <typename T>
int dstore(ns_ctx *c, const T *old, T **p, const std::vector<T> &h) {
    if (!old) return dupload(c, p, h);
    *p = const_cast<T *>(old);
    HIPCHK(hipMemcpy(*p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
    return NS_OK;
}

// CN line system for direction ax (0..2) and type A (face-normal, n+1) or B (n+2);
// transpose: the system of A^T, whose solve is the adjoint of the sweep
int make_line_sys(ns_ctx *c, int ax, int typeB, LineSys &S, bool transpose = false) {
    const ns_config &f = c->cfg;
    const int n = (ax == 0) ? f.nx : (ax == 1) ? f.ny : f.nz;
    const double h = (ax == 0) ? f.dx : (ax == 1) ? f.dy : f.dz;
    const int bc = f.bc[ax];
    const double dx2 = h * h, idx2 = 1.0 / dx2, dt = f.dt, nu = f.nu;
    // mat_CN entries (NS_functions.f90:1261-1269)
    const double dg = 1.0 + 0.5 * dt * nu * 2.0 * idx2;
    const double a = 0.5 * dt * nu * 1.0 * idx2;
    const LineSys old = S;
    S = LineSys{};
    S.a = a;
    int rc0 = NS_OK;
    std::vector<double> lo, d, up, inv, cp;
    if (bc == 1) {
        const int m = n;
        S.periodic = 1;
        S.m = m;
        S.N = typeB ? n + 2 : n + 1;
        S.off = typeB ? 1 : 0;
        S.quirk = (!typeB && !(f.flags & NS_FLAG_CN_CYCLIC)) ? 1 : 0;
        lo.assign(m, -a);
        up.assign(m, -a);
        d.assign(m, dg);
        lo[0] = 0.0;
        up[m - 1] = 0.0;
        double alpha = -a, beta = -a;  // corners (0,m-1) and (m-1,0)
        if (S.quirk) {
            d[0] = dg - a;  // (1+2a)u1 - a u_N with u_N = u_1 + r_N
            alpha = 0.0;
        }
        // the transposed system exchanges the corners (the band is symmetric)
        if (transpose) {
            std::swap(alpha, beta);
            S.adj = 1;
        }
        double gamma = -d[0];
        std::vector<double> dt_ = d;
        dt_[0] = d[0] - gamma;
        dt_[m - 1] = d[m - 1] - alpha * beta / gamma;
        thomas_factor(lo, dt_, up, inv, cp);
        std::vector<double> uvec(m, 0.0);
        uvec[0] = gamma;
        uvec[m - 1] = beta;
        std::vector<double> z = thomas_solve(lo, inv, cp, uvec);
        S.s = alpha / gamma;
        S.den = 1.0 + z[0] + S.s * z[m - 1];
        double *dz_;
        int rc = dstore(c, old.z, &dz_, z);
        if (rc) return rc;
        S.z = dz_;
        // partitioned solve of the cyclic system itself (lines of at most
        // PART_MAX_BLOCKS * PART_ROWS_LONG unknowns)
        std::vector<double> tab;
        S.K = part_tables(lo, d, up, tab, S.R, true, alpha, beta);
        if (S.K) {
            double *dt2;
            if ((rc = dstore(c, old.pt, &dt2, tab))) return rc;
            S.pt = dt2;
        }
    } else if (bc == 2 || bc == 4) {
        const int N = typeB ? n + 2 : n + 1;
        S.periodic = 0;
        S.N = N;
        S.m = N;
        S.off = 0;
        lo.assign(N, -a);
        up.assign(N, -a);
        d.assign(N, dg);
        lo[0] = 0.0;
        up[N - 1] = 0.0;
        if (!typeB || bc == 4) {
            d[0] = 1.0; up[0] = 0.0;
            d[N - 1] = 1.0; lo[N - 1] = 0.0;
        } else {
            d[0] = 0.5; up[0] = 0.5;
            d[N - 1] = 0.5; lo[N - 1] = 0.5;
        }
        if (transpose) {
            // (A^T)_{i,i-1} = A_{i-1,i}, (A^T)_{i,i+1} = A_{i+1,i}
            std::vector<double> lt(N, 0.0), ut(N, 0.0);
            for (int i = 1; i < N; i++) lt[i] = up[i - 1];
            for (int i = 0; i + 1 < N; i++) ut[i] = lo[i + 1];
            lo.swap(lt);
            up.swap(ut);
        }
        thomas_factor(lo, d, up, inv, cp);
        S.z = nullptr;
        // partitioned solve (lines of at most PART_MAX_BLOCKS * PART_ROWS_LONG rows)
        std::vector<double> tab;
        S.K = part_tables(lo, d, up, tab, S.R);
        if (S.K) {
            double *dt_;
            // the table length depends only on N, which a rebuild keeps
            if ((rc0 = dstore(c, old.pt, &dt_, tab))) return rc0;
            S.pt = dt_;
        }
        // twisted factorization at tk = N/2: the bottom-up elimination is Thomas
        // on the reversed system (sub- and super-diagonal exchanged)
        const int tk = N / 2;
        std::vector<double> loR(N, 0.0), dR(N), upR(N, 0.0), invR, cpR;
        for (int i = 0; i < N; i++) {
            loR[i] = up[N - 1 - i];
            dR[i] = d[N - 1 - i];
            upR[i] = lo[N - 1 - i];
        }
        thomas_factor(loR, dR, upR, invR, cpR);
        S.tk = tk;
        S.ak = lo[tk];
        S.ck = up[tk];
        // cp_{tk-1} (top-down) and cq_{tk+1} = cpR[N-2-tk] (bottom-up)
        S.gk = 1.0 / (d[tk] - lo[tk] * cp[tk - 1] - up[tk] * cpR[N - 2 - tk]);
        double *dlr, *dir_, *dcr;
        if ((rc0 = dstore(c, old.loR, &dlr, loR)) || (rc0 = dstore(c, old.invR, &dir_, invR)) ||
            (rc0 = dstore(c, old.cpR, &dcr, cpR)))
            return rc0;
        S.loR = dlr;
        S.invR = dir_;
        S.cpR = dcr;
    } else {
        return NS_ERR_UNSUPPORTED;
    }
    double *dl, *di, *dc;
    int rc = dstore(c, old.lo, &dl, lo);
    if (rc) return rc;
    rc = dstore(c, old.inv, &di, inv);
    if (rc) return rc;
    rc = dstore(c, old.cp, &dc, cp);
    if (rc) return rc;
    S.lo = dl;
    S.inv = di;
    S.cp = dc;
    return NS_OK;
}

bool is_dev_ptr(const void *p) {
    if (!p) return false;
    hipPointerAttribute_t at;
    hipError_t e = hipPointerGetAttributes(&at, p);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    return at.type == hipMemoryTypeDevice || at.type == hipMemoryTypeManaged;
}

// host-or-device argument staged onto the device for a mirror call; a host
// argument goes through the context's staging buffer of its position (slot)
struct Arg {
    double *dev = nullptr;
    void *user = nullptr;
    size_t n = 0;
    bool staged = false;
    int slot = -1;
};

int arg_in(ns_ctx *c, Arg &a, const double *p, size_t n, bool copy_in) {
    a.user = (void *)p;
    a.n = n;
    if (!p || a.slot < 0) return NS_ERR_INVALID;
    if (is_dev_ptr(p)) {
        a.dev = (double *)p;
        a.staged = false;
        return NS_OK;
    }
    if (c->stage.size() <= (size_t)a.slot) c->stage.resize((size_t)a.slot + 1, {nullptr, 0});
    auto &buf = c->stage[(size_t)a.slot];
    if (buf.second < n) {
        if (buf.first) {
            // earlier calls' copies through this buffer are ordered on the stream
            HIPCHK(hipStreamSynchronize(c->st));
            (void)hipFree(buf.first);
        }
        buf = {nullptr, 0};
        void *q = nullptr;
        if (hipMalloc(&q, (n ? n : 1) * sizeof(double)) != hipSuccess) {
            (void)hipGetLastError();
            return NS_ERR_NOMEM;
        }
        buf = {(double *)q, n};
    }
    a.dev = buf.first;
    a.staged = true;
    if (copy_in) HIPCHK(hipMemcpyAsync(a.dev, p, n * sizeof(double), hipMemcpyHostToDevice, c->st));
    return NS_OK;
}

int arg_out(ns_ctx *c, Arg &a) {
    if (a.staged) HIPCHK(hipMemcpyAsync(a.user, a.dev, a.n * sizeof(double), hipMemcpyDeviceToHost, c->st));
    return NS_OK;
}

// Transfers of one component's array between a caller's array in the
// reference layout (host or device memory) and the resident state (layout
// c->g): device arrays are relaid directly, host arrays through c->xfer.
// Stream-ordered on c->st; a host destination is written when the stream
// drains.
int xfer_buffer(ns_ctx *c) {
    return c->xfer ? NS_OK : dalloc(c, &c->xfer, c->gr.ref_np());
}
size_t ref_size(const Grid &g, int comp) {
    return comp == 0 ? g.ref_nu() : comp == 1 ? g.ref_nv() : comp == 2 ? g.ref_nw() : g.ref_np();
}
int field_in(ns_ctx *c, double *own, const double *ext, int comp) {
    int n[3];
    comp_dims(c->g, comp, n);
    const size_t rows = (size_t)n[1] * n[2];
    const double *src = ext;
    if (!is_dev_ptr(ext)) {
        int rc = xfer_buffer(c);
        if (rc) return rc;
        HIPCHK(hipMemcpyAsync(c->xfer, ext, ref_size(c->g, comp) * sizeof(double), hipMemcpyHostToDevice, c->st));
        src = c->xfer;
    }
    launch_relayout(src, n[0], own, comp_pitch(c->g, comp), n[0], rows, nullptr, c->st);
    return NS_OK;
}
// gauge (device, optional): subtracted from every value on the way out
int field_out(ns_ctx *c, double *ext, const double *own, int comp, const double *gauge = nullptr) {
    int n[3];
    comp_dims(c->g, comp, n);
    const size_t rows = (size_t)n[1] * n[2];
    if (is_dev_ptr(ext)) {
        launch_relayout(own, comp_pitch(c->g, comp), ext, n[0], n[0], rows, gauge, c->st);
        return NS_OK;
    }
    int rc = xfer_buffer(c);
    if (rc) return rc;
    launch_relayout(own, comp_pitch(c->g, comp), c->xfer, n[0], n[0], rows, gauge, c->st);
    HIPCHK(hipMemcpyAsync(ext, c->xfer, ref_size(c->g, comp) * sizeof(double), hipMemcpyDeviceToHost, c->st));
    return NS_OK;
}

// the arguments of one call, each with its own staging slot.  std::deque
// keeps references to earlier elements valid across push_back (a std::vector
// would reallocate and leave them dangling)
struct ArgList {
    std::deque<Arg> v;
    Arg &add() {
        v.emplace_back();
        v.back().slot = (int)v.size() - 1;
        return v.back();
    }
    Arg &back() { return v.back(); }
};

bool valid_cfg(const ns_config *f) {
    if (!f) return false;
    if (f->nx < 2 || f->ny < 2 || f->nz < 2) return false;
    if (!(f->dx > 0) || !(f->dy > 0) || !(f->dz > 0) || !(f->dt > 0) || !(f->nu >= 0)) return false;
    if (f->scheme < NS_EULER || f->scheme > NS_AB2_CN) return false;
    for (int a = 0; a < 3; a++) {
        if (f->bc[a] < 1 || f->bc[a] > 4 || f->bc[a] == 3) return false;
        if (f->pbc[a] < 1 || f->pbc[a] > 4) return false;
    }
    if (f->dp_flag != 0 && f->dp_flag != 1) return false;
    return true;
}

int run_poisson(ns_ctx *c, const double *us, const double *vs, const double *ws, const double *rhs_in,
                const double *planes, const double *mean_dev, double mean_host, double *dp_out, double *sol_out,
                hipEvent_t *ev = nullptr, bool dp_ghost_x = false, bool wrap_vel = false, double *p_acc = nullptr) {
    const ns_config &f = c->cfg;
    if (!c->poisson_ok) return NS_ERR_UNSUPPORTED;
    if (!c->pp.dct)
        poisson_fft_solve(c->g, c->pp, f.dx, f.dy, f.dz, f.dt, us, vs, ws, rhs_in, c->spec, dp_out, sol_out, c->st, ev,
                          dp_ghost_x, wrap_vel);
    else
        poisson_dct_solve(c->g, c->pp, f.dx, f.dy, f.dz, f.dt, us, vs, ws, rhs_in, planes, mean_dev, mean_host,
                          c->work, dp_out, sol_out, c->st, ev, p_acc);
    return NS_OK;
}

// capture the planes of axis ax (velocity) from u*, v*, w* into the current set
void cap_vel(ns_ctx *c, int ax) {
    const ns_config &f = c->cfg;
    const double d[3] = {f.dx, f.dy, f.dz};
    double *dst = c->cap_dst + (size_t)c->cap_len * c->cap_count;
    launch_capture(c->g, c->cap_g, c->cap_o, ax, 0, d, nullptr, c->us, c->vs, c->ws, dst, c->st);
}
void cap_pressure(ns_ctx *c, const double *gauge) {
    const ns_config &f = c->cfg;
    const double d[3] = {f.dx, f.dy, f.dz};
    double *dst = c->cap_dst + (size_t)c->cap_len * c->cap_count;
    for (int ax = 0; ax < 3; ax++)
        launch_capture(c->g, c->cap_g, c->cap_o, ax, 1, d, gauge, c->dp, nullptr, nullptr, dst, c->st);
    PlaneLayout Ls = plane_layout(c->cap_g.nx, c->cap_g.ny, c->cap_g.nz);
    launch_subbox_mean(c->g, c->cap_g, c->cap_o, gauge, c->dp, c->cap_partial, dst + Ls.off[24], c->st);
}

// the steps of this context skip the velocity ghost fill: lazy, every
// velocity axis periodic, and no ADI sweeps (AB2-CN's periodic sweeps write
// the ghost rows themselves)
bool skips_star_fill(const ns_ctx *c) {
    const ns_config &f = c->cfg;
    return c->lazy && f.bc[0] == 1 && f.bc[1] == 1 && f.bc[2] == 1 && f.scheme != NS_AB2_CN;
}

// fill the ghost entries of u*, v*, w* where the last steps left them
void fill_star(ns_ctx *c) {
    if (!c->star_fill) return;
    launch_vel_bc(c->g, c->cfg.bc, c->planes, c->us, c->vs, c->ws, c->st);
    c->star_fill = 0;
}

// lazy mode: form u, v, w, p from the current (u*, dp)
void materialize(ns_ctx *c) {
    if (!c->lazy || !c->stale) return;
    fill_star(c);
    const ns_config &f = c->cfg;
    const double *gauge = (f.flags & NS_FLAG_GAUGE_CORNER) ? c->dp : nullptr;
    launch_project(c->g, f.dx, f.dy, f.dz, f.dt, f.dp_flag, c->us, c->vs, c->ws, c->dp, gauge, c->u, c->v, c->w,
                   c->pacc ? nullptr : c->p, c->st);
    c->stale = 0;
}

// one time step on the context stream; ev (optional) receives NS_NPHASES+1
// events: before the predictor and after each phase
int one_step(ns_ctx *c, hipEvent_t *ev = nullptr) {
    const ns_config &f = c->cfg;
    const Grid &g = c->g;
    const bool cap = c->cap_on && c->cap_count < c->cap_max && ((c->cap_step + 1) % c->cap_every == 0);
    if (ev) (void)hipEventRecord(ev[0], c->st);
    // the periodic x ghosts of dp are written by the x inverse transform
    const bool pghost = !c->pp.dct && f.pbc[0] == 1;
    const bool skip = skips_star_fill(c);
    if (c->lazy) {
        const double *gauge = (f.flags & NS_FLAG_GAUGE_CORNER) ? c->dp : nullptr;
        launch_predict_tiled(f, g, c->first, true, c->us, c->vs, c->ws, c->dp, gauge, c->p, c->hx, c->hy, c->hz, c->us2,
                             c->vs2, c->ws2, c->st, skip && !c->ghost_user);
        std::swap(c->us, c->us2);
        std::swap(c->vs, c->vs2);
        std::swap(c->ws, c->ws2);
    } else {
        launch_predict_tiled(f, g, c->first, false, c->u, c->v, c->w, nullptr, nullptr, c->p, c->hx, c->hy, c->hz, c->us,
                             c->vs, c->ws, c->st);
    }
    if (ev) (void)hipEventRecord(ev[1], c->st);
    if (f.scheme != NS_AB2_CN) {
        if (!skip || cap) {
            launch_vel_bc(g, f.bc, c->planes, c->us, c->vs, c->ws, c->st);
            c->star_fill = 0;
        } else {
            c->star_fill = 1;
        }
        if (cap)
            for (int ax = 0; ax < 3; ax++) cap_vel(c, ax);
    } else {
        bool part = true;
        for (int dir = 1; dir <= 3; dir++) part = part && sweep_partitioned(c->cn[dir - 1][0], c->cn[dir - 1][1]);
        // boxes of at least 2^24 cells: each component's x, y and z sweeps touch
        // only that component (the boundary rows are read from the planes), so
        // the three components run on three streams and the device fills one
        // sweep's tail with another component's blocks; st waits for all three
        // (256^3: 1.535 -> 1.484 ms per periodic step, 1.834 -> 1.782 ms with
        // walls; smaller boxes were slower this way: 128^3 walls 0.276 ->
        // 0.295 ms, 64^3 periodic 0.100 -> 0.119 ms)
        const bool big = (double)g.nx * g.ny * g.nz >= 16777216.0;
        if (part && !cap && c->side_ok && big) {
            HIPCHK(hipEventRecord(c->fork, c->st));
            for (int k = 0; k < 2; k++) HIPCHK(hipStreamWaitEvent(c->side[k], c->fork, 0));
            for (int comp = 0; comp < 3; comp++) {
                hipStream_t s = comp == 0 ? c->st : c->side[comp - 1];
                for (int dir = 1; dir <= 3; dir++)
                    launch_adi_sweep(g, dir, c->cn[dir - 1][0], c->cn[dir - 1][1], c->us, c->vs, c->ws, s, c->planes,
                                     &c->L, nullptr, comp, 1);
            }
            for (int k = 0; k < 2; k++) {
                HIPCHK(hipEventRecord(c->join[k], c->side[k]));
                HIPCHK(hipStreamWaitEvent(c->st, c->join[k], 0));
            }
        } else {
            for (int dir = 1; dir <= 3; dir++) {
                // a partitioned sweep reads the boundary rows (vel_bc_staggered_CN2) from the planes itself
                const LineSys &sa = c->cn[dir - 1][0], &sb = c->cn[dir - 1][1];
                if (sweep_partitioned(sa, sb)) {
                    launch_adi_sweep(g, dir, sa, sb, c->us, c->vs, c->ws, c->st, c->planes, &c->L);
                } else {
                    launch_vel_bc_cn2(g, f.bc[dir - 1], dir, c->planes, c->us, c->vs, c->ws, c->st);
                    launch_adi_sweep(g, dir, sa, sb, c->us, c->vs, c->ws, c->st);
                }
                if (cap) cap_vel(c, dir - 1);  // post-sweep planes (big_periodic_3D.f90:433-489)
            }
        }
    }
    if (ev) (void)hipEventRecord(ev[2], c->st);
    double *pacc = c->pacc ? c->p : nullptr;
    int rc = run_poisson(c, c->us, c->vs, c->ws, nullptr, c->planes, c->planes + c->L.off[24], 0.0, c->dp, nullptr,
                         ev ? ev + 3 : nullptr, pghost, skip, pacc);
    c->ghost_user = 0;
    if (rc) return rc;
    launch_pr_bc(g, f.pbc, c->planes, f.dx, f.dy, f.dz, c->dp, c->st, pghost ? 1 : 0, pacc);
    if (ev) (void)hipEventRecord(ev[8], c->st);
    const double *gauge = (f.flags & NS_FLAG_GAUGE_CORNER) ? c->dp : nullptr;
    if (cap) cap_pressure(c, gauge);
    if (c->lazy) c->stale = 1;
    else
        launch_project(g, f.dx, f.dy, f.dz, f.dt, f.dp_flag, c->us, c->vs, c->ws, c->dp, gauge, c->u, c->v, c->w, c->p,
                       c->st);
    if (ev) (void)hipEventRecord(ev[9], c->st);
    c->first = 0;
    if (c->cap_on) {
        if (cap) c->cap_count++;
        c->cap_step++;
    }
    return NS_OK;
}

}  // namespace

extern "C" {

long ns_plane_set_length(int nx, int ny, int nz) { return plane_layout(nx, ny, nz).total; }

int ns_create(const ns_config *cfg, ns_ctx **out) {
    if (!out) return NS_ERR_INVALID;
    *out = nullptr;
    if (!valid_cfg(cfg)) return NS_ERR_INVALID;
    ns_ctx *c = new ns_ctx();
    c->cfg = *cfg;
    if (hipGetDevice(&c->dev) != hipSuccess) {
        (void)hipGetLastError();
        delete c;
        return NS_ERR_HIP;
    }
    c->g = dev_grid(cfg->nx, cfg->ny, cfg->nz);
    c->gr = ref_grid(cfg->nx, cfg->ny, cfg->nz);
    c->L = plane_layout(cfg->nx, cfg->ny, cfg->nz);
    int rc = NS_OK;
    auto fail = [&](int r) {
        ns_destroy(c);
        return r;
    };
    if (hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking) != hipSuccess) {
        (void)hipGetLastError();
        c->st = nullptr;
        return fail(NS_ERR_HIP);
    }
    if (cfg->scheme == NS_AB2_CN) {
        // without them the sweeps run one after another on st
        bool ok = hipEventCreateWithFlags(&c->fork, hipEventDisableTiming) == hipSuccess;
        for (int i = 0; i < 2 && ok; i++)
            ok = hipStreamCreateWithFlags(&c->side[i], hipStreamNonBlocking) == hipSuccess &&
                 hipEventCreateWithFlags(&c->join[i], hipEventDisableTiming) == hipSuccess;
        c->side_ok = ok ? 1 : 0;
        (void)hipGetLastError();
    }
    const Grid &g = c->g;
    // element indices of the ghosted arrays are 32-bit in some kernels
    if (g.np() >= ((size_t)1 << 32)) return fail(NS_ERR_UNSUPPORTED);
    if ((rc = dalloc_field(c, &c->u, g, 0)) || (rc = dalloc_field(c, &c->v, g, 1)) ||
        (rc = dalloc_field(c, &c->w, g, 2)) || (rc = dalloc_field(c, &c->p, g, 3)) ||
        (rc = dalloc_field(c, &c->hx, g, 0)) || (rc = dalloc_field(c, &c->hy, g, 1)) ||
        (rc = dalloc_field(c, &c->hz, g, 2)) || (rc = dalloc_field(c, &c->us, g, 0)) ||
        (rc = dalloc_field(c, &c->vs, g, 1)) || (rc = dalloc_field(c, &c->ws, g, 2)) ||
        (rc = dalloc_field(c, &c->dp, g, 3)) || (rc = dalloc_field(c, &c->us2, g, 0)) ||
        (rc = dalloc_field(c, &c->vs2, g, 1)) || (rc = dalloc_field(c, &c->ws2, g, 2)) ||
        (rc = dalloc(c, &c->planes, (size_t)c->L.total)) || (rc = dalloc(c, &c->partial, (size_t)4096)) ||
        (rc = dalloc(c, &c->counter, (size_t)1)))
        return fail(rc);
    const bool allp = cfg->pbc[0] == 1 && cfg->pbc[1] == 1 && cfg->pbc[2] == 1;
    const bool nonp = cfg->pbc[0] != 1 && cfg->pbc[1] != 1 && cfg->pbc[2] != 1;
    const bool dct_ok = nonp && cfg->pbc[0] != 4 && cfg->pbc[1] != 4 && cfg->pbc[2] != 4;
    c->poisson_ok = allp || dct_ok;
    if (c->poisson_ok) {
        c->pp.dct = allp ? 0 : 1;
        for (int a = 0; a < 3; a++) c->pp.pbc[a] = cfg->pbc[a];
        if ((rc = make_fft_plan(c, cfg->nx, c->pp.px)) || (rc = make_fft_plan(c, cfg->ny, c->pp.py)) ||
            (rc = make_fft_plan(c, cfg->nz, c->pp.pz)))
            return fail(rc);
        if ((rc = make_eigs(c))) return fail(rc);
        if (allp) {
            if ((rc = dalloc(c, &c->spec, (size_t)spectrum_row(cfg->nx) * cfg->ny * cfg->nz))) return fail(rc);
        } else {
            if ((rc = make_quarter_wave(c, cfg->nx, &c->pp.wq_x)) || (rc = make_quarter_wave(c, cfg->ny, &c->pp.wq_y)) ||
                (rc = make_quarter_wave(c, cfg->nz, &c->pp.wq_z)))
                return fail(rc);
        }
    }
    if ((rc = dalloc(c, &c->work, g.ni()))) return fail(rc);
    if (cfg->scheme == NS_AB2_CN) {
        for (int ax = 0; ax < 3; ax++)
            for (int t = 0; t < 2; t++)
                if ((rc = make_line_sys(c, ax, t, c->cn[ax][t]))) return fail(rc);
    }
    // zero everything so unset ghosts and history read as the reference's zero-initialised arrays
    HIPCHK(hipMemsetAsync(c->u, 0, g.nu() * 8, c->st));
    HIPCHK(hipMemsetAsync(c->v, 0, g.nv() * 8, c->st));
    HIPCHK(hipMemsetAsync(c->w, 0, g.nw() * 8, c->st));
    HIPCHK(hipMemsetAsync(c->p, 0, g.np() * 8, c->st));
    HIPCHK(hipMemsetAsync(c->hx, 0, g.nu() * 8, c->st));
    HIPCHK(hipMemsetAsync(c->hy, 0, g.nv() * 8, c->st));
    HIPCHK(hipMemsetAsync(c->hz, 0, g.nw() * 8, c->st));
    HIPCHK(hipMemsetAsync(c->us, 0, g.nu() * 8, c->st));
    HIPCHK(hipMemsetAsync(c->vs, 0, g.nv() * 8, c->st));
    HIPCHK(hipMemsetAsync(c->ws, 0, g.nw() * 8, c->st));
    HIPCHK(hipMemsetAsync(c->dp, 0, g.np() * 8, c->st));
    HIPCHK(hipMemsetAsync(c->us2, 0, g.nu() * 8, c->st));
    HIPCHK(hipMemsetAsync(c->vs2, 0, g.nv() * 8, c->st));
    HIPCHK(hipMemsetAsync(c->ws2, 0, g.nw() * 8, c->st));
    c->us_base = c->us;
    // lazy projection (dp folded into the next predictor) needs periodic pressure
    // on every periodic velocity axis: the predictor reads dp(n+1) as dp(1)
    c->lazy = (cfg->dp_flag == 0);
    // dp_flag = 1: lazy where p can be formed with dp (the DCT solve's x
    // inverse pass and the pressure fill write every entry of dp once), with
    // no corner gauge (whose value is known only after the whole solve)
    const bool nonper = cfg->pbc[0] != 1 && cfg->pbc[1] != 1 && cfg->pbc[2] != 1;
    if (cfg->dp_flag == 1 && nonper && !(cfg->flags & NS_FLAG_GAUGE_CORNER)) c->lazy = c->pacc = 1;
    for (int a = 0; a < 3; a++)
        if (cfg->bc[a] == 1 && cfg->pbc[a] != 1) c->lazy = 0;
    if (cfg->flags & NS_FLAG_EAGER_PROJECTION) c->lazy = 0;
    if (!c->lazy || !c->poisson_ok) c->lazy = c->pacc = 0;
    HIPCHK(hipMemsetAsync(c->planes, 0, (size_t)c->L.total * 8, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    *out = c;
    return NS_OK;
}

int ns_destroy(ns_ctx *c) {
    if (!c) return NS_ERR_INVALID;
    if (c->st) (void)hipStreamSynchronize(c->st);
    for (int i = 0; i < 2; i++)
        if (c->graph[i]) (void)hipGraphExecDestroy(c->graph[i]);
    for (void *q : c->allocs) (void)hipFree(q);
    for (auto &b : c->stage)
        if (b.first) (void)hipFree(b.first);
    if (c->snaps) (void)hipFree(c->snaps);
    for (int i = 0; i < 2; i++) {
        if (c->side[i]) {
            (void)hipStreamSynchronize(c->side[i]);
            (void)hipStreamDestroy(c->side[i]);
        }
        if (c->join[i]) (void)hipEventDestroy(c->join[i]);
    }
    if (c->fork) (void)hipEventDestroy(c->fork);
    if (c->st) (void)hipStreamDestroy(c->st);
    delete c;
    return NS_OK;
}

int ns_synchronize(ns_ctx *c) {
    if (!c) return NS_ERR_INVALID;
    HIPCHK(hipStreamSynchronize(c->st));
    return NS_OK;
}

int ns_upload_state(ns_ctx *c, const double *u, const double *v, const double *w, const double *p, const double *hx,
                    const double *hy, const double *hz, int first) {
    if (!c || !u || !v || !w) return NS_ERR_INVALID;
    const Grid &g = c->g;
    int rc;
    if ((rc = field_in(c, c->u, u, 0)) || (rc = field_in(c, c->v, v, 1)) || (rc = field_in(c, c->w, w, 2))) return rc;
    if (p) {
        if ((rc = field_in(c, c->p, p, 3))) return rc;
    } else {
        HIPCHK(hipMemsetAsync(c->p, 0, g.np() * 8, c->st));
    }
    const double *hs[3] = {hx, hy, hz};
    double *hd[3] = {c->hx, c->hy, c->hz};
    const size_t hn[3] = {g.nu(), g.nv(), g.nw()};
    for (int a = 0; a < 3; a++) {
        if (hs[a]) {
            if ((rc = field_in(c, hd[a], hs[a], a))) return rc;
        } else {
            HIPCHK(hipMemsetAsync(hd[a], 0, hn[a] * 8, c->st));
        }
    }
    if (c->lazy) {
        // u = u* - dt*grad(0) reproduces u exactly
        HIPCHK(hipMemcpyAsync(c->us, c->u, g.nu() * 8, hipMemcpyDeviceToDevice, c->st));
        HIPCHK(hipMemcpyAsync(c->vs, c->v, g.nv() * 8, hipMemcpyDeviceToDevice, c->st));
        HIPCHK(hipMemcpyAsync(c->ws, c->w, g.nw() * 8, hipMemcpyDeviceToDevice, c->st));
        HIPCHK(hipMemsetAsync(c->dp, 0, g.np() * 8, c->st));
        c->stale = 0;
        c->star_fill = 0;
        c->ghost_user = 1;
    }
    HIPCHK(hipStreamSynchronize(c->st));
    c->first = first ? 1 : 0;
    return NS_OK;
}

int ns_download_state(ns_ctx *c, double *u, double *v, double *w, double *p, double *hx, double *hy, double *hz) {
    if (!c) return NS_ERR_INVALID;
    materialize(c);
    double *dst[7] = {u, v, w, p, hx, hy, hz};
    const double *src[7] = {c->u, c->v, c->w, c->p, c->hx, c->hy, c->hz};
    const int comp[7] = {0, 1, 2, 3, 0, 1, 2};
    for (int i = 0; i < 7; i++) {
        if (!dst[i]) continue;
        const int rc = field_out(c, dst[i], src[i], comp[i]);
        if (rc) return rc;
    }
    HIPCHK(hipStreamSynchronize(c->st));
    return NS_OK;
}

int ns_download_work(ns_ctx *c, double *u_star, double *v_star, double *w_star, double *dp) {
    if (!c) return NS_ERR_INVALID;
    fill_star(c);
    double *dst[4] = {u_star, v_star, w_star, dp};
    const double *src[4] = {c->us, c->vs, c->ws, c->dp};
    // dp gauged like the driver's: dp - dp(1,1,1)
    const double *gauge = (c->cfg.flags & NS_FLAG_GAUGE_CORNER) ? c->dp : nullptr;
    for (int i = 0; i < 4; i++) {
        if (!dst[i]) continue;
        const int rc = field_out(c, dst[i], src[i], i, i == 3 ? gauge : nullptr);
        if (rc) return rc;
    }
    HIPCHK(hipStreamSynchronize(c->st));
    return NS_OK;
}

int ns_state_pointers(ns_ctx *c, double **p) {
    if (!c || !p) return NS_ERR_INVALID;
    fill_star(c);
    double *q[11] = {c->u, c->v, c->w, c->p, c->hx, c->hy, c->hz, c->us, c->vs, c->ws, c->dp};
    for (int i = 0; i < 11; i++) p[i] = q[i];
    return NS_OK;
}

int ns_state_pitch(ns_ctx *c, int *pitch) {
    if (!c || !pitch) return NS_ERR_INVALID;
    *pitch = c->g.pc;  // u-shaped arrays share it (dev_grid: pu == pc)
    return NS_OK;
}

int ns_set_bc_planes(ns_ctx *c, const double *planes) {
    if (!c || !planes) return NS_ERR_INVALID;
    HIPCHK(hipMemcpyAsync(c->planes, planes, (size_t)c->L.total * 8, hipMemcpyDefault, c->st));
    return NS_OK;
}

int ns_get_bc_planes(ns_ctx *c, double *planes) {
    if (!c || !planes) return NS_ERR_INVALID;
    HIPCHK(hipMemcpyAsync(planes, c->planes, (size_t)c->L.total * 8, hipMemcpyDefault, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    return NS_OK;
}

int ns_bc_interp(ns_ctx *c, const double *a, const double *b, double theta) {
    if (!c || !a || !b) return NS_ERR_INVALID;
    ArgList args;
    Arg &A = args.add();
    int rc = arg_in(c, A, a, (size_t)c->L.total, true);
    if (rc) return rc;
    Arg &B = args.add();
    rc = arg_in(c, B, b, (size_t)c->L.total, true);
    if (rc) return rc;
    launch_axpby(A.dev, B.dev, theta, c->planes, c->L.total, c->st);
    HIPCHK(hipStreamSynchronize(c->st));
    return NS_OK;
}

// capture two steady steps on the context stream into graph[parity]; the
// host-side state (u* buffer swaps, flags) is the same before and after
int build_graph(ns_ctx *c, int parity) {
    hipGraph_t gr = nullptr;
    if (hipStreamBeginCapture(c->st, hipStreamCaptureModeThreadLocal) != hipSuccess) {
        (void)hipGetLastError();
        return NS_ERR_HIP;
    }
    int rc = one_step(c);
    if (rc == NS_OK) rc = one_step(c);
    const hipError_t e = hipStreamEndCapture(c->st, &gr);
    if (rc == NS_OK && (e != hipSuccess || !gr)) rc = NS_ERR_HIP;
    if (rc == NS_OK && hipGraphInstantiate(&c->graph[parity], gr, nullptr, nullptr, 0) != hipSuccess) {
        c->graph[parity] = nullptr;
        rc = NS_ERR_HIP;
    }
    if (gr) (void)hipGraphDestroy(gr);
    (void)hipGetLastError();
    return rc;
}

int ns_step(ns_ctx *c, int nsteps) {
    if (!c || nsteps < 0) return NS_ERR_INVALID;
    int s = 0;
    while (s < nsteps) {
        const bool steady = !c->first && !c->cap_on && !c->graph_off && !(c->cfg.flags & NS_FLAG_NO_GRAPH) &&
                            !c->ghost_user;
        if (steady && nsteps - s >= 2) {
            const int parity = (c->us == c->us_base) ? 0 : 1;
            // a failed launch of an eager step before this one is reported, not
            // cleared with the capture's errors
            if (!c->graph[parity] && hipPeekAtLastError() != hipSuccess) return NS_ERR_HIP;
            if (!c->graph[parity] && build_graph(c, parity) != NS_OK) {
                c->graph_off = 1;  // run eagerly from here on
                continue;
            }
            HIPCHK(hipGraphLaunch(c->graph[parity], c->st));
            // host-side effects of the two steps (one_step's, minus the buffer
            // swaps, which cancel): u, v, w, p must be re-formed on demand
            if (c->lazy) c->stale = 1;
            if (skips_star_fill(c)) c->star_fill = 1;
            c->first = 0;
            s += 2;
            continue;
        }
        int rc = one_step(c);
        if (rc) return rc;
        s++;
    }
    if (hipGetLastError() != hipSuccess) return NS_ERR_HIP;
    return NS_OK;
}

int ns_step_profiled(ns_ctx *c, int nsteps, double *phase_ms) {
    if (!c || nsteps < 0 || !phase_ms) return NS_ERR_INVALID;
    hipEvent_t ev[NS_NPHASES + 1];
    for (int i = 0; i <= NS_NPHASES; i++) HIPCHK(hipEventCreate(&ev[i]));
    for (int i = 0; i < NS_NPHASES; i++) phase_ms[i] = 0.0;
    int rc = NS_OK;
    for (int s = 0; s < nsteps && rc == NS_OK; s++) {
        rc = one_step(c, ev);
        if (rc) break;
        if (hipEventSynchronize(ev[NS_NPHASES]) != hipSuccess) { rc = NS_ERR_HIP; break; }
        for (int i = 0; i < NS_NPHASES; i++) {
            float ms = 0.f;
            if (hipEventElapsedTime(&ms, ev[i], ev[i + 1]) != hipSuccess) { rc = NS_ERR_HIP; break; }
            phase_ms[i] += ms;
        }
    }
    for (int i = 0; i <= NS_NPHASES; i++) (void)hipEventDestroy(ev[i]);
    (void)hipGetLastError();
    return rc;
}

// snapshot history on the device: device pointers are used in place, host
// arrays are copied into the context's staging buffer
static int stage_snaps(ns_ctx *c, const double *snapshots, int nsnap, const double **base) {
    const size_t len = (size_t)c->L.total;
    *base = snapshots;
    if (is_dev_ptr(snapshots)) return NS_OK;
    size_t need = len * nsnap;
    if (c->snaps_len < need) {
        if (c->snaps) (void)hipFree(c->snaps);
        c->snaps = nullptr;
        c->snaps_len = 0;
        HIPCHK(hipMalloc(&c->snaps, need * 8));
        c->snaps_len = need;
    }
    HIPCHK(hipMemcpyAsync(c->snaps, snapshots, need * 8, hipMemcpyHostToDevice, c->st));
    *base = c->snaps;
    return NS_OK;
}

// Not-a-knot cubic spline through (t[i], y[i]), i < m, evaluated at x, as a
// linear combination sum_i w[i]*y[i] of the data (MATLAB interp1 'spline').
// Second-derivative form: interior rows
//   h[i-1] M[i-1] + 2(h[i-1]+h[i]) M[i] + h[i] M[i+1] = 6((y[i+1]-y[i])/h[i] - (y[i]-y[i-1])/h[i-1]),
// end rows from the continuity of the third derivative at t[1] and t[m-2].
// m = 3 gives the interpolating parabola, m = 2 the line; x outside
// [t[0], t[m-1]] evaluates the end pieces.
static int spline_weights(const double *t, int m, double x, double *w) {
    if (m < 2 || m > NS_WSUM_MAX) return NS_ERR_INVALID;
    for (int i = 0; i + 1 < m; i++)
        if (!(t[i + 1] > t[i])) return NS_ERR_INVALID;
    typedef long double R;
    if (m == 2) {
        const R th = ((R)x - t[0]) / ((R)t[1] - t[0]);
        w[0] = (double)(1 - th);
        w[1] = (double)th;
        return NS_OK;
    }
    if (m == 3) {
        for (int i = 0; i < 3; i++) {
            R l = 1;
            for (int j = 0; j < 3; j++)
                if (j != i) l *= ((R)x - t[j]) / ((R)t[i] - t[j]);
            w[i] = (double)l;
        }
        return NS_OK;
    }
    R h[NS_WSUM_MAX], A[NS_WSUM_MAX][NS_WSUM_MAX], Bm[NS_WSUM_MAX][NS_WSUM_MAX];
    for (int i = 0; i + 1 < m; i++) h[i] = (R)t[i + 1] - t[i];
    for (int r = 0; r < m; r++)
        for (int c = 0; c < m; c++) A[r][c] = Bm[r][c] = 0;
    // right-hand side matrix: row r of (6 * second differences) applied to unit data
    A[0][0] = h[1];
    A[0][1] = -(h[0] + h[1]);
    A[0][2] = h[0];
    for (int i = 1; i + 1 < m; i++) {
        A[i][i - 1] = h[i - 1];
        A[i][i] = 2 * (h[i - 1] + h[i]);
        A[i][i + 1] = h[i];
        Bm[i][i + 1] += 6 / h[i];
        Bm[i][i] -= 6 / h[i] + 6 / h[i - 1];
        Bm[i][i - 1] += 6 / h[i - 1];
    }
    A[m - 1][m - 3] = h[m - 2];
    A[m - 1][m - 2] = -(h[m - 3] + h[m - 2]);
    A[m - 1][m - 1] = h[m - 3];
    // solve A * Mj = Bm(:, j) for all unit vectors j at once (Gaussian elimination, partial pivoting)
    for (int k = 0; k < m; k++) {
        int piv = k;
        for (int r = k + 1; r < m; r++)
            if (fabsl(A[r][k]) > fabsl(A[piv][k])) piv = r;
        if (A[piv][k] == 0) return NS_ERR_INVALID;
        for (int c = 0; c < m; c++) {
            R ta = A[k][c]; A[k][c] = A[piv][c]; A[piv][c] = ta;
            R tb = Bm[k][c]; Bm[k][c] = Bm[piv][c]; Bm[piv][c] = tb;
        }
        for (int r = 0; r < m; r++) {
            if (r == k || A[r][k] == 0) continue;
            const R f = A[r][k] / A[k][k];
            for (int c = 0; c < m; c++) {
                A[r][c] -= f * A[k][c];
                Bm[r][c] -= f * Bm[k][c];
            }
        }
    }
    // Bm(r, j) / A(r, r) = second derivative at t[r] for unit data e_j
    int k = 0;
    while (k < m - 2 && x >= t[k + 1]) k++;
    const R hk = h[k], a = (R)t[k + 1] - x, b = (R)x - t[k];
    for (int j = 0; j < m; j++) {
        const R Mk = Bm[k][j] / A[k][k], Mk1 = Bm[k + 1][j] / A[k + 1][k + 1];
        const R yk = (j == k) ? 1 : 0, yk1 = (j == k + 1) ? 1 : 0;
        const R s = Mk * a * a * a / (6 * hk) + Mk1 * b * b * b / (6 * hk) + (yk / hk - Mk * hk / 6) * a +
                    (yk1 / hk - Mk1 * hk / 6) * b;
        w[j] = (double)s;
    }
    return NS_OK;
}

// window of the spline interpolation over snapshots at k*snap_dt: six
// snapshots, three at or before t (sub_tstep_BC_combined_interpolation.m:56-57),
// kept inside the history
static int spline_planes(ns_ctx *c, const double *base, int nsnap, double snap_dt, double t) {
    const int m = nsnap < 6 ? nsnap : 6;
    int kle = (int)floor(t / snap_dt);
    if (kle > nsnap - 1) kle = nsnap - 1;
    int s0 = kle - 2;
    if (s0 > nsnap - m) s0 = nsnap - m;
    if (s0 < 0) s0 = 0;
    double tk[NS_WSUM_MAX], w[NS_WSUM_MAX];
    for (int i = 0; i < m; i++) tk[i] = (s0 + i) * snap_dt;
    int rc = spline_weights(tk, m, t, w);
    if (rc) return rc;
    const size_t len = (size_t)c->L.total;
    launch_wsum(base + len * s0, (long)len, w, m, c->planes, (long)len, c->st);
    return NS_OK;
}

int ns_resim_advance(ns_ctx *c, const double *snapshots, int nsnap, double snap_dt, double t0, int nsteps) {
    if (!c || !snapshots || nsnap < 2 || !(snap_dt > 0) || nsteps < 0) return NS_ERR_INVALID;
    const size_t len = (size_t)c->L.total;
    const double *base = nullptr;
    int rc0 = stage_snaps(c, snapshots, nsnap, &base);
    if (rc0) return rc0;
    for (int m = 0; m < nsteps; m++) {
        double t = t0 + (m + 1) * c->cfg.dt;
        int s = (int)floor(t / snap_dt);
        if (s < 0) s = 0;
        if (s > nsnap - 2) s = nsnap - 2;
        double th = (t - s * snap_dt) / snap_dt;
        launch_axpby(base + len * s, base + len * (s + 1), th, c->planes, (long)len, c->st);
        int rc = one_step(c);
        if (rc) return rc;
    }
    if (hipGetLastError() != hipSuccess) return NS_ERR_HIP;
    return NS_OK;
}

int ns_spline_weights(const double *t, int m, double x, double *w) {
    if (!t || !w) return NS_ERR_INVALID;
    return spline_weights(t, m, x, w);
}

int ns_bc_interp_spline(ns_ctx *c, const double *snapshots, int nsnap, double snap_dt, double t) {
    if (!c || !snapshots || nsnap < 2 || !(snap_dt > 0)) return NS_ERR_INVALID;
    const double *base = nullptr;
    int rc = stage_snaps(c, snapshots, nsnap, &base);
    if (rc) return rc;
    rc = spline_planes(c, base, nsnap, snap_dt, t);
    if (rc) return rc;
    return hipGetLastError() == hipSuccess ? NS_OK : NS_ERR_HIP;
}

int ns_resim_advance_spline(ns_ctx *c, const double *snapshots, int nsnap, double snap_dt, double t0, int nsteps) {
    if (!c || !snapshots || nsnap < 2 || !(snap_dt > 0) || nsteps < 0) return NS_ERR_INVALID;
    const double *base = nullptr;
    int rc = stage_snaps(c, snapshots, nsnap, &base);
    if (rc) return rc;
    for (int m = 0; m < nsteps; m++) {
        if ((rc = spline_planes(c, base, nsnap, snap_dt, t0 + (m + 1) * c->cfg.dt))) return rc;
        if ((rc = one_step(c))) return rc;
    }
    return hipGetLastError() == hipSuccess ? NS_OK : NS_ERR_HIP;
}

int ns_set_dt(ns_ctx *c, double dt) {
    if (!c || !(dt > 0)) return NS_ERR_INVALID;
    if (dt == c->cfg.dt) return NS_OK;
    if (c->lazy && c->stale) {
        // the pending projection belongs to the old dt: form u, v, w, p and
        // carry u itself as the next predictor's input (dp = 0)
        materialize(c);
        const Grid &g = c->g;
        HIPCHK(hipMemcpyAsync(c->us, c->u, g.nu() * 8, hipMemcpyDeviceToDevice, c->st));
        HIPCHK(hipMemcpyAsync(c->vs, c->v, g.nv() * 8, hipMemcpyDeviceToDevice, c->st));
        HIPCHK(hipMemcpyAsync(c->ws, c->w, g.nw() * 8, hipMemcpyDeviceToDevice, c->st));
        HIPCHK(hipMemsetAsync(c->dp, 0, g.np() * 8, c->st));
        c->star_fill = 0;
    }
    c->cfg.dt = dt;
    c->cnT_ok = 0;
    // the systems are rebuilt in the tables queued sweeps read
    HIPCHK(hipStreamSynchronize(c->st));
    if (c->cfg.scheme == NS_AB2_CN)
        for (int ax = 0; ax < 3; ax++)
            for (int t = 0; t < 2; t++) {
                int rc = make_line_sys(c, ax, t, c->cn[ax][t]);
                if (rc) return rc;
            }
    // captured graphs hold the old dt
    for (int i = 0; i < 2; i++)
        if (c->graph[i]) {
            (void)hipGraphExecDestroy(c->graph[i]);
            c->graph[i] = nullptr;
        }
    return NS_OK;
}

int ns_history_save(ns_ctx *c) {
    if (!c) return NS_ERR_INVALID;
    const Grid &g = c->g;
    int rc;
    if (!c->hsave[0] && ((rc = dalloc(c, &c->hsave[0], g.nu())) || (rc = dalloc(c, &c->hsave[1], g.nv())) ||
                         (rc = dalloc(c, &c->hsave[2], g.nw()))))
        return rc;
    HIPCHK(hipMemcpyAsync(c->hsave[0], c->hx, g.nu() * 8, hipMemcpyDeviceToDevice, c->st));
    HIPCHK(hipMemcpyAsync(c->hsave[1], c->hy, g.nv() * 8, hipMemcpyDeviceToDevice, c->st));
    HIPCHK(hipMemcpyAsync(c->hsave[2], c->hz, g.nw() * 8, hipMemcpyDeviceToDevice, c->st));
    return NS_OK;
}

int ns_history_restore(ns_ctx *c) {
    if (!c || !c->hsave[0]) return NS_ERR_INVALID;
    const Grid &g = c->g;
    HIPCHK(hipMemcpyAsync(c->hx, c->hsave[0], g.nu() * 8, hipMemcpyDeviceToDevice, c->st));
    HIPCHK(hipMemcpyAsync(c->hy, c->hsave[1], g.nv() * 8, hipMemcpyDeviceToDevice, c->st));
    HIPCHK(hipMemcpyAsync(c->hz, c->hsave[2], g.nw() * 8, hipMemcpyDeviceToDevice, c->st));
    return NS_OK;
}

static int adj_alloc(ns_ctx *c) {
    if (c->ab[0]) return NS_OK;
    const Grid &g = c->gr;
    const size_t n[3] = {g.nu(), g.nv(), g.nw()};
    const bool cn = c->cfg.scheme == NS_AB2_CN;
    int rc;
    // the u* adjoint lives in the u adjoint's arrays: S5^T reads the u adjoint
    // once (into the dp adjoint) and the u* adjoint then starts equal to it
    for (int a = 0; a < 3; a++) {
        if ((rc = dalloc(c, &c->ab[a], n[a])) || (rc = dalloc(c, &c->ahb[a], n[a])) || (rc = dalloc(c, &c->agb[a], n[a])))
            return rc;
        c->ausb[a] = c->ab[a];
        if (cn) {
            if ((rc = dalloc(c, &c->agcb[a], n[a]))) return rc;
        } else {
            c->agcb[a] = c->agb[a];
        }
    }
    if ((rc = dalloc(c, &c->apb, g.np())) || (rc = dalloc(c, &c->adpb, g.np())) || (rc = dalloc(c, &c->ar, g.ni())) ||
        (rc = dalloc(c, &c->arsol, g.ni())) || (rc = dalloc(c, &c->apbar, (size_t)c->L.total)))
        return rc;
    if (!c->mpartial && (rc = dalloc(c, &c->mpartial, (size_t)2048 * NS_METRIC_PARTS))) return rc;
    return NS_OK;
}

// explicit schemes on any codes; AB2-CN with wall velocity codes (2, 4) on
// every axis, whose ADI systems are plain tridiagonal
// The adjoint step works on the materialized state (u, v, w, p, H) that
// ns_download_state returns and ns_upload_state takes, and on the forward
// velocity passed in, never on the context's stored forward state, so lazily
// projected contexts (which store u*, dp between steps) are supported too.
static bool adj_supported(const ns_ctx *c) { return c->poisson_ok; }

// transposed CN systems for the current dt
static int adj_cn_systems(ns_ctx *c) {
    if (c->cfg.scheme != NS_AB2_CN || c->cnT_ok) return NS_OK;
    HIPCHK(hipStreamSynchronize(c->st));  // a rebuild rewrites tables earlier sweeps read
    for (int ax = 0; ax < 3; ax++)
        for (int t = 0; t < 2; t++) {
            const int rc = make_line_sys(c, ax, t, c->cnT[ax][t], true);
            if (rc) return rc;
        }
    c->cnT_ok = 1;
    return NS_OK;
}

int ns_adjoint_set_state(ns_ctx *c, const double *ub, const double *vb, const double *wb, const double *pb,
                         const double *hxb, const double *hyb, const double *hzb) {
    if (!c) return NS_ERR_INVALID;
    if (!adj_supported(c)) return NS_ERR_UNSUPPORTED;
    int rc = adj_alloc(c);
    if (rc) return rc;
    const Grid &g = c->gr;
    const size_t n[7] = {g.nu(), g.nv(), g.nw(), g.np(), g.nu(), g.nv(), g.nw()};
    const double *src[7] = {ub, vb, wb, pb, hxb, hyb, hzb};
    double *dst[7] = {c->ab[0], c->ab[1], c->ab[2], c->apb, c->ahb[0], c->ahb[1], c->ahb[2]};
    for (int i = 0; i < 7; i++) {
        if (src[i]) HIPCHK(hipMemcpyAsync(dst[i], src[i], n[i] * 8, hipMemcpyDefault, c->st));
        else HIPCHK(hipMemsetAsync(dst[i], 0, n[i] * 8, c->st));
    }
    HIPCHK(hipStreamSynchronize(c->st));
    return NS_OK;
}

int ns_adjoint_get_state(ns_ctx *c, double *ub, double *vb, double *wb, double *pb, double *hxb, double *hyb,
                         double *hzb) {
    if (!c) return NS_ERR_INVALID;
    if (!c->ab[0]) return NS_ERR_INVALID;
    const Grid &g = c->gr;
    const size_t n[7] = {g.nu(), g.nv(), g.nw(), g.np(), g.nu(), g.nv(), g.nw()};
    double *dst[7] = {ub, vb, wb, pb, hxb, hyb, hzb};
    const double *src[7] = {c->ab[0], c->ab[1], c->ab[2], c->apb, c->ahb[0], c->ahb[1], c->ahb[2]};
    for (int i = 0; i < 7; i++)
        if (dst[i]) HIPCHK(hipMemcpyAsync(dst[i], src[i], n[i] * 8, hipMemcpyDefault, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    return NS_OK;
}

int ns_adjoint_step(ns_ctx *c, const double *u, const double *v, const double *w, int first, double *planes_bar) {
    if (!c || !u || !v || !w) return NS_ERR_INVALID;
    if (!adj_supported(c)) return NS_ERR_UNSUPPORTED;
    if (!c->ab[0]) return NS_ERR_INVALID;  // ns_adjoint_set_state first
    const ns_config &f = c->cfg;
    const Grid &g = c->gr;
    const double h[3] = {f.dx, f.dy, f.dz};
    int per[3];
    for (int a = 0; a < 3; a++) per[a] = (f.bc[a] == 1);
    ArgList args;
    Arg &AU = args.add();
    Arg &AV = args.add();
    Arg &AW = args.add();
    int rc;
    if ((rc = arg_in(c, AU, u, g.nu(), true)) || (rc = arg_in(c, AV, v, g.nv(), true)) ||
        (rc = arg_in(c, AW, w, g.nw(), true)))
        return rc;
    HIPCHK(hipMemsetAsync(c->apbar, 0, (size_t)c->L.total * 8, c->st));
    // S5^T: projection (and the corner gauge: p = dp_flag p + dp - dp(1,1,1)); the
    // sums stay on the device (block partials, then one block adds them)
    int nbs = 0;
    if (f.flags & NS_FLAG_GAUGE_CORNER) nbs = launch_block_sums(c->apb, g.np(), c->mpartial, 2048, c->st);
    launch_adj_project(g, f.dt, h, f.dp_flag, c->ab, c->ausb, c->apb, c->adpb, c->st);
    if (f.flags & NS_FLAG_GAUGE_CORNER) launch_add_partials(c->adpb, 0, c->mpartial, nbs, -1.0, c->st);
    // S4^T: pressure ghost fill
    double *const dpa[3] = {c->adpb, c->adpb, c->adpb};
    launch_adj_fill(g, 3, f.pbc, h, dpa, c->apbar, c->st);
    // S3^T: zero mode (mean_dp of the DCT solve), symmetric solve, boundary terms, divergence
    launch_adj_take_interior(g, c->adpb, c->ar, c->st);
    // mean_dp sets the zero mode, which exists only when every axis is Neumann
    if (c->pp.dct && f.pbc[0] == 3 && f.pbc[1] == 3 && f.pbc[2] == 3) {
        nbs = launch_block_sums(c->ar, g.ni(), c->mpartial, 2048, c->st);
        launch_add_partials(c->apbar, (size_t)c->L.off[24], c->mpartial, nbs, 1.0, c->st);
    }
    if ((rc = run_poisson(c, nullptr, nullptr, nullptr, c->ar, nullptr, nullptr, 0.0, nullptr, c->arsol))) return rc;
    launch_adj_rhs_div(g, f.pbc, h, f.dt, c->arsol, c->ausb, c->apbar, c->st);
    if (f.scheme != NS_AB2_CN) {
        // S2^T: velocity ghost fill
        launch_adj_fill(g, 0, f.bc, h, c->ausb, c->apbar, c->st);
    } else {
        // S2^T for AB2-CN: the sweeps in reverse, each a solve with A^T, then
        // the transpose of its boundary rows (vel_bc_staggered_CN2)
        if ((rc = adj_cn_systems(c))) return rc;
        for (int dir = 3; dir >= 1; dir--) {
            // a partitioned solve moves the boundary rows' adjoints to the planes itself
            const LineSys &sa = c->cnT[dir - 1][0], &sb = c->cnT[dir - 1][1];
            if (sweep_partitioned(sa, sb)) {
                launch_adi_sweep(g, dir, sa, sb, c->ausb[0], c->ausb[1], c->ausb[2], c->st, nullptr, &c->L,
                                 c->apbar);
            } else {
                launch_adi_sweep(g, dir, sa, sb, c->ausb[0], c->ausb[1], c->ausb[2], c->st);
                launch_adj_cn2(g, dir - 1, f.bc[dir - 1], c->ausb, c->apbar, c->st);
            }
        }
    }
    // S1^T: predictor
    const double *uu[3] = {AU.dev, AV.dev, AW.dev};
    launch_adj_predict(g, per, f.dt, f.nu, f.dp_flag, h, f.scheme, first, uu, c->ausb, c->ahb, c->agb, c->agcb, c->ab,
                       c->apb, c->st);
    if (planes_bar) HIPCHK(hipMemcpyAsync(planes_bar, c->apbar, (size_t)c->L.total * 8, hipMemcpyDefault, c->st));
    // host outputs and staged host inputs need the stream drained; device ones stay stream-ordered
    if ((planes_bar && !is_dev_ptr(planes_bar)) || AU.staged || AV.staged || AW.staged)
        HIPCHK(hipStreamSynchronize(c->st));
    return hipGetLastError() == hipSuccess ? NS_OK : NS_ERR_HIP;
}

int ns_extract_subbox(ns_ctx *c, const ns_subbox *sb, int field, double *dst) {
    if (!c || !sb || !dst || field < NS_FIELD_U || field > NS_FIELD_HZ) return NS_ERR_INVALID;
    const ns_config &f = c->cfg;
    // component (0 u, 1 v, 2 w, 3 p) and array of the field
    const int comp_of[11] = {0, 1, 2, 3, 0, 1, 2, 3, 0, 1, 2};
    const int comp = comp_of[field];
    if (field <= NS_FIELD_P) materialize(c);
    fill_star(c);
    double *src[11] = {c->u, c->v, c->w, c->p, c->us, c->vs, c->ws, c->dp, c->hx, c->hy, c->hz};
    int n[3];
    comp_dims(c->g, comp, n);
    const int nb[3] = {sb->nx, sb->ny, sb->nz};
    int e[3];
    for (int a = 0; a < 3; a++) e[a] = nb[a] + ((comp == a) ? 1 : 2);
    const int o[3] = {sb->x0 - 1, sb->y0 - 1, sb->z0 - 1};
    for (int a = 0; a < 3; a++)
        if (nb[a] < 1 || o[a] < 0 || o[a] + e[a] > n[a]) return NS_ERR_INVALID;
    const double *gauge = (field == NS_FIELD_DP && (f.flags & NS_FLAG_GAUGE_CORNER)) ? c->dp : nullptr;
    ArgList args;
    Arg &D = args.add();
    int rc = arg_in(c, D, dst, (size_t)e[0] * e[1] * e[2], false);
    if (rc) return rc;
    const int pn[3] = {comp_pitch(c->g, comp), n[1], n[2]};  // row pitch and rows of the resident array
    launch_extract_box(src[field], pn, o, e, gauge, D.dev, c->st);
    if ((rc = arg_out(c, D))) return rc;
    HIPCHK(hipStreamSynchronize(c->st));
    return NS_OK;
}

int ns_capture_begin(ns_ctx *c, const ns_subbox *sb, double *dst, int max_sets, int every) {
    if (!c || !sb || !dst || max_sets < 1 || every < 1) return NS_ERR_INVALID;
    const Grid &g = c->g;
    if (sb->nx < 2 || sb->ny < 2 || sb->nz < 2 || sb->x0 < 1 || sb->y0 < 1 || sb->z0 < 1) return NS_ERR_INVALID;
    // the sub-box's ghosted extent must lie inside the parent's ghosted arrays
    if (sb->x0 + sb->nx > g.nx + 1 || sb->y0 + sb->ny > g.ny + 1 || sb->z0 + sb->nz > g.nz + 1)
        return NS_ERR_INVALID;
    if (!is_dev_ptr(dst)) return NS_ERR_INVALID;
    if (!c->cap_partial) {
        int rc = dalloc(c, &c->cap_partial, (size_t)1024);
        if (rc) return rc;
    }
    c->cap_g = ref_grid(sb->nx, sb->ny, sb->nz);
    c->cap_o[0] = sb->x0 - 1;
    c->cap_o[1] = sb->y0 - 1;
    c->cap_o[2] = sb->z0 - 1;
    c->cap_len = plane_layout(sb->nx, sb->ny, sb->nz).total;
    c->cap_dst = dst;
    c->cap_max = max_sets;
    c->cap_every = every;
    c->cap_count = 0;
    c->cap_step = 0;
    c->cap_on = 1;
    HIPCHK(hipMemsetAsync(dst, 0, (size_t)c->cap_len * max_sets * 8, c->st));
    return NS_OK;
}

int ns_capture_end(ns_ctx *c, int *nsets) {
    if (!c) return NS_ERR_INVALID;
    HIPCHK(hipStreamSynchronize(c->st));
    if (nsets) *nsets = c->cap_count;
    c->cap_on = 0;
    c->cap_dst = nullptr;
    return NS_OK;
}

int ns_max_div(ns_ctx *c, double *out) {
    if (!c || !out) return NS_ERR_INVALID;
    materialize(c);
    const ns_config &f = c->cfg;
    int nb = launch_max_div(c->g, f.dx, f.dy, f.dz, c->u, c->v, c->w, c->partial, c->st);
    std::vector<double> h(nb);
    HIPCHK(hipMemcpyAsync(h.data(), c->partial, nb * 8, hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    double m = 0;
    for (double x : h) m = x > m ? x : m;
    *out = m;
    return NS_OK;
}

int ns_error_metrics(ns_ctx *c, const double *u_sub, const double *v_sub, const double *w_sub, const double *p_sub,
                     double *err, double *rms) {
    if (!c || !u_sub || !v_sub || !w_sub || !p_sub || !err || !rms) return NS_ERR_INVALID;
    materialize(c);
    const Grid &g = c->g;
    const int MAXB = 2048;
    if (!c->mpartial) {
        int rc = dalloc(c, &c->mpartial, (size_t)MAXB * NS_METRIC_PARTS);
        if (rc) return rc;
    }
    ArgList args;
    Arg &A0 = args.add();
    Arg &A1 = args.add();
    Arg &A2 = args.add();
    Arg &A3 = args.add();
    int rc;
    if ((rc = arg_in(c, A0, u_sub, g.ref_nu(), true)) || (rc = arg_in(c, A1, v_sub, g.ref_nv(), true)) ||
        (rc = arg_in(c, A2, w_sub, g.ref_nw(), true)) || (rc = arg_in(c, A3, p_sub, g.ref_np(), true)))
        return rc;
    const double *a[4] = {c->u, c->v, c->w, c->p};
    const double *b[4] = {A0.dev, A1.dev, A2.dev, A3.dev};
    const int nb = launch_err_partials(g, c->cfg.dx, c->cfg.dy, a, b, c->mpartial, MAXB, c->st);
    std::vector<double> h((size_t)nb * NS_METRIC_PARTS);
    HIPCHK(hipMemcpyAsync(h.data(), c->mpartial, h.size() * 8, hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    double mx[5] = {0, 0, 0, 0, 0};
    long double sabs[4] = {0, 0, 0, 0}, ssq[5] = {0, 0, 0, 0, 0};
    for (int blk = 0; blk < nb; blk++) {
        const double *q = &h[(size_t)blk * NS_METRIC_PARTS];
        for (int cc = 0; cc < 4; cc++) {
            mx[cc] = q[3 * cc] > mx[cc] ? q[3 * cc] : mx[cc];
            sabs[cc] += q[3 * cc + 1];
            ssq[cc] += q[3 * cc + 2];
        }
        mx[4] = q[12] > mx[4] ? q[12] : mx[4];
        ssq[4] += q[13];
    }
    const double n[4] = {(double)g.ref_nu(), (double)g.ref_nv(), (double)g.ref_nw(), (double)g.ref_np()};
    const double noz = (double)((size_t)(g.nx + 1) * (g.ny + 1) * (g.nz + 2));
    for (int cc = 0; cc < 4; cc++) {
        rms[cc] = sqrt((double)(ssq[cc] / n[cc]));
        err[cc] = mx[cc] / rms[cc];
    }
    for (int cc = 0; cc < 3; cc++) err[4 + cc] = (double)(sabs[cc] / n[cc]) / rms[cc];
    err[7] = mx[4] / sqrt((double)(ssq[4] / noz));
    return NS_OK;
}

int ns_error_metrics_noalloc(ns_ctx *c, const double *u_sub, const double *v_sub, const double *w_sub,
                             const double *p_sub, double *err, double *rms) {
    if (!c || !u_sub || !v_sub || !w_sub || !p_sub || !err || !rms) return NS_ERR_INVALID;
    // device arrays only: staging a host array would allocate
    if (!is_dev_ptr(u_sub) || !is_dev_ptr(v_sub) || !is_dev_ptr(w_sub) || !is_dev_ptr(p_sub)) return NS_ERR_INVALID;
    int dev = -1;
    HIPCHK(hipGetDevice(&dev));
    if (dev != c->dev) return NS_ERR_INVALID;
    double *scr = nullptr;  // this device's copy of the static scratch
    int rc = metric_scratch(&scr);
    if (rc) return rc;
    materialize(c);
    const Grid &g = c->g;
    const double *a[4] = {c->u, c->v, c->w, c->p};
    const double *b[4] = {u_sub, v_sub, w_sub, p_sub};
    // the partial kernel walks the arrays grid-stride, so every box fits in
    // NOALLOC_METRIC_BLOCKS rows; the check guards the scratch's extent
    const int nb = launch_err_partials(g, c->cfg.dx, c->cfg.dy, a, b, scr, NOALLOC_METRIC_BLOCKS, c->st);
    if (nb < 1 || nb > NOALLOC_METRIC_BLOCKS) return NS_ERR_INVALID;
    double *row = scr + (size_t)NOALLOC_METRIC_BLOCKS * NS_METRIC_PARTS;
    launch_err_combine(scr, nb, row, c->st);
    double q[NS_METRIC_PARTS];
    HIPCHK(hipMemcpyAsync(q, row, sizeof(q), hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    const double n[4] = {(double)g.ref_nu(), (double)g.ref_nv(), (double)g.ref_nw(), (double)g.ref_np()};
    const double noz = (double)((size_t)(g.nx + 1) * (g.ny + 1) * (g.nz + 2));
    for (int cc = 0; cc < 4; cc++) {
        rms[cc] = sqrt(q[3 * cc + 2] / n[cc]);
        err[cc] = q[3 * cc] / rms[cc];
    }
    for (int cc = 0; cc < 3; cc++) err[4 + cc] = q[3 * cc + 1] / n[cc] / rms[cc];
    err[7] = q[12] / sqrt(q[13] / noz);
    return NS_OK;
}

int ns_stream_copy(ns_ctx *c, const double *src, double *dst, long n) {
    if (!c || !src || !dst || n < 0) return NS_ERR_INVALID;
    if (!is_dev_ptr(src) || !is_dev_ptr(dst)) return NS_ERR_INVALID;
    launch_stream_copy(src, dst, n, c->st);
    HIPCHK(hipStreamSynchronize(c->st));
    return NS_OK;
}

int ns_check_finite(ns_ctx *c) {
    if (!c) return NS_ERR_INVALID;
    materialize(c);
    HIPCHK(hipMemsetAsync(c->counter, 0, sizeof(unsigned int), c->st));
    launch_count_nonfinite(c->u, c->g.nu(), c->counter, c->st);
    launch_count_nonfinite(c->v, c->g.nv(), c->counter, c->st);
    launch_count_nonfinite(c->w, c->g.nw(), c->counter, c->st);
    launch_count_nonfinite(c->p, c->g.np(), c->counter, c->st);
    unsigned int h = 0;
    HIPCHK(hipMemcpyAsync(&h, c->counter, sizeof(h), hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    return h ? NS_ERR_NONFINITE : NS_OK;
}

// ---------------- mirrors ----------------

int ns_cal_conv(ns_ctx *c, const double *u, const double *v, const double *w, int bc_x, int bc_y, int bc_z, double dx,
                double dy, double dz, double *conv_x, double *conv_y, double *conv_z) {
    if (!c) return NS_ERR_INVALID;
    const Grid &g = c->gr;
    int bc[3] = {bc_x, bc_y, bc_z};
    for (int a = 0; a < 3; a++)
        if (bc[a] < 1 || bc[a] > 4) return NS_ERR_INVALID;
    size_t m[3];
    for (int k = 0; k < 3; k++) {
        size_t s = g.ni();
        int n = (k == 0) ? g.nx : (k == 1) ? g.ny : g.nz;
        m[k] = (bc[k] == 1) ? s : s / n * (n - 1);
    }
    ArgList A;
    int rc;
    Arg &au = A.add(); if ((rc = arg_in(c, au, u, g.nu(), true))) return rc;
    Arg &av = A.add(); if ((rc = arg_in(c, av, v, g.nv(), true))) return rc;
    Arg &aw = A.add(); if ((rc = arg_in(c, aw, w, g.nw(), true))) return rc;
    Arg &ox = A.add(); if ((rc = arg_in(c, ox, conv_x, m[0], false))) return rc;
    Arg &oy = A.add(); if ((rc = arg_in(c, oy, conv_y, m[1], false))) return rc;
    Arg &oz = A.add(); if ((rc = arg_in(c, oz, conv_z, m[2], false))) return rc;
    launch_conv_diff(g, bc, dx, dy, dz, 0, au.dev, av.dev, aw.dev, ox.dev, oy.dev, oz.dev, c->st);
    if ((rc = arg_out(c, ox)) || (rc = arg_out(c, oy)) || (rc = arg_out(c, oz))) return rc;
    HIPCHK(hipStreamSynchronize(c->st));
    return NS_OK;
}

int ns_cal_diff(ns_ctx *c, const double *u, const double *v, const double *w, int bc_x, int bc_y, int bc_z, double dx2,
                double dy2, double dz2, double *diff_x, double *diff_y, double *diff_z) {
    if (!c) return NS_ERR_INVALID;
    const Grid &g = c->gr;
    int bc[3] = {bc_x, bc_y, bc_z};
    for (int a = 0; a < 3; a++)
        if (bc[a] < 1 || bc[a] > 4) return NS_ERR_INVALID;
    size_t m[3];
    for (int k = 0; k < 3; k++) {
        size_t s = g.ni();
        int n = (k == 0) ? g.nx : (k == 1) ? g.ny : g.nz;
        m[k] = (bc[k] == 1) ? s : s / n * (n - 1);
    }
    ArgList A;
    int rc;
    Arg &au = A.add(); if ((rc = arg_in(c, au, u, g.nu(), true))) return rc;
    Arg &av = A.add(); if ((rc = arg_in(c, av, v, g.nv(), true))) return rc;
    Arg &aw = A.add(); if ((rc = arg_in(c, aw, w, g.nw(), true))) return rc;
    Arg &ox = A.add(); if ((rc = arg_in(c, ox, diff_x, m[0], false))) return rc;
    Arg &oy = A.add(); if ((rc = arg_in(c, oy, diff_y, m[1], false))) return rc;
    Arg &oz = A.add(); if ((rc = arg_in(c, oz, diff_z, m[2], false))) return rc;
    launch_conv_diff(g, bc, dx2, dy2, dz2, 1, au.dev, av.dev, aw.dev, ox.dev, oy.dev, oz.dev, c->st);
    if ((rc = arg_out(c, ox)) || (rc = arg_out(c, oy)) || (rc = arg_out(c, oz))) return rc;
    HIPCHK(hipStreamSynchronize(c->st));
    return NS_OK;
}

int ns_vel_bc_staggered(ns_ctx *c, double *u_star, double *v_star, double *w_star, const double *planes, int bc_x,
                        int bc_y, int bc_z) {
    if (!c) return NS_ERR_INVALID;
    const Grid &g = c->gr;
    int bc[3] = {bc_x, bc_y, bc_z};
    for (int a = 0; a < 3; a++) {
        if (bc[a] < 1 || bc[a] > 4) return NS_ERR_INVALID;
        if (bc[a] != 1 && bc[a] != 3 && !planes) return NS_ERR_INVALID;
    }
    ArgList A;
    int rc;
    Arg &au = A.add(); if ((rc = arg_in(c, au, u_star, g.nu(), true))) return rc;
    Arg &av = A.add(); if ((rc = arg_in(c, av, v_star, g.nv(), true))) return rc;
    Arg &aw = A.add(); if ((rc = arg_in(c, aw, w_star, g.nw(), true))) return rc;
    const double *pl = nullptr;
    if (planes) {
        Arg &ap = A.add();
        if ((rc = arg_in(c, ap, planes, (size_t)c->L.total, true))) return rc;
        pl = ap.dev;
    }
    launch_vel_bc(g, bc, pl, au.dev, av.dev, aw.dev, c->st);
    if ((rc = arg_out(c, au)) || (rc = arg_out(c, av)) || (rc = arg_out(c, aw))) return rc;
    HIPCHK(hipStreamSynchronize(c->st));
    return NS_OK;
}

int ns_vel_bc_staggered_cn2(ns_ctx *c, double *u_star, double *v_star, double *w_star, const double *planes, int bc,
                            int dir) {
    if (!c || dir < 1 || dir > 3 || bc < 1 || bc > 4) return NS_ERR_INVALID;
    if ((bc == 2 || bc == 4) && !planes) return NS_ERR_INVALID;
    const Grid &g = c->gr;
    ArgList A;
    int rc;
    Arg &au = A.add(); if ((rc = arg_in(c, au, u_star, g.nu(), true))) return rc;
    Arg &av = A.add(); if ((rc = arg_in(c, av, v_star, g.nv(), true))) return rc;
    Arg &aw = A.add(); if ((rc = arg_in(c, aw, w_star, g.nw(), true))) return rc;
    const double *pl = nullptr;
    if (planes) {
        Arg &ap = A.add();
        if ((rc = arg_in(c, ap, planes, (size_t)c->L.total, true))) return rc;
        pl = ap.dev;
    }
    launch_vel_bc_cn2(g, bc, dir, pl, au.dev, av.dev, aw.dev, c->st);
    if ((rc = arg_out(c, au)) || (rc = arg_out(c, av)) || (rc = arg_out(c, aw))) return rc;
    HIPCHK(hipStreamSynchronize(c->st));
    return NS_OK;
}

int ns_adi_sweep(ns_ctx *c, double *u_star, double *v_star, double *w_star, int dir) {
    if (!c || dir < 1 || dir > 3) return NS_ERR_INVALID;
    if (c->cfg.scheme != NS_AB2_CN) return NS_ERR_UNSUPPORTED;
    const Grid &g = c->gr;
    ArgList A;
    int rc;
    Arg &au = A.add(); if ((rc = arg_in(c, au, u_star, g.nu(), true))) return rc;
    Arg &av = A.add(); if ((rc = arg_in(c, av, v_star, g.nv(), true))) return rc;
    Arg &aw = A.add(); if ((rc = arg_in(c, aw, w_star, g.nw(), true))) return rc;
    launch_adi_sweep(g, dir, c->cn[dir - 1][0], c->cn[dir - 1][1], au.dev, av.dev, aw.dev, c->st);
    if ((rc = arg_out(c, au)) || (rc = arg_out(c, av)) || (rc = arg_out(c, aw))) return rc;
    HIPCHK(hipStreamSynchronize(c->st));
    return NS_OK;
}

int ns_get_velpr_bc(ns_ctx *c, const double *u_bc, const double *v_bc, const double *w_bc, const double *p_bc, double dt,
                    int dp_flag, int using_ustar, const double *p_previous, double *planes) {
    if (!c || !planes) return NS_ERR_INVALID;
    const Grid &g = c->gr;
    const ns_config &f = c->cfg;
    const int corr = !(dp_flag || using_ustar);
    ArgList A;
    int rc;
    Arg &au = A.add(); if ((rc = arg_in(c, au, u_bc, g.nu(), true))) return rc;
    Arg &av = A.add(); if ((rc = arg_in(c, av, v_bc, g.nv(), true))) return rc;
    Arg &aw = A.add(); if ((rc = arg_in(c, aw, w_bc, g.nw(), true))) return rc;
    Arg &ap = A.add(); if ((rc = arg_in(c, ap, p_bc, g.np(), true))) return rc;
    const double *pprev = nullptr;
    if (corr) {
        Arg &aq = A.add();
        if ((rc = arg_in(c, aq, p_previous, g.np(), true))) return rc;
        pprev = aq.dev;
    }
    Arg &ao = A.add(); if ((rc = arg_in(c, ao, planes, (size_t)c->L.total, true))) return rc;
    launch_get_velpr_bc(g, f.bc, f.pbc, f.dx, f.dy, f.dz, dt, corr, au.dev, av.dev, aw.dev, ap.dev, pprev, ao.dev,
                        c->st);
    if ((rc = arg_out(c, ao))) return rc;
    HIPCHK(hipStreamSynchronize(c->st));
    return NS_OK;
}

int ns_get_pr_bc(ns_ctx *c, const double *p_bc, double *planes) {
    if (!c || !planes) return NS_ERR_INVALID;
    const Grid &g = c->gr;
    const ns_config &f = c->cfg;
    ArgList A;
    int rc;
    Arg &ap = A.add(); if ((rc = arg_in(c, ap, p_bc, g.np(), true))) return rc;
    Arg &ao = A.add(); if ((rc = arg_in(c, ao, planes, (size_t)c->L.total, true))) return rc;
    launch_get_pr_bc(g, f.pbc, f.dx, f.dy, f.dz, ap.dev, ao.dev, c->st);
    if ((rc = arg_out(c, ao))) return rc;
    HIPCHK(hipStreamSynchronize(c->st));
    return NS_OK;
}

int ns_divergence(ns_ctx *c, const double *u_star, const double *v_star, const double *w_star, double dt, double *rhs) {
    if (!c) return NS_ERR_INVALID;
    const Grid &g = c->gr;
    const ns_config &f = c->cfg;
    ArgList A;
    int rc;
    Arg &au = A.add(); if ((rc = arg_in(c, au, u_star, g.nu(), true))) return rc;
    Arg &av = A.add(); if ((rc = arg_in(c, av, v_star, g.nv(), true))) return rc;
    Arg &aw = A.add(); if ((rc = arg_in(c, aw, w_star, g.nw(), true))) return rc;
    Arg &ao = A.add(); if ((rc = arg_in(c, ao, rhs, g.ni(), false))) return rc;
    launch_divergence(g, f.dx, f.dy, f.dz, dt, au.dev, av.dev, aw.dev, ao.dev, c->st);
    if ((rc = arg_out(c, ao))) return rc;
    HIPCHK(hipStreamSynchronize(c->st));
    return NS_OK;
}

int ns_pr_bc_staggered_modify_rhs(ns_ctx *c, double *rhs, const double *planes) {
    if (!c || !planes) return NS_ERR_INVALID;
    const Grid &g = c->gr;
    const ns_config &f = c->cfg;
    ArgList A;
    int rc;
    Arg &ar = A.add(); if ((rc = arg_in(c, ar, rhs, g.ni(), true))) return rc;
    Arg &ap = A.add(); if ((rc = arg_in(c, ap, planes, (size_t)c->L.total, true))) return rc;
    launch_modify_rhs(g, f.pbc, ap.dev, f.dx, f.dy, f.dz, ar.dev, c->st);
    if ((rc = arg_out(c, ar))) return rc;
    HIPCHK(hipStreamSynchronize(c->st));
    return NS_OK;
}

int ns_poisson_fft(ns_ctx *c, const double *rhs, double *sol) {
    if (!c) return NS_ERR_INVALID;
    if (!c->poisson_ok || c->pp.dct) return NS_ERR_UNSUPPORTED;
    const Grid &g = c->gr;
    const ns_config &f = c->cfg;
    ArgList A;
    int rc;
    Arg &ar = A.add(); if ((rc = arg_in(c, ar, rhs, g.ni(), true))) return rc;
    Arg &ao = A.add(); if ((rc = arg_in(c, ao, sol, g.ni(), false))) return rc;
    poisson_fft_solve(g, c->pp, f.dx, f.dy, f.dz, f.dt, nullptr, nullptr, nullptr, ar.dev, c->spec, nullptr, ao.dev,
                      c->st);
    if ((rc = arg_out(c, ao))) return rc;
    HIPCHK(hipStreamSynchronize(c->st));
    return NS_OK;
}

int ns_dct_poisson_solver(ns_ctx *c, double *rhs, double mean_dp) {
    if (!c) return NS_ERR_INVALID;
    if (!c->poisson_ok || !c->pp.dct) return NS_ERR_UNSUPPORTED;
    const Grid &g = c->gr;
    const ns_config &f = c->cfg;
    ArgList A;
    int rc;
    Arg &ar = A.add(); if ((rc = arg_in(c, ar, rhs, g.ni(), true))) return rc;
    poisson_dct_solve(g, c->pp, f.dx, f.dy, f.dz, f.dt, nullptr, nullptr, nullptr, ar.dev, nullptr, nullptr, mean_dp,
                      c->work, nullptr, ar.dev, c->st);
    if ((rc = arg_out(c, ar))) return rc;
    HIPCHK(hipStreamSynchronize(c->st));
    return NS_OK;
}

int ns_pr_bc_staggered(ns_ctx *c, double *dp, const double *planes) {
    if (!c) return NS_ERR_INVALID;
    const Grid &g = c->gr;
    const ns_config &f = c->cfg;
    bool need = f.pbc[0] != 1 || f.pbc[1] != 1 || f.pbc[2] != 1;
    if (need && !planes) return NS_ERR_INVALID;
    ArgList A;
    int rc;
    Arg &ad = A.add(); if ((rc = arg_in(c, ad, dp, g.np(), true))) return rc;
    const double *pl = c->planes;
    if (planes) {
        Arg &ap = A.add();
        if ((rc = arg_in(c, ap, planes, (size_t)c->L.total, true))) return rc;
        pl = ap.dev;
    }
    launch_pr_bc(g, f.pbc, pl, f.dx, f.dy, f.dz, ad.dev, c->st);
    if ((rc = arg_out(c, ad))) return rc;
    HIPCHK(hipStreamSynchronize(c->st));
    return NS_OK;
}

int ns_project(ns_ctx *c, const double *u_star, const double *v_star, const double *w_star, const double *dp, double dt,
               int dp_flag, double *u, double *v, double *w, double *p) {
    if (!c) return NS_ERR_INVALID;
    const Grid &g = c->gr;
    const ns_config &f = c->cfg;
    ArgList A;
    int rc;
    Arg &a1 = A.add(); if ((rc = arg_in(c, a1, u_star, g.nu(), true))) return rc;
    Arg &a2 = A.add(); if ((rc = arg_in(c, a2, v_star, g.nv(), true))) return rc;
    Arg &a3 = A.add(); if ((rc = arg_in(c, a3, w_star, g.nw(), true))) return rc;
    Arg &a4 = A.add(); if ((rc = arg_in(c, a4, dp, g.np(), true))) return rc;
    Arg &o1 = A.add(); if ((rc = arg_in(c, o1, u, g.nu(), false))) return rc;
    Arg &o2 = A.add(); if ((rc = arg_in(c, o2, v, g.nv(), false))) return rc;
    Arg &o3 = A.add(); if ((rc = arg_in(c, o3, w, g.nw(), false))) return rc;
    Arg &o4 = A.add(); if ((rc = arg_in(c, o4, p, g.np(), true))) return rc;
    launch_project(g, f.dx, f.dy, f.dz, dt, dp_flag, a1.dev, a2.dev, a3.dev, a4.dev, nullptr, o1.dev, o2.dev, o3.dev,
                   o4.dev, c->st);
    if ((rc = arg_out(c, o1)) || (rc = arg_out(c, o2)) || (rc = arg_out(c, o3)) || (rc = arg_out(c, o4))) return rc;
    HIPCHK(hipStreamSynchronize(c->st));
    return NS_OK;
}

// ---- adjoint mirrors: the transpose of each linear mirror above ----

namespace {
// the plane-set adjoint staged in and out (added to)
int stage_planes_b(ns_ctx *c, ArgList &A, double *planes_b, double **dev) {
    Arg &ap = A.add();
    int rc = arg_in(c, ap, planes_b, (size_t)c->L.total, true);
    *dev = ap.dev;
    return rc;
}
}  // namespace

int ns_vel_bc_staggered_adj(ns_ctx *c, double *u_b, double *v_b, double *w_b, double *planes_b, int bc_x, int bc_y,
                            int bc_z) {
    if (!c) return NS_ERR_INVALID;
    const Grid &g = c->gr;
    const ns_config &f = c->cfg;
    int bc[3] = {bc_x, bc_y, bc_z};
    for (int a = 0; a < 3; a++) {
        if (bc[a] < 1 || bc[a] > 4) return NS_ERR_INVALID;
        if (bc[a] != 1 && bc[a] != 3 && !planes_b) return NS_ERR_INVALID;
    }
    const double h[3] = {f.dx, f.dy, f.dz};
    ArgList A;
    int rc;
    Arg &au = A.add(); if ((rc = arg_in(c, au, u_b, g.nu(), true))) return rc;
    Arg &av = A.add(); if ((rc = arg_in(c, av, v_b, g.nv(), true))) return rc;
    Arg &aw = A.add(); if ((rc = arg_in(c, aw, w_b, g.nw(), true))) return rc;
    double *pb = nullptr;
    if (planes_b && (rc = stage_planes_b(c, A, planes_b, &pb))) return rc;
    double *const a3[3] = {au.dev, av.dev, aw.dev};
    launch_adj_fill(g, 0, bc, h, a3, pb, c->st);
    if ((rc = arg_out(c, au)) || (rc = arg_out(c, av)) || (rc = arg_out(c, aw))) return rc;
    if (planes_b && (rc = arg_out(c, A.back()))) return rc;
    HIPCHK(hipStreamSynchronize(c->st));
    return NS_OK;
}

int ns_vel_bc_staggered_cn2_adj(ns_ctx *c, double *u_b, double *v_b, double *w_b, double *planes_b, int bc,
                                int dir) {
    if (!c || dir < 1 || dir > 3 || bc < 1 || bc > 4) return NS_ERR_INVALID;
    if (bc == 3) return NS_OK;  // the forward leaves the lines as they are
    if (!planes_b) return NS_ERR_INVALID;
    const Grid &g = c->gr;
    ArgList A;
    int rc;
    Arg &au = A.add(); if ((rc = arg_in(c, au, u_b, g.nu(), true))) return rc;
    Arg &av = A.add(); if ((rc = arg_in(c, av, v_b, g.nv(), true))) return rc;
    Arg &aw = A.add(); if ((rc = arg_in(c, aw, w_b, g.nw(), true))) return rc;
    double *pb = nullptr;
    if ((rc = stage_planes_b(c, A, planes_b, &pb))) return rc;
    double *const a3[3] = {au.dev, av.dev, aw.dev};
    launch_adj_cn2(g, dir - 1, bc, a3, pb, c->st);
    if ((rc = arg_out(c, au)) || (rc = arg_out(c, av)) || (rc = arg_out(c, aw)) || (rc = arg_out(c, A.back())))
        return rc;
    HIPCHK(hipStreamSynchronize(c->st));
    return NS_OK;
}

int ns_adi_sweep_adj(ns_ctx *c, double *u_b, double *v_b, double *w_b, int dir) {
    if (!c || dir < 1 || dir > 3) return NS_ERR_INVALID;
    if (c->cfg.scheme != NS_AB2_CN) return NS_ERR_UNSUPPORTED;
    int rc = adj_cn_systems(c);
    if (rc) return rc;
    const Grid &g = c->gr;
    ArgList A;
    Arg &au = A.add(); if ((rc = arg_in(c, au, u_b, g.nu(), true))) return rc;
    Arg &av = A.add(); if ((rc = arg_in(c, av, v_b, g.nv(), true))) return rc;
    Arg &aw = A.add(); if ((rc = arg_in(c, aw, w_b, g.nw(), true))) return rc;
    launch_adi_sweep(g, dir, c->cnT[dir - 1][0], c->cnT[dir - 1][1], au.dev, av.dev, aw.dev, c->st);
    if ((rc = arg_out(c, au)) || (rc = arg_out(c, av)) || (rc = arg_out(c, aw))) return rc;
    HIPCHK(hipStreamSynchronize(c->st));
    return NS_OK;
}

int ns_divergence_adj(ns_ctx *c, const double *rhs_b, double dt, double *u_b, double *v_b, double *w_b) {
    if (!c || !rhs_b) return NS_ERR_INVALID;
    const Grid &g = c->gr;
    const ns_config &f = c->cfg;
    const double h[3] = {f.dx, f.dy, f.dz};
    ArgList A;
    int rc;
    Arg &ar = A.add(); if ((rc = arg_in(c, ar, rhs_b, g.ni(), true))) return rc;
    Arg &au = A.add(); if ((rc = arg_in(c, au, u_b, g.nu(), true))) return rc;
    Arg &av = A.add(); if ((rc = arg_in(c, av, v_b, g.nv(), true))) return rc;
    Arg &aw = A.add(); if ((rc = arg_in(c, aw, w_b, g.nw(), true))) return rc;
    double *const a3[3] = {au.dev, av.dev, aw.dev};
    launch_adj_div(g, h, dt, ar.dev, a3, c->st);
    if ((rc = arg_out(c, au)) || (rc = arg_out(c, av)) || (rc = arg_out(c, aw))) return rc;
    HIPCHK(hipStreamSynchronize(c->st));
    return NS_OK;
}

int ns_pr_bc_staggered_modify_rhs_adj(ns_ctx *c, const double *rhs_b, double *planes_b) {
    if (!c || !rhs_b || !planes_b) return NS_ERR_INVALID;
    const Grid &g = c->gr;
    const ns_config &f = c->cfg;
    const double h[3] = {f.dx, f.dy, f.dz};
    ArgList A;
    int rc;
    Arg &ar = A.add(); if ((rc = arg_in(c, ar, rhs_b, g.ni(), true))) return rc;
    double *pb = nullptr;
    if ((rc = stage_planes_b(c, A, planes_b, &pb))) return rc;
    launch_adj_modify_rhs(g, f.pbc, h, ar.dev, pb, c->st);
    if ((rc = arg_out(c, A.back()))) return rc;
    HIPCHK(hipStreamSynchronize(c->st));
    return NS_OK;
}

int ns_poisson_fft_adj(ns_ctx *c, const double *sol_b, double *rhs_b) {
    // the periodic solve (modal division, zero mode to 0) is symmetric
    return ns_poisson_fft(c, sol_b, rhs_b);
}

int ns_dct_poisson_solver_adj(ns_ctx *c, double *rhs_b, double *mean_dp_b) {
    if (!c || !rhs_b) return NS_ERR_INVALID;
    if (!c->poisson_ok || !c->pp.dct) return NS_ERR_UNSUPPORTED;
    const Grid &g = c->gr;
    const ns_config &f = c->cfg;
    if (!c->mpartial) {
        int rc = dalloc(c, &c->mpartial, (size_t)2048 * NS_METRIC_PARTS);
        if (rc) return rc;
    }
    ArgList A;
    int rc;
    Arg &ar = A.add(); if ((rc = arg_in(c, ar, rhs_b, g.ni(), true))) return rc;
    // mean_dp sets the zero mode (every pressure code Neumann): its adjoint is the sum of the input
    const bool zero_mode = f.pbc[0] == 3 && f.pbc[1] == 3 && f.pbc[2] == 3;
    double sum = 0.0;
    if (zero_mode && mean_dp_b) {
        const int nbs = launch_block_sums(ar.dev, g.ni(), c->mpartial, 2048, c->st);
        HIPCHK(hipMemsetAsync(c->mpartial + 2048, 0, sizeof(double), c->st));
        launch_add_partials(c->mpartial + 2048, 0, c->mpartial, nbs, 1.0, c->st);
        HIPCHK(hipMemcpyAsync(&sum, c->mpartial + 2048, sizeof(double), hipMemcpyDeviceToHost, c->st));
    }
    // the solve with the zero mode dropped is symmetric
    poisson_dct_solve(g, c->pp, f.dx, f.dy, f.dz, f.dt, nullptr, nullptr, nullptr, ar.dev, nullptr, nullptr, 0.0,
                      c->work, nullptr, ar.dev, c->st);
    if ((rc = arg_out(c, ar))) return rc;
    HIPCHK(hipStreamSynchronize(c->st));
    if (mean_dp_b) *mean_dp_b += sum;
    return NS_OK;
}

int ns_pr_bc_staggered_adj(ns_ctx *c, double *dp_b, double *planes_b) {
    if (!c) return NS_ERR_INVALID;
    const Grid &g = c->gr;
    const ns_config &f = c->cfg;
    const bool need = f.pbc[0] != 1 || f.pbc[1] != 1 || f.pbc[2] != 1;
    if (need && !planes_b) return NS_ERR_INVALID;
    const double h[3] = {f.dx, f.dy, f.dz};
    ArgList A;
    int rc;
    Arg &ad = A.add(); if ((rc = arg_in(c, ad, dp_b, g.np(), true))) return rc;
    double *pb = nullptr;
    if (planes_b && (rc = stage_planes_b(c, A, planes_b, &pb))) return rc;
    double *const a3[3] = {ad.dev, ad.dev, ad.dev};
    launch_adj_fill(g, 3, f.pbc, h, a3, pb, c->st);
    if ((rc = arg_out(c, ad))) return rc;
    if (planes_b && (rc = arg_out(c, A.back()))) return rc;
    HIPCHK(hipStreamSynchronize(c->st));
    return NS_OK;
}

int ns_project_adj(ns_ctx *c, const double *u_b, const double *v_b, const double *w_b, double *p_b, double dt,
                   int dp_flag, double *us_b, double *vs_b, double *ws_b, double *dp_b) {
    if (!c) return NS_ERR_INVALID;
    const Grid &g = c->gr;
    const ns_config &f = c->cfg;
    const double h[3] = {f.dx, f.dy, f.dz};
    ArgList A;
    int rc;
    Arg &a1 = A.add(); if ((rc = arg_in(c, a1, u_b, g.nu(), true))) return rc;
    Arg &a2 = A.add(); if ((rc = arg_in(c, a2, v_b, g.nv(), true))) return rc;
    Arg &a3 = A.add(); if ((rc = arg_in(c, a3, w_b, g.nw(), true))) return rc;
    Arg &a4 = A.add(); if ((rc = arg_in(c, a4, p_b, g.np(), true))) return rc;
    Arg &o1 = A.add(); if ((rc = arg_in(c, o1, us_b, g.nu(), false))) return rc;
    Arg &o2 = A.add(); if ((rc = arg_in(c, o2, vs_b, g.nv(), false))) return rc;
    Arg &o3 = A.add(); if ((rc = arg_in(c, o3, ws_b, g.nw(), false))) return rc;
    Arg &o4 = A.add(); if ((rc = arg_in(c, o4, dp_b, g.np(), false))) return rc;
    const double *ub[3] = {a1.dev, a2.dev, a3.dev};
    double *const usb[3] = {o1.dev, o2.dev, o3.dev};
    launch_adj_project(g, dt, h, dp_flag, ub, usb, a4.dev, o4.dev, c->st);
    if ((rc = arg_out(c, a4)) || (rc = arg_out(c, o1)) || (rc = arg_out(c, o2)) || (rc = arg_out(c, o3)) ||
        (rc = arg_out(c, o4)))
        return rc;
    HIPCHK(hipStreamSynchronize(c->st));
    return NS_OK;
}

}  // extern "C"

// ---- batched steps of many contexts: one launch per kernel for all boxes ----

struct ns_batch {
    std::vector<ns_ctx *> ctx;
    hipStream_t st = nullptr;
    // one launch = per-box parameter array + block prefix (device), block total
    struct Phase {
        int kind = 0;    // 0 predictor, 1 velocity fill, 2 DCT pass, 3 pressure fill, 4 projection,
                         // 5 AB2-CN boundary rows, 6 ADI sweep
        int arg = 0;     // axis or pass
        void *params = nullptr;
        int *prefix = nullptr;
        int nblk = 0;
        size_t lds = 0;
        int variant = 0;  // DCT pass: kernel variant (OR over the boxes); sweep: waves per block
        int rows = 0;     // sweep: rows per lane of the partitioned kernel (the most over the boxes)
        int dense = 1;    // sweep: every box of the launch keeps the dense layout
        int nbox = 0;     // boxes of this launch (a subset for sweeps grouped by blocks per line)
    };
    // one steady (non-first) step per parity of the boxes' u* buffers: with
    // the projection folded into the predictor (lazy) each step reads one
    // buffer and writes the other, so the tables of consecutive steps alternate
    std::vector<Phase> phases[2];
    hipGraphExec_t graph[2] = {nullptr, nullptr};
    int lazy = 0;
    std::vector<double *> sig;  // each box's current u* when phases[0] was built
    // launch-shape choices per size class of the boxes (cells < 2^17, < 2^20,
    // larger): predictor z chunks sized for `rounds` rounds of resident
    // blocks, DCT lines per block reduced by 2^dct_drop, 2^adix_lsh lines per
    // x-sweep block.  None of them changes any value computed.
    // Predictor tile per class: 0 each box's shape with fewer idle lanes, or 64 / 32 wide.
    struct Knobs {
        int rounds = 8, dct_drop = 0, adix_lsh = 6, tile = 0;
    };
    Knobs knobs[3];
    std::vector<int> cls;  // size class of each box
    // Boxes dealt into groups of about equal cell counts, each a batch of its
    // own with its own stream, tables and graphs; a step enqueues every
    // group's launches and then waits for all, so the device runs one group's
    // kernels beside another's and fills the tail of each launch, where the
    // last blocks of the largest boxes would leave most of it idle (64 boxes
    // of 32..128 cells per axis: 7,700 Mcell-steps/s as one group, 8,170 as
    // two, 8,400 as four).  Empty: this batch is one group.
    std::vector<ns_batch *> parts;
};

namespace {

template <typename T>
int balloc(T **p, size_t bytes) {
    void *q = nullptr;
    if (hipMalloc(&q, bytes ? bytes : 1) != hipSuccess) {
        (void)hipGetLastError();
        return NS_ERR_NOMEM;
    }
    *p = (T *)q;
    return NS_OK;
}

// phases own their parameter tables (callers synchronise the stream before)
void free_phases(std::vector<ns_batch::Phase> &ph) {
    for (auto &p : ph) {
        if (p.params) (void)hipFree(p.params);
        if (p.prefix) (void)hipFree(p.prefix);
    }
    ph.clear();
}

// every box's sweep along ax is partitioned (one launch shape for all boxes)
bool batch_sweep_partitioned(const ns_batch *b, int ax) {
    for (const ns_ctx *c : b->ctx)
        if (!sweep_partitioned(c->cn[ax][0], c->cn[ax][1])) return false;
    return true;
}

// parameters of every box for one launch, uploaded with their block prefix
// sel: the boxes of this launch (null: all)
int add_phase(ns_batch *b, int kind, int arg, bool first_step, std::vector<ns_batch::Phase> &out,
              const std::vector<int> *sel = nullptr) {
    std::vector<int> all;
    if (!sel) {
        for (int i = 0; i < (int)b->ctx.size(); i++) all.push_back(i);
        sel = &all;
    }
    const int n = (int)sel->size();
    size_t psz = kind == 0 ? predict_params_size() : kind == 2 ? pass_params_size()
                           : kind == 4 ? proj_params_size() : kind == 6 ? sweep_params_size() : bc_params_size();
    std::vector<char> host((size_t)n * psz);
    std::vector<int> prefix(n + 1, 0);
    size_t lds = 0;
    int variant = 0, rows = 0;
    // predictor: blocks stream z chunks serially, so chunks of one length per
    // size class, short enough that the launch spans about `rounds` rounds of
    // resident blocks (a short tail), and at least 8 cells (the prologue of a
    // chunk re-reads 3 slices)
    long cols = 0;
    if (kind == 0)
        for (int bi : *sel) cols += predict_tile_columns(b->ctx[bi]->cfg, arg);
    // sweeps: one launch shape for all boxes, the partitioned solve only if every box has it
    const bool part = kind == 6 && batch_sweep_partitioned(b, arg);
    for (int i = 0; i < n; i++) {
        const int bi = (*sel)[i];
        ns_ctx *c = b->ctx[bi];
        const int cls = b->cls[bi];
        const ns_config &f = c->cfg;
        const Grid &g = c->g;
        const double d[3] = {f.dx, f.dy, f.dz};
        void *dst = host.data() + (size_t)i * psz;
        int nb = 0;
        // lazy: the predictor reads u* and dp (the previous step's) and writes
        // the other u* buffer, which the rest of the step works on; p is formed
        // where dp is (DCT x inverse pass, pressure fill) when dp_flag = 1
        double *const nus = b->lazy ? c->us2 : c->us, *const nvs = b->lazy ? c->vs2 : c->vs,
                      *const nws = b->lazy ? c->ws2 : c->ws;
        double *const pacc = (b->lazy && c->pacc) ? c->p : nullptr;
        if (kind == 0) {
            const long z = cols / ((long)predict_resident_blocks(b->lazy != 0, arg) * b->knobs[cls].rounds);
            const int zc_hint = (int)(z > 8 ? z : 8);
            const double *gauge = (b->lazy && (f.flags & NS_FLAG_GAUGE_CORNER)) ? c->dp : nullptr;
            if (b->lazy)
                nb = make_predict_params(f, c->g, first_step ? c->first : 0, true, c->us, c->vs, c->ws, c->dp, gauge, c->p,
                                         c->hx, c->hy, c->hz, nus, nvs, nws, dst, zc_hint, false, arg);
            else
                nb = make_predict_params(f, c->g, first_step ? c->first : 0, false, c->u, c->v, c->w, nullptr, nullptr,
                                         c->p, c->hx, c->hy, c->hz, nus, nvs, nws, dst, zc_hint, false, arg);
        } else if (kind == 1) {
            const double d0[3] = {0.0, 0.0, 0.0};
            nb = make_fill_params(g, false, f.bc, c->planes, d0, nus, nvs, nws, 0, dst);
        } else if (kind == 2) {
            size_t l = 0;
            nb = make_dct_pass_params(g, c->pp, f.dx, f.dy, f.dz, f.dt, nus, nvs, nws, nullptr, c->planes,
                                      c->planes + c->L.off[24], 0.0, c->work, c->dp, nullptr, arg, dst, &l,
                                      b->knobs[cls].dct_drop, pacc);
            lds = l > lds ? l : lds;
            variant |= dct_params_variant(dst);
        } else if (kind == 5) {
            nb = make_cn2_params(g, f.bc[arg], arg, c->planes, nus, nvs, nws, dst);
        } else if (kind == 6) {
            size_t l = 0;
            nb = make_sweep_params(g, arg + 1, c->cn[arg][0], c->cn[arg][1], nus, nvs, nws, dst, &l,
                                   b->knobs[cls].adix_lsh, part, part ? c->planes : nullptr, &c->L);
            lds = l > lds ? l : lds;
            const int kw = sweep_params_waves(dst), rw = sweep_params_rows(dst);
            variant = kw > variant ? kw : variant;
            rows = rw > rows ? rw : rows;
        } else if (kind == 3) {
            nb = make_fill_params(g, true, f.pbc, c->planes, d, c->dp, pacc, nullptr, 0, dst);
        } else {
            const double *gauge = (f.flags & NS_FLAG_GAUGE_CORNER) ? c->dp : nullptr;
            nb = make_project_params(g, d, f.dt, f.dp_flag, c->us, c->vs, c->ws, c->dp, gauge, c->u, c->v, c->w, c->p,
                                     dst);
        }
        prefix[i + 1] = prefix[i] + nb;
    }
    if (prefix[n] == 0) return NS_OK;
    ns_batch::Phase ph;
    ph.kind = kind;
    ph.arg = arg;
    ph.nblk = prefix[n];
    ph.lds = lds;
    ph.variant = variant;
    ph.rows = rows;
    for (int bi : *sel) ph.dense = ph.dense && b->ctx[bi]->g.dense();
    ph.nbox = n;
    int rc;
    if ((rc = balloc(&ph.params, host.size())) || (rc = balloc(&ph.prefix, prefix.size() * sizeof(int)))) {
        if (ph.params) (void)hipFree(ph.params);
        return rc;
    }
    out.push_back(ph);  // owned by `out` from here
    HIPCHK(hipMemcpy(ph.params, host.data(), host.size(), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(ph.prefix, prefix.data(), prefix.size() * sizeof(int), hipMemcpyHostToDevice));
    return NS_OK;
}

int build_phases(ns_batch *b, bool first_step, std::vector<ns_batch::Phase> &out) {
    int rc;
    // one predictor launch per tile shape (arg: the tile width)
    for (int tile : {64, 32}) {
        std::vector<int> sel;
        for (int i = 0; i < (int)b->ctx.size(); i++)
            if (predict_shape(b->ctx[i]->cfg, b->knobs[b->cls[i]].tile) == tile) sel.push_back(i);
        if (!sel.empty() && (rc = add_phase(b, 0, tile, first_step, out, &sel))) return rc;
    }
    if (b->ctx[0]->cfg.scheme == NS_AB2_CN) {
        // per axis: boundary rows, then the sweep (x, y, z); a partitioned
        // sweep reads the boundary rows from the planes itself
        for (int ax = 0; ax < 3; ax++) {
            if (!batch_sweep_partitioned(b, ax) && (rc = add_phase(b, 5, ax, first_step, out))) return rc;
            if (!batch_sweep_partitioned(b, ax)) {
                if ((rc = add_phase(b, 6, ax, first_step, out))) return rc;
                continue;
            }
            // one launch per blocks-per-line count: each sized (waves, LDS) for its own boxes
            for (int K = 1; K <= PART_MAX_BLOCKS; K++) {
                std::vector<int> sel;
                for (int i = 0; i < (int)b->ctx.size(); i++) {
                    const LineSys &sa = b->ctx[i]->cn[ax][0], &sb = b->ctx[i]->cn[ax][1];
                    if ((sa.K > sb.K ? sa.K : sb.K) == K) sel.push_back(i);
                }
                if (!sel.empty() && (rc = add_phase(b, 6, ax, first_step, out, &sel))) return rc;
            }
        }
    } else if ((rc = add_phase(b, 1, 0, first_step, out))) {  // the whole velocity fill
        return rc;
    }
    for (int pass = 0; pass < 5; pass++)
        if ((rc = add_phase(b, 2, pass, first_step, out))) return rc;
    if ((rc = add_phase(b, 3, 0, first_step, out))) return rc;  // the whole pressure fill
    if (b->lazy) return NS_OK;                                     // projected by the next predictor
    return add_phase(b, 4, 0, first_step, out);
}

void run_phases(ns_batch *b, const std::vector<ns_batch::Phase> &phases) {
    for (const auto &ph : phases) {
        const int n = ph.nbox;
        switch (ph.kind) {
            case 0: launch_predict_batch(b->lazy != 0, ph.arg, ph.params, ph.prefix, n, ph.nblk, b->st); break;
            case 1: launch_whole_fill_batch(ph.params, ph.prefix, n, ph.nblk, 3, b->st); break;
            case 2: launch_dct_batch(ph.arg, ph.variant, ph.params, ph.prefix, n, ph.nblk, ph.lds, b->st); break;
            case 3: launch_whole_fill_batch(ph.params, ph.prefix, n, ph.nblk, 1, b->st); break;
            case 5: launch_cn2_batch(ph.params, ph.prefix, n, ph.nblk, b->st); break;
            case 6:
                launch_adi_batch(ph.arg, ph.variant, ph.rows, ph.dense != 0, ph.params, ph.prefix, n, ph.nblk, ph.lds,
                                 b->st);
                break;
            default: launch_project_batch(ph.params, ph.prefix, n, ph.nblk, b->st); break;
        }
    }
}

// exchange every box's two u* buffers (host pointers; lazy steps alternate them)
void swap_star(ns_batch *b) {
    for (ns_ctx *c : b->ctx) {
        std::swap(c->us, c->us2);
        std::swap(c->vs, c->vs2);
        std::swap(c->ws, c->ws2);
    }
}

void drop_steady(ns_batch *b) {
    for (int k = 0; k < 2; k++) {
        free_phases(b->phases[k]);
        if (b->graph[k]) {
            (void)hipGraphExecDestroy(b->graph[k]);
            b->graph[k] = nullptr;
        }
    }
}

// the steady-step tables for the boxes' current u* buffers (phases[0]) and,
// lazy, for the exchanged ones (phases[1])
int build_steady(ns_batch *b) {
    drop_steady(b);
    int rc = build_phases(b, false, b->phases[0]);
    if (!rc && b->lazy) {
        swap_star(b);
        rc = build_phases(b, false, b->phases[1]);
        swap_star(b);
    }
    b->sig.clear();
    for (ns_ctx *c : b->ctx) b->sig.push_back(c->us);
    return rc;
}

// which steady tables match the boxes' buffers now: 0, 1, or -1 (none: a box
// was stepped on its own since they were built)
int steady_parity(const ns_batch *b) {
    bool p0 = true, p1 = b->lazy != 0;
    for (size_t i = 0; i < b->ctx.size(); i++) {
        p0 = p0 && b->ctx[i]->us == b->sig[i];
        p1 = p1 && b->ctx[i]->us2 == b->sig[i];
    }
    return p0 ? 0 : (p1 ? 1 : -1);
}

// host-side effects of batch steps on each box
void after_batch_steps(ns_batch *b) {
    for (ns_ctx *c : b->ctx) {
        c->first = 0;
        c->ghost_user = 0;
        if (b->lazy) c->stale = 1;
    }
}

int batch_enqueue(ns_batch *b, int nsteps);

// one group: its boxes (of one projection mode), size classes, stream and steady tables
int batch_group(ns_ctx *const *ctxs, int n, bool all_lazy, ns_batch **out) {
    ns_batch *b = new ns_batch();
    b->ctx.assign(ctxs, ctxs + n);
    b->lazy = all_lazy ? 1 : 0;
    for (ns_ctx *c : b->ctx) {
        const size_t cells = c->g.ni();
        b->cls.push_back(cells < ((size_t)1 << 17) ? 0 : cells < ((size_t)1 << 20) ? 1 : 2);
    }
    if (hipStreamCreateWithFlags(&b->st, hipStreamNonBlocking) != hipSuccess) {
        (void)hipGetLastError();
        delete b;
        return NS_ERR_HIP;
    }
    int rc = build_steady(b);
    if (rc) {
        ns_batch_destroy(b);
        return rc;
    }
    *out = b;
    return NS_OK;
}

}  // namespace

extern "C" {

int ns_batch_create(ns_ctx *const *ctxs, int n, ns_batch **out) {
    if (!ctxs || n < 1 || !out) return NS_ERR_INVALID;
    const ns_config &f0 = ctxs[0]->cfg;
    for (int i = 0; i < n; i++) {
        const ns_ctx *c = ctxs[i];
        if (!c) return NS_ERR_INVALID;
        const ns_config &f = c->cfg;
        // one kernel family per launch: same scheme and codes, DCT Poisson
        if (!c->poisson_ok || !c->pp.dct) return NS_ERR_UNSUPPORTED;
        if (f.scheme != f0.scheme || f.dp_flag != f0.dp_flag || (f.flags & ~NS_FLAG_NO_GRAPH) !=
            (f0.flags & ~NS_FLAG_NO_GRAPH))
            return NS_ERR_UNSUPPORTED;
        for (int a = 0; a < 3; a++)
            if (f.bc[a] != f0.bc[a] || f.pbc[a] != f0.pbc[a]) return NS_ERR_UNSUPPORTED;
    }
    // one projection mode per batch: where the boxes differ (a box that left
    // an earlier batch projects every step), the lazy ones form u, v, w, p and
    // project every step from here on
    bool all_lazy = true;
    for (int i = 0; i < n; i++) all_lazy = all_lazy && ctxs[i]->lazy;
    for (int i = 0; i < n && !all_lazy; i++) {
        ns_ctx *c = ctxs[i];
        if (!c->lazy) continue;
        materialize(c);
        c->lazy = c->pacc = c->star_fill = c->ghost_user = 0;
        for (int k = 0; k < 2; k++)
            if (c->graph[k]) {
                (void)hipGraphExecDestroy(c->graph[k]);
                c->graph[k] = nullptr;
            }
        HIPCHK(hipStreamSynchronize(c->st));
    }
    // groups: at most 4 (the hardware queues a process gets), of at least 8 boxes
    int ng = n / 8;
    if (ng > 4) ng = 4;
    if (ng < 1) ng = 1;
    std::vector<std::vector<ns_ctx *>> grp(ng);
    {
        std::vector<int> order(n);
        for (int i = 0; i < n; i++) order[i] = i;
        std::stable_sort(order.begin(), order.end(),
                         [&](int a, int c) { return ctxs[a]->g.ni() > ctxs[c]->g.ni(); });
        std::vector<size_t> load(ng, 0);
        for (int i : order) {
            int j = 0;
            for (int k = 1; k < ng; k++)
                if (load[k] < load[j]) j = k;
            grp[j].push_back(ctxs[i]);
            load[j] += ctxs[i]->g.ni();
        }
    }
    ns_batch *b = nullptr;
    int rc = batch_group(ctxs, n, all_lazy, &b);
    if (rc) return rc;
    if (ng > 1) {
        drop_steady(b);  // the groups run the steps
        for (int j = 0; j < ng; j++) {
            ns_batch *q = nullptr;
            if ((rc = batch_group(grp[j].data(), (int)grp[j].size(), all_lazy, &q))) {
                ns_batch_destroy(b);
                return rc;
            }
            b->parts.push_back(q);
        }
    }
    *out = b;
    return NS_OK;
}

int ns_batch_destroy(ns_batch *b) {
    if (!b) return NS_ERR_INVALID;
    for (ns_batch *q : b->parts) ns_batch_destroy(q);
    if (b->st) (void)hipStreamSynchronize(b->st);
    drop_steady(b);
    if (b->st) (void)hipStreamDestroy(b->st);
    delete b;
    return NS_OK;
}

int ns_batch_step(ns_batch *b, int nsteps) {
    if (!b || nsteps < 0) return NS_ERR_INVALID;
    if (b->parts.empty()) {
        int rc = batch_enqueue(b, nsteps);
        if (rc) return rc;
        HIPCHK(hipStreamSynchronize(b->st));
        return hipGetLastError() == hipSuccess ? NS_OK : NS_ERR_HIP;
    }
    // every group's steps queued on its own stream, then all drained
    int rc = NS_OK;
    for (ns_batch *q : b->parts)
        if (!rc) rc = batch_enqueue(q, nsteps);
    for (ns_batch *q : b->parts) HIPCHK(hipStreamSynchronize(q->st));
    if (rc) return rc;
    return hipGetLastError() == hipSuccess ? NS_OK : NS_ERR_HIP;
}

}  // extern "C"

namespace {

// the steps of one group queued on its stream (a first step is waited for)
int batch_enqueue(ns_batch *b, int nsteps) {
    // work queued on the contexts' own streams (uploads, plane sets) comes first
    for (ns_ctx *c : b->ctx) HIPCHK(hipStreamSynchronize(c->st));
    int s = 0;
    bool any_first = false;
    for (ns_ctx *c : b->ctx) any_first = any_first || c->first;
    if (any_first && nsteps > 0) {
        // the first step (Euler for the boxes that start fresh) runs from its own parameters
        std::vector<ns_batch::Phase> ph;
        int rc = build_phases(b, true, ph);
        if (!rc) run_phases(b, ph);
        const hipError_t e = hipStreamSynchronize(b->st);
        free_phases(ph);
        if (rc) return rc;
        HIPCHK(e);
        if (b->lazy) swap_star(b);
        after_batch_steps(b);
        s = 1;
    }
    if (s >= nsteps) return NS_OK;
    int par = steady_parity(b);
    if (par < 0) {
        int rc = build_steady(b);
        if (rc) return rc;
        par = 0;
    }
    const bool use_graph = !(b->ctx[0]->cfg.flags & NS_FLAG_NO_GRAPH);
    for (; s < nsteps; s++) {
        if (use_graph && !b->graph[par]) {
            hipGraph_t gr = nullptr;
            if (hipStreamBeginCapture(b->st, hipStreamCaptureModeThreadLocal) == hipSuccess) {
                run_phases(b, b->phases[par]);
                if (hipStreamEndCapture(b->st, &gr) == hipSuccess && gr) {
                    if (hipGraphInstantiate(&b->graph[par], gr, nullptr, nullptr, 0) != hipSuccess)
                        b->graph[par] = nullptr;
                }
                if (gr) (void)hipGraphDestroy(gr);
            }
            (void)hipGetLastError();
        }
        if (use_graph && b->graph[par]) HIPCHK(hipGraphLaunch(b->graph[par], b->st));
        else run_phases(b, b->phases[par]);
        if (b->lazy) {
            swap_star(b);
            par ^= 1;
        }
    }
    after_batch_steps(b);
    return NS_OK;
}

}  // namespace

extern "C" {

int ns_batch_tune(ns_batch *b, int trial_steps, int *knobs_out) {
    if (!b || trial_steps < 1) return NS_ERR_INVALID;
    // the groups that run the steps; one knob set for all of them
    std::vector<ns_batch *> G;
    if (b->parts.empty()) G.push_back(b);
    else G = b->parts;
    const int ng = (int)G.size();
    for (ns_ctx *c : b->ctx) HIPCHK(hipStreamSynchronize(c->st));
    for (ns_batch *q : G) HIPCHK(hipStreamSynchronize(q->st));
    // the state every trial starts from: u, v, w, p and the AB2 histories of each box
    // (lazy boxes: u*, dp, p and the histories, with u* in the buffers it is in now)
    std::vector<double *> saved;
    auto release = [&]() {
        for (double *q : saved) (void)hipFree(q);
        saved.clear();
    };
    auto state = [&](ns_ctx *c, double *a[8], size_t n[8]) {
        const Grid &g = c->g;
        const size_t m[8] = {g.nu(), g.nv(), g.nw(), g.np(), g.nu(), g.nv(), g.nw(), g.np()};
        double *q[8] = {b->lazy ? c->us : c->u, b->lazy ? c->vs : c->v, b->lazy ? c->ws : c->w, c->p,
                        c->hx, c->hy, c->hz, b->lazy ? c->dp : nullptr};
        for (int i = 0; i < 8; i++) {
            a[i] = q[i];
            n[i] = m[i];
        }
    };
    std::vector<std::vector<int>> first(ng);
    for (int gi = 0; gi < ng; gi++)
        for (ns_ctx *c : G[gi]->ctx) {
            double *src[8];
            size_t n[8];
            state(c, src, n);
            for (int i = 0; i < 8; i++) {
                if (!src[i]) continue;
                double *q = nullptr;
                if (balloc(&q, n[i] * sizeof(double))) {
                    release();
                    return NS_ERR_NOMEM;
                }
                saved.push_back(q);
                (void)hipMemcpyAsync(q, src[i], n[i] * sizeof(double), hipMemcpyDeviceToDevice, G[gi]->st);
            }
            first[gi].push_back(c->first);
        }
    // parity of each group's u* buffers relative to the saved state
    std::vector<int> flips(ng, 0);
    auto restore = [&]() {
        size_t k = 0;
        for (int gi = 0; gi < ng; gi++) {
            ns_batch *q = G[gi];
            if (flips[gi] & 1) swap_star(q);
            flips[gi] = 0;
            for (size_t j = 0; j < q->ctx.size(); j++) {
                ns_ctx *c = q->ctx[j];
                double *dst[8];
                size_t n[8];
                state(c, dst, n);
                for (int i = 0; i < 8; i++) {
                    if (!dst[i]) continue;
                    (void)hipMemcpyAsync(dst[i], saved[k], n[i] * sizeof(double), hipMemcpyDeviceToDevice, q->st);
                    k++;
                }
                c->first = first[gi][j];
            }
        }
    };
    std::vector<hipEvent_t> ev(ng + 1, nullptr);
    for (auto &e : ev)
        if (hipEventCreate(&e) != hipSuccess) {
            (void)hipGetLastError();
            for (auto &x : ev)
                if (x) (void)hipEventDestroy(x);
            release();
            return NS_ERR_HIP;
        }
    // ms of nst steady steps of every group at once with the current knobs
    // (best of reps runs), from an event on group 0's stream that the others
    // wait for to one after all of them
    auto trial = [&](float &ms, int nst, int reps) -> int {
        std::vector<std::vector<ns_batch::Phase>> ph(2 * ng);
        int rc = NS_OK;
        for (int gi = 0; gi < ng && !rc; gi++) {
            ns_batch *q = G[gi];
            q->knobs[0] = b->knobs[0];
            q->knobs[1] = b->knobs[1];
            q->knobs[2] = b->knobs[2];
            rc = build_phases(q, false, ph[2 * gi]);
            if (!rc && q->lazy) {
                swap_star(q);
                rc = build_phases(q, false, ph[2 * gi + 1]);
                swap_star(q);
            }
        }
        auto step_all = [&]() {
            for (int gi = 0; gi < ng; gi++) {
                ns_batch *q = G[gi];
                run_phases(q, ph[2 * gi + (flips[gi] & 1)]);
                if (q->lazy) {
                    swap_star(q);
                    flips[gi]++;
                }
            }
        };
        ms = 1e30f;
        for (int rep = 0; rep < reps && !rc; rep++) {
            step_all();  // warm-up of this shape
            (void)hipEventRecord(ev[0], G[0]->st);
            for (int gi = 1; gi < ng; gi++) (void)hipStreamWaitEvent(G[gi]->st, ev[0], 0);
            for (int s2 = 0; s2 < nst; s2++) step_all();
            for (int gi = 1; gi < ng; gi++) {
                (void)hipEventRecord(ev[1 + gi], G[gi]->st);
                (void)hipStreamWaitEvent(G[0]->st, ev[1 + gi], 0);
            }
            (void)hipEventRecord(ev[1], G[0]->st);
            restore();
            if (hipEventSynchronize(ev[1]) != hipSuccess) rc = NS_ERR_HIP;
            float t = 0.0f;
            if (!rc && hipEventElapsedTime(&t, ev[0], ev[1]) == hipSuccess && t < ms) ms = t;
        }
        for (ns_batch *q : G) (void)hipStreamSynchronize(q->st);
        for (auto &x : ph) free_phases(x);
        (void)hipGetLastError();
        return rc;
    };
    const ns_batch::Knobs start[3] = {b->knobs[0], b->knobs[1], b->knobs[2]};
    float best = 0.0f;
    int rc = trial(best, trial_steps, 2);
    bool present[3] = {false, false, false};
    for (int k : b->cls) present[k] = true;
    // coordinate descent over (size class, knob); a change is kept when it
    // saves more than 1% of the step
    const int cand_rounds[4] = {4, 8, 16, 32}, cand_drop[3] = {0, 1, 2}, cand_lsh[4] = {6, 5, 4, 3},
              cand_tile[3] = {0, 64, 32};
    // the x sweep's lines per block only shape the twisted / cyclic sweeps
    bool x_twisted = false;
    for (ns_batch *q : G)
        x_twisted = x_twisted || (q->ctx[0]->cfg.scheme == NS_AB2_CN && !batch_sweep_partitioned(q, 0));
    for (int k = 0; k < 3 && !rc; k++) {
        if (!present[k]) continue;
        for (int knob = 0; knob < 4 && !rc; knob++) {
            if (knob == 2 && !x_twisted) continue;
            int *val = knob == 0   ? &b->knobs[k].rounds
                       : knob == 1 ? &b->knobs[k].dct_drop
                       : knob == 2 ? &b->knobs[k].adix_lsh
                                   : &b->knobs[k].tile;
            const int *cand = knob == 0 ? cand_rounds : knob == 1 ? cand_drop : knob == 2 ? cand_lsh : cand_tile;
            const int nc = (knob == 1 || knob == 3) ? 3 : 4;
            const int keep = *val;
            int best_v = keep;
            for (int i = 0; i < nc && !rc; i++) {
                if (cand[i] == keep) continue;
                *val = cand[i];
                float t = 0.0f;
                rc = trial(t, trial_steps, 2);
                if (!rc && t < 0.99f * best) {
                    best = t;
                    best_v = cand[i];
                }
            }
            *val = best_v;
        }
    }
    // the short trials are noisier than the gains they look for: a choice
    // that differs from the starting shapes is kept only if longer runs of
    // both (twice the steps, best of three) still find it more than 1% faster
    bool changed = false;
    for (int k = 0; k < 3; k++)
        changed = changed || b->knobs[k].rounds != start[k].rounds || b->knobs[k].dct_drop != start[k].dct_drop ||
                  b->knobs[k].adix_lsh != start[k].adix_lsh || b->knobs[k].tile != start[k].tile;
    if (!rc && changed) {
        const ns_batch::Knobs chosen[3] = {b->knobs[0], b->knobs[1], b->knobs[2]};
        float t_new = 0.0f, t_old = 0.0f;
        rc = trial(t_new, 2 * trial_steps, 3);
        for (int k = 0; k < 3; k++) b->knobs[k] = start[k];
        if (!rc) rc = trial(t_old, 2 * trial_steps, 3);
        if (!rc && t_new < 0.99f * t_old)
            for (int k = 0; k < 3; k++) b->knobs[k] = chosen[k];
    }
    for (auto &x : ev) (void)hipEventDestroy(x);
    release();
    if (rc) return rc;
    // the steady-step tables with the chosen shapes (graphs captured on the next steps)
    for (ns_batch *q : G) {
        q->knobs[0] = b->knobs[0];
        q->knobs[1] = b->knobs[1];
        q->knobs[2] = b->knobs[2];
        if ((rc = build_steady(q))) return rc;
    }
    if (knobs_out)
        for (int k = 0; k < 3; k++) {
            knobs_out[4 * k] = b->knobs[k].rounds;
            knobs_out[4 * k + 1] = b->knobs[k].dct_drop;
            knobs_out[4 * k + 2] = b->knobs[k].adix_lsh;
            knobs_out[4 * k + 3] = b->knobs[k].tile;
        }
    return NS_OK;
}

}  // extern "C"
