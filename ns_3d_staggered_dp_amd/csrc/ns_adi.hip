// Crank-Nicolson ADI sweeps for gfx950, FP64: batched tridiagonal solves
// along x, y or z over every line of u*, v*, w* (ghost lines included), as
// the reference's getrs/dttrsb loops do (big_periodic_3D.f90:425-483,
// re_simulation_sub_tstep_interBC_combine.f90:656-701).
//
// Matrices (mat_CN NS_functions.f90:1245-1294, mat_CN_tridiagonal :1297-1342):
// walls (bc 2/4) are plain tridiagonal and solved by Thomas over the whole
// line with host-factored coefficients.  Periodic lines are reduced to the n
// owned unknowns (the ghost rows of the dense matrix are u_N = u_1 + r_N,
// v_1 = v_{n+1} + r_1, v_{n+2} = v_2 + r_{n+2}), which leaves a cyclic
// tridiagonal system solved by Thomas plus a Sherman-Morrison rank-one
// correction whose line-independent part is precomputed on the host.
#include "ns_internal.h"

namespace ns {

namespace {

// solve one line in place (global memory, stride s)
__device__ __forceinline__ void solve_line(double *x, size_t s, const LineSys &S) {
    if (!S.periodic) {
        const int N = S.N;
        double d = x[0] * S.inv[0];
        x[0] = d;
        for (int i = 1; i < N; i++) {
            d = (x[s * i] - S.lo[i] * d) * S.inv[i];
            x[s * i] = d;
        }
        double nxt = d;
        for (int i = N - 2; i >= 0; i--) {
            nxt = x[s * i] - S.cp[i] * nxt;
            x[s * i] = nxt;
        }
        return;
    }
    const int m = S.m;
    double *y = x + s * S.off;
    double r_first, r_last;
    // ghost-row right-hand sides folded into the reduced system
    double g_lo = 0.0, g_hi = 0.0;  // r of the leading / trailing ghost rows
    if (S.off == 0) {
        g_hi = x[s * m];  // A-type: row N (u_N - u_1 = r_N)
        r_first = y[0] + (S.quirk ? S.a * g_hi : 0.0);
        r_last = y[s * (m - 1)] + S.a * g_hi;
    } else {
        g_lo = x[0];
        g_hi = x[s * (m + 1)];
        r_first = y[0] + S.a * g_lo;
        r_last = y[s * (m - 1)] + S.a * g_hi;
    }
    // Thomas forward on the reduced system
    double d = r_first * S.inv[0];
    y[0] = d;
    for (int i = 1; i < m; i++) {
        double r = (i == m - 1) ? r_last : y[s * i];
        d = (r - S.lo[i] * d) * S.inv[i];
        y[s * i] = d;
    }
    double nxt = d;
    const double ylast = d;
    for (int i = m - 2; i >= 0; i--) {
        nxt = y[s * i] - S.cp[i] * nxt;
        y[s * i] = nxt;
    }
    // Sherman-Morrison correction
    const double f = (nxt + S.s * ylast) / S.den;
    double x0 = 0.0, xl = 0.0;
    for (int i = 0; i < m; i++) {
        double v = y[s * i] - f * S.z[i];
        y[s * i] = v;
        if (i == 0) x0 = v;
        if (i == m - 1) xl = v;
    }
    if (S.off == 0) {
        x[s * m] = x0 + g_hi;
    } else {
        x[0] = xl + g_lo;
        x[s * (m + 1)] = x0 + g_hi;
    }
}

struct SweepParams {
    Grid g;
    int ax;
    LineSys sys[3];  // per component (A for the normal one, B otherwise)
    double *a[3];
};

__global__ void __launch_bounds__(64) k_adi(SweepParams P) {
    const int c = blockIdx.y;
    int n[3];
    comp_dims(P.g, c, n);
    const int ax = P.ax;
    const int o1 = (ax == 0) ? 1 : 0, o2 = (ax == 2) ? 1 : 2;
    size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)n[o1] * n[o2]) return;
    int id[3];
    id[o1] = (int)(t % n[o1]) + 1;
    id[o2] = (int)(t / n[o1]) + 1;
    id[ax] = 1;
    size_t stride = (ax == 0) ? 1 : (ax == 1) ? (size_t)n[0] : (size_t)n[0] * n[1];
    double *line = P.a[c] + FIX(id[0], id[1], id[2], n[0], n[1]);
    solve_line(line, stride, P.sys[c]);
}

}  // namespace

void launch_adi_sweep(const Grid &g, int dir, const LineSys &sysA, const LineSys &sysB, double *u, double *v, double *w,
                      hipStream_t st) {
    SweepParams P;
    P.g = g;
    P.ax = dir - 1;
    for (int c = 0; c < 3; c++) P.sys[c] = (c == P.ax) ? sysA : sysB;
    P.a[0] = u; P.a[1] = v; P.a[2] = w;
    size_t mx = 0;
    for (int c = 0; c < 3; c++) {
        int n[3];
        comp_dims(g, c, n);
        int o1 = (P.ax == 0) ? 1 : 0, o2 = (P.ax == 2) ? 1 : 2;
        size_t s = (size_t)n[o1] * n[o2];
        mx = s > mx ? s : mx;
    }
    hipLaunchKernelGGL(k_adi, dim3((unsigned)((mx + 63) / 64), 3), dim3(64), 0, st, P);
}

}  // namespace ns
