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

// The sweep is one thread per line, so each elimination step would wait for
// its own load: lines are walked in register chunks of CHK elements, the next
// chunk's loads issued before the current chunk's dependent arithmetic.  The
// coefficients are line-independent and read at a wave-uniform index through
// the constant address space (scalar loads).
constexpr int CHK = 16;
typedef const __attribute__((address_space(4))) double cdouble;

// forward elimination over y[0..m) (stride s): d_i = (r_i - lo_i d_{i-1}) inv_i,
// stored in place; r_0 and (m > 1) r_{m-1} are replaced by r_first, r_last when ovr.
// Returns d_{m-1}.
__device__ __forceinline__ double thomas_fwd(double *y, size_t s, int m, const double *lo_, const double *inv_,
                                             bool ovr, double r_first, double r_last) {
    cdouble *lo = (cdouble *)lo_, *inv = (cdouble *)inv_;
    double cur[CHK], nxt[CHK];
#pragma unroll
    for (int j = 0; j < CHK; j++) cur[j] = (j < m) ? y[s * j] : 0.0;
    double d = 0.0;
    for (int i0 = 0; i0 < m; i0 += CHK) {
#pragma unroll
        for (int j = 0; j < CHK; j++) nxt[j] = (i0 + CHK + j < m) ? y[s * (i0 + CHK + j)] : 0.0;
#pragma unroll
        for (int j = 0; j < CHK; j++) {
            const int i = i0 + j;
            if (i < m) {
                double r = cur[j];
                if (ovr) {
                    if (i == 0) r = r_first;
                    else if (i == m - 1) r = r_last;
                }
                d = (i == 0) ? r * inv[0] : (r - lo[i] * d) * inv[i];
                y[s * i] = d;
            }
        }
#pragma unroll
        for (int j = 0; j < CHK; j++) cur[j] = nxt[j];
    }
    return d;
}

// back substitution x_i = d_i - cp_i x_{i+1}, i = m-2 .. 0, with x_{m-1} = last
// already in place; returns x_0
__device__ __forceinline__ double thomas_bwd(double *y, size_t s, int m, const double *cp_, double last) {
    cdouble *cp = (cdouble *)cp_;
    double cur[CHK], nxt[CHK];
    const int top = m - 2;
#pragma unroll
    for (int j = 0; j < CHK; j++) cur[j] = (top - j >= 0) ? y[s * (top - j)] : 0.0;
    double x = last;
    for (int i0 = top; i0 >= 0; i0 -= CHK) {
#pragma unroll
        for (int j = 0; j < CHK; j++) nxt[j] = (i0 - CHK - j >= 0) ? y[s * (i0 - CHK - j)] : 0.0;
#pragma unroll
        for (int j = 0; j < CHK; j++) {
            const int i = i0 - j;
            if (i >= 0) {
                x = cur[j] - cp[i] * x;
                y[s * i] = x;
            }
        }
#pragma unroll
        for (int j = 0; j < CHK; j++) cur[j] = nxt[j];
    }
    return x;
}

// solve one line in place (global memory, stride s)
__device__ __forceinline__ void solve_line(double *x, size_t s, const LineSys &S) {
    if (!S.periodic) {
        const double last = thomas_fwd(x, s, S.N, S.lo, S.inv, false, 0.0, 0.0);
        thomas_bwd(x, s, S.N, S.cp, last);
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
    // Thomas on the reduced system (rows 0 and m-1 carry the folded ghosts)
    const double ylast = thomas_fwd(y, s, m, S.lo, S.inv, true, r_first, r_last);
    const double y0 = thomas_bwd(y, s, m, S.cp, ylast);
    // Sherman-Morrison correction
    const double f = (y0 + S.s * ylast) / S.den;
    cdouble *z = (cdouble *)S.z;
    double x0 = 0.0, xl = 0.0;
    for (int i0 = 0; i0 < m; i0 += CHK) {
        double v[CHK];
#pragma unroll
        for (int j = 0; j < CHK; j++) v[j] = (i0 + j < m) ? y[s * (i0 + j)] : 0.0;
#pragma unroll
        for (int j = 0; j < CHK; j++) {
            const int i = i0 + j;
            if (i < m) {
                const double w = v[j] - f * z[i];
                y[s * i] = w;
                if (i == 0) x0 = w;
                if (i == m - 1) xl = w;
            }
        }
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
