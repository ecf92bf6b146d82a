// Crank-Nicolson ADI sweeps for gfx950, FP64: batched tridiagonal solves
// along x, y or z over every line of u*, v*, w* (ghost lines included), as
// the reference's getrs/dttrsb loops do (big_periodic_3D.f90:425-483,
// re_simulation_sub_tstep_interBC_combine.f90:656-701).
//
// Matrices (mat_CN NS_functions.f90:1245-1294, mat_CN_tridiagonal :1297-1342):
// walls (bc 2/4) are plain tridiagonal; every line of a direction and
// component has the same matrix, so what depends on the matrix alone is
// computed on the host.  Wall lines of up to PART_MAX_BLOCKS * PART_ROWS_LONG
// rows use the partitioned solve (one wave per block of rows, couplings through
// host-precomputed spikes); longer ones a twisted Thomas factorization (two
// lanes per line).  Periodic lines are reduced to the n owned unknowns (the
// ghost rows of the dense matrix are u_N = u_1 + r_N, v_1 = v_{n+1} + r_1,
// v_{n+2} = v_2 + r_{n+2}), which leaves a cyclic tridiagonal system solved by
// Thomas plus a Sherman-Morrison rank-one correction whose line-independent
// part is precomputed on the host.
#include "ns_internal.h"

namespace ns {

namespace {

// Twisted and cyclic solves: one or two lanes walk a whole line, so each
// elimination step would wait for its own loads: lines and coefficients are
// read in register chunks of CHK elements (one wait per chunk), the next
// chunk's data issued before the current chunk's dependent arithmetic.
constexpr int SB = 16;  // staging copies in flight per thread (x lines staged in LDS)

// forward elimination over y[0..m) (stride s): d_i = (r_i - lo_i d_{i-1}) inv_i,
// stored to dout (stride so; y itself for an in-place solve); r_0 and (m > 1)
// r_{m-1} are replaced by r_first, r_last when ovr.  Returns d_{m-1}.
template <int CHK>
__device__ __forceinline__ double thomas_fwd(const double *y, long s, double *dout, long so, int m,
                                             const double *lo, const double *inv, bool ovr, double r_first,
                                             double r_last) {
    double cur[CHK], nxt[CHK];
#pragma unroll
    for (int j = 0; j < CHK; j++) cur[j] = (j < m) ? y[s * j] : 0.0;
    double d = 0.0;
    for (int i0 = 0; i0 < m; i0 += CHK) {
        double clo[CHK], cinv[CHK];
#pragma unroll
        for (int j = 0; j < CHK; j++) {
            const int i = (i0 + j < m) ? i0 + j : m - 1;
            clo[j] = lo[i];
            cinv[j] = inv[i];
        }
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
                d = (i == 0) ? r * cinv[j] : (r - clo[j] * d) * cinv[j];
                dout[so * i] = d;
            }
        }
#pragma unroll
        for (int j = 0; j < CHK; j++) cur[j] = nxt[j];
    }
    return d;
}

// back substitution x_i = d_i - cp_i x_{i+1}, i = m-2 .. 0, reading d from y
// (stride s) and storing x to xout (stride so), x_{m-1} = last already stored
// there; returns x_0
template <int CHK>
__device__ __forceinline__ double thomas_bwd(const double *y, long s, double *xout, long so, int m,
                                             const double *cp, double last) {
    double cur[CHK], nxt[CHK];
    const int top = m - 2;
#pragma unroll
    for (int j = 0; j < CHK; j++) cur[j] = (top - j >= 0) ? y[s * (top - j)] : 0.0;
    double x = last;
    for (int i0 = top; i0 >= 0; i0 -= CHK) {
        double ccp[CHK];
#pragma unroll
        for (int j = 0; j < CHK; j++) ccp[j] = cp[(i0 - j >= 0) ? i0 - j : 0];
#pragma unroll
        for (int j = 0; j < CHK; j++) nxt[j] = (i0 - CHK - j >= 0) ? y[s * (i0 - CHK - j)] : 0.0;
#pragma unroll
        for (int j = 0; j < CHK; j++) {
            const int i = i0 - j;
            if (i >= 0) {
                x = cur[j] - ccp[j] * x;
                xout[so * i] = x;
            }
        }
#pragma unroll
        for (int j = 0; j < CHK; j++) cur[j] = nxt[j];
    }
    return x;
}

// solve one line in place (stride s); CHK: elements per register chunk
// (16 for lines in global memory, 8 for lines in LDS)
template <int CHK>
__device__ __forceinline__ void solve_line(double *x, long s, const LineSys &S) {
    if (!S.periodic) {
        const double last = thomas_fwd<CHK>(x, s, x, s, S.N, S.lo, S.inv, false, 0.0, 0.0);
        thomas_bwd<CHK>(x, s, x, s, S.N, S.cp, last);
        return;
    }
    const int m = S.m;
    double *y = x + s * S.off;
    double r_first, r_last;
    // ghost-row right-hand sides folded into the reduced system
    double g_lo = 0.0, g_hi = 0.0;  // r of the leading / trailing ghost rows
    if (S.off == 0) {
        g_hi = x[s * m];  // A-type: row N (u_N - u_1 = r_N)
        if (S.adj) {
            // transpose of "u_N = u_1 + r_N": the ghost's adjoint joins row 0's
            r_first = y[0] + g_hi;
            r_last = y[s * (m - 1)];
        } else {
            r_first = y[0] + (S.quirk ? S.a * g_hi : 0.0);
            r_last = y[s * (m - 1)] + S.a * g_hi;
        }
    } else {
        g_lo = x[0];
        g_hi = x[s * (m + 1)];
        if (S.adj) {
            // transpose of "v_1 = v_{n+1} + r_1, v_{n+2} = v_2 + r_{n+2}"
            r_first = y[0] + g_hi;
            r_last = y[s * (m - 1)] + g_lo;
        } else {
            r_first = y[0] + S.a * g_lo;
            r_last = y[s * (m - 1)] + S.a * g_hi;
        }
    }
    // Thomas on the reduced system (rows 0 and m-1 carry the folded ghosts)
    const double ylast = thomas_fwd<CHK>(y, s, y, s, m, S.lo, S.inv, true, r_first, r_last);
    const double y0 = thomas_bwd<CHK>(y, s, y, s, m, S.cp, ylast);
    // Sherman-Morrison correction
    const double f = (y0 + S.s * ylast) / S.den;
    const double *z = S.z;
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
    if (S.adj) {
        // adjoints of the ghost-row right-hand sides: the end rows they were folded into
        if (S.off == 0) {
            x[s * m] = ((S.quirk ? S.a * x0 : 0.0) + S.a * xl) + g_hi;
        } else {
            x[0] = S.a * x0 + g_lo;
            x[s * (m + 1)] = S.a * xl + g_hi;
        }
    } else if (S.off == 0) {
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
    int lsh, coff;   // x sweep: log2 lines per block, LDS offset of the coefficients
                     // (partitioned: of the block-end values)
    int kw;          // partitioned wall solve: waves per block (0: twisted / cyclic path)
    int rc;          // partitioned: rows per lane of the kernel (PART_ROWS or PART_ROWS_LONG)
    int xl;          // partitioned x sweep: lines per block
    int gcol;        // twisted y / z sweep: elimination results in the line itself (LDS too small)
    // partitioned solve: right-hand sides of the lines' first and last rows
    // per component (the boundary rows vel_bc_staggered_CN2 would write; one
    // value per line, in line order), or null: the array holds them
    const double *rb[3][2];
    // partitioned solve: the lines' first and last rows go to these (added, one
    // value per line) and zeros to the arrays (the transposed boundary rows of
    // the adjoint), or null: the arrays take them
    double *wb[3][2];
    // periodic lines: gz_in, the ghost rows' right-hand sides are zero (the
    // rows vel_bc_staggered_CN2 zeroes; not read); gz_out, the ghost rows are
    // stored as zeros (the transpose of that, after an adjoint solve)
    int gz_in, gz_out;
    // partitioned y / z sweeps: lanes walk the storage slots of a plane of
    // each component's lines (slot s is element s - xs[c] of the rows of
    // lines, row pitch px), so the 64 lanes of a wavefront read and write
    // whole 128-byte lines of each array row; slots of the spare elements at
    // the row ends idle.  xs[c] is the array origin's offset from its
    // line-aligned start.
    int xs[3];
    int c0;          // first component of the launch (blockIdx.y counts from it)
};

// ---- partitioned wall solve (LineSys::K blocks per line, wave k = block k) ----
//
// A block of kw waves takes 64 lines (lane = line); wave k loads rows
// [s_k, s_{k+1}) of its lane's line into registers, solves them locally,
// publishes the two end values, and after one barrier forms the couplings
// P = x_{s_k - 1}, Q = x_{e_k + 1} from the 2K end values of the line (rows
// of G) and writes x = y - P V - Q W.  Every table address is uniform over the
// wave (the block index k is), so the coefficient loads are scalar loads, and
// a lane holds at most RC rows (PART_ROWS, or PART_ROWS_LONG for long lines),
// so no LDS bounds the resident waves.

// local solve of one block in registers, y[0..R); returns y[R-1] (the last
// forward result, which back substitution keeps)
template <int RC>
__device__ __forceinline__ double part_local(double (&y)[RC], int R, const double *loL, const double *invL,
                                             const double *cpL) {
    double d = 0.0;
#pragma unroll
    for (int j = 0; j < RC; j++)
        if (j < R) {
            d = (y[j] - loL[j] * d) * invL[j];
            y[j] = d;
        }
#pragma unroll
    for (int j = RC - 2; j >= 0; j--)
        if (j + 1 < R) y[j] = y[j] - cpL[j] * y[j + 1];
    return d;
}

// rows [s_k, s_k + R) of a lane's line into registers (get(j): row s_k + j);
// the line's first and last rows come from rb when it is set (line index t)
template <int RC, typename Get>
__device__ __forceinline__ void part_load(double (&y)[RC], int R, bool valid, int k, int K,
                                          const double *const (&rb)[2], unsigned t, Get get) {
    const double *r0 = (k == 0) ? rb[0] : nullptr;  // uniform over the wave
    const double *rl = (k == K - 1) ? rb[1] : nullptr;
    const double first = (r0 && valid) ? r0[t] : 0.0;
    const double last = (rl && valid) ? rl[t] : 0.0;
#pragma unroll
    for (int j = 0; j < RC; j++) {
        const bool f = r0 && j == 0, l = rl && j == R - 1;
        double v = (j < R && valid && !f && !l) ? get(j) : 0.0;
        if (f) v = first;
        if (l) v = last;
        y[j] = v;
    }
}

// stores rows [s_k, s_k + R) of a lane's line (put(j, v): row s_k + j); the
// line's first and last rows are added to wb (line index t) when it is set and
// stored as zeros
template <int RC, typename Put>
__device__ __forceinline__ void part_store(const double (&x)[RC], int R, bool valid, int k, int K,
                                           double *const (&wb)[2], unsigned t, Put put) {
    if (!valid) return;
    double *w0 = (k == 0) ? wb[0] : nullptr;  // uniform over the wave
    double *wl = (k == K - 1) ? wb[1] : nullptr;
#pragma unroll
    for (int j = 0; j < RC; j++) {
        if (j < R) {
            double v = x[j];
            if (w0 && j == 0) {
                w0[t] += v;
                v = 0.0;
            }
            if (wl && j == R - 1) {
                wl[t] += v;
                v = 0.0;
            }
            put(j, v);
        }
    }
}

// couplings of block k from the line's 2K end values (ends[j * 64 + lane])
// (cyclic: block 0's left neighbour is block K-1, block K-1's right block 0)
__device__ __forceinline__ void part_couple(int k, int K, bool cyclic, const double *G, const double *ends, int lane,
                                            double &Pk, double &Qk) {
    const int K2 = 2 * K;
    const int kl = (k > 0) ? k - 1 : K - 1, kr = (k < K - 1) ? k + 1 : 0;
    const double *gp = G + (size_t)(2 * kl + 1) * K2;
    const double *gq = G + (size_t)(2 * kr) * K2;
    Pk = 0.0;
    Qk = 0.0;
    for (int j = 0; j < K2; j++) {
        const double e = ends[j * 64 + lane];
        Pk = Pk + gp[j] * e;
        Qk = Qk + gq[j] * e;
    }
    if (!cyclic && k == 0) Pk = 0.0;
    if (!cyclic && k == K - 1) Qk = 0.0;
}

// the line's first unknown from its 2K end values (row 0 of G: z = G ends)
__device__ __forceinline__ double part_first(int K, const double *G, const double *ends, int lane) {
    double z = 0.0;
    for (int j = 0; j < 2 * K; j++) z = z + G[j] * ends[j * 64 + lane];
    return z;
}

// Periodic lines (reduced cyclic system on the m owned unknowns, rows off ..
// off+m-1 of the line; the ghost rows' right-hand sides glo = row 0 (type B)
// and ghi = row off+m enter the end rows, as in solve_line).  Adds the folds
// to block 0's first and block K-1's last row.
template <int RC>
__device__ __forceinline__ void part_fold(double (&y)[RC], int R, int k, int K, const LineSys &S, double glo,
                                          double ghi) {
    double ff, fl;
    if (S.adj) {
        ff = ghi;
        fl = S.off ? glo : 0.0;
    } else {
        ff = S.off ? S.a * glo : (S.quirk ? S.a * ghi : 0.0);
        fl = S.a * ghi;
    }
#pragma unroll
    for (int j = 0; j < RC; j++) {
        if (k == 0 && j == 0) y[j] = y[j] + ff;
        if (k == K - 1 && j == R - 1) y[j] = y[j] + fl;
    }
}

// the ghost rows of a periodic line after its solve (solve_line's outputs):
// put(r, v) stores row r of the line; x0 / xl are the solution's first and
// last unknowns where this block holds them (z0: the first from the end
// values, for the transposed type-A row, which block K-1 writes)
template <typename Put>
__device__ __forceinline__ void part_ghosts(int k, int K, const LineSys &S, bool zero, double x0, double xl, double z0,
                                            double glo, double ghi, Put put) {
    const int m = S.m;
    if (S.off == 0) {
        if (S.adj) {
            if (k == K - 1) put(m, zero ? 0.0 : ((S.quirk ? S.a * z0 : 0.0) + S.a * xl) + ghi);
        } else if (k == 0) {
            put(m, zero ? 0.0 : x0 + ghi);
        }
        return;
    }
    if (S.adj) {
        if (k == 0) put(0, zero ? 0.0 : S.a * x0 + glo);
        if (k == K - 1) put(m + 1, zero ? 0.0 : S.a * xl + ghi);
    } else {
        if (k == 0) put(m + 1, zero ? 0.0 : x0 + ghi);
        if (k == K - 1) put(0, zero ? 0.0 : xl + glo);
    }
}

// y / z sweeps: lane = line (consecutive x), so every row load of a wave is
// 64 consecutive doubles.  DENSE: the arrays have the dense layout (lane t
// holds line t); otherwise lanes walk the storage slots of the aligned layout
// (SweepParams::xs), so a wave's row accesses start on a 128-byte line.  The
// two are separate instantiations because the slot arithmetic costs the dense
// kernel two registers and with them a resident wave per SIMD.
template <int AX, int RC, bool DENSE>
__device__ __forceinline__ void adi_part_strided_body(const SweepParams &P, const int blk) {
    extern __shared__ double sm[];
    const int c = blockIdx.y + P.c0;
    int n[3];
    comp_dims(P.g, c, n);
    const int o2 = (AX == 2) ? 1 : 2;
    const int px = DENSE ? n[0] : comp_pitch(P.g, c);
    const unsigned xs = DENSE ? 0u : (unsigned)P.xs[c];
    const int stride = (AX == 1) ? px : px * n[1];
    const LineSys &S = P.sys[c];
    const bool per = S.periodic;
    const int K = S.K, N = per ? S.m : S.N, off = per ? S.off : 0;  // N: unknowns per line
    const unsigned nl = (unsigned)px * (unsigned)n[o2] + xs;  // line slots
    if ((unsigned)blk * 64u >= nl) return;  // the whole block: this component has fewer lines
    const int k = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int lane = threadIdx.x & 63;
    const unsigned ts = (unsigned)blk * 64u + lane;
    bool valid = ts < nl;
    const bool act = k < K;
    const int s0 = act ? (k * N) / K : 0;
    const int R = act ? ((k + 1) * N) / K - s0 : 0;
    int id[3] = {1, 1, 1};
    unsigned t = ts;  // the line's index in plane order (x fastest)
    if (DENSE) {
        if (valid) {
            const unsigned q = t / (unsigned)n[0];
            id[0] = (int)(t - q * (unsigned)n[0]) + 1;
            id[o2] = (int)q + 1;
        }
    } else {
        // slot ts holds element e = ts - xs of the plane's rows of lines (e
        // wraps past nl when ts < xs); spare slots at the row ends idle
        const unsigned e = ts - xs;
        valid = e < nl - xs;
        const unsigned q = valid ? e / (unsigned)px : 0u;
        const int xr = (int)(e - q * (unsigned)px);
        valid = valid && xr < n[0];
        if (valid) {
            id[0] = xr + 1;
            id[o2] = (int)q + 1;
        }
        t = (unsigned)xr + (unsigned)n[0] * q;
    }
    double *line = P.a[c] + GIX(id[0], id[1], id[2], px, n[1]);
    double *x = line + (size_t)stride * (off + s0);
    const double *tb = S.pt;
    double glo = 0.0, ghi = 0.0;
    if (per && act && valid && !P.gz_in) {
        if (off) glo = line[0];
        ghi = line[(size_t)stride * (off + N)];
    }
    double y[RC];
    part_load(y, R, valid, k, K, P.rb[c], t, [&](int j) { return x[(size_t)stride * j]; });
    if (per) part_fold(y, R, k, K, S, glo, ghi);
    if (act) {
        const double yl = part_local(y, R, tb + s0, tb + N + s0, tb + 2 * N + s0);
        sm[(2 * k) * 64 + lane] = y[0];
        sm[(2 * k + 1) * 64 + lane] = yl;
    }
    __syncthreads();
    if (!act) return;
    double Pk, Qk;
    part_couple(k, K, per, tb + 5 * N, sm, lane, Pk, Qk);
    const double *V = tb + 3 * N + s0, *W = tb + 4 * N + s0;
    double xl = 0.0;
#pragma unroll
    for (int j = 0; j < RC; j++)
        if (j < R) {
            y[j] = (y[j] - Pk * V[j]) - Qk * W[j];
            if (j == R - 1) xl = y[j];
        }
    part_store(y, R, valid, k, K, P.wb[c], t, [&](int j, double v) { x[(size_t)stride * j] = v; });
    if (per && valid) {
        const double z0 = (S.adj && S.off == 0 && k == K - 1) ? part_first(K, tb + 5 * N, sm, lane) : 0.0;
        part_ghosts(k, K, S, P.gz_out, y[0], xl, z0, glo, ghi,
                    [&](int r, double v) { line[(size_t)stride * r] = v; });
    }
}

// x sweeps: the block stages its P.xl lines (one contiguous span; 64, or 32
// or 16 when 64 long lines would not fit the LDS) in LDS rows of odd length
// LS >= N, coalesced both ways; the block-end values follow at P.coff.
// DENSE: rows N apart; otherwise the aligned layout's storage slots are
// walked (spare slots skipped)
template <int RC, bool DENSE>
__device__ __forceinline__ void adi_part_x_body(const SweepParams &P, const int blk) {
    extern __shared__ double sm[];
    const int c = blockIdx.y + P.c0;
    int n[3];
    comp_dims(P.g, c, n);
    const int N = n[0];  // line length (the LDS rows hold whole lines)
    const LineSys &S = P.sys[c];
    const bool per = S.periodic;
    const int K = S.K, NU = per ? S.m : N, off = per ? S.off : 0;  // NU: unknowns per line
    const int nl = n[1] * n[2];
    const int l0 = blk * P.xl;
    if (l0 >= nl) return;  // the whole block: this component has fewer lines
    const int cnt = (nl - l0 < P.xl) ? nl - l0 : P.xl;
    const int LS = N | 1, px = DENSE ? N : comp_pitch(P.g, c), xs = DENSE ? 0 : P.xs[c];
    // the block's lines are rows l0.. of the array: walked as storage slots
    // from the line-aligned start before row l0 (slot g holds element
    // g - xs of the rows, row pitch px), so every wavefront access starts on
    // a 128-byte line; slots outside the rows' N entries are skipped
    double *a = P.a[c] + (size_t)l0 * px - xs;
    const int tot = cnt * px + xs;
    const float rp = 1.0f / (float)px;  // g < 2^24: float quotient within one
    auto slot = [&](int g) {  // LDS slot of storage slot g, or -1
        const int f = g - xs;
        int l = (int)((float)f * rp);
        if (l * px > f) l--;
        else if ((l + 1) * px <= f) l++;
        const int pos = f - l * px;
        if (DENSE) return l * LS + pos;
        return (f >= 0 && pos < N) ? l * LS + pos : -1;
    };
    const int nth = blockDim.x;
    for (int f0 = threadIdx.x; f0 < tot; f0 += nth * SB) {
        double v[SB];
#pragma unroll
        for (int u = 0; u < SB; u++) v[u] = (f0 + nth * u < tot) ? a[f0 + nth * u] : 0.0;
#pragma unroll
        for (int u = 0; u < SB; u++) {
            if (DENSE) {
                if (f0 + nth * u < tot) sm[slot(f0 + nth * u)] = v[u];
            } else {
                const int q = slot(f0 + nth * u);
                if (f0 + nth * u < tot && q >= 0) sm[q] = v[u];
            }
        }
    }
    __syncthreads();
    const int k = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int lane = threadIdx.x & 63;
    const bool act = k < K;
    const int s0 = act ? (k * NU) / K : 0;
    const int R = act ? ((k + 1) * NU) / K - s0 : 0;
    double *lrow = sm + (lane < cnt ? lane : 0) * LS;  // the lane's whole line
    double *row = lrow + off + s0;
    double *ends = sm + P.coff;
    const double *tb = S.pt;
    double glo = 0.0, ghi = 0.0;
    if (per && act && !P.gz_in) {
        if (off) glo = lrow[0];
        ghi = lrow[off + NU];
    }
    double y[RC];
    part_load(y, R, lane < cnt, k, K, P.rb[c], (unsigned)(l0 + lane), [&](int j) { return row[j]; });
    if (per) part_fold(y, R, k, K, S, glo, ghi);
    if (act) {
        const double yl = part_local(y, R, tb + s0, tb + NU + s0, tb + 2 * NU + s0);
        ends[(2 * k) * 64 + lane] = y[0];
        ends[(2 * k + 1) * 64 + lane] = yl;
    }
    __syncthreads();
    if (act) {
        double Pk, Qk;
        part_couple(k, K, per, tb + 5 * NU, ends, lane, Pk, Qk);
        const double *V = tb + 3 * NU + s0, *W = tb + 4 * NU + s0;
        double xl = 0.0;
#pragma unroll
        for (int j = 0; j < RC; j++)
            if (j < R) {
                y[j] = (y[j] - Pk * V[j]) - Qk * W[j];
                if (j == R - 1) xl = y[j];
            }
        part_store(y, R, lane < cnt, k, K, P.wb[c], (unsigned)(l0 + lane), [&](int j, double v) { row[j] = v; });
        if (per && lane < cnt) {
            const double z0 = (S.adj && S.off == 0 && k == K - 1) ? part_first(K, tb + 5 * NU, ends, lane) : 0.0;
            part_ghosts(k, K, S, P.gz_out, y[0], xl, z0, glo, ghi, [&](int r, double v) { lrow[r] = v; });
        }
    }
    __syncthreads();
    for (int f0 = threadIdx.x; f0 < tot; f0 += nth * SB) {
        double v[SB];
        if (DENSE) {
#pragma unroll
            for (int u = 0; u < SB; u++) v[u] = (f0 + nth * u < tot) ? sm[slot(f0 + nth * u)] : 0.0;
#pragma unroll
            for (int u = 0; u < SB; u++)
                if (f0 + nth * u < tot) a[f0 + nth * u] = v[u];
        } else {
            unsigned ok = 0;  // bit u: slot f0 + nth*u holds an entry
#pragma unroll
            for (int u = 0; u < SB; u++) {
                const int q = slot(f0 + nth * u);
                const bool in = f0 + nth * u < tot && q >= 0;
                ok |= (unsigned)in << u;
                v[u] = in ? sm[q] : 0.0;
            }
#pragma unroll
            for (int u = 0; u < SB; u++)
                if ((ok >> u) & 1u) a[f0 + nth * u] = v[u];
        }
    }
}

// y / z sweeps: one thread per line, lines walked in global memory (the
// threads of a wave hold consecutive x, so every chunk load is coalesced and
// the loads stream while the chain computes)
//
// Wall lines use the twisted factorization: lane l < 32 eliminates rows
// 0..tk-1 downward, lane l + 32 rows N-1..tk+1 upward (the same code on a
// reversed view), they exchange d_{tk-1} and e_{tk+1}, both form x_tk, and
// each substitutes back over its half: chains of about N/2 steps, two lanes
// per line.  Periodic lines: one lane per line (reduced cyclic solve).
template <int AX>
__device__ __forceinline__ void adi_strided_body(const SweepParams &P, const int blk) {
    const int c = blockIdx.y + P.c0;
    int n[3];
    comp_dims(P.g, c, n);
    const int o2 = (AX == 2) ? 1 : 2;
    const int px = comp_pitch(P.g, c);
    const long stride = (AX == 1) ? (long)px : (long)px * n[1];
    const LineSys &S = P.sys[c];
    const size_t nl = (size_t)n[0] * n[o2];
    if (S.periodic) {
        const size_t t = (size_t)blk * 64 + threadIdx.x;
        if (t >= nl) return;
        int id[3];
        id[0] = (int)(t % n[0]) + 1;
        id[o2] = (int)(t / n[0]) + 1;
        id[AX] = 1;
        solve_line<16>(P.a[c] + GIX(id[0], id[1], id[2], px, n[1]), stride, S);
        return;
    }
    const bool top = threadIdx.x < 32;
    const size_t t = (size_t)blk * 32 + (threadIdx.x & 31);
    const bool valid = t < nl;
    int id[3] = {1, 1, 1};
    if (valid) {
        id[0] = (int)(t % n[0]) + 1;
        id[o2] = (int)(t / n[0]) + 1;
    }
    const int N = S.N, tk = S.tk;
    double *x = P.a[c] + GIX(id[0], id[1], id[2], px, n[1]);
    double *base = top ? x : x + stride * (N - 1);
    const long st = top ? stride : -stride;
    const int m = valid ? (top ? tk : N - 1 - tk) : 0;
    // the elimination results of a half line stay in this lane's LDS column,
    // or (lines too long for 64 columns) overwrite the half line in place
    extern __shared__ double sm[];
    double *col = P.gcol ? base : sm + threadIdx.x;
    const long cs = P.gcol ? st : 64;
    const double dk = thomas_fwd<16>(base, st, col, cs, m, top ? S.lo : S.loR, top ? S.inv : S.invR, false, 0.0, 0.0);
    const double other = __shfl_xor(dk, 32);
    double xk = 0.0;
    if (valid) {
        const double dt = top ? dk : other, eb = top ? other : dk;
        xk = ((x[stride * tk] - S.ak * dt) - S.ck * eb) * S.gk;
    }
    if (valid && top) x[stride * tk] = xk;
    thomas_bwd<16>(col, cs, base, st, valid ? (top ? tk + 1 : N - tk) : 0, top ? S.cp : S.cpR, xk);
}

// x sweeps: a block stages L = 2^lsh lines of one component in LDS (x lines
// are contiguous rows, so the block's lines are one contiguous span of
// memory: coalesced loads and stores), with the component's coefficients
// after them; each of its first L threads solves one line there, and the
// block writes the lines back: one read and one write of every element.
__device__ __forceinline__ void adi_x_body(const SweepParams &P, const int blk) {
    extern __shared__ double sm[];
    const int lsh = P.lsh, coff = P.coff;
    const int c = blockIdx.y + P.c0;
    int n[3];
    comp_dims(P.g, c, n);
    const int N = n[0], L = 1 << lsh;
    double *a = P.a[c];
    const LineSys &S0 = P.sys[c];
    const int M = S0.m;
    double *cf = sm + coff;
    for (int f = threadIdx.x; f < M; f += 64) {
        cf[f] = S0.lo[f];
        cf[M + f] = S0.inv[f];
        cf[2 * M + f] = S0.cp[f];
        if (S0.periodic) cf[3 * M + f] = S0.z[f];
    }
    LineSys S = S0;
    S.lo = cf;
    S.inv = cf + M;
    S.cp = cf + 2 * M;
    if (S0.periodic) S.z = cf + 3 * M;
    const long nl = (long)n[1] * n[2];
    const long l0 = (long)blk << lsh;
    if (l0 >= nl) return;  // this component has fewer lines (uniform per block)
    const int cnt = (int)(nl - l0 < L ? nl - l0 : L);
    const int LS = N + 1, px = comp_pitch(P.g, c);
    // the block's rows walked as storage slots from the row origin (slot f:
    // element f of the rows, pitch px); slots past a row's N entries skipped
    const size_t base = (size_t)l0 * px;
    const int tot = cnt * px;
    const float rp = 1.0f / (float)px;  // f < 2^24: float quotient within one
    auto slot = [&](int f) {  // LDS slot of storage slot f, or -1
        int l = (int)((float)f * rp);
        if (l * px > f) l--;
        else if ((l + 1) * px <= f) l++;
        const int pos = f - l * px;
        return pos < N ? l * LS + pos : -1;
    };
    for (int f0 = threadIdx.x; f0 < tot; f0 += 64 * SB) {
        double v[SB];
#pragma unroll
        for (int u = 0; u < SB; u++) v[u] = (f0 + 64 * u < tot) ? a[base + f0 + 64 * u] : 0.0;
#pragma unroll
        for (int u = 0; u < SB; u++) {
            const int q = slot(f0 + 64 * u);
            if (f0 + 64 * u < tot && q >= 0) sm[q] = v[u];
        }
    }
    __syncthreads();
    if (S.periodic) {
        if ((int)threadIdx.x < cnt) solve_line<8>(sm + threadIdx.x * LS, 1, S);
    } else {
        // twisted factorization, two lanes per line (L <= 32): lane l < 32 the
        // top half of line l, lane l + 32 its bottom half
        const bool top = threadIdx.x < 32;
        const int l = threadIdx.x & 31;
        const bool valid = l < cnt;
        const int tk = S.tk;
        double *row = sm + (valid ? l : 0) * LS;
        double *base = top ? row : row + (N - 1);
        const long st = top ? 1 : -1;
        const int m = valid ? (top ? tk : N - 1 - tk) : 0;
        const double dk = thomas_fwd<8>(base, st, base, st, m, top ? S.lo : S.loR, top ? S.inv : S.invR, false, 0.0,
                                        0.0);
        const double other = __shfl_xor(dk, 32);
        double xk = 0.0;
        if (valid) {
            const double dt = top ? dk : other, eb = top ? other : dk;
            xk = ((row[tk] - S.ak * dt) - S.ck * eb) * S.gk;
        }
        __syncthreads();  // both lanes have read row[tk]
        if (valid && top) row[tk] = xk;
        thomas_bwd<8>(base, st, base, st, valid ? (top ? tk + 1 : N - tk) : 0, top ? S.cp : S.cpR, xk);
    }
    __syncthreads();
    for (int f0 = threadIdx.x; f0 < tot; f0 += 64 * SB) {
        double v[SB];
        unsigned ok = 0;  // bit u: slot f0 + 64*u holds an entry
#pragma unroll
        for (int u = 0; u < SB; u++) {
            const int q = slot(f0 + 64 * u);
            const bool in = f0 + 64 * u < tot && q >= 0;
            ok |= (unsigned)in << u;
            v[u] = in ? sm[q] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < SB; u++)
            if ((ok >> u) & 1u) a[base + f0 + 64 * u] = v[u];
    }
}

template <int AX>
__global__ void __launch_bounds__(64) k_adi_strided(SweepParams P) {
    adi_strided_body<AX>(P, blockIdx.x);
}
__global__ void __launch_bounds__(64) k_adi_x(SweepParams P) { adi_x_body(P, blockIdx.x); }

constexpr int PART_THREADS = 64 * PART_MAX_BLOCKS;
template <int AX, int RC, bool DENSE>
__global__ void __launch_bounds__(PART_THREADS) k_adi_part_strided(SweepParams P) {
    adi_part_strided_body<AX, RC, DENSE>(P, blockIdx.x);
}
template <int RC, bool DENSE>
__global__ void __launch_bounds__(PART_THREADS) k_adi_part_x(SweepParams P) {
    adi_part_x_body<RC, DENSE>(P, blockIdx.x);
}
template <int AX, int RC, bool DENSE>
__global__ void __launch_bounds__(PART_THREADS) k_adi_part_strided_batch(const SweepParams *__restrict__ Ps,
                                                                                  const int *__restrict__ prefix,
                                                                                  int nbox) {
    int blk;
    const SweepParams &P = batch_params(Ps, prefix, nbox, blk);
    adi_part_strided_body<AX, RC, DENSE>(P, blk);
}
template <int RC, bool DENSE>
__global__ void __launch_bounds__(PART_THREADS) k_adi_part_x_batch(const SweepParams *__restrict__ Ps,
                                                                           const int *__restrict__ prefix, int nbox) {
    int blk;
    const SweepParams &P = batch_params(Ps, prefix, nbox, blk);
    adi_part_x_body<RC, DENSE>(P, blk);
}


template <int AX>
__global__ void __launch_bounds__(64) k_adi_strided_batch(const SweepParams *__restrict__ Ps,
                                                          const int *__restrict__ prefix, int nbox) {
    int blk;
    const SweepParams &P = batch_params(Ps, prefix, nbox, blk);
    adi_strided_body<AX>(P, blk);
}
__global__ void __launch_bounds__(64) k_adi_x_batch(const SweepParams *__restrict__ Ps, const int *__restrict__ prefix,
                                                    int nbox) {
    int blk;
    const SweepParams &P = batch_params(Ps, prefix, nbox, blk);
    adi_x_body(P, blk);
}

template <typename K>
void set_lds(K kern, size_t lds) {
    if (lds > 65536) (void)hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
}

// one partitioned sweep launch with the kernel for RC rows per lane and
// the layout (DENSE: every array of the launch dense)
template <int RC, bool DENSE>
void launch_part_l(int ax, dim3 grid, dim3 blk, size_t lds, hipStream_t st, const SweepParams *P,
                   const SweepParams *Ps, const int *prefix, int nbox) {
    if (ax == 0) {
        if (P) {
            set_lds(k_adi_part_x<RC, DENSE>, lds);
            hipLaunchKernelGGL((k_adi_part_x<RC, DENSE>), grid, blk, lds, st, *P);
        } else {
            set_lds(k_adi_part_x_batch<RC, DENSE>, lds);
            hipLaunchKernelGGL((k_adi_part_x_batch<RC, DENSE>), grid, blk, lds, st, Ps, prefix, nbox);
        }
    } else if (ax == 1) {
        if (P) hipLaunchKernelGGL((k_adi_part_strided<1, RC, DENSE>), grid, blk, lds, st, *P);
        else hipLaunchKernelGGL((k_adi_part_strided_batch<1, RC, DENSE>), grid, blk, lds, st, Ps, prefix, nbox);
    } else {
        if (P) hipLaunchKernelGGL((k_adi_part_strided<2, RC, DENSE>), grid, blk, lds, st, *P);
        else hipLaunchKernelGGL((k_adi_part_strided_batch<2, RC, DENSE>), grid, blk, lds, st, Ps, prefix, nbox);
    }
}
template <int RC>
void launch_part(int ax, bool dense, dim3 grid, dim3 blk, size_t lds, hipStream_t st, const SweepParams *P,
                 const SweepParams *Ps, const int *prefix, int nbox) {
    if (dense) launch_part_l<RC, true>(ax, grid, blk, lds, st, P, Ps, prefix, nbox);
    else launch_part_l<RC, false>(ax, grid, blk, lds, st, P, Ps, prefix, nbox);
}

}  // namespace

size_t sweep_params_size() { return sizeof(SweepParams); }

// parameters of the sweep along axis dir-1; returns the block count (grid.y = 3
// components) and the dynamic LDS size
int make_sweep_params(const Grid &g, int dir, const LineSys &sysA, const LineSys &sysB, double *u, double *v, double *w,
                      void *out, size_t *lds, int x_lsh, bool part, const double *planes, const PlaneLayout *L,
                      double *planes_out) {
    SweepParams &P = *(SweepParams *)out;
    P = SweepParams{};
    P.g = g;
    P.ax = dir - 1;
    for (int c = 0; c < 3; c++) P.sys[c] = (c == P.ax) ? sysA : sysB;
    P.a[0] = u; P.a[1] = v; P.a[2] = w;
    for (int c = 0; c < 3; c++) P.xs[c] = field_origin(g, c);
    const int ax = P.ax;
    if (part && sweep_partitioned(sysA, sysB)) {
        // 64 lines per block (x: P.xl), one wave per partition block
        P.kw = sysA.K > sysB.K ? sysA.K : sysB.K;
        P.rc = sysA.R > sysB.R ? sysA.R : sysB.R;
        size_t nl = 0;
        int lsmax = 0;
        for (int c = 0; c < 3; c++) {
            int n[3];
            comp_dims(g, c, n);
            // y / z: storage slots of a plane of lines (origin offset included)
            const size_t m = (ax == 0) ? (size_t)n[1] * n[2]
                                       : (size_t)comp_pitch(g, c) * ((ax == 1) ? n[2] : n[1]) + P.xs[c];
            nl = m > nl ? m : nl;
            lsmax = (n[0] | 1) > lsmax ? (n[0] | 1) : lsmax;
        }
        // x: the staged lines and the end values within the LDS (lines of at
        // most PART_MAX_BLOCKS * PART_ROWS_LONG + 2 elements fit 16 at a time)
        P.xl = 64;
        while (ax == 0 && P.xl > 16 && ((size_t)P.xl * lsmax + (size_t)2 * P.kw * 64) * sizeof(double) > LDS_BYTES)
            P.xl /= 2;
        P.coff = (ax == 0) ? P.xl * lsmax : 0;
        *lds = ((size_t)P.coff + (size_t)2 * P.kw * 64) * sizeof(double);
        if (sysA.periodic) {
            // the boundary rows of periodic lines are their zeroed ghost rows
            P.gz_in = planes ? 1 : 0;
            P.gz_out = planes_out ? 1 : 0;
        } else {
            for (int c = 0; c < 3; c++) {
                if (planes) {
                    P.rb[c][0] = planes + L->off[6 * c + 2 * ax];
                    P.rb[c][1] = planes + L->off[6 * c + 2 * ax + 1];
                }
                if (planes_out) {
                    P.wb[c][0] = planes_out + L->off[6 * c + 2 * ax];
                    P.wb[c][1] = planes_out + L->off[6 * c + 2 * ax + 1];
                }
            }
        }
        return (int)((nl + P.xl - 1) / P.xl);
    }
    if (ax == 0) {
        // longest x line, lines per block (up to 64 within 64 KiB), blocks of the largest component
        const int nmax = g.nx + 2;
        const size_t row = (size_t)nmax + 1;
        int lsh = (x_lsh >= 0 && x_lsh < 6) ? x_lsh : 6;
        if (!sysA.periodic && lsh > 5) lsh = 5;  // wall lines: two lanes per line
        while (lsh > 0 && ((size_t)1 << lsh) * row * sizeof(double) > 65536) lsh--;
        const size_t L = (size_t)1 << lsh;
        P.lsh = lsh;
        P.coff = (int)(L * row);
        *lds = ((size_t)P.coff + 4 * (size_t)nmax) * sizeof(double);
        return (int)(((size_t)(g.ny + 2) * (g.nz + 2) + L - 1) / L);
    }
    size_t nl = 0;
    for (int c = 0; c < 3; c++) {
        int n[3];
        comp_dims(g, c, n);
        const size_t m = (size_t)n[0] * ((ax == 1) ? n[2] : n[1]);
        nl = m > nl ? m : nl;
    }
    // wall lines: two lanes per line (twisted factorization), each with an
    // LDS column for its half line's elimination results
    int nmax = 0;
    for (int c = 0; c < 3; c++) {
        int n[3];
        comp_dims(g, c, n);
        nmax = n[ax] > nmax ? n[ax] : nmax;
    }
    *lds = sysA.periodic ? 0 : (size_t)64 * (nmax / 2 + 1) * sizeof(double);
    if (*lds > LDS_BYTES) {
        P.gcol = 1;
        *lds = 0;
    }
    const size_t per = sysA.periodic ? 64 : 32;
    return (int)((nl + per - 1) / per);
}

bool sweep_partitioned(const LineSys &sysA, const LineSys &sysB) { return sysA.K > 0 && sysB.K > 0; }

int sweep_params_waves(const void *params) { return ((const SweepParams *)params)->kw; }

int sweep_params_rows(const void *params) { return ((const SweepParams *)params)->rc; }

void launch_adi_batch(int ax, int kw, int rows, bool dense, const void *params, const int *prefix, int nbox,
                      int nblk, size_t lds, hipStream_t st) {
    const SweepParams *Ps = (const SweepParams *)params;
    if (kw > 0) {
        const dim3 grid(nblk, 3), blk(64 * kw);
        if (rows > PART_ROWS) launch_part<PART_ROWS_LONG>(ax, dense, grid, blk, lds, st, nullptr, Ps, prefix, nbox);
        else launch_part<PART_ROWS>(ax, dense, grid, blk, lds, st, nullptr, Ps, prefix, nbox);
        return;
    }
    if (ax == 0) {
        set_lds(k_adi_x_batch, lds);
        hipLaunchKernelGGL(k_adi_x_batch, dim3(nblk, 3), dim3(64), lds, st, Ps, prefix, nbox);
    } else if (ax == 1) {
        set_lds(k_adi_strided_batch<1>, lds);
        hipLaunchKernelGGL(k_adi_strided_batch<1>, dim3(nblk, 3), dim3(64), lds, st, Ps, prefix, nbox);
    } else {
        set_lds(k_adi_strided_batch<2>, lds);
        hipLaunchKernelGGL(k_adi_strided_batch<2>, dim3(nblk, 3), dim3(64), lds, st, Ps, prefix, nbox);
    }
}

void launch_adi_sweep(const Grid &g, int dir, const LineSys &sysA, const LineSys &sysB, double *u, double *v, double *w,
                      hipStream_t st, const double *planes, const PlaneLayout *L, double *planes_out, int c0, int nc) {
    SweepParams P;
    size_t lds = 0;
    const int nb = make_sweep_params(g, dir, sysA, sysB, u, v, w, &P, &lds, 6, true, planes, L, planes_out);
    P.c0 = c0;
    if (P.kw > 0) {
        const dim3 grid(nb, nc), blk(64 * P.kw);
        const bool dense = g.dense();
        if (P.rc > PART_ROWS) launch_part<PART_ROWS_LONG>(dir - 1, dense, grid, blk, lds, st, &P, nullptr, nullptr, 0);
        else launch_part<PART_ROWS>(dir - 1, dense, grid, blk, lds, st, &P, nullptr, nullptr, 0);
        return;
    }
    if (dir == 1) {
        set_lds(k_adi_x, lds);
        hipLaunchKernelGGL(k_adi_x, dim3(nb, nc), dim3(64), lds, st, P);
    } else if (dir == 2) {
        set_lds(k_adi_strided<1>, lds);
        hipLaunchKernelGGL(k_adi_strided<1>, dim3(nb, nc), dim3(64), lds, st, P);
    } else {
        set_lds(k_adi_strided<2>, lds);
        hipLaunchKernelGGL(k_adi_strided<2>, dim3(nb, nc), dim3(64), lds, st, P);
    }
}

}  // namespace ns
