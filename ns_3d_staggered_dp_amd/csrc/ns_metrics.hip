// Re-simulation error metrics (re_simulation_sub_tstep_interBC_combine.f90:
// 826-833, other_utility.f90:110-129) in one pass over the ghosted arrays:
// per block, for u, v, w, p the max and sum of |x - x_sub| and the sum of
// x_sub^2, and for the z vorticity oz = diff(v,1,1)/dx - diff(u,1,2)/dy the
// max of |oz - oz_sub| and the sum of oz_sub^2.  The host combines the block
// partials (max, and sums in extended precision).
//
// ns_error_metrics_noalloc keeps its block partials in `scratch`, a
// file-scope device array (one copy per GPU, held until the process ends),
// and combines them on the device, so it needs no allocation per call.
#include <mutex>

#include "ns_internal.h"

namespace ns {

// block partials [NOALLOC_METRIC_BLOCKS][NS_METRIC_PARTS], then the combined
// row (2049 x 14 doubles = 229 KB).  In host code `scratch` names a host-side
// stand-in: host code uses only the address metric_scratch() returns.
static __device__ __align__(256) double scratch[(NOALLOC_METRIC_BLOCKS + 1) * NS_METRIC_PARTS];

namespace {

constexpr int MT = 256;

struct MetricParams {
    Grid g;              // layout of a (the run's arrays); b is dense
    double dx, dy;
    const double *a[4];  // u, v, w, p of the run
    const double *b[4];  // u_sub, v_sub, w_sub, p_sub of the parent
    size_t n[4], noz, total;  // entries of each array, of oz, and the most of them
    int n0[4], pa[4];    // row length and a's row pitch of each component
    double *partial;     // [block][NS_METRIC_PARTS]
};

__global__ void __launch_bounds__(MT) k_err_partials(MetricParams P) {
    double acc[NS_METRIC_PARTS];
#pragma unroll
    for (int q = 0; q < NS_METRIC_PARTS; q++) acc[q] = 0.0;
    const Grid g = P.g;
    const int U1 = g.nx + 1, U2 = g.ny + 2, V1 = g.nx + 2, V2 = g.ny + 1;
    const int O1 = g.nx + 1, O2 = g.ny + 1;  // oz is (nx+1, ny+1, nz+2)
    for (size_t e = (size_t)blockIdx.x * MT + threadIdx.x; e < P.total; e += (size_t)gridDim.x * MT) {
#pragma unroll
        for (int c = 0; c < 4; c++) {
            if (e < P.n[c]) {
                // entry e of the dense layout; in a, its row r = e / n0 starts pa * r
                const size_t r = e / (size_t)P.n0[c];
                const double x = P.a[c][e + r * (size_t)(P.pa[c] - P.n0[c])], y = P.b[c][e];
                const double d = fabs(x - y);
                acc[3 * c] = fmax(acc[3 * c], d);
                acc[3 * c + 1] += d;
                acc[3 * c + 2] += y * y;
            }
        }
        if (e < P.noz) {
            const int i = (int)(e % O1) + 1;
            const int j = (int)((e / O1) % O2) + 1;
            const int k = (int)(e / ((size_t)O1 * O2)) + 1;
            const size_t v0 = GIX(i, j, k, V1, V2), v1 = GIX(i + 1, j, k, V1, V2);
            const size_t u0 = GIX(i, j, k, U1, U2), u1 = GIX(i, j + 1, k, U1, U2);
            const size_t av0 = GIX(i, j, k, P.pa[1], V2), av1 = GIX(i + 1, j, k, P.pa[1], V2);
            const size_t au0 = GIX(i, j, k, P.pa[0], U2), au1 = GIX(i, j + 1, k, P.pa[0], U2);
            const double oz = (P.a[1][av1] - P.a[1][av0]) / P.dx - (P.a[0][au1] - P.a[0][au0]) / P.dy;
            const double ozs = (P.b[1][v1] - P.b[1][v0]) / P.dx - (P.b[0][u1] - P.b[0][u0]) / P.dy;
            acc[12] = fmax(acc[12], fabs(oz - ozs));
            acc[13] += ozs * ozs;
        }
    }
    __shared__ double red[MT];
    for (int q = 0; q < NS_METRIC_PARTS; q++) {
        const bool is_max = (q == 12) || (q < 12 && q % 3 == 0);
        red[threadIdx.x] = acc[q];
        __syncthreads();
        for (int s = MT / 2; s > 0; s >>= 1) {
            if ((int)threadIdx.x < s)
                red[threadIdx.x] = is_max ? fmax(red[threadIdx.x], red[threadIdx.x + s])
                                          : red[threadIdx.x] + red[threadIdx.x + s];
            __syncthreads();
        }
        if (threadIdx.x == 0) P.partial[(size_t)blockIdx.x * NS_METRIC_PARTS + q] = red[0];
        __syncthreads();
    }
}

}  // namespace

int launch_err_partials(const Grid &g, double dx, double dy, const double *const a[4], const double *const b[4],
                        double *partial, int max_blocks, hipStream_t st) {
    MetricParams P;
    P.g = g;
    P.dx = dx;
    P.dy = dy;
    const size_t n[4] = {g.ref_nu(), g.ref_nv(), g.ref_nw(), g.ref_np()};
    P.total = 0;
    for (int c = 0; c < 4; c++) {
        P.a[c] = a[c];
        P.b[c] = b[c];
        P.n[c] = n[c];
        P.n0[c] = (c == 0) ? g.nx + 1 : g.nx + 2;
        P.pa[c] = comp_pitch(g, c);
        P.total = n[c] > P.total ? n[c] : P.total;
    }
    P.noz = (size_t)(g.nx + 1) * (g.ny + 1) * (g.nz + 2);
    P.total = P.noz > P.total ? P.noz : P.total;
    P.partial = partial;
    size_t nb = (P.total + MT - 1) / MT;
    if (nb > (size_t)max_blocks) nb = max_blocks;
    hipLaunchKernelGGL(k_err_partials, dim3((unsigned)nb), dim3(MT), 0, st, P);
    return (int)nb;
}

namespace {
// one block: out[q] = max or sum over the nb rows of partial (max for the
// max|a-b| entries, sum otherwise), rows taken in a fixed order
__global__ void __launch_bounds__(MT) k_err_combine(const double *__restrict__ partial, int nb,
                                                    double *__restrict__ out) {
    __shared__ double red[MT];
    for (int q = 0; q < NS_METRIC_PARTS; q++) {
        const bool is_max = (q == 12) || (q < 12 && q % 3 == 0);
        double a = 0.0;
        for (int r = threadIdx.x; r < nb; r += MT) {
            const double x = partial[(size_t)r * NS_METRIC_PARTS + q];
            a = is_max ? fmax(a, x) : a + x;
        }
        red[threadIdx.x] = a;
        __syncthreads();
        for (int s = MT / 2; s > 0; s >>= 1) {
            if ((int)threadIdx.x < s)
                red[threadIdx.x] = is_max ? fmax(red[threadIdx.x], red[threadIdx.x + s])
                                          : red[threadIdx.x] + red[threadIdx.x + s];
            __syncthreads();
        }
        if (threadIdx.x == 0) out[q] = red[0];
        __syncthreads();
    }
}
}  // namespace

void launch_err_combine(const double *partial, int nb, double *out, hipStream_t st) {
    hipLaunchKernelGGL(k_err_combine, dim3(1), dim3(MT), 0, st, partial, nb, out);
}

namespace {
// dst[i] = src[i]: a kernel of known traffic (reads and writes 8 B per element)
__global__ void __launch_bounds__(256) k_stream_copy(const double *__restrict__ src, double *__restrict__ dst, long n) {
    for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long)gridDim.x * 256) dst[i] = src[i];
}
}  // namespace

void launch_stream_copy(const double *src, double *dst, long n, hipStream_t st) {
    long nb = (n + 255) / 256;
    if (nb > 2048) nb = 2048;
    if (nb < 1) nb = 1;
    hipLaunchKernelGGL(k_stream_copy, dim3((unsigned)nb), dim3(256), 0, st, src, dst, n);
}

int metric_scratch(double **out) {
    // device address of `scratch` on the current device, looked up once per device
    constexpr int MAXDEV = 64;
    static std::mutex mtx;
    static double *addr[MAXDEV] = {};
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess) {
        (void)hipGetLastError();
        return NS_ERR_HIP;
    }
    if (dev < 0 || dev >= MAXDEV) return NS_ERR_UNSUPPORTED;
    std::lock_guard<std::mutex> lock(mtx);
    if (!addr[dev]) {
        double *ptr = nullptr;
        if (hipGetSymbolAddress((void **)&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr) {
            (void)hipGetLastError();
            return NS_ERR_HIP;
        }
        addr[dev] = ptr;
    }
    *out = addr[dev];
    return NS_OK;
}

}  // namespace ns
