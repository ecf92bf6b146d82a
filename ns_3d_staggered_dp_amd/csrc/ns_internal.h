// Internal declarations shared by the HIP translation units of libns_hip.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "ns_hip.h"

// 1-based Fortran index into a column-major array whose x rows are px
// elements apart and whose planes hold n2 rows (px = n1 for a dense (n1,n2,*)
// array)
#define GIX(i, j, k, px, n2) ((size_t)((i) - 1) + (size_t)(px) * ((size_t)((j) - 1) + (size_t)(n2) * (size_t)((k) - 1)))

namespace ns {

// Box dimensions and the memory layout of its ghosted arrays.  pu is the x
// row pitch (in elements) of the u-shaped arrays (nx+1 entries per row), pc
// that of v, w, p (nx+2 entries); planes hold exactly ny+2 (v: ny+1) rows, so
// an array spans pitch * rows * planes elements from its origin.  The
// reference layout has pu = nx+1, pc = nx+2 (dense arrays, as a Fortran host
// holds them); a context keeps its own state in the aligned layout of
// dev_grid().
struct Grid {
    int nx, ny, nz;
    int pu, pc;
    Grid() = default;
    __host__ __device__ Grid(int x, int y, int z, int pu_, int pc_) : nx(x), ny(y), nz(z), pu(pu_), pc(pc_) {}
    // elements from an array's origin to the end of its last row
    __host__ __device__ size_t nu() const { return (size_t)pu * (ny + 2) * (nz + 2); }
    __host__ __device__ size_t nv() const { return (size_t)pc * (ny + 1) * (nz + 2); }
    __host__ __device__ size_t nw() const { return (size_t)pc * (ny + 2) * (nz + 1); }
    __host__ __device__ size_t np() const { return (size_t)pc * (ny + 2) * (nz + 2); }
    __host__ __device__ size_t ni() const { return (size_t)nx * ny * nz; }
    // entries of the reference (dense) layout
    __host__ __device__ size_t ref_nu() const { return (size_t)(nx + 1) * (ny + 2) * (nz + 2); }
    __host__ __device__ size_t ref_nv() const { return (size_t)(nx + 2) * (ny + 1) * (nz + 2); }
    __host__ __device__ size_t ref_nw() const { return (size_t)(nx + 2) * (ny + 2) * (nz + 1); }
    __host__ __device__ size_t ref_np() const { return (size_t)(nx + 2) * (ny + 2) * (nz + 2); }
    __host__ __device__ bool dense() const { return pu == nx + 1 && pc == nx + 2; }
};

// row pitch of a component's arrays: c = 0 u, 1 v, 2 w, 3 p
__host__ __device__ inline int comp_pitch(const Grid &g, int c) { return c == 0 ? g.pu : g.pc; }

// The reference layout (dense Fortran arrays).
inline Grid ref_grid(int nx, int ny, int nz) { return Grid(nx, ny, nz, nx + 1, nx + 2); }

// A context's own layout.  Row pitches are a whole number of 128-byte lines
// (an odd number, so that the rows of a plane spread over the memory
// channels), and the array origins are placed so that the first computed
// point of every row is the first element of a line: u(1) (u is computed
// from face 1 for periodic x, face 2 for walls) and v(2), w(2), p(2).  A
// wavefront's stores of 64 consecutive points of a row then cover whole
// lines; on MI355X a stream whose stores straddle line boundaries runs about
// 40% slower (3.7 against 5.3 TB/s for 7 reads and 6 writes per point).
// The spare elements at the row ends are never read as data.
constexpr int ROW_ALIGN = 16;  // doubles per 128-byte line
inline int ghost_pitch(int nx) {
    int lines = (nx + 2 + ROW_ALIGN - 1) / ROW_ALIGN;
    if ((lines & 1) == 0) lines++;
    return lines * ROW_ALIGN;
}
// The aligned layout costs pitch - (nx+2) spare elements per row.  Boxes
// narrower than 222 cells, or whose rows it would lengthen by more than 1/16,
// keep the dense layout: there the spares are a large share of every sweep's
// lanes and staged rows (128 wide: 11%, 32 wide: 41%), and such boxes'
// arrays largely stay in the 256 MB Infinity Cache, where partial-line stores
// cost little (measured at 128^3: walls AB2-CN 0.276 ms dense, 0.300 ms
// aligned; at 256^3 the aligned periodic AB2 step is 2-4% faster).
inline Grid dev_grid(int nx, int ny, int nz) {
    const int p = ghost_pitch(nx);
    if (nx < 222 || 16 * (p - (nx + 2)) > nx + 2) return ref_grid(nx, ny, nz);
    return Grid(nx, ny, nz, p, p);
}
// offset (elements) of a component's array origin from its line-aligned allocation
inline int field_origin(const Grid &g, int c) { return (g.dense() || c == 0) ? 0 : ROW_ALIGN - 1; }
// elements to allocate for a component's array (origin offset included)
inline size_t field_alloc(const Grid &g, int c) {
    const size_t n = c == 0 ? g.nu() : c == 1 ? g.nv() : c == 2 ? g.nw() : g.np();
    return n + (g.dense() ? 0 : ROW_ALIGN);
}

// x / d for x < 2^32 and 1 <= d < 2^20, with rd = 1.0/d: the double product
// is within 2^-20 of x/d, so truncation gives the quotient or, when d divides
// x, one less; one correction step fixes that
__device__ __forceinline__ unsigned udiv_rd(unsigned x, unsigned d, double rd) {
    unsigned q = (unsigned)((double)x * rd);
    const long long rem = (long long)x - (long long)q * d;
    if (rem < 0) q--;
    else if (rem >= (long long)d) q++;
    return q;
}

// Blocks are dealt to the 8 XCDs round robin (block b on XCD b mod 8), each
// with its own L2.  For a flat stencil launch of nb blocks (nb a multiple of
// 8), this logical order gives every XCD one contiguous slab of the index
// space, so the neighbour planes a block reads were fetched by the same XCD.
__device__ __forceinline__ int xcd_slab_block(int b, int nb) { return (b & 7) * (nb >> 3) + (b >> 3); }
inline unsigned int round8(unsigned int nb) { return (nb + 7u) & ~7u; }

// component dims: c=0 u, 1 v, 2 w, 3 p
// (written without a run-time array index, which would place n in scratch memory)
__host__ __device__ inline void comp_dims(const Grid &g, int c, int n[3]) {
    n[0] = g.nx + 2 - (c == 0);
    n[1] = g.ny + 2 - (c == 1);
    n[2] = g.nz + 2 - (c == 2);
}

// packed plane-set layout (18 velocity + 6 pressure planes + mean_dp)
struct PlaneLayout {
    long off[25];
    long size[24];
    long total;  // including the trailing scalar
};
PlaneLayout plane_layout(int nx, int ny, int nz);

// Stockham plan for one transform length, radices in {2,3,4,5,7,8}
struct FFTPlan {
    int n;
    int nst;
    int radix[10];
    const double2 *tw;  // device table, tw[m] = exp(-2 pi i m / n), m = 0..n-1
};

// Tridiagonal line system of one ADI direction and one matrix type
// (A: the face-normal component, n+1 long; B: the other two, n+2 long).
struct LineSys {
    int N;        // line length (n+1 or n+2)
    int periodic; // 1: reduced cyclic system on m = n unknowns
    int m;        // number of unknowns solved by the Thomas sweep
    int off;      // index (0-based) of the first unknown in the line
    double a;     // CN off-diagonal magnitude: 0.5*dt*nu/h^2
    int quirk;    // periodic A-type with mat_CN's row 1
    int adj;      // periodic: the transposed map (ghost rows added into the end rows
                  // before the solve, ghost outputs formed from the solved ends after it)
    // Thomas coefficients for the m unknowns: d'_i = (r_i - lo_i d'_{i-1}) * inv_i ; x_i = d'_i - cp_i x_{i+1}
    const double *lo, *inv, *cp;
    // Sherman-Morrison correction (periodic only): x -= f * z, f = (y_0 + s*y_{m-1}) / den
    const double *z;
    double s, den;
    // twisted factorization (walls only): rows 0..tk-1 eliminated top-down with
    // lo, inv, cp; rows N-1..tk+1 bottom-up with loR, invR, cpR indexed from the
    // bottom (loR[i'] = up_{N-1-i'}); row tk: x = (r - ak d_{tk-1} - ck e_{tk+1}) * gk
    int tk;
    double ak, ck, gk;
    const double *loR, *invR, *cpR;
    // partitioned solve: K blocks of rows [kN/K, (k+1)N/K) of the N unknowns
    // (walls: the whole line; periodic: the m owned unknowns, the first and
    // last blocks coupled), one wave per block; pt = loL | invL | cpL | V | W
    // (N each) | G (2K x 2K); K = 0: not partitioned.  R: the most rows of a
    // block (PART_ROWS, or PART_ROWS_LONG for lines of more than
    // PART_MAX_BLOCKS * PART_ROWS unknowns)
    int K;
    int R;
    const double *pt;
};
// partitioned solve: rows per block (registers per lane) and blocks per line;
// lines too long for PART_MAX_BLOCKS blocks of PART_ROWS rows take blocks of
// up to PART_ROWS_LONG rows (more registers per lane, fewer resident waves)
constexpr int PART_ROWS = 28;
constexpr int PART_ROWS_LONG = 40;
constexpr int PART_MAX_BLOCKS = 16;
// LDS of one workgroup on gfx950
constexpr size_t LDS_BYTES = 160 * 1024;

struct Ctx;

// ---- stencil / BC / projection launchers (ns_stencil.hip) ----
// fused predictor (ns_predict.hip); lazy: (u,v,w) are the previous u*,v*,w* and the
// projection with dp (and the corner gauge, if non-null) is applied on load;
// wrap (all-periodic boxes): the velocity's ghost entries are read through
// their periodic images, so they need not have been filled
void launch_predict_tiled(const ns_config &c, const Grid &g, int first, bool lazy, const double *u, const double *v, const double *w,
                          const double *dp, const double *gauge, const double *p, double *hx, double *hy, double *hz,
                          double *us, double *vs, double *ws, hipStream_t st, bool wrap = false);

// ---- batched launches: many boxes, one launch per kernel ----
// box of block blk: the last i with prefix[i] <= blk (prefix ascending, prefix[0] = 0)
// The prefix and parameter tables are not written while a batch kernel runs,
// so they are read through the constant address space: the loads are
// uniform within a block and become scalar loads, as for kernel arguments.
typedef const __attribute__((address_space(4))) int const_int;
__device__ __forceinline__ int batch_box(const int *prefix, int nbox, int blk) {
    const_int *pc = (const_int *)prefix;
    int lo = 0, hi = nbox - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (pc[mid] <= blk) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}
// parameters of block blockIdx.x's box; blk = the block's index within the box
template <typename T>
__device__ __forceinline__ const T &batch_params(const T *params, const int *prefix, int nbox, int &blk) {
    typedef const __attribute__((address_space(4))) T const_T;
    const int box = batch_box(prefix, nbox, (int)blockIdx.x);
    blk = (int)blockIdx.x - ((const_int *)prefix)[box];
    return *(const T *)((const_T *)params + box);
}
// predictor parameters of one box written to out (predict_params_size() bytes); returns its
// block count.  zc_hint: z cells per block (0: chosen for the box alone on the device)
size_t predict_params_size();
int make_predict_params(const ns_config &c, const Grid &g, int first, bool lazy, const double *u, const double *v, const double *w,
                        const double *dp, const double *gauge, const double *p, double *hx, double *hy, double *hz,
                        double *us, double *vs, double *ws, void *out, int zc_hint, bool wrap = false,
                        int tile = 0);
// tile columns x z cells of a box (the predictor's serial work units) and resident predictor blocks
// tile width of a box's predictor launch: tile 64 or 32 as asked, 0: 32 x 8
// for boxes up to 128 wide, else the shape with fewer idle lanes (64 x 4 on ties)
int predict_shape(const ns_config &c, int tile);
long predict_tile_columns(const ns_config &c, int tile = 0);
int predict_resident_blocks(bool lazy, int tile = 64);
void launch_predict_batch(bool lazy, int tile, const void *params, const int *prefix, int nbox, int nblocks,
                          hipStream_t st);
size_t bc_params_size();
size_t proj_params_size();
// fill / projection parameters of one box; return the block count (0: nothing to launch)
int make_project_params(const Grid &g, const double d[3], double dt, int dp_flag, const double *us, const double *vs,
                        const double *ws, const double *dp, const double *gauge, double *u, double *v, double *w,
                        double *p, void *out);
// a whole velocity (pres false: a0..a2 = u, v, w) or pressure (a0 = dp) ghost
// fill, axes first_axis..2, in one launch; returns the block count (grid.y = 3
// components or 1)
int make_fill_params(const Grid &g, bool pres, const int bc[3], const double *planes, const double d[3], double *a0,
                     double *a1, double *a2, int first_axis, void *out);
void launch_whole_fill_batch(const void *params, const int *prefix, int nbox, int nblk, int ncomp, hipStream_t st);
void launch_project_batch(const void *params, const int *prefix, int nbox, int nblk, hipStream_t st);
// AB2-CN boundary rows (vel_bc_staggered_CN2) along axis ax; 0 blocks: nothing to launch
int make_cn2_params(const Grid &g, int bc, int ax, const double *planes, double *u, double *v, double *w, void *out);
void launch_cn2_batch(const void *params, const int *prefix, int nbox, int nblk, hipStream_t st);
// ADI sweep along axis dir-1 (ns_adi.hip): per-box parameters, block count (grid.y = 3), LDS size
size_t sweep_params_size();
// x_lsh: at most 2^x_lsh lines per x-sweep block (6: 64)
// part: use the partitioned wall solve where both systems have one; planes
// (partitioned solve only): the lines' first and last rows are read from the
// plane set (the rows vel_bc_staggered_CN2 sets) instead of the arrays;
// planes_out (partitioned solve only): the solution's first and last rows are
// added to that plane set and zeros stored in their place (the transpose of
// those boundary rows, after an adjoint solve)
int make_sweep_params(const Grid &g, int dir, const LineSys &sysA, const LineSys &sysB, double *u, double *v, double *w,
                      void *out, size_t *lds, int x_lsh = 6, bool part = true, const double *planes = nullptr,
                      const PlaneLayout *L = nullptr, double *planes_out = nullptr);
bool sweep_partitioned(const LineSys &sysA, const LineSys &sysB);
// waves per block of a sweep's parameters (0: twisted / cyclic path)
int sweep_params_waves(const void *params);
// rows per lane of a partitioned sweep's kernel (PART_ROWS or PART_ROWS_LONG)
int sweep_params_rows(const void *params);
// kw: waves per block of the partitioned solve (every box partitioned), 0: twisted / cyclic path
// dense: every box of the launch has the dense layout (the kernels for it)
void launch_adi_batch(int ax, int kw, int rows, bool dense, const void *params, const int *prefix, int nbox, int nblk,
                      size_t lds, hipStream_t st);
struct PoissonPlans;
size_t pass_params_size();
// lsh_drop: lines per block reduced by 2^lsh_drop from the default (a launch-shape choice only)
int make_dct_pass_params(const Grid &g, const PoissonPlans &pl, double dx, double dy, double dz, double dt,
                         const double *us, const double *vs, const double *ws, const double *rhs_in,
                         const double *planes, const double *mean_dp_dev, double mean_dp_host, double *work,
                         double *dp_out, double *sol_out, int pass, void *out, size_t *lds, int lsh_drop = 0,
                         double *p_acc = nullptr);
// kernel variant of a box's pass (bit 1: odd radices, bit 0: lines over 1024);
// a batch launch uses the bitwise OR over its boxes
int dct_params_variant(const void *params);
void launch_dct_batch(int pass, int variant, const void *params, const int *prefix, int nbox, int nb, size_t lds,
                      hipStream_t st);
void launch_conv_diff(const Grid &g, const int bc[3], double h0, double h1, double h2, int which_diff,
                      const double *u, const double *v, const double *w, double *ox, double *oy, double *oz,
                      hipStream_t st);
// fills axes first_axis..2 (the earlier axes already filled by the producer)
void launch_vel_bc(const Grid &g, const int bc[3], const double *planes, double *u, double *v, double *w, hipStream_t st,
                   int first_axis = 0);
void launch_vel_bc_cn2(const Grid &g, int bc, int dir, const double *planes, double *u, double *v, double *w,
                       hipStream_t st);
// p_acc (non-periodic codes only): also p_acc += dp at every filled entry
void launch_pr_bc(const Grid &g, const int pbc[3], const double *planes, double dx, double dy, double dz, double *dp,
                  hipStream_t st, int first_axis = 0, double *p_acc = nullptr);
void launch_modify_rhs(const Grid &g, const int pbc[3], const double *planes, double dx, double dy, double dz,
                       double *rhs, hipStream_t st);
void launch_divergence(const Grid &g, double dx, double dy, double dz, double dt, const double *us, const double *vs,
                       const double *ws, double *rhs, hipStream_t st);
// gauge: if non-null, dp values are taken as dp - *gauge (the corner gauge)
void launch_project(const Grid &g, double dx, double dy, double dz, double dt, int dp_flag, const double *us,
                    const double *vs, const double *ws, const double *dp, const double *gauge, double *u, double *v,
                    double *w, double *p, hipStream_t st);
void launch_get_velpr_bc(const Grid &g, const int bc[3], const int pbc[3], double dx, double dy, double dz, double dt,
                         int corr, const double *ubc, const double *vbc, const double *wbc, const double *pbc_arr,
                         const double *pprev, double *planes, hipStream_t st);
void launch_get_pr_bc(const Grid &g, const int pbc[3], double dx, double dy, double dz, const double *p,
                      double *planes, hipStream_t st);
void launch_axpby(const double *a, const double *b, double th, double *out, long n, hipStream_t st);
// out[q] = sum_{i<m} w[i] * a[i*stride + q]
#define NS_WSUM_MAX 8
void launch_wsum(const double *a, long stride, const double *w, int m, double *out, long n, hipStream_t st);
void launch_copy_interior(const Grid &g, const double *sol, double *dp, hipStream_t st);
void launch_sub_gauge(const double *src, const double *gauge, double *dst, size_t n, hipStream_t st);
// rows of n0 entries from src (row pitch ps) to dst (row pitch pd), minus *gauge
// if gauge is set; rows * n0 < 2^32
void launch_relayout(const double *src, int ps, double *dst, int pd, int n0, size_t rows, const double *gauge,
                     hipStream_t st);
// box of extent e at 0-based origin o of a ghosted array with leading dims n[0], n[1]
void launch_extract_box(const double *src, const int n[3], const int o[3], const int e[3], const double *gauge,
                        double *dst, hipStream_t st);
// block-wise partial maxima of |div| into partial[], returns number of partials
int launch_max_div(const Grid &g, double dx, double dy, double dz, const double *u, const double *v, const double *w,
                   double *partial, hipStream_t st);
int max_div_partials(const Grid &g);
void launch_count_nonfinite(const double *a, size_t n, unsigned int *count, hipStream_t st);
// sub-box plane capture from a parent field (what 0: velocity planes of wall code 2 from
// s0..s2 = u*,v*,w*; what 1: pressure-gradient planes from s0 = dp); o = 0-based origin
void launch_capture(const Grid &gp, const Grid &gs, const int o[3], int ax, int what, const double d[3],
                    const double *gauge, const double *s0, const double *s1, const double *s2, double *planes,
                    hipStream_t st);
void launch_subbox_mean(const Grid &gp, const Grid &gs, const int o[3], const double *gauge, const double *dp,
                        double *partial, double *out, hipStream_t st);

// ---- Poisson (ns_poisson.hip) ----
struct PoissonPlans {
    FFTPlan px, py, pz;  // transform lengths nx, ny, nz
    int dct;             // 0: periodic FFT; 1: DCT/DST family
    int pbc[3];
    const double *ex, *ey, *ez;  // per-axis eigenvalue parts (FFT: nx/2+1 / ny / nz; DCT: nx/ny/nz)
    const double2 *wq_x, *wq_y, *wq_z;  // Makhoul quarter-wave twiddles exp(-i pi k/(2n)), DCT only
};
// row stride (complex) of the periodic half spectrum [z][y][kx]
int spectrum_row(int nx);
// ev (optional): 5 events recorded after the five passes
// wrap_vel: every velocity axis periodic, so the divergence reads the high
// own-axis faces through their periodic copies (the ghosts need not be filled)
void poisson_fft_solve(const Grid &g, const PoissonPlans &pl, double dx, double dy, double dz, double dt,
                       const double *us, const double *vs, const double *ws, const double *rhs_in, double2 *spec,
                       double *dp_out, double *sol_out, hipStream_t st, hipEvent_t *ev = nullptr,
                       bool dp_ghost_x = false, bool wrap_vel = false);
void poisson_dct_solve(const Grid &g, const PoissonPlans &pl, double dx, double dy, double dz, double dt,
                       const double *us, const double *vs, const double *ws, const double *rhs_in,
                       const double *planes, const double *mean_dp_dev, double mean_dp_host, double *work,
                       double *dp_out, double *sol_out, hipStream_t st, hipEvent_t *ev = nullptr,
                       double *p_acc = nullptr);

// ---- re-simulation error metrics (ns_metrics.hip) ----
// partials per block: for u, v, w, p: max|a-b|, sum|a-b|, sum b^2; then max|oz-oz_b|, sum oz_b^2
#define NS_METRIC_PARTS 14
// a (the run's u, v, w, p) in the layout of g, b (the parent's) dense
int launch_err_partials(const Grid &g, double dx, double dy, const double *const a[4], const double *const b[4],
                        double *partial, int max_blocks, hipStream_t st);
// rows of block partials held by the static scratch of ns_error_metrics_noalloc
#define NOALLOC_METRIC_BLOCKS 2048
// out[0..NS_METRIC_PARTS) := the nb partial rows combined (max or sum per entry), one block
void launch_err_combine(const double *partial, int nb, double *out, hipStream_t st);
// device address of that scratch on the current device (looked up once per device)
int metric_scratch(double **out);
// dst[i] = src[i], i < n (the traffic-counter calibration kernel)
void launch_stream_copy(const double *src, double *dst, long n, hipStream_t st);

// ---- discrete adjoint of a step (ns_adjoint.hip) ----
void launch_adj_project(const Grid &g, double dt, const double h[3], int dp_flag, const double *const ub[3],
                        double *const usb[3], double *pb, double *dpb, hipStream_t st);
// ghost-fill transpose, axes z, y, x: comp 3 = the pressure array a[0] (codes pbc),
// comp 0 = the three velocity arrays a[0..2] (codes bc), one launch per axis
void launch_adj_fill(const Grid &g, int comp, const int bc[3], const double h[3], double *const a[3], double *pbar,
                     hipStream_t st);
void launch_adj_take_interior(const Grid &g, const double *dpb, double *r, hipStream_t st);
void launch_adj_rhs_div(const Grid &g, const int pbc[3], const double h[3], double dt, const double *rb,
                        double *const usb[3], double *pbar, hipStream_t st);
// its two parts: boundary RHS terms (plane adjoints added to pbar) and divergence (added to usb)
void launch_adj_modify_rhs(const Grid &g, const int pbc[3], const double h[3], const double *rb, double *pbar,
                           hipStream_t st);
void launch_adj_div(const Grid &g, const double h[3], double dt, const double *rb, double *const usb[3],
                    hipStream_t st);
// gb / gcb: work arrays for the adjoints of the diffusive and convective parts
// of the right-hand side (the same arrays for the explicit schemes)
void launch_adj_predict(const Grid &g, const int per[3], double dt, double nu, int dp_flag, const double h[3],
                        int scheme, int first, const double *const u[3], const double *const usb[3],
                        double *const hb[3], double *const gb[3], double *const gcb[3], double *const ub[3],
                        double *pb, hipStream_t st);
// transpose of the AB2-CN boundary rows before the sweep along axis ax
// (vel_bc_staggered_CN2, walls): the line ends' adjoints move to the planes
void launch_adj_cn2(const Grid &g, int ax, int bc, double *const a[3], double *pbar, hipStream_t st);
int launch_block_sums(const double *a, size_t n, double *partial, int max_blocks, hipStream_t st);
void launch_add_at(double *a, size_t q, double v, hipStream_t st);
// a[q] += scale * sum(partial[0..nb)), on the device
void launch_add_partials(double *a, size_t q, const double *partial, int nb, double scale, hipStream_t st);

// ---- ADI (ns_adi.hip) ----
// planes, planes_out: as for make_sweep_params (used only when the sweep is partitioned)
// c0, nc: the components swept (blockIdx.y = component - c0)
void launch_adi_sweep(const Grid &g, int dir, const LineSys &sysA, const LineSys &sysB, double *u, double *v, double *w,
                      hipStream_t st, const double *planes = nullptr, const PlaneLayout *L = nullptr,
                      double *planes_out = nullptr, int c0 = 0, int nc = 3);

}  // namespace ns
