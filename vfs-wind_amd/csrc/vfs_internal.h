// vfs_internal.h -- shared declarations of the HIP implementation.
//
// Layout: one DMDA block per context, global arrays of mz*my*mx points with i
// fastest; vector fields are AoS {x,y,z} doubles exactly as the host's
// PETSc arrays, so uploads and downloads are straight copies.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>
#include <utility>
#include <vector>

#include "vfs_wind.h"

namespace vfs {

// Grid geometry and the flags the kernels branch on, passed by value.
struct Grid {
    int mx, my, mz;
    int pi, pj, pk;  // DA-periodic directions
    int bc[6];
    long long plane;  // mx*my
    long long n;      // mx*my*mz
};

__host__ __device__ inline int wrapi(int i, int m, int per)
{
    if (per) {
        if (i < 0) i += m;
        else if (i >= m) i -= m;
    }
    return i;
}

// Index of (k,j,i) with the one-rank periodic DA ghost wrap.
__host__ __device__ inline long long gidx(const Grid &g, int k, int j, int i)
{
    i = wrapi(i, g.mx, g.pi);
    j = wrapi(j, g.my, g.pj);
    k = wrapi(k, g.mz, g.pk);
    return ((long long)k * g.my + j) * g.mx + i;
}

// Source of the reference's periodic ghost fix for cell (k,j,i): cells 0 and
// m-1 of a periodic direction copy m-2 and 1 (e.g. momentum.c:638-666).
// Returns false when (k,j,i) is not such a ghost cell.
__host__ __device__ inline bool periodic_src(const Grid &g, int k, int j, int i, int &c, int &b, int &a,
                                             int &fi, int &fj, int &fk)
{
    a = i; b = j; c = k;
    fi = fj = fk = 0;
    if (g.pi && i == 0) { a = g.mx - 2; fi = 1; }
    else if (g.pi && i == g.mx - 1) { a = 1; fi = 1; }
    if (g.pj && j == 0) { b = g.my - 2; fj = 1; }
    else if (g.pj && j == g.my - 1) { b = 1; fj = 1; }
    if (g.pk && k == 0) { c = g.mz - 2; fk = 1; }
    else if (g.pk && k == g.mz - 1) { c = 1; fk = 1; }
    return fi || fj || fk;
}

// the cell metrics of point q into v[0..9]: csi xyz, eta xyz, zet xyz, aj
__device__ inline void load_cell_metrics(const double *__restrict__ csi, const double *__restrict__ eta,
                                         const double *__restrict__ zet, const double *__restrict__ aj, long long q,
                                         double v[10])
{
    for (int c = 0; c < 3; c++) {
        v[c] = csi[3 * q + c];
        v[3 + c] = eta[3 * q + c];
        v[6 + c] = zet[3 * q + c];
    }
    v[9] = aj[q];
}

// a face metric as FormMetrics builds it from the two cells the face
// separates: 0.5 X + 0.5 Y (AxByC, momentum.c:40-44) and 2 / (1/aj + 1/aj)
// (metrics.c:588-593, 696-701, 804-809)
__device__ inline void face_mean(const double a[10], const double b[10], double f[10])
{
#pragma clang fp contract(off)
    for (int c = 0; c < 9; c++) f[c] = 0.5 * a[c] + 0.5 * b[c];
    f[9] = 2. / (1. / a[9] + 1. / b[9]);
}

// inverse of [csi; eta; zet] (Calculate_Covariant_metrics, rhs2.c:595-612)
__host__ __device__ inline void covariant(const double gm[3][3], double G[3][3])
{
#pragma clang fp contract(off)
    const double a11 = gm[0][0], a12 = gm[0][1], a13 = gm[0][2];
    const double a21 = gm[1][0], a22 = gm[1][1], a23 = gm[1][2];
    const double a31 = gm[2][0], a32 = gm[2][1], a33 = gm[2][2];
    double det = a11 * (a33 * a22 - a32 * a23) - a21 * (a33 * a12 - a32 * a13) + a31 * (a23 * a12 - a22 * a13);
    G[0][0] = (a33 * a22 - a32 * a23) / det; G[0][1] = -(a33 * a12 - a32 * a13) / det; G[0][2] = (a23 * a12 - a22 * a13) / det;
    G[1][0] = -(a33 * a21 - a31 * a23) / det; G[1][1] = (a33 * a11 - a31 * a13) / det; G[1][2] = -(a23 * a11 - a21 * a13) / det;
    G[2][0] = (a32 * a21 - a31 * a22) / det; G[2][1] = -(a32 * a11 - a31 * a12) / det; G[2][2] = (a22 * a11 - a21 * a12) / det;
}

Grid make_grid(const vfs_params &p);

// Status plumbing shared by the C ABI.
int set_error(int code, const std::string &msg);
int check_hip(hipError_t e, const char *what);

#define VFS_HIP(call)                                          \
    do {                                                       \
        hipError_t e_ = (call);                                \
        if (e_ != hipSuccess) return ::vfs::check_hip(e_, #call); \
    } while (0)

#define VFS_LAUNCH_CHECK(what)                                \
    do {                                                      \
        hipError_t e_ = hipGetLastError();                    \
        if (e_ != hipSuccess) return ::vfs::check_hip(e_, what); \
    } while (0)

inline unsigned blocks_for(long long n, int threads) { return (unsigned)((n + threads - 1) / threads); }

// XCD-aware block order.  The dispatcher hands workgroup b (x fastest, then
// y, then z) to XCD b % 8, so consecutive blocks land on different XCDs and
// their shared halos miss in each XCD's own L2.  xcd_block() maps the
// hardware block id to a logical one such that every XCD works on one
// contiguous range of logical blocks (a bijection on [0, nb)).
constexpr unsigned NXCD = 8;
__device__ inline unsigned xcd_block(unsigned b, unsigned nb)
{
    const unsigned q = nb / NXCD, rem = nb % NXCD, x = b % NXCD, r = b / NXCD;
    return x * q + (x < rem ? x : rem) + r;
}
// the logical block of a 1D launch
__device__ inline unsigned xcd_block1() { return xcd_block(blockIdx.x, gridDim.x); }
// the logical (x, y, z) block of a 3D launch
__device__ inline void xcd_block3(unsigned &bx, unsigned &by, unsigned &bz)
{
    const unsigned nb = gridDim.x * gridDim.y * gridDim.z;
    const unsigned l = xcd_block(blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z), nb);
    bx = l % gridDim.x;
    by = (l / gridDim.x) % gridDim.y;
    bz = l / (gridDim.x * gridDim.y);
}


// k-marching tiles over the interior: one thread per (i, j) of a KM_I x KM_J
// tile, marching kc planes; blocks are numbered i-tile fastest, then j-tile,
// then k-chunk
constexpr int KM_I = 64, KM_J = 4, KM_K = 16;
__host__ __device__ inline unsigned km_blocks(const Grid &g, int kc)
{
    return (unsigned)(((g.mx - 2 + KM_I - 1) / KM_I) * ((g.my - 2 + KM_J - 1) / KM_J) * ((g.mz - 2 + kc - 1) / kc));
}
// this thread's interior (i, j) and planes [k0, k1); false when outside
__device__ inline bool km_tile(const Grid &g, int kc, int &i, int &j, int &k0, int &k1)
{
    const int nbi = (g.mx - 2 + KM_I - 1) / KM_I, nbj = (g.my - 2 + KM_J - 1) / KM_J;
    int b = blockIdx.x;
    const int bi = b % nbi;
    b /= nbi;
    const int bj = b % nbj, bk = b / nbj;
    i = 1 + bi * KM_I + (int)threadIdx.x % KM_I;
    j = 1 + bj * KM_J + (int)threadIdx.x / KM_I;
    k0 = 1 + bk * kc;
    k1 = k0 + kc < g.mz - 1 ? k0 + kc : g.mz - 1;
    return i < g.mx - 1 && j < g.my - 1 && k0 < k1;
}
}  // namespace vfs

// Multigrid solver state (poisson.hip).
struct vfs_poisson;

namespace vfs {
// kernels whose launches can be timed with HIP events on the context stream
enum KernelId {
    K_FACE = 0, K_CELL, K_LINE32, K_LINE64, K_APPLY64, K_C2C, K_PGRAD, K_EDDY, K_LSOLVE32, K_LSOLVE64, K_CELLFP,
    K_FF,       // a whole Formfunction_2 (face pass, periodic fix, both cell passes)
    K_LINEC32,  // the compact FP32 smoother's line right-hand side
    K_LRELAXC32,  // the compact FP32 smoother's one-pass line relaxation (right-hand side and solve)
    K_NUMK
};
}

struct vfs_ctx {
    vfs_params p;
    vfs::Grid g;
    hipStream_t stream;
    double *f[VFS_NFIELDS];
    // scratch for Formfunction_2 and the RK driver
    double *face;     // 18 doubles per point: div1..3, visc1..3 (Cmpnts each)
    double *pucont;   // RK base state
    double *snap;     // 3 doubles per point: snapshot for BC passes
    double *fpc;      // 3 planes of n doubles: Fp of every cell (Formfunction_2)
    double *red;      // reduction scratch (device)
    double *inflow_plane;  // my x mx Cmpnts: the inlet plane of -inlet 100 / -pseudo (user->ucont_plane)
    double *red_host; // pinned host mirror of reduction results
    double mean_k_area;
    double mean_k_flux;
    // true while the face metrics are the cell-metric means vfs_form_metrics
    // derived (no metric array written since): kernels may then rebuild a
    // face metric from the two cell metrics instead of reading it
    bool face_derived;
    // true while VFS_UCAT is what Contra2Cart would produce from the current
    // VFS_UCONT / VFS_UCAT_OLD / metrics / nvert (set by c2c; cleared by
    // every write to those arrays): c2c is idempotent and is then skipped
    bool ucat_fresh;
    // what is still known of VFS_UCAT while it is not fresh: UCAT_INTERIOR =
    // every interior cell holds Contra2Cart's value (only ghost cells were
    // written since), UCAT_WALLS = the same except the cells next to the wall
    // faces k_wall_flux resets; c2c then recomputes only those cells before
    // its ghost-cell passes
    int ucat_part;
    // 1 when every nvert is 0 (no blanked point: the face pass then runs
    // without blanking tests), 0 when some is not, -1 until measured; reset
    // by every write of VFS_NVERT
    int nvert_zero;
    // Newton-GMRES vectors of the implicit momentum solve, kept between calls
    // (allocated on first use, grown when a longer restart asks for more)
    double *imp_ws;
    size_t imp_ws_bytes;
    vfs_poisson *ps;
    // lCs: which value the eddy-viscosity pass reads (CS_PARAM: params.cs2,
    // CS_ZERO: 0, CS_FIELD: the VFS_CS array); scratch of the dynamic model
    int cs_mode;
    double *dyn;
    // time-step index of the step drivers (ti, tistart, rstart_flg)
    int ti, tistart, rstart;
    // per-kernel timing (vfs_profile): event pairs recorded around launches
    int prof_on;
    unsigned prof_mask;                // kernel ids recorded while prof_on (bit id)
    std::vector<int> *prof_stack;      // open regions: record index, or -1 when not recorded
    std::vector<hipEvent_t> *ev_pool;
    std::vector<std::pair<int, int>> *ev_recs;  // (kernel id, index of start event)
};

namespace vfs {
// kernels driven from the ABI layer; all enqueue on ctx->stream
int periodic_fix(vfs_ctx *c, double *f, int nc, int comp_mask_mode);
int launch_metrics(vfs_ctx *c, const Grid &g, const double *coor, double *const m[17]);
inline bool is_metric_field(int f) { return f >= VFS_CSI && f <= VFS_KAJ; }
// the arrays Contra2Cart reads or writes
inline bool is_c2c_field(int f)
{
    return f == VFS_COOR || f == VFS_NVERT || is_metric_field(f) || f == VFS_UCONT || f == VFS_UCAT || f == VFS_UCAT_OLD;
}
int c2c(vfs_ctx *c);
// whether every nvert of the grid is 0 (measured once after each write of
// VFS_NVERT); kernels then run instances without blanking tests, which give
// the same bits.  VFS_NOBLANK=0 answers false (keeps the tests).
int nvert_all_zero(vfs_ctx *c, bool *zero);
int zero_ghosts(vfs_ctx *c, double *f);  // f = 0 on the six boundary planes
enum { UCAT_STALE = 0, UCAT_INTERIOR = 1, UCAT_WALLS = 2 };
// a write that may change any cell of VFS_UCAT's Contra2Cart inputs or of
// VFS_UCAT itself
inline void ucat_stale(vfs_ctx *c)
{
    c->ucat_fresh = false;
    c->ucat_part = UCAT_STALE;
}
// a write of ghost cells of VFS_UCAT only
inline void ucat_ghosts_written(vfs_ctx *c)
{
    if (c->ucat_fresh) c->ucat_part = UCAT_INTERIOR;
    c->ucat_fresh = false;
}
// k_wall_flux reset the wall-face fluxes of VFS_UCONT
inline void ucat_walls_written(vfs_ctx *c)
{
    if (c->ucat_fresh || c->ucat_part == UCAT_INTERIOR) c->ucat_part = UCAT_WALLS;
    c->ucat_fresh = false;
}
enum { CS_PARAM = 0, CS_ZERO = 1, CS_FIELD = 2 };
// Compute_Smagorinsky_Constant_1 (les.hip); materialize_cs writes the
// current lCs into VFS_CS when it is a constant
int smagorinsky_constant(vfs_ctx *c, int ti, int tistart, int rstart);
int materialize_cs(vfs_ctx *c);
// FormBCS; lazy_ucat leaves Ucat to the next Contra2Cart where no boundary
// code reads it (form_bcs_reads_ucat)
int form_bcs(vfs_ctx *c, int outflow_scale, bool lazy_ucat = false);
bool form_bcs_reads_ucat(const vfs_ctx *c);
// RungeKutta; lazy_ucat: the last stage's Contra2Cart and FormBCS's are left
// to the next Contra2Cart (the fractional step's projection recomputes Ucat)
int runge_kutta(vfs_ctx *c, bool lazy_ucat);
int inflow_flux(vfs_ctx *c, double *flux_in);
int pseudo_periodic_plane(vfs_ctx *c);
int eddy_viscosity(vfs_ctx *c);
int pressure_gradient(vfs_ctx *c);
int calc_k_flux(vfs_ctx *c, double *out);
int mean_k_area(vfs_ctx *c, double *out);
// FormFunction_SNES's residual (momentum.c:2237-2336) formed in Formfunction_2's
// cell pass: r = (0 - u/dt + u_o/dt) + scale F(u) + 0.5 RHS_o - dP (+ F_eul),
// u = VFS_UCONT; VFS_RHS is not written.  Only on the fused path
// (formfunction2_fused).
struct CnResidual {
    const double *uo, *rhso, *feul;  // feul may be null
    double rdt;                      // 1 / dt
    double *r;
    // when rbase is set, r receives da rbase + db R(u) instead (the
    // matrix-free Jacobian product's difference quotient, da = -1/h, db = 1/h)
    const double *rbase = nullptr;
    double da = 0, db = 0;
};
int formfunction2(vfs_ctx *c, double scale, const double *pucont, double rk_alpha, double *ucont_out,
                  const CnResidual *cn = nullptr);
// whether Formfunction_2 takes the fused face pass and k_cell_rhs_pr
bool formfunction2_fused(const vfs_ctx *c);
int min_dt(vfs_ctx *c, double *out);
int poisson_rhs(vfs_ctx *c);
int poisson_apply(vfs_ctx *c, const double *x, double *y);
int poisson_solve(vfs_ctx *c, const vfs_solve_opts *o, vfs_solve_stats *st);
int poisson_solve_learned(vfs_ctx *c, vfs_unet *net, const vfs_solve_opts *o, vfs_solve_stats *st);
void poisson_free(vfs_ctx *c);
int update_pressure(vfs_ctx *c);
int projection(vfs_ctx *c);
int divergence(vfs_ctx *c, double *out);
// timing of individual launches (no-ops unless vfs_profile enabled it)
// prof_begin returns the record index (-1 when profiling is off); prof_end
// closes that record, or the most recent one when idx < 0
int prof_begin(vfs_ctx *c, int kernel_id);
void prof_end(vfs_ctx *c, int idx = -1);
}  // namespace vfs
