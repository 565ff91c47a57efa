// vfs_ctx.hip -- context management and the C ABI entry points.
//
// The drivers here follow the reference's call order: RungeKutta
// (rhs.c:1586-1708) and Flow_Solver (solvers.c:313-759).
#include <hip/hip_runtime.h>

#include <chrono>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>

#include "vfs_internal.h"

namespace {
thread_local std::string g_err;

bool is_vector_field(int f)
{
    switch (f) {
    case VFS_COOR: case VFS_CSI: case VFS_ETA: case VFS_ZET:
    case VFS_ICSI: case VFS_IETA: case VFS_IZET:
    case VFS_JCSI: case VFS_JETA: case VFS_JZET:
    case VFS_KCSI: case VFS_KETA: case VFS_KZET:
    case VFS_CENT: case VFS_UCONT: case VFS_UCAT: case VFS_UCAT_OLD:
    case VFS_DP: case VFS_RHS: case VFS_FEUL: case VFS_UCONT_O: case VFS_RHS_O:
        return true;
    default:
        return false;
    }
}

int validate(const vfs_params *p)
{
    if (!p) return vfs::set_error(VFS_EINVAL, "null params");
    if (p->mx < 4 || p->my < 4 || p->mz < 4)
        return vfs::set_error(VFS_EINVAL, "grid too small: each DA dimension needs at least 4 points");
    // point kernels decode (i, j, k) from a point index in 32-bit arithmetic
    if ((long long)p->mx * p->my * p->mz >= (1LL << 31))
        return vfs::set_error(VFS_EINVAL, "grid too large: the DA must hold fewer than 2^31 points");
    if (p->les != 0 && p->les != 1 && p->les != 2)
        return vfs::set_error(VFS_EINVAL, "only -les 0, 1 and 2 are supported");
    if (p->les == 2 && !(p->max_cs >= 0))
        return vfs::set_error(VFS_EINVAL, "max_cs must be >= 0 with -les 2");
    if (p->ren <= 0 || p->dt <= 0)
        return vfs::set_error(VFS_EINVAL, "ren and dt must be positive");
    if (p->per_j)
        return vfs::set_error(VFS_EINVAL, "a periodic j direction is not supported (wall-normal line smoother)");
    // boundary codes of the non-periodic faces (bcs.dat): 1 no-slip, 10 slip
    // (-10 on j+), 6 far field, 0 interface and 3 symmetry (they only set the
    // reference's Ubcs, which nothing on this path reads), 2 moving lid (j+;
    // on k+ it only sets Ubcs), 12 Couette wall (j-, j+), 4 outflow (k+, k-,
    // j+), 5 inlet (k-), 8 characteristic outlet (k+)
    const int per[3] = {p->per_i, p->per_j, p->per_k};
    for (int f = 0; f < 6; f++) {
        if (per[f / 2]) continue;
        const int b = p->bctype[f];
        const bool ok = b == 1 || b == 10 || b == 6 || b == 0 || b == 3 || (b == -10 && f == 3) ||
                        (b == 2 && (f == 3 || f == 5)) || (b == 12 && (f == 2 || f == 3)) ||
                        (b == 4 && (f == 5 || f == 4 || f == 3)) || (b == 5 && f == 4) || (b == 8 && f == 5);
        if (!ok)
            return vfs::set_error(VFS_EINVAL, "bcs code " + std::to_string(b) + " on face " + std::to_string(f) +
                                                  " is not supported (wall functions -1/-2, the body-fitted inflow 11 "
                                                  "and the level-set codes are outside this library)");
    }
    for (int f = 0; f < 6; f++)
        if (p->bctype[f] == 5 && (f != 4 || p->per_k))
            return vfs::set_error(VFS_EINVAL, "an inlet (bcs code 5) is supported on the k- face of a non-periodic k direction");
    if (p->bctype[4] == 5) {
        const int ip = p->inletprofile;
        if (!(ip == 0 || ip == 1 || ip == 2 || ip == 3 || ip == 10 || ip == 12 || ip == 14 || ip == 15 || ip == 21 || ip == 100))
            return vfs::set_error(VFS_EINVAL, "-inlet " + std::to_string(ip) +
                                                  " is not supported (11 adds time-seeded noise; 13, 16, 20 and 101 leave "
                                                  "the inlet velocity unset)");
    }
    if (p->pseudo_periodic && p->bctype[4] != 5)
        return vfs::set_error(VFS_EINVAL, "-pseudo needs the inlet (bcs code 5) on the k- face");
    return VFS_OK;
}
}  // namespace

namespace vfs {

int set_error(int code, const std::string &msg)
{
    g_err = msg;
    return code;
}

int check_hip(hipError_t e, const char *what)
{
    g_err = std::string(what) + ": " + hipGetErrorString(e);
    return e == hipErrorOutOfMemory ? VFS_ENOMEM : VFS_EHIP;
}

// Regions nest; each prof_begin pushes its record index (or -1 when the
// kernel is not selected) and the matching prof_end pops it.
int prof_begin(vfs_ctx *c, int kernel_id)
{
    if (!c->prof_on) return -1;
    if (!((c->prof_mask >> kernel_id) & 1u)) {
        c->prof_stack->push_back(-1);
        return -1;
    }
    size_t at = c->ev_recs->size() * 2;
    while (c->ev_pool->size() < at + 2) {
        hipEvent_t ev;
        if (hipEventCreate(&ev) != hipSuccess) { c->prof_on = 0; return -1; }
        c->ev_pool->push_back(ev);
    }
    hipEventRecord((*c->ev_pool)[at], c->stream);
    c->ev_recs->push_back({kernel_id, (int)at});
    const int idx = (int)c->ev_recs->size() - 1;
    c->prof_stack->push_back(idx);
    return idx;
}

void prof_end(vfs_ctx *c, int idx)
{
    if (!c->prof_on || c->prof_stack->empty()) return;
    const int top = c->prof_stack->back();
    c->prof_stack->pop_back();
    if (idx < 0) idx = top;
    if (idx < 0 || idx >= (int)c->ev_recs->size()) return;
    hipEventRecord((*c->ev_pool)[(*c->ev_recs)[idx].second + 1], c->stream);
}

Grid make_grid(const vfs_params &p)
{
    Grid g;
    g.mx = p.mx; g.my = p.my; g.mz = p.mz;
    g.pi = p.per_i; g.pj = p.per_j; g.pk = p.per_k;
    for (int q = 0; q < 6; q++) g.bc[q] = p.bctype[q];
    g.plane = (long long)p.mx * p.my;
    g.n = g.plane * p.mz;
    return g;
}

}  // namespace vfs

extern "C" {

const char *vfs_last_error(void) { return g_err.c_str(); }

int vfs_ctx_create(const vfs_params *params, vfs_ctx **out)
{
    if (!out) return vfs::set_error(VFS_EINVAL, "null output pointer");
    *out = nullptr;
    int rc = validate(params);
    if (rc) return rc;
    vfs_ctx *c = new vfs_ctx();
    std::memset(c->f, 0, sizeof(c->f));
    c->p = *params;
    c->g = vfs::make_grid(*params);
    c->face = c->pucont = c->snap = c->fpc = c->red = c->red_host = c->inflow_plane = nullptr;
    c->ps = nullptr;
    c->cs_mode = vfs::CS_PARAM;
    c->dyn = nullptr;
    c->ti = 0;
    c->tistart = 0;
    c->rstart = 1;
    c->mean_k_area = 0;
    c->mean_k_flux = 0;
    c->face_derived = false;
    vfs::ucat_stale(c);
    c->nvert_zero = -1;
    c->imp_ws = nullptr;
    c->imp_ws_bytes = 0;
    c->prof_on = 0;
    c->prof_mask = 0;
    c->ev_pool = new std::vector<hipEvent_t>();
    c->ev_recs = new std::vector<std::pair<int, int>>();
    c->prof_stack = new std::vector<int>();
    hipError_t e = hipStreamCreate(&c->stream);
    if (e != hipSuccess) { delete c; return vfs::check_hip(e, "hipStreamCreate"); }
    const size_t n = (size_t)c->g.n;
    auto alloc = [&](double **ptr, size_t count) -> int {
        hipError_t ee = hipMalloc((void **)ptr, count * sizeof(double));
        if (ee != hipSuccess) return vfs::check_hip(ee, "hipMalloc");
        ee = hipMemsetAsync(*ptr, 0, count * sizeof(double), c->stream);
        if (ee != hipSuccess) return vfs::check_hip(ee, "hipMemsetAsync");
        return VFS_OK;
    };
    for (int f = 0; f < VFS_NFIELDS && !rc; f++) rc = alloc(&c->f[f], is_vector_field(f) ? 3 * n : n);
    if (!rc) rc = alloc(&c->face, 18 * n);
    if (!rc) rc = alloc(&c->pucont, 3 * n);
    if (!rc) rc = alloc(&c->snap, 3 * n);
    if (!rc) rc = alloc(&c->fpc, 3 * n);
    if (!rc) rc = alloc(&c->red, 1 << 16);
    if (!rc) rc = alloc(&c->inflow_plane, 3 * (size_t)c->g.plane);
    if (!rc) {
        hipError_t ee = hipHostMalloc((void **)&c->red_host, 4096 * sizeof(double), hipHostMallocDefault);
        if (ee != hipSuccess) rc = vfs::check_hip(ee, "hipHostMalloc");
    }
    if (!rc) {
        hipError_t ee = hipStreamSynchronize(c->stream);
        if (ee != hipSuccess) rc = vfs::check_hip(ee, "hipStreamSynchronize");
    }
    if (rc) { vfs_ctx_destroy(c); return rc; }
    *out = c;
    return VFS_OK;
}

int vfs_ctx_destroy(vfs_ctx *c)
{
    if (!c) return VFS_OK;
    if (c->stream) hipStreamSynchronize(c->stream);
    vfs::poisson_free(c);
    for (int f = 0; f < VFS_NFIELDS; f++) if (c->f[f]) hipFree(c->f[f]);
    if (c->face) hipFree(c->face);
    if (c->pucont) hipFree(c->pucont);
    if (c->snap) hipFree(c->snap);
    if (c->fpc) hipFree(c->fpc);
    if (c->red) hipFree(c->red);
    if (c->inflow_plane) hipFree(c->inflow_plane);
    if (c->red_host) hipHostFree(c->red_host);
    if (c->imp_ws) hipFree(c->imp_ws);
    if (c->dyn) hipFree(c->dyn);
    if (c->ev_pool) {
        for (hipEvent_t ev : *c->ev_pool) hipEventDestroy(ev);
        delete c->ev_pool;
    }
    delete c->ev_recs;
    delete c->prof_stack;
    if (c->stream) hipStreamDestroy(c->stream);
    delete c;
    return VFS_OK;
}

int vfs_set_params(vfs_ctx *c, const vfs_params *params)
{
    if (!c) return vfs::set_error(VFS_EINVAL, "null context");
    int rc = validate(params);
    if (rc) return rc;
    if (params->mx != c->p.mx || params->my != c->p.my || params->mz != c->p.mz ||
        params->per_i != c->p.per_i || params->per_j != c->p.per_j || params->per_k != c->p.per_k)
        return vfs::set_error(VFS_EINVAL, "grid size and periodicity are fixed at creation");
    bool op_changed = params->poisson_threshold != c->p.poisson_threshold;
    for (int q = 0; q < 6; q++) op_changed = op_changed || params->bctype[q] != c->p.bctype[q];
    c->p = *params;
    c->g = vfs::make_grid(*params);
    vfs::ucat_stale(c);  // the boundary codes may have changed
    if (op_changed) vfs::poisson_free(c);  // the operator's rows depend on both
    return VFS_OK;
}

int vfs_get_params(const vfs_ctx *c, vfs_params *params)
{
    if (!c || !params) return vfs::set_error(VFS_EINVAL, "null argument");
    *params = c->p;
    return VFS_OK;
}

size_t vfs_field_size(const vfs_ctx *c, vfs_field f)
{
    if (!c || f < 0 || f >= VFS_NFIELDS) return 0;
    return (is_vector_field(f) ? 3 : 1) * (size_t)c->g.n;
}

// What a write of field f (upload, or through a handed-out device pointer)
// invalidates: the derived-face-metric flag, the fresh-Ucat state, the
// all-fluid flag, the Poisson operator hierarchy (built from the node
// coordinates, the metrics and nvert; PoissonLHSNew is rebuilt whenever the
// blanking can change, poisson_hypre.c:534-550) and the lCs mode.
static void field_written(vfs_ctx *c, int f)
{
    if (vfs::is_metric_field(f)) c->face_derived = false;
    if (vfs::is_c2c_field(f)) vfs::ucat_stale(c);
    if (f == VFS_NVERT) c->nvert_zero = -1;
    if (f == VFS_NVERT || f == VFS_COOR || vfs::is_metric_field(f)) vfs::poisson_free(c);
    if (f == VFS_CS) c->cs_mode = vfs::CS_FIELD;
}

int vfs_field_written(vfs_ctx *c, vfs_field f)
{
    if (!c || f < 0 || f >= VFS_NFIELDS) return vfs::set_error(VFS_EINVAL, "bad field");
    field_written(c, f);
    return VFS_OK;
}

int vfs_upload(vfs_ctx *c, vfs_field f, const double *host)
{
    if (!c || !host || f < 0 || f >= VFS_NFIELDS) return vfs::set_error(VFS_EINVAL, "bad upload argument");
    field_written(c, f);
    VFS_HIP(hipMemcpyAsync(c->f[f], host, vfs_field_size(c, f) * sizeof(double), hipMemcpyHostToDevice, c->stream));
    VFS_HIP(hipStreamSynchronize(c->stream));
    return VFS_OK;
}

int vfs_download(vfs_ctx *c, vfs_field f, double *host)
{
    if (!c || !host || f < 0 || f >= VFS_NFIELDS) return vfs::set_error(VFS_EINVAL, "bad download argument");
    if (f == VFS_CS) {
        int rc = vfs::materialize_cs(c);
        if (rc) return rc;
    }
    VFS_HIP(hipMemcpyAsync(host, c->f[f], vfs_field_size(c, f) * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    VFS_HIP(hipStreamSynchronize(c->stream));
    return VFS_OK;
}

double *vfs_field_device_ptr(vfs_ctx *c, vfs_field f)
{
    if (!c || f < 0 || f >= VFS_NFIELDS) return nullptr;
    if (f == VFS_CS && vfs::materialize_cs(c)) return nullptr;
    field_written(c, f);  // the caller may write through it before the next call
    return c->f[f];
}

int vfs_sync(vfs_ctx *c)
{
    if (!c) return vfs::set_error(VFS_EINVAL, "null context");
    VFS_HIP(hipStreamSynchronize(c->stream));
    return VFS_OK;
}

int vfs_form_metrics(vfs_ctx *c)
{
    if (!c) return vfs::set_error(VFS_EINVAL, "null context");
    double *const m[17] = {c->f[VFS_CSI], c->f[VFS_ETA], c->f[VFS_ZET], c->f[VFS_AJ],
                           c->f[VFS_ICSI], c->f[VFS_IETA], c->f[VFS_IZET], c->f[VFS_IAJ],
                           c->f[VFS_JCSI], c->f[VFS_JETA], c->f[VFS_JZET], c->f[VFS_JAJ],
                           c->f[VFS_KCSI], c->f[VFS_KETA], c->f[VFS_KZET], c->f[VFS_KAJ],
                           c->f[VFS_CENT]};
    int rc = vfs::launch_metrics(c, c->g, c->f[VFS_COOR], m);
    if (rc) return rc;
    c->face_derived = true;
    vfs::ucat_stale(c);
    // the multigrid hierarchy depends on the metrics; rebuild on next solve
    vfs::poisson_free(c);
    VFS_HIP(hipStreamSynchronize(c->stream));
    return VFS_OK;
}

int vfs_mean_k_area(vfs_ctx *c, double *area)
{
    if (!c) return vfs::set_error(VFS_EINVAL, "null context");
    double a = 0;
    int rc = vfs::mean_k_area(c, &a);
    if (rc) return rc;
    c->mean_k_area = a;
    if (area) *area = a;
    return VFS_OK;
}

int vfs_contra2cart(vfs_ctx *c)
{
    if (!c) return vfs::set_error(VFS_EINVAL, "null context");
    return vfs::c2c(c);
}

int vfs_form_bcs(vfs_ctx *c, int outflow_scale)
{
    if (!c) return vfs::set_error(VFS_EINVAL, "null context");
    return vfs::form_bcs(c, outflow_scale);
}

int vfs_smagorinsky_constant(vfs_ctx *c, int ti, int tistart, int rstart_flg)
{
    if (!c) return vfs::set_error(VFS_EINVAL, "null context");
    return vfs::smagorinsky_constant(c, ti, tistart, rstart_flg);
}

int vfs_set_time_index(vfs_ctx *c, int ti, int tistart, int rstart_flg)
{
    if (!c) return vfs::set_error(VFS_EINVAL, "null context");
    c->ti = ti;
    c->tistart = tistart;
    c->rstart = rstart_flg;
    return VFS_OK;
}

int vfs_eddy_viscosity(vfs_ctx *c)
{
    if (!c) return vfs::set_error(VFS_EINVAL, "null context");
    return vfs::eddy_viscosity(c);
}

int vfs_calc_k_flux(vfs_ctx *c, double *mean_k_flux)
{
    if (!c) return vfs::set_error(VFS_EINVAL, "null context");
    double v = 0;
    int rc = vfs::calc_k_flux(c, &v);
    if (rc) return rc;
    c->mean_k_flux = v;
    if (mean_k_flux) *mean_k_flux = v;
    return VFS_OK;
}

int vfs_pressure_gradient(vfs_ctx *c)
{
    if (!c) return vfs::set_error(VFS_EINVAL, "null context");
    if (c->p.per_k && c->mean_k_area == 0) {
        int rc = vfs_mean_k_area(c, nullptr);
        if (rc) return rc;
    }
    return vfs::pressure_gradient(c);
}

int vfs_formfunction2(vfs_ctx *c, double scale)
{
    if (!c) return vfs::set_error(VFS_EINVAL, "null context");
    return vfs::formfunction2(c, scale, nullptr, 0.0, nullptr);
}

int vfs_runge_kutta(vfs_ctx *c)
{
    if (!c) return vfs::set_error(VFS_EINVAL, "null context");
    return vfs::runge_kutta(c, false);
}

}  // extern "C"

int vfs::runge_kutta(vfs_ctx *c, bool lazy_ucat)
{
    static const double alfa[4] = {0.25, 1 / 3., 0.5, 1.};
    // VFS_RK_LAZY=0 computes Ucat at the end of the stages in any case
    const char *lz = std::getenv("VFS_RK_LAZY");
    const bool lazy = lazy_ucat && !(lz && lz[0] == '0') && !vfs::form_bcs_reads_ucat(c);
    int rc = vfs::inflow_flux(c, nullptr);  // rhs.c:1612
    if (!rc) rc = vfs::form_bcs(c, 0);
    if (rc) return rc;
    const size_t n3 = 3 * (size_t)c->g.n;
    VFS_HIP(hipMemcpyAsync(c->pucont, c->f[VFS_UCONT], n3 * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    for (int s = 0; s < 4; s++) {
        // the reference calls Contra2Cart at the start of every stage
        // (rhs.c:1642); from the second stage on Ucont is unchanged since the
        // previous stage's closing call and Contra2Cart is idempotent, so
        // that call would rewrite the same Ucat
        if (s == 0) rc = vfs::c2c(c);
        if (rc) return rc;
        // Rhs = 0; Formfunction_2(Rhs, 1); Rhs -= dP; Ucont = a*Rhs + pUcont  (rhs.c:1647-1658)
        rc = vfs::formfunction2(c, 1.0, c->pucont, alfa[s] * c->p.dt * c->p.st, c->f[VFS_UCONT]);
        if (rc) return rc;
        if (s < 3 || !lazy) rc = vfs::c2c(c);
        if (rc) return rc;
    }
    return vfs::form_bcs(c, 1, lazy);
}

extern "C" {

int vfs_inflow_set_plane(vfs_ctx *c, const double *host)
{
    if (!c || !host) return vfs::set_error(VFS_EINVAL, "null argument");
    VFS_HIP(hipMemcpyAsync(c->inflow_plane, host, 3 * (size_t)c->g.plane * sizeof(double), hipMemcpyHostToDevice, c->stream));
    VFS_HIP(hipStreamSynchronize(c->stream));
    return VFS_OK;
}

int vfs_inflow_get_plane(vfs_ctx *c, double *host)
{
    if (!c || !host) return vfs::set_error(VFS_EINVAL, "null argument");
    VFS_HIP(hipMemcpyAsync(host, c->inflow_plane, 3 * (size_t)c->g.plane * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    VFS_HIP(hipStreamSynchronize(c->stream));
    return VFS_OK;
}

int vfs_pseudo_periodic_plane(vfs_ctx *c)
{
    if (!c) return vfs::set_error(VFS_EINVAL, "null context");
    int rc = vfs::c2c(c);  // Ucat must be current
    return rc ? rc : vfs::pseudo_periodic_plane(c);
}

int vfs_inflow_flux(vfs_ctx *c, double *flux_in)
{
    if (!c) return vfs::set_error(VFS_EINVAL, "null context");
    return vfs::inflow_flux(c, flux_in);
}

int vfs_min_dt(vfs_ctx *c, double *dt_cfl1)
{
    if (!c || !dt_cfl1) return vfs::set_error(VFS_EINVAL, "null argument");
    return vfs::min_dt(c, dt_cfl1);
}

int vfs_poisson_rhs(vfs_ctx *c)
{
    if (!c) return vfs::set_error(VFS_EINVAL, "null context");
    return vfs::poisson_rhs(c);
}

int vfs_poisson_apply(vfs_ctx *c, const double *x, double *y)
{
    if (!c || !x || !y) return vfs::set_error(VFS_EINVAL, "null argument");
    return vfs::poisson_apply(c, x, y);
}

int vfs_poisson_solve(vfs_ctx *c, const vfs_solve_opts *o, vfs_solve_stats *st)
{
    if (!c || !o) return vfs::set_error(VFS_EINVAL, "null argument");
    return vfs::poisson_solve(c, o, st);
}

int vfs_poisson_solve_learned(vfs_ctx *c, vfs_unet *net, const vfs_solve_opts *o, vfs_solve_stats *st)
{
    if (!c || !net || !o) return vfs::set_error(VFS_EINVAL, "null argument");
    return vfs::poisson_solve_learned(c, net, o, st);
}

int vfs_update_pressure(vfs_ctx *c)
{
    if (!c) return vfs::set_error(VFS_EINVAL, "null context");
    return vfs::update_pressure(c);
}

int vfs_projection(vfs_ctx *c)
{
    if (!c) return vfs::set_error(VFS_EINVAL, "null context");
    return vfs::projection(c);
}

int vfs_divergence(vfs_ctx *c, double *maxdiv)
{
    if (!c || !maxdiv) return vfs::set_error(VFS_EINVAL, "null argument");
    return vfs::divergence(c, maxdiv);
}

static const char *k_names[vfs::K_NUMK] = {"face_fluxes",     "cell_rhs",          "line_rhs_fp32",  "line_rhs_fp64",
                                           "poisson_apply_fp64", "contra2cart",     "pressure_gradient", "eddy_viscosity",
                                           "line_solve_fp32",  "line_solve_fp64",   "cell_fp",           "formfunction2",
                                           "line_rhs_compact_fp32", "line_relax_compact_fp32"};

int vfs_profile(vfs_ctx *c, int enable)
{
    return vfs_profile_select(c, enable ? ~0u : 0u);
}

int vfs_profile_select(vfs_ctx *c, unsigned int kernel_mask)
{
    if (!c) return vfs::set_error(VFS_EINVAL, "null context");
    VFS_HIP(hipStreamSynchronize(c->stream));
    c->ev_recs->clear();
    c->prof_stack->clear();
    c->prof_mask = kernel_mask;
    c->prof_on = kernel_mask ? 1 : 0;
    return VFS_OK;
}

int vfs_profile_read(vfs_ctx *c, int kernel, double *total_ms, long long *count, const char **name)
{
    if (!c || kernel < 0 || kernel >= vfs::K_NUMK) return vfs::set_error(VFS_EINVAL, "bad profile query");
    VFS_HIP(hipStreamSynchronize(c->stream));
    double tot = 0;
    long long cnt = 0;
    for (auto &r : *c->ev_recs) {
        if (r.first != kernel) continue;
        float ms = 0;
        VFS_HIP(hipEventElapsedTime(&ms, (*c->ev_pool)[r.second], (*c->ev_pool)[r.second + 1]));
        tot += ms;
        cnt++;
    }
    if (total_ms) *total_ms = tot;
    if (count) *count = cnt;
    if (name) *name = k_names[kernel];
    return VFS_OK;
}

static int step_common(vfs_ctx *c, const vfs_solve_opts *o, const vfs_newton_opts *no, vfs_step_stats *st)
{
    using clk = std::chrono::steady_clock;
    auto t0 = clk::now();
    int rc = c->p.pseudo_periodic ? vfs_pseudo_periodic_plane(c) : VFS_OK;  // solvers.c:355-358
    if (!rc) rc = vfs_calc_k_flux(c, nullptr);              // solvers.c:363
    if (!rc && c->p.les) rc = vfs::c2c(c);                  // solvers.c:368
    const int freq = c->p.dynamic_freq < 1 ? 1 : c->p.dynamic_freq;  // main.c:1439
    if (!rc && c->p.les && (c->ti % freq == 0 || c->ti == c->tistart))
        rc = vfs::smagorinsky_constant(c, c->ti, c->tistart, c->rstart);  // solvers.c:369
    if (!rc && c->p.les) rc = vfs::eddy_viscosity(c);       // solvers.c:370
    if (!rc) rc = vfs_pressure_gradient(c);                 // solvers.c:438
    vfs_newton_stats ns;
    std::memset(&ns, 0, sizeof(ns));
    if (!rc) {
        if (no) {                                           // solvers.c:626-630, -imp 4
            rc = vfs_store_old_state(c);
            if (!rc) rc = vfs_implicit_matrixfree(c, no, &ns);
        } else {
            // Ucat after the stages is recomputed by the projection's
            // Contra2Cart before anything reads it
            rc = vfs::runge_kutta(c, true);                 // solvers.c:635
        }
    }
    if (rc) return rc;
    VFS_HIP(hipStreamSynchronize(c->stream));
    auto t1 = clk::now();
    vfs_solve_stats ps;
    std::memset(&ps, 0, sizeof(ps));
    rc = vfs::poisson_rhs(c);                               // solvers.c:650-655
    if (!rc) rc = vfs::poisson_solve(c, o, &ps);
    if (rc && rc != VFS_ENOCONV) return rc;
    int conv = rc;
    auto t2 = clk::now();
    rc = vfs::update_pressure(c);                           // solvers.c:662
    if (!rc) rc = vfs::projection(c);                       // solvers.c:663
    double md = 0;
    if (!rc) rc = vfs::divergence(c, &md);                  // solvers.c:684
    if (rc) return rc;
    auto t3 = clk::now();
    c->ti++;
    if (st) {
        st->poisson_iterations = ps.iterations;
        st->poisson_relres = ps.relres;
        st->maxdiv = md;
        st->ms_momentum = std::chrono::duration<double, std::milli>(t1 - t0).count();
        st->ms_poisson = std::chrono::duration<double, std::milli>(t2 - t1).count();
        st->ms_total = std::chrono::duration<double, std::milli>(t3 - t0).count();
        st->newton_iterations = ns.newton_iterations;
        st->linear_iterations = ns.linear_iterations;
    }
    return conv;
}

int vfs_fractional_step(vfs_ctx *c, const vfs_solve_opts *o, vfs_step_stats *st)
{
    if (!c || !o) return vfs::set_error(VFS_EINVAL, "null argument");
    return step_common(c, o, nullptr, st);
}

int vfs_fractional_step_cn(vfs_ctx *c, const vfs_newton_opts *no, const vfs_solve_opts *o, vfs_step_stats *st)
{
    if (!c || !o || !no) return vfs::set_error(VFS_EINVAL, "null argument");
    return step_common(c, o, no, st);
}

}  // extern "C"
