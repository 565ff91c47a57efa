// amg.hip -- algebraic multigrid preconditioner for the pressure Poisson
// system: the role of HYPRE BoomerAMG in PoissonSolver_Hypre
// (poisson_hypre.c:34-97: PMIS coarsening, FF1 interpolation, strength
// threshold -amg_thresh, one V-cycle per preconditioner application).
//
// Setup (host, once per operator, like the reference's PoissonLHSNew +
// HYPRE setup, poisson_hypre.c:534-565):
//   strength  S_i = {j != i : -s a_ij >= theta max_k(-s a_ik)}, s = sign(a_ii)
//   PMIS      measure |S^T_i| + a hashed fraction in (0, 1) (a fixed random
//             vector, so setups are reproducible); repeated independent sets
//   FF1       extended+i interpolation restricted as HYPRE's type 13: the
//             C-points of i's strong C neighbours, plus for a strong F
//             neighbour that shares no C-point with i its first strong C
//             neighbour; at most P_MAX entries per row, rescaled so the row
//             sum is kept
//   Galerkin  A_c = P^T A P
// Cycle (device, FP64): V(1,1) with multicolour Gauss-Seidel -- HYPRE's
// default smoother is hybrid Gauss-Seidel, sequential Gauss-Seidel inside
// one rank; here each level's rows are greedily coloured at setup so rows of
// one colour share no matrix entry, and one colour is relaxed per launch
// (forward colour order before the coarse correction, backward after).  CSR
// products run one row per thread; the coarsest level is a dense inverse of
// A_c + s 1 1^T (A's null space is the constant vector).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "vfs_internal.h"

namespace vfs {
namespace amg {

struct Csr {
    int n = 0, m = 0;  // rows, columns
    std::vector<int> ptr, col;
    std::vector<double> val;
    long long nnz() const { return (long long)col.size(); }
};

constexpr int P_MAX = 4;         // entries per interpolation row
constexpr int MAX_COARSE = 512;  // rows of the coarsest (dense) level

static double hash_unit(uint32_t i)
{
    // splitmix-style integer hash -> (0, 1)
    uint64_t z = (uint64_t)i + 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    z ^= z >> 31;
    return ((double)(z >> 11) + 0.5) / 9007199254740992.0;
}

// strength of connection (classical, sign-aware)
static void strength(const Csr &A, double theta, std::vector<int> &sp, std::vector<int> &sc)
{
    sp.assign(A.n + 1, 0);
    sc.clear();
    for (int i = 0; i < A.n; i++) {
        double diag = 0, mx = 0;
        for (int q = A.ptr[i]; q < A.ptr[i + 1]; q++)
            if (A.col[q] == i) diag = A.val[q];
        const double s = diag < 0 ? -1.0 : 1.0;
        for (int q = A.ptr[i]; q < A.ptr[i + 1]; q++)
            if (A.col[q] != i) mx = std::max(mx, -s * A.val[q]);
        if (mx > 0)
            for (int q = A.ptr[i]; q < A.ptr[i + 1]; q++)
                if (A.col[q] != i && -s * A.val[q] >= theta * mx) sc.push_back(A.col[q]);
        sp[i + 1] = (int)sc.size();
    }
}

static void transpose_pattern(int n, const std::vector<int> &sp, const std::vector<int> &sc, std::vector<int> &tp,
                              std::vector<int> &tc)
{
    tp.assign(n + 1, 0);
    for (int c : sc) tp[c + 1]++;
    for (int i = 0; i < n; i++) tp[i + 1] += tp[i];
    tc.assign(sc.size(), 0);
    std::vector<int> fill(tp.begin(), tp.end() - 1);
    for (int i = 0; i < n; i++)
        for (int q = sp[i]; q < sp[i + 1]; q++) tc[fill[sc[q]]++] = i;
}

// PMIS: cf[i] = 1 (C) or -1 (F)
static void pmis(int n, const std::vector<int> &sp, const std::vector<int> &sc, std::vector<int> &cf)
{
    std::vector<int> tp, tc;
    transpose_pattern(n, sp, sc, tp, tc);
    std::vector<double> w(n);
    cf.assign(n, 0);
    for (int i = 0; i < n; i++) {
        w[i] = (tp[i + 1] - tp[i]) + hash_unit((uint32_t)i);
        if (tp[i + 1] == tp[i]) cf[i] = -1;  // influences nobody: F
    }
    std::vector<int> newc;
    for (bool left = true; left;) {
        left = false;
        newc.clear();
        for (int i = 0; i < n; i++) {
            if (cf[i] != 0) continue;
            bool best = true;
            for (int q = sp[i]; q < sp[i + 1] && best; q++) {
                const int j = sc[q];
                if (cf[j] == 0 && w[j] > w[i]) best = false;
            }
            for (int q = tp[i]; q < tp[i + 1] && best; q++) {
                const int j = tc[q];
                if (cf[j] == 0 && w[j] > w[i]) best = false;
            }
            if (best) newc.push_back(i);
        }
        for (int i : newc) cf[i] = 1;
        // points that strongly depend on a new C-point become F
        for (int i : newc)
            for (int q = tp[i]; q < tp[i + 1]; q++)
                if (cf[tc[q]] == 0) cf[tc[q]] = -1;
        for (int i = 0; i < n; i++)
            if (cf[i] == 0) { left = true; break; }
    }
}

// FF1 (restricted extended+i) interpolation with truncation
static Csr interpolation(const Csr &A, const std::vector<int> &sp, const std::vector<int> &sc, const std::vector<int> &cf,
                         std::vector<int> &cidx)
{
    const int n = A.n;
    cidx.assign(n, -1);
    int nc = 0;
    for (int i = 0; i < n; i++)
        if (cf[i] == 1) cidx[i] = nc++;
    Csr P;
    P.n = n;
    P.m = nc;
    P.ptr.assign(n + 1, 0);
    std::vector<char> strong(n, 0);
    std::vector<int> mark(n, -1);  // position in the interpolatory set + 1, per row i
    std::vector<double> wsum(n, 0.0);
    std::vector<int> hat;
    for (int i = 0; i < n; i++) {
        if (cf[i] == 1) {
            P.col.push_back(cidx[i]);
            P.val.push_back(1.0);
            P.ptr[i + 1] = (int)P.col.size();
            continue;
        }
        // interpolatory set: strong C neighbours, then FF1 additions
        hat.clear();
        for (int q = sp[i]; q < sp[i + 1]; q++) strong[sc[q]] = 1;
        for (int q = sp[i]; q < sp[i + 1]; q++) {
            const int j = sc[q];
            if (cf[j] == 1 && mark[j] < 0) { mark[j] = (int)hat.size(); hat.push_back(j); }
        }
        const size_t nci = hat.size();
        for (int q = sp[i]; q < sp[i + 1]; q++) {
            const int k = sc[q];
            if (cf[k] != -1) continue;
            bool common = false;
            int first = -1;
            for (int r = sp[k]; r < sp[k + 1]; r++) {
                const int l = sc[r];
                if (cf[l] != 1) continue;
                if (first < 0) first = l;
                if (mark[l] >= 0 && (size_t)mark[l] < nci) common = true;
            }
            if (!common && first >= 0 && mark[first] < 0) { mark[first] = (int)hat.size(); hat.push_back(first); }
        }
        // weights: w_ij = -(a_ij + sum_k a_ik ahat_kj / sum_{l in hat u {i}} ahat_kl) / atilde_ii
        double aii = 0;
        for (int q = A.ptr[i]; q < A.ptr[i + 1]; q++)
            if (A.col[q] == i) aii = A.val[q];
        double dtil = aii;
        for (size_t h = 0; h < hat.size(); h++) wsum[h] = 0;
        for (int q = A.ptr[i]; q < A.ptr[i + 1]; q++) {
            const int j = A.col[q];
            if (j == i) continue;
            const double aij = A.val[q];
            if (mark[j] >= 0) { wsum[mark[j]] += aij; continue; }
            if (strong[j] && cf[j] == -1) {
                // distribute a_ij over hat u {i} through row j's opposite-sign entries
                double akk = 0;
                for (int r = A.ptr[j]; r < A.ptr[j + 1]; r++)
                    if (A.col[r] == j) akk = A.val[r];
                double den = 0;
                for (int r = A.ptr[j]; r < A.ptr[j + 1]; r++) {
                    const int l = A.col[r];
                    const double a = A.val[r];
                    if (l != j && (l == i || mark[l] >= 0) && (a * akk < 0)) den += a;
                }
                if (den == 0) { dtil += aij; continue; }
                for (int r = A.ptr[j]; r < A.ptr[j + 1]; r++) {
                    const int l = A.col[r];
                    const double a = A.val[r];
                    if (l == j || !(a * akk < 0)) continue;
                    if (l == i) dtil += aij * a / den;
                    else if (mark[l] >= 0) wsum[mark[l]] += aij * a / den;
                }
                continue;
            }
            dtil += aij;  // weak connections not in the interpolatory set: lumped
        }
        // row entries, truncated to the P_MAX largest (row sum kept)
        std::vector<std::pair<double, int>> ent;
        double sum_all = 0;
        for (size_t h = 0; h < hat.size(); h++) {
            const double wv = dtil != 0 ? -wsum[h] / dtil : 0.0;
            if (wv != 0) { ent.push_back({wv, cidx[hat[h]]}); sum_all += wv; }
        }
        if ((int)ent.size() > P_MAX) {
            std::stable_sort(ent.begin(), ent.end(), [](const auto &a, const auto &b) { return std::fabs(a.first) > std::fabs(b.first); });
            ent.resize(P_MAX);
            double sum_keep = 0;
            for (auto &e : ent) sum_keep += e.first;
            if (sum_keep != 0)
                for (auto &e : ent) e.first *= sum_all / sum_keep;
        }
        std::sort(ent.begin(), ent.end(), [](const auto &a, const auto &b) { return a.second < b.second; });
        for (auto &e : ent) { P.col.push_back(e.second); P.val.push_back(e.first); }
        P.ptr[i + 1] = (int)P.col.size();
        for (int q = sp[i]; q < sp[i + 1]; q++) strong[sc[q]] = 0;
        for (int h : hat) mark[h] = -1;
    }
    return P;
}

static Csr transpose(const Csr &A)
{
    Csr T;
    T.n = A.m;
    T.m = A.n;
    T.ptr.assign(T.n + 1, 0);
    for (int c : A.col) T.ptr[c + 1]++;
    for (int i = 0; i < T.n; i++) T.ptr[i + 1] += T.ptr[i];
    T.col.assign(A.col.size(), 0);
    T.val.assign(A.col.size(), 0);
    std::vector<int> fill(T.ptr.begin(), T.ptr.end() - 1);
    for (int i = 0; i < A.n; i++)
        for (int q = A.ptr[i]; q < A.ptr[i + 1]; q++) {
            const int at = fill[A.col[q]]++;
            T.col[at] = i;
            T.val[at] = A.val[q];
        }
    return T;
}

// C = A B (Gustavson, column-sorted rows)
static Csr multiply(const Csr &A, const Csr &B)
{
    Csr C;
    C.n = A.n;
    C.m = B.m;
    C.ptr.assign(A.n + 1, 0);
    std::vector<int> pos(B.m, -1), cols;
    std::vector<double> acc;
    for (int i = 0; i < A.n; i++) {
        cols.clear();
        acc.clear();
        for (int q = A.ptr[i]; q < A.ptr[i + 1]; q++) {
            const int k = A.col[q];
            const double a = A.val[q];
            for (int r = B.ptr[k]; r < B.ptr[k + 1]; r++) {
                const int j = B.col[r];
                if (pos[j] < 0) { pos[j] = (int)cols.size(); cols.push_back(j); acc.push_back(0.0); }
                acc[pos[j]] += a * B.val[r];
            }
        }
        std::vector<int> order(cols.size());
        for (size_t q = 0; q < order.size(); q++) order[q] = (int)q;
        std::sort(order.begin(), order.end(), [&](int a, int b) { return cols[a] < cols[b]; });
        for (int q : order) { C.col.push_back(cols[q]); C.val.push_back(acc[q]); }
        for (int j : cols) pos[j] = -1;
        C.ptr[i + 1] = (int)C.col.size();
    }
    return C;
}

// Greedy colouring of the symmetrised pattern of A (rows in index order,
// the smallest colour no neighbour holds).  Returns the rows grouped by
// colour (rows ascending inside a colour) and the colour offsets.
static void colouring(const Csr &A, std::vector<int> &rows, std::vector<int> &off)
{
    const int n = A.n;
    std::vector<int> tp, tc;
    transpose_pattern(n, A.ptr, A.col, tp, tc);
    std::vector<int> colour(n, -1), mark;
    int ncol = 0;
    for (int i = 0; i < n; i++) {
        mark.assign(ncol + 1, 0);
        for (int q = A.ptr[i]; q < A.ptr[i + 1]; q++)
            if (A.col[q] != i && colour[A.col[q]] >= 0) mark[colour[A.col[q]]] = 1;
        for (int q = tp[i]; q < tp[i + 1]; q++)
            if (tc[q] != i && colour[tc[q]] >= 0) mark[colour[tc[q]]] = 1;
        int c = 0;
        while (mark[c]) c++;
        colour[i] = c;
        if (c == ncol) ncol++;
    }
    off.assign(ncol + 1, 0);
    for (int i = 0; i < n; i++) off[colour[i] + 1]++;
    for (int c = 0; c < ncol; c++) off[c + 1] += off[c];
    rows.assign(n, 0);
    std::vector<int> pos(off.begin(), off.end() - 1);
    for (int i = 0; i < n; i++) rows[pos[colour[i]]++] = i;
}

struct Host {
    std::vector<Csr> A, P;
    std::vector<std::vector<int>> cf;
};

static int build(const Csr &A0, double theta, int max_levels, int max_coarse, Host &H)
{
    H.A.assign(1, A0);
    H.P.clear();
    H.cf.clear();
    while ((int)H.A.size() < max_levels && H.A.back().n > max_coarse) {
        const Csr &A = H.A.back();
        std::vector<int> sp, sc, cf, cidx;
        strength(A, theta, sp, sc);
        pmis(A.n, sp, sc, cf);
        Csr P = interpolation(A, sp, sc, cf, cidx);
        if (P.m == 0 || P.m >= A.n) break;  // no coarsening
        Csr AP = multiply(A, P);
        Csr Ac = multiply(transpose(P), AP);
        H.cf.push_back(std::move(cf));
        H.P.push_back(std::move(P));
        H.A.push_back(std::move(Ac));
    }
    return VFS_OK;
}

}  // namespace amg

// ---------------------------------------------------------------------------
// device hierarchy and cycle
// ---------------------------------------------------------------------------

constexpr int ATPB = 256;

// CSR rows are processed by groups of G lanes (G a power of two up to the
// wavefront, chosen per matrix from its mean row length): the lanes of a group
// stride over the row's entries and the partial sums are combined by a fixed
// xor-shuffle tree, so a row's result does not depend on the launch.  One lane
// per row left the coarse levels' 30-60-entry rows latency-bound (a chain of
// dependent loads per entry).

// sum over the row's entries of val * x (the diagonal left out when skip >= 0)
template <int G>
__device__ inline double group_row_dot(int r, const int *__restrict__ ptr, const int *__restrict__ col,
                                       const double *__restrict__ val, const double *__restrict__ x, int skip, int lane)
{
    double s = 0;
    for (int q = ptr[r] + lane; q < ptr[r + 1]; q += G) {
        const int j = col[q];
        if (j != skip) s += val[q] * x[j];
    }
#pragma unroll
    for (int m = G / 2; m > 0; m >>= 1) s += __shfl_xor(s, m, G);
    return s;
}

// r = b - A x (b != null) or y = A x
template <int G>
__global__ void k_csr_apply(int n, const int *__restrict__ ptr, const int *__restrict__ col, const double *__restrict__ val,
                            const double *__restrict__ x, const double *__restrict__ b, double *__restrict__ y)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = t / G, lane = t % G;
    if (i >= n) return;  // whole groups leave together (n * G threads)
    const double s = group_row_dot<G>(i, ptr, col, val, x, -1, lane);
    if (lane == 0) y[i] = b ? b[i] - s : s;
}

// y += M x (prolongation) or y = M x (restriction, add = 0)
template <int G>
__global__ void k_csr_mv(int n, int add, const int *__restrict__ ptr, const int *__restrict__ col,
                         const double *__restrict__ val, const double *__restrict__ x, double *__restrict__ y)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = t / G, lane = t % G;
    if (i >= n) return;
    const double s = group_row_dot<G>(i, ptr, col, val, x, -1, lane);
    if (lane == 0) y[i] = add ? y[i] + s : s;
}

// Gauss-Seidel on the rows of one colour: x_i = (b_i - sum_{j != i} a_ij x_j) / a_ii.
// Rows of one colour share no entry, so the pass is race-free and its result
// does not depend on the thread order.
template <int G>
__global__ void k_gs_colour(int nrows, const int *__restrict__ rows, const int *__restrict__ ptr,
                            const int *__restrict__ col, const double *__restrict__ val,
                            const double *__restrict__ dinv, const double *__restrict__ b, double *__restrict__ x)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int k = t / G, lane = t % G;
    if (k >= nrows) return;
    const int i = rows[k];
    const double s = group_row_dot<G>(i, ptr, col, val, x, i, lane);
    if (lane == 0) x[i] = (b[i] - s) * dinv[i];
}

// x = Inv b on the coarsest level (row-major inverse), one wavefront per row
__global__ void k_dense_mv(int n, const double *__restrict__ inv, const double *__restrict__ b, double *__restrict__ x)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = t / 64, lane = t % 64;
    if (i >= n) return;
    double s = 0;
    for (int q = lane; q < n; q += 64) s += inv[(long long)i * n + q] * b[q];
#pragma unroll
    for (int m = 32; m > 0; m >>= 1) s += __shfl_xor(s, m, 64);
    if (lane == 0) x[i] = s;
}

__global__ void k_gather(int n, const long long *__restrict__ idx, const double *__restrict__ v, double *__restrict__ out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = v[idx[i]];
}

__global__ void k_scatter(int n, const long long *__restrict__ idx, const double *__restrict__ x, double *__restrict__ z)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) z[idx[i]] = x[i];
}

__global__ void k_zero_d(long long n, double *__restrict__ x)
{
    const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (i < n) x[i] = 0;
}

struct AmgLevel {
    int n = 0;
    int *ap = nullptr, *ac = nullptr, *pp = nullptr, *pc = nullptr, *rp = nullptr, *rc = nullptr;
    double *av = nullptr, *pv = nullptr, *rv = nullptr, *dinv = nullptr;
    double *x = nullptr, *b = nullptr, *r = nullptr;
    int *crows = nullptr;          // rows grouped by colour
    std::vector<int> coff;         // colour offsets into crows (host)
    int ga = 1, gp = 1, gr = 1;    // lanes per row of A, P and R
};

struct AmgDev {
    std::vector<AmgLevel> L;
    double *inv = nullptr;  // coarsest level
    long long *ridx = nullptr;  // DA point of each level-0 row
    double theta = -1;
    std::vector<void *> allocs;
    double setup_ms = 0;
    hipGraphExec_t cycle = nullptr;  // the captured V-cycle on the level buffers
};

void amg_free(AmgDev *a)
{
    if (!a) return;
    if (a->cycle) hipGraphExecDestroy(a->cycle);
    for (void *p : a->allocs) hipFree(p);
    delete a;
}

// lanes per row for a matrix: the power of two at or above its mean row
// length, 2 to 64
static int lanes_for(const amg::Csr &M)
{
    const double mean = M.n ? (double)M.nnz() / M.n : 1.0;
    int g = 2;
    while (g < 64 && g < mean) g *= 2;
    return g;
}

#define VFS_AMG_GROUPED(KERN)                                                                                     \
    template <typename... Args>                                                                                   \
    static void launch_##KERN(int g, int rows, hipStream_t st, Args... args)                                      \
    {                                                                                                             \
        const unsigned nb = blocks_for((long long)rows * g, ATPB);                                                \
        switch (g) {                                                                                              \
        case 2: hipLaunchKernelGGL(KERN<2>, dim3(nb), dim3(ATPB), 0, st, args...); break;                         \
        case 4: hipLaunchKernelGGL(KERN<4>, dim3(nb), dim3(ATPB), 0, st, args...); break;                         \
        case 8: hipLaunchKernelGGL(KERN<8>, dim3(nb), dim3(ATPB), 0, st, args...); break;                         \
        case 16: hipLaunchKernelGGL(KERN<16>, dim3(nb), dim3(ATPB), 0, st, args...); break;                       \
        case 32: hipLaunchKernelGGL(KERN<32>, dim3(nb), dim3(ATPB), 0, st, args...); break;                       \
        default: hipLaunchKernelGGL(KERN<64>, dim3(nb), dim3(ATPB), 0, st, args...); break;                       \
        }                                                                                                         \
    }
VFS_AMG_GROUPED(k_csr_apply)
VFS_AMG_GROUPED(k_csr_mv)
VFS_AMG_GROUPED(k_gs_colour)
#undef VFS_AMG_GROUPED

template <typename T>
static int dev_copy(AmgDev *a, T **dst, const T *src, size_t count)
{
    hipError_t e = hipMalloc((void **)dst, (count ? count : 1) * sizeof(T));
    if (e != hipSuccess) return check_hip(e, "hipMalloc (amg)");
    a->allocs.push_back(*dst);
    if (src && count) {
        e = hipMemcpy(*dst, src, count * sizeof(T), hipMemcpyHostToDevice);
        if (e != hipSuccess) return check_hip(e, "hipMemcpy (amg)");
    } else {
        e = hipMemset(*dst, 0, (count ? count : 1) * sizeof(T));
        if (e != hipSuccess) return check_hip(e, "hipMemset (amg)");
    }
    return VFS_OK;
}

// Gauss-Jordan inverse of A + s 1 1^T (s = mean |diagonal|), row-major
static bool dense_inverse(const amg::Csr &A, std::vector<double> &inv)
{
    const int n = A.n;
    std::vector<double> M((size_t)n * 2 * n, 0.0);
    double s = 0;
    for (int i = 0; i < n; i++)
        for (int q = A.ptr[i]; q < A.ptr[i + 1]; q++)
            if (A.col[q] == i) s += std::fabs(A.val[q]);
    s = n ? s / n : 1.0;
    const int w = 2 * n;
    for (int i = 0; i < n; i++) {
        for (int j = 0; j < n; j++) M[(size_t)i * w + j] = s;
        for (int q = A.ptr[i]; q < A.ptr[i + 1]; q++) M[(size_t)i * w + A.col[q]] += A.val[q];
        M[(size_t)i * w + n + i] = 1.0;
    }
    for (int c = 0; c < n; c++) {
        int piv = c;
        for (int r = c + 1; r < n; r++)
            if (std::fabs(M[(size_t)r * w + c]) > std::fabs(M[(size_t)piv * w + c])) piv = r;
        if (M[(size_t)piv * w + c] == 0) return false;
        if (piv != c)
            for (int j = 0; j < w; j++) std::swap(M[(size_t)piv * w + j], M[(size_t)c * w + j]);
        const double d = 1.0 / M[(size_t)c * w + c];
        for (int j = 0; j < w; j++) M[(size_t)c * w + j] *= d;
        for (int r = 0; r < n; r++) {
            if (r == c) continue;
            const double f = M[(size_t)r * w + c];
            if (f == 0) continue;
            for (int j = 0; j < w; j++) M[(size_t)r * w + j] -= f * M[(size_t)c * w + j];
        }
    }
    inv.assign((size_t)n * n, 0.0);
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) inv[(size_t)i * n + j] = M[(size_t)i * w + n + j];
    return true;
}

// Builds the hierarchy from the level-0 operator: n_pts DA points, 19
// coefficient planes c64[s * n_pts + p], row mask, neighbour index of (p, s).
int amg_build(const double *c64_host, const unsigned char *row_host, long long n_pts,
              long long (*nbr)(const void *g, long long p, int s), const void *grid, double theta, AmgDev **out)
{
    *out = nullptr;
    std::vector<long long> pts;
    std::vector<int> rid(n_pts, -1);
    for (long long p = 0; p < n_pts; p++)
        if (row_host[p]) { rid[p] = (int)pts.size(); pts.push_back(p); }
    if (pts.empty()) return set_error(VFS_EINVAL, "AMG: the operator has no rows");
    if (pts.size() >= (size_t)INT32_MAX / 32) return set_error(VFS_EINVAL, "AMG: too many rows");
    amg::Csr A;
    A.n = A.m = (int)pts.size();
    A.ptr.assign(A.n + 1, 0);
    std::vector<std::pair<int, double>> ent;
    for (int r = 0; r < A.n; r++) {
        ent.clear();
        const long long p = pts[r];
        for (int s = 0; s < 19; s++) {
            const double v = c64_host[(long long)s * n_pts + p];
            if (v == 0) continue;
            const long long q = nbr(grid, p, s);
            if (!row_host[q]) continue;
            ent.push_back({rid[q], v});
        }
        std::sort(ent.begin(), ent.end(), [](const auto &a, const auto &b) { return a.first < b.first; });
        for (size_t e = 0; e < ent.size(); e++) {
            if (!A.col.empty() && (int)A.col.size() > A.ptr[r] && A.col.back() == ent[e].first) A.val.back() += ent[e].second;
            else { A.col.push_back(ent[e].first); A.val.push_back(ent[e].second); }
        }
        A.ptr[r + 1] = (int)A.col.size();
    }
    amg::Host H;
    int rc = amg::build(A, theta, 25, amg::MAX_COARSE, H);
    if (rc) return rc;
    std::vector<double> inv;
    if (!dense_inverse(H.A.back(), inv)) return set_error(VFS_EINVAL, "AMG: singular coarsest operator");
    AmgDev *a = new AmgDev();
    a->theta = theta;
    const int nl = (int)H.A.size();
    a->L.resize(nl);
    for (int l = 0; l < nl && !rc; l++) {
        AmgLevel &L = a->L[l];
        const amg::Csr &M = H.A[l];
        L.n = M.n;
        std::vector<double> dinv(M.n, 0.0);
        for (int i = 0; i < M.n; i++)
            for (int q = M.ptr[i]; q < M.ptr[i + 1]; q++)
                if (M.col[q] == i && M.val[q] != 0) dinv[i] = 1.0 / M.val[q];
        std::vector<int> crows;
        amg::colouring(M, crows, L.coff);
        L.ga = lanes_for(M);
        rc = rc ? rc : dev_copy(a, &L.crows, crows.data(), crows.size());
        rc = rc ? rc : dev_copy(a, &L.ap, M.ptr.data(), M.ptr.size());
        rc = rc ? rc : dev_copy(a, &L.ac, M.col.data(), M.col.size());
        rc = rc ? rc : dev_copy(a, &L.av, M.val.data(), M.val.size());
        rc = rc ? rc : dev_copy(a, &L.dinv, dinv.data(), dinv.size());
        rc = rc ? rc : dev_copy<double>(a, &L.x, nullptr, M.n);
        rc = rc ? rc : dev_copy<double>(a, &L.b, nullptr, M.n);
        rc = rc ? rc : dev_copy<double>(a, &L.r, nullptr, M.n);
        if (l + 1 < nl) {
            const amg::Csr &P = H.P[l];
            const amg::Csr R = amg::transpose(P);
            L.gp = lanes_for(P);
            L.gr = lanes_for(R);
            rc = rc ? rc : dev_copy(a, &L.pp, P.ptr.data(), P.ptr.size());
            rc = rc ? rc : dev_copy(a, &L.pc, P.col.data(), P.col.size());
            rc = rc ? rc : dev_copy(a, &L.pv, P.val.data(), P.val.size());
            rc = rc ? rc : dev_copy(a, &L.rp, R.ptr.data(), R.ptr.size());
            rc = rc ? rc : dev_copy(a, &L.rc, R.col.data(), R.col.size());
            rc = rc ? rc : dev_copy(a, &L.rv, R.val.data(), R.val.size());
        }
    }
    rc = rc ? rc : dev_copy(a, &a->inv, inv.data(), inv.size());
    rc = rc ? rc : dev_copy(a, &a->ridx, pts.data(), pts.size());
    if (rc) { amg_free(a); return rc; }
    *out = a;
    return VFS_OK;
}

static void amg_cycle(hipStream_t st, AmgDev *a, int l)
{
    AmgLevel &L = a->L[l];
    const unsigned nb = blocks_for(L.n, ATPB);
    if (l + 1 == (int)a->L.size()) {
        hipLaunchKernelGGL(k_dense_mv, dim3(blocks_for(64LL * L.n, ATPB)), dim3(ATPB), 0, st, L.n, a->inv, L.b, L.x);
        return;
    }
    AmgLevel &C = a->L[l + 1];
    const int ncol = (int)L.coff.size() - 1;
    auto relax = [&](int c) {
        const int cnt = L.coff[c + 1] - L.coff[c];
        launch_k_gs_colour(L.ga, cnt, st, cnt, (const int *)L.crows + L.coff[c], (const int *)L.ap, (const int *)L.ac,
                           (const double *)L.av, (const double *)L.dinv, (const double *)L.b, L.x);
    };
    // pre-smoothing (forward colours) from a zero iterate, residual, restriction
    hipLaunchKernelGGL(k_zero_d, dim3(nb), dim3(ATPB), 0, st, (long long)L.n, L.x);
    for (int c = 0; c < ncol; c++) relax(c);
    launch_k_csr_apply(L.ga, L.n, st, L.n, (const int *)L.ap, (const int *)L.ac, (const double *)L.av, (const double *)L.x,
                       (const double *)L.b, L.r);
    launch_k_csr_mv(L.gr, C.n, st, C.n, 0, (const int *)L.rp, (const int *)L.rc, (const double *)L.rv,
                    (const double *)L.r, C.b);
    amg_cycle(st, a, l + 1);
    // prolongation and post-smoothing (backward colours)
    launch_k_csr_mv(L.gp, L.n, st, L.n, 1, (const int *)L.pp, (const int *)L.pc, (const double *)L.pv,
                    (const double *)C.x, L.x);
    for (int c = ncol - 1; c >= 0; c--) relax(c);
}

// z = M^-1 v for DA-layout vectors (n_pts points; zero outside the rows).  The
// cycle itself works on the level buffers only, a fixed sequence of a few
// hundred small kernels: it is captured once into a HIP graph (kernel nodes
// only) and replayed; VFS_AMG_GRAPH=0 launches the same kernels directly.
int amg_apply(hipStream_t st, AmgDev *a, long long n_pts, const double *v, double *z)
{
    AmgLevel &L = a->L[0];
    const unsigned nb = blocks_for(L.n, ATPB);
    hipLaunchKernelGGL(k_gather, dim3(nb), dim3(ATPB), 0, st, L.n, a->ridx, v, L.b);
    const char *env = std::getenv("VFS_AMG_GRAPH");
    const bool graph = !(env && env[0] == '0');
    if (graph && !a->cycle) {
        hipGraph_t g = nullptr;
        if (hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal) == hipSuccess) {
            amg_cycle(st, a, 0);
            if (hipStreamEndCapture(st, &g) == hipSuccess && g) {
                if (hipGraphInstantiate(&a->cycle, g, nullptr, nullptr, 0) != hipSuccess) a->cycle = nullptr;
                hipGraphDestroy(g);
            }
        }
    }
    if (graph && a->cycle) hipGraphLaunch(a->cycle, st);
    else amg_cycle(st, a, 0);
    hipLaunchKernelGGL(k_zero_d, dim3(blocks_for(n_pts, ATPB)), dim3(ATPB), 0, st, n_pts, z);
    hipLaunchKernelGGL(k_scatter, dim3(nb), dim3(ATPB), 0, st, L.n, a->ridx, L.x, z);
    VFS_LAUNCH_CHECK("amg v-cycle");
    return VFS_OK;
}

int amg_levels(const AmgDev *a) { return a ? (int)a->L.size() : 0; }
double amg_theta(const AmgDev *a) { return a ? a->theta : -1; }

}  // namespace vfs

// ---------------------------------------------------------------------------
// host-side diagnostics (C ABI)
// ---------------------------------------------------------------------------
struct vfs_amg_host {
    vfs::amg::Host H;
};

extern "C" {

int vfs_amg_host_build(int n, const int *ptr, const int *col, const double *val, double theta, int max_levels,
                       int max_coarse, vfs_amg_host **out)
{
    if (n <= 0 || !ptr || !col || !val || !out || max_levels < 1 || !(theta >= 0 && theta < 1))
        return vfs::set_error(VFS_EINVAL, "vfs_amg_host_build: bad argument");
    vfs::amg::Csr A;
    A.n = A.m = n;
    A.ptr.assign(ptr, ptr + n + 1);
    A.col.assign(col, col + ptr[n]);
    A.val.assign(val, val + ptr[n]);
    for (int c : A.col)
        if (c < 0 || c >= n) return vfs::set_error(VFS_EINVAL, "vfs_amg_host_build: column index out of range");
    vfs_amg_host *h = new vfs_amg_host();
    int rc = vfs::amg::build(A, theta, max_levels, max_coarse > 0 ? max_coarse : vfs::amg::MAX_COARSE, h->H);
    if (rc) { delete h; return rc; }
    *out = h;
    return VFS_OK;
}

int vfs_amg_host_levels(const vfs_amg_host *h, int *nlevels)
{
    if (!h || !nlevels) return vfs::set_error(VFS_EINVAL, "vfs_amg_host_levels: null argument");
    *nlevels = (int)h->H.A.size();
    return VFS_OK;
}

// which = 0: the level's operator A (rows x rows); 1: its interpolation P
// (rows x next level's rows; not for the coarsest level).  With null arrays
// only the sizes are returned.
int vfs_amg_host_matrix(const vfs_amg_host *h, int level, int which, int *rows, int *cols, long long *nnz, int *ptr,
                        int *col, double *val)
{
    if (!h || level < 0 || level >= (int)h->H.A.size() || (which == 1 && level >= (int)h->H.P.size()) || which < 0 ||
        which > 1)
        return vfs::set_error(VFS_EINVAL, "vfs_amg_host_matrix: no such level or matrix");
    const vfs::amg::Csr &M = which ? h->H.P[level] : h->H.A[level];
    if (rows) *rows = M.n;
    if (cols) *cols = M.m;
    if (nnz) *nnz = M.nnz();
    if (ptr) std::copy(M.ptr.begin(), M.ptr.end(), ptr);
    if (col) std::copy(M.col.begin(), M.col.end(), col);
    if (val) std::copy(M.val.begin(), M.val.end(), val);
    return VFS_OK;
}

// C/F splitting of a level (1 = C, -1 = F), not for the coarsest level
int vfs_amg_host_splitting(const vfs_amg_host *h, int level, int *cf)
{
    if (!h || !cf || level < 0 || level >= (int)h->H.cf.size())
        return vfs::set_error(VFS_EINVAL, "vfs_amg_host_splitting: no such level");
    std::copy(h->H.cf[level].begin(), h->H.cf[level].end(), cf);
    return VFS_OK;
}

// The smoother's colouring of a level: rows grouped by colour (n entries)
// and the colour offsets (ncolours + 1 entries).  With null arrays only
// *ncolours is returned.
int vfs_amg_host_colouring(const vfs_amg_host *h, int level, int *ncolours, int *rows, int *offsets)
{
    if (!h || !ncolours || level < 0 || level >= (int)h->H.A.size())
        return vfs::set_error(VFS_EINVAL, "vfs_amg_host_colouring: no such level");
    std::vector<int> r, off;
    vfs::amg::colouring(h->H.A[level], r, off);
    *ncolours = (int)off.size() - 1;
    if (rows) std::copy(r.begin(), r.end(), rows);
    if (offsets) std::copy(off.begin(), off.end(), offsets);
    return VFS_OK;
}

int vfs_amg_host_free(vfs_amg_host *h)
{
    delete h;
    return VFS_OK;
}

}  // extern "C"
