// vfs_io.hip -- the reference's file formats on the host side of the boundary:
// PETSc binary Vec files (the vfield/ufield/pfield/nvfield checkpoints of
// main.c:376-934) and the grid readers of init.c:258-394 (grid.dat as text or
// binary, xyz.dat).  Host code only; fields move through vfs_upload /
// vfs_download, so a read invalidates whatever the context derived from the
// old contents.
#include <cerrno>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "vfs_internal.h"

namespace {

// PETSc binary files are big-endian whatever the host (PetscBinaryWrite swaps
// on little-endian machines).  A Vec is written by VecView as the class cookie
// VEC_FILE_COOKIE = 1211214 (PETSc 3.1; VEC_FILE_CLASSID in later releases),
// the global length N as a 32-bit integer, then N scalars; a DMDA vector is
// written in the DA's natural ordering, which on one rank is the global
// ordering (i fastest, dof interleaved).
constexpr int32_t VEC_COOKIE = 1211214;

uint32_t bswap32(uint32_t v) { return __builtin_bswap32(v); }
uint64_t bswap64(uint64_t v) { return __builtin_bswap64(v); }

struct File {
    FILE *f = nullptr;
    File(const char *path, const char *mode) : f(std::fopen(path, mode)) {}
    ~File()
    {
        if (f) std::fclose(f);
    }
};

int io_error(const char *what, const char *path)
{
    std::string m = std::string(what) + ": " + (path ? path : "(null)");
    if (errno) m += std::string(" (") + std::strerror(errno) + ")";
    return vfs::set_error(VFS_EINVAL, m.c_str());
}

// Reads the header of a Vec file: its length.
int vec_header(FILE *f, const char *path, long long *n)
{
    uint32_t h[2];
    if (std::fread(h, sizeof(uint32_t), 2, f) != 2) return io_error("short PETSc Vec file header", path);
    const int32_t cookie = (int32_t)bswap32(h[0]), len = (int32_t)bswap32(h[1]);
    if (cookie != VEC_COOKIE) return io_error("not a PETSc binary Vec file (bad class cookie)", path);
    if (len < 0) return io_error("negative Vec length in", path);
    *n = len;
    return VFS_OK;
}

std::string field_file(const char *dir, const char *stem, int ti, int block)
{
    char name[64];
    std::snprintf(name, sizeof(name), "%s%06d_%1d.dat", stem, ti, block);  // main.c:386
    return std::string(dir) + "/" + name;
}

// The checkpointed fields of Ucont_P_Binary_Output (main.c:684-770, qcr 0)
struct CkField {
    const char *stem;
    vfs_field field;
};
constexpr CkField CK_OUT[] = {{"vfield", VFS_UCONT}, {"ufield", VFS_UCAT}, {"pfield", VFS_P}, {"nvfield", VFS_NVERT}};

}  // namespace

extern "C" {

int vfs_vec_write(const char *path, const double *v, long long n)
{
    if (!path || (!v && n > 0) || n < 0 || n > INT32_MAX) return vfs::set_error(VFS_EINVAL, "vfs_vec_write: bad argument");
    errno = 0;
    File f(path, "wb");
    if (!f.f) return io_error("cannot open for writing", path);
    const uint32_t h[2] = {bswap32((uint32_t)VEC_COOKIE), bswap32((uint32_t)n)};
    if (std::fwrite(h, sizeof(uint32_t), 2, f.f) != 2) return io_error("write failed", path);
    std::vector<uint64_t> buf(1 << 16);
    for (long long off = 0; off < n;) {
        const long long cnt = std::min<long long>((long long)buf.size(), n - off);
        for (long long q = 0; q < cnt; q++) {
            uint64_t b;
            std::memcpy(&b, v + off + q, sizeof(b));
            buf[q] = bswap64(b);
        }
        if (std::fwrite(buf.data(), sizeof(uint64_t), (size_t)cnt, f.f) != (size_t)cnt) return io_error("write failed", path);
        off += cnt;
    }
    if (std::fflush(f.f) != 0) return io_error("write failed", path);
    return VFS_OK;
}

int vfs_vec_length(const char *path, long long *n)
{
    if (!path || !n) return vfs::set_error(VFS_EINVAL, "vfs_vec_length: bad argument");
    errno = 0;
    File f(path, "rb");
    if (!f.f) return io_error("cannot open", path);
    return vec_header(f.f, path, n);
}

int vfs_vec_read(const char *path, double *v, long long n)
{
    if (!path || (!v && n > 0) || n < 0) return vfs::set_error(VFS_EINVAL, "vfs_vec_read: bad argument");
    errno = 0;
    File f(path, "rb");
    if (!f.f) return io_error("cannot open", path);
    long long len;
    int rc = vec_header(f.f, path, &len);
    if (rc) return rc;
    // VecLoadIntoVector refuses a file whose length differs from the vector's
    if (len != n) {
        char m[160];
        std::snprintf(m, sizeof(m), "Vec length %lld in the file, %lld expected", len, n);
        return io_error(m, path);
    }
    std::vector<uint64_t> buf(1 << 16);
    for (long long off = 0; off < n;) {
        const long long cnt = std::min<long long>((long long)buf.size(), n - off);
        if (std::fread(buf.data(), sizeof(uint64_t), (size_t)cnt, f.f) != (size_t)cnt) return io_error("short PETSc Vec file", path);
        for (long long q = 0; q < cnt; q++) {
            const uint64_t b = bswap64(buf[q]);
            std::memcpy(v + off + q, &b, sizeof(b));
        }
        off += cnt;
    }
    return VFS_OK;
}

int vfs_checkpoint_write(vfs_ctx *c, const char *dir, int ti, int block)
{
    if (!c || !dir || ti < 0 || block < 0 || block > 9) return vfs::set_error(VFS_EINVAL, "vfs_checkpoint_write: bad argument");
    std::vector<double> h;
    for (const CkField &ck : CK_OUT) {
        h.resize(vfs_field_size(c, ck.field));
        int rc = vfs_download(c, ck.field, h.data());
        if (rc) return rc;
        rc = vfs_vec_write(field_file(dir, ck.stem, ti, block).c_str(), h.data(), (long long)h.size());
        if (rc) return rc;
    }
    return VFS_OK;
}

int vfs_checkpoint_read(vfs_ctx *c, const char *dir, int ti, int block, int immersed)
{
    if (!c || !dir || ti < 0 || block < 0 || block > 9) return vfs::set_error(VFS_EINVAL, "vfs_checkpoint_read: bad argument");
    std::vector<double> h;
    auto load = [&](const char *stem, vfs_field f) {
        h.resize(vfs_field_size(c, f));
        int rc = vfs_vec_read(field_file(dir, stem, ti, block).c_str(), h.data(), (long long)h.size());
        return rc ? rc : vfs_upload(c, f, h.data());
    };
    // main.c:385-405: Ucont, P, Nvert_o, Ucat in this order
    int rc = load("vfield", VFS_UCONT);
    if (!rc) rc = load("pfield", VFS_P);
    if (!rc) {
        if (immersed) {
            rc = load("nvfield", VFS_NVERT);
        } else {
            // main.c:407-410: without immersed bodies the blanking is zero
            h.assign(vfs_field_size(c, VFS_NVERT), 0.0);
            rc = vfs_upload(c, VFS_NVERT, h.data());
        }
    }
    if (!rc) rc = load("ufield", VFS_UCAT);
    if (rc) return rc;
    // main.c:418-426: Ucont_o = Ucont, lUcat_old = Ucat
    h.resize(vfs_field_size(c, VFS_UCONT));
    if ((rc = vfs_download(c, VFS_UCONT, h.data())) || (rc = vfs_upload(c, VFS_UCONT_O, h.data()))) return rc;
    if ((rc = vfs_download(c, VFS_UCAT, h.data())) || (rc = vfs_upload(c, VFS_UCAT_OLD, h.data()))) return rc;
    return VFS_OK;
}

// ---- inflow planes (rhs2.c:716-862) ---------------------------------------

int vfs_inflow_write_section(const char *dir, int ti, int tistart, int save_inflow_period, int save_inflow_minus,
                             double dt, int mx, int my, const double *plane)
{
    if (!dir || !plane || save_inflow_period <= 0 || mx <= 0 || my <= 0)
        return vfs::set_error(VFS_EINVAL, "vfs_inflow_write_section: bad argument");
    const int ti0 = ti - save_inflow_minus;
    const int ti_name = ((ti0 - 1) / save_inflow_period) * save_inflow_period + 1;  // rhs2.c:728
    char name[96];
    std::snprintf(name, sizeof(name), "inflow_%06d_dt=%g.dat", ti_name, dt);
    const std::string path = std::string(dir) + "/" + name;
    if (ti0 == tistart || ti0 % save_inflow_period == 1) std::remove(path.c_str());  // a new file (rhs2.c:731)
    errno = 0;
    File f(path.c_str(), "ab");
    if (!f.f) return io_error("cannot open for appending", path.c_str());
    // my rows of mx Cmpnts (rhs2.c:736)
    if (std::fwrite(plane, 3 * sizeof(double), (size_t)mx * my, f.f) != (size_t)mx * my)
        return io_error("write failed", path.c_str());
    return VFS_OK;
}

}  // extern "C"

struct vfs_inflow_reader {
    std::string path_inflow;
    double dt_inflow;
    int period, minus, recycle, tistart, mx, my;
    FILE *fp;
};

extern "C" {

int vfs_inflow_reader_open(const char *path_inflow, double dt_inflow, int save_inflow_period, int save_inflow_minus,
                           int recycle_period, int tistart, int mx, int my, vfs_inflow_reader **out)
{
    if (!path_inflow || !out || save_inflow_period <= 0 || recycle_period <= 0 || mx <= 0 || my <= 0)
        return vfs::set_error(VFS_EINVAL, "vfs_inflow_reader_open: bad argument");
    *out = new vfs_inflow_reader{path_inflow, dt_inflow, save_inflow_period, save_inflow_minus, recycle_period,
                                 tistart, mx, my, nullptr};
    return VFS_OK;
}

int vfs_inflow_reader_next(vfs_inflow_reader *r, int ti, double *plane)
{
    if (!r || !plane) return vfs::set_error(VFS_EINVAL, "vfs_inflow_reader_next: bad argument");
    const size_t np = (size_t)r->mx * r->my;
    // rhs2.c:797-805: the time index inside the recycled record and its file
    int ti2 = ti - r->minus;
    if (ti2 == 0) ti2 = 1;
    if (ti2 > r->recycle) ti2 -= (ti2 / r->recycle) * r->recycle;
    const int ti_name = ((ti2 - 1) / r->period) * r->period + 1;
    char name[96];
    std::snprintf(name, sizeof(name), "/inflow/inflow_%06d_dt=%g.dat", ti_name, r->dt_inflow);
    const std::string path = r->path_inflow + name;
    // rhs2.c:807-820: (re)open at the first step and at the first step of each file
    if (ti == r->tistart || (ti > r->tistart + 90 && ti2 % r->period == 1)) {
        if (ti != r->tistart && r->fp) std::fclose(r->fp);
        errno = 0;
        r->fp = std::fopen(path.c_str(), "rb");
        if (!r->fp) return io_error("cannot open the inflow file", path.c_str());
        if (ti == r->tistart)
            for (int it = 0; it < (ti2 - 1) % r->period; it++)
                if (std::fseek(r->fp, (long)(np * 3 * sizeof(double)), SEEK_CUR) != 0)
                    return io_error("cannot position in the inflow file", path.c_str());
    }
    // rhs2.c:821-823: nothing is read at the first step of a fresh start, and
    // the plane is then zero (the reference sums zeroed per-rank buffers)
    if (r->tistart == 0 && ti == 1) {
        std::memset(plane, 0, np * 3 * sizeof(double));
        return VFS_OK;
    }
    if (!r->fp) return vfs::set_error(VFS_EINVAL, "vfs_inflow_reader_next: the first call must be for ti = tistart");
    // past the end of a file (within 90 steps of a restart the reference does
    // not switch files) the values not read stay zero, as in the reference's
    // zeroed read buffer
    std::memset(plane, 0, np * 3 * sizeof(double));
    if (std::fread(plane, 3 * sizeof(double), np, r->fp) != np) std::clearerr(r->fp);
    return VFS_OK;
}

int vfs_inflow_reader_close(vfs_inflow_reader *r)
{
    if (!r) return VFS_OK;
    if (r->fp) std::fclose(r->fp);
    delete r;
    return VFS_OK;
}

// ---- grid readers (init.c:258-394) --------------------------------------

namespace {

struct GridReader {
    FILE *f;
    int format;  // VFS_GRID_TEXT, VFS_GRID_BINARY, VFS_GRID_XYZ
    bool read_int(int *v)
    {
        if (format == VFS_GRID_BINARY) return std::fread(v, sizeof(int), 1, f) == 1;
        return std::fscanf(f, "%i", v) == 1;
    }
    bool read_double(double *v)
    {
        if (format == VFS_GRID_BINARY) return std::fread(v, sizeof(double), 1, f) == 1;
        return std::fscanf(f, "%le", v) == 1;
    }
};

// Positions the reader at block `block`'s coordinates; returns its node counts.
int grid_seek(GridReader &r, const char *path, int block, int *IM, int *JM, int *KM)
{
    int nblock = 1;
    if (r.format != VFS_GRID_XYZ && !r.read_int(&nblock)) return io_error("cannot read the block count of", path);
    if (block < 0 || block >= nblock) return io_error("no such block in", path);
    for (int b = 0;; b++) {
        if (!r.read_int(IM) || !r.read_int(JM) || !r.read_int(KM)) return io_error("cannot read IM JM KM of", path);
        if (*IM < 2 || *JM < 2 || *KM < 2) return io_error("bad node counts in", path);
        if (b == block) return VFS_OK;
        // skip this block's coordinates (three values per node, or per axis for xyz.dat)
        const long long nv = 3LL * *IM * *JM * *KM;
        double d;
        for (long long q = 0; q < nv; q++)
            if (!r.read_double(&d)) return io_error("short grid file", path);
    }
}

}  // namespace

int vfs_grid_dims(const char *path, int format, int block, int *IM, int *JM, int *KM)
{
    if (!path || !IM || !JM || !KM || format < VFS_GRID_TEXT || format > VFS_GRID_XYZ)
        return vfs::set_error(VFS_EINVAL, "vfs_grid_dims: bad argument");
    errno = 0;
    File f(path, format == VFS_GRID_BINARY ? "rb" : "r");
    if (!f.f) return io_error("cannot open", path);
    GridReader r{f.f, format};
    return grid_seek(r, path, block, IM, JM, KM);
}

int vfs_grid_read(const char *path, int format, int block, double cl, double L_dim, int mx, int my, int mz, double *coor)
{
    if (!path || !coor || format < VFS_GRID_TEXT || format > VFS_GRID_XYZ || cl == 0.0)
        return vfs::set_error(VFS_EINVAL, "vfs_grid_read: bad argument");
    errno = 0;
    File f(path, format == VFS_GRID_BINARY ? "rb" : "r");
    if (!f.f) return io_error("cannot open", path);
    GridReader r{f.f, format};
    int IM, JM, KM;
    int rc = grid_seek(r, path, block, &IM, &JM, &KM);
    if (rc) return rc;
    // the DA holds IM+1 x JM+1 x KM+1 points; the nodes fill indices 0..IM-1 etc. (init.c:157-160, 320-370)
    if (mx != IM + 1 || my != JM + 1 || mz != KM + 1) {
        char m[160];
        std::snprintf(m, sizeof(m), "grid has %d x %d x %d nodes, the DA needs %d x %d x %d", IM, JM, KM, mx - 1, my - 1, mz - 1);
        return io_error(m, path);
    }
    std::memset(coor, 0, sizeof(double) * 3 * (size_t)mx * my * mz);
    auto at = [&](int k, int j, int i, int c) -> double & { return coor[3 * (((size_t)k * my + j) * mx + i) + c]; };
    if (format == VFS_GRID_XYZ) {
        // init.c:313-321: IM lines give x (first column), JM lines y (second), KM lines z (third)
        std::vector<double> X(IM), Y(JM), Z(KM);
        double t0, t1;
        for (int i = 0; i < IM; i++)
            if (std::fscanf(f.f, "%le %le %le", &X[i], &t0, &t1) != 3) return io_error("short xyz.dat", path);
        for (int j = 0; j < JM; j++)
            if (std::fscanf(f.f, "%le %le %le", &t0, &Y[j], &t1) != 3) return io_error("short xyz.dat", path);
        for (int k = 0; k < KM; k++)
            if (std::fscanf(f.f, "%le %le %le", &t0, &t1, &Z[k]) != 3) return io_error("short xyz.dat", path);
        for (int k = 0; k < KM; k++)
            for (int j = 0; j < JM; j++)
                for (int i = 0; i < IM; i++) {
                    at(k, j, i, 0) = X[i] / cl * L_dim;
                    at(k, j, i, 1) = Y[j] / cl * L_dim;
                    at(k, j, i, 2) = Z[k] / cl * L_dim;
                }
        return VFS_OK;
    }
    // init.c:337-381: all x in (k, j, i) order, then all y, then all z
    for (int c = 0; c < 3; c++)
        for (int k = 0; k < KM; k++)
            for (int j = 0; j < JM; j++)
                for (int i = 0; i < IM; i++) {
                    double v;
                    if (!r.read_double(&v)) return io_error("short grid file", path);
                    at(k, j, i, c) = v / cl * L_dim;
                }
    return VFS_OK;
}

}  // extern "C"
