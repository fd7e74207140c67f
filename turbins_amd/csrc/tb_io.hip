// tb_io.hip -- restart and field files in the reference's on-disk formats
// (host code only).
//
// PETSc binary vectors: the reference writes every restart vector with
// PetscViewerBinaryOpen + VecView and reads it back with VecLoadIntoVector
// (Resume.c:96-139; field dumps the same way, Output.c:1610-1611).  PETSc's
// binary viewer stores a Vec as
//     int32  VEC_FILE_CLASSID (1211214)
//     int32  N (global length)
//     N x    IEEE-754 double
// all big-endian, and a DA global vector in the DA's natural order (x
// fastest), which is the host layout of tb_download_field / tb_upload_field.
// The set of restart files and their names follow Resume_set_data
// (Resume.c:172-244); Resume_primary.dat is the ASCII record of
// Resume_write_primary_data / Resume_read_primary_data (Resume.c:247-319).
#include <algorithm>
#include <cerrno>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "tb_common.h"

namespace {

constexpr int32_t VEC_FILE_CLASSID = 1211214;
constexpr size_t CHUNK = 1 << 20;   // doubles converted per write / read

uint32_t be32(uint32_t v) { return __builtin_bswap32(v); }
uint64_t be64(uint64_t v) { return __builtin_bswap64(v); }

struct File {
    FILE *f = nullptr;
    File(const char *path, const char *mode) : f(fopen(path, mode)) {}
    ~File() {
        if (f) fclose(f);
    }
};

std::string join(const char *dir, const char *name) {
    std::string s = dir ? dir : ".";
    if (!s.empty() && s.back() != '/') s += '/';
    return s + name;
}

// restart vectors: file name and field (Resume.c:186-228); the
// concentration fields follow as Resume_conc<i>.bin (Resume.c:229-238)
struct ResumeVec {
    const char *name;
    int field;
};
const ResumeVec RESUME_VECS[] = {
    {"Resume_u.bin", TB_F_U},       {"Resume_v.bin", TB_F_V},       {"Resume_w.bin", TB_F_W},
    {"Resume_dpdx.bin", TB_F_DPDX}, {"Resume_dpdy.bin", TB_F_DPDY}, {"Resume_dpdz.bin", TB_F_DPDZ},
};

// saves (load = false) or loads every concentration field's Resume_conc<i>.bin;
// the selected field is selected again afterwards
int resume_conc(tb_ctx *c, const char *dir, bool load) {
    const int sel = c->conc_sel;
    for (int n = 0; n < c->nconc; n++) {
        const std::string name = "Resume_conc" + std::to_string(n) + ".bin";
        if (tb_conc_use(c, n)) return -1;
        const int rc = load ? tb_load_field(c, TB_F_C, join(dir, name.c_str()).c_str())
                            : tb_save_field(c, TB_F_C, join(dir, name.c_str()).c_str());
        if (rc) {
            tb_conc_use(c, sel);
            return -1;
        }
    }
    return tb_conc_use(c, sel);
}

}  // namespace

extern "C" int tb_petsc_write_array(const char *path, const double *a, int64_t n) {
    if (!path || n < 0 || n > INT32_MAX || (n > 0 && !a)) return TB_IO_BAD_ARGUMENT;
    File out(path, "wb");
    if (!out.f) return TB_IO_OPEN_FAILED;
    const uint32_t head[2] = {be32((uint32_t)VEC_FILE_CLASSID), be32((uint32_t)n)};
    if (fwrite(head, sizeof(uint32_t), 2, out.f) != 2) return TB_IO_SHORT;
    std::vector<uint64_t> buf(std::min<size_t>(CHUNK, (size_t)n));
    for (int64_t i0 = 0; i0 < n; i0 += (int64_t)CHUNK) {
        const size_t m = (size_t)std::min<int64_t>((int64_t)CHUNK, n - i0);
        for (size_t k = 0; k < m; k++) {
            uint64_t bits;
            memcpy(&bits, a + i0 + k, sizeof(bits));
            buf[k] = be64(bits);
        }
        if (fwrite(buf.data(), sizeof(uint64_t), m, out.f) != m) return TB_IO_SHORT;
    }
    if (fflush(out.f) != 0) return TB_IO_SHORT;
    return 0;
}

extern "C" int64_t tb_petsc_read_array(const char *path, double *a, int64_t capacity) {
    if (!path || capacity < 0 || (capacity > 0 && !a)) return TB_IO_BAD_ARGUMENT;
    File in(path, "rb");
    if (!in.f) return TB_IO_OPEN_FAILED;
    uint32_t head[2];
    if (fread(head, sizeof(uint32_t), 2, in.f) != 2) return TB_IO_SHORT;
    if ((int32_t)be32(head[0]) != VEC_FILE_CLASSID) return TB_IO_NOT_A_VEC;
    const int32_t n = (int32_t)be32(head[1]);
    if (n < 0) return TB_IO_NOT_A_VEC;
    if (n > capacity) return TB_IO_SIZE;
    std::vector<uint64_t> buf(std::min<size_t>(CHUNK, (size_t)n));
    for (int64_t i0 = 0; i0 < n; i0 += (int64_t)CHUNK) {
        const size_t m = (size_t)std::min<int64_t>((int64_t)CHUNK, (int64_t)n - i0);
        if (fread(buf.data(), sizeof(uint64_t), m, in.f) != m) return TB_IO_SHORT;
        for (size_t k = 0; k < m; k++) {
            const uint64_t bits = be64(buf[k]);
            memcpy(a + i0 + k, &bits, sizeof(bits));
        }
    }
    return n;
}

extern "C" int tb_write_resume_primary(const char *dir, const tb_resume_info *info) {
    if (!info) return TB_IO_BAD_ARGUMENT;
    const std::string p = join(dir, "Resume_primary.dat");
    File out(p.c_str(), "w");
    if (!out.f) return TB_IO_OPEN_FAILED;
    // the reference's record, its %lf (six decimals) included
    if (fprintf(out.f, "time=\n%lf\ndt=\n%lf\ndt_old=\n%lf\niteration=\n%d\noutput_iter=\n%d\n", info->time, info->dt,
                info->dt_old, info->iter, info->output_iter) < 0)
        return TB_IO_SHORT;
    return 0;
}

extern "C" int tb_read_resume_primary(const char *dir, tb_resume_info *info) {
    if (!info) return TB_IO_BAD_ARGUMENT;
    const std::string p = join(dir, "Resume_primary.dat");
    File in(p.c_str(), "r");
    if (!in.f) return TB_IO_OPEN_FAILED;
    char label[5][64];
    tb_resume_info r;
    memset(&r, 0, sizeof(r));
    const int n = fscanf(in.f, "%63s %lf %63s %lf %63s %lf %63s %d %63s %d", label[0], &r.time, label[1], &r.dt,
                         label[2], &r.dt_old, label[3], &r.iter, label[4], &r.output_iter);
    if (n != 10) return TB_IO_SHORT;
    const char *want[5] = {"time=", "dt=", "dt_old=", "iteration=", "output_iter="};
    for (int k = 0; k < 5; k++)
        if (strcmp(label[k], want[k]) != 0) return TB_IO_NOT_A_VEC;
    *info = r;
    // the record's six decimals turn a step below 5e-7 into 0: say so
    return (r.dt == 0.0 || r.dt_old == 0.0) ? TB_IO_ZERO_STEP : 0;
}

static const char *io_message(int64_t rc) {
    switch (rc) {
    case TB_IO_OPEN_FAILED: return "cannot open the file";
    case TB_IO_NOT_A_VEC: return "not a PETSc binary vector (class id or header)";
    case TB_IO_SIZE: return "vector length differs from the field";
    case TB_IO_SHORT: return "short read or write";
    default: return "bad argument";
    }
}

extern "C" int tb_save_field(tb_ctx *c, int f, const char *path) {
    if (!c || f < 0 || f >= TB_F_COUNT || !path) {
        if (c) c->err = "tb_save_field: bad argument";
        return -1;
    }
    const int64_t n = (int64_t)c->d.NX * c->d.NY * c->d.NZ;
    std::vector<double> host((size_t)n);
    if (tb_download_field(c, f, host.data())) return -1;
    const int rc = tb_petsc_write_array(path, host.data(), n);
    if (rc < 0) {
        c->err = std::string("tb_save_field ") + path + ": " + io_message(rc);
        return -1;
    }
    return 0;
}

extern "C" int tb_load_field(tb_ctx *c, int f, const char *path) {
    if (!c || f < 0 || f >= TB_F_COUNT || !path) {
        if (c) c->err = "tb_load_field: bad argument";
        return -1;
    }
    const int64_t n = (int64_t)c->d.NX * c->d.NY * c->d.NZ;
    std::vector<double> host((size_t)n);
    const int64_t got = tb_petsc_read_array(path, host.data(), n);
    if (got < 0 || got != n) {
        c->err = std::string("tb_load_field ") + path + ": " + io_message(got < 0 ? got : TB_IO_SIZE);
        return -1;
    }
    return tb_upload_field(c, f, host.data());
}

extern "C" int tb_write_resume(tb_ctx *c, const char *dir, const tb_resume_info *info) {
    if (!c || !info) {
        if (c) c->err = "tb_write_resume: bad argument";
        return -1;
    }
    for (const ResumeVec &v : RESUME_VECS)
        if (tb_save_field(c, v.field, join(dir, v.name).c_str())) return -1;
    if (resume_conc(c, dir, false)) return -1;
    const int rc = tb_write_resume_primary(dir, info);
    if (rc < 0) {
        c->err = std::string("tb_write_resume Resume_primary.dat: ") + io_message(rc);
        return -1;
    }
    return 0;
}

extern "C" int tb_read_resume(tb_ctx *c, const char *dir, tb_resume_info *info) {
    if (!c || !info) {
        if (c) c->err = "tb_read_resume: bad argument";
        return -1;
    }
    const int rc = tb_read_resume_primary(dir, info);
    if (rc < 0) {
        c->err = std::string("tb_read_resume Resume_primary.dat: ") + io_message(rc);
        return -1;
    }
    for (const ResumeVec &v : RESUME_VECS)
        if (tb_load_field(c, v.field, join(dir, v.name).c_str())) return -1;
    if (resume_conc(c, dir, true)) return -1;
    // a restart begins a new trajectory: the learned guess forgets the
    // solutions of the old one
    c->n_hist = c->gram_n = c->guess_nr = 0;
    c->phi_is_hist0 = false;
    return rc;   // TB_IO_ZERO_STEP passes through as a warning
}
