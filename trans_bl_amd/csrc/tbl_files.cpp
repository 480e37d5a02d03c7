// The reference's on-disk field files, as Fortran sequential unformatted
// records (a 4-byte length before and after each record's bytes):
//   data_RR.bin  checkpoint (escreve, escreve_par.f:6-33): record 1 the time
//                step t (integer*4), record 2 ux, uy, uz, wx, wy, wz (and th
//                for my_form 2), each complex*16 (ptsx, jmax, kfour);
//   based_RR.bin base flow (baseflow2D/main_par_2D.f:578-595, read by
//                main_par.f:847-850): one record ux, uy, wz, each real*8
//                (ptsx, jmax).
// A record longer than 2^31 - 1 bytes would need the compiler's sub-record
// convention; such files are refused.
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/trans_bl_amd.h"

void tbl_set_last_error(const char *msg, const char *file, int line);

namespace {

int fail(const std::string &msg, int rc = TBL_ERR_ARG)
{
    tbl_set_last_error(msg.c_str(), __FILE__, __LINE__);
    return rc;
}

struct File {
    std::FILE *f = nullptr;
    ~File() { if (f) std::fclose(f); }
};

// one record of `nparts` consecutive arrays of `bytes` each
int write_record(std::FILE *f, const void *const *parts, int nparts, size_t bytes)
{
    const size_t total = bytes * nparts;
    if (total > 0x7fffffffu) return fail("record above 2^31 - 1 bytes");
    const int32_t n = static_cast<int32_t>(total);
    if (std::fwrite(&n, 4, 1, f) != 1) return fail("write failed");
    for (int p = 0; p < nparts; ++p)
        if (bytes && std::fwrite(parts[p], 1, bytes, f) != bytes) return fail("write failed");
    if (std::fwrite(&n, 4, 1, f) != 1) return fail("write failed");
    return TBL_OK;
}

int read_record(std::FILE *f, void *const *parts, int nparts, size_t bytes, const char *what)
{
    int32_t n1 = 0, n2 = 0;
    if (std::fread(&n1, 4, 1, f) != 1) return fail(std::string("missing record: ") + what);
    if (n1 < 0 || static_cast<size_t>(n1) != bytes * nparts)
        return fail(std::string("record length does not match the grid: ") + what);
    for (int p = 0; p < nparts; ++p)
        if (bytes && std::fread(parts[p], 1, bytes, f) != bytes)
            return fail(std::string("truncated record: ") + what);
    if (std::fread(&n2, 4, 1, f) != 1 || n2 != n1)
        return fail(std::string("bad trailing marker: ") + what);
    return TBL_OK;
}

}  // namespace

extern "C" int tbl_checkpoint_write(const char *path, int nx, int ny, int nmodes, int t,
                                    int nfields, const double *const *fields)
{
    if (!path || !fields || nx < 1 || ny < 1 || nmodes < 1 || (nfields != 6 && nfields != 7))
        return fail("checkpoint: null pointer, empty grid or field count not 6 or 7");
    for (int p = 0; p < nfields; ++p)
        if (!fields[p]) return fail("checkpoint: null field pointer");
    File F;
    F.f = std::fopen(path, "wb");
    if (!F.f) return fail("cannot create the checkpoint file");
    const int32_t ti = t;
    const void *tp = &ti;
    int rc = write_record(F.f, &tp, 1, 4);
    if (rc != TBL_OK) return rc;
    const size_t bytes = sizeof(double) * 2 * static_cast<size_t>(nx) * ny * nmodes;
    return write_record(F.f, reinterpret_cast<const void *const *>(fields), nfields, bytes);
}

extern "C" int tbl_checkpoint_read(const char *path, int nx, int ny, int nmodes, int *t,
                                   int nfields, double *const *fields)
{
    if (!path || !fields || !t || nx < 1 || ny < 1 || nmodes < 1 ||
        (nfields != 6 && nfields != 7))
        return fail("checkpoint: null pointer, empty grid or field count not 6 or 7");
    for (int p = 0; p < nfields; ++p)
        if (!fields[p]) return fail("checkpoint: null field pointer");
    File F;
    F.f = std::fopen(path, "rb");
    if (!F.f) return fail("cannot open the checkpoint file");
    int32_t ti = 0;
    void *tp = &ti;
    int rc = read_record(F.f, &tp, 1, 4, "time step");
    if (rc != TBL_OK) return rc;
    *t = ti;
    const size_t bytes = sizeof(double) * 2 * static_cast<size_t>(nx) * ny * nmodes;
    return read_record(F.f, reinterpret_cast<void *const *>(fields), nfields, bytes, "fields");
}

extern "C" int tbl_baseflow_write(const char *path, int nx, int ny, const double *ux,
                                  const double *uy, const double *wz)
{
    if (!path || !ux || !uy || !wz || nx < 1 || ny < 1)
        return fail("base flow: null pointer or empty grid");
    File F;
    F.f = std::fopen(path, "wb");
    if (!F.f) return fail("cannot create the base-flow file");
    const void *parts[3] = {ux, uy, wz};
    return write_record(F.f, parts, 3, sizeof(double) * static_cast<size_t>(nx) * ny);
}

extern "C" int tbl_baseflow_read(const char *path, int nx, int ny, double *ux, double *uy,
                                 double *wz)
{
    if (!path || !ux || !uy || !wz || nx < 1 || ny < 1)
        return fail("base flow: null pointer or empty grid");
    File F;
    F.f = std::fopen(path, "rb");
    if (!F.f) return fail("cannot open the base-flow file");
    void *parts[3] = {ux, uy, wz};
    return read_record(F.f, parts, 3, sizeof(double) * static_cast<size_t>(nx) * ny, "base flow");
}
