// Restart files of the reference (print_restart main.f90:301-332, read by
// build_grid grid.f90:56-95): Fortran unformatted sequential files holding
// one record each, framed by 4-byte little-endian byte counts.
//   Trestart:        xg(-3:nx+3), zg(-3:nz+3), [T], [Cnew(1:ntype,...)], SF
//   restart_tracers: rtr(1:ntr), str(1:ntr), Ttype(1:ntr) (int32)
// Grid fields are the ghosted (-2:nx+2,-2:nz+2) arrays, i fastest.  Host
// code only: the codec serves initial conditions, checkpoints and fixtures.
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>

#include "tc_internal.h"

namespace {

struct Part {
    void* p;
    size_t bytes;
};

int read_record(const char* path, const std::vector<Part>& parts)
{
    size_t want = 0;
    for (const Part& q : parts) want += q.bytes;
    if (want > 0x7fffffffu) return TC_ERR_UNSUPPORTED;  // a single record only (no sub-records)
    FILE* f = std::fopen(path, "rb");
    if (!f) return TC_ERR_ARG;
    int32_t head = 0, tail = 0;
    int st = 0;
    if (std::fread(&head, 4, 1, f) != 1 || head < 0 || (size_t)head != want) st = TC_ERR_ARG;
    for (size_t j = 0; st == 0 && j < parts.size(); ++j)
        if (parts[j].bytes && std::fread(parts[j].p, 1, parts[j].bytes, f) != parts[j].bytes) st = TC_ERR_ARG;
    if (st == 0 && (std::fread(&tail, 4, 1, f) != 1 || tail != head)) st = TC_ERR_ARG;
    std::fclose(f);
    return st;
}

int write_record(const char* path, const std::vector<Part>& parts)
{
    size_t total = 0;
    for (const Part& q : parts) total += q.bytes;
    if (total > 0x7fffffffu) return TC_ERR_UNSUPPORTED;
    FILE* f = std::fopen(path, "wb");
    if (!f) return TC_ERR_ARG;
    const int32_t n = (int32_t)total;
    int st = 0;
    if (std::fwrite(&n, 4, 1, f) != 1) st = TC_ERR_ARG;
    for (size_t j = 0; st == 0 && j < parts.size(); ++j)
        if (parts[j].bytes && std::fwrite(parts[j].p, 1, parts[j].bytes, f) != parts[j].bytes) st = TC_ERR_ARG;
    if (st == 0 && std::fwrite(&n, 4, 1, f) != 1) st = TC_ERR_ARG;
    if (std::fclose(f) != 0) st = TC_ERR_ARG;
    return st;
}

std::vector<Part> restart_parts(int32_t nx, int32_t nz, int32_t has_T, int32_t ntype_c, const double* xg,
                                const double* zg, const double* T, const double* Cnew, const double* SF)
{
    const size_t plane = (size_t)(nx + 2 * tc::SPAN + 1) * (size_t)(nz + 2 * tc::SPAN + 1) * sizeof(double);
    const int ext = tc::SPAN + 1;  // span + span_interp
    std::vector<Part> v;
    v.push_back({const_cast<double*>(xg), (size_t)(nx + 2 * ext + 1) * sizeof(double)});
    v.push_back({const_cast<double*>(zg), (size_t)(nz + 2 * ext + 1) * sizeof(double)});
    if (has_T) v.push_back({const_cast<double*>(T), plane});
    if (ntype_c > 0) v.push_back({const_cast<double*>(Cnew), plane * (size_t)ntype_c});
    v.push_back({const_cast<double*>(SF), plane});
    return v;
}

bool restart_args_ok(int32_t nx, int32_t nz, int32_t has_T, int32_t ntype_c, const void* xg, const void* zg,
                     const void* T, const void* Cnew, const void* SF)
{
    return nx > 0 && nz > 0 && ntype_c >= 0 && ntype_c <= TC_MAX_TYPES && xg && zg && SF && (!has_T || T) &&
           (ntype_c == 0 || Cnew);
}

}  // namespace

extern "C" {

int tc_restart_read(const char* path, int32_t nx, int32_t nz, int32_t has_T, int32_t ntype_c, double* xg,
                    double* zg, double* T, double* Cnew, double* SF)
{
    if (!path || !restart_args_ok(nx, nz, has_T, ntype_c, xg, zg, T, Cnew, SF)) return TC_ERR_ARG;
    return read_record(path, restart_parts(nx, nz, has_T, ntype_c, xg, zg, T, Cnew, SF));
}

int tc_restart_write(const char* path, int32_t nx, int32_t nz, int32_t has_T, int32_t ntype_c, const double* xg,
                     const double* zg, const double* T, const double* Cnew, const double* SF)
{
    if (!path || !restart_args_ok(nx, nz, has_T, ntype_c, xg, zg, T, Cnew, SF)) return TC_ERR_ARG;
    return write_record(path, restart_parts(nx, nz, has_T, ntype_c, xg, zg, T, Cnew, SF));
}

int tc_restart_tracers_read(const char* path, int64_t ntr, double* rtr, double* str, int32_t* ttype)
{
    if (!path || ntr < 0 || (ntr > 0 && (!rtr || !str || !ttype))) return TC_ERR_ARG;
    const size_t n = (size_t)ntr;
    return read_record(path, {{rtr, n * sizeof(double)}, {str, n * sizeof(double)}, {ttype, n * sizeof(int32_t)}});
}

int tc_restart_tracers_write(const char* path, int64_t ntr, const double* rtr, const double* str,
                             const int32_t* ttype)
{
    if (!path || ntr < 0 || (ntr > 0 && (!rtr || !str || !ttype))) return TC_ERR_ARG;
    const size_t n = (size_t)ntr;
    return write_record(path, {{const_cast<double*>(rtr), n * sizeof(double)},
                               {const_cast<double*>(str), n * sizeof(double)},
                               {const_cast<int32_t*>(ttype), n * sizeof(int32_t)}});
}

}  // extern "C"
