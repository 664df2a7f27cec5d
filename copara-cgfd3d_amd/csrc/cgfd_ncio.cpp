// cgfd_ncio.cpp -- NetCDF output files of the time-stepping engine.
//
// The reference writes its seismograms, snapshots and peak-velocity maps
// through netCDF-C (seisnc.cpp:1229-1490, 1671-1712).  This file writes the
// same dimensions, variables and attributes in the netCDF classic format
// with 64-bit offsets (CDF-2), which every netCDF reader opens like the
// reference's files; netCDF-C itself is not needed.
//
// Classic format, as the netCDF specification defines it: big-endian;
//   header = "CDF\2" numrecs dim_list gatt_list var_list
//   dim    = name length (0 for the record dimension)
//   attr   = name nc_type nelems values (padded to 4 bytes)
//   var    = name ndims dimids vatt_list nc_type vsize begin(8 bytes)
// followed by the fixed-size variables at their `begin` offsets and then,
// per record, the slab of every record variable in definition order.
#include "cgfd3d.h"

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

namespace {

enum { NC_CHAR = 2, NC_INT = 4, NC_FLOAT = 5, NC_DOUBLE = 6 };
enum { TAG_DIM = 0x0A, TAG_VAR = 0x0B, TAG_ATT = 0x0C };

struct NcAtt {
    std::string name;
    int type;
    std::vector<unsigned char> bytes; // big-endian values
    int nelems;
};

struct NcVar {
    std::string name;
    std::vector<int> dims;
    int type;
    const void *data; // host values in native byte order, C order
    long count;       // values per record (record variables) or in total
    bool rec;
};

struct NcFile {
    std::vector<std::pair<std::string, long>> dims; // length 0 = record dimension
    std::vector<NcAtt> gatts;
    std::vector<NcVar> vars;
    long numrecs = 0;
};

int type_size(int t) { return t == NC_CHAR ? 1 : (t == NC_DOUBLE ? 8 : 4); }
long pad4(long n) { return (n + 3) & ~3L; }

void put32(std::vector<unsigned char> &b, uint32_t v)
{
    for (int s = 24; s >= 0; s -= 8) b.push_back((unsigned char)(v >> s));
}
void put64(std::vector<unsigned char> &b, uint64_t v)
{
    for (int s = 56; s >= 0; s -= 8) b.push_back((unsigned char)(v >> s));
}
void put_name(std::vector<unsigned char> &b, const std::string &n)
{
    put32(b, (uint32_t)n.size());
    b.insert(b.end(), n.begin(), n.end());
    while (b.size() % 4) b.push_back(0);
}

// values of one type, converted to big-endian
void append_values(std::vector<unsigned char> &b, int type, const void *src, long n)
{
    const unsigned char *p = (const unsigned char *)src;
    const int sz = type_size(type);
    for (long i = 0; i < n; i++)
        for (int k = sz - 1; k >= 0; k--) b.push_back(p[i * sz + k]); // little-endian host
}

NcAtt att_text(const std::string &name, const char *text, long n)
{
    NcAtt a{name, NC_CHAR, {}, (int)n};
    a.bytes.assign(text, text + n);
    return a;
}

NcAtt att_ints(const std::string &name, const int *v, int n)
{
    NcAtt a{name, NC_INT, {}, n};
    append_values(a.bytes, NC_INT, v, n);
    return a;
}

void put_atts(std::vector<unsigned char> &b, const std::vector<NcAtt> &atts)
{
    if (atts.empty()) {
        put32(b, 0);
        put32(b, 0);
        return;
    }
    put32(b, TAG_ATT);
    put32(b, (uint32_t)atts.size());
    for (const NcAtt &a : atts) {
        put_name(b, a.name);
        put32(b, (uint32_t)a.type);
        put32(b, (uint32_t)a.nelems);
        b.insert(b.end(), a.bytes.begin(), a.bytes.end());
        while (b.size() % 4) b.push_back(0);
    }
}

long var_vsize(const NcFile &f, const NcVar &v)
{
    long n = 1;
    for (int d : v.dims)
        if (f.dims[d].second > 0) n *= f.dims[d].second;
    return pad4(n * type_size(v.type));
}

std::vector<unsigned char> header(const NcFile &f, const std::vector<uint64_t> &begin)
{
    std::vector<unsigned char> b = {'C', 'D', 'F', 2};
    put32(b, (uint32_t)f.numrecs);
    if (f.dims.empty()) {
        put32(b, 0);
        put32(b, 0);
    } else {
        put32(b, TAG_DIM);
        put32(b, (uint32_t)f.dims.size());
        for (const auto &d : f.dims) {
            put_name(b, d.first);
            put32(b, (uint32_t)d.second);
        }
    }
    put_atts(b, f.gatts);
    if (f.vars.empty()) {
        put32(b, 0);
        put32(b, 0);
    } else {
        put32(b, TAG_VAR);
        put32(b, (uint32_t)f.vars.size());
        for (size_t i = 0; i < f.vars.size(); i++) {
            const NcVar &v = f.vars[i];
            put_name(b, v.name);
            put32(b, (uint32_t)v.dims.size());
            for (int d : v.dims) put32(b, (uint32_t)d);
            put32(b, 0); // no variable attributes
            put32(b, 0);
            put32(b, (uint32_t)v.type);
            put32(b, (uint32_t)var_vsize(f, v));
            put64(b, begin[i]);
        }
    }
    return b;
}

int write_file(const char *path, const NcFile &f)
{
    std::vector<uint64_t> begin(f.vars.size(), 0);
    const long hsize = (long)header(f, begin).size();
    // fixed-size variables first, then the interleaved record section
    uint64_t off = (uint64_t)hsize;
    for (size_t i = 0; i < f.vars.size(); i++)
        if (!f.vars[i].rec) {
            begin[i] = off;
            off += (uint64_t)var_vsize(f, f.vars[i]);
        }
    long recsize = 0;
    for (size_t i = 0; i < f.vars.size(); i++)
        if (f.vars[i].rec) {
            begin[i] = off + (uint64_t)recsize;
            recsize += var_vsize(f, f.vars[i]);
        }
    FILE *fp = fopen(path, "wb");
    if (!fp) return CGFD_EINVAL;
    std::vector<unsigned char> b = header(f, begin);
    bool ok = fwrite(b.data(), 1, b.size(), fp) == b.size();
    for (const NcVar &v : f.vars)
        if (!v.rec && ok) {
            b.clear();
            append_values(b, v.type, v.data, v.count);
            while ((long)b.size() < var_vsize(f, v)) b.push_back(0);
            ok = fwrite(b.data(), 1, b.size(), fp) == b.size();
        }
    for (long r = 0; r < f.numrecs && ok; r++)
        for (const NcVar &v : f.vars)
            if (v.rec && ok) {
                b.clear();
                append_values(b, v.type, (const char *)v.data + (size_t)r * v.count * type_size(v.type), v.count);
                while ((long)b.size() < var_vsize(f, v)) b.push_back(0);
                ok = fwrite(b.data(), 1, b.size(), fp) == b.size();
            }
    ok = (fclose(fp) == 0) && ok;
    return ok ? CGFD_OK : CGFD_EINVAL;
}

const char *const kWaveNames[9] = {"Vx", "Vy", "Vz", "Txx", "Tyy", "Tzz", "Txy", "Txz", "Tyz"};

} // namespace

extern "C" {

int cgfd_nc_seismo(const char *path, int precision, int nrec, int npnt, const void *const *wave,
                   const void *currtime, const int *pnum, const int *lnum, const void *posx, const void *posy,
                   const void *posz)
{
    if (!path || !wave || !currtime || nrec < 0 || npnt < 1 || (precision != CGFD_FP32 && precision != CGFD_FP64))
        return CGFD_EINVAL;
    const int T = precision == CGFD_FP64 ? NC_DOUBLE : NC_FLOAT;
    NcFile f;
    f.numrecs = nrec;
    f.dims = {{"TimeNum", 0}, {"PointNum", npnt}};
    for (int c = 0; c < 9; c++) {
        if (!wave[c]) return CGFD_EINVAL;
        f.vars.push_back({kWaveNames[c], {0, 1}, T, wave[c], npnt, true});
    }
    f.vars.push_back({"currTime", {0}, T, currtime, 1, true});
    const void *pp[5] = {pnum, lnum, posx, posy, posz};
    const char *pn[5] = {"Pnum", "Lnum", "posx", "posy", "posz"};
    for (int k = 0; k < 5; k++) {
        if (!pp[k]) return CGFD_EINVAL;
        f.vars.push_back({pn[k], {1}, k < 2 ? NC_INT : T, pp[k], npnt, false});
    }
    static const char content[] = "Velocity and Stress seismogram";
    f.gatts.push_back(att_text("Content", content, sizeof(content))); // sizeof: with its NUL, as seisnc.cpp:1263
    return write_file(path, f);
}

int cgfd_nc_snap(const char *path, int precision, const int *spec, int nrec, const void *const *wave,
                 const void *currtime)
{
    if (!path || !spec || !wave || !currtime || nrec < 0 || (precision != CGFD_FP32 && precision != CGFD_FP64))
        return CGFD_EINVAL;
    const int cmp = spec[10];
    if (cmp < 1 || cmp > 3 || spec[3] < 1 || spec[4] < 1 || spec[5] < 1) return CGFD_EINVAL;
    const int T = precision == CGFD_FP64 ? NC_DOUBLE : NC_FLOAT;
    NcFile f;
    f.numrecs = nrec;
    f.dims = {{"TimeNum", 0}, {"i", spec[3]}, {"j", spec[4]}, {"k", spec[5]}};
    const long npts = (long)spec[3] * spec[4] * spec[5];
    for (int c = 0; c < 9; c++) {
        const bool on = (c < 3 && (cmp == 1 || cmp == 3)) || (c >= 3 && (cmp == 2 || cmp == 3));
        if (!on) continue;
        if (!wave[c]) return CGFD_EINVAL;
        f.vars.push_back({kWaveNames[c], {0, 1, 2, 3}, T, wave[c], npts, true});
    }
    f.vars.push_back({"currTime", {0}, T, currtime, 1, true});
    f.gatts.push_back(att_ints("IdxStart", spec + 0, 3));
    f.gatts.push_back(att_ints("IdxNumber", spec + 3, 3));
    f.gatts.push_back(att_ints("IdxInterval", spec + 6, 3));
    f.gatts.push_back(att_ints("TimeInterval", spec + 9, 1));
    static const char s1[] = "Velocity field", s2[] = "Stress field", s3[] = "Velocity and Stress field";
    if (cmp == 1)
        f.gatts.push_back(att_text("Content", s1, sizeof(s1)));
    else if (cmp == 2)
        f.gatts.push_back(att_text("Content", s2, sizeof(s2)));
    else
        f.gatts.push_back(att_text("Content", s3, sizeof(s3)));
    return write_file(path, f);
}

int cgfd_nc_peakvel(const char *path, int precision, int ni, int nj, const void *vx, const void *vy, const void *vz)
{
    if (!path || !vx || !vy || !vz || ni < 1 || nj < 1 || (precision != CGFD_FP32 && precision != CGFD_FP64))
        return CGFD_EINVAL;
    const int T = precision == CGFD_FP64 ? NC_DOUBLE : NC_FLOAT;
    NcFile f;
    f.dims = {{"i", ni}, {"j", nj}};
    const long n = (long)ni * nj;
    f.vars.push_back({"Vx", {0, 1}, T, vx, n, false});
    f.vars.push_back({"Vy", {0, 1}, T, vy, n, false});
    f.vars.push_back({"Vz", {0, 1}, T, vz, n, false});
    return write_file(path, f);
}

} // extern "C"
