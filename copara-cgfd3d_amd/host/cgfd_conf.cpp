// cgfd_conf.cpp -- the C++ parameter-file front end (see cgfd_conf.h).
//
// Supported inputs and the reference paths they follow:
//   grid_type vmap, hybrid_grid 0 and 1       gridmesh.cpp:72-532, 1092-1210, 1446-1470
//   metrics (2nd-order central differences)   gridmesh.cpp:1476-1627
//   absorb file and ADE CFS-PML tables        absorb.cpp:20-46, 346-800
//   force / moment anchors, 8 STF types       source.cpp:12-155, 671-849, 985-1289
//   focal-process file (netCDF classic or 4)  source.cpp:43-58, 787-848, 1185-1262
//   station file                              seisplot.cpp:12-55, 330-457
//   volume media (netCDF classic or 4)        mediaparas.cpp:376-441, 1273-1425, 2140-2219
//   interface3D media                         mediaparas.cpp:34-43, 52-374, 442-914
//   sampling_point_per_cell > 1 smoothing     mediaparas.cpp:1712-1905
// Refused, with the reason in the error: grid_type point (read_point stores
// x, y and z all into crd.x, gridmesh.cpp:52-68) and media_type interface
// (readdata overwrites it with zero media via get2in1, mediaparas.cpp:2163-2200).
// netCDF-4 (HDF5) files are read through libcgfd_nc4.so, loaded at run time
// from CGFD_NC4_LIB or beside the executable.
#include "cgfd_conf.h"

#include <dlfcn.h>
#include <unistd.h>

#include <algorithm>
#include <array>
#include <cerrno>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <functional>
#include <map>
#include <memory>
#include <sstream>

namespace cgfd_conf {
namespace {

constexpr int LFD = 3;
constexpr float PI_F = 3.141592653589793f;  // constant::PI is a Real
constexpr float SEIS_ZERO = 1e-20f;
constexpr float SEIS_EQUAL = 1e-5f;
constexpr float SEIS_INF = 1e20f;
constexpr size_t STRLEN = 128;              // the fgets buffer of get_conf

[[noreturn]] void fail(const std::string &msg) { throw ConfError(msg); }

// --------------------------------------------------------------------------
// 3-D arrays in the reference layout
// --------------------------------------------------------------------------
struct Grid3 {
    int nx = 0, ny = 0, nz = 0;
    std::vector<float> v;
    Grid3() = default;
    Grid3(int a, int b, int c, float fill = 0.0f) : nx(a), ny(b), nz(c), v(size_t(a) * b * c, fill) {}
    float &operator()(int i, int j, int k) { return v[(size_t(i) * ny + j) * nz + k]; }
    float operator()(int i, int j, int k) const { return v[(size_t(i) * ny + j) * nz + k]; }
};

// --------------------------------------------------------------------------
// get_conf (common.cpp:23-66, 286-322)
// --------------------------------------------------------------------------
bool is_space(char c) { return c == ' ' || c == '\t' || c == '\n' || c == '\r' || c == '\v' || c == '\f'; }

std::vector<std::string> split_ws(const std::string &s) {
    std::vector<std::string> out;
    size_t i = 0;
    while (i < s.size()) {
        while (i < s.size() && is_space(s[i])) ++i;
        size_t j = i;
        while (j < s.size() && !is_space(s[j])) ++j;
        if (j > i) out.push_back(s.substr(i, j - i));
        i = j;
    }
    return out;
}

// fscanf("%f") over a token stream: strtof of each token.
class Scanner {
  public:
    Scanner(std::vector<std::string> toks, std::string where) : toks_(std::move(toks)), where_(std::move(where)) {}
    float real() {
        if (pos_ >= toks_.size()) fail(where_ + ": fewer values than expected");
        const char *s = toks_[pos_].c_str();
        char *end = nullptr;
        float v = std::strtof(s, &end);
        if (end == s) fail(where_ + ": '" + toks_[pos_] + "' is not a number");
        ++pos_;
        return v;
    }
    std::vector<float> reals(size_t n) {
        std::vector<float> v(n);
        for (auto &x : v) x = real();
        return v;
    }

  private:
    std::vector<std::string> toks_;
    std::string where_;
    size_t pos_ = 0;
};

// One parameter file, searched as get_conf searches it: the file is read in
// fgets pieces of at most 127 bytes, pieces starting with '#' are skipped,
// and the first piece whose first token is the key and that holds at least
// `col` tokens gives the col-th token (1-based).
class ConfFile {
  public:
    ConfFile(const std::string &path, const std::string &what) : path_(path) {
        std::ifstream f(path, std::ios::binary);
        if (!f) fail("fail to open " + what + " file " + path + ": " + std::strerror(errno));
        std::stringstream ss;
        ss << f.rdbuf();
        const std::string raw = ss.str();
        size_t start = 0;
        while (true) {
            size_t nl = raw.find('\n', start);
            std::string line = raw.substr(start, nl == std::string::npos ? std::string::npos : nl - start + 1);
            for (size_t s = 0; s < line.size(); s += STRLEN - 1) pieces_.push_back(line.substr(s, STRLEN - 1));
            if (nl == std::string::npos) break;
            start = nl + 1;
        }
    }
    const std::vector<std::string> &pieces() const { return pieces_; }
    std::string str(const std::string &key, int col = 3) const { return find(key, col).first; }
    int integer(const std::string &key, int col = 3) const { return std::atoi(find(key, col).first.c_str()); }
    float real(const std::string &key, int col = 3) const { return (float)std::atof(find(key, col).first.c_str()); }
    bool flag(const std::string &key, int col = 3) const { return find(key, col).first != "0"; }
    Scanner chunk(const std::string &anchor) const {
        const size_t n = find(anchor, 1).second;
        std::vector<std::string> toks;
        for (size_t p = n + 1; p < pieces_.size(); ++p)
            for (auto &t : split_ws(pieces_[p])) toks.push_back(t);
        return Scanner(std::move(toks), path_ + " after " + anchor);
    }

  private:
    std::pair<std::string, size_t> find(const std::string &key, int col) const {
        for (size_t n = 0; n < pieces_.size(); ++n) {
            if (!pieces_[n].empty() && pieces_[n][0] == '#') continue;
            auto toks = split_ws(pieces_[n]);
            if (!toks.empty() && toks[0] == key && (int)toks.size() >= col) return {toks[col - 1], n};
        }
        fail("there is no value found at " + std::to_string(col) + " colomn for " + key + " in " + path_);
    }
    std::string path_;
    std::vector<std::string> pieces_;
};

// --------------------------------------------------------------------------
// netCDF inputs: classic (CDF-1, CDF-2) read here, netCDF-4 through
// libcgfd_nc4.so
// --------------------------------------------------------------------------
class NcFile {
  public:
    virtual ~NcFile() = default;
    virtual long long dim(const std::string &name) = 0;
    virtual std::vector<float> var(const std::string &name) = 0;
    virtual float att(const std::string &name) = 0;
};

class NcClassic : public NcFile {
  public:
    NcClassic(std::vector<unsigned char> buf, std::string path, std::string what)
        : b_(std::move(buf)), path_(std::move(path)), what_(std::move(what)) {
        if (b_.size() < 4 || b_[3] < 1 || b_[3] > 2) bad("is not a netCDF classic (CDF-1 or CDF-2) file");
        off64_ = b_[3] == 2;
        p_ = 4;
        numrecs_ = u32();
        // dimensions
        uint32_t tag = u32(), n = u32();
        if (tag == 0x0A)
            for (uint32_t i = 0; i < n; ++i) {
                std::string nm = name();
                uint32_t len = u32();
                if (len == 0) unlim_ = (int)dims_.size();
                dims_.push_back({nm, len});
            }
        else if (tag || n) bad("has a malformed dimension list");
        // global attributes
        tag = u32();
        n = u32();
        if (tag == 0x0C)
            for (uint32_t i = 0; i < n; ++i) {
                std::string nm = name();
                int type = (int)u32();
                uint32_t cnt = u32();
                std::vector<double> vals = values(type, cnt);
                gatts_[nm] = vals;
            }
        else if (tag || n) bad("has a malformed attribute list");
        // variables
        tag = u32();
        n = u32();
        if (tag == 0x0B)
            for (uint32_t i = 0; i < n; ++i) {
                Var v;
                v.name = name();
                uint32_t nd = u32();
                for (uint32_t d = 0; d < nd; ++d) v.dimids.push_back((int)u32());
                uint32_t at = u32(), na = u32();
                if (at == 0x0C)
                    for (uint32_t a = 0; a < na; ++a) {
                        name();
                        int type = (int)u32();
                        values(type, u32());
                    }
                else if (at || na) bad("has a malformed variable attribute list");
                v.type = (int)u32();
                v.vsize = u32();
                v.begin = off64_ ? u64() : u32();
                v.rec = !v.dimids.empty() && v.dimids[0] == unlim_;
                if (v.rec) recsize_ += v.vsize;
                vars_.push_back(v);
            }
        else if (tag || n) bad("has a malformed variable list");
    }

    long long dim(const std::string &nm) override {
        for (auto &d : dims_)
            if (d.first == nm) return d.second == 0 ? numrecs_ : d.second;
        bad("has no dimension " + nm);
    }

    std::vector<float> var(const std::string &nm) override {
        const Var *v = nullptr;
        for (auto &x : vars_)
            if (x.name == nm) v = &x;
        if (!v) bad("has no variable " + nm);
        size_t per = 1;
        for (size_t d = v->rec ? 1 : 0; d < v->dimids.size(); ++d) per *= dims_.at(v->dimids[d]).second;
        const size_t nrec = v->rec ? numrecs_ : 1;
        std::vector<float> out;
        out.reserve(per * nrec);
        const int es = esize(v->type);
        for (size_t r = 0; r < nrec; ++r) {
            size_t at = size_t(v->begin) + r * (v->rec ? recsize_ : 0);
            if (at + per * es > b_.size()) bad("is truncated in variable " + nm);
            for (size_t e = 0; e < per; ++e) out.push_back((float)value(v->type, at + e * es));
        }
        return out;
    }

    float att(const std::string &nm) override {
        auto it = gatts_.find(nm);
        if (it == gatts_.end() || it->second.empty()) bad("has no global attribute " + nm);
        return (float)it->second[0];
    }

  private:
    struct Var {
        std::string name;
        std::vector<int> dimids;
        int type = 0;
        uint64_t vsize = 0, begin = 0;
        bool rec = false;
    };
    [[noreturn]] void bad(const std::string &m) const { fail(what_ + " " + path_ + " " + m); }
    void need(size_t n) const {
        if (p_ + n > b_.size()) bad("is truncated");
    }
    uint32_t u32() {
        need(4);
        uint32_t v = (uint32_t(b_[p_]) << 24) | (uint32_t(b_[p_ + 1]) << 16) | (uint32_t(b_[p_ + 2]) << 8) | b_[p_ + 3];
        p_ += 4;
        return v;
    }
    uint64_t u64() {
        uint64_t hi = u32();
        return (hi << 32) | u32();
    }
    std::string name() {
        uint32_t n = u32();
        need(n);
        std::string s(reinterpret_cast<const char *>(&b_[p_]), n);
        p_ += (n + 3) / 4 * 4;
        return s;
    }
    static int esize(int type) {
        switch (type) {
        case 1: case 2: return 1;
        case 3: return 2;
        case 4: case 5: return 4;
        case 6: return 8;
        default: return 0;
        }
    }
    double value(int type, size_t at) const {
        uint64_t u = 0;
        const int es = esize(type);
        for (int i = 0; i < es; ++i) u = (u << 8) | b_[at + i];
        switch (type) {
        case 1: return (double)(int8_t)u;
        case 2: return (double)(uint8_t)u;
        case 3: return (double)(int16_t)u;
        case 4: return (double)(int32_t)u;
        case 5: { uint32_t w = (uint32_t)u; float f; std::memcpy(&f, &w, 4); return (double)f; }
        case 6: { double d; std::memcpy(&d, &u, 8); return d; }
        default: bad("has a variable of an unknown type");
        }
    }
    std::vector<double> values(int type, uint32_t cnt) {
        const int es = esize(type);
        if (!es) bad("has an attribute of an unknown type");
        need(size_t(cnt) * es);
        std::vector<double> out;
        if (type != 2)
            for (uint32_t i = 0; i < cnt; ++i) out.push_back(value(type, p_ + size_t(i) * es));
        p_ += (size_t(cnt) * es + 3) / 4 * 4;
        return out;
    }

    std::vector<unsigned char> b_;
    std::string path_, what_;
    bool off64_ = false;
    size_t p_ = 0;
    uint32_t numrecs_ = 0;
    int unlim_ = -1;
    uint64_t recsize_ = 0;
    std::vector<std::pair<std::string, uint32_t>> dims_;
    std::map<std::string, std::vector<double>> gatts_;
    std::vector<Var> vars_;
};

struct Nc4Lib {
    int (*probe)(const char *) = nullptr;
    int (*var_dims)(const char *, const char *, int *, long long *) = nullptr;
    int (*get_var_float)(const char *, const char *, float *, long long) = nullptr;
    int (*get_att_float)(const char *, const char *, float *) = nullptr;
};

const Nc4Lib *nc4_library() {
    static Nc4Lib lib;
    static bool tried = false;
    if (!tried) {
        tried = true;
        std::string path;
        if (const char *e = std::getenv("CGFD_NC4_LIB")) {
            path = e;
        } else {
            char exe[4096];
            ssize_t n = readlink("/proc/self/exe", exe, sizeof exe - 1);
            if (n > 0) {
                exe[n] = 0;
                std::string d(exe);
                path = d.substr(0, d.rfind('/') + 1) + "libcgfd_nc4.so";
            }
        }
        void *h = path.empty() ? nullptr : dlopen(path.c_str(), RTLD_NOW | RTLD_LOCAL);
        if (h) {
            lib.probe = (int (*)(const char *))dlsym(h, "cgfd_nc4_probe");
            lib.var_dims = (int (*)(const char *, const char *, int *, long long *))dlsym(h, "cgfd_nc4_var_dims");
            lib.get_var_float = (int (*)(const char *, const char *, float *, long long))dlsym(h, "cgfd_nc4_get_var_float");
            lib.get_att_float = (int (*)(const char *, const char *, float *))dlsym(h, "cgfd_nc4_get_att_float");
        }
    }
    return lib.probe && lib.var_dims && lib.get_var_float && lib.get_att_float ? &lib : nullptr;
}

class Nc4 : public NcFile {
  public:
    Nc4(const Nc4Lib *lib, std::string path, std::string what) : lib_(lib), path_(std::move(path)), what_(std::move(what)) {}
    long long dim(const std::string &nm) override {
        auto s = shape(nm);
        return s.empty() ? 1 : s[0];
    }
    std::vector<float> var(const std::string &nm) override {
        auto s = shape(nm);
        long long n = 1;
        for (auto e : s) n *= e;
        std::vector<float> out((size_t)n);
        if (lib_->get_var_float(path_.c_str(), nm.c_str(), out.data(), n) != 0)
            fail("fail to read variable " + nm + " of " + what_);
        return out;
    }
    float att(const std::string &nm) override {
        float v = 0;
        if (lib_->get_att_float(path_.c_str(), nm.c_str(), &v) != 0) fail(what_ + " has no global attribute " + nm);
        return v;
    }

  private:
    std::vector<long long> shape(const std::string &nm) {
        int r = 0;
        long long d[8];
        if (lib_->var_dims(path_.c_str(), nm.c_str(), &r, d) != 0)
            fail(what_ + " " + path_ + " has no variable or dimension " + nm);
        return std::vector<long long>(d, d + r);
    }
    const Nc4Lib *lib_;
    std::string path_, what_;
};

std::unique_ptr<NcFile> open_netcdf(const std::string &path, const std::string &what) {
    std::ifstream f(path, std::ios::binary);
    if (!f) fail("fail to open " + what + " " + path + ": " + std::strerror(errno));
    std::vector<unsigned char> buf((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
    if (buf.size() >= 3 && buf[0] == 'C' && buf[1] == 'D' && buf[2] == 'F')
        return std::make_unique<NcClassic>(std::move(buf), path, what);
    static const unsigned char hdf[8] = {0x89, 'H', 'D', 'F', '\r', '\n', 0x1a, '\n'};
    if (buf.size() >= 8 && std::memcmp(buf.data(), hdf, 8) == 0) {
        const Nc4Lib *lib = nc4_library();
        if (!lib)
            fail(what + " " + path + " is a netCDF-4 (HDF5) file, and the netCDF-4 reader (libcgfd_nc4.so, built "
                 "against libhdf5) is not available; convert it to the classic format");
        if (lib->probe(path.c_str()) != 1) fail(what + " " + path + " has the HDF5 signature but does not open as an HDF5 file");
        return std::make_unique<Nc4>(lib, path, what);
    }
    fail(what + " " + path + " is neither a netCDF classic nor a netCDF-4 file");
}

// --------------------------------------------------------------------------
// main parameter file
// --------------------------------------------------------------------------
struct Main {
    int ni, nj, nk, nt;
    float stept, steph;
    std::string seispath, grid_type, media_type;
    int hybrid;
    float conv_depth;
    int sampling[3];
    float crit_value, crit_perc;
    int output_peakvel;
    std::string grid_file, media_file, source_file, station_file, absorb_file;
};

Main read_main(const std::string &conf, const std::string &seispath) {
    ConfFile c(conf, "main par");
    Main m;
    m.ni = c.integer("ni");
    m.nj = c.integer("nj");
    m.nk = c.integer("nk");
    m.nt = c.integer("nt");
    m.stept = c.real("stept");
    m.steph = c.real("steph");
    m.seispath = seispath.empty() ? c.str("seispath") : seispath;
    m.grid_type = c.str("grid_type");
    m.hybrid = c.integer("hybrid_grid");
    m.conv_depth = c.real("conversion_interface_depth");
    m.media_type = c.str("media_type");
    for (int i = 0; i < 3; ++i) m.sampling[i] = c.integer("sampling_point_per_cell", 3 + i);
    m.crit_value = c.real("threshold_velocity_jump_of_interface");
    m.crit_perc = c.real("threshold_percent_change_of_interface");
    m.output_peakvel = c.integer("output_peakvel");
    m.grid_file = m.seispath + "/" + c.str("grid_filename");   // sprintf("%s/%s", parpath, name)
    m.media_file = m.seispath + "/" + c.str("media_filename");
    m.source_file = m.seispath + "/" + c.str("source_filename");
    m.station_file = m.seispath + "/" + c.str("station_filename");
    m.absorb_file = m.seispath + "/" + c.str("absorb_filename");
    if (m.ni <= 0 || m.nj <= 0 || m.nk <= 0 || m.nt <= 0) fail(conf + ": ni, nj, nk and nt must be positive");
    return m;
}

// --------------------------------------------------------------------------
// vmap grid (gridmesh.cpp:72-532, 1092-1210)
// --------------------------------------------------------------------------
void set_twoside_space(float l, int na, int nb, float ha1, float ha2, float hb1, float hb2, float &ha, float &hb) {
    const float l0 = (l - (float)na * ha1) - (float)nb * hb1;
    const float la = (float)na * (ha2 - ha1);
    const float lb = (float)nb * (hb2 - hb1);
    if (std::fabs(l0) < SEIS_EQUAL) { ha = ha1; hb = hb1; return; }
    if (std::fabs(la) < SEIS_EQUAL) { ha = ha1; hb = hb1 + l0 / (float)nb; return; }
    if (std::fabs(lb) < SEIS_EQUAL) { ha = ha1 + l0 / (float)na; hb = hb1; return; }
    ha = ha1 + ((l0 * la) / (la + lb)) / (float)na;
    hb = hb1 + ((l0 * lb) / (la + lb)) / (float)nb;
}

// n+1 depths from zb (bottom) to about za whose spacing changes from hb at
// the bottom to ha at the top.
std::vector<float> smooth_zdepth(float za, float zb, float ha, float hb, int n) {
    const float l = za - zb;
    const float h0 = l / (float)n;
    std::vector<float> h(n + 2, 0.0f);
    h[0] = hb;
    h[n + 1] = ha;
    if ((h0 >= ha && h0 <= hb) || (h0 <= ha && h0 >= hb)) {
        int nb = 1, na = n, n0 = n;
        bool fitted = false;
        for (int it = 0; it < 4 * n + 8; ++it) {
            float l0 = 0.0f;
            const float f = (float)(std::pow((double)(ha / hb), 1.0 / (n0 + 1)) - 1.0);
            for (int k = 1; k < nb; ++k) h[k] = hb;
            for (int k = na + 1; k < n + 2; ++k) h[k] = ha;
            for (int k = nb; k < na + 1; ++k) {
                h[k] = h[k - 1] * (1.0f + f);
                l0 = l0 + h[k];
            }
            const float ha1 = (float)((double)ha * std::min(1.0, 1.0 - (double)f));
            const float ha2 = (float)((double)ha * std::max(1.0, 1.0 - (double)f));
            const float hb1 = (float)((double)hb * std::min(1.0, 1.0 + (double)f));
            const float hb2 = (float)((double)hb * std::max(1.0, 1.0 + (double)f));
            const float la1 = (float)(n - na) * ha1, la2 = (float)(n - na) * ha2;
            const float lb1 = (float)(nb - 1) * hb1, lb2 = (float)(nb - 1) * hb2;
            const float rest = l - l0;
            if (la1 + lb1 <= rest && rest <= la2 + lb2) {
                float ha0, hb0;
                set_twoside_space(rest, n - na, nb - 1, ha1, ha2, hb1, hb2, ha0, hb0);
                for (int k = na + 1; k < n + 1; ++k) h[k] = ha0;
                for (int k = 1; k < nb; ++k) h[k] = hb0;
                fitted = true;
                break;
            }
            if (rest < la1 + lb1) {
                if (ha < hb) { na -= 1; n0 -= 1; } else { nb += 1; n0 -= 1; }
            } else {
                if (ha < hb) { nb += 1; n0 -= 1; } else { na -= 1; n0 -= 1; }
            }
            if (n0 < 0 || na < nb - 1) break;
        }
        if (!fitted || n0 < 0 || na < nb - 1) fail("smooth_zdepth: no transition spacing fits this layer");
    } else {
        const int na = n / 4, nb = n / 4;
        const int n0 = n - na - nb;
        const float la1 = (float)((na + 1) * na / 2.0);
        const float lb1 = (float)((nb + 1) * nb / 2.0);
        const float la2 = (la1 + ((float)na * lb1) / (float)nb) + (float)(na * n0);
        float t = l - (float)na * ha;
        t = t - (float)nb * hb;
        t = t - (float)n0 * ha;
        const float l0 = t - (la1 * (ha - hb)) / (float)nb;
        const float ha0 = l0 / la2;
        const float hb0 = (((float)na * ha0 + ha) - hb) / (float)nb;
        for (int k = 1; k < nb + 1; ++k) h[k] = hb + (float)k * hb0;
        for (int k = nb + 1; k < nb + n0 + 1; ++k) h[k] = h[k - 1];
        for (int k = nb + n0 + 1; k < n + 1; ++k) h[k] = ha + (float)(n - k + 1) * ha0;
    }
    std::vector<float> z(n + 1, 0.0f);
    for (int k = 0; k < n; ++k) z[k + 1] = z[k] + h[k + 1];
    for (auto &v : z) v = v + zb;
    return z;
}

struct Vmap {
    int ni, nj, ns;
    std::vector<float> posx, posy, posz;  // [ni][nj], [ni][nj], [ni][nj][ns]
    std::vector<int> gp;
    std::vector<bool> eq;
    float z(int i, int j, int k) const { return posz[(size_t(i) * nj + j) * ns + k]; }
};

Vmap read_vmap(const std::string &path, int ni, int nj, int nk) {
    ConfFile c(path, "gridvmap data");
    int dims[3];
    for (int i = 0; i < 3; ++i) dims[i] = c.integer("vmap_dims", 3 + i);
    const int ns = dims[2];
    if (ns < 2) fail(path + ": vmap_dims needs at least 2 surfaces");
    Vmap v{dims[0], dims[1], ns, {}, {}, {}, {}, {}};
    for (int i = 0; i < ns - 1; ++i) v.gp.push_back(c.integer("vmap_gridpoints", 3 + i));
    for (int i = 0; i < ns - 1; ++i) v.eq.push_back(c.flag("vmap_equalspacing", 3 + i));
    if (dims[0] <= 0 || dims[1] <= 0) fail(path + ": vmap_dims must be positive");
    Scanner s = c.chunk("<vmap_anchor>");
    auto vals = s.reals(size_t(dims[0]) * dims[1] * (ns + 2));
    v.posx.resize(size_t(dims[0]) * dims[1]);
    v.posy.resize(v.posx.size());
    v.posz.resize(v.posx.size() * ns);
    for (size_t p = 0; p < v.posx.size(); ++p) {
        v.posx[p] = vals[p * (ns + 2)];
        v.posy[p] = vals[p * (ns + 2) + 1];
        for (int k = 0; k < ns; ++k) v.posz[p * ns + k] = vals[p * (ns + 2) + 2 + k];
    }
    if (dims[0] != ni) fail("The X value number(dims[0]) obtained from gridfile is different with configure number(cdx.ni)");
    if (dims[1] != nj) fail("The Y value number(dims[0]) obtained from gridfile is different with configure number(cdx.nj)");
    if (!v.eq[0] || !v.eq[ns - 2]) fail("The first layer and last layer should be equal spacing");
    for (int i = 1; i < ns - 1; ++i)
        if (!v.eq[i - 1] && !v.eq[i]) fail("There is only one layer(in two adjacent layers) need to adjust the grid spacing");
    int sum = 0;
    for (int g : v.gp) sum += g;
    if (sum + 1 != nk)
        fail("The Z value number[ sum(gridpoints)+1 ] obtained from gridfile is different with configure number(cdx.nk)");
    for (int i = 0; i < ni; ++i)
        for (int j = 0; j < nj; ++j)
            for (int k = 1; k < ns; ++k)
                if (v.z(i, j, k) > v.z(i, j, k - 1))
                    fail("The Z direction depth should be monotonically decreasing, while position[" + std::to_string(i) +
                         "][" + std::to_string(j) + "][" + std::to_string(k) + "] should deeper than [" +
                         std::to_string(i) + "][" + std::to_string(j) + "][" + std::to_string(k - 1) + "]");
    return v;
}

// interfaces zdep [ni][nj][L+1], plane counts, spacing flags and plane
// indices of the layers between them
struct Layers {
    int nl;
    std::vector<float> zdep;
    std::vector<int> gps;
    std::vector<bool> space;
    std::vector<int> index;
    float z(int nj, int i, int j, int k) const { return zdep[(size_t(i) * nj + j) * (nl + 1) + k]; }
};

// hybrid layering (gridmesh.cpp:243-375): the layer that holds the conversion
// depth ends there, one more layer runs flat from it to the shallowest point
// of the last interface, and every layer's share of the nk planes follows
// from its thickness.
Layers hybrid_layers(const Vmap &v, int nk, float cd, int &conindex) {
    const int ni = v.ni, nj = v.nj, ns = v.ns;
    auto zmin = [&](int k) {
        float m = v.z(0, 0, k);
        for (int i = 0; i < ni; ++i)
            for (int j = 0; j < nj; ++j) m = std::min(m, v.z(i, j, k));
        return m;
    };
    auto zmax = [&](int k) {
        float m = v.z(0, 0, k);
        for (int i = 0; i < ni; ++i)
            for (int j = 0; j < nj; ++j) m = std::max(m, v.z(i, j, k));
        return m;
    };
    int con = -1;
    for (int k = 0; k < ns - 1; ++k)
        if (zmin(k + 1) <= cd && cd <= zmax(k)) {
            con = k + 1;
            break;
        }
    if (con < 0) {
        char buf[64];
        std::snprintf(buf, sizeof buf, "%g", (double)cd);
        fail(std::string("conversion_interface_depth ") + buf + " m lies in no layer of the vmap grid");
    }
    const float top = cd, bottom = zmin(ns - 1);
    Layers L;
    L.nl = con + 1;
    for (int k = 0; k < con; ++k) L.space.push_back(v.eq[k]);
    L.space.push_back(true);
    L.zdep.resize(size_t(ni) * nj * (con + 2));
    for (int i = 0; i < ni; ++i)
        for (int j = 0; j < nj; ++j) {
            float *p = &L.zdep[(size_t(i) * nj + j) * (con + 2)];
            for (int k = 0; k < con; ++k) p[k] = v.z(i, j, k);
            p[con] = top;
            p[con + 1] = bottom;
        }
    std::vector<std::pair<float, float>> layer;
    for (int k = 0; k < con; ++k) layer.push_back({zmin(k + 1), zmax(k)});
    layer.push_back({bottom, top});
    layer[con - 1].first = top;
    float total = 0.0f;
    for (int k = 0; k < con + 1; ++k) total = total + (layer[k].second - layer[k].first);
    L.gps.assign(con + 1, 0);
    L.gps[con] = nk - 1;
    for (int k = 0; k < con; ++k) {
        const float part = layer[k].second - layer[k].first;
        L.gps[k] = (int)(((float)nk * part) / total);   // int = int*Real/Real
        L.gps[con] -= L.gps[k];
    }
    L.index.assign(con + 2, 0);
    L.index[con + 1] = LFD;
    for (int k = con; k >= 0; --k) L.index[k] = L.index[k + 1] + L.gps[k];
    conindex = L.index[con + 1] + L.gps[con];
    return L;
}

// column depths between the interfaces (gridmesh.cpp:395-414, 469-496)
Grid3 fill_z(const Layers &L, int ni, int nj, int nk) {
    Grid3 zv(ni, nj, nk + 2 * LFD);
    for (int k = 0; k < (int)L.gps.size(); ++k) {
        const int hi = L.index[k], lo = L.index[k + 1];
        for (int i = 0; i < ni; ++i)
            for (int j = 0; j < nj; ++j) {
                if (L.space[k]) {
                    const float step = (float)((double)(L.z(nj, i, j, k) - L.z(nj, i, j, k + 1)) / (1.0 * L.gps[k]));
                    for (int m = lo; m < hi + 1; ++m) zv(i, j, m) = L.z(nj, i, j, k) + (float)(m - hi) * step;
                } else {
                    const float zs0 = (float)((double)(L.z(nj, i, j, k - 1) - L.z(nj, i, j, k)) / (1.0 * L.gps[k - 1]));
                    const float zs1 = (float)((double)(L.z(nj, i, j, k + 1) - L.z(nj, i, j, k + 2)) / (1.0 * L.gps[k + 1]));
                    auto zz = smooth_zdepth(L.z(nj, i, j, k), L.z(nj, i, j, k + 1), zs0, zs1, L.gps[k]);
                    for (int m = 0; m < L.gps[k] + 1; ++m) zv(i, j, lo + m) = zz[m];
                }
                zv(i, j, hi) = L.z(nj, i, j, k);
                zv(i, j, lo) = L.z(nj, i, j, k + 1);
            }
    }
    return zv;
}

// linear extrapolation into the ghosts, z then y then x (common.cpp:67-99)
void interpolated_extend(Grid3 &v) {
    const int NX = v.nx, NY = v.ny, NZ = v.nz;
    for (int k = 1; k <= LFD; ++k)
        for (int i = 0; i < NX; ++i)
            for (int j = 0; j < NY; ++j) {
                v(i, j, LFD - k) = 2.0f * v(i, j, LFD) - v(i, j, LFD + k);
                v(i, j, NZ - LFD + k - 1) = 2.0f * v(i, j, NZ - LFD - 1) - v(i, j, NZ - LFD - k - 1);
            }
    for (int j = 1; j <= LFD; ++j)
        for (int i = 0; i < NX; ++i)
            for (int k = 0; k < NZ; ++k) {
                v(i, LFD - j, k) = 2.0f * v(i, LFD, k) - v(i, LFD + j, k);
                v(i, NY - LFD + j - 1, k) = 2.0f * v(i, NY - LFD - 1, k) - v(i, NY - LFD - j - 1, k);
            }
    for (int i = 1; i <= LFD; ++i)
        for (int j = 0; j < NY; ++j)
            for (int k = 0; k < NZ; ++k) {
                v(LFD - i, j, k) = 2.0f * v(LFD, j, k) - v(LFD + i, j, k);
                v(NX - LFD + i - 1, j, k) = 2.0f * v(NX - LFD - 1, j, k) - v(NX - LFD - i - 1, j, k);
            }
}

// the valid region mirrored into the face ghosts, no corners (common.cpp:100-135)
void mirror_extend(Grid3 &v) {
    const int a = LFD, bx = v.nx - LFD, by = v.ny - LFD, bz = v.nz - LFD;
    for (int i = 1; i <= LFD; ++i)
        for (int j = a; j < by; ++j)
            for (int k = a; k < bz; ++k) {
                v(a - i, j, k) = v(a + i, j, k);
                v(bx - 1 + i, j, k) = v(bx - 1 - i, j, k);
            }
    for (int j = 1; j <= LFD; ++j)
        for (int i = a; i < bx; ++i)
            for (int k = a; k < bz; ++k) {
                v(i, a - j, k) = v(i, a + j, k);
                v(i, by - 1 + j, k) = v(i, by - 1 - j, k);
            }
    for (int k = 1; k <= LFD; ++k)
        for (int i = a; i < bx; ++i)
            for (int j = a; j < by; ++j) {
                v(i, j, a - k) = v(i, j, a + k);
                v(i, j, bz - 1 + k) = v(i, j, bz - 1 - k);
            }
}

// ghosts of the media, with its y mirror (common.cpp:136-166)
void equivalence_extend(Grid3 &v) {
    const int a = LFD, bx = v.nx - LFD, by = v.ny - LFD, bz = v.nz - LFD;
    for (int i = 1; i <= LFD; ++i)
        for (int j = 0; j < v.ny; ++j)
            for (int k = 0; k < v.nz; ++k) {
                v(a - i, j, k) = v(a, j, k);
                v(bx + i - 1, j, k) = v(bx - 1, j, k);
            }
    for (int j = 1; j <= LFD; ++j)
        for (int i = 0; i < v.nx; ++i)
            for (int k = 0; k < v.nz; ++k) {
                v(i, a - j, k) = v(i, a + j, k);
                v(i, by + j - 1, k) = v(i, by - 1, k);
            }
    for (int k = 1; k <= LFD; ++k)
        for (int i = 0; i < v.nx; ++i)
            for (int j = 0; j < v.ny; ++j) {
                v(i, j, a - k) = v(i, j, a + k);
                v(i, j, bz + k - 1) = v(i, j, bz - 1);
            }
}

// --------------------------------------------------------------------------
// metrics (gridmesh.cpp:1476-1627): differences in float, /2.0/steph in
// double, cross products and the Jacobian in float
// --------------------------------------------------------------------------
void metrics_from_coords(const Grid3 &x, const Grid3 &y, const Grid3 &z, float dh, std::vector<float> out[10]) {
    const double steph = (double)dh;
    const int NX = x.nx, NY = x.ny, NZ = x.nz;
    Grid3 m[10];
    for (auto &g : m) g = Grid3(NX, NY, NZ);
    auto dc = [&](const Grid3 &c, int ax, int i, int j, int k) {
        const float d = ax == 0 ? c(i + 1, j, k) - c(i - 1, j, k)
                      : ax == 1 ? c(i, j + 1, k) - c(i, j - 1, k) : c(i, j, k + 1) - c(i, j, k - 1);
        return (float)((double)d / 2.0 / steph);
    };
    for (int i = LFD; i < NX - LFD; ++i)
        for (int j = LFD; j < NY - LFD; ++j)
            for (int k = LFD; k < NZ - LFD; ++k) {
                const float x_xi = dc(x, 0, i, j, k), y_xi = dc(y, 0, i, j, k), z_xi = dc(z, 0, i, j, k);
                const float x_et = dc(x, 1, i, j, k), y_et = dc(y, 1, i, j, k), z_et = dc(z, 1, i, j, k);
                const float x_zt = dc(x, 2, i, j, k), y_zt = dc(y, 2, i, j, k), z_zt = dc(z, 2, i, j, k);
                float c0 = y_xi * z_et - z_xi * y_et;
                float c1 = z_xi * x_et - x_xi * z_et;
                float c2 = x_xi * y_et - y_xi * x_et;
                const float jac = ((0.0f + c0 * x_zt) + c1 * y_zt) + c2 * z_zt;
                m[9](i, j, k) = jac;
                m[6](i, j, k) = c0 / jac;
                m[7](i, j, k) = c1 / jac;
                m[8](i, j, k) = c2 / jac;
                c0 = y_zt * z_xi - z_zt * y_xi;
                c1 = z_zt * x_xi - x_zt * z_xi;
                c2 = x_zt * y_xi - y_zt * x_xi;
                m[3](i, j, k) = c0 / jac;
                m[4](i, j, k) = c1 / jac;
                m[5](i, j, k) = c2 / jac;
                c0 = y_et * z_zt - z_et * y_zt;
                c1 = z_et * x_zt - x_et * z_zt;
                c2 = x_et * y_zt - y_et * x_zt;
                m[0](i, j, k) = c0 / jac;
                m[1](i, j, k) = c1 / jac;
                m[2](i, j, k) = c2 / jac;
            }
    for (int q = 0; q < 10; ++q) {
        mirror_extend(m[q]);
        out[q] = std::move(m[q].v);
    }
}

// --------------------------------------------------------------------------
// absorbing layers (absorb.cpp:20-46, 346-800)
// --------------------------------------------------------------------------
void cfs_pml_tables(int ni, int nj, int nk, const int nabs[6], const float vel[6], const float bmax[6],
                    const float amax[6], float dh, const std::vector<float> met[10], std::vector<float> tabs[9]) {
    const int n[3] = {ni, nj, nk}, N[3] = {ni + 6, nj + 6, nk + 6};
    const size_t NY = nj + 6, NZ = nk + 6;
    const float steph = dh, PD = 2.0f, PB = 2.0f, PA = 1.0f;
    std::vector<float> A[3], B[3], D[3];
    for (int ax = 0; ax < 3; ++ax) {
        A[ax].assign(N[ax], 0.0f);
        B[ax].assign(N[ax], 1.0f);
        D[ax].assign(N[ax], 0.0f);
    }
    auto grad_norm = [&](int ax, int i) {
        int idx[3] = {LFD, LFD, LFD};
        idx[ax] = i;
        const size_t c = (size_t(idx[0]) * NY + idx[1]) * NZ + idx[2];
        float dp = 0.0f;
        for (int q = 0; q < 3; ++q) {
            const float g = met[3 * ax + q][c];
            dp = dp + g * g;
        }
        return std::sqrt(dp);
    };
    for (int side = 0; side < 6; ++side) {
        const int ax = side / 2, hi = side % 2, nl = nabs[side];
        if (nl <= 0) continue;
        const float Rpp = (float)std::pow(10.0, -(std::log10(nl * 1.0) - 1) / std::log10(2.0) - 3.0);
        std::vector<float> L(N[ax] + 1, 0.0f);
        float L0;
        int i0, i1;
        if (hi == 0) {
            const int i2 = LFD + nl - 1;
            for (int i = i2; i > LFD - 1; --i) L[i] = L[i + 1] + steph / grad_norm(ax, i);
            L0 = L[LFD];
            i0 = LFD;
            i1 = i2 + 1;
        } else {
            const int ia = LFD + n[ax] - nl;
            for (int i = ia; i < LFD + n[ax]; ++i) L[i] = L[i - 1] + steph / grad_norm(ax, i);
            L0 = L[LFD + n[ax] - 1];
            i0 = ia;
            i1 = LFD + n[ax];
        }
        const float dmax = (float)(-(double)vel[side] / 2.0 / (double)L0 * (double)logf(Rpp) * ((double)PD + 1.0));
        for (int i = i0; i < i1; ++i) {
            const float r = L[i] / L0;
            D[ax][i] = dmax * powf(r, PD);
            B[ax][i] = (float)(1.0 + ((double)bmax[side] - 1.0) * (double)powf(r, PB));
            A[ax][i] = (float)((double)amax[side] * (1.0 - (double)powf(r, PA)));
        }
    }
    for (int ax = 0; ax < 3; ++ax) {
        tabs[ax].resize(N[ax]);
        tabs[3 + ax] = B[ax];
        tabs[6 + ax].resize(N[ax]);
        for (int i = 0; i < N[ax]; ++i) {
            tabs[6 + ax][i] = D[ax][i] / B[ax][i];
            tabs[ax][i] = A[ax][i] + D[ax][i] / B[ax][i];
        }
    }
}

// --------------------------------------------------------------------------
// source time functions (mathfunc.cpp:481-610), float build
// --------------------------------------------------------------------------
const float SQRT_PI_F = std::sqrt(PI_F);  // sqrt(PI) binds sqrtf: PI is a float

float stf_value(const std::string &name, float x, float t0, float f0) {
    if (name == "Gauss" || name == "gauss") {
        if (std::fabs(t0) >= SEIS_ZERO && (x <= 0.0f || (double)x >= 2.0 * (double)t0)) return 0.0f;
        const float d = x - t0;
        const float g = expf(-((d * d) / (f0 * f0))) / (SQRT_PI_F * f0);
        if (name == "Gauss") return g;
        return (float)((double)g * (-2.0 * (double)d / (double)(f0 * f0)));
    }
    if (name == "Ricker" || name == "ricker") {
        if (x <= 0.0f) return 0.0f;
        const float u = (float)((double)(x - t0) * 2.0 * (double)PI_F * (double)f0);
        const double uu = (double)(u * u);
        const double e = std::exp(-uu / 4.0);
        if (name == "Ricker") return (float)((uu / 4.0 - 0.5) * e * (double)SQRT_PI_F / 2.0);
        return (float)((double)u * (1.5 - uu / 4.0) * e * (double)SQRT_PI_F / 2.0 * (double)PI_F * (double)f0);
    }
    if (name == "Bell" || name == "bell" || name == "BELL") {
        const float xb = x - t0;
        const double w = 2.0 * (double)PI_F * (double)xb / (double)f0;
        if (name == "Bell") return (0.0f < xb && xb < f0) ? (float)((1.0 - std::cos(w)) / (double)f0) : 0.0f;
        if (name == "bell") return (0.0f < xb && xb < f0) ? (float)(2.0 * (double)PI_F * std::sin(w) / (double)f0) : 0.0f;
        if (xb <= 0.0f) return 0.0f;
        if (xb < f0) return (float)((double)(xb / f0) - std::sin(w) / (2.0 * (double)PI_F));
        return 1.0f;
    }
    if (name == "Step") return x > 0.0f ? 1.0f : 0.0f;
    fail("Wrong type(" + name + ") for source time function");
}

// the 2(nt+1) half-step samples of one source (source.cpp:72-132)
void stf_samples(const std::string &name, float start, float t0, float f0, int nt, float stept, bool moment,
                 float *out) {
    const int nstf = 2 * (nt + 1);
    const float dt = (float)((double)stept / 2.0);
    for (int i = 0; i < nstf; ++i) {
        const float s = stf_value(name, (float)i * dt - start, t0, f0);
        out[i] = (moment && std::fabs(s) < SEIS_ZERO) ? 0.0f : s;
    }
}

// source::angle2moment (source.cpp:134-155), float build
void angle2moment_f(float strike, float dip, float rake, float m[6]) {
    const float S = (float)((double)strike / 180.0 * (double)PI_F);
    const float D = (float)((double)dip / 180.0 * (double)PI_F);
    const float R = (float)((double)rake / 180.0 * (double)PI_F);
    const float sD = sinf(D), cD = cosf(D), sR = sinf(R), cR = cosf(R), sS = sinf(S), cS = cosf(S);
    const double s2S = std::sin(2.0 * (double)S), c2S = std::cos(2.0 * (double)S);
    const double s2D = std::sin(2.0 * (double)D), c2D = std::cos(2.0 * (double)D);
    m[0] = (float)(-((double)(sD * cR) * s2S + s2D * (double)sR * (double)sS * (double)sS));
    m[1] = (float)((double)(sD * cR) * s2S - s2D * (double)sR * (double)cS * (double)cS);
    m[2] = -(m[0] + m[1]);
    m[3] = (float)((double)(sD * cR) * c2S + s2D * (double)sR * s2S * 0.5);
    m[4] = (float)(-((double)((cD * cR) * cS) + c2D * (double)sR * (double)sS));
    m[5] = (float)(-((double)((cD * cR) * sS) - c2D * (double)sR * (double)cS));
}

// --------------------------------------------------------------------------
// point location (common.cpp:225-268): the nearest grid point in float
// squared distance, first in (i, j, k) order among equals; points above
// `hyp` go to the surface row by horizontal distance and take its z
// --------------------------------------------------------------------------
void reposition(float px, float py, float pz, const Grid3 &x, const Grid3 &y, const Grid3 &z, float hyp, int loc[3],
                float &pzout) {
    const int ni = x.nx - 2 * LFD, nj = x.ny - 2 * LFD, nk = x.nz - 2 * LFD;
    if (pz > hyp) {
        const int kz = LFD + nk - 1;
        float best = 0;
        int bi = -1, bj = -1;
        for (int i = 0; i < ni; ++i)
            for (int j = 0; j < nj; ++j) {
                const float dx = x(i + LFD, j + LFD, kz) - px, dy = y(i + LFD, j + LFD, kz) - py;
                const float p = dx * dx + dy * dy;
                if (bi < 0 || p < best) { best = p; bi = i; bj = j; }
            }
        if (!(best < SEIS_INF)) {
            loc[0] = 0; loc[1] = 0; loc[2] = kz;
            pzout = z(0, 0, kz);
            return;
        }
        loc[0] = bi + LFD; loc[1] = bj + LFD; loc[2] = kz;
        pzout = z(bi + LFD, bj + LFD, kz);
        return;
    }
    float best = SEIS_INF;
    loc[0] = loc[1] = loc[2] = 0;
    for (int i = 0; i < ni; ++i)
        for (int j = 0; j < nj; ++j)
            for (int k = 0; k < nk; ++k) {
                const float dx = x(i + LFD, j + LFD, k + LFD) - px;
                const float dy = y(i + LFD, j + LFD, k + LFD) - py;
                const float dz = z(i + LFD, j + LFD, k + LFD) - pz;
                const float p = (dx * dx + dy * dy) + dz * dz;
                if (p < best) {
                    best = p;
                    loc[0] = i + LFD; loc[1] = j + LFD; loc[2] = k + LFD;
                }
            }
    pzout = pz;
}

bool inside(const int loc[3], int ni, int nj, int nk) {
    return loc[0] >= LFD && loc[0] < LFD + ni && loc[1] >= LFD && loc[1] < LFD + nj && loc[2] >= LFD && loc[2] < LFD + nk;
}

// --------------------------------------------------------------------------
// volume media
// --------------------------------------------------------------------------
// mathfunc::LocValue1d (mathfunc.cpp:294-328): left index of the bracketing
// pair and the value clamped to the array's ends
int loc_value_1d(float p, const std::vector<float> &arr, float &val) {
    const int n = (int)arr.size();
    val = p;
    if (p <= arr[0]) { val = arr[0]; return 0; }
    if (p >= arr[n - 1]) { val = arr[n - 1]; return n - 2; }
    int L = int(std::upper_bound(arr.begin(), arr.end(), p) - arr.begin()) - 1;
    return std::min(std::max(L, 0), n - 2);
}

struct Volume {
    std::vector<float> X, Y, Z;
    std::vector<float> F[3];  // Vp Vs rho on [x][y][z] of the cut
};

Volume volume_table(const std::string &path, const Grid3 &x, const Grid3 &y, const Grid3 &z, bool zwmedia) {
    auto f = open_netcdf(path, "media volume file");
    Volume v;
    auto X = f->var("x"), Y = f->var("y"), Z = f->var(zwmedia ? "depth" : "z");
    std::vector<float> full[3] = {f->var("Vp"), f->var("Vs"), f->var("rho")};
    for (auto &a : full)
        if (a.size() != X.size() * Y.size() * Z.size()) fail("media volume file " + path + ": Vp/Vs/rho do not match x, y, z");
    const int ni = x.nx - 2 * LFD, nj = x.ny - 2 * LFD, nk = x.nz - 2 * LFD;
    const Grid3 *crd[3] = {&x, &y, &z};
    const std::vector<float> *axes[3] = {&X, &Y, &Z};
    int lo[3], hi[3];
    for (int a = 0; a < 3; ++a) {
        float mn = (*crd[a])(LFD, LFD, LFD), mx = mn;
        for (int i = LFD; i < LFD + ni; ++i)
            for (int j = LFD; j < LFD + nj; ++j)
                for (int k = LFD; k < LFD + nk; ++k) {
                    mn = std::min(mn, (*crd[a])(i, j, k));
                    mx = std::max(mx, (*crd[a])(i, j, k));
                }
        float dummy;
        lo[a] = loc_value_1d(mn, *axes[a], dummy);
        hi[a] = loc_value_1d(mx, *axes[a], dummy) + 2;
    }
    v.X.assign(X.begin() + lo[0], X.begin() + hi[0]);
    v.Y.assign(Y.begin() + lo[1], Y.begin() + hi[1]);
    v.Z.assign(Z.begin() + lo[2], Z.begin() + hi[2]);
    for (int q = 0; q < 3; ++q)
        for (int i = lo[0]; i < hi[0]; ++i)
            for (int j = lo[1]; j < hi[1]; ++j)
                for (int k = lo[2]; k < hi[2]; ++k) v.F[q].push_back(full[q][(size_t(i) * Y.size() + j) * Z.size() + k]);
    return v;
}

// volume_discrete (mediaparas.cpp:376-441): two-node Lagrange interpolation
// in each direction
void volume_sample(const Volume &vt, float px, float py, float pz, float out[3]) {
    float xv, yv, zv;
    const int i1 = loc_value_1d(px, vt.X, xv), j1 = loc_value_1d(py, vt.Y, yv), k1 = loc_value_1d(pz, vt.Z, zv);
    const float lx[2] = {(vt.X[i1 + 1] - xv) / (vt.X[i1 + 1] - vt.X[i1]), (vt.X[i1] - xv) / (vt.X[i1] - vt.X[i1 + 1])};
    const float ly[2] = {(vt.Y[j1 + 1] - yv) / (vt.Y[j1 + 1] - vt.Y[j1]), (vt.Y[j1] - yv) / (vt.Y[j1] - vt.Y[j1 + 1])};
    const float lz[2] = {(vt.Z[k1 + 1] - zv) / (vt.Z[k1 + 1] - vt.Z[k1]), (vt.Z[k1] - zv) / (vt.Z[k1] - vt.Z[k1 + 1])};
    const size_t ny = vt.Y.size(), nz = vt.Z.size();
    for (int q = 0; q < 3; ++q) {
        float acc = 0.0f;
        for (int a = 0; a < 2; ++a)
            for (int b = 0; b < 2; ++b)
                for (int c = 0; c < 2; ++c) {
                    const float w = (lx[a] * ly[b]) * lz[c];
                    acc = acc + w * vt.F[q][(size_t(i1 + a) * ny + j1 + b) * nz + k1 + c];
                }
        out[q] = acc;
    }
}

// TwoinOne over samples, then average (mediaparas.cpp:838-914); s[n] =
// (vp0, vs0, dc0, vp1, vs1, dc1).  The reference's test `nsamp>0 &&
// muval==0` compares pointers and is never true, so mu's harmonic sum
// restarts at 0 after any fluid sample.
void two_in_one_acc(const std::vector<std::array<float, 6>> &s, float &density, float &mu, float &lamb) {
    float rhoval = 0.0f, lamval = 0.0f, muval = 0.0f;
    for (auto &p : s) {
        const float vp0 = p[0], vs0 = p[1], dc0 = p[2], vp1 = p[3], vs1 = p[4], dc1 = p[5];
        const float m0 = (dc0 * vs0) * vs0, m1 = (dc1 * vs1) * vs1;
        const float l0 = (float)((double)((vp0 * vp0) * dc0) - 2.0 * (double)m0);
        const float l1 = (float)((double)((vp1 * vp1) * dc1) - 2.0 * (double)m1);
        rhoval = (float)((double)rhoval + 0.5 * (double)(dc0 + dc1));
        const bool fluid = m0 < SEIS_ZERO || m1 < SEIS_ZERO;
        const float acc = (float)(((double)muval + 0.5 / (double)m0) + 0.5 / (double)m1);
        lamval = (float)(((double)lamval + 0.5 / (double)l0) + 0.5 / (double)l1);
        muval = fluid ? 0.0f : acc;
    }
    const float n = (float)s.size();
    density = rhoval / n;
    mu = muval > 0 ? (float)((double)n * 1.0 / (double)muval) : 0.0f;
    lamb = (float)((double)n * 1.0 / (double)lamval);
}

// set_vel3d (mediaparas.cpp:916-923)
float vel3d(float vp0, float vp1, float cv, float cp) {
    const float d = std::fabs(vp0 - vp1);
    const bool smooth = d < cv && d / std::min(vp0, vp1) < cp;
    const double mean = (double)(vp0 + vp1) / 2.0;
    return (float)(smooth ? mean : mean + 9000e3);
}

// three-node Lagrange weights on nodes 0, 2, 4 (interp3d with vec3)
void lag3(float x0, float L[3]) {
    const float X[3] = {0.0f, 2.0f, 4.0f};
    for (int i = 0; i < 3; ++i) {
        float w = 1.0f;
        for (int j = 0; j < 3; ++j) w = w * (j == i ? 1.0f : (X[j] - x0) / (X[j] - X[i]));
        L[i] = w;
    }
}

using SamplePairs = std::function<std::array<float, 6>(float, float, float)>;

// ApplySmooth (mediaparas.cpp:1712-1824) when a sampling factor is above 1,
// MediaStatistics (1837-1905) and the ghost extension of readdata
void finish_media(Grid3 &density, Grid3 &miu, Grid3 &lamb, const Grid3 *v3d, const Grid3 &x, const Grid3 &y,
                  const Grid3 &z, const int samp[3], float cv, float cp, const SamplePairs &pairs, Scenario &sc) {
    const int NX = x.nx, NY = x.ny, NZ = x.nz;
    const int ni = NX - 2 * LFD, nj = NY - 2 * LFD, nk = NZ - 2 * LFD;
    if (v3d && std::max({samp[0], samp[1], samp[2]}) > 1) {
        auto over = [&](float a, float b) {
            const float d = std::fabs(a - b);
            return d >= cv || d / std::min(a, b) >= cp;
        };
        const Grid3 &V = *v3d;
        for (int i = LFD; i < LFD + ni; ++i)
            for (int j = LFD; j < LFD + nj; ++j)
                for (int k = LFD; k < LFD + nk; ++k) {
                    const float c = V(i, j, k);
                    const int nx = (over(c, V(i - 1, j, k)) || over(c, V(i + 1, j, k))) ? samp[0] : 1;
                    const int ny = (over(c, V(i, j - 1, k)) || over(c, V(i, j + 1, k))) ? samp[1] : 1;
                    const int nz = (over(c, V(i, j, k - 1)) || over(c, V(i, j, k + 1))) ? samp[2] : 1;
                    if (nx <= 1 && ny <= 1 && nz <= 1) continue;
                    std::vector<std::array<float, 6>> s;
                    for (int mi = 1; mi <= nx; ++mi)
                        for (int mj = 1; mj <= ny; ++mj)
                            for (int mk = 1; mk <= nz; ++mk) {
                                float lx[3], ly[3], lz[3];
                                lag3((float)(1.0 + 2.0 / nx * (mi - 1 + 0.5)), lx);
                                lag3((float)(1.0 + 2.0 / ny * (mj - 1 + 0.5)), ly);
                                lag3((float)(1.0 + 2.0 / nz * (mk - 1 + 0.5)), lz);
                                float pos[3];
                                const Grid3 *crd[3] = {&x, &y, &z};
                                for (int q = 0; q < 3; ++q) {
                                    float acc = 0.0f;
                                    for (int a = 0; a < 3; ++a)
                                        for (int b = 0; b < 3; ++b)
                                            for (int g = 0; g < 3; ++g) {
                                                const float w = (lx[a] * ly[b]) * lz[g];
                                                acc = acc + w * (*crd[q])(i - 1 + a, j - 1 + b, k - 1 + g);
                                            }
                                    pos[q] = acc;
                                }
                                s.push_back(pairs(pos[0], pos[1], pos[2]));
                            }
                    two_in_one_acc(s, density(i, j, k), miu(i, j, k), lamb(i, j, k));
                }
    }
    Grid3 alpha(NX, NY, NZ), beta(NX, NY, NZ), rho(NX, NY, NZ);
    for (int i = LFD; i < LFD + ni; ++i)
        for (int j = LFD; j < LFD + nj; ++j)
            for (int k = LFD; k < LFD + nk; ++k) {
                beta(i, j, k) = std::sqrt(miu(i, j, k) / density(i, j, k));
                alpha(i, j, k) = (float)std::sqrt(((double)lamb(i, j, k) + 2.0 * (double)miu(i, j, k)) /
                                                  (double)density(i, j, k));
                rho(i, j, k) = density(i, j, k);
            }
    equivalence_extend(alpha);
    equivalence_extend(beta);
    equivalence_extend(rho);
    sc.alpha = std::move(alpha.v);
    sc.beta = std::move(beta.v);
    sc.rho = std::move(rho.v);
}

void volume_media(const Main &m, const Grid3 &x, const Grid3 &y, const Grid3 &z, bool zwmedia, Scenario &sc) {
    const Volume vt = volume_table(m.media_file, x, y, z, zwmedia);
    const int NX = x.nx, NY = x.ny, NZ = x.nz;
    Grid3 density(NX, NY, NZ), miu(NX, NY, NZ), lamb(NX, NY, NZ), v3d(NX, NY, NZ);
    for (size_t c = 0; c < x.v.size(); ++c) {
        float s[3];
        volume_sample(vt, x.v[c], y.v[c], z.v[c], s);
        // get2in1 (mediaparas.cpp:872-886) with equal upper and lower values
        const float mu = (s[2] * s[1]) * s[1];
        const float lam = (float)((double)((s[0] * s[0]) * s[2]) - 2.0 * (double)mu);
        density.v[c] = (float)(((double)s[2] + s[2]) / 2.0);
        miu.v[c] = (float)(((double)mu + mu) / 2.0);
        lamb.v[c] = (float)(((double)lam + lam) / 2.0);
        v3d.v[c] = vel3d(s[0], s[0], m.crit_value, m.crit_perc);
    }
    SamplePairs pairs = [&](float px, float py, float pz) {
        float s[3];
        volume_sample(vt, px, py, pz, s);
        return std::array<float, 6>{s[0], s[1], s[2], s[0], s[1], s[2]};
    };
    finish_media(density, miu, lamb, &v3d, x, y, z, m.sampling, m.crit_value, m.crit_perc, pairs, sc);
}

// --------------------------------------------------------------------------
// 3-D interface media (mediaparas.cpp:52-374, 442-836)
// --------------------------------------------------------------------------
struct Interface3D {
    int nx, ny, nint;
    std::vector<float> Z;       // [nx][ny][nint]
    std::vector<float> V[3];    // Vp Vs Dc [2][nx][ny][nint]
    std::vector<float> X, Y;
    std::vector<bool> flat, layer_const;
    std::vector<std::array<bool, 2>> cnst;
    std::vector<std::array<float, 2>> range;
    std::vector<int> idx;       // [nint][nx][ny][2]
    float z(int i, int j, int n) const { return Z[(size_t(i) * ny + j) * nint + n]; }
    float v(int q, int m, int i, int j, int n) const { return V[q][((size_t(m) * nx + i) * ny + j) * nint + n]; }
    int &ix(int n, int i, int j, int s) { return idx[((size_t(n) * nx + i) * ny + j) * 2 + s]; }
    int ix(int n, int i, int j, int s) const { return idx[((size_t(n) * nx + i) * ny + j) * 2 + s]; }
};

Interface3D read_interface3d(const std::string &path, int nx, int ny, int x0, int y0, int dx, int dy, int nint) {
    ConfFile c(path, "media interface");
    std::vector<std::string> toks;
    for (auto &p : c.pieces())
        for (auto &t : split_ws(p)) toks.push_back(t);
    Scanner s(std::move(toks), path);
    Interface3D t;
    t.nx = nx; t.ny = ny; t.nint = nint;
    t.Z.resize(size_t(nx) * ny * nint);
    auto zr = s.reals(size_t(nint) * ny * nx);   // file order [n][j][i]
    for (int n = 0; n < nint; ++n)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) t.Z[(size_t(i) * ny + j) * nint + n] = zr[(size_t(n) * ny + j) * nx + i];
    for (int q = 0; q < 3; ++q) {
        auto r = s.reals(size_t(nint) * 2 * ny * nx);   // file order [n][m][j][i]
        t.V[q].resize(r.size());
        for (int n = 0; n < nint; ++n)
            for (int m = 0; m < 2; ++m)
                for (int j = 0; j < ny; ++j)
                    for (int i = 0; i < nx; ++i)
                        t.V[q][((size_t(m) * nx + i) * ny + j) * nint + n] = r[((size_t(n) * 2 + m) * ny + j) * nx + i];
    }
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j)
            for (int n = 0; n + 1 < nint; ++n)
                if (t.z(i, j, n) < t.z(i, j, n + 1))
                    fail("at location (" + std::to_string(i) + "," + std::to_string(j) + "," + std::to_string(n) +
                         ") interface depth[" + std::to_string(n) + "] is lower than depth[" + std::to_string(n + 1) + "]");
    for (int m = 0; m < 2; ++m)
        for (int i = 0; i < nx; ++i)
            for (int j = 0; j < ny; ++j)
                for (int n = 0; n < nint; ++n) {
                    const double vp = t.v(0, m, i, j, n), vs = t.v(1, m, i, j, n);
                    const bool air = m == 0 && n == 0 && t.v(0, 0, i, j, 0) == 0 && t.v(1, 0, i, j, 0) == 0;
                    if (vp * vp <= 2.0 * vs * vs && !air) fail("there is a velocity error: Vp^2 <= 2 Vs^2");
                }
    for (int i = 0; i < nx; ++i) t.X.push_back((float)(x0 + i * dx));
    for (int j = 0; j < ny; ++j) t.Y.push_back((float)(y0 + j * dy));
    for (int n = 0; n < nint; ++n) {
        bool f = true;
        for (int i = 0; i < nx; ++i)
            for (int j = 0; j < ny; ++j) f = f && std::fabs(t.z(0, 0, n) - t.z(i, j, n)) <= SEIS_ZERO;
        t.flat.push_back(f);
        std::array<bool, 2> cc{};
        for (int m = 0; m < 2; ++m) {
            bool k = true;
            for (int q = 0; q < 3; ++q)
                for (int i = 0; i < nx; ++i)
                    for (int j = 0; j < ny; ++j) k = k && std::fabs(t.v(q, m, 0, 0, n) - t.v(q, m, i, j, n)) <= SEIS_ZERO;
            cc[m] = k;
        }
        t.cnst.push_back(cc);
    }
    for (int n = 0; n < nint; ++n) {
        bool lc = false;
        if (n == nint - 1 && t.cnst[n][1]) {
            lc = true;
        } else if (t.cnst[n][1] && n + 1 < nint && t.cnst[n + 1][0]) {
            lc = true;
            for (int q = 0; q < 3; ++q) lc = lc && std::fabs(t.v(q, 1, 0, 0, n) - t.v(q, 0, 0, 0, n + 1)) < SEIS_ZERO;
        }
        t.layer_const.push_back(lc);
    }
    t.range.resize(nint);
    auto zmax = [&](int n) {
        float m = t.z(0, 0, n);
        for (int i = 0; i < nx; ++i)
            for (int j = 0; j < ny; ++j) m = std::max(m, t.z(i, j, n));
        return m;
    };
    auto zmin = [&](int n) {
        float m = t.z(0, 0, n);
        for (int i = 0; i < nx; ++i)
            for (int j = 0; j < ny; ++j) m = std::min(m, t.z(i, j, n));
        return m;
    };
    for (int n = 0; n + 1 < nint; ++n)
        t.range[n] = {std::max(t.z(0, 0, n), zmax(n)), std::min(t.z(0, 0, n + 1), zmin(n + 1))};
    t.range[nint - 1] = {zmax(nint - 1), (float)((double)zmin(nint - 1) - 6371e3)};
    t.idx.assign(size_t(nint) * nx * ny * 2, 0);
    for (int i = 0; i + 1 < nx; ++i)
        for (int j = 0; j + 1 < ny; ++j) {
            auto h = [&](int n1, int n2) {
                float m = 0.0f;
                for (int a = 0; a < 2; ++a)
                    for (int b = 0; b < 2; ++b) m = std::max(m, std::fabs(t.z(i + a, j + b, n1) - t.z(i + a, j + b, n2)));
                return m;
            };
            t.ix(nint - 1, i, j, 1) = nint;
            for (int n = nint - 2; n >= 0; --n) t.ix(n, i, j, 1) = h(n, n + 1) < SEIS_ZERO ? t.ix(n + 1, i, j, 1) : n + 1;
            const bool top_air = t.cnst[0][0] && t.v(0, 0, 0, 0, 0) == 0;
            t.ix(0, i, j, 0) = top_air ? -t.ix(0, i, j, 1) : 1;
            for (int n = 1; n < nint; ++n) t.ix(n, i, j, 0) = h(n, n - 1) < SEIS_ZERO ? t.ix(n - 1, i, j, 0) : n + 1;
        }
    for (int n = 0; n < nint; ++n)
        for (int s2 = 0; s2 < 2; ++s2) {
            for (int i = 0; i < nx; ++i) t.ix(n, i, ny - 1, s2) = t.ix(n, i, ny - 2, s2);
            for (int j = 0; j < ny; ++j) t.ix(n, nx - 1, j, s2) = t.ix(n, nx - 2, j, s2);
        }
    return t;
}

using Triple = std::array<float, 3>;

struct Column {
    int i1, j1;
    std::vector<float> z;
    std::map<std::pair<int, int>, Triple> side, inner;
};

float interp2d(const float lx[2], const float ly[2], float z00, float z01, float z10, float z11) {
    const float Z[2][2] = {{z00, z01}, {z10, z11}};
    float acc = 0.0f;
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b) acc = acc + (lx[a] * ly[b]) * Z[a][b];
    return acc;
}

Column i3d_column(const Interface3D &t, float x0, float y0) {
    Column col;
    float xv, yv;
    col.i1 = loc_value_1d(x0, t.X, xv);
    col.j1 = loc_value_1d(y0, t.Y, yv);
    const int i1 = col.i1, j1 = col.j1;
    const float lx[2] = {(t.X[i1 + 1] - xv) / (t.X[i1 + 1] - t.X[i1]), (t.X[i1] - xv) / (t.X[i1] - t.X[i1 + 1])};
    const float ly[2] = {(t.Y[j1 + 1] - yv) / (t.Y[j1 + 1] - t.Y[j1]), (t.Y[j1] - yv) / (t.Y[j1] - t.Y[j1 + 1])};
    auto iz = [&](int n) {
        return interp2d(lx, ly, t.z(i1, j1, n), t.z(i1, j1 + 1, n), t.z(i1 + 1, j1, n), t.z(i1 + 1, j1 + 1, n));
    };
    auto iv = [&](int q, int m, int n) {
        return interp2d(lx, ly, t.v(q, m, i1, j1, n), t.v(q, m, i1, j1 + 1, n), t.v(q, m, i1 + 1, j1, n),
                        t.v(q, m, i1 + 1, j1 + 1, n));
    };
    for (int n = 0; n < t.nint; ++n) col.z.push_back(t.flat[n] ? t.z(i1, j1, n) : iz(n));
    for (int n = 0; n < t.nint; ++n)
        for (int m = 0; m < 2; ++m) {
            Triple s, in;
            for (int q = 0; q < 3; ++q) {
                in[q] = iv(q, m, n);
                s[q] = t.cnst[n][m] ? t.v(q, m, i1, j1, n) : in[q];
            }
            col.side[{n, m}] = s;
            col.inner[{n, m}] = in;
        }
    return col;
}

// interface3d_discrete (mediaparas.cpp:442-836) at depth z0 of a column:
// the (upper, lower) pair; `prev` when no layer matches
std::array<Triple, 2> i3d_point(const Interface3D &t, const Column &col, float z0, const std::array<Triple, 2> &prev) {
    const int nint = t.nint, i1 = col.i1, j1 = col.j1;
    auto upper = [&](int n) {
        int k = t.ix(n, i1, j1, 0), s = 0;
        if (k < 0) { k = -k; s = 1; }
        return col.side.at({k - 1, s});
    };
    auto lower = [&](int n) { return col.side.at({t.ix(n, i1, j1, 1) - 1, 1}); };
    int n1 = nint + 1, n2 = -1;
    for (int n = 0; n < nint; ++n)
        if (t.range[n][1] <= z0 && z0 <= t.range[n][0]) {
            n1 = std::min(n1, n);
            n2 = std::max(n2, n + 1);
        }
    n2 = std::min(n2, nint - 1);
    if (n1 > n2) {
        auto v = upper(0);
        return {v, v};
    }
    for (int n = n1; n < n2 + 1; ++n) {
        if (t.ix(n, i1, j1, 1) != n + 1) continue;
        const float z1 = col.z[n];
        if (std::fabs(z0 - z1) < SEIS_ZERO) return {upper(n), lower(n)};
        if (z0 > z1) {
            auto v = upper(n);
            return {v, v};
        }
        if (n == nint - 1) {
            auto v = lower(n);
            return {v, v};
        }
        const float z2 = col.z[n + 1];
        if (std::fabs(z0 - z2) < SEIS_ZERO) return {upper(n + 1), lower(n + 1)};
        if (z0 > z2) {
            if (t.layer_const[n]) {
                Triple v{t.v(0, 1, i1, j1, n), t.v(1, 1, i1, j1, n), t.v(2, 1, i1, j1, n)};
                return {v, v};
            }
            const float L = z1 - z2;
            const float L1 = (z0 - z2) / L;
            const float L2 = 1.0f - L1;
            const Triple &a = col.inner.at({n, 1}), &b = col.inner.at({n + 1, 0});
            Triple v;
            for (int q = 0; q < 3; ++q) v[q] = a[q] * L1 + b[q] * L2;
            return {v, v};
        }
    }
    return prev;
}

void interface3d_media(const Main &m, const ConfFile &c, const Grid3 &x, const Grid3 &y, const Grid3 &z, Scenario &sc) {
    int par[7];
    for (int q = 0; q < 7; ++q) par[q] = c.integer("interface3dpar", 3 + q);
    const Interface3D t = read_interface3d(m.media_file, par[0], par[1], par[2], par[3], par[4], par[5], par[6]);
    const int NX = x.nx, NY = x.ny, NZ = x.nz;
    Grid3 side[6];
    for (auto &g : side) g = Grid3(NX, NY, NZ);
    std::array<Triple, 2> prev{};
    for (int i = 0; i < NX; ++i)
        for (int j = 0; j < NY; ++j) {
            std::map<std::pair<float, float>, Column> cols;
            for (int k = 0; k < NZ; ++k) {
                const std::pair<float, float> key{x(i, j, k), y(i, j, k)};
                auto it = cols.find(key);
                if (it == cols.end()) it = cols.emplace(key, i3d_column(t, key.first, key.second)).first;
                prev = i3d_point(t, it->second, z(i, j, k), prev);
                for (int q = 0; q < 6; ++q) side[q](i, j, k) = prev[q / 3][q % 3];
            }
        }
    Grid3 density(NX, NY, NZ), miu(NX, NY, NZ), lamb(NX, NY, NZ), v3d(NX, NY, NZ);
    for (size_t cidx = 0; cidx < x.v.size(); ++cidx) {
        std::vector<std::array<float, 6>> s{{side[0].v[cidx], side[1].v[cidx], side[2].v[cidx], side[3].v[cidx],
                                             side[4].v[cidx], side[5].v[cidx]}};
        two_in_one_acc(s, density.v[cidx], miu.v[cidx], lamb.v[cidx]);
        v3d.v[cidx] = vel3d(side[0].v[cidx], side[3].v[cidx], m.crit_value, m.crit_perc);
    }
    SamplePairs pairs = [&](float px, float py, float pz) {
        auto p = i3d_point(t, i3d_column(t, px, py), pz, prev);
        return std::array<float, 6>{p[0][0], p[0][1], p[0][2], p[1][0], p[1][1], p[1][2]};
    };
    finish_media(density, miu, lamb, &v3d, x, y, z, m.sampling, m.crit_value, m.crit_perc, pairs, sc);
}

// --------------------------------------------------------------------------
// sources (source.cpp:12-70, 671-849, 985-1289)
// --------------------------------------------------------------------------
void load_sources(const Main &m, const Grid3 &x, const Grid3 &y, const Grid3 &z, Scenario &sc) {
    ConfFile c(m.source_file, "source configure");
    const int stype = c.integer("source_type");
    const float dst2m = c.real("distance2meter");
    const float hyp = c.real("source_hyper_height") * dst2m;
    const float fd0 = c.real("simulate_damp");  // read for every source type (source.cpp:786-788)
    const int ni = m.ni, nj = m.nj, nk = m.nk, nt = m.nt;
    auto mu_at = [&](const int l[3]) {
        const size_t q = (size_t(l[0]) * (nj + 6) + l[1]) * (nk + 6) + l[2];
        return (sc.rho[q] * sc.beta[q]) * sc.beta[q];
    };
    if (stype == 1) {
        auto f = open_netcdf(m.seispath + "/" + c.str("focal_process_NC_file"), "focal process file");
        const long long npnt = f->dim("NP"), fnt = f->dim("NT");
        const float fdt = f->att("DT");
        const char *names[6] = {"Mxx", "Myy", "Mzz", "Mxy", "Mxz", "Myz"};
        std::vector<float> ser[6];
        for (int q = 0; q < 6; ++q) {
            ser[q] = f->var(names[q]);
            if ((long long)ser[q].size() != npnt * fnt) fail("focal process file: " + std::string(names[q]) + " is not NP x NT");
            for (auto &v : ser[q]) v = v * fd0;
        }
        auto px = f->var("posx"), py = f->var("posy"), pz = f->var("posz");
        std::vector<int> keep;
        for (long long p = 0; p < npnt; ++p) {
            int l[3];
            float zz;
            reposition(px[p], py[p], pz[p], x, y, z, hyp, l, zz);
            if (inside(l, ni, nj, nk)) {
                keep.push_back((int)p);
                sc.focal_loc.insert(sc.focal_loc.end(), l, l + 3);
            }
        }
        const size_t nf = keep.size();
        sc.focal_series.assign(6 * nf * size_t(fnt), 0.0f);
        for (size_t n = 0; n < nf; ++n) {
            const float mu = mu_at(&sc.focal_loc[3 * n]);
            for (int q = 0; q < 6; ++q)
                for (long long s = 0; s < fnt; ++s)
                    sc.focal_series[(q * nf + n) * fnt + s] = mu * ser[q][size_t(keep[n]) * fnt + s];
        }
        sc.focal_nt = (int)fnt;
        sc.focal_dt = fdt;
        return;
    }
    if (stype != 2 && stype != 3) return;  // any other type leaves the run without sources (source.cpp:39-66)
    const int nfrc = c.integer("number_of_force_source"), nmnt = c.integer("number_of_moment_source");
    const int nstf = 2 * (nt + 1);
    if (nfrc) {
        const int nwin = c.integer("force_stf_window");
        Scanner s = c.chunk("<anchor_force>");
        sc.force_stf.assign(size_t(nfrc) * nstf, 0.0f);
        for (int i = 0; i < nfrc; ++i) {
            auto r = s.reals(8);
            int l[3];
            float zz;
            reposition(r[0] * dst2m, r[1] * dst2m, r[2] * dst2m, x, y, z, hyp, l, zz);
            if (!inside(l, ni, nj, nk)) {
                char buf[160];
                std::snprintf(buf, sizeof buf, "The %d force source at position(%g,%g,%g), exceed the coord boundary", i + 1,
                              (double)(r[0] * dst2m), (double)(r[1] * dst2m), (double)(r[2] * dst2m));
                fail(buf);
            }
            sc.force_loc.insert(sc.force_loc.end(), l, l + 3);
            sc.force_vec.insert(sc.force_vec.end(), {r[5] * r[4], r[6] * r[4], r[7] * r[4]});
            for (int k = 0; k < nwin; ++k)  // the last window is what the reference keeps
                stf_samples(c.str("force_stf_type", 3 + k), r[3], c.real("force_stf_timefactor", 3 + k),
                            c.real("force_stf_freqfactor", 3 + k), nt, m.stept, false, &sc.force_stf[size_t(i) * nstf]);
        }
    }
    if (nmnt) {
        const int nwin = c.integer("moment_stf_window");
        const std::string mech = c.str("moment_mech_input");
        if (mech != "moment" && mech != "angle") fail("moment mechanics input wrong = " + mech);
        const bool angle = mech == "angle";
        Scanner s = c.chunk("<anchor_moment>");
        sc.moment_stf.assign(size_t(nmnt) * nstf, 0.0f);
        for (int i = 0; i < nmnt; ++i) {
            auto r = s.reals(angle ? 8 : 11);
            float mv[6];
            if (angle) angle2moment_f(r[5], r[6], r[7], mv);
            else std::copy(r.begin() + 5, r.begin() + 11, mv);
            int l[3];
            float zz;
            reposition(r[0] * dst2m, r[1] * dst2m, r[2] * dst2m, x, y, z, hyp, l, zz);
            if (!inside(l, ni, nj, nk)) {
                char buf[160];
                std::snprintf(buf, sizeof buf, "The %d moment source at position(%g,%g,%g), exceed the coord boundary", i + 1,
                              (double)(r[0] * dst2m), (double)(r[1] * dst2m), (double)(r[2] * dst2m));
                fail(buf);
            }
            sc.moment_loc.insert(sc.moment_loc.end(), l, l + 3);
            const float mu = angle ? mu_at(l) : 1.0f;
            for (int q = 0; q < 6; ++q) {
                const float v = mv[q] * r[4];
                sc.moment_vec.push_back(angle ? v * mu : v);
            }
            for (int k = 0; k < nwin; ++k)
                stf_samples(c.str("moment_stf_type", 3 + k), r[3], c.real("moment_stf_timefactor", 3 + k),
                            c.real("moment_stf_freqfactor", 3 + k), nt, m.stept, true, &sc.moment_stf[size_t(i) * nstf]);
        }
    }
}

// --------------------------------------------------------------------------
// stations (seisplot.cpp:12-55, 330-457)
// --------------------------------------------------------------------------
void load_stations(const Main &m, const Grid3 &x, const Grid3 &y, const Grid3 &z, Scenario &sc) {
    ConfFile c(m.station_file, "seisplot configure");
    const int nsnap = c.integer("number_of_snap"), nline = c.integer("number_of_inline"), nrecv = c.integer("number_of_recv");
    const float hyp = c.real("topo_hyper_height");
    sc.seismo_tinv = c.integer("tinv_of_seismo");
    for (int i = 0; i < nsnap; ++i) {
        char key[32];
        std::snprintf(key, sizeof key, "snap_%03d", i + 1);
        for (int col = 3; col < 14; ++col) sc.snapshots.push_back(c.integer(key, col));
    }
    for (int i = 0; i < nrecv; ++i) {
        char key[32];
        std::snprintf(key, sizeof key, "recv_%03d", i + 1);
        sc.recv_pos.insert(sc.recv_pos.end(), {c.real(key, 3), c.real(key, 4), c.real(key, 5)});
        sc.recv_pnum.push_back(i);
        sc.recv_lnum.push_back(0);
    }
    for (int i = 0; i < nline; ++i) {
        char key[32];
        std::snprintf(key, sizeof key, "line_%03d", i + 1);
        const float x0 = c.real(key, 3), y0 = c.real(key, 4), z0 = c.real(key, 5);
        const float dx = c.real(key, 7), dy = c.real(key, 8), dz = c.real(key, 9);
        const float num = c.real(key, 11);
        for (int j = 0; (float)j < num; ++j) {
            sc.recv_pos.insert(sc.recv_pos.end(), {x0 + (float)j * dx, y0 + (float)j * dy, z0 + (float)j * dz});
            sc.recv_pnum.push_back(j);
            sc.recv_lnum.push_back(i + 1);
        }
    }
    const int n = (int)sc.recv_pnum.size();
    for (int i = 0; i < n; ++i) {
        int l[3];
        float zz;
        reposition(sc.recv_pos[3 * i], sc.recv_pos[3 * i + 1], sc.recv_pos[3 * i + 2], x, y, z, hyp, l, zz);
        if (!inside(l, m.ni, m.nj, m.nk)) fail("No." + std::to_string(i) + " reciver is out of Computing Boundary (No Virtual)!");
        sc.recv_pos[3 * i + 2] = zz;
        sc.recv_loc.insert(sc.recv_loc.end(), l, l + 3);
    }
}

} // namespace

Scenario load(const std::string &conf, const std::string &seispath, bool zwmedia) {
    const Main m = read_main(conf, seispath);
    if (m.grid_type == "point")
        fail("grid_type point: the reference's read_point stores x, y and z all into crd.x (gridmesh.cpp:52-68), so "
             "it yields no valid grid; use a vmap grid");
    if (m.grid_type != "vmap") fail("gridtype configure " + m.grid_type + " is wrong when reading grid data");
    Scenario sc;
    sc.ni = m.ni; sc.nj = m.nj; sc.nk = m.nk; sc.nt = m.nt;
    sc.dh = m.steph; sc.dt = m.stept;
    sc.pgv = m.output_peakvel != 0;
    sc.hybrid = m.hybrid != 0;
    const int ni = m.ni, nj = m.nj, nk = m.nk;
    const int NX = ni + 6, NY = nj + 6, NZ = nk + 6;

    // grid
    const Vmap vm = read_vmap(m.grid_file, ni, nj, nk);
    Layers L;
    int conindex = LFD;
    if (m.hybrid) {
        L = hybrid_layers(vm, nk, m.conv_depth, conindex);
    } else {
        L.nl = vm.ns - 1;
        L.zdep = vm.posz;
        L.gps = vm.gp;
        L.space = vm.eq;
        L.index.assign(vm.ns, 0);
        L.index[vm.ns - 1] = LFD;
        for (int i = vm.ns - 2; i >= 0; --i) L.index[i] = L.index[i + 1] + vm.gp[i];
    }
    const Grid3 zv = fill_z(L, ni, nj, nk);
    Grid3 x(NX, NY, NZ), y(NX, NY, NZ), z(NX, NY, NZ);
    for (int i = 0; i < ni; ++i)
        for (int j = 0; j < nj; ++j) {
            for (int k = 0; k < NZ; ++k) {
                x(i + LFD, j + LFD, k) = vm.posx[size_t(i) * nj + j];
                y(i + LFD, j + LFD, k) = vm.posy[size_t(i) * nj + j];
            }
            for (int k = LFD; k < LFD + nk; ++k) z(i + LFD, j + LFD, k) = zv(i, j, k);
        }
    interpolated_extend(x);
    interpolated_extend(y);
    interpolated_extend(z);
    sc.conindex = conindex;
    metrics_from_coords(x, y, z, m.steph, sc.metrics);

    // media
    if (m.media_type == "volume") {
        volume_media(m, x, y, z, zwmedia, sc);
    } else if (m.media_type == "interface3D") {
        ConfFile c(conf, "main par");
        interface3d_media(m, c, x, y, z, sc);
    } else if (m.media_type == "interface") {
        fail("media_type interface: the reference's readdata overwrites the interface media with zero Vp/Vs/rho "
             "(get2in1, mediaparas.cpp:2163-2200) and MediaStatistics turns them into NaN; use volume media");
    } else {
        fail("media_type " + m.media_type + " is not supported by this front end");
    }

    // absorbing layers
    {
        ConfFile c(m.absorb_file, "absorb configure");
        float vel[6], bmax[6], amax[6];
        for (int i = 0; i < 6; ++i) {
            sc.nabs[i] = c.integer("abs_number", 3 + i);
            vel[i] = c.real("abs_velocity", 3 + i);
            bmax[i] = c.real("bmax", 3 + i);
            amax[i] = c.real("amax", 3 + i);
        }
        cfs_pml_tables(ni, nj, nk, sc.nabs, vel, bmax, amax, m.steph, sc.metrics, sc.pml);
    }

    load_sources(m, x, y, z, sc);
    load_stations(m, x, y, z, sc);
    sc.x = std::move(x.v);
    sc.y = std::move(y.v);
    sc.z = std::move(z.v);
    return sc;
}

} // namespace cgfd_conf
