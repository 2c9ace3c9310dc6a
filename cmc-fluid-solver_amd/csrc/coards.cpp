// COARDS result files: the reference's OutputNetCDF3D_header and
// OutputNetCDF3D_layer (src/Common/IO.h:136-388), written as netCDF classic
// with 64-bit offsets (format version 2) directly from the published format
// specification, so that no netCDF library is needed.  The reference creates a
// netCDF-4 (HDF5) file through libnetcdf; the data model here is the same --
// dimensions x, y, z and the unlimited t, coordinate variables x(x), y(y), z(z)
// (float) and time(t) (double), the requested output variables u, v, w, T as
// double (t, x, y, z) and d as float (x, y), with the reference's attributes
// (units, actual_range, valid_range, missing_value 99999, long_name, var_desc)
// and global attributes (Conventions = "COARDS", ...) -- and any netCDF reader
// opens it.
//
// Layout of a classic file: header (magic "CDF\2", record count, dimension,
// global attribute and variable lists, each variable with its dimensions,
// attributes, type, size and file offset), then the fixed-size variables in
// definition order, then the records: per record, each record variable's slice
// in definition order.  Everything big-endian, padded to 4 bytes.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/cmcx.h"

namespace cmcx {
void set_error(const std::string& msg);
}

namespace {

enum : uint32_t { NC_CHAR = 2, NC_FLOAT = 5, NC_DOUBLE = 6 };
enum : uint32_t { TAG_DIMENSION = 10, TAG_VARIABLE = 11, TAG_ATTRIBUTE = 12 };
constexpr double kFillDouble = 9.9692099683868690e+36;  // netCDF's default fill value
constexpr float kMissing = 99999.0f;                    // the reference's MISSING_VALUE

struct Buf {
    std::vector<uint8_t> b;
    void u32(uint32_t v) {
        for (int s = 24; s >= 0; s -= 8) b.push_back((uint8_t)(v >> s));
    }
    void u64(uint64_t v) {
        for (int s = 56; s >= 0; s -= 8) b.push_back((uint8_t)(v >> s));
    }
    void f32(float v) {
        uint32_t u;
        std::memcpy(&u, &v, 4);
        u32(u);
    }
    void f64(double v) {
        uint64_t u;
        std::memcpy(&u, &v, 8);
        u64(u);
    }
    void pad() {
        while (b.size() % 4) b.push_back(0);
    }
    void name(const std::string& s) {
        u32((uint32_t)s.size());
        b.insert(b.end(), s.begin(), s.end());
        pad();
    }
};

struct Att {
    std::string name;
    uint32_t type;
    std::string text;
    std::vector<double> vals;  // FLOAT or DOUBLE values
};

Att text_att(const char* n, const std::string& t) { return Att{n, NC_CHAR, t, {}}; }
Att num_att(const char* n, uint32_t type, std::vector<double> v) { return Att{n, type, "", v}; }

void put_atts(Buf& h, const std::vector<Att>& atts) {
    if (atts.empty()) {
        h.u32(0);
        h.u32(0);
        return;
    }
    h.u32(TAG_ATTRIBUTE);
    h.u32((uint32_t)atts.size());
    for (const Att& a : atts) {
        h.name(a.name);
        h.u32(a.type);
        if (a.type == NC_CHAR) {
            h.u32((uint32_t)a.text.size());
            h.b.insert(h.b.end(), a.text.begin(), a.text.end());
        } else {
            h.u32((uint32_t)a.vals.size());
            for (double v : a.vals) a.type == NC_FLOAT ? h.f32((float)v) : h.f64(v);
        }
        h.pad();
    }
}

struct Var {
    std::string name;
    std::vector<uint32_t> dims;  // dimension ids
    std::vector<Att> atts;
    uint32_t type;
    uint64_t vsize = 0;  // bytes of the variable (per record for record variables)
    uint64_t begin = 0;
    bool record = false;
};

// Output variables of the reference, in its order (IO.h:138-139)
const char* kShort[5] = {"u", "v", "w", "T", "d"};
const char* kLong[5] = {"x-velocity", "y-velocity", "z-velocity", "temperature", "depth"};

std::vector<std::string> split_vars(const char* s) {
    std::vector<std::string> out;
    std::string cur;
    for (const char* p = s ? s : ""; ; p++) {
        if (*p == ' ' || *p == ',' || *p == '\t' || *p == 0) {
            if (!cur.empty()) out.push_back(cur);
            cur.clear();
            if (*p == 0) break;
        } else {
            cur.push_back(*p);
        }
    }
    return out;
}

bool has(const std::vector<std::string>& v, const char* s) {
    for (const auto& x : v)
        if (x == s) return true;
    return false;
}

// Header of the file (begin offsets filled by the caller's second pass).
void put_header(Buf& h, uint32_t numrecs, const uint32_t dimlen[4], const std::vector<Var>& vars,
                const std::vector<Att>& gatts) {
    const char* dn[4] = {"x", "y", "z", "t"};
    h.b.insert(h.b.end(), {'C', 'D', 'F', 2});
    h.u32(numrecs);
    h.u32(TAG_DIMENSION);
    h.u32(4);
    for (int d = 0; d < 4; d++) {
        h.name(dn[d]);
        h.u32(dimlen[d]);
    }
    put_atts(h, gatts);
    h.u32(TAG_VARIABLE);
    h.u32((uint32_t)vars.size());
    for (const Var& v : vars) {
        h.name(v.name);
        h.u32((uint32_t)v.dims.size());
        for (uint32_t d : v.dims) h.u32(d);
        put_atts(h, v.atts);
        h.u32(v.type);
        h.u32((uint32_t)std::min<uint64_t>(v.vsize, 0xffffffffu));
        h.u64(v.begin);
    }
}

// What a layer write needs from a file's header.
struct Parsed {
    uint32_t numrecs = 0;
    uint64_t numrecs_pos = 4;
    std::vector<uint32_t> dimlen;
    struct V {
        std::string name;
        std::vector<uint32_t> dims;
        uint32_t type;
        uint64_t vsize, begin;
    };
    std::vector<V> vars;
    uint64_t recsize = 0;
};

struct Reader {
    FILE* f;
    uint32_t u32() {
        uint8_t b[4];
        if (fread(b, 1, 4, f) != 4) throw std::runtime_error("truncated header");
        return (uint32_t)b[0] << 24 | (uint32_t)b[1] << 16 | (uint32_t)b[2] << 8 | b[3];
    }
    uint64_t u64() {
        uint64_t hi = u32();
        return hi << 32 | u32();
    }
    std::string name() {
        uint32_t n = u32();
        std::string s(n, '\0');
        if (n && fread(&s[0], 1, n, f) != n) throw std::runtime_error("truncated header");
        skip_pad(n);
        return s;
    }
    void skip_pad(uint64_t n) {
        const long p = (long)((4 - n % 4) % 4);
        if (p) fseek(f, p, SEEK_CUR);
    }
    void skip_atts() {
        uint32_t tag = u32(), n = u32();
        if (tag == 0 && n == 0) return;
        if (tag != TAG_ATTRIBUTE) throw std::runtime_error("bad attribute list");
        for (uint32_t i = 0; i < n; i++) {
            name();
            uint32_t type = u32(), cnt = u32();
            const uint64_t sz = type == NC_CHAR ? 1 : type == NC_FLOAT ? 4 : type == NC_DOUBLE ? 8 : 0;
            if (!sz) throw std::runtime_error("unexpected attribute type");
            fseek(f, (long)(sz * cnt), SEEK_CUR);
            skip_pad(sz * cnt);
        }
    }
};

Parsed parse(FILE* f) {
    Reader r{f};
    Parsed p;
    char magic[4];
    if (fread(magic, 1, 4, f) != 4 || std::memcmp(magic, "CDF\2", 4) != 0)
        throw std::runtime_error("not a netCDF classic (64-bit offset) file");
    p.numrecs = r.u32();
    uint32_t tag = r.u32(), n = r.u32();
    if (tag != TAG_DIMENSION) throw std::runtime_error("bad dimension list");
    for (uint32_t i = 0; i < n; i++) {
        r.name();
        p.dimlen.push_back(r.u32());
    }
    r.skip_atts();
    tag = r.u32();
    n = r.u32();
    if (tag != TAG_VARIABLE) throw std::runtime_error("bad variable list");
    for (uint32_t i = 0; i < n; i++) {
        Parsed::V v;
        v.name = r.name();
        uint32_t nd = r.u32();
        for (uint32_t d = 0; d < nd; d++) v.dims.push_back(r.u32());
        r.skip_atts();
        v.type = r.u32();
        v.vsize = r.u32();
        v.begin = r.u64();
        const bool rec = !v.dims.empty() && p.dimlen.at(v.dims[0]) == 0;
        if (rec) p.recsize += v.vsize;
        p.vars.push_back(v);
    }
    return p;
}

bool is_record(const Parsed& p, const Parsed::V& v) {
    return !v.dims.empty() && p.dimlen.at(v.dims[0]) == 0;
}

void write_be_doubles(FILE* f, const double* x, size_t n) {
    std::vector<uint8_t> b(n * 8);
    for (size_t i = 0; i < n; i++) {
        uint64_t u;
        std::memcpy(&u, &x[i], 8);
        for (int k = 0; k < 8; k++) b[8 * i + k] = (uint8_t)(u >> (56 - 8 * k));
    }
    if (fwrite(b.data(), 1, b.size(), f) != b.size()) throw std::runtime_error("write failed");
}

int fail(int code, const std::string& msg) {
    cmcx::set_error(msg);
    return code;
}

}  // namespace

extern "C" int cmcx_coards_write_header(const char* path, const cmcx_coards_header* hd) {
    if (!path || !hd) return CMCX_ERR_ARG;
    for (int a = 0; a < 3; a++)
        if (hd->outdims[a] < 1) return fail(CMCX_ERR_ARG, "cmcx_coards_write_header: output dims must be >= 1");
    if (!(hd->timestep > 0) || !(hd->time >= 0))
        return fail(CMCX_ERR_ARG, "cmcx_coards_write_header: timestep must be > 0 and time >= 0");
    const std::vector<std::string> req = split_vars(hd->vars);
    for (const auto& v : req)
        if (v != "u" && v != "v" && v != "w" && v != "T" && v != "d")
            return fail(CMCX_ERR_ARG, "cmcx_coards_write_header: unknown output variable '" + v + "'");
    const uint64_t nx = hd->outdims[0], ny = hd->outdims[1], nz = hd->outdims[2];
    if (8 * nx * ny * nz > 0xffffffffull)
        return fail(CMCX_ERR_TOO_LARGE, "cmcx_coards_write_header: a record slice must be below 4 GiB");
    bool use[5];
    for (int i = 0; i < 5; i++) use[i] = has(req, kShort[i]);
    if (use[4] && !hd->depth)
        return fail(CMCX_ERR_ARG, "cmcx_coards_write_header: variable d needs the depth array");

    // variables in the reference's definition order: x, y, z, time, u, v, w, T, d
    std::vector<Var> vars;
    const bool deg = hd->xy_degree_units != 0;
    vars.push_back({"x", {0},
                    {num_att("actual_range", NC_FLOAT, {hd->bbox_min[0], hd->bbox_max[0]}),
                     text_att("long_name", "x coord"), text_att("units", deg ? "degree_north" : "metres")},
                    NC_FLOAT});
    vars.push_back({"y", {1},
                    {num_att("actual_range", NC_FLOAT, {hd->bbox_min[1], hd->bbox_max[1]}),
                     text_att("long_name", "y coord"), text_att("units", deg ? "degree_east" : "metres")},
                    NC_FLOAT});
    vars.push_back({"z", {2},
                    {text_att("units", "metres"),
                     num_att("actual_range", NC_FLOAT, {hd->bbox_min[2], hd->bbox_max[2]}),
                     text_att("long_name", "z coord")},
                    NC_FLOAT});
    vars.push_back({"time", {3},
                    {text_att("units", "s"), num_att("actual_range", NC_DOUBLE, {0.0, hd->time}),
                     text_att("long_name", "time")},
                    NC_DOUBLE});
    for (int i = 0; i < 5; i++) {
        if (!use[i]) continue;
        Var v;
        v.name = kShort[i];
        v.type = i == 4 ? NC_FLOAT : NC_DOUBLE;
        v.dims = i == 4 ? std::vector<uint32_t>{0, 1} : std::vector<uint32_t>{3, 0, 1, 2};
        const char* units = i == 3 ? "tmp" : i == 4 ? "m" : "m/s";
        v.atts = {text_att("units", units), num_att("actual_range", NC_DOUBLE, {-1.0, 1.0}),
                  num_att("valid_range", NC_DOUBLE, {-1.0, 1.0}),
                  num_att("missing_value", NC_FLOAT, {kMissing}), text_att("long_name", kLong[i]),
                  text_att("var_desc", kShort[i])};
        vars.push_back(v);
    }
    const uint64_t dimlen[3] = {nx, ny, nz};
    for (Var& v : vars) {
        v.record = v.dims[0] == 3;
        uint64_t n = v.type == NC_DOUBLE ? 8 : 4;
        for (uint32_t d : v.dims)
            if (d != 3) n *= dimlen[d];
        v.vsize = (n + 3) / 4 * 4;
    }
    std::vector<Att> gatts = {text_att("Conventions", "COARDS"),
                              text_att("title", "cmc-fluid-solver results"),
                              text_att("history", "created by using cmc-fluid-solver"),
                              text_att("description", "Test data"), text_att("platform", "Model")};
    // time values of the records the header announces (IO.h:287-293)
    const uint32_t nrec = (uint32_t)(hd->time / hd->timestep);
    const uint32_t dl[4] = {(uint32_t)nx, (uint32_t)ny, (uint32_t)nz, 0};
    Buf h;
    put_header(h, nrec, dl, vars, gatts);
    uint64_t off = h.b.size();
    for (Var& v : vars)
        if (!v.record) {
            v.begin = off;
            off += v.vsize;
        }
    uint64_t recsize = 0;
    for (Var& v : vars)
        if (v.record) {
            v.begin = off + recsize;
            recsize += v.vsize;
        }
    Buf hh;
    put_header(hh, nrec, dl, vars, gatts);
    if (hh.b.size() != h.b.size()) return fail(CMCX_ERR_ARG, "cmcx_coards_write_header: header size changed");

    FILE* f = fopen(path, "wb");
    if (!f) return fail(CMCX_ERR_ARG, std::string("cmcx_coards_write_header: cannot create ") + path);
    try {
        Buf data;
        data.b = hh.b;
        for (const Var& v : vars) {
            if (v.record) continue;
            if (v.name == "d") {
                for (uint64_t k = 0; k < nx * ny; k++) data.f32(hd->depth[k]);
            } else {
                const int a = v.name == "x" ? 0 : v.name == "y" ? 1 : 2;
                const float lo = (float)hd->bbox_min[a], hi = (float)hd->bbox_max[a];
                const float dd = (float)(hi - lo) / (float)dimlen[a];
                for (uint64_t k = 0; k < dimlen[a]; k++) data.f32(lo + dd * (float)k);
            }
            data.pad();
        }
        if (fwrite(data.b.data(), 1, data.b.size(), f) != data.b.size()) throw std::runtime_error("write failed");
        // records: time = i * timestep, the field slices filled with netCDF's fill value
        std::vector<double> fill(nx * ny * nz, kFillDouble);
        for (uint32_t r = 0; r < nrec; r++)
            for (const Var& v : vars) {
                if (!v.record) continue;
                if (v.name == "time") {
                    const double tv = r * hd->timestep;
                    write_be_doubles(f, &tv, 1);
                } else {
                    write_be_doubles(f, fill.data(), fill.size());
                }
            }
    } catch (const std::exception& e) {
        fclose(f);
        return fail(CMCX_ERR_ARG, std::string("cmcx_coards_write_header: ") + e.what());
    }
    if (fclose(f) != 0) return fail(CMCX_ERR_ARG, "cmcx_coards_write_header: close failed");
    return CMCX_OK;
}

extern "C" int cmcx_coards_write_layer(const char* path, const void* vel, int vel_dtype,
                                       const double* T, int time_index, int dimx, int dimy,
                                       int dimz, const char* vars) {
    if (!path || time_index < 0 || dimx < 1 || dimy < 1 || dimz < 1) return CMCX_ERR_ARG;
    if (vel_dtype != CMCX_F32 && vel_dtype != CMCX_F64) return CMCX_ERR_ARG;
    const std::vector<std::string> req = split_vars(vars);
    FILE* f = fopen(path, "r+b");
    if (!f) return fail(CMCX_ERR_ARG, std::string("cmcx_coards_write_layer: cannot open ") + path);
    try {
        Parsed p = parse(f);
        const uint64_t n = (uint64_t)dimx * dimy * dimz;
        // records up to this one: new records take the fill value (time too)
        if ((uint32_t)time_index >= p.numrecs) {
            std::vector<double> fill;
            for (uint32_t r = p.numrecs; r <= (uint32_t)time_index; r++)
                for (const auto& v : p.vars) {
                    if (!is_record(p, v)) continue;
                    const uint64_t m = v.vsize / 8;
                    fill.assign(m, kFillDouble);
                    fseek(f, (long)(v.begin + (uint64_t)r * p.recsize), SEEK_SET);
                    write_be_doubles(f, fill.data(), m);
                }
            p.numrecs = (uint32_t)time_index + 1;
            fseek(f, (long)p.numrecs_pos, SEEK_SET);
            Buf b;
            b.u32(p.numrecs);
            fwrite(b.b.data(), 1, 4, f);
        }
        std::vector<double> dp(n);
        for (const auto& name : req) {
            if (name == "d") continue;  // written with the header (IO.h:348-349)
            const Parsed::V* v = nullptr;
            for (const auto& x : p.vars)
                if (x.name == name) v = &x;
            if (!v) throw std::runtime_error("variable '" + name + "' is not in the file");
            if (!is_record(p, *v) || v->type != NC_DOUBLE || v->dims.size() != 4 ||
                p.dimlen[v->dims[1]] != (uint32_t)dimx || p.dimlen[v->dims[2]] != (uint32_t)dimy ||
                p.dimlen[v->dims[3]] != (uint32_t)dimz)
                throw std::runtime_error("variable '" + name + "' does not have shape (t, x, y, z) of the layer");
            const int c = name == "u" ? 0 : name == "v" ? 1 : name == "w" ? 2 : 3;
            if (c < 3 && !vel) throw std::runtime_error("velocity array missing");
            if (c == 3 && !T) throw std::runtime_error("temperature array missing");
            for (uint64_t k = 0; k < n; k++)
                dp[k] = c == 3 ? T[k]
                               : vel_dtype == CMCX_F32 ? (double)static_cast<const float*>(vel)[3 * k + c]
                                                       : static_cast<const double*>(vel)[3 * k + c];
            fseek(f, (long)(v->begin + (uint64_t)time_index * p.recsize), SEEK_SET);
            write_be_doubles(f, dp.data(), n);
        }
    } catch (const std::exception& e) {
        fclose(f);
        return fail(CMCX_ERR_ARG, std::string("cmcx_coards_write_layer: ") + e.what());
    }
    if (fclose(f) != 0) return fail(CMCX_ERR_ARG, "cmcx_coards_write_layer: close failed");
    return CMCX_OK;
}
