// Mesh reader / writer behind include/fwn_io.h: OBJ and PLY positions + polygonal facets.
//
// Reading mirrors what lagrange::io::load_mesh_obj / load_mesh_ply keep of a file for a
// triangle-mesh consumer (modules/io/include/lagrange/io/load_mesh_obj.h, load_mesh_ply.h):
// vertex positions and facet corner lists; everything else in the file is skipped.  The parsers
// work on one in-memory buffer (the whole file is read first) with std::from_chars.
#include <fwn_io.h>

#include <algorithm>
#include <cctype>
#include <charconv>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

struct fwn_mesh {
    std::vector<double> V;        // n x 3
    std::vector<int64_t> offsets; // n_facets + 1
    std::vector<int32_t> corners;
};

namespace {

thread_local std::string g_error;

struct IoError : std::runtime_error {
    int status;
    IoError(int s, const std::string& m) : std::runtime_error(m), status(s) {}
};

[[noreturn]] void parse_fail(const std::string& m) { throw IoError(FWN_IO_PARSE_ERROR, m); }

template <typename F>
int guarded(F&& f)
{
    try {
        f();
        return FWN_IO_OK;
    } catch (const IoError& e) {
        g_error = e.what();
        return e.status;
    } catch (const std::bad_alloc&) {
        g_error = "out of memory";
        return FWN_IO_FILE_ERROR;
    } catch (const std::exception& e) {
        g_error = e.what();
        return FWN_IO_PARSE_ERROR;
    }
}

// Append facet `c` (vertex ids) to m, as a fan of triangles when triangulate is set.
void add_facet(fwn_mesh& m, const std::vector<int32_t>& c, bool triangulate)
{
    if (c.size() < 3) parse_fail("facet with fewer than 3 corners");
    if (!triangulate || c.size() == 3) {
        m.corners.insert(m.corners.end(), c.begin(), c.end());
        m.offsets.push_back((int64_t)m.corners.size());
        return;
    }
    for (size_t i = 1; i + 1 < c.size(); ++i) {
        m.corners.push_back(c[0]);
        m.corners.push_back(c[i]);
        m.corners.push_back(c[i + 1]);
        m.offsets.push_back((int64_t)m.corners.size());
    }
}

// ---- OBJ ------------------------------------------------------------------------------------
inline bool is_space(char c) { return c == ' ' || c == '\t' || c == '\r'; }

const char* skip_space(const char* p, const char* e)
{
    while (p < e && is_space(*p)) ++p;
    return p;
}

template <typename T>
const char* parse_num(const char* p, const char* e, T& v, const char* what, int64_t line)
{
    if (p < e && *p == '+') ++p; // from_chars takes no leading '+'
    auto r = std::from_chars(p, e, v);
    if (r.ec != std::errc()) parse_fail(std::string("bad ") + what + " on line " + std::to_string(line));
    return r.ptr;
}

void parse_obj(const char* p, const char* e, bool triangulate, fwn_mesh& m)
{
    std::vector<int32_t> facet;
    int64_t line = 0;
    m.offsets.assign(1, 0);
    while (p < e) {
        ++line;
        const char* eol = static_cast<const char*>(std::memchr(p, '\n', (size_t)(e - p)));
        if (!eol) eol = e;
        const char* q = skip_space(p, eol);
        if (q + 1 < eol && q[0] == 'v' && is_space(q[1])) {
            double xyz[3];
            q += 2;
            for (double& x : xyz) {
                q = skip_space(q, eol);
                q = parse_num(q, eol, x, "vertex coordinate", line);
            }
            m.V.insert(m.V.end(), xyz, xyz + 3);
        } else if (q + 1 < eol && q[0] == 'f' && is_space(q[1])) {
            q += 2;
            facet.clear();
            const int64_t nv = (int64_t)(m.V.size() / 3);
            for (;;) {
                q = skip_space(q, eol);
                if (q >= eol || *q == '#') break;
                int64_t idx = 0;
                q = parse_num(q, eol, idx, "facet index", line);
                // skip /t, //n, /t/n
                while (q < eol && !is_space(*q)) ++q;
                if (idx == 0) parse_fail("facet index 0 on line " + std::to_string(line));
                const int64_t v = idx > 0 ? idx - 1 : nv + idx;
                if (v < 0 || v >= nv) parse_fail("facet index out of range on line " + std::to_string(line));
                if (v > INT32_MAX) parse_fail("more than 2^31 vertices");
                facet.push_back((int32_t)v);
            }
            add_facet(m, facet, triangulate);
        }
        // every other record (comments, vt, vn, o, g, usemtl, mtllib, s, l, p) is skipped
        p = eol < e ? eol + 1 : e;
    }
}

// ---- PLY ------------------------------------------------------------------------------------
enum class PlyType { I8, U8, I16, U16, I32, U32, F32, F64 };

PlyType ply_type(const std::string& s)
{
    if (s == "char" || s == "int8") return PlyType::I8;
    if (s == "uchar" || s == "uint8") return PlyType::U8;
    if (s == "short" || s == "int16") return PlyType::I16;
    if (s == "ushort" || s == "uint16") return PlyType::U16;
    if (s == "int" || s == "int32") return PlyType::I32;
    if (s == "uint" || s == "uint32") return PlyType::U32;
    if (s == "float" || s == "float32") return PlyType::F32;
    if (s == "double" || s == "float64") return PlyType::F64;
    parse_fail("unknown PLY type '" + s + "'");
}

int ply_size(PlyType t)
{
    switch (t) {
    case PlyType::I8:
    case PlyType::U8: return 1;
    case PlyType::I16:
    case PlyType::U16: return 2;
    case PlyType::I32:
    case PlyType::U32:
    case PlyType::F32: return 4;
    case PlyType::F64: return 8;
    }
    return 0;
}

struct PlyProp {
    std::string name;
    bool list = false;
    PlyType count_type = PlyType::U8, type = PlyType::F32;
};
struct PlyElement {
    std::string name;
    int64_t count = 0;
    std::vector<PlyProp> props;
};

// Reads values of a PLY body: text tokens or binary scalars of either byte order.
bool host_little_endian()
{
    const uint16_t x = 1;
    unsigned char c;
    std::memcpy(&c, &x, 1);
    return c == 1;
}

struct PlyReader {
    const char* p;
    const char* e;
    int mode;  // 0 ascii, 1 little endian, 2 big endian
    bool swap; // binary byte order differs from the host's

    double read(PlyType t)
    {
        if (mode == 0) {
            while (p < e && std::isspace((unsigned char)*p)) ++p;
            if (p >= e) parse_fail("PLY body ends early");
            double v = 0;
            if (t == PlyType::F32 || t == PlyType::F64) {
                if (*p == '+') ++p;
                auto r = std::from_chars(p, e, v);
                if (r.ec != std::errc()) parse_fail("bad PLY number");
                p = r.ptr;
            } else {
                int64_t i = 0;
                if (*p == '+') ++p;
                auto r = std::from_chars(p, e, i);
                if (r.ec != std::errc()) parse_fail("bad PLY integer");
                p = r.ptr;
                v = (double)i;
            }
            return v;
        }
        const int n = ply_size(t);
        if (e - p < n) parse_fail("PLY body ends early");
        unsigned char b[8];
        std::memcpy(b, p, (size_t)n);
        p += n;
        if (swap) std::reverse(b, b + n);
        switch (t) {
        case PlyType::I8: { int8_t v; std::memcpy(&v, b, 1); return v; }
        case PlyType::U8: { uint8_t v; std::memcpy(&v, b, 1); return v; }
        case PlyType::I16: { int16_t v; std::memcpy(&v, b, 2); return v; }
        case PlyType::U16: { uint16_t v; std::memcpy(&v, b, 2); return v; }
        case PlyType::I32: { int32_t v; std::memcpy(&v, b, 4); return v; }
        case PlyType::U32: { uint32_t v; std::memcpy(&v, b, 4); return v; }
        case PlyType::F32: { float v; std::memcpy(&v, b, 4); return v; }
        case PlyType::F64: { double v; std::memcpy(&v, b, 8); return v; }
        }
        return 0;
    }
};

void parse_ply(const char* p, const char* e, bool triangulate, fwn_mesh& m)
{
    // header
    std::vector<PlyElement> elems;
    int mode = -1;
    bool first = true;
    for (;;) {
        const char* eol = static_cast<const char*>(std::memchr(p, '\n', (size_t)(e - p)));
        if (!eol) parse_fail("PLY header without end_header");
        std::string line(p, eol);
        if (!line.empty() && line.back() == '\r') line.pop_back();
        p = eol + 1;
        std::vector<std::string> tok;
        for (size_t i = 0; i < line.size();) {
            while (i < line.size() && std::isspace((unsigned char)line[i])) ++i;
            size_t j = i;
            while (j < line.size() && !std::isspace((unsigned char)line[j])) ++j;
            if (j > i) tok.emplace_back(line.substr(i, j - i));
            i = j;
        }
        if (first) {
            if (tok.size() != 1 || tok[0] != "ply") parse_fail("not a PLY file");
            first = false;
            continue;
        }
        if (tok.empty() || tok[0] == "comment" || tok[0] == "obj_info") continue;
        if (tok[0] == "end_header") break;
        if (tok[0] == "format") {
            if (tok.size() < 2) parse_fail("bad PLY format line");
            if (tok[1] == "ascii") mode = 0;
            else if (tok[1] == "binary_little_endian") mode = 1;
            else if (tok[1] == "binary_big_endian") mode = 2;
            else parse_fail("unknown PLY format " + tok[1]);
        } else if (tok[0] == "element") {
            if (tok.size() != 3) parse_fail("bad PLY element line");
            PlyElement el;
            el.name = tok[1];
            el.count = std::stoll(tok[2]);
            if (el.count < 0) parse_fail("negative PLY element count");
            elems.push_back(el);
        } else if (tok[0] == "property") {
            if (elems.empty()) parse_fail("PLY property before any element");
            PlyProp pr;
            if (tok.size() == 5 && tok[1] == "list") {
                pr.list = true;
                pr.count_type = ply_type(tok[2]);
                pr.type = ply_type(tok[3]);
                pr.name = tok[4];
            } else if (tok.size() == 3) {
                pr.type = ply_type(tok[1]);
                pr.name = tok[2];
            } else {
                parse_fail("bad PLY property line");
            }
            elems.back().props.push_back(pr);
        } else {
            parse_fail("unknown PLY header keyword " + tok[0]);
        }
    }
    if (mode < 0) parse_fail("PLY header without format");
    PlyReader r{p, e, mode, mode != 0 && (mode == 1) != host_little_endian()};
    m.offsets.assign(1, 0);
    std::vector<int32_t> facet;
    int64_t n_vertices = -1;
    for (const PlyElement& el : elems) {
        const bool is_vertex = el.name == "vertex", is_face = el.name == "face";
        int ix = -1, iy = -1, iz = -1, iface = -1;
        for (int k = 0; k < (int)el.props.size(); ++k) {
            const PlyProp& pr = el.props[k];
            if (is_vertex && !pr.list) {
                if (pr.name == "x") ix = k;
                if (pr.name == "y") iy = k;
                if (pr.name == "z") iz = k;
            }
            if (is_face && pr.list && (pr.name == "vertex_indices" || pr.name == "vertex_index")) iface = k;
        }
        if (is_vertex) {
            if (ix < 0 || iy < 0 || iz < 0) parse_fail("PLY vertex element without x, y, z");
            m.V.resize((size_t)el.count * 3);
            n_vertices = el.count;
        }
        if (is_face && iface < 0) parse_fail("PLY face element without vertex_indices");
        for (int64_t i = 0; i < el.count; ++i) {
            for (int k = 0; k < (int)el.props.size(); ++k) {
                const PlyProp& pr = el.props[k];
                if (!pr.list) {
                    const double v = r.read(pr.type);
                    if (is_vertex) {
                        if (k == ix) m.V[3 * i] = v;
                        else if (k == iy) m.V[3 * i + 1] = v;
                        else if (k == iz) m.V[3 * i + 2] = v;
                    }
                    continue;
                }
                const double cnt = r.read(pr.count_type);
                if (cnt < 0 || cnt != (double)(int64_t)cnt) parse_fail("bad PLY list length");
                const int64_t c = (int64_t)cnt;
                if (k == iface) facet.clear();
                for (int64_t j = 0; j < c; ++j) {
                    const double v = r.read(pr.type);
                    if (k == iface) {
                        if (n_vertices < 0) parse_fail("PLY face element before the vertex element");
                        if (v < 0 || v >= (double)n_vertices || v != (double)(int64_t)v)
                            parse_fail("PLY facet index out of range");
                        facet.push_back((int32_t)v);
                    }
                }
                if (k == iface) add_facet(m, facet, triangulate);
            }
        }
    }
}

std::string lower_ext(const std::string& path)
{
    const size_t dot = path.find_last_of('.');
    std::string ext = dot == std::string::npos ? "" : path.substr(dot + 1);
    for (char& c : ext) c = (char)std::tolower((unsigned char)c);
    return ext;
}

void load_into(const char* data, int64_t size, int32_t format, bool triangulate, fwn_mesh& m)
{
    if (format == FWN_IO_FORMAT_OBJ) parse_obj(data, data + size, triangulate, m);
    else if (format == FWN_IO_FORMAT_PLY) parse_ply(data, data + size, triangulate, m);
    else throw IoError(FWN_IO_INVALID_ARGUMENT, "format must be OBJ or PLY");
    if (m.V.size() / 3 > (size_t)INT32_MAX) parse_fail("more than 2^31 vertices");
}

// ---- writers --------------------------------------------------------------------------------
void check_mesh_args(const double* V, int64_t nv, const int64_t* off, const int32_t* corners, int64_t nf)
{
    if (nv < 0 || nf < 0) throw IoError(FWN_IO_INVALID_ARGUMENT, "negative size");
    if ((nv > 0 && !V) || (nf > 0 && (!off || !corners))) throw IoError(FWN_IO_INVALID_ARGUMENT, "NULL array");
    if (nf > 0 && off[0] != 0) throw IoError(FWN_IO_INVALID_ARGUMENT, "offsets[0] must be 0");
    for (int64_t f = 0; f < nf; ++f) {
        if (off[f + 1] - off[f] < 3) throw IoError(FWN_IO_INVALID_ARGUMENT, "facet with fewer than 3 corners");
        for (int64_t c = off[f]; c < off[f + 1]; ++c)
            if (corners[c] < 0 || corners[c] >= nv) throw IoError(FWN_IO_INVALID_ARGUMENT, "corner index out of range");
    }
}

void put_double(std::string& s, double v)
{
    char buf[40];
    const int n = std::snprintf(buf, sizeof(buf), "%.17g", v);
    s.append(buf, (size_t)n);
}

std::string write_obj(const double* V, int64_t nv, const int64_t* off, const int32_t* corners, int64_t nf)
{
    std::string s;
    s.reserve((size_t)(nv * 64 + nf * 32));
    for (int64_t i = 0; i < nv; ++i) {
        s += "v ";
        put_double(s, V[3 * i]);
        s += ' ';
        put_double(s, V[3 * i + 1]);
        s += ' ';
        put_double(s, V[3 * i + 2]);
        s += '\n';
    }
    for (int64_t f = 0; f < nf; ++f) {
        s += 'f';
        for (int64_t c = off[f]; c < off[f + 1]; ++c) {
            s += ' ';
            s += std::to_string((int64_t)corners[c] + 1);
        }
        s += '\n';
    }
    return s;
}

std::string write_ply(const double* V, int64_t nv, const int64_t* off, const int32_t* corners, int64_t nf,
                      bool binary)
{
    std::string s = "ply\nformat ";
    s += binary ? "binary_little_endian" : "ascii";
    s += " 1.0\nelement vertex " + std::to_string(nv) +
         "\nproperty double x\nproperty double y\nproperty double z\nelement face " + std::to_string(nf) +
         "\nproperty list uchar int vertex_indices\nend_header\n";
    for (int64_t f = 0; f < nf; ++f)
        if (off[f + 1] - off[f] > 255) throw IoError(FWN_IO_INVALID_ARGUMENT, "PLY facet with more than 255 corners");
    if (!binary) {
        for (int64_t i = 0; i < nv; ++i) {
            put_double(s, V[3 * i]);
            s += ' ';
            put_double(s, V[3 * i + 1]);
            s += ' ';
            put_double(s, V[3 * i + 2]);
            s += '\n';
        }
        for (int64_t f = 0; f < nf; ++f) {
            s += std::to_string(off[f + 1] - off[f]);
            for (int64_t c = off[f]; c < off[f + 1]; ++c) {
                s += ' ';
                s += std::to_string(corners[c]);
            }
            s += '\n';
        }
        return s;
    }
    // binary little endian (this writer runs on little-endian hosts only)
    const uint16_t probe = 1;
    unsigned char lo;
    std::memcpy(&lo, &probe, 1);
    if (lo != 1) throw IoError(FWN_IO_INVALID_ARGUMENT, "binary PLY writing needs a little-endian host");
    s.append(reinterpret_cast<const char*>(V), sizeof(double) * 3 * (size_t)nv);
    for (int64_t f = 0; f < nf; ++f) {
        s += (char)(unsigned char)(off[f + 1] - off[f]);
        s.append(reinterpret_cast<const char*>(corners + off[f]), sizeof(int32_t) * (size_t)(off[f + 1] - off[f]));
    }
    return s;
}

} // namespace

extern "C" {

int fwn_mesh_load(const char* path, int32_t format, int32_t triangulate, fwn_mesh** out)
{
    return guarded([&] {
        if (!path || !out) throw IoError(FWN_IO_INVALID_ARGUMENT, "NULL path or out");
        *out = nullptr;
        if (format == FWN_IO_FORMAT_AUTO) {
            const std::string ext = lower_ext(path);
            if (ext == "obj") format = FWN_IO_FORMAT_OBJ;
            else if (ext == "ply") format = FWN_IO_FORMAT_PLY;
            else throw IoError(FWN_IO_INVALID_ARGUMENT, std::string("unknown mesh file extension: ") + path);
        }
        std::ifstream f(path, std::ios::binary);
        if (!f) throw IoError(FWN_IO_FILE_ERROR, std::string("cannot open ") + path);
        f.seekg(0, std::ios::end);
        const std::streamoff size = f.tellg();
        f.seekg(0, std::ios::beg);
        std::string data((size_t)std::max<std::streamoff>(size, 0), '\0');
        if (size > 0 && !f.read(&data[0], size)) throw IoError(FWN_IO_FILE_ERROR, std::string("cannot read ") + path);
        auto m = std::make_unique<fwn_mesh>();
        load_into(data.data(), (int64_t)data.size(), format, triangulate != 0, *m);
        *out = m.release();
    });
}

int fwn_mesh_load_buffer(const char* data, int64_t size, int32_t format, int32_t triangulate, fwn_mesh** out)
{
    return guarded([&] {
        if (!out || size < 0 || (size > 0 && !data)) throw IoError(FWN_IO_INVALID_ARGUMENT, "bad buffer arguments");
        *out = nullptr;
        auto m = std::make_unique<fwn_mesh>();
        load_into(data, size, format, triangulate != 0, *m);
        *out = m.release();
    });
}

int fwn_mesh_sizes(const fwn_mesh* m, int64_t* nv, int64_t* nf, int64_t* nc)
{
    return guarded([&] {
        if (!m) throw IoError(FWN_IO_INVALID_ARGUMENT, "NULL mesh");
        if (nv) *nv = (int64_t)(m->V.size() / 3);
        if (nf) *nf = (int64_t)m->offsets.size() - 1;
        if (nc) *nc = (int64_t)m->corners.size();
    });
}

int fwn_mesh_copy(const fwn_mesh* m, double* V, int64_t* offsets, int32_t* corners)
{
    return guarded([&] {
        if (!m) throw IoError(FWN_IO_INVALID_ARGUMENT, "NULL mesh");
        if (V) std::copy(m->V.begin(), m->V.end(), V);
        if (offsets) std::copy(m->offsets.begin(), m->offsets.end(), offsets);
        if (corners) std::copy(m->corners.begin(), m->corners.end(), corners);
    });
}

int fwn_mesh_free(fwn_mesh* m)
{
    delete m;
    return FWN_IO_OK;
}

int fwn_mesh_save(const char* path, int32_t format, const double* V, int64_t nv, const int64_t* offsets,
                  const int32_t* corners, int64_t nf)
{
    return guarded([&] {
        if (!path) throw IoError(FWN_IO_INVALID_ARGUMENT, "NULL path");
        check_mesh_args(V, nv, offsets, corners, nf);
        std::string s;
        if (format == FWN_IO_FORMAT_OBJ) s = write_obj(V, nv, offsets, corners, nf);
        else if (format == FWN_IO_FORMAT_PLY) s = write_ply(V, nv, offsets, corners, nf, false);
        else if (format == FWN_IO_FORMAT_PLY_BINARY) s = write_ply(V, nv, offsets, corners, nf, true);
        else throw IoError(FWN_IO_INVALID_ARGUMENT, "format must be OBJ, PLY or PLY_BINARY");
        std::ofstream f(path, std::ios::binary);
        if (!f || !f.write(s.data(), (std::streamsize)s.size()))
            throw IoError(FWN_IO_FILE_ERROR, std::string("cannot write ") + path);
    });
}

const char* fwn_io_last_error(void) { return g_error.c_str(); }

} // extern "C"
