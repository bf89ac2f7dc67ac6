// Autodesk FBX models (binary or ASCII FBX, version 7100 and later) as the reference
// loads them through assimp (Model::Model, src/model.cpp:172-215): the
// importer runs aiProcess_Triangulate | aiProcess_PreTransformVertices (with
// AI_CONFIG_PP_PTV_KEEP_HIERARCHY = 1) | aiProcess_SortByPType |
// aiProcess_FixInfacingNormals, then the reference takes per material its
// name, opacity and texture paths (Model::processMaterial,
// src/model.cpp:124-160; Material::loadMaterialProperties,
// src/material.cpp:300-333) and per mesh its vertices, first uv channel and
// normals (Model::processMesh, src/model.cpp:88-122).
//
// assimp is a third-party dependency that the reference tree neither vendors
// nor pins to a version.  This file restates its published FBX importer and
// post-processing steps:
//   - the binary record format: a 27-byte header ("Kaydara FBX Binary  ",
//     0x1A 0x00, version), node records (end offset, property count,
//     property bytes, name) with 32-bit fields before version 7500 and
//     64-bit ones from 7500, nested lists closed by an all-zero record,
//     typed properties, arrays raw or zlib-deflated;
//   - the ASCII form of the same tree: "Name: value, value, ... { children }"
//     with ';' comments, numbers, quoted strings, bare words and arrays
//     "*count { a: v, v, ... }", object names "Class::Name" where the binary
//     form has "Name\x00\x01Class", and the version in
//     FBXHeaderExtension / FBXVersion;
//   - objects and connections: a Model's geometries, its materials (in
//     connection order) and its child models; a Material's textures by the
//     property they connect to;
//   - per geometry, converted once: every polygon vertex its own vertex;
//     the normal and uv layers the Layer records name (typed index 0)
//     resolved ByPolygonVertex / ByVertice, Direct / IndexToDirect; the
//     material layer AllSame or ByPolygon; one mesh per distinct material
//     index in order of first appearance among the polygons;
//   - materials converted in order of first use by a mesh; a default
//     material for a material index the model has no material for;
//   - the node transform T * Roff * Rp * Rpre * R * Rpost^-1 * Rp^-1 * Soff *
//     Sp * S * Sp^-1 with Euler rotations in the model's rotation order, and
//     for the node's own meshes the geometric transform Gt * Gr * Gs after
//     it; properties a model leaves out come from the file's property
//     templates;
//   - PreTransformVertices with the hierarchy kept: each mesh moved to world
//     space by its node's transform (normals by the inverse transpose, then
//     normalised; polygon winding reversed under a mirroring transform); a
//     mesh that another node uses with a different transform is copied to
//     the end of the mesh list;
//   - Triangulate: a quad is split along the diagonal from its concave
//     corner (the first corner whose two angles to that diagonal add up to
//     more than pi), else from corner 0; larger polygons are ear-clipped in
//     the plane of their Newell normal;
//   - SortByPType: points and lines leave the triangle meshes; they are
//     dropped here (the reference reads three indices of every face);
//   - FixInfacingNormals: a mesh whose bounding box (of its vertices each
//     moved by its normal) has a smaller volume than the box of its plain
//     vertices, and which is not planar, gets its normals negated and its
//     triangles reversed.
// The transforms are evaluated in double precision and rounded to float
// once, where assimp composes float matrices, so positions can differ from
// the reference's in their last bits; the split of polygons with more than
// four corners can differ as well.
#include <zlib.h>
#include <algorithm>
#include <cctype>
#include <cerrno>
#include <cstdlib>
#include <array>
#include <cmath>
#include <cstring>
#include <string>
#include <unordered_map>
#include <vector>
#include "assets.h"
#include "rm_internal.h"

namespace rm {
namespace {

struct FbxProp {
    char type = 0;
    double num = 0.0;        // Y C I L F D
    int64_t iv = 0;          // Y C I L, exactly (object ids are 64-bit)
    std::string str;         // S R
    std::vector<double> f;   // f d arrays
    std::vector<int64_t> i;  // i l b arrays
    bool numeric() const { return strchr("YCILFD", type) != nullptr && type != 0; }
};

struct FbxNode {
    std::string name;
    std::vector<FbxProp> props;
    std::vector<FbxNode> kids;
    const FbxNode *kid(const char *n) const {
        for (const FbxNode &k : kids)
            if (k.name == n) return &k;
        return nullptr;
    }
    // first property as a string ("" when absent)
    std::string str0() const { return props.empty() ? std::string() : props[0].str; }
};

constexpr int FBX_MAX_DEPTH = 64;

class FbxReader {
  public:
    explicit FbxReader(const std::vector<uint8_t> &b) : buf(b) {}

    int32_t read(const std::string &path, std::vector<FbxNode> &top) {
        static const char magic[] = "Kaydara FBX Binary  ";
        if (buf.size() < 27 || memcmp(buf.data(), magic, 21) != 0)
            return fail(RM_ERR_INVALID, "%s: not a binary FBX file", path.c_str());
        uint32_t version;
        memcpy(&version, buf.data() + 23, 4);
        if (version < 7100)
            return fail(RM_ERR_UNSUPPORTED, "%s: FBX version %u (7100 or later is read)", path.c_str(), version);
        wide = version >= 7500;
        pos = 27;
        while (have(wide ? 25 : 13)) {
            FbxNode n;
            bool null_rec = false;
            int32_t rc = node(n, null_rec, buf.size(), 0);
            if (rc) {
                const std::string why = rm_last_error();  // fail() rewrites that buffer
                return fail(rc, "%s: %s", path.c_str(), why.c_str());
            }
            if (null_rec) break;
            top.push_back(std::move(n));
        }
        return RM_OK;
    }

  private:
    const std::vector<uint8_t> &buf;
    size_t pos = 0;
    bool wide = false;

    bool have(size_t k) const { return pos <= buf.size() && buf.size() - pos >= k; }
    template <class T>
    T take() {
        T v;
        memcpy(&v, buf.data() + pos, sizeof(T));
        pos += sizeof(T);
        return v;
    }
    uint64_t field() { return wide ? take<uint64_t>() : (uint64_t)take<uint32_t>(); }

    int32_t prop(FbxProp &p) {
        if (!have(1)) return fail(RM_ERR_INVALID, "property past the end of the file");
        const char t = (char)buf[pos++];
        p.type = t;
        switch (t) {
            case 'Y':
                if (!have(2)) break;
                p.iv = take<int16_t>();
                p.num = (double)p.iv;
                return RM_OK;
            case 'C':
                if (!have(1)) break;
                p.iv = buf[pos++] ? 1 : 0;
                p.num = (double)p.iv;
                return RM_OK;
            case 'I':
                if (!have(4)) break;
                p.iv = take<int32_t>();
                p.num = (double)p.iv;
                return RM_OK;
            case 'F':
                if (!have(4)) break;
                p.num = take<float>();
                return RM_OK;
            case 'D':
                if (!have(8)) break;
                p.num = take<double>();
                return RM_OK;
            case 'L':
                if (!have(8)) break;
                p.iv = take<int64_t>();
                p.num = (double)p.iv;
                return RM_OK;
            case 'S':
            case 'R': {
                if (!have(4)) break;
                const uint32_t len = take<uint32_t>();
                if (!have(len)) break;
                p.str.assign((const char *)buf.data() + pos, len);
                pos += len;
                return RM_OK;
            }
            case 'f':
            case 'd':
            case 'l':
            case 'i':
            case 'b': {
                if (!have(12)) break;
                const uint32_t n = take<uint32_t>(), enc = take<uint32_t>(), clen = take<uint32_t>();
                const size_t es = (t == 'd' || t == 'l') ? 8 : t == 'b' ? 1 : 4;
                if (n > (1u << 30) / es) return fail(RM_ERR_INVALID, "array of %u elements", n);
                const size_t raw = (size_t)n * es;
                if (!have(clen)) break;
                std::vector<uint8_t> tmp;
                const uint8_t *src = buf.data() + pos;
                if (enc == 0) {
                    if (clen != raw) return fail(RM_ERR_INVALID, "raw array of %zu bytes stored in %u", raw, clen);
                } else if (enc == 1) {
                    tmp.resize(raw);
                    uLongf got = (uLongf)raw;
                    if (raw && (uncompress(tmp.data(), &got, src, clen) != Z_OK || got != raw))
                        return fail(RM_ERR_INVALID, "compressed array does not inflate to %zu bytes", raw);
                    src = tmp.data();
                } else {
                    return fail(RM_ERR_INVALID, "array encoding %u", enc);
                }
                pos += clen;
                if (t == 'f' || t == 'd') {
                    p.f.resize(n);
                    for (uint32_t k = 0; k < n; ++k) {
                        if (t == 'f') {
                            float v;
                            memcpy(&v, src + 4 * k, 4);
                            p.f[k] = v;
                        } else {
                            memcpy(&p.f[k], src + 8 * k, 8);
                        }
                    }
                } else {
                    p.i.resize(n);
                    for (uint32_t k = 0; k < n; ++k) {
                        if (t == 'l') {
                            memcpy(&p.i[k], src + 8 * k, 8);
                        } else if (t == 'i') {
                            int32_t v;
                            memcpy(&v, src + 4 * k, 4);
                            p.i[k] = v;
                        } else {
                            p.i[k] = src[k];
                        }
                    }
                }
                return RM_OK;
            }
            default:
                return fail(RM_ERR_INVALID, "unknown property type 0x%02x", (unsigned)(uint8_t)t);
        }
        return fail(RM_ERR_INVALID, "property past the end of the file");
    }

    // One node record ending at most at `limit`; an all-zero record (the end
    // of a nested list) sets null_rec.
    int32_t node(FbxNode &n, bool &null_rec, size_t limit, int depth) {
        const size_t hs = wide ? 25 : 13, start = pos;
        if (depth > FBX_MAX_DEPTH) return fail(RM_ERR_INVALID, "records nested deeper than %d", FBX_MAX_DEPTH);
        if (!have(hs)) return fail(RM_ERR_INVALID, "record past the end of the file");
        const uint64_t end = field(), np = field(), plen = field();
        const uint8_t nlen = buf[pos++];
        null_rec = end == 0 && np == 0 && plen == 0 && nlen == 0;
        if (null_rec) return RM_OK;
        if (end <= start || end > limit) return fail(RM_ERR_INVALID, "record at %zu ends outside its parent", start);
        if (!have(nlen) || pos + nlen > end) return fail(RM_ERR_INVALID, "record name past its end");
        n.name.assign((const char *)buf.data() + pos, nlen);
        pos += nlen;
        const size_t ps = pos;
        if (plen > end - pos || np > plen) return fail(RM_ERR_INVALID, "record %s: property list past its end", n.name.c_str());
        n.props.resize((size_t)np);
        for (FbxProp &p : n.props) {
            int32_t rc = prop(p);
            if (rc) return rc;
            if (pos > ps + plen) return fail(RM_ERR_INVALID, "record %s: property list longer than stated", n.name.c_str());
        }
        if (pos != ps + plen) return fail(RM_ERR_INVALID, "record %s: property list shorter than stated", n.name.c_str());
        while (pos < end) {
            FbxNode k;
            bool nr = false;
            int32_t rc = node(k, nr, end, depth + 1);
            if (rc) return rc;
            if (nr) break;
            n.kids.push_back(std::move(k));
        }
        if (pos != end) return fail(RM_ERR_INVALID, "record %s: nested list does not end at the record's end", n.name.c_str());
        return RM_OK;
    }
};

// The ASCII form: the same node tree as text.
class FbxAsciiReader {
  public:
    explicit FbxAsciiReader(const std::vector<uint8_t> &b) : s((const char *)b.data()), n(b.size()) {}

    int32_t read(const std::string &path, std::vector<FbxNode> &top) {
        int32_t rc = list(top, false, 0);
        if (rc) {
            const std::string why = rm_last_error();  // fail() rewrites that buffer
            return fail(rc, "%s: line %zu: %s", path.c_str(), line, why.c_str());
        }
        return RM_OK;
    }

  private:
    const char *s;
    size_t n, pos = 0, line = 1;

    bool eof() const { return pos >= n; }
    // spaces, tabs and (with nl) line ends and ';' comments
    void skip(bool nl) {
        while (pos < n) {
            const char c = s[pos];
            if (c == ' ' || c == '\t' || c == '\r') {
                ++pos;
            } else if (nl && c == '\n') {
                ++pos;
                ++line;
            } else if (nl && c == ';') {
                while (pos < n && s[pos] != '\n') ++pos;
            } else {
                break;
            }
        }
    }
    static bool word_char(char c) { return isalnum((unsigned char)c) || c == '_' || c == '|' || c == '-' || c == '.'; }

    // one number as a property: integers exactly (ids are 64-bit), others as double
    int32_t number(FbxProp &p) {
        const char *b = s + pos;
        char *e = nullptr;
        size_t k = pos;
        bool real = false;
        if (k < n && (s[k] == '-' || s[k] == '+')) ++k;
        while (k < n && (isdigit((unsigned char)s[k]) || s[k] == '.' || s[k] == 'e' || s[k] == 'E' ||
                         ((s[k] == '-' || s[k] == '+') && (s[k - 1] == 'e' || s[k - 1] == 'E')))) {
            if (!isdigit((unsigned char)s[k])) real = true;
            ++k;
        }
        if (k == pos) return fail(RM_ERR_INVALID, "expected a value");
        const std::string tok(b, k - pos);
        if (!real) {
            errno = 0;
            const long long v = strtoll(tok.c_str(), &e, 10);
            if (*e || errno) return fail(RM_ERR_INVALID, "bad integer '%s'", tok.c_str());
            p.type = 'L';
            p.iv = v;
            p.num = (double)v;
        } else {
            const double v = strtod(tok.c_str(), &e);
            if (*e) return fail(RM_ERR_INVALID, "bad number '%s'", tok.c_str());
            p.type = 'D';
            p.num = v;
        }
        pos = k;
        return RM_OK;
    }

    int32_t value(FbxProp &p) {
        if (eof()) return fail(RM_ERR_INVALID, "value past the end of the file");
        const char c = s[pos];
        if (c == '"') {
            const size_t e = std::string(s + pos + 1, n - pos - 1).find('"');
            if (e == std::string::npos) return fail(RM_ERR_INVALID, "unterminated string");
            p.type = 'S';
            p.str.assign(s + pos + 1, e);
            pos += e + 2;
            return RM_OK;
        }
        if (c == '*') {  // array: *count { a: v, v, ... }
            ++pos;
            FbxProp cnt;
            int32_t rc = number(cnt);
            if (rc) return rc;
            if (cnt.type != 'L' || cnt.iv < 0 || cnt.iv > (1ll << 30)) return fail(RM_ERR_INVALID, "bad array size");
            skip(true);
            if (eof() || s[pos] != '{') return fail(RM_ERR_INVALID, "array without '{'");
            ++pos;
            skip(true);
            if (pos + 1 < n && s[pos] == 'a' && s[pos + 1] == ':') {
                pos += 2;
                skip(true);
            }
            p.type = 'd';
            bool all_int = true;
            p.f.reserve((size_t)cnt.iv);
            while (!eof() && s[pos] != '}') {
                FbxProp v;
                int32_t r2 = number(v);
                if (r2) return r2;
                if (v.type != 'L') all_int = false;
                p.f.push_back(v.num);
                p.i.push_back(v.iv);
                skip(true);
                if (!eof() && s[pos] == ',') {
                    ++pos;
                    skip(true);
                }
            }
            if (eof()) return fail(RM_ERR_INVALID, "unterminated array");
            ++pos;
            if ((int64_t)p.f.size() != cnt.iv)
                return fail(RM_ERR_INVALID, "array of %lld values holds %zu", (long long)cnt.iv, p.f.size());
            if (!all_int) p.i.clear();  // an index array must hold integers
            return RM_OK;
        }
        if (c == '-' || c == '+' || c == '.' || isdigit((unsigned char)c)) return number(p);
        size_t k = pos;
        while (k < n && word_char(s[k])) ++k;
        if (k == pos) return fail(RM_ERR_INVALID, "unexpected '%c'", c);
        p.type = 'S';  // a bare word (T, Y, W, ...)
        p.str.assign(s + pos, k - pos);
        pos = k;
        return RM_OK;
    }

    int32_t node(FbxNode &nd, int depth) {
        size_t k = pos;
        while (k < n && word_char(s[k])) ++k;
        if (k == pos || k >= n || s[k] != ':') return fail(RM_ERR_INVALID, "expected 'Name:'");
        nd.name.assign(s + pos, k - pos);
        pos = k + 1;
        skip(false);
        while (!eof() && s[pos] != '\n' && s[pos] != '{' && s[pos] != '}' && s[pos] != ';') {
            FbxProp p;
            int32_t rc = value(p);
            if (rc) return rc;
            nd.props.push_back(std::move(p));
            skip(false);
            if (!eof() && s[pos] == ',') {
                ++pos;
                skip(true);  // values go on after a comma, across lines
            }
        }
        if (!eof() && s[pos] == '{') {
            ++pos;
            return list(nd.kids, true, depth + 1);
        }
        return RM_OK;
    }

    int32_t list(std::vector<FbxNode> &out, bool braced, int depth) {
        if (depth > FBX_MAX_DEPTH) return fail(RM_ERR_INVALID, "records nested deeper than %d", FBX_MAX_DEPTH);
        while (true) {
            skip(true);
            if (eof()) return braced ? fail(RM_ERR_INVALID, "unterminated '{'") : RM_OK;
            if (s[pos] == '}') {
                if (!braced) return fail(RM_ERR_INVALID, "unmatched '}'");
                ++pos;
                return RM_OK;
            }
            FbxNode nd;
            int32_t rc = node(nd, depth);
            if (rc) return rc;
            out.push_back(std::move(nd));
        }
    }
};

// ---------------------------------------------------------------- transforms
using M4 = std::array<double, 16>;  // row-major; points are columns: p' = M p

M4 ident() { return M4{1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1}; }
M4 mul(const M4 &a, const M4 &b) {
    M4 r{};
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) {
            double s = 0.0;
            for (int k = 0; k < 4; ++k) s += a[i * 4 + k] * b[k * 4 + j];
            r[i * 4 + j] = s;
        }
    return r;
}
M4 translate(const double v[3], double sign = 1.0) {
    M4 m = ident();
    m[3] = sign * v[0];
    m[7] = sign * v[1];
    m[11] = sign * v[2];
    return m;
}
M4 scale(const double v[3]) {
    M4 m = ident();
    m[0] = v[0];
    m[5] = v[1];
    m[10] = v[2];
    return m;
}
M4 rot_axis(int axis, double deg) {
    const double a = deg * (M_PI / 180.0), c = cos(a), s = sin(a);
    M4 m = ident();
    const int i = (axis + 1) % 3, j = (axis + 2) % 3;  // the plane the axis turns
    m[i * 4 + i] = c;
    m[i * 4 + j] = -s;
    m[j * 4 + i] = s;
    m[j * 4 + j] = c;
    return m;
}
// Euler angles (degrees) in FBX rotation order `order` (0 XYZ, 1 XZY, 2 YZX,
// 3 YXZ, 4 ZXY, 5 ZYX: the first letter's rotation is applied first);
// 6 (spherical) is not supported by the importer and gives the identity.
M4 euler(const double deg[3], int order) {
    static const int seq[6][3] = {{2, 1, 0}, {1, 2, 0}, {0, 2, 1}, {2, 0, 1}, {1, 0, 2}, {0, 1, 2}};
    M4 m = ident();
    if (order < 0 || order > 5) return m;
    for (int k = 0; k < 3; ++k) {
        const int ax = seq[order][k];
        if (deg[ax] != 0.0) m = mul(m, rot_axis(ax, deg[ax]));
    }
    return m;
}
M4 transpose(const M4 &a) {
    M4 r{};
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) r[i * 4 + j] = a[j * 4 + i];
    return r;
}
double det3(const M4 &m) {
    return m[0] * (m[5] * m[10] - m[6] * m[9]) - m[1] * (m[4] * m[10] - m[6] * m[8]) +
           m[2] * (m[4] * m[9] - m[5] * m[8]);
}
// inverse transpose of the upper 3x3 (row-major 3x3 in the first 9 entries)
std::array<double, 9> normal_matrix(const M4 &m) {
    const double d = det3(m);
    std::array<double, 9> c{};
    const double a00 = m[0], a01 = m[1], a02 = m[2], a10 = m[4], a11 = m[5], a12 = m[6], a20 = m[8], a21 = m[9],
                 a22 = m[10];
    // cofactor matrix / det = (A^-1)^T
    c[0] = (a11 * a22 - a12 * a21) / d;
    c[1] = -(a10 * a22 - a12 * a20) / d;
    c[2] = (a10 * a21 - a11 * a20) / d;
    c[3] = -(a01 * a22 - a02 * a21) / d;
    c[4] = (a00 * a22 - a02 * a20) / d;
    c[5] = -(a00 * a21 - a01 * a20) / d;
    c[6] = (a01 * a12 - a02 * a11) / d;
    c[7] = -(a00 * a12 - a02 * a10) / d;
    c[8] = (a00 * a11 - a01 * a10) / d;
    return c;
}

// A Properties70 block with the file's property template behind it.
struct Props {
    const FbxNode *own = nullptr, *tmpl = nullptr;
    const FbxNode *find(const char *name) const {
        for (const FbxNode *t : {own, tmpl})
            if (t)
                for (const FbxNode &p : t->kids)
                    if (p.name == "P" && !p.props.empty() && p.props[0].str == name) return &p;
        return nullptr;
    }
    bool num(const char *name, double &v) const {
        const FbxNode *p = find(name);
        if (!p || p->props.size() < 5 || !p->props[4].numeric()) return false;
        v = p->props[4].num;
        return true;
    }
    void vec3(const char *name, double v[3], double def) const {
        v[0] = v[1] = v[2] = def;
        const FbxNode *p = find(name);
        if (!p || p->props.size() < 7) return;
        for (int k = 0; k < 3; ++k)
            if (p->props[4 + k].numeric()) v[k] = p->props[4 + k].num;
    }
};

// -------------------------------------------------------------------- meshes
struct Mesh {
    std::vector<float> pos, nrm, uv;  // per vertex (3, 3, 2 floats)
    bool has_n = false, has_uv = false;
    std::vector<std::vector<uint32_t>> polys;  // vertex indices per polygon
    int material = -1;                         // scene material index
    bool placed = false;
    M4 world{};
    int src = -1;  // the mesh a copy was made of
};

struct Conn {
    bool op;
    int64_t src, dst;
    std::string prop;
};

struct Scene {
    std::unordered_map<int64_t, const FbxNode *> obj;  // Objects children by id
    std::vector<Conn> conns;
    std::unordered_map<int64_t, std::vector<size_t>> to;  // connections by destination, in file order
    std::unordered_map<std::string, const FbxNode *> tmpl;  // "Model.FbxNode" -> Properties70
    const FbxNode *find_tmpl(const std::string &type, const std::string &name) const {
        auto it = tmpl.find(type + "." + name);
        return it == tmpl.end() ? nullptr : it->second;
    }
    const FbxNode *first_tmpl(const std::string &type) const {
        auto it = tmpl.find(type + ".");
        return it == tmpl.end() ? nullptr : it->second;
    }
    // objects of node type `type` connected object-to-object into id
    std::vector<const FbxNode *> sources(int64_t id, const char *type) const {
        std::vector<const FbxNode *> out;
        auto it = to.find(id);
        if (it == to.end()) return out;
        for (size_t c : it->second) {
            const Conn &k = conns[c];
            if (k.op) continue;
            auto o = obj.find(k.src);
            if (o != obj.end() && o->second->name == type) out.push_back(o->second);
        }
        return out;
    }
};

int64_t id_of(const FbxNode &n) { return n.props.empty() ? 0 : n.props[0].iv; }

// "Name\x00\x01Class" (binary) or "Class::Name" (ASCII) -> "Name"
std::string object_name(const FbxNode &n) {
    if (n.props.size() < 2) return std::string();
    const std::string &s = n.props[1].str;
    const size_t k = s.find(std::string("\x00\x01", 2));
    if (k != std::string::npos) return s.substr(0, k);
    const size_t c = s.find("::");
    return c == std::string::npos ? s : s.substr(c + 2);
}

struct Converter {
    const std::string &path;
    Scene &sc;
    std::vector<Mesh> meshes;
    std::unordered_map<int64_t, std::vector<int>> geo_meshes;  // converted geometries
    std::vector<MtlDef> defs;
    std::unordered_map<int64_t, int> mat_index;
    int default_mat = -1;
    struct Use {
        int mesh;
        M4 world;
    };
    std::vector<Use> uses;  // (mesh, node transform) in node order

    Converter(const std::string &p, Scene &s) : path(p), sc(s) {}

    int material_for(const std::vector<const FbxNode *> &mats, int64_t index) {
        if (index < 0 || index >= (int64_t)mats.size()) {
            if (default_mat < 0) {
                MtlDef d;
                d.name = "DefaultMaterial";
                d.m = default_material();
                defs.push_back(d);
                default_mat = (int)defs.size() - 1;
            }
            return default_mat;
        }
        const FbxNode &M = *mats[(size_t)index];
        auto it = mat_index.find(id_of(M));
        if (it != mat_index.end()) return it->second;
        MtlDef d;
        d.name = object_name(M);
        d.m = default_material();
        // template by shading model ("phong" -> FbxSurfacePhong)
        Props pr;
        pr.own = M.kid("Properties70");
        std::string shading;
        if (const FbxNode *s = M.kid("ShadingModel")) shading = s->str0();
        if (!shading.empty()) {
            std::string t = "FbxSurface" + shading;
            for (const auto &kv : sc.tmpl) {
                const std::string &key = kv.first;
                if (key.size() == 9 + t.size() && key.compare(0, 9, "Material.") == 0 &&
                    std::equal(t.begin(), t.end(), key.begin() + 9,
                               [](char a, char b) { return tolower((unsigned char)a) == tolower((unsigned char)b); }))
                    pr.tmpl = kv.second;
            }
        }
        if (!pr.tmpl) pr.tmpl = sc.first_tmpl("Material");
        double opacity = 1.0;
        bool have = pr.num("Opacity", opacity);
        if (!have) {
            // 1 - factor * mean(transparent colour), when that colour is set
            const FbxNode *tc = pr.find("TransparentColor");
            if (tc && tc->props.size() >= 7) {
                double c[3], f = 1.0;
                pr.vec3("TransparentColor", c, 0.0);
                pr.num("TransparencyFactor", f);
                const double o = 1.0 - f * ((c[0] + c[1] + c[2]) / 3.0);
                if ((float)o != 1.0f) {
                    opacity = o;
                    have = true;
                }
            }
        }
        if (have && (float)opacity < 0.99f) glass_by_name(d);
        // textures by the property they connect to; later names win
        static const struct {
            const char *prop;
            int slot;  // 0 diffuse, 1 normals, 2 emissive, 3 specular
        } maps[] = {{"DiffuseColor", 0},       {"EmissiveColor", 2},        {"SpecularColor", 3},
                    {"SpecularFactor", 3},     {"NormalMap", 1},            {"EmissiveFactor", 2},
                    {"Maya|DiffuseTexture", 0}, {"Maya|NormalTexture", 1}, {"Maya|SpecularTexture", 3}};
        auto it2 = sc.to.find(id_of(M));
        for (const auto &mp : maps) {
            if (it2 == sc.to.end()) break;
            for (size_t c : it2->second) {
                const Conn &k = sc.conns[c];
                if (!k.op || k.prop != mp.prop) continue;
                auto o = sc.obj.find(k.src);
                if (o == sc.obj.end() || o->second->name != "Texture") continue;
                std::string file;
                if (const FbxNode *rf = o->second->kid("RelativeFilename")) file = rf->str0();
                std::string *slot = mp.slot == 0 ? &d.map_kd : mp.slot == 1 ? &d.map_norm : mp.slot == 2 ? &d.map_ke : &d.map_ks;
                *slot = file;
            }
        }
        defs.push_back(d);
        const int idx = (int)defs.size() - 1;
        mat_index[id_of(M)] = idx;
        return idx;
    }

    // per-polygon-vertex values of a layer element (C components); false when
    // the element gives no data the importer would use
    int32_t layer(const FbxNode &e, const char *data_name, const char *index_name, int C,
                  const std::vector<uint32_t> &cp_of, size_t ncp, std::vector<float> &out, bool &ok) {
        ok = false;
        const FbxNode *mi = e.kid("MappingInformationType"), *ri = e.kid("ReferenceInformationType");
        const std::string mapping = mi ? mi->str0() : "", ref = ri ? ri->str0() : "";
        const FbxNode *dn = e.kid(data_name), *in = e.kid(index_name);
        bool direct = ref == "Direct", indexed = ref == "IndexToDirect" || ref == "Index";
        if (indexed && !in) direct = true, indexed = false;
        if (!dn || dn->props.empty()) return RM_OK;
        const std::vector<double> &data = dn->props[0].f;
        if (data.size() % C) return fail(RM_ERR_INVALID, "%s: %s holds %zu values, not a multiple of %d", path.c_str(), data_name, data.size(), C);
        const size_t nd = data.size() / C, nv = cp_of.size();
        const std::vector<int64_t> empty;
        const std::vector<int64_t> &ix = (indexed && !in->props.empty()) ? in->props[0].i : empty;
        out.assign(nv * C, 0.0f);
        auto put = [&](size_t v, int64_t d) { for (int c = 0; c < C; ++c) out[v * C + c] = (float)data[(size_t)d * C + c]; };
        if (mapping == "ByVertice" || mapping == "ByVertex") {
            if (direct) {
                if (nd != ncp) return RM_OK;
                for (size_t v = 0; v < nv; ++v) put(v, cp_of[v]);
            } else if (indexed) {
                if (ix.size() != ncp) return RM_OK;
                for (size_t v = 0; v < nv; ++v) {
                    const int64_t d = ix[cp_of[v]];
                    if (d < 0 || (size_t)d >= nd) return fail(RM_ERR_INVALID, "%s: %s index %lld out of range", path.c_str(), index_name, (long long)d);
                    put(v, d);
                }
            } else {
                return RM_OK;
            }
        } else if (mapping == "ByPolygonVertex") {
            if (direct) {
                if (nd != nv) return RM_OK;
                for (size_t v = 0; v < nv; ++v) put(v, (int64_t)v);
            } else if (indexed) {
                if (ix.size() < nv) return RM_OK;
                for (size_t v = 0; v < nv; ++v) {
                    const int64_t d = ix[v];
                    if (d == -1) continue;  // an unset value: zero
                    if (d < 0 || (size_t)d >= nd) return fail(RM_ERR_INVALID, "%s: %s index %lld out of range", path.c_str(), index_name, (long long)d);
                    put(v, d);
                }
            } else {
                return RM_OK;
            }
        } else {
            return RM_OK;
        }
        ok = true;
        return RM_OK;
    }

    int32_t convert_geometry(const FbxNode &g, const std::vector<const FbxNode *> &mats, std::vector<int> &out) {
        out.clear();
        const FbxNode *V = g.kid("Vertices"), *P = g.kid("PolygonVertexIndex");
        if (!V || !P || V->props.empty() || P->props.empty())
            return fail(RM_ERR_INVALID, "%s: geometry without Vertices or PolygonVertexIndex", path.c_str());
        const std::vector<double> &vd = V->props[0].f;
        const std::vector<int64_t> &pv = P->props[0].i;
        if (vd.size() % 3) return fail(RM_ERR_INVALID, "%s: Vertices holds %zu values, not a multiple of 3", path.c_str(), vd.size());
        const size_t ncp = vd.size() / 3;
        if (ncp == 0 || pv.empty()) return RM_OK;  // an empty geometry is skipped
        std::vector<uint32_t> cp_of;                  // control point of each polygon vertex
        std::vector<uint32_t> poly_size;
        cp_of.reserve(pv.size());
        uint32_t count = 0;
        for (int64_t x : pv) {
            const int64_t a = x < 0 ? -x - 1 : x;
            if (a >= (int64_t)ncp) return fail(RM_ERR_INVALID, "%s: polygon vertex index %lld out of range", path.c_str(), (long long)a);
            cp_of.push_back((uint32_t)a);
            ++count;
            if (x < 0) {
                poly_size.push_back(count);
                count = 0;
            }
        }
        cp_of.resize(cp_of.size() - count);  // an unterminated last polygon is dropped
        if (poly_size.empty()) return RM_OK;
        // layers named by the Layer records
        std::vector<float> nrm, uv;
        bool has_n = false, has_uv = false, has_m = false;
        std::vector<int64_t> mindex;
        for (const FbxNode &L : g.kids) {
            if (L.name != "Layer") continue;
            for (const FbxNode &le : L.kids) {
                if (le.name != "LayerElement") continue;
                const FbxNode *ty = le.kid("Type"), *ti = le.kid("TypedIndex");
                if (!ty || !ti || ti->props.empty()) continue;
                const std::string type = ty->str0();
                const int64_t typed = ti->props[0].iv;
                const FbxNode *src = nullptr;
                for (const FbxNode &k : g.kids)
                    if (k.name == type && !k.props.empty() && k.props[0].iv == typed) {
                        src = &k;
                        break;
                    }
                if (!src) continue;
                int32_t rc = RM_OK;
                if (type == "LayerElementNormal" && !has_n) {
                    rc = layer(*src, "Normals", "NormalsIndex", 3, cp_of, ncp, nrm, has_n);
                } else if (type == "LayerElementUV" && typed == 0 && !has_uv) {
                    rc = layer(*src, "UV", "UVIndex", 2, cp_of, ncp, uv, has_uv);
                } else if (type == "LayerElementMaterial" && !has_m) {
                    const FbxNode *mn = src->kid("Materials");
                    if (!mn || mn->props.empty()) return fail(RM_ERR_INVALID, "%s: material layer without Materials", path.c_str());
                    const std::vector<int64_t> &m = mn->props[0].i;
                    const FbxNode *mi = src->kid("MappingInformationType"), *ri = src->kid("ReferenceInformationType");
                    const std::string mapping = mi ? mi->str0() : "", ref = ri ? ri->str0() : "";
                    std::vector<int64_t> per_poly;
                    if (mapping == "AllSame" && !m.empty()) {
                        per_poly.assign(poly_size.size(), m[0]);
                    } else if (mapping == "ByPolygon" && ref == "IndexToDirect") {
                        per_poly = m;
                        per_poly.resize(poly_size.size(), 0);
                    }
                    // a layer of only negative entries means "no material"
                    if (!per_poly.empty() && std::any_of(per_poly.begin(), per_poly.end(), [](int64_t x) { return x >= 0; })) {
                        mindex.swap(per_poly);
                        has_m = true;
                    }
                }
                if (rc) return rc;
            }
        }
        // one mesh per distinct material index, in order of first appearance
        std::vector<int64_t> order;
        if (mindex.empty()) {
            order.push_back(0);
        } else {
            for (int64_t x : mindex)
                if (std::find(order.begin(), order.end(), x) == order.end()) order.push_back(x);
        }
        for (int64_t want : order) {
            Mesh M;
            M.has_n = has_n;
            M.has_uv = has_uv;
            M.material = material_for(mats, want);
            size_t v0 = 0;
            for (size_t p = 0; p < poly_size.size(); ++p) {
                const uint32_t n = poly_size[p];
                if (mindex.empty() || mindex[p] == want) {
                    std::vector<uint32_t> idx(n);
                    for (uint32_t k = 0; k < n; ++k) {
                        const size_t v = v0 + k;
                        idx[k] = (uint32_t)(M.pos.size() / 3);
                        for (int c = 0; c < 3; ++c) M.pos.push_back((float)vd[(size_t)cp_of[v] * 3 + c]);
                        for (int c = 0; c < 3; ++c) M.nrm.push_back(has_n ? nrm[v * 3 + c] : 0.0f);
                        for (int c = 0; c < 2; ++c) M.uv.push_back(has_uv ? uv[v * 2 + c] : 0.0f);
                    }
                    M.polys.push_back(std::move(idx));
                }
                v0 += n;
            }
            meshes.push_back(std::move(M));
            out.push_back((int)meshes.size() - 1);
        }
        return RM_OK;
    }

    int32_t convert_nodes(int64_t parent_id, const M4 &parent_world, int depth) {
        if (depth > FBX_MAX_DEPTH) return fail(RM_ERR_INVALID, "%s: models nested deeper than %d", path.c_str(), FBX_MAX_DEPTH);
        for (const FbxNode *mp : sc.sources(parent_id, "Model")) {
            const FbxNode &m = *mp;
            Props pr;
            pr.own = m.kid("Properties70");
            pr.tmpl = sc.find_tmpl("Model", "FbxNode");
            double T[3], R[3], S[3], Rpre[3], Rpost[3], Rp[3], Roff[3], Sp[3], Soff[3], Gt[3], Gr[3], Gs[3], ro = 0.0;
            pr.vec3("Lcl Translation", T, 0.0);
            pr.vec3("Lcl Rotation", R, 0.0);
            pr.vec3("Lcl Scaling", S, 1.0);
            pr.vec3("PreRotation", Rpre, 0.0);
            pr.vec3("PostRotation", Rpost, 0.0);
            pr.vec3("RotationPivot", Rp, 0.0);
            pr.vec3("RotationOffset", Roff, 0.0);
            pr.vec3("ScalingPivot", Sp, 0.0);
            pr.vec3("ScalingOffset", Soff, 0.0);
            pr.vec3("GeometricTranslation", Gt, 0.0);
            pr.vec3("GeometricRotation", Gr, 0.0);
            pr.vec3("GeometricScaling", Gs, 1.0);
            pr.num("RotationOrder", ro);
            const int order = (int)ro;
            M4 L = translate(T);
            L = mul(L, translate(Roff));
            L = mul(L, translate(Rp));
            L = mul(L, euler(Rpre, order));
            L = mul(L, euler(R, order));
            L = mul(L, transpose(euler(Rpost, order)));  // inverse of a rotation
            L = mul(L, translate(Rp, -1.0));
            L = mul(L, translate(Soff));
            L = mul(L, translate(Sp));
            L = mul(L, scale(S));
            L = mul(L, translate(Sp, -1.0));
            const M4 world = mul(parent_world, L);
            M4 G = mul(mul(translate(Gt), euler(Gr, order)), scale(Gs));
            const M4 mesh_world = mul(world, G);
            const std::vector<const FbxNode *> mats = sc.sources(id_of(m), "Material");
            for (const FbxNode *g : sc.sources(id_of(m), "Geometry")) {
                if (g->props.size() < 3 || g->props[2].str != "Mesh") continue;  // lines, shapes, ...
                auto it = geo_meshes.find(id_of(*g));
                std::vector<int> idx;
                if (it != geo_meshes.end()) {
                    idx = it->second;
                } else {
                    int32_t rc = convert_geometry(*g, mats, idx);
                    if (rc) return rc;
                    geo_meshes[id_of(*g)] = idx;
                }
                for (int k : idx) uses.push_back(Use{k, mesh_world});
            }
            int32_t rc = convert_nodes(id_of(m), world, depth + 1);
            if (rc) return rc;
        }
        return RM_OK;
    }
};

bool is_identity(const M4 &m) { return m == ident(); }

// Quad split: the first corner whose angles to the opposite diagonal add up
// to more than pi starts both triangles (float arithmetic, as the importer).
void split_quad(const Mesh &M, const std::vector<uint32_t> &q, std::vector<std::array<uint32_t, 3>> &tris) {
    auto at = [&](uint32_t v) { return std::array<float, 3>{M.pos[v * 3], M.pos[v * 3 + 1], M.pos[v * 3 + 2]}; };
    auto unit = [](std::array<float, 3> a) {
        const float l = std::sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
        if (l != 0.0f) {
            const float inv = 1.0f / l;
            for (float &x : a) x *= inv;
        }
        return a;
    };
    auto sub = [](const std::array<float, 3> &a, const std::array<float, 3> &b) {
        return std::array<float, 3>{a[0] - b[0], a[1] - b[1], a[2] - b[2]};
    };
    auto dot = [](const std::array<float, 3> &a, const std::array<float, 3> &b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; };
    unsigned start = 0;
    for (unsigned i = 0; i < 4; ++i) {
        const auto v = at(q[i]);
        const auto left = unit(sub(at(q[(i + 3) % 4]), v)), diag = unit(sub(at(q[(i + 2) % 4]), v)),
                   right = unit(sub(at(q[(i + 1) % 4]), v));
        const float angle = std::acos(dot(left, diag)) + std::acos(dot(right, diag));
        if (angle > (float)M_PI) {
            start = i;
            break;
        }
    }
    tris.push_back({q[start], q[(start + 1) % 4], q[(start + 2) % 4]});
    tris.push_back({q[start], q[(start + 2) % 4], q[(start + 3) % 4]});
}

// Ear clipping of a polygon in the plane of its Newell normal.
void clip_ears(const Mesh &M, const std::vector<uint32_t> &poly, std::vector<std::array<uint32_t, 3>> &tris) {
    const size_t n = poly.size();
    double N[3] = {0, 0, 0};
    for (size_t i = 0; i < n; ++i) {
        const float *a = &M.pos[poly[i] * 3], *b = &M.pos[poly[(i + 1) % n] * 3];
        N[0] += ((double)a[1] - b[1]) * ((double)a[2] + b[2]);
        N[1] += ((double)a[2] - b[2]) * ((double)a[0] + b[0]);
        N[2] += ((double)a[0] - b[0]) * ((double)a[1] + b[1]);
    }
    int drop = 2;
    if (fabs(N[0]) >= fabs(N[1]) && fabs(N[0]) >= fabs(N[2])) drop = 0;
    else if (fabs(N[1]) >= fabs(N[2])) drop = 1;
    const int ax = (drop + 1) % 3, ay = (drop + 2) % 3;
    const double sgn = N[drop] < 0 ? -1.0 : 1.0;  // counter-clockwise in (ax, ay)
    std::vector<std::array<double, 2>> p(n);
    for (size_t i = 0; i < n; ++i) p[i] = {M.pos[poly[i] * 3 + ax], sgn * M.pos[poly[i] * 3 + ay]};
    auto cross = [](const std::array<double, 2> &o, const std::array<double, 2> &a, const std::array<double, 2> &b) {
        return (a[0] - o[0]) * (b[1] - o[1]) - (a[1] - o[1]) * (b[0] - o[0]);
    };
    std::vector<size_t> live(n);
    for (size_t i = 0; i < n; ++i) live[i] = i;
    while (live.size() > 3) {
        const size_t m = live.size();
        bool cut = false;
        for (size_t k = 0; k < m && !cut; ++k) {
            const size_t a = live[(k + m - 1) % m], b = live[k], c = live[(k + 1) % m];
            if (cross(p[a], p[b], p[c]) <= 0.0) continue;  // reflex or degenerate corner
            bool inside = false;
            for (size_t j = 0; j < m && !inside; ++j) {
                const size_t q = live[j];
                if (q == a || q == b || q == c) continue;
                inside = cross(p[a], p[b], p[q]) >= 0.0 && cross(p[b], p[c], p[q]) >= 0.0 && cross(p[c], p[a], p[q]) >= 0.0;
            }
            if (inside) continue;
            tris.push_back({poly[a], poly[b], poly[c]});
            live.erase(live.begin() + (long)k);
            cut = true;
        }
        if (!cut) break;  // no ear (self-intersecting or degenerate): fan the rest
    }
    for (size_t k = 1; k + 1 < live.size(); ++k) tris.push_back({poly[live[0]], poly[live[k]], poly[live[k + 1]]});
}

}  // namespace

int32_t parse_fbx(const std::string &folder, const std::string &name, LoadedModel &S) {
    const std::string path = folder + name;
    std::vector<uint8_t> bytes;
    {
        FILE *f = fopen(path.c_str(), "rb");
        if (!f) return fail(RM_ERR_INVALID, "cannot read model %s", path.c_str());
        uint8_t buf[1 << 16];
        size_t k;
        while ((k = fread(buf, 1, sizeof(buf), f)) > 0) bytes.insert(bytes.end(), buf, buf + k);
        fclose(f);
    }
    std::vector<FbxNode> top;
    static const char magic[] = "Kaydara FBX Binary  ";
    int32_t rc;
    if (bytes.size() >= 21 && memcmp(bytes.data(), magic, 21) == 0) {
        rc = FbxReader(bytes).read(path, top);
    } else {
        const std::string head(bytes.begin(), bytes.begin() + std::min<size_t>(bytes.size(), 4096));
        if (head.find("FBXHeaderExtension") == std::string::npos && head.rfind("; FBX", 0) != 0)
            return fail(RM_ERR_INVALID, "%s: not an FBX file", path.c_str());
        rc = FbxAsciiReader(bytes).read(path, top);
        if (!rc) {  // FBXHeaderExtension / FBXVersion
            int64_t version = 0;
            for (const FbxNode &t : top)
                if (t.name == "FBXHeaderExtension")
                    if (const FbxNode *v = t.kid("FBXVersion"))
                        if (!v->props.empty()) version = v->props[0].iv;
            if (version < 7100)
                return fail(RM_ERR_UNSUPPORTED, "%s: FBX version %lld (7100 or later is read)", path.c_str(),
                            (long long)version);
        }
    }
    if (rc) return rc;
    std::vector<uint8_t>().swap(bytes);
    Scene sc;
    for (const FbxNode &n : top) {
        if (n.name == "Objects") {
            for (const FbxNode &o : n.kids)
                if (!o.props.empty() && o.props[0].type == 'L') sc.obj[o.props[0].iv] = &o;
        } else if (n.name == "Connections") {
            for (const FbxNode &c : n.kids) {
                if (c.name != "C" || c.props.size() < 3) continue;
                Conn k{c.props[0].str == "OP", c.props[1].iv, c.props[2].iv,
                       c.props.size() > 3 ? c.props[3].str : std::string()};
                sc.to[k.dst].push_back(sc.conns.size());
                sc.conns.push_back(k);
            }
        } else if (n.name == "Definitions") {
            for (const FbxNode &t : n.kids) {
                if (t.name != "ObjectType") continue;
                for (const FbxNode &pt : t.kids) {
                    if (pt.name != "PropertyTemplate") continue;
                    const FbxNode *p70 = pt.kid("Properties70");
                    if (!p70) continue;
                    sc.tmpl[t.str0() + "." + pt.str0()] = p70;
                    sc.tmpl.emplace(t.str0() + ".", p70);  // the first template of the type
                }
            }
        }
    }
    Converter cv(path, sc);
    if ((rc = cv.convert_nodes(0, ident(), 0))) return rc;
    // pre-transform: each mesh takes its first user's transform; other users
    // with another transform get copies appended in node order
    const size_t n_in = cv.meshes.size();
    for (const auto &u : cv.uses) {
        Mesh &m = cv.meshes[(size_t)u.mesh];
        if (!m.placed) {
            m.placed = true;
            m.world = u.world;
            continue;
        }
        if (m.world == u.world) continue;
        bool found = false;
        for (size_t k = n_in; k < cv.meshes.size() && !found; ++k)
            found = cv.meshes[k].src == u.mesh && cv.meshes[k].world == u.world;
        if (found) continue;
        Mesh c = cv.meshes[(size_t)u.mesh];
        c.src = u.mesh;
        c.world = u.world;
        cv.meshes.push_back(std::move(c));
    }
    bool any = false;
    for (size_t mi = 0; mi < cv.meshes.size(); ++mi) {
        Mesh &M = cv.meshes[mi];
        if (!M.placed && M.src < 0) continue;  // a geometry no node places (not reached)
        const size_t nv = M.pos.size() / 3;
        if (!is_identity(M.world)) {
            const M4 &W = M.world;
            if (det3(W) < 0.0)
                for (auto &p : M.polys) std::reverse(p.begin(), p.end());
            const auto Nm = normal_matrix(W);
            for (size_t v = 0; v < nv; ++v) {
                const double x = M.pos[v * 3], y = M.pos[v * 3 + 1], z = M.pos[v * 3 + 2];
                for (int r = 0; r < 3; ++r)
                    M.pos[v * 3 + r] = (float)(W[r * 4] * x + W[r * 4 + 1] * y + W[r * 4 + 2] * z + W[r * 4 + 3]);
                if (M.has_n) {
                    const double a = M.nrm[v * 3], b = M.nrm[v * 3 + 1], c = M.nrm[v * 3 + 2];
                    float t[3];
                    for (int r = 0; r < 3; ++r) t[r] = (float)(Nm[r * 3] * a + Nm[r * 3 + 1] * b + Nm[r * 3 + 2] * c);
                    const float l = std::sqrt(t[0] * t[0] + t[1] * t[1] + t[2] * t[2]);
                    if (l != 0.0f) {
                        const float inv = 1.0f / l;
                        for (float &x2 : t) x2 *= inv;
                    }
                    for (int r = 0; r < 3; ++r) M.nrm[v * 3 + r] = t[r];
                }
            }
        }
        // triangulate; points and lines are dropped
        std::vector<std::array<uint32_t, 3>> tris;
        for (const auto &p : M.polys) {
            if (p.size() == 3) tris.push_back({p[0], p[1], p[2]});
            else if (p.size() == 4) split_quad(M, p, tris);
            else if (p.size() > 4) clip_ears(M, p, tris);
        }
        if (tris.empty()) continue;
        // inward-facing normals: compare the boxes of the triangles' vertices
        // with and without the normals added
        if (M.has_n) {
            float lo0[3] = {1e10f, 1e10f, 1e10f}, hi0[3] = {-1e10f, -1e10f, -1e10f};
            float lo1[3] = {1e10f, 1e10f, 1e10f}, hi1[3] = {-1e10f, -1e10f, -1e10f};
            std::vector<char> used(nv, 0);
            for (const auto &t : tris)
                for (uint32_t v : t) used[v] = 1;
            for (size_t v = 0; v < nv; ++v) {
                if (!used[v]) continue;
                for (int c = 0; c < 3; ++c) {
                    const float p = M.pos[v * 3 + c], q = p + M.nrm[v * 3 + c];
                    lo1[c] = std::min(lo1[c], p);
                    hi1[c] = std::max(hi1[c], p);
                    lo0[c] = std::min(lo0[c], q);
                    hi0[c] = std::max(hi0[c], q);
                }
            }
            const float d0x = hi0[0] - lo0[0], d0y = hi0[1] - lo0[1], d0z = hi0[2] - lo0[2];
            const float d1x = hi1[0] - lo1[0], d1y = hi1[1] - lo1[1], d1z = hi1[2] - lo1[2];
            bool flip = (d0x > 0.0f) == (d1x > 0.0f) && (d0y > 0.0f) == (d1y > 0.0f) && (d0z > 0.0f) == (d1z > 0.0f);
            const float d1yz = d1y * d1z;
            if (flip && (d1x < 0.05f * std::sqrt(d1yz) || d1y < 0.05f * std::sqrt(d1z * d1x) ||
                         d1z < 0.05f * std::sqrt(d1y * d1x)))
                flip = false;  // planar
            if (flip && std::fabs(d0x * d0y * d0z) < std::fabs(d1x * d1yz)) {
                for (float &x : M.nrm) x = -x;
                for (auto &t : tris) std::swap(t[0], t[2]);
            }
        }
        const int mesh_id = S.mesh.empty() ? 0 : S.mesh.back() + 1;
        for (const auto &t : tris) {
            for (uint32_t v : t) {
                for (int c = 0; c < 3; ++c) S.verts.push_back(M.pos[v * 3 + c]);
                S.uvs.push_back(M.uv[v * 2]);
                S.uvs.push_back(M.uv[v * 2 + 1]);
                for (int c = 0; c < 3; ++c) S.normals.push_back(M.nrm[v * 3 + c]);
            }
            S.material.push_back(M.material);
            S.mesh.push_back(mesh_id);
        }
        if (!M.has_n) S.all_normals = false;
        any = true;
    }
    if (!any) return fail(RM_ERR_EMPTY, "FBX: no triangles in %s", path.c_str());
    return load_material_maps(folder, cv.defs, S);
}

}  // namespace rm
