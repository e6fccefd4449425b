// sms_io_assets.cpp -- asset files of the reference's scenes, read on the host (no device code):
// Wavefront OBJ meshes (src/shapes/obj.cpp), and the images its normal maps come from -- OpenEXR
// scanline files (uncompressed, RLE, ZIPS, ZIP; HALF / FLOAT / UINT channels) and PNG (8 / 16
// bits, grey, grey + alpha, RGB, RGBA, palette) -- as float RGB (src/librender/normalmap.cpp:7-35
// reads .exr as stored and .png as the 8-bit code values / 255, without linearisation).
#include "io_common.h"

#include <zlib.h>

#include <cmath>
#include <cstring>
#include <fstream>
#include <memory>
#include <string>
#include <vector>

namespace smsio {

bool read_file(const std::string &path, std::vector<unsigned char> &out, std::string &err) {
    std::ifstream f(path, std::ios::binary);
    if (!f) { err = "cannot open \"" + path + "\""; return false; }
    f.seekg(0, std::ios::end);
    const std::streamoff size = f.tellg();
    if (size < 0) { err = "cannot read \"" + path + "\""; return false; }
    f.seekg(0, std::ios::beg);
    out.resize((size_t) size);
    if (size > 0 && !f.read((char *) out.data(), size)) { err = "cannot read \"" + path + "\""; return false; }
    return true;
}

// ------------------------------------------------------------------ affine transforms
Mat4 mat_identity() {
    Mat4 m;
    for (int i = 0; i < 16; ++i) m.a[i] = (i % 5 == 0) ? 1.0 : 0.0;
    return m;
}

Mat4 mat_mul(const Mat4 &x, const Mat4 &y) {
    Mat4 r;
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) {
            double s = 0.0;
            for (int k = 0; k < 4; ++k) s += x.a[4 * i + k] * y.a[4 * k + j];
            r.a[4 * i + j] = s;
        }
    return r;
}

bool mat_inverse(const Mat4 &m, Mat4 &out) {
    // Gauss-Jordan with partial pivoting on [m | I]
    double a[4][8];
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 8; ++j) a[i][j] = j < 4 ? m.a[4 * i + j] : (j - 4 == i ? 1.0 : 0.0);
    for (int c = 0; c < 4; ++c) {
        int p = c;
        for (int r = c + 1; r < 4; ++r) if (std::fabs(a[r][c]) > std::fabs(a[p][c])) p = r;
        if (a[p][c] == 0.0) return false;
        if (p != c) for (int j = 0; j < 8; ++j) std::swap(a[p][j], a[c][j]);
        const double inv = 1.0 / a[c][c];
        for (int j = 0; j < 8; ++j) a[c][j] *= inv;
        for (int r = 0; r < 4; ++r) {
            if (r == c) continue;
            const double f = a[r][c];
            for (int j = 0; j < 8; ++j) a[r][j] -= f * a[c][j];
        }
    }
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) out.a[4 * i + j] = a[i][j + 4];
    return true;
}

void mat_point(const Mat4 &m, const double p[3], double out[3]) {
    for (int i = 0; i < 3; ++i) out[i] = m.a[4 * i] * p[0] + m.a[4 * i + 1] * p[1] + m.a[4 * i + 2] * p[2] + m.a[4 * i + 3];
}

void mat_vector(const Mat4 &m, const double v[3], double out[3]) {
    for (int i = 0; i < 3; ++i) out[i] = m.a[4 * i] * v[0] + m.a[4 * i + 1] * v[1] + m.a[4 * i + 2] * v[2];
}

// ------------------------------------------------------------------ OBJ (src/shapes/obj.cpp:95-345)
namespace {

struct Key { uint32_t v, t, n; };

// unit_angle of two unit vectors (the numerically stable form: 2 asin(|a -/+ b| / 2))
double unit_angle(const double a[3], const double b[3]) {
    const double d = a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
    if (d < 0.0) {
        const double s[3] = {a[0] + b[0], a[1] + b[1], a[2] + b[2]};
        return M_PI - 2.0 * std::asin(std::sqrt(s[0] * s[0] + s[1] * s[1] + s[2] * s[2]) * 0.5);
    }
    const double s[3] = {a[0] - b[0], a[1] - b[1], a[2] - b[2]};
    return 2.0 * std::asin(std::sqrt(s[0] * s[0] + s[1] * s[1] + s[2] * s[2]) * 0.5);
}

void normalize3(double v[3]) {
    const double l = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    if (l > 0.0) { v[0] /= l; v[1] /= l; v[2] /= l; }
}

// Mesh::recompute_vertex_normals (src/librender/mesh.cpp:420-470): face normals weighted by the
// face angle at each vertex (Thuermer and Wuethrich, JGT 1998); a vertex whose sum vanishes
// gets (1, 0, 0)
void recompute_vertex_normals(MeshData &m) {
    std::vector<double> acc((size_t) m.n_vertices * 3, 0.0);
    for (int64_t f = 0; f < m.n_faces; ++f) {
        const uint32_t *fi = &m.faces[3 * f];
        double v[3][3];
        for (int j = 0; j < 3; ++j)
            for (int c = 0; c < 3; ++c) v[j][c] = m.positions[3 * (size_t) fi[j] + c];
        double e0[3], e1[3], n[3];
        for (int c = 0; c < 3; ++c) { e0[c] = v[1][c] - v[0][c]; e1[c] = v[2][c] - v[0][c]; }
        n[0] = e0[1] * e1[2] - e0[2] * e1[1];
        n[1] = e0[2] * e1[0] - e0[0] * e1[2];
        n[2] = e0[0] * e1[1] - e0[1] * e1[0];
        const double l2 = n[0] * n[0] + n[1] * n[1] + n[2] * n[2];
        if (!(l2 > 0.0)) continue;
        for (int c = 0; c < 3; ++c) n[c] /= std::sqrt(l2);
        for (int j = 0; j < 3; ++j) {
            double s1[3], s2[3];
            const double *a = v[j], *b = v[(j + 1) % 3], *d = v[(j + 2) % 3];
            for (int c = 0; c < 3; ++c) { s1[c] = b[c] - a[c]; s2[c] = d[c] - a[c]; }
            normalize3(s1);
            normalize3(s2);
            const double w = unit_angle(s1, s2);
            for (int c = 0; c < 3; ++c) acc[3 * (size_t) fi[j] + c] += n[c] * w;
        }
    }
    m.normals.resize((size_t) m.n_vertices * 3);
    for (int64_t i = 0; i < m.n_vertices; ++i) {
        double *a = &acc[3 * i];
        const double l = std::sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
        for (int c = 0; c < 3; ++c) m.normals[3 * i + c] = l != 0.0 ? (float) (a[c] / l) : (c == 0 ? 1.f : 0.f);
    }
}

} // namespace

sms_status load_obj(const std::string &path, const Mat4 *to_world, int flags, MeshData &m, std::string &err) {
    std::vector<unsigned char> buf;
    if (!read_file(path, buf, err)) return SMS_E_INVALID;
    const bool flip_uv = (flags & SMS_OBJ_FLIP_TEXCOORDS) != 0, face_normals = (flags & SMS_OBJ_FACE_NORMALS) != 0;
    Mat4 M = to_world ? *to_world : mat_identity(), Mn;
    if (!mat_inverse(M, Mn)) { err = "OBJ \"" + path + "\": singular to_world"; return SMS_E_INVALID; }
    std::vector<float> pos, nrm, uv;
    std::vector<std::vector<std::pair<Key, uint32_t>>> vmap;   // per position index: (key, vertex id)
    std::vector<Key> keys;                                     // vertex id -> key
    std::vector<uint32_t> tris;
    size_t line_no = 0, i = 0;
    auto fail = [&](const std::string &what) {
        err = "OBJ \"" + path + "\" line " + std::to_string(line_no) + ": " + what;
        return SMS_E_INVALID;
    };
    while (i < buf.size()) {
        size_t e = i;
        while (e < buf.size() && buf[e] != '\n') ++e;
        std::string line((const char *) buf.data() + i, e - i);
        i = e + 1;
        ++line_no;
        size_t s = line.find_first_not_of(" \t\r");
        if (s == std::string::npos) continue;
        const char *cur = line.c_str() + s;
        auto floats = [&](int count, double *out) -> bool {
            for (int k = 0; k < count; ++k) {
                char *end = nullptr;
                out[k] = std::strtod(cur, &end);
                if (end == cur) return false;
                cur = end;
            }
            return true;
        };
        if (cur[0] == 'v' && (cur[1] == ' ' || cur[1] == '\t')) {
            cur += 2;
            double p[3], q[3];
            if (!floats(3, p)) return fail("could not parse a vertex position");
            mat_point(M, p, q);
            for (int c = 0; c < 3; ++c) {
                if (!std::isfinite(q[c])) return fail("invalid vertex position data");
                pos.push_back((float) q[c]);
            }
        } else if (cur[0] == 'v' && cur[1] == 'n' && (cur[2] == ' ' || cur[2] == '\t')) {
            cur += 3;
            double n[3], q[3];
            if (!floats(3, n)) return fail("could not parse a vertex normal");
            // normals transform with the inverse transpose of the linear part
            for (int r = 0; r < 3; ++r) q[r] = Mn.a[r] * n[0] + Mn.a[4 + r] * n[1] + Mn.a[8 + r] * n[2];
            normalize3(q);
            for (int c = 0; c < 3; ++c) {
                if (!std::isfinite(q[c])) return fail("invalid vertex normal data");
                nrm.push_back((float) q[c]);
            }
        } else if (cur[0] == 'v' && cur[1] == 't' && (cur[2] == ' ' || cur[2] == '\t')) {
            cur += 3;
            double t[2];
            if (!floats(2, t)) return fail("could not parse a texture coordinate");
            uv.push_back((float) t[0]);
            uv.push_back(flip_uv ? 1.f - (float) t[1] : (float) t[1]);
        } else if (cur[0] == 'f' && (cur[1] == ' ' || cur[1] == '\t')) {
            cur += 2;
            uint32_t tri[3] = {0, 0, 0};
            int corner = 0;
            while (true) {
                while (*cur == ' ' || *cur == '\t' || *cur == '\r') ++cur;
                if (*cur == '\0') break;
                Key k = {0, 0, 0};
                uint32_t *slot[3] = {&k.v, &k.t, &k.n};
                int field = 0;
                while (true) {
                    if (*cur == '-') return fail("negative (relative) indices are not supported");
                    char *end = nullptr;
                    const unsigned long val = std::strtoul(cur, &end, 10);
                    if (end != cur) {
                        if (field > 2) return fail("more than three indices in a face corner");
                        *slot[field] = (uint32_t) val;
                        cur = end;
                    }
                    if (*cur == '/') { ++field; ++cur; continue; }
                    break;
                }
                if (!(*cur == ' ' || *cur == '\t' || *cur == '\r' || *cur == '\0'))
                    return fail("could not parse a face");
                if (k.v == 0 || k.v > pos.size() / 3) return fail("reference to invalid vertex " + std::to_string(k.v));
                if (k.t > uv.size() / 2) return fail("reference to invalid texture coordinate " + std::to_string(k.t));
                if (k.n > nrm.size() / 3) return fail("reference to invalid normal " + std::to_string(k.n));
                if (vmap.size() < pos.size() / 3) vmap.resize(pos.size() / 3);
                auto &bucket = vmap[k.v - 1];
                uint32_t id = UINT32_MAX;
                for (auto &kv : bucket)
                    if (kv.first.v == k.v && kv.first.t == k.t && kv.first.n == k.n) { id = kv.second; break; }
                if (id == UINT32_MAX) {
                    id = (uint32_t) keys.size();
                    keys.push_back(k);
                    bucket.emplace_back(k, id);
                }
                // fan triangulation: (0, 1, 2), (0, 2, 3), ...
                if (corner < 3) tri[corner] = id;
                else { tri[1] = tri[2]; tri[2] = id; }
                ++corner;
                if (corner >= 3) { tris.push_back(tri[0]); tris.push_back(tri[1]); tris.push_back(tri[2]); }
            }
        }
    }
    m.n_vertices = (int64_t) keys.size();
    m.n_faces = (int64_t) tris.size() / 3;
    if (m.n_faces == 0) { err = "OBJ \"" + path + "\": no faces"; return SMS_E_INVALID; }
    m.positions.resize((size_t) m.n_vertices * 3);
    const bool has_uv = !uv.empty(), has_n = !nrm.empty() && !face_normals;
    if (has_uv) m.texcoords.assign((size_t) m.n_vertices * 2, 0.f);
    if (has_n) m.normals.assign((size_t) m.n_vertices * 3, 0.f);
    for (int64_t v = 0; v < m.n_vertices; ++v) {
        const Key &k = keys[(size_t) v];
        for (int c = 0; c < 3; ++c) m.positions[3 * v + c] = pos[3 * (size_t) (k.v - 1) + c];
        if (has_uv && k.t) { m.texcoords[2 * v] = uv[2 * (size_t) (k.t - 1)]; m.texcoords[2 * v + 1] = uv[2 * (size_t) (k.t - 1) + 1]; }
        if (has_n && k.n) for (int c = 0; c < 3; ++c) m.normals[3 * v + c] = nrm[3 * (size_t) (k.n - 1) + c];
    }
    m.faces = std::move(tris);
    if (!face_normals && nrm.empty()) recompute_vertex_normals(m);
    return SMS_OK;
}

// ------------------------------------------------------------------ images
namespace {

uint32_t be32(const unsigned char *p) { return ((uint32_t) p[0] << 24) | ((uint32_t) p[1] << 16) | ((uint32_t) p[2] << 8) | p[3]; }
uint32_t le32(const unsigned char *p) { return (uint32_t) p[0] | ((uint32_t) p[1] << 8) | ((uint32_t) p[2] << 16) | ((uint32_t) p[3] << 24); }
uint64_t le64(const unsigned char *p) { return (uint64_t) le32(p) | ((uint64_t) le32(p + 4) << 32); }

// IEEE 754 binary16 -> binary32
float half_to_float(uint16_t h) {
    const uint32_t sign = (uint32_t) (h >> 15) << 31, exp = (h >> 10) & 0x1f, man = h & 0x3ff;
    uint32_t bits;
    if (exp == 0) {
        if (man == 0) bits = sign;
        else {                                   // subnormal: renormalise
            int e = -1;
            uint32_t m = man;
            do { ++e; m <<= 1; } while ((m & 0x400) == 0);
            bits = sign | ((uint32_t) (127 - 15 - e) << 23) | ((m & 0x3ff) << 13);
        }
    } else if (exp == 31) bits = sign | 0x7f800000u | (man << 13);
    else bits = sign | ((exp - 15 + 127) << 23) | (man << 13);
    float f;
    std::memcpy(&f, &bits, 4);
    return f;
}

bool inflate_all(const unsigned char *src, size_t n, std::vector<unsigned char> &out, size_t expect) {
    out.resize(expect);
    uLongf len = (uLongf) expect;
    const int rc = uncompress(out.data(), &len, src, (uLong) n);
    return rc == Z_OK && len == expect;
}

// OpenEXR's byte reordering of ZIP / RLE blocks, undone: a delta predictor over the bytes
// (t[i] += t[i-1] - 128), then the two halves (even-position bytes first) interleaved again
void exr_unpredict(std::vector<unsigned char> &t) {
    for (size_t i = 1; i < t.size(); ++i) t[i] = (unsigned char) ((int) t[i - 1] + (int) t[i] - 128);
    std::vector<unsigned char> o(t.size());
    const size_t half = (t.size() + 1) / 2;
    for (size_t i = 0; i < t.size(); ++i) o[i] = (i & 1) ? t[half + i / 2] : t[i / 2];
    t.swap(o);
}

bool exr_unrle(const unsigned char *src, size_t n, std::vector<unsigned char> &out, size_t expect) {
    out.clear();
    size_t i = 0;
    while (i < n) {
        const int c = (signed char) src[i++];
        if (c < 0) {
            const size_t cnt = (size_t) -c;
            if (i + cnt > n) return false;
            out.insert(out.end(), src + i, src + i + cnt);
            i += cnt;
        } else {
            if (i >= n) return false;
            out.insert(out.end(), (size_t) c + 1, src[i++]);
        }
        if (out.size() > expect) return false;
    }
    return out.size() == expect;
}

sms_status load_exr(const std::string &path, const std::vector<unsigned char> &b, Image &img, std::string &err) {
    auto bad = [&](const std::string &w) { err = "EXR \"" + path + "\": " + w; return SMS_E_INVALID; };
    auto unsupported = [&](const std::string &w) { err = "EXR \"" + path + "\": " + w; return SMS_E_UNSUPPORTED; };
    if (b.size() < 8 || le32(b.data()) != 20000630u) return bad("not an OpenEXR file");
    const uint32_t version = le32(b.data() + 4);
    if ((version & 0xff) != 2) return unsupported("format version " + std::to_string(version & 0xff));
    if (version & 0x200) return unsupported("tiled images are not supported");
    if (version & 0x1800) return unsupported("deep or multi-part files are not supported");
    size_t p = 8;
    struct Ch { std::string name; int type; };
    std::vector<Ch> chans;
    int compression = -1;
    int32_t dw[4] = {0, 0, -1, -1};
    bool have_dw = false;
    auto cstr = [&](std::string &s) -> bool {
        size_t e = p;
        while (e < b.size() && b[e] != 0) ++e;
        if (e >= b.size()) return false;
        s.assign((const char *) b.data() + p, e - p);
        p = e + 1;
        return true;
    };
    while (true) {
        if (p >= b.size()) return bad("truncated header");
        if (b[p] == 0) { ++p; break; }
        std::string name, type;
        if (!cstr(name) || !cstr(type) || p + 4 > b.size()) return bad("truncated header");
        const uint32_t size = le32(b.data() + p);
        p += 4;
        if (p + size > b.size()) return bad("truncated attribute " + name);
        const unsigned char *v = b.data() + p;
        if (name == "channels" && type == "chlist") {
            size_t q = 0;
            while (q < size && v[q] != 0) {
                size_t e = q;
                while (e < size && v[e] != 0) ++e;
                if (e + 17 > size) return bad("bad channel list");
                Ch c;
                c.name.assign((const char *) v + q, e - q);
                c.type = (int) le32(v + e + 1);
                if (le32(v + e + 9) != 1 || le32(v + e + 13) != 1) return unsupported("subsampled channels");
                chans.push_back(c);
                q = e + 17;
            }
        } else if (name == "compression" && size >= 1) {
            compression = v[0];
        } else if (name == "dataWindow" && size >= 16) {
            for (int k = 0; k < 4; ++k) dw[k] = (int32_t) le32(v + 4 * k);
            have_dw = true;
        }
        p += size;
    }
    if (chans.empty() || !have_dw || compression < 0) return bad("missing channels, dataWindow or compression");
    const int64_t w = (int64_t) dw[2] - dw[0] + 1, h = (int64_t) dw[3] - dw[1] + 1;
    if (w <= 0 || h <= 0 || w > (1 << 16) || h > (1 << 16)) return bad("bad data window");
    int lines;
    switch (compression) {
        case 0: case 1: case 2: lines = 1; break;      // NONE, RLE, ZIPS
        case 3: lines = 16; break;                     // ZIP
        default: return unsupported("compression " + std::to_string(compression) + " (only NONE, RLE, ZIPS, ZIP)");
    }
    size_t line_bytes = 0;
    for (auto &c : chans) {
        if (c.type < 0 || c.type > 2) return bad("channel type");
        line_bytes += (size_t) w * (c.type == 1 ? 2 : 4);
    }
    auto chan_index = [&](const char *n) {
        for (size_t k = 0; k < chans.size(); ++k) if (chans[k].name == n) return (int) k;
        return -1;
    };
    int src_of[3] = {chan_index("R"), chan_index("G"), chan_index("B")};
    if (src_of[0] < 0 || src_of[1] < 0 || src_of[2] < 0) {
        const int y = chan_index("Y");
        if (y < 0) return unsupported("no R, G, B or Y channels");
        src_of[0] = src_of[1] = src_of[2] = y;
    }
    const int64_t chunks = (h + lines - 1) / lines;
    if (p + 8 * (size_t) chunks > b.size()) return bad("truncated offset table");
    img.width = (int) w; img.height = (int) h; img.channels = 3;
    img.pixels.assign((size_t) w * h * 3, 0.f);
    std::vector<unsigned char> raw, tmp;
    for (int64_t c = 0; c < chunks; ++c) {
        const uint64_t off = le64(b.data() + p + 8 * c);
        if (off + 8 > b.size()) return bad("bad chunk offset");
        const int32_t y0 = (int32_t) le32(b.data() + off);
        const uint32_t n = le32(b.data() + off + 4);
        if (off + 8 + n > b.size()) return bad("truncated chunk");
        const int64_t ly = (int64_t) y0 - dw[1];
        if (ly < 0 || ly >= h) return bad("chunk outside the data window");
        const int64_t nl = std::min<int64_t>(lines, h - ly);
        const size_t expect = line_bytes * (size_t) nl;
        const unsigned char *src = b.data() + off + 8;
        if (compression == 0 || n == expect) raw.assign(src, src + expect);     // stored as is
        else if (compression == 1) {
            if (!exr_unrle(src, n, raw, expect)) return bad("corrupt RLE chunk");
            exr_unpredict(raw);
        } else {
            if (!inflate_all(src, n, raw, expect)) return bad("corrupt ZIP chunk");
            exr_unpredict(raw);
        }
        if (raw.size() < expect) return bad("short chunk");
        size_t q = 0;
        for (int64_t l = 0; l < nl; ++l) {
            const int64_t y = ly + l;
            for (size_t k = 0; k < chans.size(); ++k) {
                const int t = chans[k].type;
                const size_t bytes = t == 1 ? 2 : 4;
                for (int dst = 0; dst < 3; ++dst) {
                    if (src_of[dst] != (int) k) continue;
                    for (int64_t x = 0; x < w; ++x) {
                        const unsigned char *e = raw.data() + q + (size_t) x * bytes;
                        float val;
                        if (t == 1) val = half_to_float((uint16_t) (e[0] | (e[1] << 8)));
                        else if (t == 2) { uint32_t u = le32(e); std::memcpy(&val, &u, 4); }
                        else val = (float) le32(e);
                        img.pixels[((size_t) y * w + x) * 3 + dst] = val;
                    }
                }
                q += (size_t) w * bytes;
            }
        }
    }
    return SMS_OK;
}

sms_status load_png(const std::string &path, const std::vector<unsigned char> &b, Image &img, std::string &err) {
    auto bad = [&](const std::string &w) { err = "PNG \"" + path + "\": " + w; return SMS_E_INVALID; };
    auto unsupported = [&](const std::string &w) { err = "PNG \"" + path + "\": " + w; return SMS_E_UNSUPPORTED; };
    static const unsigned char sig[8] = {0x89, 'P', 'N', 'G', 0x0d, 0x0a, 0x1a, 0x0a};
    if (b.size() < 8 || std::memcmp(b.data(), sig, 8) != 0) return bad("not a PNG file");
    size_t p = 8;
    uint32_t w = 0, h = 0;
    int depth = 0, ctype = -1, interlace = 0;
    std::vector<unsigned char> idat, plte;
    bool ended = false;
    while (p + 12 <= b.size() && !ended) {
        const uint32_t len = be32(b.data() + p);
        const std::string type((const char *) b.data() + p + 4, 4);
        if (p + 12 + (size_t) len > b.size()) return bad("truncated chunk " + type);
        const unsigned char *d = b.data() + p + 8;
        if (type == "IHDR") {
            if (len < 13) return bad("short IHDR");
            w = be32(d); h = be32(d + 4); depth = d[8]; ctype = d[9]; interlace = d[12];
        } else if (type == "PLTE") plte.assign(d, d + len);
        else if (type == "IDAT") idat.insert(idat.end(), d, d + len);
        else if (type == "IEND") ended = true;
        p += 12 + (size_t) len;
    }
    if (w == 0 || h == 0 || w > (1u << 16) || h > (1u << 16)) return bad("bad size");
    if (interlace) return unsupported("interlaced images are not supported");
    int spp;
    switch (ctype) {
        case 0: spp = 1; break;
        case 2: spp = 3; break;
        case 3: spp = 1; break;
        case 4: spp = 2; break;
        case 6: spp = 4; break;
        default: return bad("colour type " + std::to_string(ctype));
    }
    const bool depth_ok = (ctype == 0 || ctype == 3) ? (depth == 1 || depth == 2 || depth == 4 || depth == 8 ||
                                                         (ctype == 0 && depth == 16))
                                                      : (depth == 8 || depth == 16);
    if (!depth_ok) return bad("bit depth " + std::to_string(depth));
    if (ctype == 3 && plte.size() < 3) return bad("palette image without PLTE");
    const size_t row_bits = (size_t) w * spp * depth, row_bytes = (row_bits + 7) / 8;
    const size_t bpp = std::max<size_t>(1, (size_t) spp * depth / 8);
    std::vector<unsigned char> raw;
    if (!inflate_all(idat.data(), idat.size(), raw, (row_bytes + 1) * h)) return bad("corrupt image data");
    // scanline filters (None, Sub, Up, Average, Paeth)
    std::vector<unsigned char> cur(row_bytes), prev(row_bytes, 0);
    img.width = (int) w; img.height = (int) h; img.channels = 3;
    img.pixels.assign((size_t) w * h * 3, 0.f);
    const double maxv = (double) ((1u << depth) - 1u);
    for (uint32_t y = 0; y < h; ++y) {
        const unsigned char *r = raw.data() + (size_t) y * (row_bytes + 1);
        const int f = r[0];
        for (size_t i = 0; i < row_bytes; ++i) {
            const int a = i >= bpp ? cur[i - bpp] : 0, up = prev[i], c = i >= bpp ? prev[i - bpp] : 0;
            int v = r[1 + i];
            switch (f) {
                case 0: break;
                case 1: v += a; break;
                case 2: v += up; break;
                case 3: v += (a + up) / 2; break;
                case 4: {
                    const int pp = a + up - c, pa = std::abs(pp - a), pb = std::abs(pp - up), pc = std::abs(pp - c);
                    v += (pa <= pb && pa <= pc) ? a : (pb <= pc ? up : c);
                    break;
                }
                default: return bad("filter type " + std::to_string(f));
            }
            cur[i] = (unsigned char) v;
        }
        for (uint32_t x = 0; x < w; ++x) {
            double s[4] = {0, 0, 0, 0};
            for (int k = 0; k < spp; ++k) {
                const size_t bit = ((size_t) x * spp + k) * depth;
                uint32_t v;
                if (depth == 16) v = ((uint32_t) cur[bit / 8] << 8) | cur[bit / 8 + 1];
                else if (depth == 8) v = cur[bit / 8];
                else v = (cur[bit / 8] >> (8 - depth - (bit % 8))) & ((1u << depth) - 1u);
                s[k] = (double) v;
            }
            float *o = &img.pixels[((size_t) y * w + x) * 3];
            if (ctype == 3) {
                const size_t idx = (size_t) s[0];
                if (3 * idx + 2 >= plte.size()) return bad("palette index out of range");
                for (int k = 0; k < 3; ++k) o[k] = (float) plte[3 * idx + k] / 255.f;
            } else if (spp >= 3) {
                for (int k = 0; k < 3; ++k) o[k] = (float) s[k] / (float) maxv;
            } else {
                o[0] = o[1] = o[2] = (float) s[0] / (float) maxv;
            }
        }
        std::swap(cur, prev);
    }
    return SMS_OK;
}

} // namespace

sms_status load_image(const std::string &path, Image &img, std::string &err) {
    std::vector<unsigned char> b;
    if (!read_file(path, b, err)) return SMS_E_INVALID;
    if (b.size() >= 4 && le32(b.data()) == 20000630u) return load_exr(path, b, img, err);
    if (b.size() >= 8 && b[0] == 0x89 && b[1] == 'P' && b[2] == 'N' && b[3] == 'G') return load_png(path, b, img, err);
    if (b.size() >= 3 && b[0] == 0xff && b[1] == 0xd8 && b[2] == 0xff) {
        err = "\"" + path + "\": JPEG images are not supported (OpenEXR and PNG are)";
        return SMS_E_UNSUPPORTED;
    }
    err = "\"" + path + "\": unknown image format (OpenEXR and PNG are supported)";
    return SMS_E_UNSUPPORTED;
}

} // namespace smsio
