// sms_io_scene.cpp -- the Mitsuba 2 scene-file subset the SMS experiments use, read on the host
// into an sms_scene_desc (src/libcore/xml.cpp: <default> parameters and $-substitution
// (:447-456, :615-630), transform composition, each op multiplied on the left (:843-926),
// id / <ref> references; the plugins' properties as src/shapes/{sphere,rectangle,obj}.cpp,
// src/bsdfs/{diffuse,dielectric,conductor,roughconductor,normalmap,twosided}.cpp,
// src/emitters/{area,point,directional}.cpp, src/sensors/perspective.cpp, src/films/hdrfilm.cpp,
// src/rfilters/gaussian.cpp and src/samplers/independent.cpp read them).
//
// Plugins outside the solver path (bitmap textures, environment, spot and constant emitters,
// plastic and blended BSDFs, cylinders, ...) are refused with SMS_E_UNSUPPORTED, or, with
// SMS_LOAD_SUBSTITUTE, replaced by the nearest built equivalent (constant textures, diffuse
// BSDFs; emitters and shapes that cannot be replaced are dropped), each replacement recorded
// as a note.
#include "io_common.h"

#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <new>
#include <stdexcept>
#include <string>
#include <vector>

namespace smsio {
namespace {

// ------------------------------------------------------------------ minimal XML DOM
struct Node {
    std::string tag;
    std::vector<std::pair<std::string, std::string>> attrs;
    std::vector<std::unique_ptr<Node>> children;
    int line = 0;
    const std::string *attr(const std::string &k) const {
        for (auto &a : attrs) if (a.first == k) return &a.second;
        return nullptr;
    }
};

struct LoadError : std::runtime_error {
    sms_status code;
    LoadError(sms_status c, const std::string &m) : std::runtime_error(m), code(c) {}
};

class XmlReader {
public:
    XmlReader(const std::string &text, const std::string &name) : s(text), file(name) {}
    std::unique_ptr<Node> parse() {
        std::unique_ptr<Node> root;
        while (true) {
            skip_misc();
            if (p >= s.size()) break;
            if (root) fail("content after the root element");
            root = element();
        }
        if (!root) fail("no root element");
        return root;
    }

private:
    const std::string &s;
    std::string file;
    size_t p = 0;
    int line = 1;

    [[noreturn]] void fail(const std::string &m) {
        throw LoadError(SMS_E_INVALID, file + ":" + std::to_string(line) + ": " + m);
    }
    void adv(size_t n) {
        for (size_t i = 0; i < n && p < s.size(); ++i, ++p) if (s[p] == '\n') ++line;
    }
    bool starts(const char *t) const { return s.compare(p, std::strlen(t), t) == 0; }
    void skip_ws() { while (p < s.size() && std::isspace((unsigned char) s[p])) adv(1); }
    void skip_until(const char *t) {
        const size_t e = s.find(t, p);
        if (e == std::string::npos) fail(std::string("unterminated construct, expected ") + t);
        adv(e + std::strlen(t) - p);
    }
    // whitespace, comments, processing instructions and declarations between elements
    void skip_misc() {
        while (true) {
            skip_ws();
            if (starts("<!--")) skip_until("-->");
            else if (starts("<?")) skip_until("?>");
            else if (starts("<!")) skip_until(">");
            else return;
        }
    }
    std::string name() {
        const size_t b = p;
        while (p < s.size() && (std::isalnum((unsigned char) s[p]) || s[p] == '_' || s[p] == '-' || s[p] == ':' || s[p] == '.'))
            adv(1);
        if (p == b) fail("expected a name");
        return s.substr(b, p - b);
    }
    static std::string unescape(const std::string &v) {
        std::string o;
        for (size_t i = 0; i < v.size(); ++i) {
            if (v[i] == '&') {
                static const char *ents[][2] = {{"&lt;", "<"}, {"&gt;", ">"}, {"&amp;", "&"}, {"&quot;", "\""}, {"&apos;", "'"}};
                bool done = false;
                for (auto &e : ents)
                    if (v.compare(i, std::strlen(e[0]), e[0]) == 0) { o += e[1]; i += std::strlen(e[0]) - 1; done = true; break; }
                if (!done) o += v[i];
            } else o += v[i];
        }
        return o;
    }
    std::unique_ptr<Node> element() {
        if (!starts("<")) fail("expected an element");
        auto n = std::make_unique<Node>();
        n->line = line;
        adv(1);
        n->tag = name();
        while (true) {
            skip_ws();
            if (starts("/>")) { adv(2); return n; }
            if (starts(">")) { adv(1); break; }
            std::string k = name();
            skip_ws();
            if (!starts("=")) fail("expected '=' after attribute " + k);
            adv(1);
            skip_ws();
            if (p >= s.size() || (s[p] != '"' && s[p] != '\'')) fail("expected a quoted value for " + k);
            const char q = s[p];
            adv(1);
            const size_t e = s.find(q, p);
            if (e == std::string::npos) fail("unterminated attribute value");
            n->attrs.emplace_back(k, unescape(s.substr(p, e - p)));
            adv(e + 1 - p);
        }
        while (true) {
            // text content is not used by scene files: skipped
            while (p < s.size() && s[p] != '<') adv(1);
            if (p >= s.size()) fail("unterminated element <" + n->tag + ">");
            if (starts("<!--")) { skip_until("-->"); continue; }
            if (starts("<?")) { skip_until("?>"); continue; }
            if (starts("<![CDATA[")) { skip_until("]]>"); continue; }
            if (starts("</")) {
                adv(2);
                const std::string t = name();
                if (t != n->tag) fail("</" + t + "> closes <" + n->tag + ">");
                skip_ws();
                if (!starts(">")) fail("expected '>'");
                adv(1);
                return n;
            }
            n->children.push_back(element());
        }
    }
};

// ------------------------------------------------------------------ values
std::vector<double> parse_numbers(const std::string &v, const std::string &what) {
    std::vector<double> out;
    const char *c = v.c_str();
    while (*c) {
        while (*c && (std::isspace((unsigned char) *c) || *c == ',')) ++c;
        if (!*c) break;
        char *end = nullptr;
        const double x = std::strtod(c, &end);
        if (end == c) throw LoadError(SMS_E_INVALID, "could not parse \"" + v + "\" as numbers (" + what + ")");
        out.push_back(x);
        c = end;
    }
    return out;
}

double parse_number(const std::string &v, const std::string &what) {
    auto n = parse_numbers(v, what);
    if (n.size() != 1) throw LoadError(SMS_E_INVALID, "expected one number for " + what + ", got \"" + v + "\"");
    return n[0];
}

bool parse_bool(const std::string &v, const std::string &what) {
    std::string l;
    for (char c : v) l += (char) std::tolower((unsigned char) c);
    while (!l.empty() && std::isspace((unsigned char) l.back())) l.pop_back();
    if (l == "true" || l == "1") return true;
    if (l == "false" || l == "0") return false;
    throw LoadError(SMS_E_INVALID, "could not parse \"" + v + "\" as a boolean (" + what + ")");
}

// Transform4f::rotate (axis, degrees), Rodrigues form
Mat4 mat_rotate(const double axis[3], double deg) {
    double a[3] = {axis[0], axis[1], axis[2]};
    const double l = std::sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
    if (!(l > 0.0)) throw LoadError(SMS_E_INVALID, "rotation about a zero axis");
    for (double &x : a) x /= l;
    const double th = deg * M_PI / 180.0, c = std::cos(th), s = std::sin(th), t = 1.0 - c;
    Mat4 m = mat_identity();
    m.a[0] = a[0] * a[0] * t + c;        m.a[1] = a[0] * a[1] * t - a[2] * s; m.a[2] = a[0] * a[2] * t + a[1] * s;
    m.a[4] = a[1] * a[0] * t + a[2] * s; m.a[5] = a[1] * a[1] * t + c;        m.a[6] = a[1] * a[2] * t - a[0] * s;
    m.a[8] = a[2] * a[0] * t - a[1] * s; m.a[9] = a[2] * a[1] * t + a[0] * s; m.a[10] = a[2] * a[2] * t + c;
    return m;
}

// Transform4f::look_at: columns (left, up', dir, origin), left = up x dir
Mat4 mat_look_at(const double o[3], const double t[3], const double up[3]) {
    double d[3] = {t[0] - o[0], t[1] - o[1], t[2] - o[2]};
    auto nrm = [](double *v) {
        const double l = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        if (!(l > 0.0)) throw LoadError(SMS_E_INVALID, "invalid lookat transformation");
        for (int i = 0; i < 3; ++i) v[i] /= l;
    };
    nrm(d);
    double left[3] = {up[1] * d[2] - up[2] * d[1], up[2] * d[0] - up[0] * d[2], up[0] * d[1] - up[1] * d[0]};
    nrm(left);
    double nu[3] = {d[1] * left[2] - d[2] * left[1], d[2] * left[0] - d[0] * left[2], d[0] * left[1] - d[1] * left[0]};
    Mat4 m = mat_identity();
    for (int i = 0; i < 3; ++i) { m.a[4 * i] = left[i]; m.a[4 * i + 1] = nu[i]; m.a[4 * i + 2] = d[i]; m.a[4 * i + 3] = o[i]; }
    return m;
}

// ------------------------------------------------------------------ plugin objects
struct Prop {
    enum Kind { Bool, Int, Float, String, Rgb, Point, Vector, Transform, Texture } kind;
    std::string text;                 // String; Texture: the texture plugin type
    double num = 0.0;
    double v3[3] = {0, 0, 0};
    Mat4 m = mat_identity();
    std::string filename;             // Texture (bitmap): its file
    bool used = false;
};

struct Obj {
    std::string tag, type, id;
    int line = 0;
    std::map<std::string, Prop> props;
    std::vector<std::shared_ptr<Obj>> children;
};

} // namespace

struct SceneFile {
    // storage of the description
    std::vector<sms_shape_desc> shapes;
    std::vector<sms_emitter_desc> emitters;
    std::vector<sms_normalmap_desc> normalmaps;
    std::vector<std::unique_ptr<MeshData>> meshes;
    std::vector<std::unique_ptr<std::vector<float>>> nm_pixels;
    sms_scene_desc desc;
    // integrator, sampler, film
    std::string integrator = "path";
    std::vector<sms_property> iprops;
    int spp = 4;
    float filter_stddev = 0.5f;
    int filter_gaussian = 1;
    std::vector<std::string> notes;
};

namespace {

class Loader {
public:
    Loader(const std::string &path, std::map<std::string, std::string> params, int flags)
        : path_(path), params_(std::move(params)), flags_(flags) {
        const size_t sl = path.find_last_of('/');
        dir_ = sl == std::string::npos ? std::string() : path.substr(0, sl + 1);
    }

    void load(SceneFile &out) {
        std::vector<unsigned char> raw;
        std::string err;
        if (!read_file(path_, raw, err)) throw LoadError(SMS_E_INVALID, err);
        const std::string text(raw.begin(), raw.end());
        XmlReader rd(text, path_);
        std::unique_ptr<Node> root = rd.parse();
        if (root->tag != "scene") throw LoadError(SMS_E_INVALID, path_ + ": root element is <" + root->tag + ">, not <scene>");
        // <default> parameters first (only where no define overrides them)
        for (auto &c : root->children)
            if (c->tag == "default") {
                const std::string *n = c->attr("name"), *v = c->attr("value");
                if (!n || !v) throw LoadError(SMS_E_INVALID, where(*c) + "<default> needs name and value");
                if (!params_.count(*n)) params_[*n] = *v;
            }
        std::vector<std::shared_ptr<Obj>> top;
        for (auto &c : root->children) {
            if (c->tag == "default") continue;
            if (!is_plugin(c->tag)) throw LoadError(SMS_E_INVALID, where(*c) + "unexpected <" + c->tag + "> in <scene>");
            top.push_back(object(*c));
        }
        for (auto &o : top) {
            if (o->tag == "integrator") integrator(*o, out);
            else if (o->tag == "sensor") sensor(*o, out);
            else if (o->tag == "shape") shape(*o, out);
            else if (o->tag == "emitter") emitter(*o, out, -1);
            else if (o->tag == "bsdf" || o->tag == "texture" || o->tag == "spectrum") {
                if (o->id.empty()) note(where_obj(*o) + "unreferenced top-level <" + o->tag + "> ignored");
            } else if (o->tag == "sampler" || o->tag == "film") {
                if (o->id.empty()) note(where_obj(*o) + "top-level <" + o->tag + "> outside a sensor ignored");
            } else unsupported(*o, "top-level <" + o->tag + ">");
        }
        if (!have_sensor_) throw LoadError(SMS_E_INVALID, path_ + ": the scene has no sensor");
        out.desc.n_shapes = (int32_t) out.shapes.size();
        out.desc.shapes = out.shapes.data();
        out.desc.n_emitters = (int32_t) out.emitters.size();
        out.desc.emitters = out.emitters.data();
        out.desc.n_normalmaps = (int32_t) out.normalmaps.size();
        out.desc.normalmaps = out.normalmaps.data();
    }

private:
    std::string path_, dir_;
    std::map<std::string, std::string> params_;
    int flags_;
    std::map<std::string, std::shared_ptr<Obj>> ids_;
    std::map<std::string, int> nm_index_;
    bool have_sensor_ = false;
    SceneFile *sf_ = nullptr;

    std::string where(const Node &n) const { return path_ + ":" + std::to_string(n.line) + ": "; }
    std::string where_obj(const Obj &o) const { return path_ + ":" + std::to_string(o.line) + ": "; }
    bool substitute() const { return (flags_ & SMS_LOAD_SUBSTITUTE) != 0; }
    void note(const std::string &m) { if (sf_) sf_->notes.push_back(m); }
    void unsupported(const Obj &o, const std::string &what) {
        throw LoadError(SMS_E_UNSUPPORTED, where_obj(o) + what + " is not supported");
    }
    static bool is_plugin(const std::string &t) {
        return t == "integrator" || t == "sensor" || t == "shape" || t == "emitter" || t == "bsdf" || t == "texture" ||
               t == "spectrum" || t == "sampler" || t == "film" || t == "rfilter" || t == "medium" || t == "phase" ||
               t == "volume";
    }

    // $name substitution in an attribute value (xml.cpp:447-456)
    std::string value(const Node &n, const char *key, bool required = true) {
        const std::string *v = n.attr(key);
        if (!v) {
            if (required) throw LoadError(SMS_E_INVALID, where(n) + "<" + n.tag + "> needs the attribute " + key);
            return std::string();
        }
        std::string s = *v;
        if (s.find('$') != std::string::npos)
            for (auto &kv : params_) {
                const std::string k = "$" + kv.first;
                size_t pos = 0;
                while ((pos = s.find(k, pos)) != std::string::npos) { s.replace(pos, k.size(), kv.second); pos += kv.second.size(); }
            }
        if (s.find('$') != std::string::npos)
            throw LoadError(SMS_E_INVALID, where(n) + "unresolved parameter in \"" + s + "\"");
        return s;
    }

    void xyz(const Node &n, double out[3], double def) {
        const std::string *v = n.attr("value");
        if (v) {
            auto nums = parse_numbers(value(n, "value"), n.tag);
            if (nums.size() == 1) { out[0] = out[1] = out[2] = nums[0]; return; }
            if (nums.size() != 3) throw LoadError(SMS_E_INVALID, where(n) + "expected 1 or 3 values");
            for (int i = 0; i < 3; ++i) out[i] = nums[(size_t) i];
            return;
        }
        const char *k[3] = {"x", "y", "z"};
        for (int i = 0; i < 3; ++i) out[i] = n.attr(k[i]) ? parse_number(value(n, k[i]), k[i]) : def;
    }

    Mat4 transform(const Node &t) {
        Mat4 m = mat_identity();
        for (auto &c : t.children) {
            const Node &op = *c;
            if (op.tag == "translate") {
                double v[3]; xyz(op, v, 0.0);
                Mat4 x = mat_identity(); x.a[3] = v[0]; x.a[7] = v[1]; x.a[11] = v[2];
                m = mat_mul(x, m);
            } else if (op.tag == "scale") {
                double v[3]; xyz(op, v, 1.0);
                Mat4 x = mat_identity(); x.a[0] = v[0]; x.a[5] = v[1]; x.a[10] = v[2];
                m = mat_mul(x, m);
            } else if (op.tag == "rotate") {
                double v[3]; xyz(op, v, 0.0);
                m = mat_mul(mat_rotate(v, parse_number(value(op, "angle"), "angle")), m);
            } else if (op.tag == "lookat") {
                auto o = parse_numbers(value(op, "origin"), "origin"), tg = parse_numbers(value(op, "target"), "target");
                std::vector<double> up = op.attr("up") ? parse_numbers(value(op, "up"), "up") : std::vector<double>{0, 1, 0};
                if (o.size() != 3 || tg.size() != 3 || up.size() != 3) throw LoadError(SMS_E_INVALID, where(op) + "lookat needs 3-vectors");
                m = mat_mul(mat_look_at(o.data(), tg.data(), up.data()), m);
            } else if (op.tag == "matrix") {
                auto v = parse_numbers(value(op, "value"), "matrix");
                Mat4 x = mat_identity();
                if (v.size() == 16) for (int i = 0; i < 16; ++i) x.a[i] = v[(size_t) i];
                else if (v.size() == 9) for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) x.a[4 * i + j] = v[(size_t) (3 * i + j)];
                else throw LoadError(SMS_E_INVALID, where(op) + "matrix: expected 16 or 9 values");
                m = mat_mul(x, m);
            } else throw LoadError(SMS_E_INVALID, where(op) + "unknown transform operation <" + op.tag + ">");
        }
        return m;
    }

    std::shared_ptr<Obj> object(const Node &n) {
        auto o = std::make_shared<Obj>();
        o->tag = n.tag;
        o->line = n.line;
        o->type = value(n, "type", n.tag != "ref");
        if (n.attr("id")) o->id = value(n, "id");
        for (auto &cp : n.children) {
            const Node &c = *cp;
            if (c.tag == "ref") {
                const std::string id = value(c, "id");
                auto it = ids_.find(id);
                if (it == ids_.end()) throw LoadError(SMS_E_INVALID, where(c) + "reference to unknown id \"" + id + "\"");
                o->children.push_back(it->second);
                continue;
            }
            const bool named_value = (c.tag == "spectrum" || c.tag == "texture") && c.attr("name");
            if (is_plugin(c.tag) && !named_value) {
                auto ch = object(c);
                if (c.attr("name")) ch->id = ch->id.empty() ? value(c, "name") : ch->id;
                o->children.push_back(ch);
                continue;
            }
            const std::string name = value(c, "name");
            Prop p;
            if (c.tag == "boolean") { p.kind = Prop::Bool; p.num = parse_bool(value(c, "value"), name) ? 1.0 : 0.0; }
            else if (c.tag == "integer") { p.kind = Prop::Int; p.num = parse_number(value(c, "value"), name); }
            else if (c.tag == "float") { p.kind = Prop::Float; p.num = parse_number(value(c, "value"), name); }
            else if (c.tag == "string") { p.kind = Prop::String; p.text = value(c, "value"); }
            else if (c.tag == "rgb" || c.tag == "color") { p.kind = Prop::Rgb; xyz(c, p.v3, 0.0); }
            else if (c.tag == "point") { p.kind = Prop::Point; xyz(c, p.v3, 0.0); }
            else if (c.tag == "vector") { p.kind = Prop::Vector; xyz(c, p.v3, 0.0); }
            else if (c.tag == "transform") { p.kind = Prop::Transform; p.m = transform(c); }
            else if (c.tag == "spectrum" || c.tag == "texture") {
                if (c.attr("type")) {
                    // a spectrum-valued texture (bitmap) given inline
                    p.kind = Prop::Texture; p.text = value(c, "type");
                    for (auto &g : c.children)
                        if (g->tag == "string" && g->attr("name") && *g->attr("name") == "filename") p.filename = value(*g, "value");
                } else {
                    const std::string v = value(c, "value");
                    if (v.find(':') != std::string::npos)
                        throw LoadError(SMS_E_UNSUPPORTED, where(c) + "tabulated spectra (wavelength:value) are not supported");
                    p.kind = Prop::Rgb; xyz(c, p.v3, 0.0);
                }
            } else throw LoadError(SMS_E_INVALID, where(c) + "unknown element <" + c.tag + ">");
            o->props[name] = p;
        }
        if (!o->id.empty() && n.tag != "ref") ids_[o->id] = o;
        return o;
    }

    // ---------------------------------------------------------------- property access
    Prop *get(Obj &o, const std::string &k, Prop::Kind kind) {
        auto it = o.props.find(k);
        if (it == o.props.end()) return nullptr;
        Prop &p = it->second;
        const bool ok = p.kind == kind || (kind == Prop::Float && p.kind == Prop::Int) ||
                        (kind == Prop::Rgb && (p.kind == Prop::Float || p.kind == Prop::Int));
        if (!ok) throw LoadError(SMS_E_INVALID, where_obj(o) + "property \"" + k + "\" of <" + o.tag + " type=\"" + o.type + "\"> has another type");
        p.used = true;
        return &p;
    }
    bool b(Obj &o, const std::string &k, bool d) { Prop *p = get(o, k, Prop::Bool); return p ? p->num != 0.0 : d; }
    double f(Obj &o, const std::string &k, double d) { Prop *p = get(o, k, Prop::Float); return p ? p->num : d; }
    int64_t i(Obj &o, const std::string &k, int64_t d) { Prop *p = get(o, k, Prop::Int); return p ? (int64_t) p->num : d; }
    std::string s(Obj &o, const std::string &k, const std::string &d) { Prop *p = get(o, k, Prop::String); return p ? p->text : d; }
    Mat4 xf(Obj &o, const std::string &k) { Prop *p = get(o, k, Prop::Transform); return p ? p->m : mat_identity(); }
    // a colour: rgb / uniform spectrum / float, or a texture (bitmap: the image's mean under
    // SMS_LOAD_SUBSTITUTE; other textures refused)
    void rgb(Obj &o, const std::string &k, const double d[3], float out[3]) {
        auto it = o.props.find(k);
        if (it != o.props.end() && it->second.kind == Prop::Texture) {
            Prop &p = it->second;
            p.used = true;
            if (!substitute()) throw LoadError(SMS_E_UNSUPPORTED, where_obj(o) + "the " + p.text + " texture of \"" + k + "\" is not supported");
            double mean[3] = {0.5, 0.5, 0.5};
            std::string how = "0.5 grey";
            if (p.text == "bitmap" && !p.filename.empty()) {
                Image img;
                std::string err;
                if (load_image(resolve(p.filename), img, err) == SMS_OK && !img.pixels.empty()) {
                    double acc[3] = {0, 0, 0};
                    for (size_t q = 0; q < img.pixels.size(); q += 3) for (int c = 0; c < 3; ++c) acc[c] += img.pixels[q + c];
                    const double npx = (double) img.pixels.size() / 3.0;
                    for (int c = 0; c < 3; ++c) mean[c] = acc[c] / npx;
                    how = "the image's mean colour";
                } else how += " (" + err + ")";
            }
            note(where_obj(o) + p.text + " texture \"" + k + "\" replaced by " + how);
            for (int c = 0; c < 3; ++c) out[c] = (float) mean[c];
            return;
        }
        Prop *p = get(o, k, Prop::Rgb);
        if (!p) { for (int c = 0; c < 3; ++c) out[c] = (float) d[c]; return; }
        if (p->kind == Prop::Rgb) for (int c = 0; c < 3; ++c) out[c] = (float) p->v3[c];
        else for (int c = 0; c < 3; ++c) out[c] = (float) p->num;
    }
    void check_used(Obj &o) {
        for (auto &kv : o.props)
            if (!kv.second.used) note(where_obj(o) + "property \"" + kv.first + "\" of <" + o.tag + " type=\"" + o.type + "\"> ignored");
    }
    std::string resolve(const std::string &fn) const {
        if (!fn.empty() && fn[0] == '/') return fn;
        return dir_ + fn;
    }

    // ---------------------------------------------------------------- plugins
    void integrator(Obj &o, SceneFile &out) {
        out.integrator = o.type;
        for (auto &kv : o.props) {
            const Prop &p = kv.second;
            sms_property q;
            std::memset(&q, 0, sizeof(q));
            std::snprintf(q.name, sizeof(q.name), "%s", kv.first.c_str());
            if (p.kind == Prop::Bool) q.type = SMS_PROP_BOOL;
            else if (p.kind == Prop::Int) q.type = SMS_PROP_INT;
            else if (p.kind == Prop::Float) q.type = SMS_PROP_FLOAT;
            else if (p.kind == Prop::String) { q.type = SMS_PROP_STRING; std::snprintf(q.text, sizeof(q.text), "%s", p.text.c_str()); }
            else throw LoadError(SMS_E_INVALID, where_obj(o) + "integrator property \"" + kv.first + "\" has an unexpected type");
            q.value = p.num;
            out.iprops.push_back(q);
        }
        for (auto &c : o.children) note(where_obj(*c) + "<" + c->tag + "> inside the integrator ignored");
    }

    void sensor(Obj &o, SceneFile &out) {
        if (o.type != "perspective") { unsupported(o, "sensor \"" + o.type + "\""); }
        if (have_sensor_) { note(where_obj(o) + "second sensor ignored"); return; }
        have_sensor_ = true;
        int w = 768, h = 576;
        for (auto &cp : o.children) {
            Obj &c = *cp;
            if (c.tag == "film") {
                if (c.type != "hdrfilm") note(where_obj(c) + "film \"" + c.type + "\" read as hdrfilm");
                w = (int) i(c, "width", 768);
                h = (int) i(c, "height", 576);
                for (const char *k : {"crop_offset_x", "crop_offset_y", "crop_width", "crop_height"}) {
                    const int64_t v = i(c, k, -1);
                    const bool full = v < 0 || (std::string(k) == "crop_width" && v == w) ||
                                      (std::string(k) == "crop_height" && v == h) || (v == 0 && std::string(k).find("offset") != std::string::npos);
                    if (!full) throw LoadError(SMS_E_UNSUPPORTED, where_obj(c) + "cropped films are not supported");
                }
                (void) s(c, "pixel_format", "rgb");
                for (auto &rp : c.children) {
                    Obj &r = *rp;
                    if (r.tag != "rfilter") continue;
                    if (r.type == "gaussian") { out.filter_gaussian = 1; out.filter_stddev = (float) f(r, "stddev", 0.5); }
                    else if (r.type == "box") { out.filter_gaussian = 0; out.filter_stddev = 0.f; }
                    else unsupported(r, "reconstruction filter \"" + r.type + "\"");
                    check_used(r);
                }
                check_used(c);
            } else if (c.tag == "sampler") {
                if (c.type != "independent") note(where_obj(c) + "sampler \"" + c.type + "\" read as independent");
                out.spp = (int) i(c, "sample_count", 4);
                (void) i(c, "seed", 0);
                check_used(c);
            } else note(where_obj(c) + "<" + c.tag + "> inside the sensor ignored");
        }
        const Mat4 m = xf(o, "to_world");
        double fov = f(o, "fov", -1.0);
        if (fov < 0.0) {
            const double fl = std::stod(s(o, "focal_length", "50mm").c_str());
            fov = 2.0 * std::atan(36.0 / (2.0 * fl)) * 180.0 / M_PI;   // 35 mm film, horizontal
        }
        const std::string axis = s(o, "fov_axis", "x");
        (void) f(o, "near_clip", 1e-2);
        (void) f(o, "far_clip", 1e4);
        // the camera's horizontal field of view from the one fov_axis names (perspective.cpp)
        const double aspect = (double) w / (double) h, t = std::tan(0.5 * fov * M_PI / 180.0);
        double tx;
        if (axis == "x") tx = t;
        else if (axis == "y") tx = t * aspect;
        else if (axis == "diagonal") tx = t * aspect / std::sqrt(1.0 + aspect * aspect);
        else if (axis == "smaller") tx = aspect >= 1.0 ? t * aspect : t;
        else if (axis == "larger") tx = aspect >= 1.0 ? t : t * aspect;
        else throw LoadError(SMS_E_INVALID, where_obj(o) + "unknown fov_axis \"" + axis + "\"");
        sms_camera_desc &cam = out.desc.camera;
        const double o0[3] = {0, 0, 0}, z[3] = {0, 0, 1}, y[3] = {0, 1, 0};
        double po[3], pz[3], py[3];
        mat_point(m, o0, po);
        mat_vector(m, z, pz);
        mat_vector(m, y, py);
        for (int c = 0; c < 3; ++c) { cam.origin[c] = (float) po[c]; cam.target[c] = (float) (po[c] + pz[c]); cam.up[c] = (float) py[c]; }
        cam.fov_deg = (float) (2.0 * std::atan(tx) * 180.0 / M_PI);
        cam.width = w; cam.height = h;
        check_used(o);
    }

    // BSDF of a shape into its description
    void bsdf(Obj &o, sms_shape_desc &d, SceneFile &out, bool nested = false) {
        const std::string &t = o.type;
        if (t == "twosided") {
            std::shared_ptr<Obj> in;
            for (auto &c : o.children) if (c->tag == "bsdf") { in = c; break; }
            if (!in) throw LoadError(SMS_E_INVALID, where_obj(o) + "twosided without a nested bsdf");
            note(where_obj(o) + "twosided: the nested bsdf is used on both sides");
            bsdf(*in, d, out, nested);
            return;
        }
        if (t == "diffuse") {
            d.bsdf = SMS_BSDF_DIFFUSE;
            const double def[3] = {0.5, 0.5, 0.5};
            rgb(o, "reflectance", def, d.reflectance);
        } else if (t == "dielectric") {
            d.bsdf = SMS_BSDF_DIELECTRIC;
            d.eta = (float) (ior(o, "int_ior", 1.5046) / ior(o, "ext_ior", 1.000277));
            const double one[3] = {1, 1, 1};
            float tmp[3];
            rgb(o, "specular_reflectance", one, tmp);
            rgb(o, "specular_transmittance", one, tmp);
        } else if (t == "conductor" || t == "roughconductor") {
            d.bsdf = SMS_BSDF_CONDUCTOR;
            const std::string mat = s(o, "material", "none");
            if (mat != "none") {
                if (!substitute()) throw LoadError(SMS_E_UNSUPPORTED, where_obj(o) + "conductor material \"" + mat + "\" (tabulated spectra) is not supported");
                note(where_obj(o) + "conductor material \"" + mat + "\" replaced by a perfect mirror (eta 0, k 1)");
            }
            const double e0[3] = {0, 0, 0}, k1[3] = {1, 1, 1};
            float eta[3], k[3];
            rgb(o, "eta", e0, eta);
            rgb(o, "k", k1, k);
            if (eta[0] != eta[1] || eta[1] != eta[2] || k[0] != k[1] || k[1] != k[2])
                throw LoadError(SMS_E_UNSUPPORTED, where_obj(o) + "coloured conductor eta / k are not supported");
            d.eta = eta[0]; d.k = k[0];
            const double one[3] = {1, 1, 1};
            float tmp[3];
            rgb(o, "specular_reflectance", one, tmp);
            if (t == "roughconductor") {
                const std::string dist = s(o, "distribution", "beckmann");
                if (dist != "beckmann") {
                    if (!substitute()) throw LoadError(SMS_E_UNSUPPORTED, where_obj(o) + "microfacet distribution \"" + dist + "\" is not supported");
                    note(where_obj(o) + "microfacet distribution \"" + dist + "\" replaced by beckmann");
                }
                d.roughness = (float) f(o, "alpha", 0.1);
                (void) b(o, "sample_visible", true);
            }
        } else if (t == "normalmap") {
            std::shared_ptr<Obj> in;
            for (auto &c : o.children) if (c->tag == "bsdf") { in = c; break; }
            if (!in) throw LoadError(SMS_E_INVALID, where_obj(o) + "normalmap without a nested bsdf");
            bsdf(*in, d, out, true);
            if (d.bsdf == SMS_BSDF_DIFFUSE) {
                // a normal-mapped diffuse surface shades as the diffuse surface here
                note(where_obj(o) + "normal map over a diffuse bsdf ignored");
                (void) s(o, "filename", "");
                (void) f(o, "tiles", 1.0);
                (void) b(o, "use_slopes", false);
                check_used(o);
                return;
            }
            if (d.roughness > 0.f) throw LoadError(SMS_E_UNSUPPORTED, where_obj(o) + "normal maps over rough bsdfs are not supported");
            if (d.bsdf == SMS_BSDF_DIELECTRIC) d.k = 0.f;
            d.bsdf = SMS_BSDF_NORMALMAP;
            // use_slopes false (the default) interpolates the encoded normals (normalmap.cpp:226-258)
            const bool encoded = !b(o, "use_slopes", false);
            d.normalmap = normalmap(o, resolve(s(o, "filename", "")), encoded, out);
            d.tiles = (float) f(o, "tiles", 1.0);
            (void) b(o, "lean_fallback", false);
            (void) f(o, "lean_fallback_alpha", 0.0);
        } else {
            if (!substitute()) unsupported(o, "bsdf \"" + t + "\"");
            // plastic-like and blended BSDFs: their diffuse colour, as a diffuse BSDF
            d.bsdf = SMS_BSDF_DIFFUSE;
            const double def[3] = {0.5, 0.5, 0.5};
            if (o.props.count("diffuse_reflectance")) rgb(o, "diffuse_reflectance", def, d.reflectance);
            else if (o.props.count("reflectance")) rgb(o, "reflectance", def, d.reflectance);
            else for (int c = 0; c < 3; ++c) d.reflectance[c] = 0.5f;
            note(where_obj(o) + "bsdf \"" + t + "\" replaced by a diffuse bsdf");
            for (auto &kv : o.props) kv.second.used = true;
        }
        (void) nested;
        check_used(o);
    }

    double ior(Obj &o, const std::string &k, double d) {
        auto it = o.props.find(k);
        if (it == o.props.end()) return d;
        Prop &p = it->second;
        p.used = true;
        if (p.kind == Prop::Float || p.kind == Prop::Int) return p.num;
        if (p.kind == Prop::String) {
            // the named indices of refraction of src/libcore/... (a few common ones)
            static const std::map<std::string, double> named = {
                {"vacuum", 1.0}, {"air", 1.000277}, {"water", 1.3330}, {"bk7", 1.5046}, {"glass", 1.5046},
                {"acrylic glass", 1.49}, {"polypropylene", 1.49}, {"diamond", 2.419}, {"fused quartz", 1.458}};
            auto n = named.find(p.text);
            if (n != named.end()) return n->second;
        }
        throw LoadError(SMS_E_UNSUPPORTED, where_obj(o) + "index of refraction \"" + k + "\" not understood");
    }

    // normal map image -> float RGB in the description (one entry per file and interpolation)
    int normalmap(Obj &o, const std::string &file, bool encoded, SceneFile &out) {
        const std::string key = (encoded ? "encoded:" : "slopes:") + file;
        auto it = nm_index_.find(key);
        if (it != nm_index_.end()) return it->second;
        Image img;
        std::string err;
        const sms_status st = load_image(file, img, err);
        if (st) throw LoadError(st, where_obj(o) + err);
        if (img.width != img.height || (img.width & (img.width - 1)) != 0)
            throw LoadError(SMS_E_INVALID, where_obj(o) + "normal map \"" + file + "\" must be square power-of-two");
        auto px = std::make_unique<std::vector<float>>(std::move(img.pixels));
        sms_normalmap_desc nd;
        std::memset(&nd, 0, sizeof(nd));
        nd.res = img.width;
        nd.rgb32f = px->data();
        nd.encoded_normals = encoded ? 1 : 0;
        out.nm_pixels.push_back(std::move(px));
        out.normalmaps.push_back(nd);
        const int idx = (int) out.normalmaps.size() - 1;
        nm_index_[key] = idx;
        return idx;
    }

    void shape(Obj &o, SceneFile &out) {
        sms_shape_desc d;
        std::memset(&d, 0, sizeof(d));
        d.emitter = -1; d.normalmap = -1; d.tiles = 1.f; d.eta = 1.f;
        d.bsdf = SMS_BSDF_DIFFUSE;
        for (int c = 0; c < 3; ++c) d.reflectance[c] = 0.5f;
        const std::string &t = o.type;
        bool sphere_rotated = false;
        const Mat4 m = xf(o, "to_world");
        if (b(o, "flip_normals", false)) throw LoadError(SMS_E_UNSUPPORTED, where_obj(o) + "flip_normals is not supported");
        if (t == "sphere") {
            d.type = SMS_SHAPE_SPHERE;
            // sphere.cpp: to_world * translate(center) * scale(radius); centre and radius of the result
            double c0[3] = {0, 0, 0};
            if (Prop *p = get(o, "center", Prop::Point)) for (int k = 0; k < 3; ++k) c0[k] = p->v3[k];
            const double r = f(o, "radius", 1.0);
            double cw[3], ex[3], ey[3], ez[3];
            mat_point(m, c0, cw);
            const double ux[3] = {r, 0, 0}, uy[3] = {0, r, 0}, uz[3] = {0, 0, r};
            mat_vector(m, ux, ex); mat_vector(m, uy, ey); mat_vector(m, uz, ez);
            const double rx = std::sqrt(ex[0] * ex[0] + ex[1] * ex[1] + ex[2] * ex[2]);
            const double ry = std::sqrt(ey[0] * ey[0] + ey[1] * ey[1] + ey[2] * ey[2]);
            const double rz = std::sqrt(ez[0] * ez[0] + ez[1] * ez[1] + ez[2] * ez[2]);
            if (std::fabs(rx - ry) > 1e-6 * rx || std::fabs(rx - rz) > 1e-6 * rx)
                throw LoadError(SMS_E_UNSUPPORTED, where_obj(o) + "non-uniformly scaled spheres are not supported");
            for (int k = 0; k < 3; ++k) d.center[k] = (float) cw[k];
            d.radius = (float) rx;
            // the sphere's uv and tangents follow its object axes, which the solver takes to be
            // the world axes
            for (int k = 0; k < 3; ++k)
                if (std::fabs(ex[k] / rx - (k == 0)) > 1e-6 || std::fabs(ey[k] / rx - (k == 1)) > 1e-6 ||
                    std::fabs(ez[k] / rx - (k == 2)) > 1e-6) sphere_rotated = true;
        } else if (t == "rectangle") {
            d.type = SMS_SHAPE_RECTANGLE;
            for (int r = 0; r < 3; ++r) for (int c = 0; c < 4; ++c) d.to_world[4 * r + c] = (float) m.a[4 * r + c];
        } else if (t == "obj") {
            d.type = SMS_SHAPE_MESH;
            int fl = 0;
            if (b(o, "flip_tex_coords", true)) fl |= SMS_OBJ_FLIP_TEXCOORDS;
            if (b(o, "face_normals", false)) fl |= SMS_OBJ_FACE_NORMALS;
            auto mesh = std::make_unique<MeshData>();
            std::string err;
            const sms_status st = load_obj(resolve(s(o, "filename", "")), &m, fl, *mesh, err);
            if (st) throw LoadError(st, where_obj(o) + err);
            if (mesh->n_vertices > INT32_MAX || mesh->n_faces > INT32_MAX) throw LoadError(SMS_E_UNSUPPORTED, where_obj(o) + "mesh too large");
            d.n_vertices = (int32_t) mesh->n_vertices; d.n_faces = (int32_t) mesh->n_faces;
            d.positions = mesh->positions.data();
            d.normals = mesh->normals.empty() ? nullptr : mesh->normals.data();
            d.texcoords = mesh->texcoords.empty() ? nullptr : mesh->texcoords.data();
            d.faces = mesh->faces.data();
            out.meshes.push_back(std::move(mesh));
        } else {
            if (!substitute()) unsupported(o, "shape \"" + t + "\"");
            note(where_obj(o) + "shape \"" + t + "\" dropped");
            return;
        }
        if (b(o, "caustic_caster_single", false)) d.flags |= SMS_FLAG_CAUSTIC_CASTER_SS;
        if (b(o, "caustic_caster_multi", false)) d.flags |= SMS_FLAG_CAUSTIC_CASTER_MS;
        if (b(o, "caustic_receiver", false)) d.flags |= SMS_FLAG_CAUSTIC_RECEIVER;
        if (b(o, "caustic_bouncer", false)) d.flags |= SMS_FLAG_CAUSTIC_BOUNCER;
        if (b(o, "glinty", false)) d.flags |= SMS_FLAG_GLINTY;
        const int index = (int) out.shapes.size();
        bool have_bsdf = false;
        for (auto &c : o.children) {
            if (c->tag == "bsdf") {
                if (have_bsdf) throw LoadError(SMS_E_INVALID, where_obj(o) + "shape with two bsdfs");
                bsdf(*c, d, out);
                have_bsdf = true;
            }
        }
        if ((d.flags & SMS_FLAG_GLINTY) && d.bsdf == SMS_BSDF_NORMALMAP && out.normalmaps[d.normalmap].encoded_normals)
            throw LoadError(SMS_E_UNSUPPORTED, where_obj(o) + "glinty shapes need a normal map with use_slopes");
        if (sphere_rotated) {
            if (d.bsdf == SMS_BSDF_NORMALMAP)
                throw LoadError(SMS_E_UNSUPPORTED, where_obj(o) + "rotated normal-mapped spheres are not supported");
            note(where_obj(o) + "sphere rotation dropped: its uv and tangents follow the world axes");
        }
        out.shapes.push_back(d);
        for (auto &c : o.children) {
            if (c->tag == "emitter") emitter(*c, out, index);
            else if (c->tag != "bsdf") note(where_obj(*c) + "<" + c->tag + "> inside a shape ignored");
        }
        check_used(o);
    }

    void emitter(Obj &o, SceneFile &out, int shape) {
        sms_emitter_desc e;
        std::memset(&e, 0, sizeof(e));
        e.shape = -1;
        const double zero[3] = {0, 0, 0}, one[3] = {1, 1, 1};
        if (o.type == "area") {
            if (shape < 0) throw LoadError(SMS_E_INVALID, where_obj(o) + "area emitter outside a shape");
            e.type = SMS_EMITTER_AREA;
            e.shape = shape;
            rgb(o, "radiance", one, e.radiance);
        } else if (o.type == "point") {
            e.type = SMS_EMITTER_POINT;
            double p[3] = {0, 0, 0};
            if (Prop *pp = get(o, "position", Prop::Point)) for (int k = 0; k < 3; ++k) p[k] = pp->v3[k];
            else { const Mat4 m = xf(o, "to_world"); const double z[3] = {0, 0, 0}; mat_point(m, z, p); }
            for (int k = 0; k < 3; ++k) e.position[k] = (float) p[k];
            rgb(o, "intensity", one, e.intensity);
        } else if (o.type == "directional") {
            e.type = SMS_EMITTER_DIRECTIONAL;
            double dir[3] = {0, 0, 1};
            if (Prop *pp = get(o, "direction", Prop::Vector)) for (int k = 0; k < 3; ++k) dir[k] = pp->v3[k];
            else { const Mat4 m = xf(o, "to_world"); const double z[3] = {0, 0, 1}; mat_vector(m, z, dir); }
            const double l = std::sqrt(dir[0] * dir[0] + dir[1] * dir[1] + dir[2] * dir[2]);
            if (!(l > 0.0)) throw LoadError(SMS_E_INVALID, where_obj(o) + "zero direction");
            for (int k = 0; k < 3; ++k) e.direction[k] = (float) (dir[k] / l);
            rgb(o, "irradiance", one, e.irradiance);
        } else {
            if (!substitute()) unsupported(o, "emitter \"" + o.type + "\"");
            note(where_obj(o) + "emitter \"" + o.type + "\" dropped");
            for (auto &kv : o.props) kv.second.used = true;
            return;
        }
        (void) zero;
        e.caustic_ss = b(o, "caustic_emitter_single", false) ? 1 : 0;
        e.caustic_ms = b(o, "caustic_emitter_multi", false) ? 1 : 0;
        if (shape >= 0) out.shapes[(size_t) shape].emitter = (int32_t) out.emitters.size();
        out.emitters.push_back(e);
        check_used(o);
    }

public:
    void bind(SceneFile *sf) { sf_ = sf; }
};

} // namespace

sms_status load_scene(const std::string &path, const std::map<std::string, std::string> &defines, int flags,
                      SceneFile &out, std::string &err) {
    try {
        std::memset(&out.desc, 0, sizeof(out.desc));
        Loader ld(path, defines, flags);
        ld.bind(&out);
        ld.load(out);
        return SMS_OK;
    } catch (const LoadError &e) {
        err = e.what();
        return e.code;
    } catch (const std::bad_alloc &) {
        err = "out of host memory while loading \"" + path + "\"";
        return SMS_E_OOM;
    } catch (const std::exception &e) {
        err = e.what();
        return SMS_E_INVALID;
    }
}

} // namespace smsio

// ------------------------------------------------------------------ C ABI
namespace {

void put_err(char *err, int32_t len, const std::string &m) {
    if (err && len > 0) std::snprintf(err, (size_t) len, "%s", m.c_str());
}

template <class F> sms_status guarded_io(char *err, int32_t len, F &&f) {
    try {
        return f();
    } catch (const std::bad_alloc &) {
        put_err(err, len, "out of host memory");
        return SMS_E_OOM;
    } catch (const std::exception &e) {
        put_err(err, len, e.what());
        return SMS_E_INVALID;
    } catch (...) {
        put_err(err, len, "unknown error");
        return SMS_E_INVALID;
    }
}

} // namespace

struct sms_scene_file { smsio::SceneFile f; };
struct sms_mesh_holder { sms_mesh_data pub; smsio::MeshData m; };
struct sms_image_holder { sms_image pub; smsio::Image img; };

extern "C" {

sms_status sms_obj_load(const char *path, const float *to_world, int32_t flags, sms_mesh_data **out, char *err,
                        int32_t err_len) {
    return guarded_io(err, err_len, [&]() -> sms_status {
        if (!path || !out) { put_err(err, err_len, "sms_obj_load: null argument"); return SMS_E_INVALID; }
        *out = nullptr;
        smsio::Mat4 m = smsio::mat_identity();
        if (to_world) for (int i = 0; i < 16; ++i) m.a[i] = to_world[i];
        auto *h = new sms_mesh_holder();
        std::string e;
        const sms_status st = smsio::load_obj(path, &m, flags, h->m, e);
        if (st) { delete h; put_err(err, err_len, e); return st; }
        h->pub.n_vertices = h->m.n_vertices;
        h->pub.n_faces = h->m.n_faces;
        h->pub.positions = h->m.positions.data();
        h->pub.normals = h->m.normals.empty() ? nullptr : h->m.normals.data();
        h->pub.texcoords = h->m.texcoords.empty() ? nullptr : h->m.texcoords.data();
        h->pub.faces = h->m.faces.data();
        *out = &h->pub;
        return SMS_OK;
    });
}

void sms_mesh_data_free(sms_mesh_data *m) {
    if (m) delete reinterpret_cast<sms_mesh_holder *>(m);
}

sms_status sms_image_load(const char *path, sms_image **out, char *err, int32_t err_len) {
    return guarded_io(err, err_len, [&]() -> sms_status {
        if (!path || !out) { put_err(err, err_len, "sms_image_load: null argument"); return SMS_E_INVALID; }
        *out = nullptr;
        auto *h = new sms_image_holder();
        std::string e;
        const sms_status st = smsio::load_image(path, h->img, e);
        if (st) { delete h; put_err(err, err_len, e); return st; }
        h->pub.width = h->img.width; h->pub.height = h->img.height; h->pub.channels = h->img.channels;
        h->pub.pixels = h->img.pixels.data();
        *out = &h->pub;
        return SMS_OK;
    });
}

void sms_image_free(sms_image *img) {
    if (img) delete reinterpret_cast<sms_image_holder *>(img);
}

sms_status sms_scene_load(const char *path, const char *const *defines, int32_t n_defines, int32_t flags,
                          sms_scene_file **out, char *err, int32_t err_len) {
    return guarded_io(err, err_len, [&]() -> sms_status {
        if (!path || !out || n_defines < 0 || (n_defines > 0 && !defines)) {
            put_err(err, err_len, "sms_scene_load: bad argument");
            return SMS_E_INVALID;
        }
        *out = nullptr;
        std::map<std::string, std::string> defs;
        for (int32_t i = 0; i < n_defines; ++i) {
            const std::string d = defines[i] ? defines[i] : "";
            const size_t eq = d.find('=');
            if (eq == std::string::npos || eq == 0) {
                put_err(err, err_len, "sms_scene_load: define \"" + d + "\" is not name=value");
                return SMS_E_INVALID;
            }
            defs[d.substr(0, eq)] = d.substr(eq + 1);
        }
        auto *h = new sms_scene_file();
        std::string e;
        const sms_status st = smsio::load_scene(path, defs, flags, h->f, e);
        if (st) { delete h; put_err(err, err_len, e); return st; }
        *out = h;
        return SMS_OK;
    });
}

const sms_scene_desc *sms_scene_file_desc(const sms_scene_file *f) { return f ? &f->f.desc : nullptr; }

sms_status sms_scene_file_info(const sms_scene_file *f, sms_scene_info *info) {
    if (!f || !info) return SMS_E_INVALID;
    std::memset(info, 0, sizeof(*info));
    std::snprintf(info->integrator, sizeof(info->integrator), "%s", f->f.integrator.c_str());
    info->n_properties = (int32_t) f->f.iprops.size();
    info->spp = f->f.spp;
    info->filter_gaussian = f->f.filter_gaussian;
    info->filter_stddev = f->f.filter_stddev;
    info->n_notes = (int32_t) f->f.notes.size();
    return SMS_OK;
}

sms_status sms_scene_file_property(const sms_scene_file *f, int32_t i, sms_property *p) {
    if (!f || !p || i < 0 || i >= (int32_t) f->f.iprops.size()) return SMS_E_INVALID;
    *p = f->f.iprops[(size_t) i];
    return SMS_OK;
}

const char *sms_scene_file_note(const sms_scene_file *f, int32_t i) {
    if (!f || i < 0 || i >= (int32_t) f->f.notes.size()) return nullptr;
    return f->f.notes[(size_t) i].c_str();
}

void sms_scene_file_free(sms_scene_file *f) { delete f; }

} // extern "C"
