// .psy scene reader: text -> data tree -> psy_scene_desc / psy_world_desc /
// psy_camera_desc / psy_render_settings (include/psy_file.h).
//
// The data tree grammar is the reference's (src/parse/data_tree.rs:292-563):
// nodes are `Type $ident { children }`, `Type { children }` or
// `Type [leaf contents]`; whitespace is space, tab, CR and LF; `#` starts a
// comment to the end of the line; `{ } [ ] $ # \` are reserved; identifiers
// and leaf contents keep backslash escapes verbatim.  A node's byte offset
// is the position just after its type name, as in the reference.  The scene
// sections follow src/parse/psy.rs, psy_assembly.rs, psy_mesh_surface.rs,
// psy_surface_shader.rs and psy_light.rs.
#include <algorithm>
#include <deque>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <memory>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/psy_file.h"
#include "color_host.h"
#include "psymath.h"
#include "scene.h"

namespace psy {
namespace api {
int fail(int code, const std::string& msg);  // api.hip: records psy_last_error()
}  // namespace api
}  // namespace psy

namespace {

// ---------------------------------------------------------------- errors

struct ParseFail {
    std::string kind;
    uint64_t offset;
    std::string message;
};

[[noreturn]] void raise(const char* kind, uint64_t offset, const std::string& message) {
    throw ParseFail{kind, offset, message};
}

// ---------------------------------------------------------------- data tree

struct Node {
    bool internal = false;
    bool has_ident = false;
    std::string type, ident, contents;
    uint64_t offset = 0;
    std::vector<Node> children;
};

inline bool is_ws(char c) { return c == '\n' || c == '\r' || c == '\t' || c == ' '; }
inline bool is_reserved(char c) {
    return c == '{' || c == '}' || c == '[' || c == ']' || c == '$' || c == '#' || c == '\\';
}
// UTF-8 continuation and lead bytes are >= 0x80: neither whitespace nor
// reserved, so byte-wise classification splits text where the reference's
// char-wise one does.
inline bool is_ident_char(char c) { return !is_ws(c) && !is_reserved(c); }

enum class Tok { OpenInner, CloseInner, OpenLeaf, CloseLeaf, TypeName, Ident, End, Unknown };

class TreeReader {
public:
    TreeReader(const char* s, uint64_t n) : s_(s), n_(n) {}

    Node read_root() {
        Node root;
        root.internal = true;
        root.type = "ROOT";
        uint64_t pos = read_children(0, root.children);
        pos = skip(pos);
        if (pos != n_) raise("Other", 0, "Failed to parse the entire string.");
        return root;
    }

private:
    // Nesting deeper than this is refused instead of recursing further (the
    // reference recurses without a limit and overflows its stack).
    static constexpr int MAX_NESTING = 2048;
    const char* s_;
    uint64_t n_;
    int depth_ = 0;

    // Whitespace and comments (data_tree.rs:536-553).
    uint64_t skip(uint64_t p) const {
        for (;;) {
            const uint64_t start = p;
            while (p < n_ && is_ws(s_[p])) p++;
            if (p < n_ && s_[p] == '#')
                while (p < n_ && s_[p] != '\n' && s_[p] != '\r') p++;
            if (p == start) return p;
        }
    }

    // One token from p: its kind, its text [b, e) and the position after it.
    Tok token(uint64_t p, uint64_t& b, uint64_t& e) const {
        p = skip(p);
        b = p;
        if (p >= n_) {
            e = p;
            return Tok::End;
        }
        const char c = s_[p];
        e = p + 1;
        switch (c) {
            case '{': return Tok::OpenInner;
            case '}': return Tok::CloseInner;
            case '[': return Tok::OpenLeaf;
            case ']': return Tok::CloseLeaf;
            case '$': {
                uint64_t q = p + 1;
                bool esc = false;
                for (; q < n_; q++) {
                    if (esc) esc = false;
                    else if (s_[q] == '\\') esc = true;
                    else if (!is_ident_char(s_[q])) break;
                }
                e = q;
                return Tok::Ident;
            }
            default:
                if (is_ident_char(c)) {
                    uint64_t q = p;
                    while (q < n_ && is_ident_char(s_[q])) q++;
                    e = q;
                    return Tok::TypeName;
                }
                e = p;
                return Tok::Unknown;
        }
    }

    uint64_t read_children(uint64_t p, std::vector<Node>& out) {
        for (;;) {
            Node n;
            uint64_t next;
            if (!read_node(p, n, next)) return p;
            out.push_back(std::move(n));
            p = next;
        }
    }

    // A node at p; false (consuming nothing) if no type name starts there.
    bool read_node(uint64_t p, Node& node, uint64_t& after) {
        uint64_t b, e;
        if (token(p, b, e) != Tok::TypeName) return false;
        node.type.assign(s_ + b, e - b);
        node.offset = e;
        const uint64_t p1 = e;
        const Tok t2 = token(p1, b, e);
        if (t2 == Tok::Ident || t2 == Tok::OpenInner) {
            node.internal = true;
            uint64_t p2 = e;
            if (t2 == Tok::Ident) {
                node.has_ident = true;
                node.ident.assign(s_ + b, e - b);
                if (token(p2, b, e) != Tok::OpenInner) raise("MissingOpenInternal", p2, "expected '{'");
                p2 = e;
            }
            if (++depth_ > MAX_NESTING) raise("NestingTooDeep", p1, "nodes nested more than 2048 deep");
            const uint64_t rem = read_children(p2, node.children);
            --depth_;
            if (token(rem, b, e) != Tok::CloseInner) raise("MissingCloseInternal", rem, "expected '}'");
            after = e;
            return true;
        }
        if (t2 == Tok::OpenLeaf) {
            uint64_t q = e;
            bool esc = false;
            for (; q < n_; q++) {
                if (esc) esc = false;
                else if (s_[q] == '\\') esc = true;
                else if (s_[q] == ']') break;
            }
            node.contents.assign(s_ + e, q - e);
            if (token(q, b, e) != Tok::CloseLeaf) raise("MissingCloseLeaf", q, "expected ']'");
            after = e;
            return true;
        }
        raise("MissingOpener", p1, "expected '{', '[' or an identifier");
    }
};

void dump(const Node& n, int depth, std::string& out) {
    out.append(2 * depth, ' ');
    if (n.internal) {
        out += "I " + n.type + " " + (n.has_ident ? n.ident : std::string("-")) + " " + std::to_string(n.offset) + "\n";
        for (const Node& c : n.children) dump(c, depth + 1, out);
        return;
    }
    out += "L " + n.type + " " + std::to_string(n.offset) + " ";
    for (char c : n.contents) {
        if (c == '\\') out += "\\\\";
        else if (c == '\n') out += "\\n";
        else if (c == '\r') out += "\\r";
        else out += c;
    }
    out += "\n";
}

// ---------------------------------------------------------------- numbers
// Leaf numbers (src/parse/basics.rs:16-26): optional spaces, tabs, CR or LF
// around each value.  Floats: sign, digits with an optional fraction (or a
// fraction alone), an optional exponent with digits, or nan / inf /
// infinity in any case; correctly rounded to f32.

class Cursor {
public:
    explicit Cursor(const std::string& s) : s_(&s) {}
    uint64_t pos() const { return p_; }
    bool at_end() const { return p_ == s_->size(); }
    void space() {
        const std::string& s = *s_;
        while (p_ < s.size() && (s[p_] == ' ' || s[p_] == '\t' || s[p_] == '\r' || s[p_] == '\n')) p_++;
    }
    bool f32(float& v) {
        const uint64_t save = p_;
        space();
        if (!float_token(v)) {
            p_ = save;
            return false;
        }
        space();
        return true;
    }
    bool uint(uint64_t limit, uint64_t& v) {
        const std::string& s = *s_;
        const uint64_t save = p_;
        space();
        uint64_t q = p_;
        while (q < s.size() && s[q] >= '0' && s[q] <= '9') q++;
        if (q == p_) {
            p_ = save;
            return false;
        }
        uint64_t acc = 0;
        for (uint64_t i = p_; i < q; i++) {
            const uint64_t d = (uint64_t)(s[i] - '0');
            if (acc > (limit - d) / 10) {
                p_ = save;
                return false;
            }
            acc = acc * 10 + d;
        }
        v = acc;
        p_ = q;
        space();
        return true;
    }

private:
    const std::string* s_;
    uint64_t p_ = 0;

    static bool lower_eq(const std::string& s, uint64_t at, const char* w) {
        const uint64_t n = std::strlen(w);
        if (at + n > s.size()) return false;
        for (uint64_t i = 0; i < n; i++) {
            char c = s[at + i];
            if (c >= 'A' && c <= 'Z') c = (char)(c - 'A' + 'a');
            if (c != w[i]) return false;
        }
        return true;
    }
    bool float_token(float& v) {
        const std::string& s_ = *this->s_;
        uint64_t q = p_;
        bool neg = false;
        if (q < s_.size() && (s_[q] == '+' || s_[q] == '-')) {
            neg = s_[q] == '-';
            q++;
        }
        for (const char* w : {"infinity", "inf", "nan"}) {
            if (lower_eq(s_, q, w)) {
                v = w[0] == 'n' ? std::numeric_limits<float>::quiet_NaN() : std::numeric_limits<float>::infinity();
                if (neg) v = -v;
                p_ = q + std::strlen(w);
                return true;
            }
        }
        auto digits = [&](uint64_t i) {
            while (i < s_.size() && s_[i] >= '0' && s_[i] <= '9') i++;
            return i;
        };
        const uint64_t d0 = q;
        q = digits(q);
        const bool int_digits = q > d0;
        if (q < s_.size() && s_[q] == '.') {
            const uint64_t f0 = q + 1;
            const uint64_t f1 = digits(f0);
            if (!int_digits && f1 == f0) return false;
            q = f1;
        } else if (!int_digits) {
            return false;
        }
        if (q < s_.size() && (s_[q] == 'e' || s_[q] == 'E')) {
            uint64_t m = q + 1;
            if (m < s_.size() && (s_[m] == '+' || s_[m] == '-')) m++;
            const uint64_t e1 = digits(m);
            if (e1 == m) return false;
            q = e1;
        }
        const std::string text = s_.substr(p_, q - p_);
        v = std::strtof(text.c_str(), nullptr);
        p_ = q;
        return true;
    }
};

// all_consuming(tuple((ws_f32 x n)))
bool all_f32(const std::string& s, int n, float* out) {
    Cursor c(s);
    for (int i = 0; i < n; i++)
        if (!c.f32(out[i])) return false;
    return c.at_end();
}

// ---------------------------------------------------------------- scene

psy_color parse_color(const std::string& contents, uint64_t offset, const char* kind) {
    // psy.rs:577-615: "<space>, <values>" split on commas, both trimmed.
    std::vector<std::string> items;
    uint64_t start = 0;
    for (;;) {
        const uint64_t comma = contents.find(',', start);
        items.push_back(contents.substr(start, comma == std::string::npos ? std::string::npos : comma - start));
        if (comma == std::string::npos) break;
        start = comma + 1;
    }
    auto trim = [](const std::string& t) {
        // str::trim: Unicode whitespace; the ASCII set covers .psy text.
        uint64_t a = 0, b = t.size();
        while (a < b && (t[a] == ' ' || t[a] == '\t' || t[a] == '\n' || t[a] == '\r' || t[a] == '\f' || t[a] == '\v')) a++;
        while (b > a && (t[b - 1] == ' ' || t[b - 1] == '\t' || t[b - 1] == '\n' || t[b - 1] == '\r' ||
                         t[b - 1] == '\f' || t[b - 1] == '\v'))
            b--;
        return t.substr(a, b - a);
    };
    for (auto& it : items) it = trim(it);
    const std::string msg = "Color should be '<rec709|blackbody|color_temperature>, <values>'.";
    if (items.size() != 2) raise(kind, offset, msg);
    Cursor c(items[1]);
    psy_color out{};
    if (items[0] == "rec709") {
        float rgb[3];
        for (float& v : rgb)
            if (!c.f32(v)) raise(kind, offset, msg);
        // rec709_e_to_xyz (sub_crates/color/build.rs:132-153)
        float xyz[3];
        psy::color::apply(psy::color::matrices().rec709_e_to_xyz, rgb[0], rgb[1], rgb[2], xyz);
        out.kind = PSY_COLOR_XYZ;
        out.a = xyz[0];
        out.b = xyz[1];
        out.c = xyz[2];
        return out;
    }
    if (items[0] == "blackbody" || items[0] == "color_temperature") {
        float tf[2];
        for (float& v : tf)
            if (!c.f32(v)) raise(kind, offset, msg);
        out.kind = items[0] == "blackbody" ? PSY_COLOR_BLACKBODY : PSY_COLOR_TEMPERATURE;
        out.a = tf[0];
        out.b = tf[1];
        return out;
    }
    raise(kind, offset, msg);
}

// parse_matrix (psy.rs:556-572): sixteen column-major values; the fourth row
// is dropped.
void parse_matrix(const std::string& contents, uint64_t offset, const char* kind, float out[12]) {
    float v[16];
    if (!all_f32(contents, 16, v))
        raise(kind, offset, "Transform should be sixteen integers specified in the form '[# # # # # # # # # # # # # # # #]'.");
    const int order[12] = {0, 4, 8, 12, 1, 5, 9, 13, 2, 6, 10, 14};
    for (int i = 0; i < 12; i++) out[i] = v[order[i]];
}

std::vector<const Node*> children_of(const Node& n, const char* type, int want /* 0 any, 1 internal, 2 leaf */) {
    std::vector<const Node*> out;
    for (const Node& c : n.children)
        if (c.type == type && (want == 0 || (want == 1) == c.internal)) out.push_back(&c);
    return out;
}

struct ObjectRef {
    uint32_t kind;   // PSY_OBJECT_*
    uint32_t index;  // into meshes / rect lights / sphere lights
};

struct MeshData {
    uint32_t vert_count = 0, time_samples = 0;
    std::vector<float> verts, normals;
    std::vector<uint32_t> indices;
};

}  // namespace

// Storage for light time samples: the descriptors point into it, and a
// deque never moves its elements, so the pointers stay valid as it grows
// and when its vectors are moved into the file's scene data.
struct SamplePool {
    std::deque<std::vector<float>> floats;
    std::deque<std::vector<psy_color>> colours;
};

// One assembly as the reference's AssemblyBuilder holds it (scene/
// assembly.rs:150-286): its own shaders and objects, sub-assemblies sharing
// the object namespace, and instances of objects or sub-assemblies.
constexpr uint32_t KIND_ASSEMBLY = 3;
struct AsmDef {
    std::vector<psy_shader_desc> shaders;
    std::vector<MeshData> meshes;
    std::vector<psy_rect_light_desc> rects;
    std::vector<psy_sphere_light_desc> spheres;
    SamplePool pool;  // light time samples
    std::vector<std::unique_ptr<AsmDef>> subs;
    struct Inst {
        uint32_t kind;    // PSY_OBJECT_* or KIND_ASSEMBLY
        uint32_t index;   // into meshes / rects / spheres / subs
        uint32_t shader;  // into shaders, PSY_NONE if unbound
        std::vector<psy::Xf> xf;
    };
    std::vector<Inst> instances;
    uint32_t shader_base = 0, mesh_base = 0;  // positions in the scene's shader and mesh lists
};

// An assembly as descriptors (psy_scene_desc) over the arrays it owns; its
// sub-assemblies' descriptors are contiguous, as psy_scene_desc.assemblies
// requires.
struct AsmOut {
    std::vector<psy_mesh_desc> meshes;
    std::vector<psy_instance_desc> instances;
    std::vector<float> xforms;
    std::vector<std::unique_ptr<AsmOut>> subs;
    std::vector<psy_scene_desc> sub_descs;
    psy_scene_desc desc{};
};

struct psy_file_scene_data {
    std::string name, output;
    std::vector<MeshData> mesh_data;   // every assembly's meshes, depth first in definition order
    std::unique_ptr<AsmDef> root_def;  // the parsed assemblies (their lights and light samples)
    std::unique_ptr<AsmOut> root_out;  // and their descriptors, nested as in the file
    std::vector<psy_shader_desc> shaders;
    std::vector<psy_distant_light_desc> distant;
    psy_color background{};
    std::vector<float> cam_xforms;
    std::vector<float> fovs, apertures, focuses;  // lens time samples
    SamplePool pool;                              // light time samples
    psy_render_settings settings{};
};

struct psy_file {
    std::vector<std::unique_ptr<psy_file_scene_data>> scenes;
};

namespace {

void parse_output(const Node& n, psy_file_scene_data& sd) {
    if (!n.internal) raise("ExpectedInternalNode", n.offset, "Output section should be an internal node, containing at least a Path.");
    bool found = false;
    for (const Node& c : n.children) {
        if (c.internal || c.type != "Path") continue;
        std::string tc = c.contents;
        uint64_t a = 0, b = tc.size();
        while (a < b && std::strchr(" \t\r\n\f\v", tc[a])) a++;
        while (b > a && std::strchr(" \t\r\n\f\v", tc[b - 1])) b--;
        tc = tc.substr(a, b - a);
        if (tc.size() < 2) raise("IncorrectLeafData", c.offset, "File path format is incorrect.");
        if (tc.front() != '"' || tc.back() != '"') raise("IncorrectLeafData", c.offset, "File paths must be surrounded by quotes.");
        sd.output = tc.substr(1, tc.size() - 2);
        found = true;
    }
    if (!found) raise("MissingNode", n.offset, "Output section must contain a Path.");
}

void parse_render_settings(const Node& n, psy_file_scene_data& sd) {
    if (!n.internal)
        raise("ExpectedInternalNode", n.offset, "RenderSettings section should be an internal node, containing at least Resolution and SamplesPerPixel.");
    bool res = false, spp = false;
    uint64_t w = 0, h = 0, s = 0, seed = 0;
    for (const Node& c : n.children) {
        if (c.internal) continue;
        Cursor cur(c.contents);
        if (c.type == "Resolution") {
            if (!(cur.uint(0xffffffffull, w) && cur.uint(0xffffffffull, h) && cur.at_end()))
                raise("IncorrectLeafData", c.offset, "Resolution should be specified with two integers in the form '[width height]'.");
            res = true;
        } else if (c.type == "SamplesPerPixel") {
            if (!(cur.uint(0xffffffffull, s) && cur.at_end()))
                raise("IncorrectLeafData", c.offset, "SamplesPerPixel should be an integer specified in the form '[samples]'.");
            spp = true;
        } else if (c.type == "Seed") {
            if (!(cur.uint(0xffffffffull, seed) && cur.at_end()))
                raise("IncorrectLeafData", c.offset, "Seed should be an integer specified in the form '[samples]'.");
        }
    }
    if (!res || !spp) raise("MissingNode", n.offset, "RenderSettings must have both Resolution and SamplesPerPixel specified.");
    sd.settings.width = (uint32_t)w;
    sd.settings.height = (uint32_t)h;
    sd.settings.spp = (uint32_t)s;
    sd.settings.seed = (uint32_t)seed;
    sd.settings.max_bounces = 2;  // renderer.rs:599
}

void parse_camera(const Node& n, psy_file_scene_data& sd) {
    if (!n.internal) raise("ExpectedInternalNode", n.offset, "Camera section should be an internal node, containing at least Fov and Transform.");
    std::vector<float> fovs, fds, ars;
    for (const Node& c : n.children) {
        if (c.internal) continue;
        float v;
        if (c.type == "Fov" || c.type == "FocalDistance" || c.type == "ApertureRadius") {
            if (!all_f32(c.contents, 1, &v))
                raise("IncorrectLeafData", c.offset, c.type + " should be a decimal number specified in the form '[value]'.");
            if (c.type == "Fov") {
                const float k = 3.14159265358979323846f / 180.0f;  // f32::consts::PI / 180.0 in f32
                fovs.push_back(v * k);
            } else if (c.type == "FocalDistance") {
                fds.push_back(v);
            } else {
                ars.push_back(v);
            }
        } else if (c.type == "Transform") {
            float m[12];
            parse_matrix(c.contents, c.offset, "IncorrectLeafData", m);
            sd.cam_xforms.insert(sd.cam_xforms.end(), m, m + 12);
        }
    }
    // Camera::new (camera.rs:29-57): no aperture, no focus distance or a
    // zero focus distance among the samples disables the lens.
    if (sd.cam_xforms.empty()) raise("MissingNode", n.offset, "Camera has no transform(s).");
    if (fovs.empty()) raise("MissingNode", n.offset, "Camera has no fov(s).");
    if (fovs.size() > 64 || fds.size() > 64 || ars.size() > 64)
        raise("UnknownVariant", n.offset, "more than 64 time samples of a camera lens parameter are not supported");
    sd.fovs = fovs;
    if (ars.empty() || fds.empty() || std::find(fds.begin(), fds.end(), 0.0f) != fds.end()) {
        sd.apertures.assign(1, 0.0f);
        sd.focuses.assign(1, 1.0f);
    } else {
        sd.apertures = ars;
        sd.focuses = fds;
    }
}

// A light property's time samples (each `width` floats): the first goes to
// `one`; two or more are kept in the pool and returned as (count, pointer).
// No sample is an error, as a light without the property cannot render.
uint32_t light_samples(const std::vector<std::vector<float>>& v, uint32_t width, const Node& n, const char* what,
                       float* one, SamplePool& pool, const float** list) {
    if (v.empty()) raise("MissingNode", n.offset, std::string("light has no ") + what);
    if (v.size() > 64) raise("UnknownVariant", n.offset, std::string("more than 64 time samples of light ") + what);
    for (uint32_t k = 0; k < width; k++) one[k] = v[0][k];
    *list = nullptr;
    if (v.size() == 1) return 0;
    pool.floats.emplace_back();
    for (const auto& s : v) pool.floats.back().insert(pool.floats.back().end(), s.begin(), s.end());
    *list = pool.floats.back().data();
    return (uint32_t)v.size();
}
uint32_t light_colours(const std::vector<psy_color>& v, const Node& n, psy_color* one, SamplePool& pool,
                       const psy_color** list) {
    if (v.empty()) raise("MissingNode", n.offset, "light has no colour");
    if (v.size() > 64) raise("UnknownVariant", n.offset, "more than 64 time samples of light colour");
    for (const psy_color& c : v)
        if (c.kind != v[0].kind) raise("UnknownVariant", n.offset, "Cannot lerp colors with different representations.");
    *one = v[0];
    *list = nullptr;
    if (v.size() == 1) return 0;
    pool.colours.push_back(v);
    *list = pool.colours.back().data();
    return (uint32_t)v.size();
}

// parse_distant_disk_light / parse_sphere_light / parse_rectangle_light
// (psy_light.rs:16-184): Radius, Direction, Dimensions and Color leaves,
// each possibly repeated as time samples.
struct LightFields {
    std::vector<std::vector<float>> radius, direction, dims;
    std::vector<psy_color> color;
};

LightFields parse_light(const Node& n, const char* which) {
    if (!n.internal) raise("UnknownError", n.offset, std::string(which) + " should be an internal node");
    const std::string w = which;
    LightFields f;
    for (const Node& c : n.children) {
        if (c.internal) continue;
        float v[3];
        if (c.type == "Radius" && (w == "DistantDiskLight" || w == "SphereLight")) {
            if (!all_f32(c.contents, 1, v)) raise("UnknownError", c.offset, "Radius should be one number");
            f.radius.push_back({v[0]});
        } else if (c.type == "Direction" && w == "DistantDiskLight") {
            if (!all_f32(c.contents, 3, v)) raise("UnknownError", c.offset, "Direction should be three numbers");
            f.direction.push_back({v[0], v[1], v[2]});
        } else if (c.type == "Dimensions" && w == "RectangleLight") {
            if (!all_f32(c.contents, 2, v)) raise("UnknownError", c.offset, "Dimensions should be two numbers");
            f.dims.push_back({v[0], v[1]});
        } else if (c.type == "Color") {
            f.color.push_back(parse_color(c.contents, c.offset, "UnknownError"));
        }
    }
    return f;
}

void parse_world(const Node& n, psy_file_scene_data& sd) {
    if (!n.internal) raise("ExpectedInternalNode", n.offset, "World section should be an internal node, containing at least a BackgroundShader.");
    const auto bgs = children_of(n, "BackgroundShader", 0);
    if (bgs.size() != 1) raise("WrongNodeCount", n.offset, "World should have precisely one BackgroundShader section.  Found: " + std::to_string(bgs.size()));
    const Node& bg = *bgs[0];
    const auto types = children_of(bg, "Type", 0);
    if (types.size() != 1) raise("WrongNodeCount", bg.offset, "BackgroundShader should have precisely one Type specified.  Found: " + std::to_string(types.size()));
    if (types[0]->internal) raise("ExpectedLeafNode", bg.offset, "BackgroundShader's Type should be a leaf node.");
    std::string t = types[0]->contents;
    uint64_t a = 0, b = t.size();
    while (a < b && std::strchr(" \t\r\n\f\v", t[a])) a++;
    while (b > a && std::strchr(" \t\r\n\f\v", t[b - 1])) b--;
    if (t.substr(a, b - a) != "Color") raise("UnknownVariant", bg.offset, "The specified BackgroundShader Type isn't a recognized type.");
    const auto cols = children_of(bg, "Color", 0);
    if (cols.empty() || cols[0]->internal) raise("MissingNode", bg.offset, "BackgroundShader's Type is Color, but no Color is specified.");
    sd.background = parse_color(cols[0]->contents, cols[0]->offset, "IncorrectLeafData");
    for (const Node& c : n.children) {
        if (!(c.internal && c.type == "DistantDiskLight")) continue;
        const LightFields f = parse_light(c, "DistantDiskLight");
        psy_distant_light_desc d{};
        d.radius_count = light_samples(f.radius, 1, c, "radius", &d.radius, sd.pool, &d.radii);
        d.direction_count = light_samples(f.direction, 3, c, "direction", d.direction, sd.pool, &d.directions);
        d.color_count = light_colours(f.color, c, &d.color, sd.pool, &d.colors);
        sd.distant.push_back(d);
    }
}

psy_shader_desc parse_surface_shader(const Node& n) {
    const auto types = children_of(n, "Type", 2);
    if (types.empty()) raise("MissingNode", n.offset, "Expected a Type field in SurfaceShader.");
    std::string t = types[0]->contents;
    uint64_t a = 0, b = t.size();
    while (a < b && std::strchr(" \t\r\n\f\v", t[a])) a++;
    while (b > a && std::strchr(" \t\r\n\f\v", t[b - 1])) b--;
    t = t.substr(a, b - a);
    auto color = [&](const char* label) {
        const auto cs = children_of(n, "Color", 2);
        if (cs.empty()) raise("MissingNode", n.offset, std::string("Expected a Color field in ") + label + " SurfaceShader.");
        return parse_color(cs[0]->contents, cs[0]->offset, "UnknownError");
    };
    auto scalar = [&](const char* name) {
        const auto cs = children_of(n, name, 2);
        if (cs.empty()) raise("MissingNode", n.offset, std::string("Expected a ") + name + " field in GTR SurfaceShader.");
        float v;
        if (!all_f32(cs[0]->contents, 1, &v)) raise("UnknownError", cs[0]->offset, std::string(name) + " should be one number");
        return v;
    };
    psy_shader_desc s{};
    if (t == "Lambert") {
        s.kind = PSY_SHADER_LAMBERT;
        s.color = color("Lambert");
    } else if (t == "GGX") {
        s.kind = PSY_SHADER_GGX;
        s.color = color("GTR");
        s.roughness = scalar("Roughness");
        s.fresnel = scalar("Fresnel");
    } else if (t == "Emit") {
        s.kind = PSY_SHADER_EMIT;
        s.color = color("Emit");
    } else {
        raise("UnknownVariant", n.offset, "unknown SurfaceShader Type '" + t + "'");
    }
    return s;
}

// parse_mesh_surface (psy_mesh_surface.rs:25-127).
MeshData parse_mesh_surface(const Node& n) {
    MeshData m;
    std::vector<std::vector<float>> verts, normals;
    auto triples = [](const std::string& s, bool normalize) {
        std::vector<float> out;
        Cursor c(s);
        for (;;) {
            float v[3];
            Cursor probe = c;
            if (!(probe.f32(v[0]) && probe.f32(v[1]) && probe.f32(v[2]))) break;
            c = probe;
            if (normalize) {
                // Normal::normalized: v * (1 / sqrt((x*x + y*y) + z*z))
                const float inv = 1.0f / std::sqrt((v[0] * v[0] + v[1] * v[1]) + v[2] * v[2]);
                v[0] *= inv;
                v[1] *= inv;
                v[2] *= inv;
            }
            out.insert(out.end(), v, v + 3);
        }
        return out;
    };
    for (const Node* c : children_of(n, "Vertices", 2)) verts.push_back(triples(c->contents, false));
    if (verts.empty()) raise("MissingNode", n.offset, "MeshSurface has no Vertices");
    const uint64_t vc = verts[0].size() / 3;
    for (const auto& v : verts)
        if (v.size() / 3 != vc) raise("IncorrectLeafData", n.offset, "every Vertices time sample must have the same vertex count");
    for (const Node* c : children_of(n, "Normals", 2)) normals.push_back(triples(c->contents, true));
    if (!normals.empty()) {
        if (normals.size() != verts.size()) raise("IncorrectLeafData", n.offset, "Normals must have one time sample per Vertices sample");
        for (const auto& v : normals)
            if (v.size() / 3 != vc) raise("IncorrectLeafData", n.offset, "Normals must have one normal per vertex");
    }
    auto uints = [&](const char* type) {
        std::vector<uint64_t> out;
        const auto cs = children_of(n, type, 2);
        if (cs.empty()) return out;
        Cursor c(cs[0]->contents);
        uint64_t v;
        while (c.uint(~0ull, v)) out.push_back(v);
        return out;
    };
    const std::vector<uint64_t> counts = uints("FaceVertCounts"), idx = uints("FaceVertIndices");
    uint64_t ii = 0;
    for (uint64_t fvc : counts) {
        if (fvc < 3) raise("IncorrectLeafData", n.offset, "Cannot handle polygons with less than three vertices.");
        if (ii + fvc > idx.size() || ii + fvc < ii) raise("IncorrectLeafData", n.offset, "FaceVertIndices is shorter than FaceVertCounts requires");
        for (uint64_t k = 0; k + 2 < fvc; k++) {
            const uint64_t t[3] = {idx[ii], idx[ii + k + 1], idx[ii + k + 2]};
            for (uint64_t x : t) {
                if (x >= vc) raise("IncorrectLeafData", n.offset, "vertex index out of range");
                m.indices.push_back((uint32_t)x);
            }
        }
        ii += fvc;
    }
    if (vc > 0xffffffffull) raise("IncorrectLeafData", n.offset, "too many vertices");
    m.vert_count = (uint32_t)vc;
    m.time_samples = (uint32_t)verts.size();
    for (const auto& v : verts) m.verts.insert(m.verts.end(), v.begin(), v.end());
    for (const auto& v : normals) m.normals.insert(m.normals.end(), v.begin(), v.end());
    return m;
}

// parse_assembly (psy_assembly.rs:18-188).
std::unique_ptr<AsmDef> parse_assembly(const Node& n, bool root) {
    if (!n.internal) raise("UnknownError", n.offset, "Assembly should be an internal node");
    auto def = std::make_unique<AsmDef>();
    std::unordered_map<std::string, uint32_t> shader_map;
    std::unordered_map<std::string, ObjectRef> object_map;
    for (const Node& c : n.children) {
        if (c.type == "Assembly") {
            if (!c.internal || !c.has_ident) raise("UnknownError", c.offset, "Assembly should be an internal node with a name");
            if (object_map.count(c.ident))
                raise("UnknownError", c.offset, "Attempted to add assembly to another assembly with a name that already exists.");
            object_map[c.ident] = ObjectRef{KIND_ASSEMBLY, (uint32_t)def->subs.size()};
            def->subs.push_back(parse_assembly(c, false));
        } else if (c.type == "Instance") {
            if (!c.internal) raise("UnknownError", c.offset, "Instance should be an internal node");
            const auto data = children_of(c, "Data", 2);
            if (data.size() != 1) raise("UnknownError", c.offset, "Instance should have precisely one Data leaf");
            const std::string& name = data[0]->contents;
            const auto binds = children_of(c, "SurfaceShaderBind", 2);
            AsmDef::Inst inst{};
            for (const Node* t : children_of(c, "Transform", 2)) {
                float m[12];
                parse_matrix(t->contents, 0, "UnknownError", m);
                inst.xf.push_back(psy::xf_from_values(m));
            }
            auto it = object_map.find(name);
            if (it == object_map.end())
                raise("InstancedMissingData", data[0]->offset,
                      "Attempted to add instance for data with a name that doesn't exist. Data name: '" + name + "'");
            inst.kind = it->second.kind;
            inst.index = it->second.index;
            inst.shader = PSY_NONE;
            if (!binds.empty()) {
                auto sh = shader_map.find(binds[0]->contents);
                if (sh == shader_map.end())
                    raise("UnknownVariant", c.offset, "Unknown surface shader '" + binds[0]->contents + "'.");
                // An assembly instance's binding is checked but, as in
                // tracer.rs:114-128, not used.
                if (inst.kind == PSY_OBJECT_MESH) inst.shader = sh->second;
            }
            def->instances.push_back(std::move(inst));
        } else if (c.type == "SurfaceShader") {
            if (!c.internal || !c.has_ident) raise("UnknownError", c.offset, "SurfaceShader should be an internal node with a name");
            if (shader_map.count(c.ident))
                raise("UnknownError", c.offset, "Attempted to add surface shader to assembly with a name that already exists.");
            shader_map[c.ident] = (uint32_t)def->shaders.size();
            def->shaders.push_back(parse_surface_shader(c));
        } else if (c.type == "MeshSurface" || c.type == "SphereLight" || c.type == "RectangleLight") {
            if (!c.internal || !c.has_ident) raise("UnknownError", c.offset, c.type + " should be an internal node with a name");
            if (object_map.count(c.ident))
                raise("UnknownError", c.offset, "Attempted to add object to assembly with a name that already exists.");
            ObjectRef ref{};
            if (c.type == "MeshSurface") {
                ref = ObjectRef{PSY_OBJECT_MESH, (uint32_t)def->meshes.size()};
                def->meshes.push_back(parse_mesh_surface(c));
            } else if (c.type == "SphereLight") {
                const LightFields f = parse_light(c, "SphereLight");
                ref = ObjectRef{PSY_OBJECT_SPHERE_LIGHT, (uint32_t)def->spheres.size()};
                psy_sphere_light_desc l{};
                l.radius_count = light_samples(f.radius, 1, c, "radius", &l.radius, def->pool, &l.radii);
                l.color_count = light_colours(f.color, c, &l.color, def->pool, &l.colors);
                def->spheres.push_back(l);
            } else {
                const LightFields f = parse_light(c, "RectangleLight");
                ref = ObjectRef{PSY_OBJECT_RECT_LIGHT, (uint32_t)def->rects.size()};
                psy_rect_light_desc l{};
                float wh[2];
                l.dims_count = light_samples(f.dims, 2, c, "dimensions", wh, def->pool, &l.dims);
                l.width = wh[0];
                l.height = wh[1];
                l.color_count = light_colours(f.color, c, &l.color, def->pool, &l.colors);
                def->rects.push_back(l);
            }
            object_map[c.ident] = ref;
        }
        // Other node types are ignored, as in the reference.
    }
    return def;
}

// Shaders and meshes of every assembly go to the scene's lists depth first
// in definition order (the root's first); shader indices are global, as the
// renderer takes one shader list.
void place_definitions(AsmDef& d, psy_file_scene_data& sd) {
    d.shader_base = (uint32_t)sd.shaders.size();
    sd.shaders.insert(sd.shaders.end(), d.shaders.begin(), d.shaders.end());
    d.mesh_base = (uint32_t)sd.mesh_data.size();
    for (MeshData& m : d.meshes) sd.mesh_data.push_back(std::move(m));
    for (auto& sub : d.subs) place_definitions(*sub, sd);
}

// The descriptors of an assembly and its sub-assemblies, nested as in the
// file (AssemblyBuilder::add_assembly, assembly.rs:197-282): meshes, lights
// and sub-assemblies by their index in this assembly, each instance with its
// own transforms.  A light instance of the root assembly without a
// transform gets the identity (the reference would lerp an empty transform
// list); elsewhere it keeps the enclosing assembly's transforms.
std::unique_ptr<AsmOut> build_out(AsmDef& d, bool root, const psy_file_scene_data& sd) {
    auto o = std::make_unique<AsmOut>();
    for (uint32_t k = 0; k < (uint32_t)d.meshes.size(); k++) {
        const MeshData& m = sd.mesh_data[d.mesh_base + k];
        o->meshes.push_back(psy_mesh_desc{m.vert_count, m.time_samples, m.verts.data(),
                                          m.normals.empty() ? nullptr : m.normals.data(),
                                          (uint32_t)(m.indices.size() / 3), m.indices.data()});
    }
    for (auto& sub : d.subs) o->subs.push_back(build_out(*sub, false, sd));
    for (auto& sub : o->subs) o->sub_descs.push_back(sub->desc);
    for (const AsmDef::Inst& inst : d.instances) {
        psy_instance_desc out{};
        out.kind = inst.kind;  // KIND_ASSEMBLY is PSY_OBJECT_ASSEMBLY
        out.mesh = inst.index;
        out.shader = inst.shader == PSY_NONE ? PSY_NONE : d.shader_base + inst.shader;
        std::vector<psy::Xf> xf = inst.xf;
        if (xf.empty() && root && inst.kind != PSY_OBJECT_MESH && inst.kind != KIND_ASSEMBLY)
            xf.push_back(psy::xf_identity());
        out.xform_start = (uint32_t)(o->xforms.size() / 12);
        out.xform_count = (uint32_t)xf.size();
        for (const psy::Xf& x : xf)
            for (int r = 0; r < 3; r++)
                for (int c = 0; c < 4; c++) o->xforms.push_back(x.m[r][c]);
        o->instances.push_back(out);
    }
    psy_scene_desc& ds = o->desc;
    ds.mesh_count = (uint32_t)o->meshes.size();
    ds.meshes = o->meshes.data();
    ds.instance_count = (uint32_t)o->instances.size();
    ds.instances = o->instances.data();
    ds.xform_count = (uint32_t)(o->xforms.size() / 12);
    ds.xforms = o->xforms.data();
    ds.rect_light_count = (uint32_t)d.rects.size();
    ds.rect_lights = d.rects.data();
    ds.sphere_light_count = (uint32_t)d.spheres.size();
    ds.sphere_lights = d.spheres.data();
    ds.assembly_count = (uint32_t)o->sub_descs.size();
    ds.assemblies = o->sub_descs.data();
    return o;
}

void read_assembly(const Node& n, psy_file_scene_data& sd) {
    sd.root_def = parse_assembly(n, true);
    place_definitions(*sd.root_def, sd);
    sd.root_out = build_out(*sd.root_def, true, sd);
}

// parse_scene (psy.rs:99-200).
std::unique_ptr<psy_file_scene_data> parse_scene(const Node& n) {
    auto sd = std::make_unique<psy_file_scene_data>();
    static const struct {
        const char* type;
        const char* msg;
    } sections[] = {
        {"Output", "Scene should have precisely one Output section."},
        {"RenderSettings", "Scene should have precisely one RenderSettings section."},
        {"Camera", "Scene should have precisely one Camera section."},
        {"World", "Scene should have precisely one World section."},
        {"Assembly", "Scene should have precisely one Root Assembly section."},
    };
    for (const auto& s : sections) {
        const auto found = children_of(n, s.type, 0);
        if (found.size() != 1) raise("WrongNodeCount", n.offset, std::string(s.msg) + "  Found: " + std::to_string(found.size()));
    }
    parse_output(*children_of(n, "Output", 0)[0], *sd);
    parse_render_settings(*children_of(n, "RenderSettings", 0)[0], *sd);
    parse_camera(*children_of(n, "Camera", 0)[0], *sd);
    parse_world(*children_of(n, "World", 0)[0], *sd);
    read_assembly(*children_of(n, "Assembly", 0)[0], *sd);
    if (n.has_ident) sd->name = n.ident.substr(1);
    return sd;
}

uint64_t line_of(const char* text, uint64_t len, uint64_t offset) {
    uint64_t line = 1;
    for (uint64_t i = 0; i < offset && i < len; i++)
        if (text[i] == '\n') line++;
    return line;
}

}  // namespace

extern "C" {

int psy_file_parse(const char* text, uint64_t len, psy_file** out) {
    using psy::api::fail;
    if (!out) return fail(PSY_E_INVALID, "out is NULL");
    *out = nullptr;
    if (!text && len) return fail(PSY_E_INVALID, "text is NULL");
    try {
        TreeReader reader(text ? text : "", len);
        const Node root = reader.read_root();
        auto f = std::make_unique<psy_file>();
        for (const Node& c : root.children)
            if (c.type == "Scene") f->scenes.push_back(parse_scene(c));
        *out = f.release();
        return PSY_OK;
    } catch (const ParseFail& e) {
        return fail(PSY_E_PARSE, e.kind + " at byte " + std::to_string(e.offset) + " (line " +
                                     std::to_string(line_of(text, len, e.offset)) + "): " + e.message);
    } catch (const std::bad_alloc&) {
        return fail(PSY_E_NOMEM, "out of host memory while reading the scene file");
    }
}

int psy_file_destroy(psy_file* f) {
    delete f;
    return PSY_OK;
}

uint32_t psy_file_scene_count(const psy_file* f) { return f ? (uint32_t)f->scenes.size() : 0u; }

int psy_file_get_scene(const psy_file* f, uint32_t index, psy_file_scene* out) {
    using psy::api::fail;
    if (!f || !out) return fail(PSY_E_INVALID, "NULL argument");
    if (index >= f->scenes.size()) return fail(PSY_E_RANGE, "scene index out of range");
    const psy_file_scene_data& sd = *f->scenes[index];
    psy_file_scene s{};
    s.name = sd.name.c_str();
    s.output_path = sd.output.c_str();
    s.scene = sd.root_out->desc;
    s.world.shader_count = (uint32_t)sd.shaders.size();
    s.world.shaders = sd.shaders.data();
    s.world.background = sd.background;
    s.world.distant_light_count = (uint32_t)sd.distant.size();
    s.world.distant_lights = sd.distant.data();
    s.camera.xform_count = (uint32_t)(sd.cam_xforms.size() / 12);
    s.camera.xforms = sd.cam_xforms.data();
    s.camera.fov = sd.fovs[0];
    s.camera.aperture_radius = sd.apertures[0];
    s.camera.focus_distance = sd.focuses[0];
    s.camera.fov_count = (uint32_t)sd.fovs.size();
    s.camera.fovs = sd.fovs.data();
    s.camera.aperture_count = (uint32_t)sd.apertures.size();
    s.camera.aperture_radii = sd.apertures.data();
    s.camera.focus_count = (uint32_t)sd.focuses.size();
    s.camera.focus_distances = sd.focuses.data();
    s.settings = sd.settings;
    *out = s;
    return PSY_OK;
}

int64_t psy_file_dump_tree(const char* text, uint64_t len, char* out, uint64_t cap) {
    using psy::api::fail;
    if (!text && len) return fail(PSY_E_INVALID, "text is NULL");
    if (!out && cap) return fail(PSY_E_INVALID, "out is NULL");
    try {
        TreeReader reader(text ? text : "", len);
        std::string s;
        dump(reader.read_root(), 0, s);
        if (cap) std::memcpy(out, s.data(), std::min<uint64_t>(cap, s.size()));
        return (int64_t)s.size();
    } catch (const ParseFail& e) {
        return fail(PSY_E_PARSE, e.kind + " at byte " + std::to_string(e.offset) + " (line " +
                                     std::to_string(line_of(text, len, e.offset)) + "): " + e.message);
    }
}

}  // extern "C"
