// The .pbrt scene front end: tokenizer, parameter lists, graphics-state stack
// and the directives the path tracer consumes, producing a pbrtm_scene_desc.
//
// Mirrors the reference's SceneBuilder (scene/scene_builder.cu:316-460 directive
// dispatch, :460-720 handlers) and ParameterDictionary (scene/parameter_dictionary.cu)
// for: LookAt, Camera "perspective", Film "rgb", Sampler, Integrator "path",
// PixelFilter "mitchell", WorldBegin, AttributeBegin/End, Translate, Scale,
// Rotate, Transform, ConcatTransform, Identity, CoordSysTransform, ReverseOrientation,
// Material / MakeNamedMaterial / NamedMaterial ("diffuse", "conductor",
// "dielectric"), AreaLightSource "diffuse", LightSource "infinite" (uniform),
// Shape "trianglemesh" / "sphere" / "disk", ObjectBegin / ObjectEnd /
// ObjectInstance (triangle meshes), Include.
// Anything else that changes the image is reported as unsupported rather than
// silently approximated.
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <new>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/pbrtm.h"
#include "pathtrace.h"

namespace {

struct Tok {
    enum Kind { Ident, Str, Num, Open, Close } kind;
    std::string s;
    int line;
};

std::vector<Tok> tokenize(const std::string &text) {
    std::vector<Tok> out;
    int line = 1;
    size_t i = 0;
    while (i < text.size()) {
        const char c = text[i];
        if (c == '\n') {
            ++line;
            ++i;
        } else if (c == ' ' || c == '\t' || c == '\r') {
            ++i;
        } else if (c == '#') {
            while (i < text.size() && text[i] != '\n') ++i;
        } else if (c == '[') {
            out.push_back({Tok::Open, "[", line});
            ++i;
        } else if (c == ']') {
            out.push_back({Tok::Close, "]", line});
            ++i;
        } else if (c == '"') {
            size_t j = i + 1;
            while (j < text.size() && text[j] != '"') {
                if (text[j] == '\n') throw std::runtime_error("line " + std::to_string(line) + ": unterminated string");
                ++j;
            }
            if (j >= text.size()) throw std::runtime_error("line " + std::to_string(line) + ": unterminated string");
            out.push_back({Tok::Str, text.substr(i + 1, j - i - 1), line});
            i = j + 1;
        } else if ((c >= '0' && c <= '9') || c == '-' || c == '+' || c == '.') {
            size_t j = i;
            while (j < text.size() && (std::isalnum((unsigned char)text[j]) || text[j] == '.' || text[j] == '-' ||
                                       text[j] == '+'))
                ++j;
            out.push_back({Tok::Num, text.substr(i, j - i), line});
            i = j;
        } else if (std::isalpha((unsigned char)c) || c == '_') {
            size_t j = i;
            while (j < text.size() && (std::isalnum((unsigned char)text[j]) || text[j] == '_')) ++j;
            std::string w = text.substr(i, j - i);
            // bare true/false are string values (scene/lexer.h)
            out.push_back({(w == "true" || w == "false") ? Tok::Str : Tok::Ident, w, line});
            i = j;
        } else {
            throw std::runtime_error("line " + std::to_string(line) + ": unexpected character '" + std::string(1, c) + "'");
        }
    }
    return out;
}

double to_num(const Tok &t) {
    try {
        size_t used = 0;
        const double v = std::stod(t.s, &used);
        if (used != t.s.size()) throw std::invalid_argument(t.s);
        return v;
    } catch (const std::exception &) {
        throw std::runtime_error("line " + std::to_string(t.line) + ": bad number '" + t.s + "'");
    }
}

struct Param {
    std::string type, name;
    std::vector<double> nums;
    std::vector<std::string> strs;
    bool used = false;
};

struct Params {
    std::vector<Param> list;
    Param *find(const std::string &name, const char *type = nullptr) {
        for (auto &p : list)
            if (p.name == name && (!type || p.type == type)) {
                p.used = true;
                return &p;
            }
        return nullptr;
    }
    float get_float(const std::string &n, float def) {
        Param *p = find(n, "float");
        return p && !p->nums.empty() ? float(p->nums[0]) : def;
    }
    int get_int(const std::string &n, int def) {
        Param *p = find(n, "integer");
        return p && !p->nums.empty() ? int(p->nums[0]) : def;
    }
    bool get_bool(const std::string &n, bool def) {
        Param *p = find(n, "bool");
        return p && !p->strs.empty() ? p->strs[0] == "true" : def;
    }
    std::string get_string(const std::string &n, const std::string &def) {
        Param *p = find(n, "string");
        return p && !p->strs.empty() ? p->strs[0] : def;
    }
    void check_used(const std::string &what, const std::vector<std::string> &ignorable = {}) {
        for (auto &p : list) {
            if (p.used) continue;
            bool ok = false;
            for (auto &n : ignorable) ok = ok || n == p.name;
            if (!ok) throw std::runtime_error(what + ": parameter \"" + p.type + " " + p.name + "\" is not supported");
        }
    }
};

// Row-major 4x4 products with the reference's FMA chain (squared_matrix.h:98-109).
struct M4 {
    float m[16];
};
M4 ident() {
    M4 r{};
    r.m[0] = r.m[5] = r.m[10] = r.m[15] = 1.f;
    return r;
}
M4 mul(const M4 &a, const M4 &b) {
    M4 r;
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) {
            float acc = 0;
            for (int k = 0; k < 4; ++k) acc = std::fma(a.m[4 * i + k], b.m[4 * k + j], acc);
            r.m[4 * i + j] = acc;
        }
    return r;
}

struct Material {
    pbrtm_material m;
};

M4 transpose(const M4 &a) {
    M4 r;
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) r.m[4 * i + j] = a.m[4 * j + i];
    return r;
}

// The CTM keeps its inverse as the reference's Transform does (m, inv_m): products
// compose inverses in reverse order; Translate, Scale and Rotate have their own
// (transform.h:61-132), a general matrix the 4x4 inverse (squared_matrix.cu).
struct GState {
    M4 ctm = ident();
    M4 ctm_inv = ident();
    bool ctm_identity = true;
    int material = -1;  // index into materials; -1 = the default diffuse 0.5
    bool has_light = false;
    pbrtm_area_light light{};
    bool reverse = false;
};

constexpr float kPiF = 3.14159274101257324f;

}  // namespace

struct pbrtm_pbrt_file {
    pbrtm_scene_desc desc{};
    std::vector<float> tri, xforms, xforms_inv;
    std::vector<uint32_t> tri_material, tri_xform, tri_shape;
    std::vector<pbrtm_shape> shapes;
    std::vector<int32_t> tri_light;
    std::vector<uint8_t> tri_reverse;
    std::vector<pbrtm_material> materials;
    std::vector<pbrtm_area_light> lights;
    std::vector<pbrtm_infinite_light> infinite_lights;
    uint32_t spp = 4, max_depth = 5, integrator = PBRTM_INTEGRATOR_PATH;
    std::string filename, error;
};

namespace {

class Builder {
  public:
    explicit Builder(pbrtm_pbrt_file &f) : F(f) {}

    void parse_file(const std::string &path) {
        std::ifstream in(path);
        if (!in) throw std::runtime_error("cannot open " + path);
        std::stringstream ss;
        ss << in.rdbuf();
        const std::string dir = path.find('/') == std::string::npos ? "" : path.substr(0, path.rfind('/') + 1);
        run(tokenize(ss.str()), dir);
    }

    void finish() {
        if (!have_film) throw std::runtime_error("no Film directive");
        pbrtm_scene_desc &d = F.desc;
        for (int k = 0; k < 3; ++k) d.eye[k] = eye[k], d.look[k] = look[k], d.up[k] = up[k];
        d.fov = fov;
        d.width = width;
        d.height = height;
        d.num_tris = F.tri_material.size();
        d.tri = F.tri.data();
        d.tri_material = F.tri_material.data();
        d.tri_light = F.tri_light.data();
        d.tri_reverse = F.tri_reverse.data();
        d.num_materials = uint32_t(F.materials.size());
        d.materials = F.materials.data();
        d.num_lights = uint32_t(F.lights.size());
        d.lights = F.lights.empty() ? nullptr : F.lights.data();
        d.num_xforms = uint32_t(F.xforms.size() / 16);
        d.xforms = F.xforms.empty() ? nullptr : F.xforms.data();
        d.xforms_inv = F.xforms_inv.empty() ? nullptr : F.xforms_inv.data();
        d.tri_xform = F.tri_xform.data();
        d.num_shapes = uint32_t(F.shapes.size());
        d.shapes = F.shapes.empty() ? nullptr : F.shapes.data();
        d.tri_shape = F.shapes.empty() ? nullptr : F.tri_shape.data();
        d.num_infinite_lights = uint32_t(F.infinite_lights.size());
        d.infinite_lights = F.infinite_lights.empty() ? nullptr : F.infinite_lights.data();
        const int s = int(std::sqrt(double(spp_asked)));  // stratified: spp rounded to a square (scene_builder.cu:283-287)
        F.spp = uint32_t(s * s);
        F.max_depth = uint32_t(max_depth);
    }

  private:
    pbrtm_pbrt_file &F;
    GState gs;
    std::vector<GState> stack;
    std::map<std::string, int> named_materials;
    std::map<std::string, std::pair<M4, M4>> named_coords;
    bool world = false, have_film = false, camera_set = false, lookat_only = true;
    float eye[3] = {0, 0, 0}, look[3] = {0, 0, 1}, up[3] = {0, 1, 0};
    float fov = 90.f;
    uint32_t width = 1280, height = 720;
    int spp_asked = 4, max_depth = 5;

    static std::string where(const Tok &t) { return "line " + std::to_string(t.line) + ": "; }

    // "type name" value-or-[values] pairs up to the next directive (parameter_dictionary.cu).
    static Params read_params(const std::vector<Tok> &T, size_t &i) {
        Params P;
        while (i < T.size() && T[i].kind == Tok::Str) {
            std::istringstream decl(T[i].s);
            Param p;
            decl >> p.type >> p.name;
            if (p.name.empty()) throw std::runtime_error(where(T[i]) + "expected \"type name\", got \"" + T[i].s + "\"");
            ++i;
            auto take = [&](const Tok &v) {
                if (v.kind == Tok::Num) p.nums.push_back(to_num(v));
                else if (v.kind == Tok::Str) p.strs.push_back(v.s);
                else throw std::runtime_error(where(v) + "bad parameter value");
            };
            if (i < T.size() && T[i].kind == Tok::Open) {
                ++i;
                while (i < T.size() && T[i].kind != Tok::Close) take(T[i++]);
                if (i >= T.size()) throw std::runtime_error("unterminated [");
                ++i;
            } else if (i < T.size()) {
                take(T[i++]);
            }
            P.list.push_back(std::move(p));
        }
        return P;
    }
    static std::vector<double> read_nums(const std::vector<Tok> &T, size_t &i, size_t n) {
        std::vector<double> v;
        const bool br = i < T.size() && T[i].kind == Tok::Open;
        if (br) ++i;
        while (i < T.size() && T[i].kind == Tok::Num && v.size() < n) v.push_back(to_num(T[i++]));
        if (br) {
            if (i >= T.size() || T[i].kind != Tok::Close) throw std::runtime_error("expected ]");
            ++i;
        }
        if (v.size() != n) throw std::runtime_error("expected " + std::to_string(n) + " numbers");
        return v;
    }
    static std::string read_str(const std::vector<Tok> &T, size_t &i, const char *what) {
        if (i >= T.size() || T[i].kind != Tok::Str) throw std::runtime_error(std::string(what) + ": expected a quoted name");
        return T[i++].s;
    }

    void concat(const M4 &t, const M4 &t_inv) {
        gs.ctm = mul(gs.ctm, t);
        gs.ctm_inv = mul(t_inv, gs.ctm_inv);
        gs.ctm_identity = false;
        if (!world) lookat_only = false;
    }
    static M4 inverse_of(const M4 &m) {
        M4 r;
        if (!pbrtm::pt::mat4_inverse_ref(m.m, r.m)) throw std::runtime_error("singular transform matrix");
        return r;
    }
    // The CTM as a transform index (UINT32_MAX = identity).
    uint32_t ctm_index() {
        if (gs.ctm_identity) return 0xFFFFFFFFu;
        const uint32_t xf = uint32_t(F.xforms.size() / 16);
        F.xforms.insert(F.xforms.end(), gs.ctm.m, gs.ctm.m + 16);
        F.xforms_inv.insert(F.xforms_inv.end(), gs.ctm_inv.m, gs.ctm_inv.m + 16);
        return xf;
    }
    // One primitive slot (the triangle list orders the aggregate and the lights).
    void add_slot(const float *v9, uint32_t mat, uint32_t xf, uint32_t shape) {
        F.tri.insert(F.tri.end(), v9, v9 + 9);
        F.tri_material.push_back(mat);
        F.tri_xform.push_back(xf);
        F.tri_reverse.push_back(gs.reverse ? 1 : 0);
        F.tri_shape.push_back(shape);
        if (gs.has_light) {
            F.tri_light.push_back(int32_t(F.lights.size()));
            F.lights.push_back(gs.light);
        } else {
            F.tri_light.push_back(-1);
        }
    }
    // A triangle of an ObjectBegin block, kept until ObjectInstance.
    struct ProtoTri {
        float v[9];
        uint32_t mat, xf;
        bool reverse;
    };
    struct Definition {
        std::vector<ProtoTri> tris;
    };
    std::map<std::string, Definition> definitions;
    Definition *active_def = nullptr;
    std::string active_name;

    pbrtm_material make_material(const std::string &type, Params &P, const Tok &at) {
        pbrtm_material m{};
        m.remap_roughness = 1;
        if (type == "diffuse") {
            m.type = PBRTM_MATERIAL_DIFFUSE;
            Param *r = P.find("reflectance", "rgb");
            if (r) {
                if (r->nums.size() != 3) throw std::runtime_error(where(at) + "rgb reflectance needs 3 values");
                for (int k = 0; k < 3; ++k) m.rgb[k] = float(r->nums[k]);
            } else {
                for (int k = 0; k < 3; ++k) m.rgb[k] = 0.5f;  // scene_builder.cu:173-174 / diffuse_material.cu:11-13
            }
            P.check_used("Material \"diffuse\"");
        } else if (type == "conductor") {
            m.type = PBRTM_MATERIAL_CONDUCTOR;
            static const char *metals[4] = {"Ag", "Al", "Au", "Cu"};
            int me = 3, mk = 3;  // default copper (conductor_material.cu:23-33)
            auto metal_of = [&](const char *key, const char *suffix) {
                Param *p = P.find(key, "spectrum");
                if (!p) return -1;
                if (p->strs.size() != 1) throw std::runtime_error(where(at) + "only named metal spectra are supported");
                for (int k = 0; k < 4; ++k)
                    if (p->strs[0] == std::string("metal-") + metals[k] + suffix) return k;
                throw std::runtime_error(where(at) + "unknown spectrum \"" + p->strs[0] + "\"");
            };
            const int e = metal_of("eta", "-eta"), k = metal_of("k", "-k");
            if (e >= 0) me = e;
            if (k >= 0) mk = k;
            if (me != mk) throw std::runtime_error(where(at) + "eta and k must name the same metal");
            m.metal = uint32_t(me);
            const float r = P.get_float("roughness", 0.f);
            m.uroughness = P.get_float("uroughness", r);
            m.vroughness = P.get_float("vroughness", r);
            m.remap_roughness = P.get_bool("remaproughness", true) ? 1 : 0;
            P.check_used("Material \"conductor\"");
        } else if (type == "dielectric") {
            m.type = PBRTM_MATERIAL_DIELECTRIC;
            m.eta = 1.5f;
            if (Param *p = P.find("eta", "float")) m.eta = float(p->nums.at(0));
            if (Param *p = P.find("eta", "spectrum")) {
                if (p->strs.size() == 1 && p->strs[0] == "glass-BK7") m.eta_named = 1;
                else throw std::runtime_error(where(at) + "dielectric eta: only a float or \"glass-BK7\"");
            }
            const float r = P.get_float("roughness", 0.f);
            m.uroughness = P.get_float("uroughness", r);
            m.vroughness = P.get_float("vroughness", r);
            m.remap_roughness = P.get_bool("remaproughness", true) ? 1 : 0;
            P.check_used("Material \"dielectric\"");
        } else {
            throw std::runtime_error(where(at) + "material \"" + type + "\" is not supported");
        }
        return m;
    }

    int default_material() {
        if (default_mat < 0) {
            pbrtm_material m{};
            m.type = PBRTM_MATERIAL_DIFFUSE;
            m.rgb[0] = m.rgb[1] = m.rgb[2] = 0.5f;
            m.remap_roughness = 1;
            default_mat = int(F.materials.size());
            F.materials.push_back(m);
        }
        return default_mat;
    }
    int default_mat = -1;

    void shape(const std::string &type, Params &P, const Tok &at) {
        if (type == "sphere" || type == "disk") {
            // Sphere::create (shapes/sphere.cu:8-22), Disk::create (shapes/disk.cu:6-25)
            if (active_def) throw std::runtime_error(where(at) + "only triangle meshes are supported in ObjectBegin");
            pbrtm_shape q{};
            q.radius = P.get_float("radius", 1.f);
            q.phimax = P.get_float("phimax", 360.f);
            if (type == "sphere") {
                q.type = PBRTM_SHAPE_SPHERE;
                q.zmin = P.get_float("zmin", -q.radius);
                q.zmax = P.get_float("zmax", q.radius);
            } else {
                q.type = PBRTM_SHAPE_DISK;
                q.height = P.get_float("height", 0.f);
                q.innerradius = P.get_float("innerradius", 0.f);
            }
            P.check_used("Shape \"" + type + "\"", {"alpha"});
            q.xform = ctm_index();
            F.shapes.push_back(q);
            const float zero[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
            add_slot(zero, uint32_t(gs.material >= 0 ? gs.material : default_material()), q.xform,
                     uint32_t(F.shapes.size() - 1));
            return;
        }
        if (type != "trianglemesh") throw std::runtime_error(where(at) + "shape \"" + type + "\" is not supported");
        Param *idx = P.find("indices", "integer");
        Param *pts = P.find("P", "point3");
        if (!pts) pts = P.find("P", "point");
        if (!pts || pts->nums.size() % 3) throw std::runtime_error(where(at) + "trianglemesh needs \"point3 P\"");
        std::vector<long> ind;
        if (idx) {
            for (double v : idx->nums) ind.push_back(long(v));
        } else if (pts->nums.size() == 9) {
            ind = {0, 1, 2};
        } else {
            throw std::runtime_error(where(at) + "trianglemesh needs \"integer indices\"");
        }
        if (ind.size() % 3) throw std::runtime_error(where(at) + "indices are not a triangle list");
        P.check_used("Shape \"trianglemesh\"");
        const long nv = long(pts->nums.size() / 3);
        if (active_def && gs.has_light)
            throw std::runtime_error(where(at) + "area lights are not supported with object instancing");
        const uint32_t xf = ctm_index();
        const uint32_t mat = uint32_t(gs.material >= 0 ? gs.material : default_material());
        for (size_t t = 0; t < ind.size(); t += 3) {
            float v9[9];
            for (int v = 0; v < 3; ++v) {
                const long k = ind[t + v];
                if (k < 0 || k >= nv) throw std::runtime_error(where(at) + "vertex index out of range");
                for (int c = 0; c < 3; ++c) v9[3 * v + c] = float(pts->nums[3 * k + c]);
            }
            if (active_def) {
                ProtoTri p;
                std::memcpy(p.v, v9, sizeof(v9));
                p.mat = mat, p.xf = xf, p.reverse = gs.reverse;
                active_def->tris.push_back(p);
            } else {
                add_slot(v9, mat, xf, 0xFFFFFFFFu);
            }
        }
    }

    // ObjectInstance (scene_builder.cu:775-810): the definition's primitives with
    // render_from_instance = render_from_world * CTM * world_from_render; the library
    // adds them as they are when that is the identity, else as TransformedPrimitives.
    void instance(const std::string &name, const Tok &at) {
        auto it = definitions.find(name);
        if (it == definitions.end()) throw std::runtime_error(where(at) + "object \"" + name + "\" not found");
        if (it->second.tris.empty()) return;
        pbrtm_shape q{};
        q.type = PBRTM_SHAPE_INSTANCE;
        q.xform = ctm_index();
        F.shapes.push_back(q);
        const uint32_t sk = uint32_t(F.shapes.size() - 1);
        const GState saved = gs;
        gs.has_light = false;
        for (const ProtoTri &p : it->second.tris) {
            gs.reverse = p.reverse;
            add_slot(p.v, p.mat, p.xf, sk);
        }
        gs = saved;
    }

    void run(const std::vector<Tok> &T, const std::string &dir) {
        size_t i = 0;
        while (i < T.size()) {
            const Tok &t = T[i];
            if (t.kind != Tok::Ident) throw std::runtime_error(where(t) + "expected a directive, got \"" + t.s + "\"");
            const std::string &k = t.s;
            ++i;
            if (k == "LookAt") {
                const auto v = read_nums(T, i, 9);
                if (camera_set || !gs.ctm_identity || world) lookat_only = false;
                for (int c = 0; c < 3; ++c) eye[c] = float(v[c]), look[c] = float(v[3 + c]), up[c] = float(v[6 + c]);
                gs.ctm_identity = false;
            } else if (k == "Camera") {
                const std::string type = read_str(T, i, "Camera");
                Params P = read_params(T, i);
                if (type != "perspective") throw std::runtime_error(where(t) + "camera \"" + type + "\" is not supported");
                if (!lookat_only) throw std::runtime_error(where(t) + "only a LookAt may precede the Camera");
                fov = P.get_float("fov", 90.f);
                if (P.get_float("lensradius", 0.f) > 0.f) throw std::runtime_error(where(t) + "thin-lens cameras are not supported");
                P.find("focaldistance");
                P.check_used("Camera", {"screenwindow", "frameaspectratio"});
                camera_set = true;
            } else if (k == "Film") {
                read_str(T, i, "Film");
                Params P = read_params(T, i);
                width = uint32_t(P.get_int("xresolution", 1280));
                height = uint32_t(P.get_int("yresolution", 720));
                F.filename = P.get_string("filename", "pbrt.png");
                P.check_used("Film", {"iso", "whitebalance", "sensor", "savefp16"});
                have_film = true;
            } else if (k == "Sampler") {
                read_str(T, i, "Sampler");
                Params P = read_params(T, i);
                spp_asked = P.get_int("pixelsamples", 4);
                P.check_used("Sampler", {"jitter", "seed", "randomization"});
            } else if (k == "Integrator") {
                const std::string type = read_str(T, i, "Integrator");
                Params P = read_params(T, i);
                // scene_builder.cu:289-313: "path"/"volpath" is the wavefront integrator,
                // the megakernel ones go to MegakernelIntegrator::create
                if (pbrtm_integrator_from_name(type.c_str(), &F.integrator) != PBRTM_OK)
                    throw std::runtime_error(where(t) + "integrator \"" + type + "\" is not supported");
                max_depth = P.get_int("maxdepth", 5);
                if (P.get_bool("regularize", false)) throw std::runtime_error(where(t) + "regularize is not supported");
                P.check_used("Integrator", {"lightsampler", "cossample", "maxdistance"});
            } else if (k == "PixelFilter") {
                const std::string type = read_str(T, i, "PixelFilter");
                Params P = read_params(T, i);
                if (type != "mitchell" || !P.list.empty())
                    throw std::runtime_error(where(t) + "only the default Mitchell filter is supported");
            } else if (k == "WorldBegin") {
                world = true;
                gs.ctm = gs.ctm_inv = ident();
                gs.ctm_identity = true;
                named_coords["world"] = {gs.ctm, gs.ctm_inv};
            } else if (k == "AttributeBegin" || k == "TransformBegin") {
                stack.push_back(gs);
            } else if (k == "AttributeEnd" || k == "TransformEnd") {
                if (stack.empty()) throw std::runtime_error(where(t) + "unmatched " + k);
                if (k == "TransformEnd") {
                    gs.ctm = stack.back().ctm;
                    gs.ctm_inv = stack.back().ctm_inv;
                    gs.ctm_identity = stack.back().ctm_identity;
                } else {
                    gs = stack.back();
                }
                stack.pop_back();
            } else if (k == "Translate") {
                const auto v = read_nums(T, i, 3);
                M4 m = ident(), mi = ident();
                m.m[3] = float(v[0]), m.m[7] = float(v[1]), m.m[11] = float(v[2]);
                mi.m[3] = -float(v[0]), mi.m[7] = -float(v[1]), mi.m[11] = -float(v[2]);
                concat(m, mi);
            } else if (k == "Scale") {
                const auto v = read_nums(T, i, 3);
                M4 m = ident(), mi = ident();
                m.m[0] = float(v[0]), m.m[5] = float(v[1]), m.m[10] = float(v[2]);
                mi.m[0] = float(1.0) / float(v[0]), mi.m[5] = float(1.0) / float(v[1]), mi.m[10] = float(1.0) / float(v[2]);
                concat(m, mi);
            } else if (k == "Rotate") {
                // Transform::rotate (transform.h:61-88); sin and cos rounded from double
                const auto v = read_nums(T, i, 4);
                const float rad = float(double(kPiF) / 180.0 * double(float(v[0])));
                const float s = float(std::sin(double(rad))), c = float(std::cos(double(rad)));
                float ax = float(v[1]), ay = float(v[2]), az = float(v[3]);
                const float len = std::sqrt(ax * ax + ay * ay + az * az);
                ax /= len, ay /= len, az /= len;
                M4 m = ident();
                m.m[0] = float(ax * ax + (1.0 - ax * ax) * c);
                m.m[1] = float(ax * ay * (1.0 - c) - az * s);
                m.m[2] = float(ax * az * (1.0 - c) + ay * s);
                m.m[4] = float(ax * ay * (1.0 - c) + az * s);
                m.m[5] = float(ay * ay + (1.0 - ay * ay) * c);
                m.m[6] = float(ay * az * (1.0 - c) - ax * s);
                m.m[8] = float(ax * az * (1.0 - c) - ay * s);
                m.m[9] = float(ay * az * (1.0 - c) + ax * s);
                m.m[10] = float(az * az + (1.0 - az * az) * c);
                concat(m, transpose(m));
            } else if (k == "Transform" || k == "ConcatTransform") {
                const auto v = read_nums(T, i, 16);
                M4 m;  // the file lists the matrix column-major (scene_builder.cu: transpose)
                for (int r = 0; r < 4; ++r)
                    for (int c = 0; c < 4; ++c) m.m[4 * r + c] = float(v[4 * c + r]);
                if (k == "Transform") {
                    gs.ctm = m;
                    gs.ctm_inv = inverse_of(m);
                    gs.ctm_identity = false;
                    if (!world) lookat_only = false;
                } else {
                    concat(m, inverse_of(m));
                }
            } else if (k == "Identity") {
                gs.ctm = gs.ctm_inv = ident();
                gs.ctm_identity = true;
            } else if (k == "CoordSysTransform") {
                const std::string name = read_str(T, i, "CoordSysTransform");
                auto it = named_coords.find(name);
                if (it == named_coords.end()) throw std::runtime_error(where(t) + "coordinate system \"" + name + "\" not available");
                gs.ctm = it->second.first;
                gs.ctm_inv = it->second.second;
                gs.ctm_identity = name == "world";
            } else if (k == "CoordinateSystem") {
                named_coords[read_str(T, i, "CoordinateSystem")] = {gs.ctm, gs.ctm_inv};
            } else if (k == "ReverseOrientation") {
                gs.reverse = !gs.reverse;
            } else if (k == "Material") {
                const std::string type = read_str(T, i, "Material");
                Params P = read_params(T, i);
                F.materials.push_back(make_material(type, P, t));
                gs.material = int(F.materials.size()) - 1;
            } else if (k == "MakeNamedMaterial") {
                const std::string name = read_str(T, i, "MakeNamedMaterial");
                Params P = read_params(T, i);
                Param *ty = P.find("type", "string");
                if (!ty || ty->strs.empty()) throw std::runtime_error(where(t) + "MakeNamedMaterial needs \"string type\"");
                const std::string type = ty->strs[0];
                F.materials.push_back(make_material(type, P, t));
                named_materials[name] = int(F.materials.size()) - 1;
            } else if (k == "NamedMaterial") {
                const std::string name = read_str(T, i, "NamedMaterial");
                auto it = named_materials.find(name);
                if (it == named_materials.end()) throw std::runtime_error(where(t) + "unknown material \"" + name + "\"");
                gs.material = it->second;
            } else if (k == "AreaLightSource") {
                const std::string type = read_str(T, i, "AreaLightSource");
                Params P = read_params(T, i);
                if (type != "diffuse") throw std::runtime_error(where(t) + "area light \"" + type + "\" is not supported");
                pbrtm_area_light L{};
                Param *l = P.find("L", "rgb");
                if (l) {
                    if (l->nums.size() != 3) throw std::runtime_error(where(t) + "rgb L needs 3 values");
                    for (int c = 0; c < 3; ++c) L.rgb[c] = float(l->nums[c]);
                } else {
                    throw std::runtime_error(where(t) + "area lights need \"rgb L\"");
                }
                L.scale = P.get_float("scale", 1.f);
                L.two_sided = P.get_bool("twosided", false) ? 1 : 0;
                if (P.find("power")) throw std::runtime_error(where(t) + "\"power\" is not supported");
                P.check_used("AreaLightSource \"diffuse\"");
                gs.has_light = true;
                gs.light = L;
            } else if (k == "LightSource") {
                // parse_light_source (scene_builder.cu:480-495) -> Light::create: only the
                // uniform "infinite" light (lights/uniform_infinite_light.cu:4-30)
                const std::string type = read_str(T, i, "LightSource");
                Params P = read_params(T, i);
                if (type != "infinite") throw std::runtime_error(where(t) + "LightSource \"" + type + "\" is not supported");
                if (P.find("filename")) throw std::runtime_error(where(t) + "image infinite lights are not supported");
                pbrtm_infinite_light L{};
                if (Param *l = P.find("L", "rgb")) {
                    if (l->nums.size() != 3) throw std::runtime_error(where(t) + "rgb L needs 3 values");
                    for (int c = 0; c < 3; ++c) L.rgb[c] = float(l->nums[c]);
                    L.has_rgb = 1;
                }
                L.scale = P.get_float("scale", 1.f);
                L.illuminance = P.get_float("illuminance", -1.f);
                uint32_t emissive = 0;  // area lights are created per triangle as shapes are parsed
                for (int32_t v : F.tri_light) emissive += v >= 0 ? 1u : 0u;
                L.light_index = emissive;
                if (F.infinite_lights.size() >= 4) throw std::runtime_error(where(t) + "at most 4 infinite lights");
                P.check_used("LightSource \"infinite\"");
                F.infinite_lights.push_back(L);
            } else if (k == "Shape") {
                const std::string type = read_str(T, i, "Shape");
                Params P = read_params(T, i);
                if (!world) throw std::runtime_error(where(t) + "Shape before WorldBegin");
                shape(type, P, t);
            } else if (k == "ObjectBegin") {
                // scene_builder.cu:742-756: the graphics state is pushed, shapes go to the definition
                const std::string name = read_str(T, i, "ObjectBegin");
                if (active_def) throw std::runtime_error(where(t) + "ObjectBegin called inside of instance definition");
                stack.push_back(gs);
                active_name = name;
                definitions[name] = Definition{};
                active_def = &definitions[name];
            } else if (k == "ObjectEnd") {
                if (!active_def) throw std::runtime_error(where(t) + "ObjectEnd called before an instance defined");
                active_def = nullptr;
                if (stack.empty()) throw std::runtime_error(where(t) + "unmatched ObjectEnd");
                gs = stack.back();
                stack.pop_back();
            } else if (k == "ObjectInstance") {
                const std::string name = read_str(T, i, "ObjectInstance");
                if (active_def) throw std::runtime_error(where(t) + "ObjectInstance inside of instance definition");
                instance(name, t);
            } else if (k == "Include" || k == "Import") {
                const std::string f = read_str(T, i, "Include");
                parse_file(f.size() && f[0] == '/' ? f : dir + f);
            } else if (k == "MakeNamedMedium" || k == "MediumInterface") {
                // ignored, as the reference does (scene_builder.cu:451-458)
                if (k == "MakeNamedMedium") read_str(T, i, "MakeNamedMedium");
                else {
                    read_str(T, i, "MediumInterface");
                    if (i < T.size() && T[i].kind == Tok::Str) ++i;
                }
                read_params(T, i);
            } else if (k == "WorldEnd") {
            } else {
                throw std::runtime_error(where(t) + "directive \"" + k + "\" is not supported");
            }
        }
    }
};

}  // namespace

extern "C" {

int pbrtm_pbrt_parse(const char *path, pbrtm_pbrt_file **out, char *err, uint64_t errlen) {
    if (!path || !out) return PBRTM_ERR_INVALID;
    *out = nullptr;
    auto *f = new (std::nothrow) pbrtm_pbrt_file();
    if (!f) return PBRTM_ERR_NOMEM;
    try {
        Builder b(*f);
        b.parse_file(path);
        b.finish();
    } catch (const std::exception &e) {
        if (err && errlen) std::snprintf(err, size_t(errlen), "%s", e.what());
        delete f;
        return PBRTM_ERR_INVALID;
    }
    *out = f;
    return PBRTM_OK;
}

const pbrtm_scene_desc *pbrtm_pbrt_desc(const pbrtm_pbrt_file *f) { return f ? &f->desc : nullptr; }

int pbrtm_pbrt_render_params(const pbrtm_pbrt_file *f, pbrtm_render_params *out) {
    if (!f || !out) return PBRTM_ERR_INVALID;
    *out = pbrtm_render_params{};
    out->spp = f->spp;
    out->max_depth = f->max_depth;
    out->flags = f->integrator;
    return PBRTM_OK;
}

const char *pbrtm_pbrt_output(const pbrtm_pbrt_file *f) { return f ? f->filename.c_str() : ""; }

void pbrtm_pbrt_free(pbrtm_pbrt_file *f) { delete f; }

}  // extern "C"
