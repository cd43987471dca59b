// Plugin translation and generation of the device program's model code.
//
// The generated compute_<k>() follows the ray-shooting semantics of the code
// Lensed generates (src/kernel.c:65-111, compute_kernel 235-399): objects are
// visited in list order; deflections of consecutive lenses are summed at the
// current ray position and applied when the list moves on to a source, a
// non-finite total deflection sends the ray to 1e10; sources add their
// brightness at the (possibly deflected) ray position; foregrounds always
// use the image-plane position.  set_params_<k>() follows the parameter
// setter (src/kernel.c:113-150, set_params_kernel 401-656), including
// image-plane priors traced back through the lenses of earlier planes.
#include "program.h"

#include <dlfcn.h>
#include <sys/stat.h>

#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <map>
#include <mutex>
#include <regex>
#include <sstream>

namespace lsd {

std::string mangle(const std::string& name) {
    std::string out;
    for (char c : name) {
        bool ok = (c >= 'a' && c <= 'z') || (c >= 'A' && c <= 'Z') || (c >= '0' && c <= '9') || c == '_';
        out += ok ? c : '_';
    }
    if (out.empty() || (out[0] >= '0' && out[0] <= '9')) out = "_" + out;
    return out;
}

static bool is_ident_start(char c) { return (c >= 'a' && c <= 'z') || (c >= 'A' && c <= 'Z') || c == '_'; }
static bool is_ident_char(char c) { return is_ident_start(c) || (c >= '0' && c <= '9'); }

static bool is_vector_type(const std::string& s) {
    static const char* names[] = {"float2", "float3", "float4", "float8", "float16", "mat22",
                                  "int2", "int3", "int4", "uint2", "uint3", "uint4", nullptr};
    for (const char** n = names; *n; ++n)
        if (s == *n) return true;
    return false;
}

std::string translate_object(const std::string& src, const std::string& mangled) {
    (void)mangled;
    std::string out;
    out.reserve(src.size() + 64);
    const size_t n = src.size();
    size_t i = 0;
    auto skip_ws = [&](size_t j) {
        while (j < n && (src[j] == ' ' || src[j] == '\t' || src[j] == '\n' || src[j] == '\r')) ++j;
        return j;
    };
    while (i < n) {
        char c = src[i];
        if (c == '/' && i + 1 < n && src[i + 1] == '/') {  // line comment
            size_t e = src.find('\n', i);
            if (e == std::string::npos) e = n;
            out.append(src, i, e - i);
            i = e;
        } else if (c == '/' && i + 1 < n && src[i + 1] == '*') {  // block comment
            size_t e = src.find("*/", i + 2);
            e = (e == std::string::npos) ? n : e + 2;
            out.append(src, i, e - i);
            i = e;
        } else if (c == '"' || c == '\'') {  // string or character literal
            size_t j = i + 1;
            while (j < n && src[j] != c) j += (src[j] == '\\') ? 2 : 1;
            j = j < n ? j + 1 : n;
            out.append(src, i, j - i);
            i = j;
        } else if (is_ident_start(c) && (i == 0 || !is_ident_char(src[i - 1]))) {
            size_t j = i;
            while (j < n && is_ident_char(src[j])) ++j;
            std::string id = src.substr(i, j - i);
            if (id == "this")
                out += "self";
            else if (id == "local" || id == "global" || id == "private" || id == "__local" ||
                     id == "__global" || id == "__private" || id == "__kernel")
                ;  // address spaces: plain (generic) pointers in HIP C++
            else if (id == "constant" || id == "__constant")
                out += "const";
            else
                out += id;
            i = j;
        } else if (c >= '0' && c <= '9') {  // numeric literal, copied whole
            size_t j = i;
            while (j < n && (is_ident_char(src[j]) || src[j] == '.' ||
                             ((src[j] == '+' || src[j] == '-') && (src[j - 1] == 'e' || src[j - 1] == 'E' ||
                                                                   src[j - 1] == 'p' || src[j - 1] == 'P'))))
                ++j;
            out.append(src, i, j - i);
            i = j;
        } else if (c == '(') {
            // OpenCL vector literal "(T)(a, b, ...)": in C++ that would be a
            // comma expression cast to T, so rewrite it as the constructor
            // call "T(a, b, ...)".  The argument list is left untouched.
            size_t j = skip_ws(i + 1);
            size_t k = j;
            while (k < n && is_ident_char(src[k])) ++k;
            std::string id = src.substr(j, k - j);
            if (!id.empty() && is_vector_type(id)) {
                size_t m = skip_ws(k);
                if (m < n && src[m] == ')') {
                    size_t p = skip_ws(m + 1);
                    if (p < n && src[p] == '(') {
                        out += id;
                        i = p;
                        continue;
                    }
                }
            }
            out += c;
            ++i;
        } else {
            out += c;
            ++i;
        }
    }
    return out;
}

static const char* const kObjectMacros[] = {"type",         "params",    "data",          "deflection",
                                             "brightness",   "foreground", "set",          "frame_centre",
                                             "frame_map",    "frame_profile", "frame_cutoff"};

static std::string undef_macros() {
    std::string s;
    for (const char* m : kObjectMacros) s += std::string("#undef ") + m + "\n";
    return s;
}

// a source may describe itself in its own frame u = frame_map(x - frame_centre)
// with frame_map linear: brightness(x) = frame_profile(u)
static std::string frame_macros(const std::string& m) {
    return "#define frame_centre frame_centre_" + m + "\n#define frame_map frame_map_" + m +
           "\n#define frame_profile frame_profile_" + m + "\n#define frame_cutoff frame_cutoff_" + m + "\n";
}

std::string wrap_object(const std::string& translated, const std::string& m, const std::string& name) {
    std::ostringstream o;
    o << "// ---------------------------------------------------------------- objects/" << name << ".cl\n"
      << "#define type static const int type_" << m << "\n"
      << "#define params static const struct param parlst_" << m << "[] =\n"
      << "#define data struct data_" << m << "\n"
      << "#define deflection deflection_" << m << "\n"
      << "#define brightness brightness_" << m << "\n"
      << "#define foreground foreground_" << m << "\n"
      << "#define set set_" << m << "\n"
      << frame_macros(m)
      << "#line 1 \"objects/" << name << ".cl\"\n"
      << translated << "\n"
      << undef_macros();
    return o.str();
}

std::string wrap_object_meta(const std::string& translated, const std::string& m, const std::string& name) {
    std::ostringstream o;
    o << "#define type extern \"C\" __device__ const int lensed_meta_type\n"
      << "#define params extern \"C\" __device__ const struct param lensed_meta_params[] =\n"
      << "#define data struct data_" << m << "\n"
      << "#define deflection deflection_" << m << "\n"
      << "#define brightness brightness_" << m << "\n"
      << "#define foreground foreground_" << m << "\n"
      << "#define set set_" << m << "\n"
      << frame_macros(m)
      << "#line 1 \"objects/" << name << ".cl\"\n"
      << translated << "\n"
      << undef_macros()
      << "extern \"C\" __device__ const unsigned long long lensed_meta_size = sizeof(struct data_" << m << ");\n"
      << "extern \"C\" __device__ const unsigned long long lensed_meta_align = alignof(struct data_" << m << ");\n"
      << "extern \"C\" __device__ const unsigned long long lensed_meta_npar = "
         "sizeof(lensed_meta_params)/sizeof(struct param);\n";
    return o.str();
}

bool read_text(const std::string& path, std::string* out) {
    std::ifstream f(path, std::ios::binary);
    if (!f) return false;
    std::stringstream ss;
    ss << f.rdbuf();
    *out = ss.str();
    return true;
}

static bool file_exists(const std::string& p) {
    struct stat st;
    return stat(p.c_str(), &st) == 0 && S_ISREG(st.st_mode);
}

// The package directory: the folder holding this shared library.
static std::string package_dir() {
    Dl_info info;
    if (dladdr((void*)&package_dir, &info) && info.dli_fname) {
        std::string p = info.dli_fname;
        size_t s = p.rfind('/');
        if (s != std::string::npos) return p.substr(0, s);
    }
    return ".";
}

std::string kernel_dir() {
    const char* env = getenv("LENSED_KERNEL_PATH");
    if (env && *env) return env;
    return package_dir() + "/kernels";
}

bool load_object_source(const std::string& name, std::string* source, std::string* path) {
    std::vector<std::string> dirs;
    if (const char* env = getenv("LENSED_OBJECTS_PATH")) {
        std::string s = env;
        size_t b = 0;
        while (b <= s.size()) {
            size_t e = s.find(':', b);
            if (e == std::string::npos) e = s.size();
            if (e > b) dirs.push_back(s.substr(b, e - b));
            b = e + 1;
        }
    }
    dirs.push_back(package_dir() + "/objects");
    for (const auto& d : dirs) {
        std::string p = d + "/" + name + ".cl";
        if (file_exists(p) && read_text(p, source)) {
            if (path) *path = p;
            return true;
        }
    }
    return false;
}

size_t model_data_words(const std::vector<ModelObject>& objs) {
    size_t d = 0;
    for (const auto& o : objs) d += o.size;
    return d;
}

static std::string data_ptr(const std::string& m, size_t off, const char* base) {
    std::ostringstream o;
    o << "(struct data_" << m << "*)(" << base << " + " << off << ")";
    return o.str();
}

// the plugin's source text, empty if there is none (read on every call: the
// object search path may change between contexts)
static std::string plugin_text(const std::string& name) {
    std::string src;
    if (!load_object_source(name, &src, nullptr)) src.clear();
    return src;
}

// whether a plugin defines the functions named by the patterns
static bool plugin_defines(const std::string& name, const std::vector<const char*>& patterns) {
    const std::string src = plugin_text(name);
    bool ok = !src.empty();
    for (const char* re : patterns) ok = ok && std::regex_search(src, std::regex(re));
    return ok;
}

// whether the plugin's data block declares every field of the patterns
static bool data_declares(const std::string& name, const std::vector<const char*>& fields) {
    // comments removed first: they may hold braces
    const std::string src =
        std::regex_replace(std::regex_replace(plugin_text(name), std::regex(R"(/\*[\s\S]*?\*/)"), " "),
                           std::regex(R"(//[^\n]*)"), " ");
    std::smatch m;
    if (!std::regex_search(src, m, std::regex(R"(\bdata\s*\{([^}]*)\})"))) return false;
    const std::string block = m[1].str();
    for (const char* re : fields)
        if (!std::regex_search(block, std::regex(re))) return false;
    return true;
}

bool object_tabulates(const std::string& name) {
    return data_declares(name, {R"(\bfloat2\s+c\s*;)", R"(\bmat22\s+fwd\s*;)", R"(\bmat22\s+bwd\s*;)",
                                R"(\bfloat\s+coef\s*\[\s*10\s*\]\s*;)", R"(\bfloat\s+hexp\s*;)"});
}

bool object_table_shear(const std::string& name) {
    return object_tabulates(name) && data_declares(name, {R"(\bmat22\s+shear\s*;)"});
}

static bool source_has_frame(const std::string& name) {
    return plugin_defines(name, {R"(static\s+float2\s+frame_centre\s*\()", R"(static\s+float2\s+frame_map\s*\()",
                                 R"(static\s+float\s+frame_profile\s*\()"});
}

static bool source_has_cutoff(const std::string& name) {
    return plugin_defines(name, {R"(static\s+float\s+frame_cutoff\s*\()"});
}

// The lens-frame form of a tabulated EPL lens applies when that lens is the
// only one and a single source behind it has a frame (source_has_frame):
// its index, or -1.
static int frame_source(const std::vector<ModelObject>& objs) {
    int nlens = 0, table = -1, nbehind = 0, src = -1;
    for (size_t i = 0; i < objs.size(); ++i) {
        if (objs[i].type == OBJ_LENS) {
            ++nlens;
            if (objs[i].table) table = (int)i;
        } else if (objs[i].type == OBJ_SOURCE && nlens > 0) {
            ++nbehind;
            src = (int)i;
        }
    }
    if (nlens != 1 || table < 0 || nbehind != 1 || !source_has_frame(objs[src].name)) return -1;
    return src;
}

bool model_culls(const std::vector<ModelObject>& objs) {
    const int fsrc = frame_source(objs);
    return fsrc >= 0 && objs.size() == 2 && source_has_cutoff(objs[fsrc].name);
}

std::string compute_function(const std::vector<ModelObject>& objs, int k) {
    std::ostringstream o;
    const int fsrc = frame_source(objs);
    size_t lens_off = 0, src_off = 0;
    std::string lm, sm, pre_fn = "lensed_epl_pre";
    {
        size_t dd = 0;
        for (size_t i = 0; i < objs.size(); ++i) {
            if (objs[i].table && lm.empty()) {
                lens_off = dd;
                lm = mangle(objs[i].name);
                if (object_table_shear(objs[i].name)) pre_fn += "_shear";
            }
            if ((int)i == fsrc) src_off = dd, sm = mangle(objs[i].name);
            dd += objs[i].size;
        }
    }
    // a tabulated EPL lens (at most one per model) starts from the pair
    // (F (x - c), x - S (x - c)) that the render kernel forms incrementally
    // (pixel part at d = p, node part at d = q_n); in the lens-frame form
    // the second half is taken on into the source's frame, T (x - S (x - c) - c_s)
    o << "static inline float4 epl_pre_" << k << "(uint* data, float2 d, bool pixel)\n{\n"
      << "    (void)data; (void)d; (void)pixel;\n";
    if (!lm.empty()) {
        o << "    const struct data_" << lm << "* s = " << data_ptr(lm, lens_off, "data") << ";\n"
          << "    const float4 e = " << pre_fn << "(s, pixel ? d - s->c : d, d);\n";
        if (fsrc >= 0)
            o << "    struct data_" << sm << "* g = " << data_ptr(sm, src_off, "data") << ";\n"
              << "    const float2 X = float2(e.z, e.w);\n"
              << "    const float2 P = frame_map_" << sm << "(g, pixel ? X - frame_centre_" << sm << "(g) : X);\n"
              << "    return float4(e.x, e.y, P.x, P.y);\n";
        else
            o << "    return e;\n";
    } else {
        o << "    return float4(0.0f, 0.0f, 0.0f, 0.0f);\n";
    }
    o << "}\n";
    o << "static inline float compute_" << k
      << "(uint* data, float2 x, const float4* tab, float4 pre = float4(0.0f, 0.0f, 0.0f, 0.0f))\n{\n"
      << "    (void)tab; (void)pre;\n"
      << "    float2 y = x;\n"
      << "    float f = 0;\n";
    bool plane_open = false, lin = false;
    int trigger = 0;
    size_t d = 0;
    // with a tabulated EPL lens, pre.zw is the ray position with the lens's
    // linear (shear) deflection already taken off, x - S (x - c); the rest
    // of the deflection is subtracted from it.  The guard then sees the
    // deflection without the shear, which is finite wherever the shear is
    // (finite parameters), so the rays it rejects are the same.
    auto close_plane = [&]() {
        o << "        y = " << (lin ? "float2(pre.z, pre.w)" : "y")
          << " - (lensed_deflection_ok(a) ? a : float2(1E10f, 1E10f));\n    }\n";
        plane_open = false;
    };
    for (size_t i = 0; i < objs.size(); ++i) {
        const ModelObject& ob = objs[i];
        std::string m = mangle(ob.name);
        if (ob.type != trigger && ob.type != OBJ_FOREGROUND) {
            if (trigger == OBJ_LENS && plane_open) close_plane();
            trigger = ob.type;
        }
        if (fsrc >= 0 && ob.table) {
            // lens-frame form: the deflection stays in the lens frame, a = bwd v,
            // and the guard clamps v (lensed_clamp_deflection)
            o << "    const struct data_" << m << "* s = " << data_ptr(m, d, "data") << ";\n"
              << "    const float2 v = lensed_clamp_deflection(lensed_epl_table_vector_pre(s, pre, tab));\n";
        } else if ((int)i == fsrc) {
            o << "    {\n        struct data_" << m << "* g = " << data_ptr(m, d, "data") << ";\n"
              << "        const float2 n0 = frame_map_" << m << "(g, float2(-s->bwd.x, -s->bwd.z));\n"
              << "        const float2 n1 = frame_map_" << m << "(g, float2(-s->bwd.y, -s->bwd.w));\n"
              << "        f += frame_profile_" << m << "(g, lensed_frame_trace(pre, n0, n1, v));\n    }\n";
        } else if (ob.type == OBJ_LENS) {
            if (!plane_open) {
                o << "    {\n        float2 a = 0;\n";
                plane_open = true;
            }
            if (ob.table) {
                o << "        a += lensed_epl_table_deflection_pre(" << data_ptr(m, d, "data") << ", pre, tab);\n";
                lin = true;
            } else
                o << "        a += deflection_" << m << "(" << data_ptr(m, d, "data") << ", y);\n";
        } else if (ob.type == OBJ_SOURCE) {
            o << "    f += brightness_" << m << "(" << data_ptr(m, d, "data") << ", y);\n";
        } else if (ob.type == OBJ_FOREGROUND) {
            o << "    f += foreground_" << m << "(" << data_ptr(m, d, "data") << ", x);\n";
        }
        d += ob.size;
    }
    if (plane_open) close_plane();
    o << "    return f;\n}\n";
    // pixel culling (lensed_frame_cull): only when the model is the lens and
    // the framed source alone, so a skipped pixel's value is exactly 0; the
    // candidate's constants are formed once (cull_consts_k, in set_params)
    const bool culls = model_culls(objs);
    o << "static inline void cull_consts_" << k << "(uint* data, float4* cc)\n{\n"
      << "    (void)data; (void)cc;\n";
    if (culls) {
        const bool shear = pre_fn == "lensed_epl_pre_shear";
        o << "    const struct data_" << lm << "* s = " << data_ptr(lm, lens_off, "data") << ";\n"
          << "    struct data_" << sm << "* g = " << data_ptr(sm, src_off, "data") << ";\n"
          << "    const float2 n0 = frame_map_" << sm << "(g, float2(-s->bwd.x, -s->bwd.z));\n"
          << "    const float2 n1 = frame_map_" << sm << "(g, float2(-s->bwd.y, -s->bwd.w));\n"
          << "    const float2 t0 = frame_map_" << sm << "(g, float2(1.0f, 0.0f));\n"
          << "    const float2 t1 = frame_map_" << sm << "(g, float2(0.0f, 1.0f));\n"
          << "    const float lnorm = "
          << (shear ? "lensed_frob(float4(1.0f - s->shear.x, -s->shear.y, -s->shear.z, 1.0f - s->shear.w))"
                    : "1.41421356f")
          << ";\n"
          << "    lensed_frame_cull_consts(s, n0, n1, sqrtf(dot(t0, t0) + dot(t1, t1)), lnorm, frame_cutoff_" << sm
          << "(g), cc);\n";
    }
    o << "}\n";
    o << "static inline bool cull_" << k << "(uint* data, float4 base, const float4* tab, const float4* cc, float rho)\n{\n"
      << "    (void)data; (void)base; (void)tab; (void)cc; (void)rho;\n";
    if (culls)
        o << "    return lensed_frame_cull(" << data_ptr(lm, lens_off, "data") << ", base, tab, cc, rho);\n";
    else
        o << "    return false;\n";
    o << "}\n";
    return o.str();
}

std::string set_params_function(const std::vector<ModelObject>& objs, int k) {
    std::ostringstream o;
    o << "static inline void set_params_" << k << "(uint* data, const float* params)\n{\n"
      << "    float2 x = 0;\n"
      << "    float2 a = 0;\n"
      << "    (void)x; (void)a;\n";
    int trigger = 0;
    size_t plane = 0;
    size_t d = 0, p = 0;
    for (size_t i = 0; i < objs.size(); ++i) {
        const auto& ob = objs[i];
        std::string m = mangle(ob.name);
        if (ob.type != trigger && ob.type != OBJ_FOREGROUND) {
            if (trigger == OBJ_LENS && ob.type == OBJ_SOURCE) plane = i;
            trigger = ob.type;
        }
        // image-plane priors: trace the observed position back through the
        // lenses of all earlier planes
        for (size_t j = 0; j < ob.npars; ++j) {
            if (!ob.ipp[j] || ob.ptype[j] != PAR_POSITION_X) continue;
            o << "    x = float2(params[" << p + j << "], params[" << p + j + 1 << "]);\n";
            int trig2 = 0;
            size_t d2 = 0;
            for (size_t q = 0; q < plane; ++q) {
                if (objs[q].type != trig2 && objs[q].type != OBJ_FOREGROUND) {
                    if (trig2 == OBJ_LENS)
                        o << "    x -= dot(a, a) < HUGE_VALF ? a : float2(1E10f, 1E10f);\n    a = 0;\n";
                    trig2 = objs[q].type;
                }
                if (objs[q].type == OBJ_LENS) {
                    std::string mq = mangle(objs[q].name);
                    o << "    a += deflection_" << mq << "(" << data_ptr(mq, d2, "data") << ", x);\n";
                }
                d2 += objs[q].size;
            }
            if (trig2 == OBJ_LENS)
                o << "    x -= dot(a, a) < HUGE_VALF ? a : float2(1E10f, 1E10f);\n    a = 0;\n";
        }
        o << "    set_" << m << "(" << data_ptr(m, d, "data");
        for (size_t j = 0; j < ob.npars; ++j) {
            if (ob.ipp[j] && ob.ptype[j] == PAR_POSITION_X)
                o << ", x.x";
            else if (ob.ipp[j] && ob.ptype[j] == PAR_POSITION_Y)
                o << ", x.y";
            else if (ob.ipp[j])
                o << ", 0";
            else
                o << ", params[" << p + j << "]";
        }
        o << ");\n";
        d += ob.size;
        p += ob.npars;
    }
    o << "}\n";
    return o.str();
}

}  // namespace lsd
