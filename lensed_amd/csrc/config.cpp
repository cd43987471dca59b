// Configuration: INI files and command-line options, as Lensed reads them.
//
// Groups [options] (the default), [objects], [priors], [labels]
// (src/input/ini.c:109-284); the option table with its types, defaults and
// conditional requirements (src/input/options.c:63-228); `--name=value`
// arguments (src/input.c:83-100); objects loaded from their plugins as
// they are named, priors with the `wrap` and `image` keywords
// (src/input/objects.c:328-359), labels; and the final checks of
// src/input.c:214-239 (required options, some objects, a prior for every
// parameter).  Relative paths in a file are taken from the file's directory.
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/lensed.h"

namespace lsd {
void set_error(const std::string& msg);
}

namespace {

enum Kind { STRING, BOOL, INT, REAL, PATH, PATH_OR_REAL };
enum Req { OPTIONAL, REQUIRED, REQ_IF_SET, REQ_IF_NOT };

struct OptionDef {
    const char* name;
    const char* help;
    Kind kind;
    Req req;
    const char* depend;  // for REQ_IF_SET / REQ_IF_NOT
    const char* defval;  // nullptr: none
};

// the option table of src/input/options.c:63-228, in its order (which
// decides abbreviations: a prefix selects the first option it matches)
const OptionDef kOptions[] = {
    {"device", "Select computation device", STRING, OPTIONAL, nullptr, "auto"},
    {"output", "Output results", BOOL, OPTIONAL, nullptr, "true"},
    {"root", "Root element for all output paths", STRING, REQ_IF_SET, "output", nullptr},
    {"image", "Input image, FITS file in counts/sec", PATH, REQUIRED, nullptr, nullptr},
    {"gain", "Conversion factor to counts", PATH_OR_REAL, REQ_IF_NOT, "weight", nullptr},
    {"offset", "Subtracted flat-field offset", REAL, OPTIONAL, nullptr, "0"},
    {"bscale", "Scale input pixel values", REAL, OPTIONAL, nullptr, "0"},
    {"weight", "Weight map in 1/(counts/sec)^2", PATH_OR_REAL, OPTIONAL, nullptr, nullptr},
    {"xweight", "Extra weight map multiplier", PATH_OR_REAL, OPTIONAL, nullptr, nullptr},
    {"mask", "Input mask, FITS file", PATH, OPTIONAL, nullptr, nullptr},
    {"psf", "Point-spread function, FITS file", PATH, OPTIONAL, nullptr, nullptr},
    {"rule", "Rule for numerical integration", STRING, OPTIONAL, nullptr, "g3k7"},
    {"nlive", "Number of live points", INT, OPTIONAL, nullptr, "300"},
    {"ins", "Use importance nested sampling", BOOL, OPTIONAL, nullptr, "true"},
    {"mmodal", "Mode separation (if ins = false)", BOOL, OPTIONAL, nullptr, "true"},
    {"ceff", "Constant efficiency mode", BOOL, OPTIONAL, nullptr, "true"},
    {"acc", "Target acceptance rate", REAL, OPTIONAL, nullptr, "0.05"},
    {"tol", "Tolerance in log-evidence", REAL, OPTIONAL, nullptr, "0.1"},
    {"shf", "Shrinking factor", REAL, OPTIONAL, nullptr, "0.8"},
    {"maxmodes", "Maximum number of expected modes", INT, OPTIONAL, nullptr, "100"},
    {"feedback", "Show feedback from MultiNest", BOOL, OPTIONAL, nullptr, "false"},
    {"updint", "Update interval for output", INT, OPTIONAL, nullptr, "100"},
    {"seed", "Random number seed for sampling", INT, OPTIONAL, nullptr, "-1"},
    {"resume", "Resume from last checkpoint", BOOL, OPTIONAL, nullptr, "false"},
    {"maxiter", "Maximum number of iterations", INT, OPTIONAL, nullptr, "0"},
};
constexpr int kNumOptions = sizeof(kOptions) / sizeof(kOptions[0]);

int find_option(const std::string& name) {
    if (name.empty()) return -1;
    for (int i = 0; i < kNumOptions; ++i)
        if (std::strncmp(name.c_str(), kOptions[i].name, name.size()) == 0) return i;
    return -1;
}

bool parse_bool(const std::string& v, bool& out) {
    if (v == "true" || v == "TRUE" || v == "1") out = true;
    else if (v == "false" || v == "FALSE" || v == "0") out = false;
    else return false;
    return true;
}

bool parse_int(const std::string& v, long& out) {
    if (v.empty()) return false;
    char* end;
    out = std::strtol(v.c_str(), &end, 10);
    return !*end && out >= -2147483648L && out <= 2147483647L;
}

bool parse_real(const std::string& v, double& out) {
    if (v.empty()) return false;
    char* end;
    out = std::strtod(v.c_str(), &end);
    return !*end;
}

const char* kWS = " \t\n\v\f\r";

std::string trim(const std::string& s) {
    size_t a = s.find_first_not_of(kWS), b = s.find_last_not_of(kWS);
    return a == std::string::npos ? "" : s.substr(a, b - a + 1);
}

const char* kind_name(Kind k) {
    switch (k) {
        case STRING: return "string";
        case BOOL: return "bool";
        case INT: return "int";
        case REAL: return "real";
        case PATH: return "path";
        case PATH_OR_REAL: return "path_or_real";
    }
    return "?";
}

}  // namespace

struct lensed_config {
    std::string value[kNumOptions];
    bool has[kNumOptions] = {};   // a value (given or default)
    bool given[kNumOptions] = {};
    std::vector<object> objs;
    int last_type = 0, nplanes = 0;
    std::string cwd;                  // directory of the file being read

    lensed_config() {
        for (int i = 0; i < kNumOptions; ++i)
            if (kOptions[i].defval) {
                value[i] = kOptions[i].defval;
                has[i] = true;
            }
    }
    ~lensed_config() {
        for (auto& o : objs) lensed_object_free(&o);
    }

    object* find_object(const std::string& id) {
        for (auto& o : objs)
            if (o.id && id == o.id) return &o;
        return nullptr;
    }

    // src/input/options.c:274-326
    bool set_option(const std::string& name, std::string v, std::string& err) {
        int k = find_option(name);
        if (k < 0) {
            err = "invalid option: " + name;
            return false;
        }
        const OptionDef& d = kOptions[k];
        if (v.empty()) {
            if (d.kind != BOOL) {
                err = "option " + name + ": no value given";
                return false;
            }
            v = "true";
        }
        bool ok = true, b;
        long iv;
        double rv;
        switch (d.kind) {
            case BOOL: ok = parse_bool(v, b); if (ok) v = b ? "true" : "false"; break;
            case INT: ok = parse_int(v, iv); break;
            case REAL: ok = parse_real(v, rv); break;
            case PATH: v = absolute(v); break;
            case PATH_OR_REAL: if (!parse_real(v, rv)) v = absolute(v); break;
            case STRING: break;
        }
        if (!ok) {
            err = "option " + name + ": invalid value: " + v + " (should be " + kind_name(d.kind) + ")";
            return false;
        }
        value[k] = v;
        has[k] = given[k] = true;
        return true;
    }

    std::string absolute(const std::string& p) const {
        if (p.empty() || p[0] == '/' || cwd.empty()) return p;
        return cwd + "/" + p;
    }

    bool add_object(const std::string& id, const std::string& name, std::string& err) {
        if (find_object(id)) {
            err = "duplicate object name: " + id;
            return false;
        }
        object o;
        std::memset(&o, 0, sizeof o);
        if (lensed_object_load(&o, id.c_str(), name.c_str())) {
            err = std::string("object ") + id + ": " + lensed_last_error();
            return false;
        }
        // one lens plane only (src/input/ini.c:249-260)
        if (o.type != last_type && o.type != OBJ_FOREGROUND) {
            if (o.type == OBJ_LENS && ++nplanes > 1) {
                lensed_object_free(&o);
                err = "multiple lensing planes are not supported";
                return false;
            }
            last_type = o.type;
        }
        objs.push_back(o);
        return true;
    }

    param* find_param(const std::string& full, std::string& err) {
        size_t dot = full.find('.');
        if (dot == std::string::npos) {
            err = "invalid parameter name (should be <object>.<param>)";
            return nullptr;
        }
        std::string oid = trim(full.substr(0, dot)), pname = trim(full.substr(dot + 1));
        if (pname.empty()) {
            err = "object " + oid + ": no parameter given (should be " + oid + ".<param>)";
            return nullptr;
        }
        object* o = find_object(oid);
        if (!o) {
            err = "unknown object: " + oid + " (check [objects] group)";
            return nullptr;
        }
        for (size_t j = 0; j < o->npars; ++j)
            if (o->pars[j].name && pname == o->pars[j].name) return &o->pars[j];
        err = "object " + oid + ": unknown parameter " + pname;
        return nullptr;
    }

    // src/input/objects.c:328-359
    bool set_prior(param* par, const std::string& text, std::string& err) {
        const char* s = text.c_str();
        for (;;) {
            size_t len = std::strcspn(s, kWS);
            if (len && std::strncmp(s, "wrap", len) == 0 && len <= 4) par->wrap = 1;
            else if (len && std::strncmp(s, "image", len) == 0 && len <= 5) par->ipp = 1;
            else break;
            s += len;
            s += std::strspn(s, kWS);
        }
        if (par->pri) prior_free(par->pri);
        par->pri = nullptr;
        if (*s) {
            par->pri = prior_read(s);
            if (!par->pri) {
                err = std::string("invalid prior: ") + s;
                return false;
            }
        }
        par->defval = 0;
        return true;
    }
};

extern "C" {

lensed_config* lensed_config_create(void) { return new lensed_config(); }

void lensed_config_free(lensed_config* c) { delete c; }

int lensed_config_set(lensed_config* c, const char* name, const char* value) {
    if (!c || !name) {
        lsd::set_error("lensed_config_set: null argument");
        return -1;
    }
    std::string err;
    c->cwd.clear();  // command-line paths are relative to the working directory
    if (!c->set_option(name, value ? value : "", err)) {
        lsd::set_error(err);
        return -1;
    }
    return 0;
}

int lensed_config_arg(lensed_config* c, const char* arg) {
    if (!c || !arg) {
        lsd::set_error("lensed_config_arg: null argument");
        return -1;
    }
    std::string a = arg;
    if (a.rfind("--", 0) == 0) a = a.substr(2);
    size_t eq = a.find('=');
    return lensed_config_set(c, a.substr(0, eq).c_str(), eq == std::string::npos ? "" : a.substr(eq + 1).c_str());
}

int lensed_config_read_ini(lensed_config* c, const char* path) {
    if (!c || !path) {
        lsd::set_error("lensed_config_read_ini: null argument");
        return -1;
    }
    FILE* f = std::fopen(path, "r");
    if (!f) {
        lsd::set_error(std::string(path) + ": could not open file");
        return -1;
    }
    std::string p = path;
    size_t slash = p.rfind('/');
    c->cwd = slash == std::string::npos ? "" : p.substr(0, slash);
    enum { OPTIONS, OBJECTS, PRIORS, LABELS } grp = OPTIONS;
    char buf[1024];
    size_t line = 0;
    std::string err;
    auto fail = [&](const std::string& msg) {
        std::fclose(f);
        c->cwd.clear();
        lsd::set_error(std::string(path) + ":" + std::to_string(line) + ": " + msg);
        return -1;
    };
    while (std::fgets(buf, sizeof buf, f)) {
        line += 1;
        size_t len = std::strcspn(buf, ";#\n");
        if (buf[len] == '\0' && !std::feof(f)) return fail("line too long (max. 1024 characters)");
        buf[len] = '\0';
        std::string s = trim(buf);
        if (s.empty()) continue;
        if (s[0] == '[') {
            if (s.back() != ']') return fail("missing closing ']' character for group");
            std::string g = trim(s.substr(1, s.size() - 2));
            if (g == "options") grp = OPTIONS;
            else if (g == "objects") grp = OBJECTS;
            else if (g == "priors") grp = PRIORS;
            else if (g == "labels") grp = LABELS;
            else return fail("unknown group: " + g);
            continue;
        }
        size_t eq = s.find_first_of("=:");
        if (eq == std::string::npos) return fail("line does not assign anything to \"" + s + "\"");
        std::string name = trim(s.substr(0, eq)), value = trim(s.substr(eq + 1));
        param* par = nullptr;
        if (grp == PRIORS || grp == LABELS) {
            par = c->find_param(name, err);
            if (!par) return fail(err);
        }
        switch (grp) {
            case OPTIONS:
                if (!c->set_option(name, value, err)) return fail(err);
                break;
            case OBJECTS:
                if (!c->add_object(name, value, err)) return fail(err);
                break;
            case PRIORS:
                if (!c->set_prior(par, value, err)) return fail(err);
                break;
            case LABELS: {  // owned by the parameter, freed with its object
                char* lab = (char*)std::malloc(value.size() + 1);
                std::memcpy(lab, value.c_str(), value.size() + 1);
                std::free((void*)par->label);
                par->label = lab;
                break;
            }
        }
    }
    bool bad = std::ferror(f);
    std::fclose(f);
    c->cwd.clear();
    if (bad) {
        lsd::set_error(std::string(path) + ": read error");
        return -1;
    }
    return 0;
}

int lensed_config_check(const lensed_config* c) {
    if (!c) {
        lsd::set_error("invalid configuration");
        return -1;
    }
    for (int i = 0; i < kNumOptions; ++i) {
        const OptionDef& d = kOptions[i];
        bool need = false;
        if (d.req == REQUIRED) need = true;
        else if (d.req == REQ_IF_SET || d.req == REQ_IF_NOT) {
            int k = find_option(d.depend);
            // a flag is set when true; any other option when it has a value
            bool set = kOptions[k].kind == BOOL ? c->value[k] == "true" : c->has[k];
            need = d.req == REQ_IF_SET ? set : !set;
        }
        if (need && !c->given[i]) {
            lsd::set_error(std::string("missing required option: ") + d.name);
            return -1;
        }
    }
    if (c->objs.empty()) {
        lsd::set_error("no objects were given (check [objects] section)");
        return -1;
    }
    for (const auto& o : c->objs)
        for (size_t j = 0; j < o.npars; ++j)
            if (!o.pars[j].pri) {
                lsd::set_error(std::string("missing prior: ") + (o.pars[j].id ? o.pars[j].id : "?") +
                               " (check [priors] section)");
                return -1;
            }
    return 0;
}

const char* lensed_config_get(const lensed_config* c, const char* name) {
    if (!c || !name) return nullptr;
    for (int i = 0; i < kNumOptions; ++i)
        if (std::strcmp(name, kOptions[i].name) == 0) return c->has[i] ? c->value[i].c_str() : nullptr;
    return nullptr;
}

int lensed_config_option(int i, const char** name, const char** type, const char** help, const char** defval) {
    if (i < 0 || i >= kNumOptions) return -1;
    if (name) *name = kOptions[i].name;
    if (type) *type = kind_name(kOptions[i].kind);
    if (help) *help = kOptions[i].help;
    if (defval) *defval = kOptions[i].defval;
    return 0;
}

size_t lensed_config_nobjects(const lensed_config* c) { return c ? c->objs.size() : 0; }

object* lensed_config_objects(lensed_config* c) { return c && !c->objs.empty() ? c->objs.data() : nullptr; }

int lensed_config_data(const lensed_config* c, lensed_data* d) {
    if (!c) {
        lsd::set_error("invalid configuration");
        return -1;
    }
    lensed_data_options o;
    std::memset(&o, 0, sizeof o);
    auto get = [&](const char* n) { return lensed_config_get(c, n); };
    o.image = get("image");
    o.weight = get("weight");
    o.gain = get("gain");
    o.offset = std::atof(get("offset"));
    o.bscale = std::atof(get("bscale"));
    o.xweight = get("xweight");
    o.mask = get("mask");
    o.psf = get("psf");
    return lensed_data_read(&o, d);
}

}  // extern "C"
