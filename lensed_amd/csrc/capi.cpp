// C ABI of the program generator (src/kernel.h) and the object metadata
// loader (src/input/objects.c:15-266).
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/lensed.h"
#include "program.h"
#include "rtc.h"

namespace lsd {
void set_error(const std::string& msg);
std::string device_arch(int device);
}  // namespace lsd

using namespace lsd;

namespace {

char* dup(const std::string& s) {
    char* p = (char*)malloc(s.size() + 1);
    memcpy(p, s.c_str(), s.size() + 1);
    return p;
}

std::string prelude_source() {
    std::string p;
    read_text(kernel_dir() + "/prelude.h", &p);
    return p;
}

// layout of struct param in the device program (kernels/prelude.h)
struct DevParam {
    char name[16];
    int type;
    float bounds[2];
    float defval;
};
static_assert(sizeof(DevParam) == 32, "struct param layout");

struct Meta {
    int type = 0;
    size_t size = 0;  // bytes
    std::vector<DevParam> pars;
};

std::mutex g_meta_mutex;
std::map<std::string, Meta> g_meta_cache;  // keyed by object source

std::string meta_program(const std::string& name, const std::string& src) {
    std::string m = mangle(name);
    return prelude_source() + "\n#pragma clang force_cuda_host_device begin\nnamespace lcl {\n" +
           wrap_object_meta(translate_object(src, m), m, name) +
           "}\n#pragma clang force_cuda_host_device end\n";
}

bool object_meta(const std::string& name, Meta* meta, std::string* err) {
    std::string src, path;
    if (!load_object_source(name, &src, &path)) {
        *err = "could not load object \"" + name + "\"";
        return false;
    }
    std::lock_guard<std::mutex> lock(g_meta_mutex);
    auto it = g_meta_cache.find(src);
    if (it != g_meta_cache.end()) {
        *meta = it->second;
        return true;
    }
    std::vector<char> code;
    std::string log;
    // metadata do not depend on the GPU architecture; no device is needed
    if (!rtc_compile(meta_program(name, src), name + ".cl", device_flags("gfx950"), &code, &log)) {
        *err = "object " + name + ": failed to build program\n" + log;
        return false;
    }
    std::vector<char> b;
    Meta M;
    if (!elf_symbol(code, "lensed_meta_type", &b) || b.size() != 4) {
        *err = "object " + name + ": no type";
        return false;
    }
    memcpy(&M.type, b.data(), 4);
    if (!elf_symbol(code, "lensed_meta_size", &b) || b.size() != 8) {
        *err = "object " + name + ": no data size";
        return false;
    }
    unsigned long long sz;
    memcpy(&sz, b.data(), 8);
    M.size = (size_t)sz;
    if (!elf_symbol(code, "lensed_meta_params", &b) || b.size() % sizeof(DevParam)) {
        *err = "object " + name + ": no parameter list";
        return false;
    }
    M.pars.resize(b.size() / sizeof(DevParam));
    memcpy(M.pars.data(), b.data(), b.size());
    g_meta_cache[src] = M;
    *meta = M;
    return true;
}

}  // namespace

extern "C" {

void object_program(const char* name, size_t* nkernels, const char*** kernels) {
    std::string src;
    std::vector<std::string> parts;
    parts.push_back(prelude_source());
    if (name && load_object_source(name, &src, nullptr))
        parts.push_back("#pragma clang force_cuda_host_device begin\nnamespace lcl {\n" +
                        wrap_object_meta(translate_object(src, mangle(name)), mangle(name), name) +
                        "}\n#pragma clang force_cuda_host_device end\n");
    *nkernels = parts.size();
    *kernels = (const char**)malloc(parts.size() * sizeof(const char*));
    for (size_t i = 0; i < parts.size(); ++i) (*kernels)[i] = dup(parts[i]);
}

static std::vector<ModelObject> model_of(size_t nobjs, const object* objs) {
    std::vector<ModelObject> model;
    for (size_t i = 0; i < nobjs; ++i) {
        ModelObject mo;
        mo.name = objs[i].name;
        mo.type = objs[i].type;
        mo.size = (objs[i].size + 3) / 4 * 4;
        mo.npars = objs[i].npars;
        for (size_t j = 0; j < objs[i].npars; ++j) {
            mo.ptype.push_back(objs[i].pars ? objs[i].pars[j].type : 0);
            mo.ipp.push_back(objs[i].pars ? objs[i].pars[j].ipp != 0 : 0);
        }
        model.push_back(mo);
    }
    return model;
}

void main_program(size_t nobjs, object objs[], size_t* nkernels, const char*** kernels) {
    std::string err;
    std::string all = assemble_program({model_of(nobjs, objs)}, &err);
    *nkernels = all.empty() ? 0 : 1;
    *kernels = (const char**)malloc(sizeof(const char*));
    if (*nkernels) (*kernels)[0] = dup(all);
    if (!all.empty()) return;
    set_error(err);
}

char* lensed_model_source(size_t nobjs, const object* objs) {
    std::string err;
    std::string src = model_source({model_of(nobjs, objs)}, &err);
    if (src.empty()) {
        set_error(err);
        return nullptr;
    }
    return dup(src);
}

char* lensed_model_source_tabulated(size_t nobjs, const object* objs) {
    std::vector<ModelObject> model = model_of(nobjs, objs);
    for (auto& mo : model)
        if (mo.type == OBJ_LENS && object_tabulates(mo.name)) {
            mo.table = true;
            break;
        }
    std::string err;
    std::string src = model_source({model}, &err);
    if (src.empty()) {
        set_error(err);
        return nullptr;
    }
    return dup(src);
}

const char* kernel_options(size_t width, size_t height, int psf, size_t psfw, size_t psfh, size_t nq,
                           const char* flags[]) {
    char buf[256];
    snprintf(buf, sizeof buf,
             " -DIMAGE_SIZE=%zu -DIMAGE_WIDTH=%zu -DIMAGE_HEIGHT=%zu -DPSF=%d -DPSF_WIDTH=%zu -DPSF_HEIGHT=%zu "
             "-DQUAD_POINTS=%zu",
             width * height, width, height, psf ? 1 : 0, psfw, psfh, nq);
    std::string s = buf;
    if (flags)
        for (const char** f = flags; *f; ++f) s += std::string(" ") + *f;
    return dup(s);
}

char* kernel_name(const char* prefix, const char* name) { return dup(std::string(prefix) + name); }

long lensed_program_compile(const char* source, const char* arch, char** log) {
    std::vector<char> code;
    std::string lg;
    bool ok = rtc_compile(source ? source : "", "lensed_program.hip", device_flags(arch ? arch : "gfx950"), &code, &lg);
    if (log) *log = dup(lg);
    return ok ? (long)code.size() : -1;
}

int lensed_object_load(object* obj, const char* id, const char* name) {
    if (!obj || !name) {
        set_error("lensed_object_load: null argument");
        return -1;
    }
    Meta M;
    std::string err;
    if (!object_meta(name, &M, &err)) {
        set_error(err);
        return -1;
    }
    if (M.type != OBJ_LENS && M.type != OBJ_SOURCE && M.type != OBJ_FOREGROUND) {
        set_error(std::string("object ") + name + ": invalid type (should be LENS, SOURCE or FOREGROUND)");
        return -1;
    }
    std::string oid = id ? id : name;
    obj->type = M.type;
    // size in 4-byte units, rounded up (src/input/objects.c:139)
    obj->size = (M.size + 3) / 4;
    obj->id = dup(oid);
    obj->name = dup(name);
    obj->npars = M.pars.size();
    obj->pars = (param*)calloc(M.pars.size() ? M.pars.size() : 1, sizeof(param));
    for (size_t i = 0; i < M.pars.size(); ++i) {
        const DevParam& d = M.pars[i];
        char nm[17];
        memcpy(nm, d.name, 16);
        nm[16] = '\0';
        param& p = obj->pars[i];
        p.name = dup(nm);
        p.id = dup(oid + "." + nm);
        p.type = d.type;
        p.lower = d.bounds[0];
        p.upper = d.bounds[1];
        // a positive or negatively signed default fixes the parameter
        // (src/input/objects.c:225-226)
        p.pri = (d.defval > 0 || std::signbit(d.defval)) ? prior_default(d.defval) : nullptr;
        p.defval = p.pri ? 1 : 0;
    }
    return 0;
}

void lensed_object_free(object* obj) {
    if (!obj) return;
    for (size_t i = 0; i < obj->npars; ++i) {
        free((void*)obj->pars[i].name);
        free((void*)obj->pars[i].id);
        free((void*)obj->pars[i].label);
        prior_free(obj->pars[i].pri);
    }
    free(obj->pars);
    free((void*)obj->id);
    free((void*)obj->name);
    memset(obj, 0, sizeof *obj);
}

}  // extern "C"
