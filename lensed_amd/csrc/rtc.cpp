// hipRTC compilation of generated device programs and symbol lookup in the
// resulting code objects (ELF), used for the objects' metadata.
#include "rtc.h"

#include <hip/hiprtc.h>

#include <dlfcn.h>
#include <link.h>
#include <sys/stat.h>

#include <algorithm>
#include <cstdint>
#include <cstdlib>
#include <cstring>

namespace lsd {

std::vector<std::string> device_flags(const std::string& arch) {
    // Counterpart of "-cl-denorms-are-zero -cl-fast-relaxed-math"
    // (src/lensed.c:744-748) minus finite-math-only: the ray-shooting code
    // relies on Inf/NaN comparisons for its non-finite deflection guard, and
    // the chi^2 accumulation is compensated, so no reassociation either.
    return {"--offload-arch=" + arch,
            "-O3",
            "-std=c++17",
            "-fgpu-flush-denormals-to-zero",
            "-fno-math-errno",
            "-fno-trapping-math",
            "-ffp-contract=fast",
            "-fapprox-func",
            "-freciprocal-math",
            "-fno-signed-zeros",
            "-fno-strict-aliasing",
            // packed FP32 instructions are kept out of the generated code
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"};
}

bool rtc_compile(const std::string& source, const std::string& name,
                 const std::vector<std::string>& options, std::vector<char>* code,
                 std::string* log) {
    hiprtcProgram prog;
    if (hiprtcCreateProgram(&prog, source.c_str(), name.c_str(), 0, nullptr, nullptr) != HIPRTC_SUCCESS) {
        if (log) *log = "hiprtcCreateProgram failed";
        return false;
    }
    std::vector<const char*> opts;
    for (const auto& o : options) opts.push_back(o.c_str());
    hiprtcResult r = hiprtcCompileProgram(prog, (int)opts.size(), opts.data());
    size_t ls = 0;
    hiprtcGetProgramLogSize(prog, &ls);
    if (log) {
        log->assign(ls, '\0');
        if (ls) hiprtcGetProgramLog(prog, &(*log)[0]);
        while (!log->empty() && log->back() == '\0') log->pop_back();
    }
    bool ok = r == HIPRTC_SUCCESS;
    if (ok && code) {
        size_t cs = 0;
        hiprtcGetCodeSize(prog, &cs);
        code->resize(cs);
        hiprtcGetCode(prog, code->data());
    }
    hiprtcDestroyProgram(&prog);
    return ok;
}

namespace {
std::string file_identity(const char* path) {
    std::string s = path;
    struct stat st;
    if (stat(path, &st) == 0)
        s += ":" + std::to_string((long long)st.st_size) + ":" + std::to_string((long long)st.st_mtime);
    return s;
}

std::vector<std::string> loaded_comgr() {
    std::vector<std::string> found;
    dl_iterate_phdr(
        [](dl_phdr_info* info, size_t, void* out) {
            if (info->dlpi_name && strstr(info->dlpi_name, "amd_comgr"))
                static_cast<std::vector<std::string>*>(out)->push_back(info->dlpi_name);
            return 0;
        },
        &found);
    std::sort(found.begin(), found.end());
    return found;
}
}  // namespace

std::string compiler_identity() {
    // hipRTC loads its compiler library at its first build: a tiny one
    // settles which copy that is before anything is looked up by it
    if (loaded_comgr().empty())
        rtc_compile("extern \"C\" __global__ void lensed_probe() {}", "probe.hip", {"-O0"}, nullptr, nullptr);
    std::string id;
    Dl_info info;
    if (dladdr(reinterpret_cast<void*>(&hiprtcCompileProgram), &info) && info.dli_fname)
        id += "hiprtc " + file_identity(info.dli_fname);
    for (const auto& c : loaded_comgr()) id += " comgr " + file_identity(c.c_str());
    return id;
}

namespace {
template <class T>
T rd(const std::vector<char>& b, size_t off) {
    T v;
    memcpy(&v, b.data() + off, sizeof(T));
    return v;
}
}  // namespace

bool elf_symbol(const std::vector<char>& elf, const char* sym, std::vector<char>* bytes) {
    if (elf.size() < 64 || memcmp(elf.data(), "\x7f" "ELF", 4) != 0 || elf[4] != 2) return false;
    const uint64_t shoff = rd<uint64_t>(elf, 0x28);
    const uint16_t shentsize = rd<uint16_t>(elf, 0x3A);
    const uint16_t shnum = rd<uint16_t>(elf, 0x3C);
    if (shoff + (uint64_t)shnum * shentsize > elf.size()) return false;
    struct Sec { uint32_t type; uint64_t addr, off, size; uint32_t link; uint64_t entsize; };
    std::vector<Sec> secs(shnum);
    for (uint16_t i = 0; i < shnum; ++i) {
        size_t b = shoff + (size_t)i * shentsize;
        secs[i] = {rd<uint32_t>(elf, b + 4), rd<uint64_t>(elf, b + 16), rd<uint64_t>(elf, b + 24),
                   rd<uint64_t>(elf, b + 32), rd<uint32_t>(elf, b + 40), rd<uint64_t>(elf, b + 56)};
    }
    for (const Sec& st : secs) {
        if (st.type != 2 && st.type != 11) continue;  // SHT_SYMTAB, SHT_DYNSYM
        if (st.link >= shnum || st.entsize < 24) continue;
        const Sec& str = secs[st.link];
        for (uint64_t o = 0; o + 24 <= st.size; o += st.entsize) {
            size_t b = st.off + o;
            uint32_t name = rd<uint32_t>(elf, b);
            uint16_t shndx = rd<uint16_t>(elf, b + 6);
            uint64_t value = rd<uint64_t>(elf, b + 8);
            uint64_t size = rd<uint64_t>(elf, b + 16);
            if (str.off + name >= elf.size()) continue;
            if (strcmp(elf.data() + str.off + name, sym) != 0) continue;
            if (shndx == 0 || shndx >= shnum) return false;
            const Sec& s = secs[shndx];
            bytes->assign(size, 0);
            if (s.type == 8) return true;  // SHT_NOBITS: zero-initialised
            uint64_t fo = s.off + (value - s.addr);
            if (fo + size > elf.size()) return false;
            memcpy(bytes->data(), elf.data() + fo, size);
            return true;
        }
    }
    return false;
}

}  // namespace lsd
