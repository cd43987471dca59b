// hipRTC helpers (no device needed for compilation).
#pragma once

#include <string>
#include <vector>

namespace lsd {

// compile options for device programs on the given GPU architecture
std::vector<std::string> device_flags(const std::string& arch);

// compile `source` to a code object; false with the build log on failure
bool rtc_compile(const std::string& source, const std::string& name,
                 const std::vector<std::string>& options, std::vector<char>* code,
                 std::string* log);

// the compiler that rtc_compile runs in this process: the hipRTC library
// this library's calls resolve to and the code-object compiler libraries
// (amd_comgr) the process has loaded, each as path, size and mtime.  A
// process can hold another ROCm's copies (a framework imported before this
// library brings its own), and the same source built by two of them need
// not round alike, so whatever caches code objects keys them by this too.
std::string compiler_identity();

// bytes of a named data symbol in a code object (ELF64)
bool elf_symbol(const std::vector<char>& elf, const char* sym, std::vector<char>* bytes);

}  // namespace lsd
