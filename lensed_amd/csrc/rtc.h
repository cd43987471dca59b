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

// bytes of a named data symbol in a code object (ELF64)
bool elf_symbol(const std::vector<char>& elf, const char* sym, std::vector<char>* bytes);

}  // namespace lsd
