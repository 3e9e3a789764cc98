// Per-order code objects for the gfx950 kernels: hipRTC compilation of the
// embedded kernel source, an in-process cache of loaded modules and an
// on-disk cache of code objects.
//
// Reference behaviour followed (src/unified/unified_solver.cpp):
//   - three layers, in-process weak_ptr map -> disk -> compile (:522-597),
//   - cache directory $SUNNY_GENEIG_CACHE_DIR -> $XDG_CACHE_HOME -> ~/.cache
//     -> /tmp (:403-420),
//   - disk entries validated (size, ELF magic, successful module load) and
//     deleted + recompiled when bad (:434-479),
//   - atomic write through a temporary file and rename (:481-520),
//   - FNV-1a-64 content hash in the key (:321-371).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

namespace sgeig {

struct HipError : std::runtime_error {
    hipError_t code;
    HipError(hipError_t c, const std::string& what) : std::runtime_error(what), code(c) {}
};
struct CompileError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

void hip_check(hipError_t e, const char* what);
#define SG_HIP(expr) ::sgeig::hip_check((expr), #expr)

// What a compiled module is specialised for.
struct ModuleKey {
    int n = 0;
    int block_dim = 0;
    std::string arch;   // base gfx name, e.g. "gfx950"
    bool fp32 = false;  // FP32 build of the split stages (mixed-precision path)
    bool staged = false; // the HBM-staged pipeline at an order up to 64
    bool operator<(const ModuleKey& o) const {
        if (n != o.n) return n < o.n;
        if (block_dim != o.block_dim) return block_dim < o.block_dim;
        if (fp32 != o.fp32) return fp32 < o.fp32;
        if (staged != o.staged) return staged < o.staged;
        return arch < o.arch;
    }
};

// A loaded module; unloads on destruction.
struct LoadedModule {
    hipModule_t module = nullptr;
    int device = 0;
    std::map<std::string, hipFunction_t> functions;
    ~LoadedModule();
    hipFunction_t get(const char* name);
    // nullptr where the module has no such kernel (built only for some
    // orders or precisions).
    hipFunction_t find(const char* name);
};

enum class CacheLayer { InProcess, Disk, Compile };
const char* layer_name(CacheLayer l);

// Compile the kernel source for `key` with hipRTC; throws CompileError.
std::vector<char> compile_code_object(const ModuleKey& key);

// Resolved cache directory (created on demand).
std::string cache_dir();
// On-disk file name for a key (includes the source/options/compiler hash).
std::string code_object_path(const ModuleKey& key);
// Write `code` for `key` to the disk cache atomically; failures are logged and
// ignored (a cache is an optimisation).
void write_disk(const ModuleKey& key, const std::vector<char>& code);

enum class CacheMode { Auto, NoDisk, NoCache };

// Acquire a loaded module for `key` on the current device, honouring `mode`.
std::shared_ptr<LoadedModule> acquire_module(const ModuleKey& key, CacheMode mode,
                                             CacheLayer* layer, double* ms);

// Base architecture name of the current device ("gfx950" from
// "gfx950:sramecc+:xnack-").
std::string device_arch(int device);

}  // namespace sgeig
