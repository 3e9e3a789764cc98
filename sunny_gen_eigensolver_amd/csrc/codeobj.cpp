#include "codeobj.hpp"

#include <hip/hiprtc.h>

#include <sys/stat.h>
#include <unistd.h>

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>

// Generated at build time from kernels/geneig_kernels.hip: defines
// `static const char kGeneigKernelSource[]`.
#include "_gen/kernel_source.inc"

namespace sgeig {

void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) {
        std::string msg = std::string(what) + ": " + hipGetErrorString(e);
        throw HipError(e, msg);
    }
}

const char* layer_name(CacheLayer l) {
    switch (l) {
        case CacheLayer::InProcess: return "in-process";
        case CacheLayer::Disk: return "disk";
        default: return "compile";
    }
}

LoadedModule::~LoadedModule() {
    if (module) {
        int cur = 0;
        if (hipGetDevice(&cur) == hipSuccess && cur != device) (void)hipSetDevice(device);
        (void)hipModuleUnload(module);
        if (cur != device) (void)hipSetDevice(cur);
    }
}

hipFunction_t LoadedModule::get(const char* name) {
    auto it = functions.find(name);
    if (it != functions.end()) return it->second;
    hipFunction_t f = nullptr;
    SG_HIP(hipModuleGetFunction(&f, module, name));
    functions[name] = f;
    return f;
}

hipFunction_t LoadedModule::find(const char* name) {
    auto it = functions.find(name);
    if (it != functions.end()) return it->second;
    hipFunction_t f = nullptr;
    if (hipModuleGetFunction(&f, module, name) != hipSuccess) {
        (void)hipGetLastError();
        return nullptr;
    }
    functions[name] = f;
    return f;
}

namespace {

uint64_t fnv1a(const void* data, size_t len, uint64_t h = 0xcbf29ce484222325ULL) {
    const unsigned char* p = static_cast<const unsigned char*>(data);
    for (size_t i = 0; i < len; ++i) {
        h ^= p[i];
        h *= 0x100000001b3ULL;
    }
    return h;
}

std::vector<std::string> compile_options(const ModuleKey& key) {
    std::vector<std::string> o = {
        "--offload-arch=" + key.arch,
        "-O3",
        "-std=c++17",
        "-DGENEIG_N=" + std::to_string(key.n),
        "-DGENEIG_NT=" + std::to_string(key.block_dim),
        // Keep the compiler from emitting packed FP32 instructions.
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    };
    if (key.fp32) o.push_back("-DGENEIG_FP32");
    if (key.staged) o.push_back("-DGENEIG_STAGED");
    return o;
}

int rtc_version() {
    int maj = 0, min = 0;
    hiprtcVersion(&maj, &min);
    return maj * 1000 + min;
}

uint64_t key_hash(const ModuleKey& key) {
    uint64_t h = fnv1a(kGeneigKernelSource, sizeof(kGeneigKernelSource));
    for (const auto& o : compile_options(key)) h = fnv1a(o.data(), o.size(), h);
    int v = rtc_version();
    h = fnv1a(&v, sizeof(v), h);
    return h;
}

bool mkdirs(const std::string& path) {
    if (path.empty()) return false;
    std::string cur;
    std::stringstream ss(path);
    std::string part;
    if (path[0] == '/') cur = "/";
    while (std::getline(ss, part, '/')) {
        if (part.empty()) continue;
        cur += part + "/";
        mkdir(cur.c_str(), 0755);
    }
    struct stat st;
    return stat(path.c_str(), &st) == 0 && S_ISDIR(st.st_mode);
}

std::mutex g_cache_mutex;
std::map<std::pair<ModuleKey, int>, std::weak_ptr<LoadedModule>> g_in_process;

std::shared_ptr<LoadedModule> load_code(const std::vector<char>& code, int device) {
    auto m = std::make_shared<LoadedModule>();
    m->device = device;
    hipError_t e = hipModuleLoadData(&m->module, code.data());
    if (e != hipSuccess) {
        m->module = nullptr;
        throw HipError(e, std::string("hipModuleLoadData: ") + hipGetErrorString(e));
    }
    return m;
}

bool read_file(const std::string& path, std::vector<char>& out) {
    std::ifstream f(path, std::ios::binary | std::ios::ate);
    if (!f) return false;
    std::streamsize sz = f.tellg();
    if (sz <= 0) return false;
    out.resize((size_t)sz);
    f.seekg(0);
    f.read(out.data(), sz);
    return (bool)f;
}

}  // namespace

std::string device_arch(int device) {
    hipDeviceProp_t prop;
    SG_HIP(hipGetDeviceProperties(&prop, device));
    std::string a = prop.gcnArchName;
    auto colon = a.find(':');
    if (colon != std::string::npos) a = a.substr(0, colon);
    return a;
}

std::string cache_dir() {
    const char* env = std::getenv("SUNNY_GENEIG_CACHE_DIR");
    std::string dir;
    if (env && env[0]) {
        dir = env;
    } else if (const char* xdg = std::getenv("XDG_CACHE_HOME"); xdg && xdg[0]) {
        dir = std::string(xdg) + "/sunny_geneig";
    } else if (const char* home = std::getenv("HOME"); home && home[0]) {
        dir = std::string(home) + "/.cache/sunny_geneig";
    } else {
        dir = "/tmp/sunny_geneig";
    }
    mkdirs(dir);
    return dir;
}

std::string code_object_path(const ModuleKey& key) {
    char name[160];
    std::snprintf(name, sizeof(name), "geneig_n%d_wg%d_%s%s%s_%016llx.hsaco", key.n, key.block_dim,
                  key.arch.c_str(), key.fp32 ? "_fp32" : "", key.staged ? "_staged" : "",
                  (unsigned long long)key_hash(key));
    return cache_dir() + "/" + name;
}

std::vector<char> compile_code_object(const ModuleKey& key) {
    hiprtcProgram prog;
    if (hiprtcCreateProgram(&prog, kGeneigKernelSource, "geneig_kernels.hip", 0, nullptr,
                            nullptr) != HIPRTC_SUCCESS)
        throw CompileError("hiprtcCreateProgram failed");
    auto opts = compile_options(key);
    std::vector<const char*> argv;
    for (auto& o : opts) argv.push_back(o.c_str());
    hiprtcResult r = hiprtcCompileProgram(prog, (int)argv.size(), argv.data());
    if (r != HIPRTC_SUCCESS) {
        size_t ls = 0;
        hiprtcGetProgramLogSize(prog, &ls);
        std::string log(ls, '\0');
        if (ls) hiprtcGetProgramLog(prog, &log[0]);
        hiprtcDestroyProgram(&prog);
        throw CompileError("hipRTC compile of N=" + std::to_string(key.n) + " failed: " +
                           hiprtcGetErrorString(r) + "\n" + log);
    }
    size_t cs = 0;
    hiprtcGetCodeSize(prog, &cs);
    std::vector<char> code(cs);
    hiprtcGetCode(prog, code.data());
    hiprtcDestroyProgram(&prog);
    return code;
}

void write_disk(const ModuleKey& key, const std::vector<char>& code) {
    const std::string path = code_object_path(key);
    char tmp_suffix[64];
    std::snprintf(tmp_suffix, sizeof(tmp_suffix), ".tmp.%d.%lld", (int)getpid(),
                  (long long)std::chrono::steady_clock::now().time_since_epoch().count());
    const std::string tmp = path + tmp_suffix;
    {
        std::ofstream f(tmp, std::ios::binary | std::ios::trunc);
        if (!f) {
            std::fprintf(stderr, "[sunny_geneig cache] cannot write %s\n", tmp.c_str());
            return;
        }
        f.write(code.data(), (std::streamsize)code.size());
        if (!f) {
            std::fprintf(stderr, "[sunny_geneig cache] short write %s\n", tmp.c_str());
            std::remove(tmp.c_str());
            return;
        }
    }
    if (std::rename(tmp.c_str(), path.c_str()) != 0) {
        std::fprintf(stderr, "[sunny_geneig cache] rename to %s failed\n", path.c_str());
        std::remove(tmp.c_str());
    }
}

std::shared_ptr<LoadedModule> acquire_module(const ModuleKey& key, CacheMode mode,
                                             CacheLayer* layer, double* ms) {
    auto t0 = std::chrono::steady_clock::now();
    int device = 0;
    SG_HIP(hipGetDevice(&device));
    auto finish = [&](std::shared_ptr<LoadedModule> m, CacheLayer l) {
        if (layer) *layer = l;
        if (ms)
            *ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0)
                      .count();
        return m;
    };
    const auto mkey = std::make_pair(key, device);
    if (mode != CacheMode::NoCache) {
        std::lock_guard<std::mutex> lk(g_cache_mutex);
        auto it = g_in_process.find(mkey);
        if (it != g_in_process.end()) {
            if (auto sp = it->second.lock()) return finish(sp, CacheLayer::InProcess);
        }
    }
    if (mode == CacheMode::Auto) {
        const std::string path = code_object_path(key);
        std::vector<char> code;
        if (read_file(path, code)) {
            bool ok = code.size() >= 64 && code[0] == 0x7f && code[1] == 'E' && code[2] == 'L' &&
                      code[3] == 'F';
            std::shared_ptr<LoadedModule> m;
            if (ok) {
                try {
                    m = load_code(code, device);
                } catch (const HipError&) {
                    ok = false;
                }
            }
            if (ok) {
                std::lock_guard<std::mutex> lk(g_cache_mutex);
                g_in_process[mkey] = m;
                return finish(m, CacheLayer::Disk);
            }
            std::fprintf(stderr, "[sunny_geneig cache] invalid code object %s: recompiling\n",
                         path.c_str());
            std::remove(path.c_str());
        }
    }
    std::vector<char> code = compile_code_object(key);
    auto m = load_code(code, device);
    if (mode == CacheMode::Auto) write_disk(key, code);
    if (mode != CacheMode::NoCache) {
        std::lock_guard<std::mutex> lk(g_cache_mutex);
        g_in_process[mkey] = m;
    }
    return finish(m, CacheLayer::Compile);
}

}  // namespace sgeig
