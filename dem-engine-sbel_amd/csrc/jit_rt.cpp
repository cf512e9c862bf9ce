// Runtime specialisation of the force kernel through hipRTC, in place of the reference's jitify programs
// (JitHelper.cpp:44-48 substitutes placeholders into kernel sources; dT.cpp:1385-1388 launches them).
//
// The template and its headers are embedded in the library. A model string (or explicit placeholder values) is
// substituted in, the simulation constants are written in as literals, and the result is compiled for the device's
// architecture. Code objects are cached process-wide by source text (mutex-guarded), modules per context.
#include <hip/hiprtc.h>

#include <cstring>
#include <mutex>
#include <unordered_map>

#include "deme_internal.h"
#include "jit/jit_gen.h"

namespace deme_jit_text {
extern const char* const force_kernel_tpl;
extern const char* const owner_kernel_tpl;
extern const char* const inspect_kernel_tpl;
extern const char* const jit_prelude_h;
extern const char* const deme_device_h;
extern const char* const deme_types_h;
extern const char* const hertz_frictional_model;
extern const char* const hertz_frictionless_model;
}  // namespace deme_jit_text

namespace deme_int {
namespace {

std::mutex g_cacheMu;
std::unordered_map<std::string, std::vector<char>> g_codeCache;

std::string flat_includes(std::string s) {
    // the embedded headers are handed to hipRTC under flat names
    s = deme_jitgen::render(s, {{"#include \"../deme_device.h\"", "#include \"deme_device.h\""},
                                {"#include \"../../include/deme/deme_types.h\"", "#include \"deme_types.h\""}});
    return s;
}

}  // namespace

const char* jit_builtin_model(int which) {
    return which == 1 ? deme_jit_text::hertz_frictionless_model : deme_jit_text::hertz_frictional_model;
}

bool jit_force_source(const std::string& model, const std::vector<std::string>& wildcards, const deme_sim_params* p,
                      const std::vector<std::pair<std::string, std::string>>& overrides, std::string& src,
                      std::string& err) {
    deme_jitgen::Subst s;
    deme_jitgen::param_substitutions(p, s);
    if (!deme_jitgen::model_substitutions(model, wildcards, s, err)) return false;
    for (const auto& kv : overrides) s.push_back(kv);
    src = deme_jitgen::render(deme_jit_text::force_kernel_tpl, s);
    return true;
}

std::string jit_owner_source(const std::vector<std::pair<std::string, std::string>>& values) {
    deme_jitgen::Subst s;
    std::string vel, pos, fam;
    for (const auto& kv : values) {
        if (kv.first == "_velPrescriptionStrategy_") vel = kv.second;
        else if (kv.first == "_posPrescriptionStrategy_") pos = kv.second;
        else if (kv.first == "_familyChangeRules_") fam = kv.second;
    }
    deme_jitgen::owner_substitutions(vel, pos, fam, s);
    return deme_jitgen::render(deme_jit_text::owner_kernel_tpl, s);
}

std::string jit_inspect_source(const std::string& region) {
    deme_jitgen::Subst s;
    deme_jitgen::inspect_substitutions(region, s);
    return deme_jitgen::render(deme_jit_text::inspect_kernel_tpl, s);
}

bool jit_is_owner_key(const std::string& k) {
    return k == "_velPrescriptionStrategy_" || k == "_posPrescriptionStrategy_" || k == "_familyChangeRules_";
}

int jit_compile(const std::string& src, const std::string& arch, std::vector<char>& code, std::string& log) {
    const std::string key = arch + '\n' + src;
    {
        std::lock_guard<std::mutex> g(g_cacheMu);
        auto it = g_codeCache.find(key);
        if (it != g_codeCache.end()) {
            code = it->second;
            return 0;
        }
    }
    const std::string hdrText[3] = {flat_includes(deme_jit_text::jit_prelude_h),
                                    flat_includes(deme_jit_text::deme_device_h),
                                    flat_includes(deme_jit_text::deme_types_h)};
    const char* hdrs[3] = {hdrText[0].c_str(), hdrText[1].c_str(), hdrText[2].c_str()};
    const char* names[3] = {"jit_prelude.h", "deme_device.h", "deme_types.h"};
    hiprtcProgram prog;
    if (hiprtcCreateProgram(&prog, src.c_str(), "deme_force_jit.hip", 3, hdrs, names) != HIPRTC_SUCCESS) {
        log = "hiprtcCreateProgram failed";
        return 1;
    }
    const std::string archOpt = "--offload-arch=" + arch;
    const char* opts[] = {archOpt.c_str(), "-O3", "-std=c++17", "-ffp-contract=off", "-fno-fast-math",
                          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"};
    const hiprtcResult rc = hiprtcCompileProgram(prog, (int)(sizeof(opts) / sizeof(opts[0])), opts);
    size_t ls = 0;
    hiprtcGetProgramLogSize(prog, &ls);
    log.assign(ls, '\0');
    if (ls) hiprtcGetProgramLog(prog, &log[0]);
    while (!log.empty() && log.back() == '\0') log.pop_back();
    if (rc != HIPRTC_SUCCESS) {
        hiprtcDestroyProgram(&prog);
        if (log.empty()) log = hiprtcGetErrorString(rc);
        return 1;
    }
    size_t cs = 0;
    hiprtcGetCodeSize(prog, &cs);
    code.resize(cs);
    hiprtcGetCode(prog, code.data());
    hiprtcDestroyProgram(&prog);
    std::lock_guard<std::mutex> g(g_cacheMu);
    g_codeCache[key] = code;
    return 0;
}

void JitForce::release() {
    if (mod) (void)hipModuleUnload(mod);
    mod = nullptr;
    full = terms = nullptr;
}

hipError_t JitForce::load(const std::vector<char>& code) {
    release();
    hipError_t e = hipModuleLoadData(&mod, code.data());
    if (e != hipSuccess) {
        mod = nullptr;
        return e;
    }
    if ((e = hipModuleGetFunction(&full, mod, "deme_force_full")) != hipSuccess) return e;
    return hipModuleGetFunction(&terms, mod, "deme_force_terms");
}

hipError_t JitForce::launch(bool wantFull, const KParams& P, uint32_t nC, const uint32_t* idA, const uint32_t* idB,
                            const uint8_t* ctype, const SphereRec* sph, const OwnerRec* own, const float4* moi,
                            const deme_anal_obj* anal, const PlanePose* pose, const deme_material* mats,
                            const float4* histIn, const uint32_t* map, float4* histOut, const uint32_t* mirror,
                            float* aTerm, float* bTerm, float* gTerm, uint8_t* aOn, uint8_t* bOn, ForceRec* out,
                            hipStream_t st, SpecGuard guard) const {
    if (nC == 0) return hipSuccess;
    KParams p = P;
    void* args[] = {&p,    &nC,     &idA, &idB,     &ctype,  &sph,   &own,   &moi,   &anal, &pose, &mats,
                    &histIn, &map, &histOut, &mirror, &aTerm, &bTerm, &gTerm, &aOn, &bOn, &out, &guard};
    const uint32_t grid = (nC + 255) / 256;
    return hipModuleLaunchKernel(wantFull ? full : terms, grid, 1, 1, 256, 1, 1, 0, st, args, nullptr);
}

void JitOwner::release() {
    if (mod) (void)hipModuleUnload(mod);
    mod = nullptr;
    fn = nullptr;
}

hipError_t JitOwner::load(const std::vector<char>& code) {
    release();
    hipError_t e = hipModuleLoadData(&mod, code.data());
    if (e != hipSuccess) {
        mod = nullptr;
        return e;
    }
    return hipModuleGetFunction(&fn, mod, "deme_owner_update");
}

hipError_t JitOwner::launch(const KParams& P, uint32_t nO, uint32_t nClumps, uint32_t firstAnal, uint32_t nAnalOwners,
                            const uint32_t* ownerSphStart, const uint32_t* rowStart, const uint32_t* bRowStart,
                            const float* aTerm, const float* bTerm, const uint8_t* aOn, const uint8_t* bOn,
                            const float* analAcc, const deme_family_rule* rules, const OwnerRec* own,
                            OwnerRec* ownOut, float* acc, int doIntegrate, const double* clock,
                            hipStream_t st, SpecGuard guard) const {
    if (nO == 0) return hipSuccess;
    KParams p = P;
    void* args[] = {&p,     &nO,   &nClumps, &firstAnal, &nAnalOwners, &ownerSphStart, &rowStart, &bRowStart,
                    &aTerm, &bTerm, &aOn,    &bOn,       &analAcc,     &rules,         &own,      &ownOut,
                    &acc,   &doIntegrate, &clock, &guard};
    return hipModuleLaunchKernel(fn, (nO + 255) / 256, 1, 1, 256, 1, 1, 0, st, args, nullptr);
}

void JitInspect::release() {
    if (mod) (void)hipModuleUnload(mod);
    mod = nullptr;
    fn = nullptr;
}

hipError_t JitInspect::load(const std::vector<char>& code) {
    release();
    hipError_t e = hipModuleLoadData(&mod, code.data());
    if (e != hipSuccess) {
        mod = nullptr;
        return e;
    }
    return hipModuleGetFunction(&fn, mod, "deme_inspect_kernel");
}

hipError_t JitInspect::launch(const KParams& P, int quantity, int flavor, uint32_t n, const SphereRec* sph,
                              const OwnerRec* own, const float* vol, const float lbf[3], double* partial,
                              hipStream_t st) const {
    KParams p = P;
    float lx = lbf[0], ly = lbf[1], lz = lbf[2];
    void* args[] = {&p, &quantity, &flavor, &n, &sph, &own, &vol, &lx, &ly, &lz, &partial};
    return hipModuleLaunchKernel(fn, inspect_blocks(n), 1, 1, INSPECT_BLOCK, 1, 1, 0, st, args, nullptr);
}

}  // namespace deme_int

// ---- context-free entry points (render and compile; no device needed) ----
extern "C" int deme_jit_render(const char* model, const char* const* wildcards, int nWildcards,
                               const deme_sim_params* p, char* out, uint64_t cap, uint64_t* needed) {
    std::vector<std::string> w;
    for (int k = 0; k < nWildcards; k++) w.emplace_back(wildcards[k]);
    std::string src, err;
    const std::string m = model ? model : deme_int::jit_builtin_model(0);
    if (!model) w = {"delta_tan_x", "delta_tan_y", "delta_tan_z", "delta_time"};
    if (!deme_int::jit_force_source(m, w, p, {}, src, err)) return DEME_ERR_INVALID_ARGUMENT;
    if (needed) *needed = src.size() + 1;
    if (out && cap) {
        const size_t n = std::min<size_t>(cap - 1, src.size());
        std::memcpy(out, src.data(), n);
        out[n] = '\0';
    }
    return DEME_OK;
}

extern "C" int deme_jit_render_owner(const char* velStrategy, const char* posStrategy, const char* familyChangeRules,
                                     char* out, uint64_t cap, uint64_t* needed) {
    std::vector<std::pair<std::string, std::string>> v;
    if (velStrategy) v.emplace_back("_velPrescriptionStrategy_", velStrategy);
    if (posStrategy) v.emplace_back("_posPrescriptionStrategy_", posStrategy);
    if (familyChangeRules) v.emplace_back("_familyChangeRules_", familyChangeRules);
    const std::string src = deme_int::jit_owner_source(v);
    if (needed) *needed = src.size() + 1;
    if (out && cap) {
        const size_t n = std::min<size_t>(cap - 1, src.size());
        std::memcpy(out, src.data(), n);
        out[n] = '\0';
    }
    return DEME_OK;
}

extern "C" int deme_jit_render_inspect(const char* region, char* out, uint64_t cap, uint64_t* needed) {
    const std::string src = deme_int::jit_inspect_source(region ? region : "");
    if (needed) *needed = src.size() + 1;
    if (out && cap) {
        const size_t n = std::min<size_t>(cap - 1, src.size());
        std::memcpy(out, src.data(), n);
        out[n] = '\0';
    }
    return DEME_OK;
}

extern "C" int deme_jit_compile(const char* source, char* log, uint64_t cap) {
    if (!source) return DEME_ERR_INVALID_ARGUMENT;
    std::vector<char> code;
    std::string lg;
    const int rc = deme_int::jit_compile(source, "gfx950", code, lg);
    if (log && cap) {
        const size_t n = std::min<size_t>(cap - 1, lg.size());
        std::memcpy(log, lg.data(), n);
        log[n] = '\0';
    }
    return rc ? DEME_ERR_JIT : DEME_OK;
}
