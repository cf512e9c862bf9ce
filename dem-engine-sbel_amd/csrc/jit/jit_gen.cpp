// Source generation for the force kernel template; see jit_gen.h.
#include "jit_gen.h"

#include <algorithm>
#include <cstdio>
#include <map>

namespace deme_jitgen {

static bool ident_char(char c) {
    return (c >= 'a' && c <= 'z') || (c >= 'A' && c <= 'Z') || (c >= '0' && c <= '9') || c == '_';
}

bool whole_word(const std::string& text, const std::string& word) {
    if (word.empty()) return false;
    for (size_t p = text.find(word); p != std::string::npos; p = text.find(word, p + 1)) {
        const bool left = p == 0 || !ident_char(text[p - 1]);
        const size_t e = p + word.size();
        const bool right = e >= text.size() || !ident_char(text[e]);
        if (left && right) return true;
    }
    return false;
}

static bool valid_identifier(const std::string& s) {
    if (s.empty() || (s[0] >= '0' && s[0] <= '9')) return false;
    return std::all_of(s.begin(), s.end(), ident_char);
}

bool model_substitutions(const std::string& model, std::vector<std::string> wildcards, Subst& out, std::string& err) {
    std::sort(wildcards.begin(), wildcards.end());
    wildcards.erase(std::unique(wildcards.begin(), wildcards.end()), wildcards.end());
    if (wildcards.size() > MAX_CONTACT_WILDCARDS) {
        err = "at most 4 contact wildcards are kept per contact";
        return false;
    }
    for (const std::string& w : wildcards)
        if (!valid_identifier(w)) {
            err = "contact wildcard '" + w + "' is not a valid identifier";
            return false;
        }
    std::string def, acqA, acqB;
    if (whole_word(model, "ts")) def += "const float ts = P.h;\n";
    if (whole_word(model, "AOwnerFamily")) {
        def += "deme::family_t AOwnerFamily;\n";
        acqA += "AOwnerFamily = myOwnerRec.fam;\n";
    }
    if (whole_word(model, "BOwnerFamily")) {
        def += "deme::family_t BOwnerFamily;\n";
        acqB += "BOwnerFamily = myOwnerRec.fam;\n";
    }
    if (whole_word(model, "ALinVel") || whole_word(model, "BLinVel")) {
        def += "float3 ALinVel, BLinVel;\n";
        acqA += "ALinVel = make_float3(myOwnerRec.vx, myOwnerRec.vy, myOwnerRec.vz);\n";
        acqB += "BLinVel = make_float3(myOwnerRec.vx, myOwnerRec.vy, myOwnerRec.vz);\n";
    }
    if (whole_word(model, "ARotVel") || whole_word(model, "BRotVel")) {
        def += "float3 ARotVel, BRotVel;\n";
        acqA += "ARotVel = make_float3(myOwnerRec.wx, myOwnerRec.wy, myOwnerRec.wz);\n";
        acqB += "BRotVel = make_float3(myOwnerRec.wx, myOwnerRec.wy, myOwnerRec.wz);\n";
    }
    static const char* comp[4] = {"x", "y", "z", "w"};
    std::string wAcq, wWrite = "wc = make_float4(", wDestroy;
    for (size_t k = 0; k < MAX_CONTACT_WILDCARDS; k++) {
        if (k < wildcards.size()) {
            wAcq += "float " + wildcards[k] + " = wc." + comp[k] + ";\n";
            wDestroy += wildcards[k] + " = 0.f;\n";
            wWrite += wildcards[k];
        } else {
            wWrite += "0.f";
        }
        wWrite += k + 1 < MAX_CONTACT_WILDCARDS ? ", " : ");\n";
    }
    out.push_back({"_materialDefs_", "const deme_jit::MatProp E{mats, 0}, nu{mats, 1}, CoR{mats, 2}, mu{mats, 3}, "
                                     "Crr{mats, 4}"});
    out.push_back({"_forceModelIngredientDefinition_", def});
    out.push_back({"_forceModelIngredientAcqForA_", acqA});
    out.push_back({"_forceModelIngredientAcqForB_", acqB});
    out.push_back({"_forceModelContactWildcardAcq_", wAcq});
    out.push_back({"_forceModelContactWildcardWrite_", wWrite});
    out.push_back({"_forceModelContactWildcardDestroy_", wDestroy});
    out.push_back({"_DEMForceModel_", model});
    return true;
}

void owner_substitutions(const std::string& velStrategy, const std::string& posStrategy,
                         const std::string& familyChangeRules, Subst& out) {
    out.push_back({"_familyChangeRules_", familyChangeRules.empty() ? " " : familyChangeRules});
    out.push_back({"_velPrescriptionStrategy_", velStrategy.empty() ? " " : velStrategy});
    out.push_back({"_posPrescriptionStrategy_", posStrategy.empty() ? " " : posStrategy});
}

void inspect_substitutions(const std::string& region, Subst& out) {
    out.push_back({"_inRegionPolicy_", region.empty() ? "return true;" : region});
}

static std::string exact(double v) {
    char b[64];
    std::snprintf(b, sizeof b, "%.17g", v);
    std::string s(b);
    if (s.find_first_of(".eEn") == std::string::npos) s += ".0";
    return s;
}

void param_substitutions(const deme_sim_params* p, Subst& out) {
    if (!p) {
        out.push_back({"_nvXp2_", "P.nvX"});
        out.push_back({"_nvYp2_", "P.nvY"});
        out.push_back({"_voxelSize_", "P.voxelSize"});
        out.push_back({"_l_", "P.l"});
        return;
    }
    out.push_back({"_nvXp2_", std::to_string((unsigned)p->nvXp2) + "u"});
    out.push_back({"_nvYp2_", std::to_string((unsigned)p->nvYp2) + "u"});
    out.push_back({"_voxelSize_", exact(p->voxelSize)});
    out.push_back({"_l_", exact(p->l)});
}

std::string render(const std::string& tpl, const Subst& s) {
    // the last value given for a key wins; keys are applied in the order of their first appearance
    std::vector<std::string> order;
    std::map<std::string, std::string> value;
    for (const auto& kv : s) {
        if (!value.count(kv.first)) order.push_back(kv.first);
        value[kv.first] = kv.second;
    }
    std::string code = tpl;
    for (const std::string& key : order) {
        const std::string& v = value[key];
        std::string outS;
        outS.reserve(code.size());
        size_t pos = 0;
        for (size_t p = code.find(key); p != std::string::npos; p = code.find(key, pos)) {
            outS.append(code, pos, p - pos);
            outS += v;
            pos = p + key.size();
        }
        outS.append(code, pos, std::string::npos);
        code.swap(outS);
    }
    return code;
}

}  // namespace deme_jitgen
