// Build-time generator of the ahead-of-time kernels from their templates:
//   jitgen force <template> <model> <out.hip>   the force kernel with the default (frictional Hertz-Mindlin) model
//                                               and the simulation constants read from the kernel parameters
//   jitgen owner <template> <out.hip>           the owner update kernel with no motion strategies
//   jitgen inspect <template> <out.hip>         the inspection kernels with a region that counts everything
#include <fstream>
#include <iostream>
#include <sstream>

#include "jit_gen.h"

int main(int argc, char** argv) {
    const bool owner = argc == 4 && std::string(argv[1]) == "owner";
    const bool inspect = argc == 4 && std::string(argv[1]) == "inspect";
    if (!(owner || inspect || (argc == 5 && std::string(argv[1]) == "force"))) {
        std::cerr << "usage: jitgen force <template> <model> <out.hip> | jitgen owner|inspect <template> <out.hip>\n";
        return 2;
    }
    auto slurp = [](const char* path, std::string& s) {
        std::ifstream f(path);
        if (!f) return false;
        std::stringstream b;
        b << f.rdbuf();
        s = b.str();
        return true;
    };
    std::string tpl, model, err;
    if (owner || inspect) {
        if (!slurp(argv[2], tpl)) {
            std::cerr << "jitgen: cannot read inputs\n";
            return 1;
        }
        deme_jitgen::Subst s;
        if (owner)
            deme_jitgen::owner_substitutions("", "", "", s);
        else
            deme_jitgen::inspect_substitutions("", s);
        std::ofstream o(argv[3]);
        o << "// generated from " << argv[2] << " by jitgen; do not edit\n" << deme_jitgen::render(tpl, s);
        return o ? 0 : 1;
    }
    argv++;
    if (!slurp(argv[1], tpl) || !slurp(argv[2], model)) {
        std::cerr << "jitgen: cannot read inputs\n";
        return 1;
    }
    deme_jitgen::Subst s;
    deme_jitgen::param_substitutions(nullptr, s);
    if (!deme_jitgen::model_substitutions(model, {"delta_tan_x", "delta_tan_y", "delta_tan_z", "delta_time"}, s, err)) {
        std::cerr << "jitgen: " << err << "\n";
        return 1;
    }
    std::ofstream o(argv[3]);
    o << "// generated from " << argv[1] << " and " << argv[2] << " by jitgen; do not edit\n"
      << deme_jitgen::render(tpl, s);
    return o ? 0 : 1;
}
