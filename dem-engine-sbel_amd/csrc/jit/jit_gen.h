// Source generation for the force kernel template (host-only C++, no HIP): placeholder values for a force model and
// for the simulation constants, and the substitution itself. Shared by the build-time generator of the
// ahead-of-time kernel and by the runtime hipRTC path, so both specialise the one template the same way.
#pragma once

#include <string>
#include <utility>
#include <vector>

#include "../../../include/deme/deme_types.h"

namespace deme_jitgen {

using Subst = std::vector<std::pair<std::string, std::string>>;

// Contact wildcards a model keeps per contact (at most this many, stored as one float4 per contact).
constexpr size_t MAX_CONTACT_WILDCARDS = 4;

// Placeholder values of a force model: the model body, the optional per-body quantities it names (ts, ALinVel,
// BLinVel, ARotVel, BRotVel, AOwnerFamily, BOwnerFamily, by whole-word match as Models.h:197-231 does), the contact
// wildcard load / store / reset code for the sorted wildcard names, and the material arrays. Returns false (with a
// reason) for more than MAX_CONTACT_WILDCARDS wildcards or a name that is not an identifier.
bool model_substitutions(const std::string& model, std::vector<std::string> wildcards, Subst& out, std::string& err);

// Simulation constants: read from the kernel parameters (p == nullptr) or written into the source as literals that
// parse back to the exact values.
void param_substitutions(const deme_sim_params* p, Subst& out);

// Placeholder values of the owner update kernel: the prescribed-velocity and prescribed-position strategies
// (switch cases on the family, as equipFamilyPrescribedMotions writes them, APIPrivate.cpp:1005-1048) and the family
// change rules (statements setting newFamily, APIPrivate.cpp:980-1003). Empty strings: none (the numeric table only).
void owner_substitutions(const std::string& velStrategy, const std::string& posStrategy,
                         const std::string& familyChangeRules, Subst& out);

// Placeholder value of the inspection kernel: the region code of an inspector (a function body over X, Y, Z that
// returns a bool, AuxClasses.cpp:121-139); empty: every entity counts.
void inspect_substitutions(const std::string& region, Subst& out);

// Replace every occurrence of each key in order; later entries with the same key as an earlier one win (overrides).
std::string render(const std::string& tpl, const Subst& s);

// True when `word` occurs in `text` delimited by non-identifier characters.
bool whole_word(const std::string& text, const std::string& word);

}  // namespace deme_jitgen
