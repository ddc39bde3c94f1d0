// Host side of the kinetic-law programs: the formula compiler and the mass-action shape.
#pragma once

#include <map>
#include <string>
#include <vector>

#include "host.h"
#include "law.h"

namespace gpgmp {

// Names a formula may use: the species ids (int counts in stochastic programs, float
// concentrations in deterministic ones), the user parameters (`const Real <id> = <value>;`,
// Solver.cpp:834-841), the header constants (PhysicalCellWidth, ...; Solver.cpp:494-505),
// GridX / GridY, and the internal _h0 and _mask<i>.
struct LawEnv {
    std::vector<std::string> species;
    std::map<std::string, float> params;  // parameters and header constants, as float values
    std::vector<std::string> fields;      // field parameters (drift/diffusivity methods)
    bool deterministic = false;
};

struct LawProgram {
    std::vector<uint32_t> code;
    LawType type = LT_I32;
};

// Compiles the C expression `text`; throws Error(GPGMP_E_INVALID) with the reason.
LawProgram compile_law(const std::string &text, const LawEnv &env);

// The mass-action shape c * X1 * (X2 - m2) * ... (left to right, c a constant subexpression of
// type int, float or double, possibly as the second element after one species factor, or
// absent): the reaction table evaluates exactly this chain.  False for any other formula.
struct MassAction {
    LawType type = LT_I32;
    double value = 1.0;
    std::vector<int> spec, off;
    bool commuted = false;  // written X * c * ...: the constant second
};
bool mass_action_shape(const std::string &text, const LawEnv &env, MassAction &ma);

// A drift/diffusivity method compiled for every species: programs[f * S + s] gives field f
// (0 DiffusivityX, 1 DiffusivityY, 2 DriftX, 3 DriftY) of species s -- the value the method
// leaves in species s's struct when run for s (`All` = s), or its current value when the
// method does not assign it (Solver.cpp:659-777 saves the function species' fields only).
std::vector<LawProgram> compile_method(const std::string &text, const LawEnv &env);
// What a compiled method reads besides the cell's position, the constants and the field
// parameters: the counts (X->State, central differences), the clock, the current fields.
// A method that reads none of them gives the same fields every time it is evaluated (a field
// it does not assign keeps its value, which is not a dependency).  progs: compile_method's,
// for S species.
constexpr uint32_t kMethodReadsState = 1, kMethodReadsTime = 2, kMethodReadsFields = 4;
uint32_t method_dependencies(const std::vector<LawProgram> &progs, int S);

// The names a model's laws may use, and reaction i's law text (custom or generated).
LawEnv law_env(const Model &m, bool deterministic);
std::string law_text(const Model &m, int i, bool deterministic);

// Host evaluation of what the kernels compute (for tests of the compiler): reaction i's
// `_h[0] + [mask *] <law>` for state x; how = 0 as the solver would (table or program),
// 1 always the program, 2 always the table.  law_rhs_host: species s's deterministic
// expression (which 0 RK4 derivative, 1 alpha-QSS production, 2 destruction) in float state y.
float law_hazard_host(const Model &m, int i, const int32_t *x, int gx, int gy, float h0, uint32_t act, int how);
double law_rhs_host(const Model &m, int s, int which, const float *y, int gx, int gy, uint32_t act, float *out);

}  // namespace gpgmp
