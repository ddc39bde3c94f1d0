// Compiler of kinetic-law formulas (the C expression text of Reaction::setCustomLaw,
// Reaction.cpp:108-111) into the typed stack programs of law.h, and recognition of the
// mass-action shape that the reaction table evaluates directly.
#include "law_compile.h"

#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>

namespace gpgmp {

namespace {

struct Node {
    // CMP: comparison (op '<' '>' 'l' (<=) 'g' (>=) '=' '!'), SEL: a ? b : c, CVT: a to `type`;
    // STATE/CUR/CDX/CDY/FIELD/PX/PY/TIME: the inputs of drift/diffusivity methods (LawExtra)
    enum Kind { NUM, SPEC, GX, GY, H0, MASK, BIN, NEG, CALL, CMP, SEL, CVT, STATE, CUR, CDX, CDY, FIELD, PX, PY,
                TIME } kind;
    LawType type = LT_I32;
    double value = 0;   // NUM
    int idx = 0;        // SPEC / MASK / STATE / CUR / CDX / CDY / FIELD: species, reaction, field
    char op = 0;        // BIN: + - * /
    uint32_t fn = 0;    // CALL: LawFn, or 100 = pow, 101 = abs (int)
    int a = -1, b = -1, c = -1; // children
};

LawType promote(LawType a, LawType b) {
    if (a == LT_F64 || b == LT_F64) return LT_F64;
    if (a == LT_F32 || b == LT_F32) return LT_F32;
    return LT_I32;
}

class Parser {
  public:
    Parser(const std::string &text, const LawEnv &env) : s_(text), env_(env) {}

    int parse() {
        const int n = expr();
        skip();
        if (p_ != s_.size()) fail("unexpected '" + s_.substr(p_, 8) + "'");
        return n;
    }
    std::vector<Node> nodes;

    // A drift/diffusivity method (DiffusionModel::computeDriftDiffusivityMethod, pasted into
    // computeDiffusionConstants_forSpecies_<id>, Solver.cpp:659-777) run for species `all`:
    // `Real x = e;` locals, `X->DiffusivityX = e;` (and Y, DriftX, DriftY) assignments,
    // `x = e;` reassignments, comments.  Returns the final expression of each of the four
    // fields of every species (cur_out[f * S + s]); -1 where the method does not assign it.
    void method(int all, std::vector<int> &cur_out) {
        all_ = all;
        const int S = (int)env_.species.size();
        cur_.assign((size_t)4 * S, -1);
        for (;;) {
            skip_ws_comments();
            if (p_ >= s_.size()) break;
            std::string id = ident_token();
            if (id.empty()) fail("expected a statement");
            if (id == "Real" || id == "float" || id == "double" || id == "int") {
                skip_ws_comments();
                const std::string name = ident_token();
                if (name.empty()) fail("expected a variable name");
                if (!eat('=')) fail("expected '=' in the declaration of " + name);
                const LawType t = (id == "double") ? LT_F64 : (id == "int") ? LT_I32 : LT_F32;
                locals_[name] = convert(expr(), t);
            } else if (peek("->")) {
                p_ += 2;
                const int sp = species_ref(id);
                const int f = field_index(ident_token());
                if (!eat('=')) fail("expected '='");
                cur_[(size_t)f * S + sp] = convert(expr(), LT_F32);
            } else {
                auto it = locals_.find(id);
                if (it == locals_.end()) fail("assignment to unknown variable " + id);
                if (!eat('=')) fail("expected '='");
                it->second = convert(expr(), nodes[it->second].type);
            }
            if (!eat(';')) fail("expected ';'");
        }
        cur_out = cur_;
    }

  private:
    int all_ = -1;
    std::map<std::string, int> locals_;
    std::vector<int> cur_;

    void skip_ws_comments() {
        for (;;) {
            skip();
            if (p_ + 1 < s_.size() && s_[p_] == '/' && s_[p_ + 1] == '/') {
                while (p_ < s_.size() && s_[p_] != '\n') ++p_;
            } else if (p_ + 1 < s_.size() && s_[p_] == '/' && s_[p_ + 1] == '*') {
                const size_t e = s_.find("*/", p_ + 2);
                p_ = (e == std::string::npos) ? s_.size() : e + 2;
            } else {
                return;
            }
        }
    }
    bool peek(const char *t) {
        skip();
        return s_.compare(p_, std::strlen(t), t) == 0;
    }
    std::string ident_token() {
        skip();
        const size_t st = p_;
        if (p_ < s_.size() && (std::isalpha((unsigned char)s_[p_]) || s_[p_] == '_'))
            while (p_ < s_.size() && (std::isalnum((unsigned char)s_[p_]) || s_[p_] == '_')) ++p_;
        return s_.substr(st, p_ - st);
    }
    int species_ref(const std::string &id) {
        if (id == "All" && all_ >= 0) return all_;
        for (size_t i = 0; i < env_.species.size(); ++i)
            if (env_.species[i] == id) return (int)i;
        fail("unknown species " + id);
    }
    int field_index(const std::string &f) {
        static const char *names[] = {"DiffusivityX", "DiffusivityY", "DriftX", "DriftY"};
        for (int k = 0; k < 4; ++k)
            if (f == names[k]) return k;
        if (f == "State") return 4;
        if (f == "MeanX") fail("MeanX (the moments of computeMoments models) is not supported");
        fail("unknown species field " + f);
    }
    int convert(int n, LawType t) {
        if (nodes[n].type == t) return n;
        Node c;
        c.kind = Node::CVT;
        c.type = t;
        c.a = n;
        return add(c);
    }
    [[noreturn]] void fail(const std::string &m) {
        if (all_ >= 0)
            throw Error(GPGMP_E_INVALID, "drift/diffusivity method, near \"" + s_.substr(p_ > 20 ? p_ - 20 : 0, 40) +
                                             "\": " + m);
        throw Error(GPGMP_E_INVALID, "kinetic law \"" + s_ + "\": " + m);
    }
    void skip() {
        while (p_ < s_.size() && std::isspace((unsigned char)s_[p_])) ++p_;
    }
    bool eat(char c) {
        skip();
        if (p_ < s_.size() && s_[p_] == c) {
            ++p_;
            return true;
        }
        return false;
    }
    int add(Node n) {
        nodes.push_back(n);
        if (nodes.size() > 2048) fail("expression too large");
        return (int)nodes.size() - 1;
    }
    int bin(char op, int a, int b) {
        Node n;
        n.kind = Node::BIN;
        n.op = op;
        n.a = a;
        n.b = b;
        n.type = promote(nodes[a].type, nodes[b].type);
        return add(n);
    }
    int expr() {  // conditional-expression
        const int c = equality();
        if (eat('?')) {
            const int a = expr();
            if (!eat(':')) fail("expected ':'");
            const int b = expr();
            Node n;
            n.kind = Node::SEL;
            n.a = c;
            n.b = a;
            n.c = b;
            n.type = promote(nodes[a].type, nodes[b].type);
            return add(n);
        }
        return c;
    }
    int cmp(char op, int a, int b) {
        Node n;
        n.kind = Node::CMP;
        n.op = op;
        n.a = a;
        n.b = b;
        n.type = LT_I32;
        return add(n);
    }
    int equality() {
        int n = relational();
        for (;;) {
            if (peek("==")) {
                p_ += 2;
                n = cmp('=', n, relational());
            } else if (peek("!=")) {
                p_ += 2;
                n = cmp('!', n, relational());
            } else {
                return n;
            }
        }
    }
    int relational() {
        int n = additive();
        for (;;) {
            if (peek("<=")) {
                p_ += 2;
                n = cmp('l', n, additive());
            } else if (peek(">=")) {
                p_ += 2;
                n = cmp('g', n, additive());
            } else if (peek("<")) {
                ++p_;
                n = cmp('<', n, additive());
            } else if (peek(">")) {
                ++p_;
                n = cmp('>', n, additive());
            } else {
                return n;
            }
        }
    }
    int additive() {
        int n = term();
        for (;;) {
            if (eat('+')) n = bin('+', n, term());
            else if (eat('-')) n = bin('-', n, term());
            else return n;
        }
    }
    int term() {
        int n = unary();
        for (;;) {
            if (eat('*')) n = bin('*', n, unary());
            else if (eat('/')) n = bin('/', n, unary());
            else return n;
        }
    }
    int unary() {
        if (eat('-')) {
            const int c = unary();
            Node n;
            n.kind = Node::NEG;
            n.a = c;
            n.type = nodes[c].type;
            return add(n);
        }
        if (eat('+')) return unary();
        return primary();
    }
    int number() {
        // a C floating or integer literal: digits, '.', exponent, optional f suffix
        const size_t st = p_;
        bool flt = false;
        while (p_ < s_.size() && std::isdigit((unsigned char)s_[p_])) ++p_;
        if (p_ < s_.size() && s_[p_] == '.') {
            flt = true;
            ++p_;
            while (p_ < s_.size() && std::isdigit((unsigned char)s_[p_])) ++p_;
        }
        if (p_ < s_.size() && (s_[p_] == 'e' || s_[p_] == 'E')) {
            flt = true;
            ++p_;
            if (p_ < s_.size() && (s_[p_] == '+' || s_[p_] == '-')) ++p_;
            if (p_ >= s_.size() || !std::isdigit((unsigned char)s_[p_])) fail("bad exponent");
            while (p_ < s_.size() && std::isdigit((unsigned char)s_[p_])) ++p_;
        }
        const std::string lit = s_.substr(st, p_ - st);
        Node n;
        n.kind = Node::NUM;
        if (p_ < s_.size() && (s_[p_] == 'f' || s_[p_] == 'F')) {
            if (!flt) fail("bad float literal");
            ++p_;
            n.type = LT_F32;
            n.value = (double)std::strtof(lit.c_str(), nullptr);
        } else if (flt) {
            n.type = LT_F64;
            n.value = std::strtod(lit.c_str(), nullptr);
        } else {
            const unsigned long long v = std::strtoull(lit.c_str(), nullptr, 10);
            if (v > 0x7fffffffull) fail("integer literal " + lit + " does not fit an int");
            n.type = LT_I32;
            n.value = (double)v;
        }
        if (p_ < s_.size() && (std::isalnum((unsigned char)s_[p_]) || s_[p_] == '_')) fail("bad literal " + lit);
        return add(n);
    }
    int call(const std::string &name) {
        std::vector<int> args;
        if (!eat(')')) {
            for (;;) {
                args.push_back(expr());
                if (eat(')')) break;
                if (!eat(',')) fail("expected ',' or ')'");
            }
        }
        Node n;
        if (name == "getCentralDifferenceX" || name == "getCentralDifferenceY") {
            // args were parsed as expressions: the one argument must be a species reference
            if (args.size() != 1 || nodes[args[0]].kind != Node::STATE) fail(name + " takes a species");
            n.kind = (name == "getCentralDifferenceX") ? Node::CDX : Node::CDY;
            n.idx = nodes[args[0]].idx;
            n.type = LT_F32;
            return add(n);
        }
        n.kind = Node::CALL;
        static const char *unary_fns[] = {"exp", "log", "log10", "sqrt", "fabs", "floor", "ceil"};
        if (name == "pow") {
            if (args.size() != 2) fail("pow takes two arguments");
            n.fn = 100;
            n.a = args[0];
            n.b = args[1];
            n.type = (nodes[args[0]].type == LT_F64 || nodes[args[1]].type == LT_F64) ? LT_F64 : LT_F32;
            return add(n);
        }
        if (name == "abs" && args.size() == 1 && nodes[args[0]].type == LT_I32) {
            n.fn = 101;
            n.a = args[0];
            n.type = LT_I32;
            return add(n);
        }
        for (uint32_t f = 0; f < 7; ++f) {
            if (name == unary_fns[f] || (name == "abs" && f == LF_FABS)) {
                if (args.size() != 1) fail(name + " takes one argument");
                n.fn = f;
                n.a = args[0];
                n.type = (nodes[args[0]].type == LT_F64) ? LT_F64 : LT_F32;
                return add(n);
            }
        }
        fail("unknown function " + name);
    }
    int primary() {
        skip();
        if (p_ >= s_.size()) fail("unexpected end");
        const char c = s_[p_];
        if (c == '(') {
            ++p_;
            const int n = expr();
            if (!eat(')')) fail("expected ')'");
            return n;
        }
        if (std::isdigit((unsigned char)c) || c == '.') return number();
        if (std::isalpha((unsigned char)c) || c == '_') {
            const size_t st = p_;
            while (p_ < s_.size() && (std::isalnum((unsigned char)s_[p_]) || s_[p_] == '_')) ++p_;
            const std::string id = s_.substr(st, p_ - st);
            if (eat('(')) return call(id);
            return ident(id);
        }
        fail(std::string("unexpected '") + c + "'");
    }
    int ident(const std::string &id) {
        Node n;
        if (all_ >= 0) {  // method scope
            auto l = locals_.find(id);
            if (l != locals_.end()) return l->second;
            if (peek("->")) {
                p_ += 2;
                const int sp = species_ref(id);
                const int f = field_index(ident_token());
                const int S = (int)env_.species.size();
                if (f == 4) {
                    n.kind = Node::STATE;
                    n.idx = sp;
                    n.type = LT_I32;
                    return add(n);
                }
                if (cur_[(size_t)f * S + sp] >= 0) return cur_[(size_t)f * S + sp];
                n.kind = Node::CUR;
                n.idx = sp * 4 + f;
                n.type = LT_F32;
                return add(n);
            }
            if (id == "All" || [&] {
                    for (const auto &x : env_.species)
                        if (x == id) return true;
                    return false;
                }()) {
                // a bare species: the argument of getCentralDifferenceX/Y (a _species_t *)
                n.kind = Node::STATE;
                n.idx = species_ref(id);
                n.type = LT_I32;
                return add(n);
            }
            if (id == "PhysicalX" || id == "PhysicalY" || id == "PhysicalSimTime") {
                n.kind = (id == "PhysicalX") ? Node::PX : (id == "PhysicalY") ? Node::PY : Node::TIME;
                n.type = LT_F32;
                return add(n);
            }
            for (size_t k = 0; k < env_.fields.size(); ++k)
                if (env_.fields[k] == id) {
                    n.kind = Node::FIELD;
                    n.idx = (int)k;
                    n.type = LT_F32;
                    return add(n);
                }
        }
        for (size_t i = 0; i < env_.species.size(); ++i) {
            if (env_.species[i] == id) {
                n.kind = Node::SPEC;
                n.idx = (int)i;
                n.type = env_.deterministic ? LT_F32 : LT_I32;
                return add(n);
            }
        }
        if (id == "GridX" || id == "GridY") {
            n.kind = (id == "GridX") ? Node::GX : Node::GY;
            n.type = LT_I32;
            return add(n);
        }
        if (id == "_h0") {
            n.kind = Node::H0;
            n.type = LT_F32;
            return add(n);
        }
        if (id.compare(0, 5, "_mask") == 0 && id.size() > 5) {
            n.kind = Node::MASK;
            n.idx = std::atoi(id.c_str() + 5);
            n.type = LT_F32;
            return add(n);
        }
        auto p = env_.params.find(id);
        if (p != env_.params.end()) {
            n.kind = Node::NUM;
            n.type = LT_F32;
            n.value = (double)p->second;
            return add(n);
        }
        fail("unknown identifier " + id);
    }

    std::string s_;
    const LawEnv &env_;
    size_t p_ = 0;
};

bool is_const(const std::vector<Node> &N, int i) {
    const Node &n = N[i];
    switch (n.kind) {
        case Node::NUM: return true;
        case Node::BIN: return is_const(N, n.a) && is_const(N, n.b);
        case Node::NEG: return is_const(N, n.a);
        case Node::CALL: return is_const(N, n.a) && (n.b < 0 || is_const(N, n.b));
        case Node::CMP: return is_const(N, n.a) && is_const(N, n.b);
        case Node::SEL: return is_const(N, n.a) && is_const(N, n.b) && is_const(N, n.c);
        case Node::CVT: return is_const(N, n.a);
        default: return false;
    }
}

class Emitter {
  public:
    explicit Emitter(const std::vector<Node> &N) : N_(N) {}
    std::vector<uint32_t> code;
    int depth = 0, max_depth = 0;

    void emit(int i) {
        const Node &n = N_[i];
        switch (n.kind) {
            case Node::NUM: {
                uint64_t b;
                std::memcpy(&b, &n.value, 8);
                word(LO_CONST, 0);
                code.push_back((uint32_t)b);
                code.push_back((uint32_t)(b >> 32));
                grow();
                break;
            }
            case Node::SPEC: word(LO_SPEC, (uint32_t)n.idx); grow(); break;
            case Node::GX: word(LO_GRIDX, 0); grow(); break;
            case Node::GY: word(LO_GRIDY, 0); grow(); break;
            case Node::H0: word(LO_H0, 0); grow(); break;
            case Node::MASK: word(LO_MASK, (uint32_t)n.idx); grow(); break;
            case Node::STATE: word(LO_STATE, (uint32_t)n.idx); grow(); break;
            case Node::FIELD: word(LO_FIELD, (uint32_t)n.idx); grow(); break;
            case Node::CUR: word(LO_CUR, (uint32_t)n.idx); grow(); break;
            case Node::CDX: word(LO_CDX, (uint32_t)n.idx); grow(); break;
            case Node::CDY: word(LO_CDY, (uint32_t)n.idx); grow(); break;
            case Node::PX: word(LO_PHYSX, 0); grow(); break;
            case Node::PY: word(LO_PHYSY, 0); grow(); break;
            case Node::TIME: word(LO_SIMTIME, 0); grow(); break;
            case Node::CVT: {
                const LawType from = N_[n.a].type;
                emit(n.a);
                // int -> double and float -> double are exact; double/int -> float round
                if (n.type == LT_F32 && from == LT_F64) word(LO_D2F, 0);
                else if (n.type == LT_F32 && from == LT_I32) word(LO_I2F, 0);
                else if (n.type == LT_I32) throw Error(GPGMP_E_INVALID, "conversion to int is not supported");
                break;
            }
            case Node::CMP: {
                const LawType t = promote(N_[n.a].type, N_[n.b].type);
                operand(n.a, t);
                operand(n.b, t);
                const uint32_t base = (t == LT_I32) ? LO_LT_I : (t == LT_F32) ? LO_LT_F : LO_LT_D;
                const uint32_t k = (n.op == '<') ? 0 : (n.op == '>') ? 1 : (n.op == 'l') ? 2 : (n.op == 'g') ? 3
                                 : (n.op == '=') ? 4 : 5;
                word(base + k, 0);
                --depth;
                break;
            }
            case Node::SEL:
                emit(n.a);
                operand(n.b, n.type);
                operand(n.c, n.type);
                word(LO_SEL, 0);
                depth -= 2;
                break;
            case Node::NEG:
                emit(n.a);
                word(n.type == LT_I32 ? LO_NEG_I : n.type == LT_F32 ? LO_NEG_F : LO_NEG_D, 0);
                break;
            case Node::BIN: {
                operand(n.a, n.type);
                operand(n.b, n.type);
                const uint32_t base = (n.type == LT_I32) ? LO_ADD_I : (n.type == LT_F32) ? LO_ADD_F : LO_ADD_D;
                const uint32_t k = (n.op == '+') ? 0 : (n.op == '-') ? 1 : (n.op == '*') ? 2 : 3;
                word(base + k, 0);
                --depth;
                break;
            }
            case Node::CALL:
                if (n.fn == 100) {
                    operand(n.a, n.type);
                    operand(n.b, n.type);
                    word(n.type == LT_F32 ? LO_POW_F : LO_POW_D, 0);
                    --depth;
                } else if (n.fn == 101) {
                    emit(n.a);
                    word(LO_ABS_I, 0);
                } else {
                    operand(n.a, n.type);
                    word(n.type == LT_F32 ? LO_FN_F : LO_FN_D, n.fn);
                }
                break;
        }
    }

  private:
    // a child converted to the operation's type: only int -> float changes the value
    // (int -> double and float -> double are exact on the double stack)
    void operand(int i, LawType t) {
        emit(i);
        if (t == LT_F32 && N_[i].type == LT_I32) word(LO_I2F, 0);
    }
    void word(uint32_t op, uint32_t arg) { code.push_back(op | (arg << 8)); }
    void grow() {
        if (++depth > max_depth) max_depth = depth;
    }
    const std::vector<Node> &N_;
};

// Replaces every constant subtree by its value (evaluated with the same typed operations).
void fold(std::vector<Node> &N, int i) {
    Node &n = N[i];
    if (n.kind == Node::NUM) return;
    if (is_const(N, i)) {
        Emitter e(N);
        e.emit(i);
        e.code.push_back(LO_END);
        const int32_t x[1] = {0};
        const double v = law_run<int32_t, 1>(e.code.data(), x, LawCell{0, 0, 0u, 0.f});
        const LawType t = n.type;
        n = Node();
        n.kind = Node::NUM;
        n.type = t;
        n.value = v;
        return;
    }
    if (n.a >= 0) fold(N, n.a);
    if (n.b >= 0) fold(N, n.b);
    if (n.c >= 0) fold(N, n.c);
}

// Factor of the mass-action shape: species X, or X - m with an int literal 0 <= m <= 255.
bool factor(const std::vector<Node> &N, int i, int &spec, int &off) {
    const Node &n = N[i];
    if (n.kind == Node::SPEC && n.type == LT_I32) {
        spec = n.idx;
        off = 0;
        return true;
    }
    if (n.kind == Node::BIN && n.op == '-' && N[n.a].kind == Node::SPEC && N[n.a].type == LT_I32 &&
        N[n.b].kind == Node::NUM && N[n.b].type == LT_I32 && N[n.b].value >= 0 && N[n.b].value <= 255) {
        spec = N[n.a].idx;
        off = (int)N[n.b].value;
        return true;
    }
    return false;
}

}  // namespace

LawProgram compile_law(const std::string &text, const LawEnv &env) {
    Parser p(text, env);
    const int root = p.parse();
    std::vector<Node> &N = p.nodes;
    fold(N, root);
    Emitter e(N);
    e.emit(root);
    e.code.push_back(LO_END);
    if (e.max_depth > kLawStack)
        throw Error(GPGMP_E_INVALID, "kinetic law \"" + text + "\" needs a deeper stack than " +
                                         std::to_string(kLawStack));
    LawProgram out;
    out.code = std::move(e.code);
    out.type = N[root].type;
    return out;
}

std::vector<LawProgram> compile_method(const std::string &text, const LawEnv &env) {
    const int S = (int)env.species.size();
    std::vector<LawProgram> out((size_t)4 * S);
    for (int s = 0; s < S; ++s) {
        Parser p(text, env);
        std::vector<int> cur;
        p.method(s, cur);
        std::vector<Node> &N = p.nodes;
        for (int f = 0; f < 4; ++f) {
            int root = cur[(size_t)f * S + s];
            if (root < 0) {  // not assigned: the field keeps its current value
                Node n;
                n.kind = Node::CUR;
                n.idx = s * 4 + f;
                n.type = LT_F32;
                N.push_back(n);
                root = (int)N.size() - 1;
            }
            fold(N, root);
            Emitter e(N);
            e.emit(root);
            e.code.push_back(LO_END);
            if (e.max_depth > kLawStack)
                throw Error(GPGMP_E_INVALID, "drift/diffusivity method needs a deeper stack than " +
                                                 std::to_string(kLawStack));
            out[(size_t)f * S + s].code = std::move(e.code);
            out[(size_t)f * S + s].type = N[root].type;
        }
    }
    return out;
}

uint32_t method_dependencies(const std::vector<LawProgram> &progs, int S) {
    uint32_t deps = 0;
    for (size_t i = 0; i < progs.size(); ++i) {
        const LawProgram &p = progs[i];
        // a field the method does not assign keeps its value: that program reads only itself
        const uint32_t f = (uint32_t)(i / (size_t)S), s = (uint32_t)(i % (size_t)S);
        if (p.code.size() == 2 && p.code[0] == (LO_CUR | ((s * 4 + f) << 8)) && p.code[1] == LO_END) continue;
        for (size_t pc = 0; pc < p.code.size(); ++pc) {
            const uint32_t op = p.code[pc] & 0xffu;
            if (op == LO_END) break;
            if (op == LO_CONST) pc += 2;  // the double's two words
            if (op == LO_STATE || op == LO_CDX || op == LO_CDY) deps |= kMethodReadsState;
            if (op == LO_SIMTIME) deps |= kMethodReadsTime;
            if (op == LO_CUR) deps |= kMethodReadsFields;
        }
    }
    return deps;
}

bool mass_action_shape(const std::string &text, const LawEnv &env, MassAction &ma) {
    if (env.deterministic) return false;
    Parser p(text, env);
    const int root = p.parse();
    std::vector<Node> &N = p.nodes;
    fold(N, root);
    // the left-associative product chain e0 * e1 * ... * en
    std::vector<int> chain;
    int cur = root;
    while (N[cur].kind == Node::BIN && N[cur].op == '*') {
        chain.push_back(N[cur].b);
        cur = N[cur].a;
    }
    chain.push_back(cur);
    std::vector<int> elems(chain.rbegin(), chain.rend());
    ma = MassAction();
    size_t first = 0;
    if (N[elems[0]].kind == Node::NUM) {
        first = 1;
        ma.type = N[elems[0]].type;
        ma.value = N[elems[0]].value;
    } else if (elems.size() >= 2 && N[elems[1]].kind == Node::NUM) {
        // X * k * ...: the first product is commutative (exact int, float or double product)
        int s, o;
        if (!factor(N, elems[0], s, o)) return false;
        ma.type = N[elems[1]].type;
        ma.value = N[elems[1]].value;
        ma.spec.push_back(s);
        ma.off.push_back(o);
        ma.commuted = true;
        first = 2;
    } else {
        ma.type = LT_I32;  // no constant: an int product, the same as 1 * X * ...
        ma.value = 1.0;
    }
    for (size_t k = first; k < elems.size(); ++k) {
        int s, o;
        if (!factor(N, elems[k], s, o)) return false;
        ma.spec.push_back(s);
        ma.off.push_back(o);
    }
    return true;
}

}  // namespace gpgmp

namespace gpgmp {

std::string species_id(const Model &m, int s) {
    if (s < (int)m.species_ids.size() && !m.species_ids[s].empty()) return m.species_ids[s];
    return "S" + std::to_string(s);
}

// `const Real <id> = <value>;` prints the parameter with the stream's default 6 significant
// digits (Solver.cpp:834-841); the float is the printed literal's value.
static float printed_real(float v) {
    char buf[64];
    std::snprintf(buf, sizeof buf, "%g", (double)v);
    return (float)std::strtod(buf, nullptr);
}

LawEnv law_env(const Model &m, bool deterministic) {
    LawEnv env;
    env.deterministic = deterministic;
    for (int s = 0; s < m.desc.num_species; ++s) env.species.push_back(species_id(m, s));
    // header constants (Solver.cpp:494-505)
    env.params["PhysicalCellWidth"] = m.cell_width;
    env.params["PhysicalCellHeight"] = m.cell_height;
    env.params["PhysicalModelWidth"] = header_float(m.desc.physical_length);
    env.params["PhysicalModelHeight"] = header_float(m.desc.physical_length);
    env.params["StayProbability"] = m.probs.stay;
    for (const auto &p : m.params) env.params[p.first] = printed_real(p.second);
    return env;
}

// Reaction.cpp:119-162: the generated laws as text -- the rate as `ss << rate` prints it, then
// per reactant " * X" once and, stochastic, " * (X-k)" for its further units, deterministic
// " * X" per unit (species in index order).
static std::string generated_law(const Model &m, int i, bool deterministic) {
    const gpgmp_reaction_desc &rd = m.reactions[i];
    char buf[64];
    std::snprintf(buf, sizeof buf, "%g", (double)rd.rate);
    std::string t = buf;
    for (int s = 0; s < m.desc.num_species; ++s) {
        const int n = rd.reactant_stoich[s];
        for (int k = 0; k < n; ++k) {
            if (k == 0 || deterministic) t += " * " + species_id(m, s);
            else t += " * (" + species_id(m, s) + "-" + std::to_string(k) + ")";
        }
    }
    return t;
}

std::string law_text(const Model &m, int i, bool deterministic) {
    if (i < (int)m.custom_law.size() && !m.custom_law[i].empty()) return m.custom_law[i];
    return generated_law(m, i, deterministic);
}

static void append(std::vector<uint32_t> &code, const LawProgram &p, uint16_t &off, uint8_t &type) {
    GPGMP_REQUIRE(code.size() + p.code.size() <= (size_t)kLawMaxWords, GPGMP_E_CAPACITY, "kinetic laws too long");
    off = (uint16_t)code.size();
    type = (uint8_t)p.type;
    code.insert(code.end(), p.code.begin(), p.code.end());
}

std::vector<uint32_t> build_law_programs(const Model &m, RxTable &T) {
    std::vector<uint32_t> code;
    T.custom = 0;
    T.det_custom = 0;
    bool any_custom = false;
    for (const auto &c : m.custom_law) any_custom = any_custom || !c.empty();
    if (!any_custom) return code;
    const LawEnv env = law_env(m, false);
    for (int i = 0; i < T.R; ++i) {
        if (i >= (int)m.custom_law.size() || m.custom_law[i].empty()) continue;
        const std::string &law = m.custom_law[i];
        MassAction ma;
        // with a mask the chain starts at w: w * X * k would convert X to float first, so a
        // localized law must lead with its constant to be a table entry
        const bool loc = (T.localized >> i) & 1u;
        if (mass_action_shape(law, env, ma) && (int)ma.spec.size() <= kMaxFactors && !(loc && ma.commuted)) {
            T.lit_is_int &= ~(1u << i);
            T.lit_is_float &= ~(1u << i);
            if (ma.type == LT_I32) {
                T.lit_is_int |= 1u << i;
                T.lit_int[i] = (int32_t)ma.value;
            } else if (ma.type == LT_F32) {
                T.lit_is_float |= 1u << i;
                T.lit_flt[i] = (float)ma.value;
            }
            T.lit_dbl[i] = ma.value;
            T.nf[i] = (uint8_t)ma.spec.size();
            for (size_t f = 0; f < ma.spec.size(); ++f) {
                T.fs[i][f] = (uint8_t)ma.spec[f];
                T.fo[i][f] = (uint8_t)ma.off[f];
            }
            continue;
        }
        // Solver.cpp:849-855: `_h[i+1] = _h[0] + [mask *] <law>;` compiled as one expression
        std::string text = "_h0 + ";
        if ((T.localized >> i) & 1u) text += "_mask" + std::to_string(i) + " * ";
        text += law;
        append(code, compile_law(text, env), T.prog_off[i], T.prog_type[i]);
        T.custom |= 1u << i;
    }
    // deterministic right-hand sides (Solver.cpp:918-1010), generated as the reference does
    const LawEnv denv = law_env(m, true);
    for (int s = 0; s < T.S; ++s) {
        std::string rk = "0.", q = "0.", d = "0.";
        for (int i = 0; i < T.R; ++i) {
            const int n = T.net[i][s];
            if (n == 0) continue;
            std::string term = std::to_string(n > 0 ? n : -n) + ". * ";
            if ((T.localized >> i) & 1u) term += "_mask" + std::to_string(i) + " * ";
            term += law_text(m, i, true);
            rk += (n > 0 ? " + " : " - ") + term;
            (n > 0 ? q : d) += " + " + term;
        }
        append(code, compile_law(rk, denv), T.det_rk[s], T.det_rk_t[s]);
        append(code, compile_law(q, denv), T.det_q[s], T.det_q_t[s]);
        append(code, compile_law(d, denv), T.det_d[s], T.det_d_t[s]);
    }
    T.det_custom = 1;
    T.compact = (T.R <= kCompactReactions && !T.localized && !T.custom) ? 1 : 0;
    for (int i = 0; i < T.R; ++i)
        if (T.nf[i] > kCompactFactors) T.compact = 0;
    return code;
}

}  // namespace gpgmp

namespace gpgmp {

// _h[i+1] of the reaction table on the host (the device's add_law for table entries).
static float table_hazard_host(const RxTable &T, int i, float h0, const int32_t *x, uint32_t act) {
    const int nf = T.nf[i];
    const bool loc = (T.localized >> i) & 1u;
    const float w = ((act >> i) & 1u) ? 1.f : 0.f;
    if ((T.lit_is_int >> i) & 1u) {
        if (loc) {
            float v = w * (float)T.lit_int[i];
            for (int f = 0; f < nf; ++f) v = v * (float)(x[T.fs[i][f]] - (int32_t)T.fo[i][f]);
            return h0 + v;
        }
        uint32_t v = (uint32_t)T.lit_int[i];
        for (int f = 0; f < nf; ++f) v *= (uint32_t)(x[T.fs[i][f]] - (int32_t)T.fo[i][f]);
        return h0 + (float)(int32_t)v;
    }
    if ((T.lit_is_float >> i) & 1u) {
        float v = loc ? w * T.lit_flt[i] : T.lit_flt[i];
        for (int f = 0; f < nf; ++f) v = v * (float)(x[T.fs[i][f]] - (int32_t)T.fo[i][f]);
        return h0 + v;
    }
    double v = loc ? (double)w * T.lit_dbl[i] : T.lit_dbl[i];
    for (int f = 0; f < nf; ++f) v = v * (double)(x[T.fs[i][f]] - (int32_t)T.fo[i][f]);
    return (float)((double)h0 + v);
}

float law_hazard_host(const Model &m, int i, const int32_t *x, int gx, int gy, float h0, uint32_t act, int how) {
    RxTable T = m.table;
    const std::vector<uint32_t> code = build_law_programs(m, T);
    if (how == 1 || (how == 0 && ((T.custom >> i) & 1u))) {
        const LawEnv env = law_env(m, false);
        std::string text = "_h0 + ";
        if ((T.localized >> i) & 1u) text += "_mask" + std::to_string(i) + " * ";
        text += law_text(m, i, false);
        const LawProgram p = compile_law(text, env);
        int32_t xs[kMaxSpecies] = {};
        for (int s = 0; s < T.S; ++s) xs[s] = x[s];
        const double v = law_run<int32_t, kMaxSpecies>(p.code.data(), xs, LawCell{gx, gy, act, h0});
        return law_to_float(v, p.type);
    }
    return table_hazard_host(T, i, h0, x, act);
}

double law_rhs_host(const Model &m, int s, int which, const float *y, int gx, int gy, uint32_t act, float *out) {
    RxTable T = m.table;
    const std::vector<uint32_t> code = build_law_programs(m, T);
    GPGMP_REQUIRE(T.det_custom, GPGMP_E_STATE, "the model has no custom laws");
    float ys[kMaxSpecies] = {};
    for (int k = 0; k < T.S; ++k) ys[k] = y[k];
    const uint16_t off = (which == 0) ? T.det_rk[s] : (which == 1) ? T.det_q[s] : T.det_d[s];
    const uint8_t t = (which == 0) ? T.det_rk_t[s] : (which == 1) ? T.det_q_t[s] : T.det_d_t[s];
    const double v = law_run<float, kMaxSpecies>(code.data() + off, ys, LawCell{gx, gy, act, 0.f});
    *out = law_to_float(v, t);
    return v;
}

}  // namespace gpgmp
