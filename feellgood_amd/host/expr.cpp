// expr.cpp -- compiler and evaluator of the settings' arithmetic expressions (see expr.h).
#include "expr.h"

#include <cctype>
#include <cmath>
#include <cstdlib>
#include <cstring>

namespace fgh {
namespace {

enum Op {
    LIT, VAR, NEG, NOT, ADD, SUB, MUL, DIV, MOD, POW, LT, LE, GT, GE, EQ, NE, AND, OR, COND, CALL1, CALL2, MINV, MAXV
};

enum Fn {
    F_ABS, F_ACOS, F_ACOSH, F_ASIN, F_ASINH, F_ATAN, F_ATANH, F_CBRT, F_CEIL, F_COS, F_COSH, F_EXP, F_EXPM1,
    F_FLOOR, F_LOG, F_LOG1P, F_LOG10, F_LOG2, F_ROUND, F_SIGN, F_SIN, F_SINH, F_SQRT, F_TAN, F_TANH, F_TRUNC,
    F_ATAN2, F_POW, F_HYPOT
};

struct FnName {
    const char *name;
    int id, nargs;  // nargs -1: min/max (any count >= 1)
};
const FnName FNS[] = {
    {"abs", F_ABS, 1},   {"acos", F_ACOS, 1},   {"acosh", F_ACOSH, 1}, {"asin", F_ASIN, 1},   {"asinh", F_ASINH, 1},
    {"atan", F_ATAN, 1}, {"atanh", F_ATANH, 1}, {"cbrt", F_CBRT, 1},   {"ceil", F_CEIL, 1},   {"cos", F_COS, 1},
    {"cosh", F_COSH, 1}, {"exp", F_EXP, 1},     {"expm1", F_EXPM1, 1}, {"floor", F_FLOOR, 1}, {"log", F_LOG, 1},
    {"log1p", F_LOG1P, 1}, {"log10", F_LOG10, 1}, {"log2", F_LOG2, 1}, {"round", F_ROUND, 1}, {"sign", F_SIGN, 1},
    {"sin", F_SIN, 1},   {"sinh", F_SINH, 1},   {"sqrt", F_SQRT, 1},   {"tan", F_TAN, 1},     {"tanh", F_TANH, 1},
    {"trunc", F_TRUNC, 1}, {"atan2", F_ATAN2, 2}, {"pow", F_POW, 2},   {"hypot", F_HYPOT, 2}, {"min", -1, -1},
    {"max", -2, -1},
};

struct Const {
    const char *name;
    double v;
};
const Const CONSTS[] = {{"PI", M_PI},          {"E", M_E},           {"LN2", M_LN2},         {"LN10", M_LN10},
                        {"LOG2E", M_LOG2E},    {"LOG10E", M_LOG10E}, {"SQRT2", M_SQRT2},     {"SQRT1_2", M_SQRT1_2},
                        {"Infinity", INFINITY}, {"NaN", NAN}};

double call1(int fn, double x)
{
    switch (fn) {
    case F_ABS: return std::fabs(x);
    case F_ACOS: return std::acos(x);
    case F_ACOSH: return std::acosh(x);
    case F_ASIN: return std::asin(x);
    case F_ASINH: return std::asinh(x);
    case F_ATAN: return std::atan(x);
    case F_ATANH: return std::atanh(x);
    case F_CBRT: return std::cbrt(x);
    case F_CEIL: return std::ceil(x);
    case F_COS: return std::cos(x);
    case F_COSH: return std::cosh(x);
    case F_EXP: return std::exp(x);
    case F_EXPM1: return std::expm1(x);
    case F_FLOOR: return std::floor(x);
    case F_LOG: return std::log(x);
    case F_LOG1P: return std::log1p(x);
    case F_LOG10: return std::log10(x);
    case F_LOG2: return std::log2(x);
    case F_ROUND: return std::floor(x + 0.5);  // JavaScript rounds halves towards +infinity
    case F_SIGN: return std::isnan(x) ? x : (x > 0) - (x < 0) + 0.0 * x;
    case F_SIN: return std::sin(x);
    case F_SINH: return std::sinh(x);
    case F_SQRT: return std::sqrt(x);
    case F_TAN: return std::tan(x);
    case F_TANH: return std::tanh(x);
    default: return std::trunc(x);
    }
}

double call2(int fn, double x, double y)
{
    switch (fn) {
    case F_ATAN2: return std::atan2(x, y);
    case F_POW: return std::pow(x, y);
    default: return std::hypot(x, y);
    }
}

bool truthy(double v) { return v != 0 && !std::isnan(v); }

class Parser {
public:
    Parser(const std::string &s, const std::vector<std::string> &vars, std::vector<Expr::N> &out)
        : s_(s), vars_(vars), n_(out) {}

    int parse()
    {
        const int r = cond();
        skip();
        if (p_ != s_.size()) fail("unexpected '" + s_.substr(p_, 1) + "'");
        return r;
    }

private:
    const std::string &s_;
    const std::vector<std::string> &vars_;
    std::vector<Expr::N> &n_;
    size_t p_ = 0;

    [[noreturn]] void fail(const std::string &why) const
    {
        throw ExprError("in \"" + s_ + "\" at " + std::to_string(p_ + 1) + ": " + why);
    }
    void skip()
    {
        while (p_ < s_.size() && std::isspace((unsigned char)s_[p_])) p_++;
    }
    bool eat(const char *tok)
    {
        skip();
        const size_t L = std::strlen(tok);
        if (s_.compare(p_, L, tok) != 0) return false;
        // do not split longer operators: "*" is not the start of "**", "<" not of "<=", ...
        if (L == 1 && p_ + 1 < s_.size()) {
            const char c = tok[0], d = s_[p_ + 1];
            if ((c == '*' && d == '*') || ((c == '<' || c == '>' || c == '=' || c == '!') && d == '=') ||
                (c == '&' && d == '&') || (c == '|' && d == '|'))
                return false;
        }
        p_ += L;
        return true;
    }
    int node(int op, int a = -1, int b = -1, int c = -1, double v = 0, int fn = 0)
    {
        Expr::N x;
        x.op = op;
        x.a = a;
        x.b = b;
        x.c = c;
        x.v = v;
        x.fn = fn;
        n_.push_back(x);
        return (int)n_.size() - 1;
    }
    int cond()
    {
        const int c = lor();
        if (eat("?")) {
            const int a = cond();
            if (!eat(":")) fail("expected ':'");
            const int b = cond();
            return node(COND, c, a, b);
        }
        return c;
    }
    int lor()
    {
        int a = land();
        while (eat("||")) a = node(OR, a, land());
        return a;
    }
    int land()
    {
        int a = equality();
        while (eat("&&")) a = node(AND, a, equality());
        return a;
    }
    int equality()
    {
        int a = rel();
        for (;;) {
            if (eat("===") || eat("=="))
                a = node(EQ, a, rel());
            else if (eat("!==") || eat("!="))
                a = node(NE, a, rel());
            else
                return a;
        }
    }
    int rel()
    {
        int a = add();
        for (;;) {
            if (eat("<="))
                a = node(LE, a, add());
            else if (eat(">="))
                a = node(GE, a, add());
            else if (eat("<"))
                a = node(LT, a, add());
            else if (eat(">"))
                a = node(GT, a, add());
            else
                return a;
        }
    }
    int add()
    {
        int a = mul();
        for (;;) {
            if (eat("+"))
                a = node(ADD, a, mul());
            else if (eat("-"))
                a = node(SUB, a, mul());
            else
                return a;
        }
    }
    int mul()
    {
        int a = unary();
        for (;;) {
            if (eat("*"))
                a = node(MUL, a, unary());
            else if (eat("/"))
                a = node(DIV, a, unary());
            else if (eat("%"))
                a = node(MOD, a, unary());
            else
                return a;
        }
    }
    int unary()
    {
        if (eat("-")) return node(NEG, unary());
        if (eat("+")) return unary();
        if (eat("!")) return node(NOT, unary());
        const int a = primary();
        if (eat("**")) return node(POW, a, unary());  // right-associative
        return a;
    }
    int primary()
    {
        skip();
        if (p_ >= s_.size()) fail("unexpected end");
        const char c = s_[p_];
        if (c == '(') {
            p_++;
            const int a = cond();
            if (!eat(")")) fail("expected ')'");
            return a;
        }
        if (std::isdigit((unsigned char)c) || c == '.') {
            const char *b = s_.c_str() + p_;
            char *e = nullptr;
            const double v = std::strtod(b, &e);
            if (e == b) fail("bad number");
            p_ += (size_t)(e - b);
            return node(LIT, -1, -1, -1, v);
        }
        if (std::isalpha((unsigned char)c) || c == '_') {
            std::string id = ident();
            if (id == "Math") {
                if (!eat(".")) fail("expected '.' after Math");
                id = ident();
            } else {
                for (size_t k = 0; k < vars_.size(); k++)
                    if (id == vars_[k]) return node(VAR, -1, -1, -1, (double)k);
            }
            for (const Const &k : CONSTS)
                if (id == k.name) return node(LIT, -1, -1, -1, k.v);
            for (const FnName &f : FNS)
                if (id == f.name) return call(f);
            p_ -= id.size();
            fail("unknown name '" + id + "'");
        }
        fail("unexpected '" + std::string(1, c) + "'");
    }
    std::string ident()
    {
        skip();
        const size_t b = p_;
        while (p_ < s_.size() && (std::isalnum((unsigned char)s_[p_]) || s_[p_] == '_')) p_++;
        if (p_ == b) fail("expected a name");
        return s_.substr(b, p_ - b);
    }
    int call(const FnName &f)
    {
        if (!eat("(")) fail(std::string("expected '(' after ") + f.name);
        std::vector<int> args;
        if (!eat(")")) {
            do args.push_back(cond());
            while (eat(","));
            if (!eat(")")) fail("expected ')'");
        }
        if (f.nargs < 0) {  // min, max: JavaScript gives +inf / -inf for no argument
            if (args.empty()) return node(LIT, -1, -1, -1, f.id == -1 ? INFINITY : -INFINITY);
            int a = args[0];
            for (size_t k = 1; k < args.size(); k++) a = node(f.id == -1 ? MINV : MAXV, a, args[k]);
            return a;
        }
        if ((int)args.size() != f.nargs)
            fail(std::string(f.name) + " takes " + std::to_string(f.nargs) + " argument(s)");
        return f.nargs == 1 ? node(CALL1, args[0], -1, -1, 0, f.id) : node(CALL2, args[0], args[1], -1, 0, f.id);
    }
};

}  // namespace

Expr Expr::compile(const std::string &src, const std::vector<std::string> &vars)
{
    Expr e;
    Parser p(src, vars, e.nodes_);
    e.root_ = p.parse();
    return e;
}

bool Expr::uses_variables() const
{
    for (const N &n : nodes_)
        if (n.op == VAR) return true;
    return false;
}

double Expr::eval(const double *vals) const { return ev(root_, vals); }

double Expr::ev(int i, const double *vals) const
{
    const N &n = nodes_[i];
    switch (n.op) {
    case LIT: return n.v;
    case VAR: return vals[(int)n.v];
    case NEG: return -ev(n.a, vals);
    case NOT: return truthy(ev(n.a, vals)) ? 0.0 : 1.0;
    case ADD: return ev(n.a, vals) + ev(n.b, vals);
    case SUB: return ev(n.a, vals) - ev(n.b, vals);
    case MUL: return ev(n.a, vals) * ev(n.b, vals);
    case DIV: return ev(n.a, vals) / ev(n.b, vals);
    case MOD: return std::fmod(ev(n.a, vals), ev(n.b, vals));
    case POW: return std::pow(ev(n.a, vals), ev(n.b, vals));
    case LT: return ev(n.a, vals) < ev(n.b, vals);
    case LE: return ev(n.a, vals) <= ev(n.b, vals);
    case GT: return ev(n.a, vals) > ev(n.b, vals);
    case GE: return ev(n.a, vals) >= ev(n.b, vals);
    case EQ: return ev(n.a, vals) == ev(n.b, vals);
    case NE: return ev(n.a, vals) != ev(n.b, vals);
    case AND: {
        const double a = ev(n.a, vals);
        return truthy(a) ? ev(n.b, vals) : a;
    }
    case OR: {
        const double a = ev(n.a, vals);
        return truthy(a) ? a : ev(n.b, vals);
    }
    case COND: return truthy(ev(n.a, vals)) ? ev(n.b, vals) : ev(n.c, vals);
    case CALL1: return call1(n.fn, ev(n.a, vals));
    case CALL2: return call2(n.fn, ev(n.a, vals), ev(n.b, vals));
    case MINV: {
        const double a = ev(n.a, vals), b = ev(n.b, vals);
        return std::isnan(a) || std::isnan(b) ? NAN : std::fmin(a, b);
    }
    default: {
        const double a = ev(n.a, vals), b = ev(n.b, vals);
        return std::isnan(a) || std::isnan(b) ? NAN : std::fmax(a, b);
    }
    }
}

}  // namespace fgh
