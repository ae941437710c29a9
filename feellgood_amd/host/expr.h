// expr.h -- arithmetic expressions of the settings file, evaluated on the host.
//
// FeeLLGood evaluates the components of initial_magnetization (functions of x, y, z, in metres) and
// of Bext (functions of t, in seconds) as JavaScript with the members of Math copied to the global
// object plus a few hyperbolic helpers (src/expression_parser.cpp:9-33, src/settings.cpp:454-543).
// This is the arithmetic subset of that language, compiled once and evaluated per node or per step:
// numbers, the named variables, + - * / % ** (right-associative), unary - + !, comparisons,
// && || (with JavaScript's value semantics), ?:, parentheses, the Math constants and functions with
// or without the "Math." prefix, and cosh sinh tanh acosh asinh atanh expm1 log1p.
#pragma once

#include <stdexcept>
#include <string>
#include <vector>

namespace fgh {

struct ExprError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

class Expr {
public:
    // Compile src with the given variable names (their values are passed to eval in that order).
    // Throws ExprError with the position and reason on a syntax error or an unknown name.
    static Expr compile(const std::string &src, const std::vector<std::string> &vars);
    double eval(const double *vals) const;
    // Whether the expression reads any variable (false for constants such as "0" or "2*PI").
    bool uses_variables() const;

    struct N {
        int op;
        int a = -1, b = -1, c = -1;  // operand nodes
        double v = 0;                // literal value, or variable index
        int fn = 0;                  // function id
    };

private:
    std::vector<N> nodes_;
    int root_ = -1;
    double ev(int i, const double *vals) const;
};

}  // namespace fgh
