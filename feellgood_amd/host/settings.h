// settings.h -- the subset of YAML that FeeLLGood settings files use (block mappings and
// sequences by indentation, flow sequences and mappings, plain and quoted scalars, comments),
// which includes JSON: the reference's own test feeds its settings as JSON on stdin
// (ci-tests/full_test.py:53). No anchors, tags or multi-document streams.
#pragma once

#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace fgh {

struct Node {
    enum Kind { Null, Scalar, Seq, Map } kind = Null;
    std::string text;                                  // Scalar
    std::vector<Node> items;                           // Seq
    std::vector<std::pair<std::string, Node>> fields;  // Map, in document order
    bool quoted = false;

    bool is_null() const { return kind == Null || (kind == Scalar && !quoted && (text.empty() || text == "~" || text == "null")); }
    const Node *get(const std::string &k) const
    {
        if (kind != Map) return nullptr;
        for (auto &f : fields)
            if (f.first == k) return &f.second;
        return nullptr;
    }
};

class SettingsError : public std::runtime_error {
public:
    using std::runtime_error::runtime_error;
};

Node parse_settings(const std::string &text);

// Typed access with the reference's defaults; `where` names the key in error messages.
double as_double(const Node &n, const std::string &where);
long as_long(const Node &n, const std::string &where);
bool as_bool(const Node &n, const std::string &where);
std::string as_string(const Node &n, const std::string &where);
void as_vec3(const Node &n, double out[3], const std::string &where);

}  // namespace fgh
