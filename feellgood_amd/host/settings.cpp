// settings.cpp -- reader of the YAML subset described in settings.h.
#include "settings.h"

#include <cctype>
#include <cerrno>
#include <cstdlib>
#include <sstream>

namespace fgh {
namespace {

std::string trim(const std::string &s)
{
    size_t a = 0, b = s.size();
    while (a < b && std::isspace((unsigned char)s[a])) a++;
    while (b > a && std::isspace((unsigned char)s[b - 1])) b--;
    return s.substr(a, b - a);
}

std::string strip_comment(const std::string &l)
{
    bool sq = false, dq = false;
    for (size_t i = 0; i < l.size(); i++) {
        const char c = l[i];
        if (c == '\'' && !dq) sq = !sq;
        else if (c == '"' && !sq) dq = !dq;
        else if (c == '#' && !sq && !dq && (i == 0 || l[i - 1] == ' ' || l[i - 1] == '\t')) return l.substr(0, i);
    }
    return l;
}

Node scalar(const std::string &t, bool quoted = false)
{
    Node n;
    n.kind = Node::Scalar;
    n.text = t;
    n.quoted = quoted;
    return n;
}

// Flow collections and scalars ("[...]", "{...}", quoted or plain), JSON included.
class Flow {
public:
    Flow(const std::string &s, int line) : s(s), line(line) {}
    Node parse_all()
    {
        Node n = value(false);
        ws();
        if (p != s.size()) fail("unexpected text after a value");
        return n;
    }

private:
    const std::string &s;
    size_t p = 0;
    int line;

    [[noreturn]] void fail(const std::string &what) const
    {
        std::ostringstream o;
        o << "settings, line " << line << ": " << what;
        throw SettingsError(o.str());
    }
    void ws()
    {
        while (p < s.size() && std::isspace((unsigned char)s[p])) p++;
    }
    std::string quoted_string()
    {
        const char q = s[p++];
        std::string out;
        while (p < s.size() && s[p] != q) {
            char c = s[p++];
            if (q == '"' && c == '\\' && p < s.size()) {
                const char e = s[p++];
                switch (e) {
                case 'n': c = '\n'; break;
                case 't': c = '\t'; break;
                case '"': c = '"'; break;
                case '\\': c = '\\'; break;
                case '/': c = '/'; break;
                default: fail(std::string("unsupported escape \\") + e);
                }
            } else if (q == '\'' && c == '\'' && p < s.size() && s[p] == '\'') {
                p++;  // '' inside a single-quoted scalar
            }
            out += c;
        }
        if (p >= s.size()) fail("unterminated string");
        p++;
        return out;
    }
    Node value(bool in_collection)
    {
        ws();
        if (p >= s.size()) return Node();
        if (s[p] == '[') {
            p++;
            Node n;
            n.kind = Node::Seq;
            ws();
            if (p < s.size() && s[p] == ']') {
                p++;
                return n;
            }
            for (;;) {
                n.items.push_back(value(true));
                ws();
                if (p < s.size() && s[p] == ',') {
                    p++;
                    continue;
                }
                if (p < s.size() && s[p] == ']') {
                    p++;
                    return n;
                }
                fail("expected ',' or ']'");
            }
        }
        if (s[p] == '{') {
            p++;
            Node n;
            n.kind = Node::Map;
            ws();
            if (p < s.size() && s[p] == '}') {
                p++;
                return n;
            }
            for (;;) {
                ws();
                std::string key;
                if (p < s.size() && (s[p] == '"' || s[p] == '\'')) key = quoted_string();
                else {
                    const size_t a = p;
                    while (p < s.size() && s[p] != ':' && s[p] != ',' && s[p] != '}') p++;
                    key = trim(s.substr(a, p - a));
                }
                ws();
                if (p >= s.size() || s[p] != ':') fail("expected ':' after key '" + key + "'");
                p++;
                for (auto &f : n.fields)
                    if (f.first == key) fail("duplicate key '" + key + "'");
                n.fields.emplace_back(key, value(true));
                ws();
                if (p < s.size() && s[p] == ',') {
                    p++;
                    continue;
                }
                if (p < s.size() && s[p] == '}') {
                    p++;
                    return n;
                }
                fail("expected ',' or '}'");
            }
        }
        if (s[p] == '"' || s[p] == '\'') return scalar(quoted_string(), true);
        const size_t a = p;
        if (in_collection)
            while (p < s.size() && s[p] != ',' && s[p] != ']' && s[p] != '}') p++;
        else
            p = s.size();
        return scalar(trim(s.substr(a, p - a)));
    }
};

struct Line {
    int indent;
    std::string s;
    int no;
};

bool is_seq_item(const std::string &s) { return s == "-" || (s.size() > 1 && s[0] == '-' && s[1] == ' '); }

// Position of the ':' that ends a mapping key (followed by a space or the end of the line), or npos.
size_t key_colon(const std::string &s)
{
    bool sq = false, dq = false;
    int depth = 0;
    for (size_t i = 0; i < s.size(); i++) {
        const char c = s[i];
        if (c == '\'' && !dq) sq = !sq;
        else if (c == '"' && !sq) dq = !dq;
        else if (!sq && !dq) {
            if (c == '[' || c == '{') depth++;
            else if (c == ']' || c == '}') depth--;
            else if (c == ':' && depth == 0 && (i + 1 == s.size() || s[i + 1] == ' ')) return i;
        }
    }
    return std::string::npos;
}

class Block {
public:
    explicit Block(std::vector<Line> l) : L(std::move(l)) {}
    Node parse_document()
    {
        if (L.empty()) return Node();
        Node n = parse(L[0].indent);
        if (i < L.size()) fail(L[i], "unexpected indentation");
        return n;
    }

private:
    std::vector<Line> L;
    size_t i = 0;

    [[noreturn]] static void fail(const Line &l, const std::string &what)
    {
        std::ostringstream o;
        o << "settings, line " << l.no << ": " << what;
        throw SettingsError(o.str());
    }
    static Node value_of(const std::string &v, int no)
    {
        if (!v.empty() && (v[0] == '[' || v[0] == '{' || v[0] == '"' || v[0] == '\'')) return Flow(v, no).parse_all();
        return scalar(v);
    }
    Node parse(int indent)
    {
        if (i >= L.size() || L[i].indent < indent) return Node();
        return is_seq_item(L[i].s) ? parse_seq(L[i].indent) : parse_map(L[i].indent);
    }
    Node parse_seq(int ind)
    {
        Node n;
        n.kind = Node::Seq;
        while (i < L.size() && L[i].indent == ind && is_seq_item(L[i].s)) {
            const std::string rest = L[i].s == "-" ? "" : trim(L[i].s.substr(2));
            if (rest.empty()) {
                i++;
                n.items.push_back(i < L.size() && L[i].indent > ind ? parse(L[i].indent) : Node());
            } else if (rest[0] != '[' && rest[0] != '{' && rest[0] != '"' && rest[0] != '\'' &&
                       key_colon(rest) != std::string::npos) {
                // "- key: value" starts a mapping item whose further keys sit two columns in
                L[i].indent = ind + 2;
                L[i].s = rest;
                n.items.push_back(parse_map(ind + 2));
            } else {
                n.items.push_back(value_of(rest, L[i].no));
                i++;
            }
        }
        return n;
    }
    Node parse_map(int ind)
    {
        Node n;
        n.kind = Node::Map;
        while (i < L.size() && L[i].indent == ind && !is_seq_item(L[i].s)) {
            const Line &ln = L[i];
            const size_t c = key_colon(ln.s);
            if (c == std::string::npos) fail(ln, "expected 'key: value'");
            std::string key = trim(ln.s.substr(0, c));
            if (key.size() >= 2 && (key[0] == '"' || key[0] == '\'') && key.back() == key[0])
                key = key.substr(1, key.size() - 2);
            const std::string rest = trim(ln.s.substr(c + 1));
            for (auto &f : n.fields)
                if (f.first == key) fail(ln, "duplicate key '" + key + "'");
            i++;
            Node v;
            if (!rest.empty()) v = value_of(rest, ln.no);
            else if (i < L.size() && L[i].indent > ind) v = parse(L[i].indent);
            else if (i < L.size() && L[i].indent == ind && is_seq_item(L[i].s)) v = parse_seq(ind);
            n.fields.emplace_back(key, std::move(v));
        }
        if (i < L.size() && L[i].indent > ind) fail(L[i], "unexpected indentation");
        return n;
    }
};

}  // namespace

Node parse_settings(const std::string &text)
{
    size_t a = 0;
    while (a < text.size() && std::isspace((unsigned char)text[a])) a++;
    if (a < text.size() && (text[a] == '{' || text[a] == '[')) {
        std::string s = text.substr(a);
        return Flow(s, 1).parse_all();
    }
    std::vector<Line> lines;
    std::istringstream in(text);
    std::string raw;
    int no = 0;
    while (std::getline(in, raw)) {
        no++;
        if (!raw.empty() && raw.back() == '\r') raw.pop_back();
        std::string l = strip_comment(raw);
        if (trim(l).empty() || trim(l) == "---") continue;
        int ind = 0;
        while (ind < (int)l.size() && l[ind] == ' ') ind++;
        if (ind < (int)l.size() && l[ind] == '\t') {
            std::ostringstream o;
            o << "settings, line " << no << ": tab in indentation";
            throw SettingsError(o.str());
        }
        std::string body = l.substr(ind);
        while (!body.empty() && std::isspace((unsigned char)body.back())) body.pop_back();
        lines.push_back({ind, body, no});
    }
    return Block(std::move(lines)).parse_document();
}

double as_double(const Node &n, const std::string &where)
{
    if (n.kind != Node::Scalar) throw SettingsError(where + " should be a number");
    const char *s = n.text.c_str();
    char *end = nullptr;
    errno = 0;
    const double v = std::strtod(s, &end);
    if (end == s || *end != '\0' || errno == ERANGE) throw SettingsError(where + " should be a number, not '" + n.text + "'");
    return v;
}

long as_long(const Node &n, const std::string &where)
{
    const double v = as_double(n, where);
    if (v != (double)(long)v) throw SettingsError(where + " should be an integer");
    return (long)v;
}

bool as_bool(const Node &n, const std::string &where)
{
    if (n.kind == Node::Scalar) {
        const std::string &t = n.text;
        if (t == "true" || t == "True" || t == "TRUE" || t == "yes" || t == "on") return true;
        if (t == "false" || t == "False" || t == "FALSE" || t == "no" || t == "off") return false;
    }
    throw SettingsError(where + " should be true or false");
}

std::string as_string(const Node &n, const std::string &where)
{
    if (n.kind != Node::Scalar) throw SettingsError(where + " should be a string");
    return n.text;
}

void as_vec3(const Node &n, double out[3], const std::string &where)
{
    if (n.kind != Node::Seq || n.items.size() != 3) throw SettingsError(where + " should be a list of three numbers");
    for (int k = 0; k < 3; k++) out[k] = as_double(n.items[k], where);
}

}  // namespace fgh
