#include "yaml_lite.hpp"

#include <cstdlib>
#include <fstream>
#include <sstream>

namespace cuibm {

namespace {

struct Line {
    int indent;
    std::string text;  // without indentation and comment
    int number;
};

std::string strip(const std::string &s) {
    size_t a = s.find_first_not_of(" \t\r");
    if (a == std::string::npos) return "";
    size_t b = s.find_last_not_of(" \t\r");
    return s.substr(a, b - a + 1);
}

// remove a trailing comment that is not inside quotes
std::string drop_comment(const std::string &s) {
    bool sq = false, dq = false;
    for (size_t i = 0; i < s.size(); i++) {
        const char c = s[i];
        if (c == '\'' && !dq) sq = !sq;
        else if (c == '"' && !sq) dq = !dq;
        else if (c == '#' && !sq && !dq && (i == 0 || s[i - 1] == ' ' || s[i - 1] == '\t')) return s.substr(0, i);
    }
    return s;
}

std::string unquote(const std::string &s) {
    if (s.size() >= 2 && ((s.front() == '"' && s.back() == '"') || (s.front() == '\'' && s.back() == '\'')))
        return s.substr(1, s.size() - 2);
    return s;
}

YamlNode make_scalar(const std::string &s) {
    YamlNode n;
    n.kind = YamlNode::SCALAR;
    n.scalar = unquote(strip(s));
    return n;
}

YamlNode parse_flow(const std::string &s, int line) {
    // "[a, b, [c, d]]"
    YamlNode n;
    n.kind = YamlNode::SEQ;
    if (s.size() < 2 || s.front() != '[' || s.back() != ']')
        throw YamlError("line " + std::to_string(line) + ": malformed flow sequence " + s);
    std::string body = s.substr(1, s.size() - 2);
    int depth = 0;
    std::string cur;
    for (char c : body) {
        if (c == '[') depth++;
        if (c == ']') depth--;
        if (c == ',' && depth == 0) {
            std::string t = strip(cur);
            if (!t.empty()) n.items.push_back(t[0] == '[' ? parse_flow(t, line) : make_scalar(t));
            cur.clear();
        } else {
            cur += c;
        }
    }
    std::string t = strip(cur);
    if (!t.empty()) n.items.push_back(t[0] == '[' ? parse_flow(t, line) : make_scalar(t));
    return n;
}

YamlNode value_of(const std::string &s, int line) {
    std::string t = strip(s);
    if (!t.empty() && t[0] == '[') return parse_flow(t, line);
    return make_scalar(t);
}

class Parser {
  public:
    explicit Parser(std::vector<Line> lines) : L(std::move(lines)) {}
    YamlNode parse() {
        if (L.empty()) return YamlNode();
        return block(L[0].indent);
    }

  private:
    std::vector<Line> L;
    size_t pos = 0;

    // a block whose lines start at exactly `indent`
    YamlNode block(int indent) {
        if (pos < L.size() && L[pos].text.rfind("- ", 0) == 0) return sequence(indent);
        if (pos < L.size() && L[pos].text == "-") return sequence(indent);
        return mapping(indent);
    }

    YamlNode sequence(int indent) {
        YamlNode n;
        n.kind = YamlNode::SEQ;
        while (pos < L.size() && L[pos].indent == indent &&
               (L[pos].text.rfind("- ", 0) == 0 || L[pos].text == "-")) {
            Line &ln = L[pos];
            std::string rest = ln.text == "-" ? "" : strip(ln.text.substr(2));
            if (rest.empty()) {
                pos++;
                if (pos < L.size() && L[pos].indent > indent) n.items.push_back(block(L[pos].indent));
                else n.items.push_back(YamlNode());
                continue;
            }
            // "- key: value" starts a map whose keys sit at indent+2
            const size_t colon = key_colon(rest);
            if (colon != std::string::npos) {
                const int inner = indent + (int)(ln.text.size() - rest.size());
                ln.indent = inner;
                ln.text = rest;
                n.items.push_back(mapping(inner));
            } else {
                n.items.push_back(value_of(rest, ln.number));
                pos++;
            }
        }
        return n;
    }

    static size_t key_colon(const std::string &s) {
        if (!s.empty() && (s[0] == '[' || s[0] == '"' || s[0] == '\'')) return std::string::npos;
        for (size_t i = 0; i < s.size(); i++)
            if (s[i] == ':' && (i + 1 == s.size() || s[i + 1] == ' ')) return i;
        return std::string::npos;
    }

    YamlNode mapping(int indent) {
        YamlNode n;
        n.kind = YamlNode::MAP;
        while (pos < L.size() && L[pos].indent == indent) {
            const Line ln = L[pos];
            const size_t colon = key_colon(ln.text);
            if (colon == std::string::npos)
                throw YamlError("line " + std::to_string(ln.number) + ": expected 'key: value', got '" + ln.text + "'");
            const std::string key = strip(ln.text.substr(0, colon));
            const std::string rest = strip(ln.text.substr(colon + 1));
            pos++;
            if (!rest.empty()) {
                n.fields.push_back({key, value_of(rest, ln.number)});
            } else if (pos < L.size() && L[pos].indent > indent) {
                n.fields.push_back({key, block(L[pos].indent)});
            } else if (pos < L.size() && L[pos].indent == indent && L[pos].text.rfind("- ", 0) == 0) {
                // a sequence written at the same indentation as its key
                n.fields.push_back({key, sequence(indent)});
            } else {
                n.fields.push_back({key, YamlNode()});
            }
        }
        if (pos < L.size() && L[pos].indent > indent)
            throw YamlError("line " + std::to_string(L[pos].number) + ": unexpected indentation");
        return n;
    }
};

}  // namespace

bool YamlNode::has(const std::string &key) const {
    if (kind != MAP) return false;
    for (auto &f : fields)
        if (f.first == key) return true;
    return false;
}

const YamlNode &YamlNode::operator[](const std::string &key) const {
    if (kind == MAP)
        for (auto &f : fields)
            if (f.first == key) return f.second;
    throw YamlError("missing key '" + key + "'");
}

const YamlNode &YamlNode::operator[](size_t i) const {
    if (kind != SEQ || i >= items.size()) throw YamlError("sequence index " + std::to_string(i) + " out of range");
    return items[i];
}

double YamlNode::as_double() const {
    if (kind != SCALAR) throw YamlError("expected a number");
    char *end = nullptr;
    const double v = std::strtod(scalar.c_str(), &end);
    if (end == scalar.c_str()) throw YamlError("not a number: '" + scalar + "'");
    return v;
}

int YamlNode::as_int() const {
    const double v = as_double();
    return (int)v;
}

bool YamlNode::as_bool() const {
    if (kind != SCALAR) throw YamlError("expected a boolean");
    const std::string &s = scalar;
    if (s == "true" || s == "True" || s == "TRUE" || s == "1" || s == "yes" || s == "on") return true;
    if (s == "false" || s == "False" || s == "FALSE" || s == "0" || s == "no" || s == "off") return false;
    throw YamlError("not a boolean: '" + s + "'");
}

const std::string &YamlNode::as_string() const {
    if (kind != SCALAR) throw YamlError("expected a string");
    return scalar;
}

YamlNode yaml_parse(const std::string &text) {
    std::vector<Line> lines;
    std::istringstream in(text);
    std::string raw;
    int number = 0;
    while (std::getline(in, raw)) {
        number++;
        std::string s = drop_comment(raw);
        if (strip(s).empty()) continue;
        if (strip(s) == "---" || strip(s) == "...") continue;
        int ind = 0;
        while (ind < (int)s.size() && s[ind] == ' ') ind++;
        lines.push_back({ind, strip(s), number});
    }
    Parser p(std::move(lines));
    return p.parse();
}

YamlNode yaml_load(const std::string &path) {
    std::ifstream f(path);
    if (!f) throw YamlError("cannot open " + path);
    std::stringstream ss;
    ss << f.rdbuf();
    try {
        return yaml_parse(ss.str());
    } catch (const YamlError &e) {
        throw YamlError(path + ": " + e.what());
    }
}

}  // namespace cuibm
