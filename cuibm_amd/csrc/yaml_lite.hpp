// Minimal YAML reader for cuIBM case files (simParams.yaml, flow.yaml, domain.yaml,
// bodies.yaml). It covers the subset those files use (src/io/parse*.cu read them with
// yaml-cpp 0.2): block sequences and maps nested by indentation, flow sequences
// "[a, b]", plain or quoted scalars, '#' comments and blank lines.
#pragma once
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace cuibm {

struct YamlError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

class YamlNode {
  public:
    enum Kind { NONE, SCALAR, SEQ, MAP };
    Kind kind = NONE;
    std::string scalar;
    std::vector<YamlNode> items;                             // SEQ
    std::vector<std::pair<std::string, YamlNode>> fields;    // MAP, in file order

    bool has(const std::string &key) const;
    const YamlNode &operator[](const std::string &key) const;  // throws if missing
    const YamlNode &operator[](size_t i) const;                // throws if out of range
    size_t size() const { return kind == SEQ ? items.size() : (kind == MAP ? fields.size() : 0); }
    double as_double() const;
    int as_int() const;
    bool as_bool() const;
    const std::string &as_string() const;
};

// parse a whole document from text / from a file
YamlNode yaml_parse(const std::string &text);
YamlNode yaml_load(const std::string &path);

}  // namespace cuibm
