// Flat YAML reader for render configs: one `key: value` mapping per line,
// scalars, flow sequences `[a, b, c]` or block sequences (`- a` lines under a
// key), `#` comments, single/double quoted strings. This is the subset the
// reference's configs use (path_tracer/config/*.yaml, core/config.h:51-63).
#pragma once
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

namespace host {

struct YamlValue {
    std::string scalar;               // raw scalar text (unquoted)
    std::vector<std::string> seq;     // sequence items
    bool is_seq = false;
    int line = 0;
};

class FlatYaml {
  public:
    static FlatYaml load_file(const std::string& path);
    static FlatYaml parse(const std::string& text);

    bool has(const std::string& key) const { return map_.count(key) != 0; }
    const std::map<std::string, YamlValue>& entries() const { return map_; }
    const std::vector<std::string>& keys_in_order() const { return order_; }

    int as_int(const std::string& key) const;
    float as_float(const std::string& key) const;
    bool as_bool(const std::string& key) const;
    std::string as_string(const std::string& key) const;
    std::vector<float> as_floats(const std::string& key) const;

  private:
    const YamlValue& at(const std::string& key) const;
    std::map<std::string, YamlValue> map_;
    std::vector<std::string> order_;
};

// scalar conversions with yaml-cpp's accepted spellings
int parse_int(const std::string& s, const std::string& what);
float parse_float(const std::string& s, const std::string& what);
bool parse_bool(const std::string& s, const std::string& what);

}  // namespace host
