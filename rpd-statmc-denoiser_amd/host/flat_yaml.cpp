#include "flat_yaml.h"

#include <cctype>
#include <cerrno>
#include <cmath>
#include <cstdlib>
#include <fstream>
#include <limits>
#include <sstream>

namespace host {

namespace {

std::string trim(const std::string& s) {
    size_t a = 0, b = s.size();
    while (a < b && std::isspace(static_cast<unsigned char>(s[a]))) ++a;
    while (b > a && std::isspace(static_cast<unsigned char>(s[b - 1]))) --b;
    return s.substr(a, b - a);
}

// strips a trailing comment that is not inside quotes
std::string strip_comment(const std::string& s) {
    char q = 0;
    for (size_t i = 0; i < s.size(); ++i) {
        const char c = s[i];
        if (q) {
            if (c == q) q = 0;
        } else if (c == '"' || c == '\'') {
            q = c;
        } else if (c == '#' && (i == 0 || std::isspace(static_cast<unsigned char>(s[i - 1])))) {
            return s.substr(0, i);
        }
    }
    return s;
}

std::string unquote(const std::string& s, int line) {
    if (s.size() >= 2 && ((s.front() == '"' && s.back() == '"') || (s.front() == '\'' && s.back() == '\'')))
        return s.substr(1, s.size() - 2);
    if (!s.empty() && (s.front() == '"' || s.front() == '\''))
        throw std::runtime_error("unterminated quoted string at line " + std::to_string(line));
    return s;
}

std::vector<std::string> split_flow(const std::string& body, int line) {
    std::vector<std::string> out;
    std::string cur;
    char q = 0;
    for (char c : body) {
        if (q) {
            cur += c;
            if (c == q) q = 0;
        } else if (c == '"' || c == '\'') {
            q = c;
            cur += c;
        } else if (c == ',') {
            out.push_back(unquote(trim(cur), line));
            cur.clear();
        } else if (c == '[' || c == ']' || c == '{' || c == '}') {
            throw std::runtime_error("nested collections are not supported (line " + std::to_string(line) + ")");
        } else {
            cur += c;
        }
    }
    const std::string last = trim(cur);
    if (!last.empty() || !out.empty()) out.push_back(unquote(last, line));
    return out;
}

}  // namespace

FlatYaml FlatYaml::load_file(const std::string& path) {
    std::ifstream f(path);
    if (!f) throw std::runtime_error("bad file: " + path);
    std::stringstream ss;
    ss << f.rdbuf();
    return parse(ss.str());
}

FlatYaml FlatYaml::parse(const std::string& text) {
    FlatYaml y;
    std::istringstream in(text);
    std::string raw;
    int line = 0;
    std::string open_block;  // key whose block sequence is being read
    while (std::getline(in, raw)) {
        ++line;
        if (!raw.empty() && raw.back() == '\r') raw.pop_back();
        const std::string s = strip_comment(raw);
        const std::string t = trim(s);
        if (t.empty() || t == "---" || t == "...") continue;
        const bool indented = !s.empty() && std::isspace(static_cast<unsigned char>(s[0]));
        if (t[0] == '-' && (t.size() == 1 || t[1] == ' ')) {
            if (open_block.empty()) throw std::runtime_error("sequence item without a key at line " + std::to_string(line));
            YamlValue& v = y.map_[open_block];
            v.is_seq = true;
            v.seq.push_back(unquote(trim(t.substr(1)), line));
            continue;
        }
        if (indented) throw std::runtime_error("nested mappings are not supported (line " + std::to_string(line) + ")");
        const size_t colon = t.find(':');
        if (colon == std::string::npos || colon == 0)
            throw std::runtime_error("expected 'key: value' at line " + std::to_string(line));
        const std::string key = unquote(trim(t.substr(0, colon)), line);
        const std::string val = trim(t.substr(colon + 1));
        if (y.map_.count(key)) throw std::runtime_error("duplicate key '" + key + "' at line " + std::to_string(line));
        YamlValue v;
        v.line = line;
        open_block.clear();
        if (val.empty()) {
            v.is_seq = true;  // block sequence follows (or an empty value)
            open_block = key;
        } else if (val.front() == '[') {
            if (val.back() != ']') throw std::runtime_error("unterminated sequence at line " + std::to_string(line));
            v.is_seq = true;
            v.seq = split_flow(val.substr(1, val.size() - 2), line);
        } else {
            v.scalar = unquote(val, line);
        }
        y.map_[key] = v;
        y.order_.push_back(key);
    }
    return y;
}

const YamlValue& FlatYaml::at(const std::string& key) const {
    auto it = map_.find(key);
    if (it == map_.end()) throw std::runtime_error("missing required key: " + key);
    return it->second;
}

int parse_int(const std::string& s, const std::string& what) {
    errno = 0;
    char* end = nullptr;
    const long v = std::strtol(s.c_str(), &end, 0);
    if (s.empty() || *end != 0 || errno || v < std::numeric_limits<int>::min() || v > std::numeric_limits<int>::max())
        throw std::runtime_error("bad conversion: '" + s + "' is not an integer (" + what + ")");
    return int(v);
}

float parse_float(const std::string& s, const std::string& what) {
    std::string t = s;
    for (auto& c : t) c = char(std::tolower(static_cast<unsigned char>(c)));
    if (t == ".inf" || t == "+.inf") return std::numeric_limits<float>::infinity();
    if (t == "-.inf") return -std::numeric_limits<float>::infinity();
    if (t == ".nan") return std::numeric_limits<float>::quiet_NaN();
    errno = 0;
    char* end = nullptr;
    const float v = std::strtof(s.c_str(), &end);
    if (s.empty() || *end != 0) throw std::runtime_error("bad conversion: '" + s + "' is not a number (" + what + ")");
    return v;
}

bool parse_bool(const std::string& s, const std::string& what) {
    std::string t = s;
    for (auto& c : t) c = char(std::tolower(static_cast<unsigned char>(c)));
    if (t == "true" || t == "yes" || t == "on" || t == "y") return true;
    if (t == "false" || t == "no" || t == "off" || t == "n") return false;
    throw std::runtime_error("bad conversion: '" + s + "' is not a boolean (" + what + ")");
}

int FlatYaml::as_int(const std::string& key) const {
    const YamlValue& v = at(key);
    if (v.is_seq) throw std::runtime_error("bad conversion: " + key + " is a sequence");
    return parse_int(v.scalar, key);
}
float FlatYaml::as_float(const std::string& key) const {
    const YamlValue& v = at(key);
    if (v.is_seq) throw std::runtime_error("bad conversion: " + key + " is a sequence");
    return parse_float(v.scalar, key);
}
bool FlatYaml::as_bool(const std::string& key) const {
    const YamlValue& v = at(key);
    if (v.is_seq) throw std::runtime_error("bad conversion: " + key + " is a sequence");
    return parse_bool(v.scalar, key);
}
std::string FlatYaml::as_string(const std::string& key) const {
    const YamlValue& v = at(key);
    if (v.is_seq) throw std::runtime_error("bad conversion: " + key + " is a sequence");
    return v.scalar;
}
std::vector<float> FlatYaml::as_floats(const std::string& key) const {
    const YamlValue& v = at(key);
    if (!v.is_seq) throw std::runtime_error("bad conversion: " + key + " is not a sequence");
    std::vector<float> out;
    for (const auto& s : v.seq) out.push_back(parse_float(s, key));
    return out;
}

}  // namespace host
