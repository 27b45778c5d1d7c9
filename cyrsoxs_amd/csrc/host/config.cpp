// libconfig-subset parser (see config.hpp).
#include "config.hpp"

#include <cctype>
#include <cstdlib>
#include <fstream>
#include <sstream>

namespace cyrhost {

double Setting::as_double() const {
  if (kind == FLOAT) return f;
  if (kind == INT) return (double)i;
  throw ConfigError("setting '" + name + "' is not a number");
}

int64_t Setting::as_int() const {
  if (kind == INT) return i;
  if (kind == FLOAT) return (int64_t)f;
  if (kind == BOOL) return b ? 1 : 0;
  throw ConfigError("setting '" + name + "' is not an integer");
}

bool Setting::as_bool() const {
  if (kind == BOOL) return b;
  if (kind == INT) return i != 0;
  throw ConfigError("setting '" + name + "' is not a boolean");
}

const Setting *Setting::find(const std::string &key) const {
  for (const auto &c : children)
    if (c->name == key) return c.get();
  return nullptr;
}

const Setting &Setting::at(size_t idx) const {
  if (idx >= children.size()) throw ConfigError("index out of range in '" + name + "'");
  return *children[idx];
}

namespace {

class Parser {
 public:
  Parser(const std::string &t, const std::string &origin) : s_(t), origin_(origin) {}

  void parse_settings(Setting &group, bool until_brace) {
    for (;;) {
      skip();
      if (eof()) {
        if (until_brace) error("missing '}'");
        return;
      }
      if (until_brace && peek() == '}') {
        ++p_;
        return;
      }
      auto st = std::make_shared<Setting>();
      st->name = parse_name();
      skip();
      if (eof() || (peek() != '=' && peek() != ':')) error("expected '=' or ':' after '" + st->name + "'");
      ++p_;
      parse_value(*st);
      skip();
      if (!eof() && (peek() == ';' || peek() == ',')) ++p_;
      group.children.push_back(st);
    }
  }

 private:
  const std::string &s_;
  std::string origin_;
  size_t p_ = 0;

  bool eof() const { return p_ >= s_.size(); }
  char peek() const { return s_[p_]; }

  [[noreturn]] void error(const std::string &msg) const {
    size_t line = 1;
    for (size_t i = 0; i < p_ && i < s_.size(); i++)
      if (s_[i] == '\n') line++;
    throw ConfigError("Parse error at " + origin_ + ":" + std::to_string(line) + " - " + msg);
  }

  void skip() {
    while (!eof()) {
      char c = peek();
      if (std::isspace((unsigned char)c)) {
        ++p_;
      } else if (c == '#') {
        while (!eof() && peek() != '\n') ++p_;
      } else if (c == '/' && p_ + 1 < s_.size() && s_[p_ + 1] == '/') {
        while (!eof() && peek() != '\n') ++p_;
      } else if (c == '/' && p_ + 1 < s_.size() && s_[p_ + 1] == '*') {
        size_t e = s_.find("*/", p_ + 2);
        if (e == std::string::npos) error("unterminated comment");
        p_ = e + 2;
      } else {
        break;
      }
    }
  }

  std::string parse_name() {
    size_t b = p_;
    if (eof() || !(std::isalpha((unsigned char)peek()) || peek() == '*' || peek() == '_')) error("expected a name");
    while (!eof() && (std::isalnum((unsigned char)peek()) || peek() == '_' || peek() == '-' || peek() == '*')) ++p_;
    return s_.substr(b, p_ - b);
  }

  void parse_value(Setting &st) {
    skip();
    if (eof()) error("expected a value");
    char c = peek();
    if (c == '[' || c == '(') {
      const char close = c == '[' ? ']' : ')';
      st.kind = c == '[' ? Setting::ARRAY : Setting::LIST;
      ++p_;
      skip();
      if (!eof() && peek() == close) {
        ++p_;
        return;
      }
      for (;;) {
        auto el = std::make_shared<Setting>();
        el->name = st.name + "[" + std::to_string(st.children.size()) + "]";
        parse_value(*el);
        if (st.kind == Setting::ARRAY && (el->kind == Setting::ARRAY || el->kind == Setting::LIST ||
                                          el->kind == Setting::GROUP))
          error("arrays hold scalars only");
        st.children.push_back(el);
        skip();
        if (eof()) error("unterminated array/list");
        if (peek() == ',') {
          ++p_;
          continue;
        }
        if (peek() == close) {
          ++p_;
          return;
        }
        error("expected ',' or closing bracket");
      }
    }
    if (c == '{') {
      ++p_;
      st.kind = Setting::GROUP;
      parse_settings(st, true);
      return;
    }
    if (c == '"') {
      st.kind = Setting::STRING;
      ++p_;
      std::string out;
      while (!eof() && peek() != '"') {
        if (peek() == '\\' && p_ + 1 < s_.size()) {
          ++p_;
          char e = peek();
          out += e == 'n' ? '\n' : e == 't' ? '\t' : e;
        } else {
          out += peek();
        }
        ++p_;
      }
      if (eof()) error("unterminated string");
      ++p_;
      st.s = out;
      return;
    }
    // bare word: boolean or number
    size_t b = p_;
    while (!eof() && !std::isspace((unsigned char)peek()) && peek() != ';' && peek() != ',' && peek() != ']' &&
           peek() != ')' && peek() != '}' && peek() != '#')
      ++p_;
    std::string w = s_.substr(b, p_ - b);
    std::string lw;
    for (char ch : w) lw += (char)std::tolower((unsigned char)ch);
    if (lw == "true" || lw == "false") {
      st.kind = Setting::BOOL;
      st.b = lw == "true";
      return;
    }
    if (w.empty()) error("expected a value");
    // integers (decimal / hex, optional L suffix) or floats
    std::string num = w;
    if (!num.empty() && (num.back() == 'L' || num.back() == 'l')) num.pop_back();
    if (!num.empty() && (num.back() == 'L' || num.back() == 'l')) num.pop_back();
    char *end = nullptr;
    bool is_float = num.find_first_of(".eE") != std::string::npos && num.find("0x") != 0 && num.find("0X") != 0;
    if (!is_float) {
      long long v = std::strtoll(num.c_str(), &end, 0);
      if (end && *end == '\0') {
        st.kind = Setting::INT;
        st.i = v;
        return;
      }
    }
    double d = std::strtod(num.c_str(), &end);
    if (!end || *end != '\0') error("bad value '" + w + "'");
    st.kind = Setting::FLOAT;
    st.f = d;
  }
};

}  // namespace

void Config::read_string(const std::string &text, const std::string &origin) {
  root_ = Setting();
  root_.kind = Setting::GROUP;
  Parser p(text, origin);
  p.parse_settings(root_, false);
}

void Config::read_file(const std::string &path) {
  std::ifstream in(path);
  if (!in) throw ConfigError("Cannot read " + path);
  std::stringstream ss;
  ss << in.rdbuf();
  read_string(ss.str(), path);
}

const Setting *Config::lookup(const std::string &path) const {
  const Setting *cur = &root_;
  size_t b = 0;
  while (b <= path.size()) {
    size_t e = path.find('.', b);
    std::string part = path.substr(b, e == std::string::npos ? std::string::npos : e - b);
    if (part.empty()) return nullptr;
    if (part[0] == '[') {
      size_t idx = std::strtoul(part.c_str() + 1, nullptr, 10);
      if (idx >= cur->children.size()) return nullptr;
      cur = cur->children[idx].get();
    } else {
      cur = cur->find(part);
      if (!cur) return nullptr;
    }
    if (e == std::string::npos) break;
    b = e + 1;
  }
  return cur;
}

}  // namespace cyrhost
