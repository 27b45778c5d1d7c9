// Parser for the libconfig subset used by CyRSoXS input files (config.txt and
// Material<N>.txt; include/Input/InputData.h:258-379 reads them through
// libconfig++, which this image does not have).
//
// Grammar accepted:
//   file     := setting*
//   setting  := name ('=' | ':') value (';' | ',')?
//   value    := scalar | '[' scalars ']' | '(' values ')' | '{' setting* '}'
//   scalar   := integer | float | true | false | "string"
// with '#', '//' and '/* */' comments.
#pragma once

#include <cstdint>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace cyrhost {

struct ConfigError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

struct Setting {
  enum Kind { INT, FLOAT, BOOL, STRING, ARRAY, LIST, GROUP } kind = GROUP;
  std::string name;
  int64_t i = 0;
  double f = 0;
  bool b = false;
  std::string s;
  std::vector<std::shared_ptr<Setting>> children;  // array / list / group members (in order)

  bool is_number() const { return kind == INT || kind == FLOAT; }
  double as_double() const;
  int64_t as_int() const;
  bool as_bool() const;
  // group member lookup (nullptr if absent)
  const Setting *find(const std::string &key) const;
  const Setting &at(size_t idx) const;
  size_t size() const { return children.size(); }
};

class Config {
 public:
  void read_file(const std::string &path);
  void read_string(const std::string &text, const std::string &origin = "<string>");
  const Setting &root() const { return root_; }
  // dotted path lookup ("EnergyData0.Energy"); nullptr if absent
  const Setting *lookup(const std::string &path) const;
  bool exists(const std::string &path) const { return lookup(path) != nullptr; }

 private:
  Setting root_;
};

}  // namespace cyrhost
