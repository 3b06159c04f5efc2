// config.h -- a self-contained reader for the libconfig grammar JAMS configs
// are written in (the reference links libconfig++ v1.8.1 with AutoConvert,
// ScientificNotation and SemicolonSeparators, helpers/defaults.h:16-19).
//
// Supported: groups { }, lists ( ), arrays [ ], '=' or ':' assignment,
// ';' / ',' separators, # // /* */ comments, integers (decimal, hex, L
// suffix), floats with exponents, booleans, adjacent string concatenation.
// Files and inline strings merge left to right with the reference's patch
// rules (interface/config.cc:145-: groups merge by name, lists and arrays
// merge element by element).
#pragma once

#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace jams {

class ConfigError : public std::runtime_error {
 public:
  using std::runtime_error::runtime_error;
};

class Setting {
 public:
  enum Type { kNone, kGroup, kList, kArray, kInt, kInt64, kFloat, kBool, kString };

  Setting() = default;
  explicit Setting(Type t) : type_(t) {}

  Type type() const { return type_; }
  bool is_group() const { return type_ == kGroup; }
  bool is_list() const { return type_ == kList; }
  bool is_array() const { return type_ == kArray; }
  bool is_aggregate() const { return type_ == kGroup || type_ == kList || type_ == kArray; }
  bool is_number() const { return type_ == kInt || type_ == kInt64 || type_ == kFloat; }
  bool is_string() const { return type_ == kString; }
  bool is_bool() const { return type_ == kBool; }

  int length() const { return (int)children_.size(); }
  const std::string &name() const { return name_; }
  std::string path() const;

  bool exists(const std::string &key) const;
  const Setting &operator[](const std::string &key) const;
  const Setting &operator[](const char *key) const { return (*this)[std::string(key)]; }
  const Setting &operator[](int index) const;
  Setting &at(const std::string &key);
  Setting &at(int index);
  const Setting *find(const std::string &key) const;

  // Value access with libconfig's AutoConvert (int -> float).
  double as_double() const;
  int as_int() const;
  int64_t as_int64() const;
  bool as_bool() const;
  const std::string &as_string() const;


  // Lookup helpers in the reference's style (config_optional/required).
  template <class T> T get(const std::string &key) const;
  template <class T> T get(const std::string &key, const T &fallback) const;

  // construction
  Setting &add(const std::string &name, Type t);
  Setting &append(Type t);
  void set_int(int64_t v, bool wide) { type_ = wide ? kInt64 : kInt; i_ = v; f_ = (double)v; }
  void set_float(double v) { type_ = kFloat; f_ = v; i_ = (int64_t)v; }
  void set_bool(bool v) { type_ = kBool; b_ = v; }
  void set_string(const std::string &s) { type_ = kString; s_ = s; }

  // merge `patch` into this setting (interface/config.cc:145-)
  void merge(const Setting &patch);
  std::string to_string(int indent = 0) const;

 private:
  friend class Config;
  friend class ConfigParser;
  Type type_ = kNone;
  std::string name_;
  const Setting *parent_ = nullptr;
  std::vector<std::unique_ptr<Setting>> children_;
  int64_t i_ = 0;
  double f_ = 0.0;
  bool b_ = false;
  std::string s_;
  void reparent();
  void copy_from(const Setting &o);
};

class Config {
 public:
  Config() : root_(Setting::kGroup) {}
  // Parse text; `origin` names the source in error messages.
  void read_string(const std::string &text, const std::string &origin = "<string>");
  void read_file(const std::string &path);
  // Merge each entry (a file path if it exists, else config text) in order.
  void merge_sources(const std::vector<std::string> &sources);
  const Setting &root() const { return root_; }
  Setting &root() { return root_; }
  const Setting &lookup(const std::string &path) const;
  bool exists(const std::string &path) const;

 private:
  Setting root_;
};

// typed access -------------------------------------------------------------
template <> inline double Setting::get<double>(const std::string &k) const { return (*this)[k].as_double(); }
template <> inline int Setting::get<int>(const std::string &k) const { return (*this)[k].as_int(); }
template <> inline bool Setting::get<bool>(const std::string &k) const { return (*this)[k].as_bool(); }
template <> inline std::string Setting::get<std::string>(const std::string &k) const { return (*this)[k].as_string(); }
template <class T> T Setting::get(const std::string &k, const T &fallback) const {
  return exists(k) ? get<T>(k) : fallback;
}

}  // namespace jams
