// config.cc -- libconfig-grammar reader and merge (see config.h).
#include "config.h"

#include <cctype>
#include <cerrno>
#include <cstdlib>
#include <fstream>
#include <sstream>
#include <sys/stat.h>

namespace jams {

// ---- Setting ----------------------------------------------------------------

std::string Setting::path() const {
  if (!parent_) return name_.empty() ? std::string("<root>") : name_;
  std::string p = parent_->parent_ ? parent_->path() : std::string();
  if (!name_.empty()) return p.empty() ? name_ : p + "." + name_;
  // element of a list/array: find its index
  for (int i = 0; i < parent_->length(); ++i)
    if (parent_->children_[i].get() == this) return p + ".[" + std::to_string(i) + "]";
  return p + ".[?]";
}

const Setting *Setting::find(const std::string &key) const {
  if (type_ != kGroup) return nullptr;
  for (const auto &c : children_)
    if (c->name_ == key) return c.get();
  return nullptr;
}

bool Setting::exists(const std::string &key) const { return find(key) != nullptr; }

const Setting &Setting::operator[](const std::string &key) const {
  const Setting *s = find(key);
  if (!s) throw ConfigError("setting not found: " + (parent_ ? path() + "." : std::string()) + key);
  return *s;
}

const Setting &Setting::operator[](int index) const {
  if (!is_aggregate() || index < 0 || index >= length())
    throw ConfigError("index " + std::to_string(index) + " out of range in setting " + path());
  return *children_[index];
}

Setting &Setting::at(const std::string &key) { return const_cast<Setting &>((*this)[key]); }
Setting &Setting::at(int index) { return const_cast<Setting &>((*this)[index]); }

double Setting::as_double() const {
  if (type_ == kFloat) return f_;
  if (type_ == kInt || type_ == kInt64) return (double)i_;
  throw ConfigError("setting " + path() + " is not a number");
}

int Setting::as_int() const {
  if (type_ == kInt || type_ == kInt64) return (int)i_;
  if (type_ == kFloat) return (int)f_;
  throw ConfigError("setting " + path() + " is not a number");
}

int64_t Setting::as_int64() const {
  if (type_ == kInt || type_ == kInt64) return i_;
  if (type_ == kFloat) return (int64_t)f_;
  throw ConfigError("setting " + path() + " is not a number");
}

bool Setting::as_bool() const {
  if (type_ == kBool) return b_;
  throw ConfigError("setting " + path() + " is not a boolean");
}

const std::string &Setting::as_string() const {
  if (type_ == kString) return s_;
  throw ConfigError("setting " + path() + " is not a string");
}

Setting &Setting::add(const std::string &name, Type t) {
  auto s = std::make_unique<Setting>(t);
  s->name_ = name;
  s->parent_ = this;
  children_.push_back(std::move(s));
  return *children_.back();
}

Setting &Setting::append(Type t) { return add(std::string(), t); }

void Setting::reparent() {
  for (auto &c : children_) {
    c->parent_ = this;
    c->reparent();
  }
}

void Setting::copy_from(const Setting &o) {
  type_ = o.type_;
  i_ = o.i_;
  f_ = o.f_;
  b_ = o.b_;
  s_ = o.s_;
  children_.clear();
  for (const auto &c : o.children_) {
    auto n = std::make_unique<Setting>();
    n->copy_from(*c);
    n->name_ = c->name_;
    n->parent_ = this;
    children_.push_back(std::move(n));
  }
}

void Setting::merge(const Setting &patch) {
  if (patch.type_ == kGroup && type_ == kGroup) {
    for (const auto &pc : patch.children_) {
      Setting *mine = nullptr;
      for (auto &c : children_)
        if (c->name_ == pc->name_) mine = c.get();
      if (mine && mine->is_aggregate() && pc->is_aggregate() && mine->type_ == pc->type_) {
        mine->merge(*pc);
      } else if (mine) {
        mine->copy_from(*pc);
      } else {
        Setting &n = add(pc->name_, pc->type_);
        n.copy_from(*pc);
      }
    }
    return;
  }
  if ((patch.type_ == kList || patch.type_ == kArray) && type_ == patch.type_) {
    for (int i = 0; i < patch.length(); ++i) {
      const Setting &pc = *patch.children_[i];
      if (i < length()) {
        Setting &mine = *children_[i];
        if (mine.is_aggregate() && pc.is_aggregate() && mine.type_ == pc.type_) mine.merge(pc);
        else mine.copy_from(pc);
      } else {
        Setting &n = append(pc.type_);
        n.copy_from(pc);
      }
    }
    return;
  }
  copy_from(patch);
}

std::string Setting::to_string(int indent) const {
  std::ostringstream os;
  const std::string pad(indent, ' ');
  auto value = [&](const Setting &s, int ind) -> std::string { return s.to_string(ind); };
  switch (type_) {
    case kGroup: {
      os << "{\n";
      for (const auto &c : children_) os << pad << "  " << c->name_ << " = " << value(*c, indent + 2) << ";\n";
      os << pad << "}";
      break;
    }
    case kList:
    case kArray: {
      os << (type_ == kList ? "( " : "[ ");
      for (int i = 0; i < length(); ++i) {
        if (i) os << ", ";
        os << value(*children_[i], indent + 2);
      }
      os << (type_ == kList ? " )" : " ]");
      break;
    }
    case kInt: os << i_; break;
    case kInt64: os << i_ << "L"; break;
    case kFloat: {
      char b[64];
      std::snprintf(b, sizeof b, "%.17g", f_);
      std::string t(b);
      if (t.find_first_of(".eEn") == std::string::npos) t += ".0";
      os << t;
      break;
    }
    case kBool: os << (b_ ? "true" : "false"); break;
    case kString: {
      os << '"';
      for (char ch : s_) {
        if (ch == '"' || ch == '\\') os << '\\';
        os << ch;
      }
      os << '"';
      break;
    }
    default: os << "<none>";
  }
  return os.str();
}

// ---- parser -------------------------------------------------------------------

class ConfigParser {
 public:
  ConfigParser(const std::string &text, const std::string &origin) : t_(text), origin_(origin) {}

  void parse_root(Setting &root) {
    skip();
    while (pos_ < t_.size()) {
      parse_setting(root);
      skip();
    }
  }

 private:
  const std::string &t_;
  std::string origin_;
  size_t pos_ = 0;
  int line_ = 1;

  [[noreturn]] void fail(const std::string &msg) {
    throw ConfigError("parse error in " + origin_ + " line " + std::to_string(line_) + ": " + msg);
  }

  char peek() const { return pos_ < t_.size() ? t_[pos_] : '\0'; }
  char get() {
    char c = t_[pos_++];
    if (c == '\n') ++line_;
    return c;
  }

  void skip() {
    for (;;) {
      while (pos_ < t_.size() && std::isspace((unsigned char)t_[pos_])) get();
      if (pos_ >= t_.size()) return;
      if (t_[pos_] == '#' || (t_[pos_] == '/' && pos_ + 1 < t_.size() && t_[pos_ + 1] == '/')) {
        while (pos_ < t_.size() && t_[pos_] != '\n') get();
        continue;
      }
      if (t_[pos_] == '/' && pos_ + 1 < t_.size() && t_[pos_ + 1] == '*') {
        pos_ += 2;
        while (pos_ + 1 < t_.size() && !(t_[pos_] == '*' && t_[pos_ + 1] == '/')) get();
        if (pos_ + 1 >= t_.size()) fail("unterminated comment");
        pos_ += 2;
        continue;
      }
      return;
    }
  }

  static bool name_start(char c) { return std::isalpha((unsigned char)c) || c == '*'; }
  static bool name_char(char c) {
    return std::isalnum((unsigned char)c) || c == '_' || c == '-' || c == '*';
  }

  void parse_setting(Setting &group) {
    if (!name_start(peek())) fail(std::string("expected a setting name, found '") + peek() + "'");
    size_t a = pos_;
    while (name_char(peek())) get();
    std::string name = t_.substr(a, pos_ - a);
    skip();
    if (peek() != '=' && peek() != ':') fail("expected '=' or ':' after " + name);
    get();
    skip();
    if (group.exists(name)) fail("duplicate setting " + name);
    Setting &s = group.add(name, Setting::kNone);
    parse_value(s);
    skip();
    if (peek() == ';' || peek() == ',') get();
  }

  void parse_value(Setting &s) {
    skip();
    char c = peek();
    if (c == '{') {
      get();
      s.type_ = Setting::kGroup;
      skip();
      while (peek() != '}') {
        if (pos_ >= t_.size()) fail("unterminated group");
        parse_setting(s);
        skip();
      }
      get();
      return;
    }
    if (c == '(' || c == '[') {
      const bool list = c == '(';
      const char close = list ? ')' : ']';
      get();
      s.type_ = list ? Setting::kList : Setting::kArray;
      skip();
      while (peek() != close) {
        if (pos_ >= t_.size()) fail("unterminated list/array");
        Setting &e = s.append(Setting::kNone);
        parse_value(e);
        if (!list && e.is_aggregate()) fail("arrays may only hold scalars");
        skip();
        if (peek() == ',') {
          get();
          skip();
        } else if (peek() != close) {
          fail(std::string("expected ',' or '") + close + "'");
        }
      }
      get();
      if (!list && s.length() > 1) {
        // AutoConvert: an array mixing int and float becomes float
        bool any_float = false;
        for (auto &e : s.children_) any_float |= e->type_ == Setting::kFloat;
        for (auto &e : s.children_) {
          if (any_float && (e->type_ == Setting::kInt || e->type_ == Setting::kInt64))
            e->set_float((double)e->i_);
          else if (e->type_ != s.children_[0]->type_ && !(e->is_number() && s.children_[0]->is_number()))
            fail("array elements must have the same type");
        }
      }
      return;
    }
    if (c == '"') {
      std::string out;
      while (peek() == '"') {
        get();
        while (peek() != '"') {
          if (pos_ >= t_.size()) fail("unterminated string");
          char ch = get();
          if (ch == '\\') {
            char e = get();
            switch (e) {
              case 'n': out += '\n'; break;
              case 't': out += '\t'; break;
              case 'r': out += '\r'; break;
              case 'f': out += '\f'; break;
              default: out += e;
            }
          } else {
            out += ch;
          }
        }
        get();
        skip();
      }
      s.set_string(out);
      return;
    }
    // scalar word: bool or number
    size_t a = pos_;
    while (pos_ < t_.size() && (std::isalnum((unsigned char)peek()) || peek() == '.' || peek() == '+' ||
                                peek() == '-' || peek() == '_'))
      get();
    std::string w = t_.substr(a, pos_ - a);
    if (w.empty()) fail(std::string("unexpected character '") + c + "'");
    std::string lw;
    for (char ch : w) lw += (char)std::tolower((unsigned char)ch);
    if (lw == "true" || lw == "false") {
      s.set_bool(lw == "true");
      return;
    }
    bool wide = false;
    std::string num = w;
    if (!num.empty() && (num.back() == 'L' || num.back() == 'l') &&
        num.find_first_of(".eE") == std::string::npos) {
      wide = true;
      num.pop_back();
      if (!num.empty() && (num.back() == 'L' || num.back() == 'l')) num.pop_back();
    }
    const bool hex = num.size() > 2 && num[0] == '0' && (num[1] == 'x' || num[1] == 'X');
    const bool is_float = !hex && num.find_first_of(".eE") != std::string::npos;
    char *end = nullptr;
    errno = 0;
    if (is_float) {
      double v = std::strtod(num.c_str(), &end);
      if (*end != '\0') fail("bad number '" + w + "'");
      s.set_float(v);
    } else {
      long long v = std::strtoll(num.c_str(), &end, hex ? 16 : 10);
      if (*end != '\0' || errno == ERANGE) fail("bad number '" + w + "'");
      if (v > 2147483647LL || v < -2147483648LL) wide = true;
      s.set_int(v, wide);
    }
  }
};

namespace {

bool file_exists(const std::string &p) {
  struct stat st;
  return stat(p.c_str(), &st) == 0 && S_ISREG(st.st_mode);
}

}  // namespace

void Config::read_string(const std::string &text, const std::string &origin) {
  Setting fresh(Setting::kGroup);
  ConfigParser(text, origin).parse_root(fresh);
  root_.copy_from(fresh);
  root_.reparent();
}

void Config::read_file(const std::string &path) {
  std::ifstream f(path);
  if (!f) throw ConfigError("IO error opening config file: " + path);
  std::stringstream ss;
  ss << f.rdbuf();
  read_string(ss.str(), path);
}

void Config::merge_sources(const std::vector<std::string> &sources) {
  for (const auto &src : sources) {
    Config patch;
    if (file_exists(src)) patch.read_file(src);
    else patch.read_string(src, "config string");
    root_.merge(patch.root_);
    root_.reparent();
  }
}

const Setting &Config::lookup(const std::string &path) const {
  const Setting *s = &root_;
  size_t a = 0;
  while (a <= path.size()) {
    size_t b = path.find('.', a);
    std::string part = path.substr(a, b == std::string::npos ? std::string::npos : b - a);
    if (!part.empty()) {
      if (part.front() == '[' && part.back() == ']') s = &(*s)[std::stoi(part.substr(1, part.size() - 2))];
      else s = &(*s)[part];
    }
    if (b == std::string::npos) break;
    a = b + 1;
  }
  return *s;
}

bool Config::exists(const std::string &path) const {
  try {
    lookup(path);
    return true;
  } catch (const ConfigError &) {
    return false;
  }
}

}  // namespace jams
