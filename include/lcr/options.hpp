// lcr/options.hpp -- a small command-line parser accepting the flag forms that
// Boost.program_options accepts in the reference's tools
// (fine_registration.cpp:44-58, normal_estimation.cpp:40-52): --name value,
// --name=value, -n value, -nvalue, boolean switches, and defaults.  Errors are
// thrown as lcr::options_error with Boost's wording.
#pragma once
#include <map>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace lcr {

struct options_error : std::runtime_error {
  using std::runtime_error::runtime_error;
};

class Options {
 public:
  struct Opt {
    std::string long_name, short_name, help, def;
    bool takes_value = true, has_default = false;
  };

  explicit Options(std::string caption) : caption_(std::move(caption)) {}

  // spec: "long,s" or "long"
  Options& add(const std::string& spec, const std::string& help, bool takes_value = true,
               const std::string& def = "", bool has_default = false) {
    Opt o;
    const auto comma = spec.find(',');
    o.long_name = spec.substr(0, comma);
    if (comma != std::string::npos) o.short_name = spec.substr(comma + 1);
    o.help = help;
    o.takes_value = takes_value;
    o.def = def;
    o.has_default = has_default;
    opts_.push_back(o);
    return *this;
  }

  void parse(int argc, char** argv) {
    for (int i = 1; i < argc; ++i) {
      std::string a = argv[i];
      const Opt* o = nullptr;
      std::string value;
      bool inline_value = false;
      if (a.rfind("--", 0) == 0) {
        std::string name = a.substr(2);
        const auto eq = name.find('=');
        if (eq != std::string::npos) {
          value = name.substr(eq + 1);
          name = name.substr(0, eq);
          inline_value = true;
        }
        o = find_long(name);
        if (!o) throw options_error("unrecognised option '--" + name + "'");
      } else if (a.size() >= 2 && a[0] == '-' && a[1] != '-') {
        o = find_short(a.substr(1, 1));
        if (!o) throw options_error("unrecognised option '" + a.substr(0, 2) + "'");
        if (a.size() > 2) {
          value = a.substr(2);
          inline_value = true;
        }
      } else {
        throw options_error("too many positional options have been specified on the command line");
      }
      if (o->takes_value) {
        if (!inline_value) {
          if (i + 1 >= argc)
            throw options_error("the required argument for option '--" + o->long_name +
                                "' is missing");
          value = argv[++i];
        }
        if (values_.count(o->long_name))
          throw options_error("option '--" + o->long_name + "' cannot be specified more than once");
        values_[o->long_name] = value;
      } else {
        if (inline_value) throw options_error("option '--" + o->long_name + "' does not take any arguments");
        values_[o->long_name] = "";
      }
    }
  }

  size_t count(const std::string& name) const {
    if (values_.count(name)) return 1;
    for (const Opt& o : opts_)
      if (o.long_name == name && o.has_default) return 1;
    return 0;
  }

  std::string str(const std::string& name) const {
    auto it = values_.find(name);
    if (it != values_.end()) return it->second;
    for (const Opt& o : opts_)
      if (o.long_name == name && o.has_default) return o.def;
    throw options_error("option '--" + name + "' is missing");
  }

  template <class T>
  T as(const std::string& name) const {
    const std::string s = str(name);
    std::istringstream ss(s);
    T v{};
    ss >> v;
    if (ss.fail() || !ss.eof())
      throw options_error("the argument ('" + s + "') for option '--" + name + "' is invalid");
    return v;
  }

  std::string help() const {
    std::ostringstream os;
    os << caption_ << ":\n";
    for (const Opt& o : opts_) {
      std::string f = "  ";
      f += o.short_name.empty() ? "--" + o.long_name : "-" + o.short_name + " [ --" + o.long_name + " ]";
      if (o.takes_value) f += o.has_default ? " arg (=" + o.def + ")" : " arg";
      os << f;
      for (size_t p = f.size(); p < 46; ++p) os << ' ';
      os << o.help << "\n";
    }
    return os.str();
  }

 private:
  const Opt* find_long(const std::string& n) const {
    for (const Opt& o : opts_)
      if (o.long_name == n) return &o;
    return nullptr;
  }
  const Opt* find_short(const std::string& n) const {
    for (const Opt& o : opts_)
      if (!o.short_name.empty() && o.short_name == n) return &o;
    return nullptr;
  }
  std::string caption_;
  std::vector<Opt> opts_;
  std::map<std::string, std::string> values_;
};

}  // namespace lcr
