// Command-line flags in the gflags spelling the reference's CLIs use
// (--name=value or --name value); unknown flags are an error.
#ifndef DTS_CLI_FLAGS_H_
#define DTS_CLI_FLAGS_H_

#include <cstdio>
#include <cstdlib>
#include <map>
#include <string>

namespace dts_cli {

class Flags {
 public:
  // name -> (default value, help)
  void Define(const std::string& name, const std::string& def, const std::string& help) {
    values_[name] = def;
    help_[name] = help;
  }
  // Returns false after printing a message on bad input; true with
  // *help_requested when --help is given.
  bool Parse(int argc, char** argv, bool* help_requested) {
    *help_requested = false;
    for (int i = 1; i < argc; i++) {
      std::string a = argv[i];
      if (a == "--help" || a == "-h") {
        *help_requested = true;
        return true;
      }
      if (a.rfind("--", 0) != 0) {
        std::fprintf(stderr, "unexpected argument %s\n", a.c_str());
        return false;
      }
      a = a.substr(2);
      std::string name = a, value;
      const size_t eq = a.find('=');
      if (eq != std::string::npos) {
        name = a.substr(0, eq);
        value = a.substr(eq + 1);
      } else if (i + 1 < argc) {
        value = argv[++i];
      } else {
        std::fprintf(stderr, "flag --%s needs a value\n", name.c_str());
        return false;
      }
      if (!values_.count(name)) {
        std::fprintf(stderr, "unknown flag --%s\n", name.c_str());
        return false;
      }
      values_[name] = value;
    }
    return true;
  }
  void PrintHelp(const char* prog) const {
    std::fprintf(stderr, "usage: %s [flags]\n", prog);
    for (const auto& kv : help_)
      std::fprintf(stderr, "  --%s (default \"%s\")  %s\n", kv.first.c_str(),
                   values_.at(kv.first).c_str(), kv.second.c_str());
  }
  std::string Get(const std::string& name) const { return values_.at(name); }
  int GetInt(const std::string& name) const { return std::atoi(values_.at(name).c_str()); }
  double GetDouble(const std::string& name) const {
    return std::atof(values_.at(name).c_str());
  }

 private:
  std::map<std::string, std::string> values_, help_;
};

}  // namespace dts_cli

#endif  // DTS_CLI_FLAGS_H_
