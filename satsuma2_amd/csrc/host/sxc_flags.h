// sxc_flags.h -- command-line flags with the behaviour of the reference's
// commandLineParser (base/CommandLineParser.h:76-420), so scripts that launch the
// reference's executables launch these unchanged:
//   * arguments come as "-name value"; a value that starts with '-' is not taken (the
//     flag gets an empty value, which means "use the default");
//   * "-h" anywhere, or no argument at all while some flag has no default, prints the help
//     and exits 0; an unknown "-name" prints "Invalid command-line arg" and the help and
//     exits 0; "-print-command-line" echoes the command line;
//   * a required flag that is missing prints "need to specify -name" and exits 255;
//   * ints are read with atoi, doubles with atof, bools are false only for "0"/"false";
//   * the first occurrence of a repeated flag wins.
#pragma once
#include <stdio.h>
#include <stdlib.h>

#include <map>
#include <string>
#include <vector>

namespace sxc_host {

class Flags {
 public:
  Flags(int argc, char** argv, std::string desc) : argc_(argc), argv_(argv), desc_(std::move(desc)) {}

  void add(const std::string& name, const std::string& type, const std::string& help,
           const char* def /* nullptr: required */) {
    specs_.push_back({name, type, help, def ? std::string(def) : std::string(), def != nullptr});
  }

  void parse() {
    bool all_default = true;
    for (const Spec& s : specs_) all_default &= s.has_def;
    if (argc_ == 1 && !all_default) help_exit();
    for (int i = 1; i < argc_; ++i)
      if (std::string(argv_[i]) == "-h") help_exit();
    int i = 1;
    while (i < argc_) {
      const std::string n = argv_[i];
      const std::string v = i < argc_ - 1 ? argv_[i + 1] : "";
      if (!find(n)) {
        if (n == "-print-command-line") {
          for (int k = 0; k < argc_; ++k) printf("%s ", argv_[k]);
          printf("\n");
          if (argc_ == 2) help_exit();
          ++i;
          continue;
        }
        printf("\nInvalid command-line arg: %s\n", n.c_str());
        help_exit();
      }
      if (!v.empty() && v[0] == '-' && !second_token_is_number(v)) {
        vals_.emplace(n, "");
        i += 1;
      } else {
        vals_.emplace(n, v);
        i += 2;
      }
    }
  }

  std::string str(const std::string& name) const { return raw(name, false); }
  int i32(const std::string& name) const { return atoi(raw(name, true).c_str()); }
  double f64(const std::string& name) const { return atof(raw(name, true).c_str()); }
  bool boolean(const std::string& name) const {
    const std::string v = raw(name, false);
    return !(v == "0" || v == "false");
  }

 private:
  struct Spec {
    std::string name, type, help, def;
    bool has_def;
  };

  // A value starting with '-' is still taken when it splits (on blanks) into two or more
  // items whose second is made of digits and dots (base/CommandLineParser.h:222-227).
  static bool second_token_is_number(const std::string& v) {
    std::vector<std::string> t;
    std::string cur;
    for (char c : v + " ") {
      if (c == ' ' || c == '\t') {
        if (!cur.empty()) t.push_back(cur);
        cur.clear();
      } else {
        cur.push_back(c);
      }
    }
    if (t.size() < 2) return false;
    for (char c : t[1])
      if (!((c >= '0' && c <= '9') || c == '.')) return false;
    return true;
  }

  const Spec* find(const std::string& n) const {
    for (const Spec& s : specs_)
      if (s.name == n) return &s;
    return nullptr;
  }

  std::string raw(const std::string& name, bool empty_means_default) const {
    const Spec* s = find(name);
    auto it = vals_.find(name);
    if (it != vals_.end() && !(empty_means_default && it->second.empty())) return it->second;
    if (!s || !s->has_def) {
      printf("need to specify %s\n", name.c_str());
      fflush(stdout);
      exit(-1);
    }
    return s->def;
  }

  [[noreturn]] void help_exit() const {
    printf("\n%s: %s\n\n\nAvailable arguments:\n\n", argv_[0], desc_.c_str());
    for (const Spec& s : specs_) {
      printf("%s<%s> : %s", s.name.c_str(), s.type.c_str(), s.help.c_str());
      if (s.has_def) printf(" (def=%s)", s.def.c_str());
      printf("\n");
    }
    fflush(stdout);
    exit(0);
  }

  int argc_;
  char** argv_;
  std::string desc_;
  std::vector<Spec> specs_;
  std::multimap<std::string, std::string> vals_;
};

}  // namespace sxc_host
