// Command-line driver of the NMODL lowering used by the library build:
//   nmodlc hh.mod pas.mod ... -o mech_gen.h
// writes the mech_gen.h header of the builtin mechanisms;
//   nmodlc --macro WilsonCowan.mod ... -o macro_gen.h
// the header of the macro step's MIND-role mechanisms.
#include <stdio.h>

#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "nmodl.h"

int main(int argc, char** argv) {
  std::vector<std::pair<std::string, std::string>> sources;
  std::string out;
  bool macro = false;
  for (int k = 1; k < argc; ++k) {
    const std::string a = argv[k];
    if (a == "--macro") {
      macro = true;
      continue;
    }
    if (a == "-o" && k + 1 < argc) {
      out = argv[++k];
      continue;
    }
    std::ifstream f(a, std::ios::binary);
    if (!f) {
      fprintf(stderr, "nmodlc: cannot read %s\n", a.c_str());
      return 1;
    }
    std::stringstream ss;
    ss << f.rdbuf();
    sources.push_back({a, ss.str()});
  }
  if (out.empty() || sources.empty()) {
    fprintf(stderr, "usage: nmodlc [--macro] file.mod ... -o header.h\n");
    return 2;
  }
  try {
    const std::string text =
        macro ? mind::nmodl::macro_header(sources) : mind::nmodl::builtin_header(sources);
    std::ofstream o(out, std::ios::binary);
    o << text;
    if (!o) {
      fprintf(stderr, "nmodlc: cannot write %s\n", out.c_str());
      return 1;
    }
  } catch (const mind::nmodl::ModError& e) {
    fprintf(stderr, "%s\n", e.what());
    return 1;
  }
  return 0;
}
