// atus_params -- prints what ParameterHandler reads from an XML parameter
// file as JSON, or evaluates a formula with the file's constants:
//   atus_params file.xml
//   atus_params file.xml --eval "<formula>" [x=.. y=.. z=.. t=..]
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <sstream>

#include "ParameterHandler.h"

namespace {

std::string jstr(const std::string& s) {
  std::string o = "\"";
  for (char c : s) {
    if (c == '"' || c == '\\') {
      o += '\\';
      o += c;
    } else if (c == '\n') {
      o += "\\n";
    } else if ((unsigned char)c < 0x20) {
      o += ' ';
    } else {
      o += c;
    }
  }
  return o + "\"";
}

std::string jnum(double v) {
  if (!std::isfinite(v)) return "null";
  char b[64];
  snprintf(b, sizeof b, "%.17g", v);
  return b;
}

// exposes the maps for printing
class Dump : public ParameterHandler {
 public:
  using ParameterHandler::ParameterHandler;
  void print(std::ostream& o) {
    o << "{\"constants\":{";
    bool first = true;
    for (auto& kv : m_map_constants) {
      o << (first ? "" : ",") << jstr(kv.first) << ":" << jnum(kv.second);
      first = false;
    }
    o << "},\"vconstants\":{";
    first = true;
    for (auto& kv : m_map_vconstants) {
      o << (first ? "" : ",") << jstr(kv.first) << ":[";
      for (size_t i = 0; i < kv.second.size(); ++i) o << (i ? "," : "") << jnum(kv.second[i]);
      o << "]";
      first = false;
    }
    o << "},\"algorithm\":{";
    first = true;
    for (auto& kv : m_map_algorithm) {
      o << (first ? "" : ",") << jstr(kv.first) << ":" << jstr(kv.second);
      first = false;
    }
    o << "},\"simulation\":{";
    first = true;
    for (auto& kv : m_map_simulation) {
      o << (first ? "" : ",") << jstr(kv.first) << ":" << jstr(kv.second);
      first = false;
    }
    o << "},\"getters\":{\"dt\":" << jnum(Get_dt()) << ",\"NX\":" << Get_NX() << ",\"NY\":" << Get_NY()
      << ",\"NZ\":" << Get_NZ() << ",\"xMin\":" << jnum(Get_xMin()) << ",\"xMax\":" << jnum(Get_xMax())
      << ",\"epsilon\":" << jnum(Get_epsilon()) << "}";
    o << ",\"sequence\":[";
    for (size_t i = 0; i < m_sequence.size(); ++i) {
      const sequence_item& s = m_sequence[i];
      o << (i ? "," : "") << "{\"name\":" << jstr(s.name) << ",\"content\":" << jstr(s.content) << ",\"dt\":" << jnum(s.dt)
        << ",\"Nk\":" << s.Nk << ",\"comp\":" << s.comp << ",\"output_freq\":" << s.output_freq
        << ",\"rabi_output_freq\":" << s.rabi_output_freq << ",\"pn_freq\":" << s.compute_pn_freq
        << ",\"custom_freq\":" << s.custom_freq << ",\"chirp_mode\":" << s.chirp_mode
        << ",\"no_of_chirps\":" << s.no_of_chirps << ",\"chirp_min\":" << jnum(s.chirp_min)
        << ",\"chirp_max\":" << jnum(s.chirp_max) << ",\"duration\":[";
      for (size_t j = 0; j < s.duration.size(); ++j) o << (j ? "," : "") << jnum(s.duration[j]);
      o << "]}";
    }
    o << "]}" << std::endl;
  }
};

}  // namespace

int main(int argc, char* argv[]) {
  if (argc < 2) {
    std::cerr << "usage: atus_params file.xml [--eval formula [x=.. y=.. z=.. t=..]]\n";
    return EXIT_FAILURE;
  }
  try {
    Dump p(argv[1]);
    if (argc >= 4 && std::string(argv[2]) == "--eval") {
      double vars[4] = {0, 0, 0, 0};
      const char* names[4] = {"x", "y", "z", "t"};
      for (int i = 4; i < argc; ++i) {
        std::string a = argv[i];
        auto eq = a.find('=');
        if (eq == std::string::npos) continue;
        for (int v = 0; v < 4; ++v)
          if (a.substr(0, eq) == names[v]) vars[v] = std::stod(a.substr(eq + 1));
      }
      atus_host::Expr e;
      p.Setup_muParser(e);
      e.SetExpr(argv[3]);
      for (int v = 0; v < 4; ++v) e.DefineVar(names[v], &vars[v]);
      std::cout << jnum(e.Eval()) << std::endl;
    } else {
      p.print(std::cout);
    }
  } catch (std::string& s) {
    std::cout << "ERROR: " << s << std::endl;
    return EXIT_FAILURE;
  }
  return EXIT_SUCCESS;
}
