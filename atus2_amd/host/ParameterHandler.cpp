// ParameterHandler.cpp -- see ParameterHandler.h.  Section queries and
// defaults follow ParameterHandler.cpp:49-107, 183-312, 359-563 of the
// reference.
#include "ParameterHandler.h"

#include <algorithm>
#include <cmath>
#include <iostream>
#include <stdexcept>

#include "xml.h"

using atus_host::XmlNode;

namespace {

// comma-separated tokens (the reference splits with strtk::parse(s, ",", v))
std::vector<std::string> split_commas(const std::string& s) {
  std::vector<std::string> out;
  std::string cur;
  for (char c : s) {
    if (c == ',') {
      out.push_back(cur);
      cur.clear();
    } else {
      cur += c;
    }
  }
  out.push_back(cur);
  // strtk drops a trailing empty token produced by a trailing delimiter
  if (out.size() > 1 && out.back().find_first_not_of(" \t\r\n") == std::string::npos) out.pop_back();
  return out;
}

const XmlNode* root_simulation(const XmlNode* doc) {
  for (const XmlNode* e : doc->elements())
    if (e->name == "SIMULATION") return e;
  return nullptr;
}

// elements matched by /SIMULATION//<section>//*, in document order
std::vector<const XmlNode*> section_items(const XmlNode* doc, const std::string& section) {
  std::vector<const XmlNode*> out;
  const XmlNode* sim = root_simulation(doc);
  if (!sim) return out;
  std::vector<const XmlNode*> all;
  sim->descendants(all);
  for (const XmlNode* s : all) {
    if (s->name != section) continue;
    std::vector<const XmlNode*> sub;
    s->descendants(sub);
    for (const XmlNode* e : sub)
      if (std::find(out.begin(), out.end(), e) == out.end()) out.push_back(e);
  }
  return out;
}

double attr_double(const XmlNode* n, const char* key, double def) {
  const std::string* v = n->attr(key);
  if (!v) return def;
  char* e = nullptr;
  double d = std::strtod(v->c_str(), &e);
  return e == v->c_str() ? def : d;
}

int attr_int(const XmlNode* n, const char* key, int def) {
  const std::string* v = n->attr(key);
  if (!v) return def;
  char* e = nullptr;
  long d = std::strtol(v->c_str(), &e, 0);
  return e == v->c_str() ? def : (int)d;
}

std::string attr_string(const XmlNode* n, const char* key, const char* def) {
  const std::string* v = n->attr(key);
  return v ? *v : std::string(def);
}

bool attr_bool(const XmlNode* n, const char* key, bool def) {
  const std::string* v = n->attr(key);
  if (!v || v->empty()) return def;
  char c = (*v)[0];
  return c == '1' || c == 't' || c == 'T' || c == 'y' || c == 'Y';
}

}  // namespace

ParameterHandler::ParameterHandler(const std::string filename) {
  try {
    m_xml_doc = atus_host::xml_parse_file(filename);
  } catch (const std::string& e) {
    std::cerr << "Critical error occured during loading of file " << filename << std::endl;
    throw std::string("Critical error occured during loading of file " + filename + ": " + e);
  }
  m_map_freq["none"] = freq::none;
  m_map_freq["each"] = freq::each;
  m_map_freq["last"] = freq::last;
  m_map_freq["packed"] = freq::packed;
  populate_constants();
  populate_vconstants();
  populate_algorithm();
  populate_simulation();
  populate_sequence();
  populate_analyze();
  populate_zernike();
}

ParameterHandler::~ParameterHandler() = default;

void ParameterHandler::populate_sequence() {
  m_sequence.clear();
  for (const XmlNode* n : section_items(m_xml_doc.get(), "SEQUENCE")) {
    sequence_item item{};
    item.name = n->name;
    item.content = n->child_value();
    item.dt = attr_double(n, "dt", 0.001);
    item.Nk = attr_int(n, "Nk", 100);
    item.comp = attr_int(n, "comp", 0);
    item.output_freq = m_map_freq[attr_string(n, "output_freq", "none")];
    item.rabi_output_freq = m_map_freq[attr_string(n, "rabi_output_freq", "none")];
    item.compute_pn_freq = m_map_freq[attr_string(n, "pn_freq", "last")];
    item.custom_freq = m_map_freq[attr_string(n, "custom_freq", "none")];
    item.analyze = m_map_freq[attr_string(n, "analyze", "none")];
    item.chirp_mode = attr_int(n, "chirp_mode", 0);
    item.no_of_chirps = attr_int(n, "no_of_chirps", 1);
    item.chirp_min = attr_double(n, "chirp_min", 0);
    item.chirp_max = attr_double(n, "chirp_max", 0);
    for (const std::string& tok : split_commas(item.content)) {
      double val;
      try {
        val = std::stod(tok);
      } catch (const std::exception&) {
        val = 0.0;
      }
      item.duration.push_back(val);
    }
    m_sequence.push_back(item);
  }
}

void ParameterHandler::populate_analyze() {
  m_analyze.clear();
  for (const XmlNode* n : section_items(m_xml_doc.get(), "ANALYZE")) {
    analyze_item item{};
    item.name = n->name;
    std::string tmp = n->child_value();
    std::transform(tmp.begin(), tmp.end(), tmp.begin(), ::toupper);
    item.content = tmp == "TRUE";
    item.separate = attr_bool(n, "separate", false);
    m_analyze.push_back(item);
  }
}

void ParameterHandler::populate_zernike() {
  m_zernike.clear();
  for (const XmlNode* n : section_items(m_xml_doc.get(), "ZERNIKE")) {
    std::vector<std::string> vec = split_commas(n->child_value());
    zernike_polynomial z{};
    try {
      z.n = std::stoi(vec.at(0));
      z.m = std::stoi(vec.at(1));
      z.alpha = std::stod(vec.at(2));
    } catch (const std::exception&) {
      z = zernike_polynomial{};
      std::cerr << "Error Parsing xml file: " << n->name << " must contain 3 numbers in section ZERNIKE\n";
    }
    m_zernike.push_back(z);
  }
}

void ParameterHandler::populate_constants() {
  m_map_constants.clear();
  for (const XmlNode* n : section_items(m_xml_doc.get(), "CONSTANTS")) {
    std::string tmp = n->child_value();
    try {
      double val = std::stod(tmp);
      m_map_constants.insert({n->name, val});  // first definition wins, as std::map::insert
    } catch (const std::exception&) {
      std::cerr << "Error Parsing xml file: Unable to convert " << tmp << " to double for element <" << n->name
                << "> in section CONSTANTS\n";
    }
  }
}

void ParameterHandler::populate_vconstants() {
  m_map_vconstants.clear();
  for (const XmlNode* n : section_items(m_xml_doc.get(), "VCONSTANTS")) {
    std::string tmp = n->child_value();
    std::vector<double> data;
    for (const std::string& tok : split_commas(tmp)) {
      try {
        data.push_back(std::stod(tok));
      } catch (const std::exception&) {
        std::cerr << "Error Parsing xml file: Unable to convert " << tmp << " to double for element <" << n->name
                  << "> in section VCONSTANTS\n";
        throw std::string("Error Parsing xml file: element <" + n->name + "> in section VCONSTANTS");
      }
    }
    m_map_vconstants.insert({n->name, data});
  }
}

void ParameterHandler::populate_algorithm() {
  m_map_algorithm.clear();
  for (const XmlNode* n : section_items(m_xml_doc.get(), "ALGORITHM")) m_map_algorithm.insert({n->name, n->child_value()});
}

void ParameterHandler::populate_simulation() {
  m_map_simulation.clear();
  const XmlNode* sim = root_simulation(m_xml_doc.get());
  if (!sim) return;
  for (const XmlNode* e : sim->elements())
    if (e->has_text_child()) m_map_simulation[e->name] = e->child_value();
}

void ParameterHandler::Setup_muParser(atus_host::Expr& mup) {
  mup.ClearConst();
  for (const auto& kv : m_map_constants) mup.DefineConst(kv.first, kv.second);
}

std::string ParameterHandler::Get_simulation(const std::string k) {
  auto it = m_map_simulation.find(k);
  if (it == m_map_simulation.end()) throw std::string("Error: Could not find the key: " + k + " in section SIMULATION.");
  return it->second;
}

double ParameterHandler::Get_Constant(const std::string k) {
  auto it = m_map_constants.find(k);
  if (it == m_map_constants.end()) throw std::string("Error: Could not find the key: " + k + " in section CONSTANT.");
  return it->second;
}

double ParameterHandler::Get_VConstant(const std::string k, const int p) {
  auto it = m_map_vconstants.find(k);
  if (it == m_map_vconstants.end()) throw std::string("Error: Could not find the key " + k + " in section VCONSTANTS.");
  if (p < 0 || (size_t)p >= it->second.size())
    throw std::string("Error: element " + std::to_string(p) + " of " + k + " missing in section VCONSTANTS.");
  return it->second[p];
}

std::vector<double> ParameterHandler::Get_VConstant(const std::string k) {
  auto it = m_map_vconstants.find(k);
  if (it == m_map_vconstants.end()) throw std::string("Error: Could not find the key " + k + " in section VCONSTANTS.");
  return it->second;
}

bool ParameterHandler::Get_MomentumStates(std::vector<double>& vect, const int p) {
  vect.clear();
  auto it = m_map_vconstants.find("momentum_state_" + std::to_string(p));
  if (it == m_map_vconstants.end()) return false;
  vect = it->second;
  return true;
}

namespace {
template <class T, class F>
T algo(const std::map<std::string, std::string>& m, const char* key, T def, F conv) {
  auto it = m.find(key);
  return it == m.end() ? def : conv(it->second);
}
double to_d(const std::string& s) { return std::stod(s); }
int to_i(const std::string& s) { return std::stoi(s); }
}  // namespace

double ParameterHandler::Get_dt() { return algo(m_map_algorithm, "DT", 0.001, to_d); }
double ParameterHandler::Get_t() { return algo(m_map_algorithm, "T", 0.0, to_d); }
double ParameterHandler::Get_epsilon() { return algo(m_map_algorithm, "EPSILON", 1e-5, to_d); }
double ParameterHandler::Get_stepsize() { return algo(m_map_algorithm, "STEPSIZE", 0.001, to_d); }
double ParameterHandler::Get_xMax() { return algo(m_map_algorithm, "XMAX", 10.0, to_d); }
double ParameterHandler::Get_xMin() { return algo(m_map_algorithm, "XMIN", -10.0, to_d); }
double ParameterHandler::Get_yMax() { return algo(m_map_algorithm, "YMAX", 10.0, to_d); }
double ParameterHandler::Get_yMin() { return algo(m_map_algorithm, "YMIN", -10.0, to_d); }
double ParameterHandler::Get_zMax() { return algo(m_map_algorithm, "ZMAX", 10.0, to_d); }
double ParameterHandler::Get_zMin() { return algo(m_map_algorithm, "ZMIN", -10.0, to_d); }
int ParameterHandler::Get_NX() { return algo(m_map_algorithm, "NX", 256, to_i); }
int ParameterHandler::Get_NY() { return algo(m_map_algorithm, "NY", 256, to_i); }
int ParameterHandler::Get_NZ() { return algo(m_map_algorithm, "NZ", 256, to_i); }
int ParameterHandler::Get_NA() { return algo(m_map_algorithm, "NA", 10, to_i); }
int ParameterHandler::Get_NK() { return algo(m_map_algorithm, "NK", 10, to_i); }
int ParameterHandler::Get_MaxIter() { return algo(m_map_algorithm, "MAXITER", 10, to_i); }

// ParameterHandler::Get_Header (ParameterHandler.cpp:488-563)
void ParameterHandler::Get_Header(generic_header& header, bool bcomplex) {
  header = {};
  long long nDims = (long long)std::stod(Get_simulation("DIM"));
  header.nself = sizeof(generic_header);
  header.nDatatyp = bcomplex ? 2 * sizeof(double) : sizeof(double);
  header.bComplex = bcomplex ? 1 : 0;
  header.bAtom = 1;
  header.nDims = nDims;
  header.nDimX = Get_NX();
  header.nDimY = nDims > 1 ? Get_NY() : 1;
  header.nDimZ = nDims > 2 ? Get_NZ() : 1;
  header.xMax = Get_xMax();
  header.xMin = Get_xMin();
  header.dx = std::fabs(header.xMax - header.xMin) / double(header.nDimX);
  header.dkx = 2 * M_PI / std::fabs(header.xMax - header.xMin);
  if (nDims > 1) {
    header.yMax = Get_yMax();
    header.yMin = Get_yMin();
    header.dy = std::fabs(header.yMax - header.yMin) / double(header.nDimY);
    header.dky = 2 * M_PI / std::fabs(header.yMax - header.yMin);
  }
  if (nDims > 2) {
    header.zMax = Get_zMax();
    header.zMin = Get_zMin();
    header.dz = std::fabs(header.zMax - header.zMin) / double(header.nDimZ);
    header.dkz = 2 * M_PI / std::fabs(header.zMax - header.zMin);
  }
  header.dt = 0.001;
  header.nself_and_data = header.nself + (header.nDimX + header.nDimY) * header.nDatatyp;
}
