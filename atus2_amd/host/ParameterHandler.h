// ParameterHandler.h -- the XML parameter-file interface of ATUS-2
// (ParameterHandler.h:14-106 of the reference): the same sections, keys,
// defaults and sequence items, read with this project's XML reader; formulae
// go to atus_host::Expr instead of muParser.
#ifndef ATUS_AMD_PARAMETERHANDLER_H
#define ATUS_AMD_PARAMETERHANDLER_H

#include <map>
#include <memory>
#include <string>
#include <vector>

#include "expr.h"
#include "my_structs.h"

namespace atus_host {
struct XmlNode;
}

enum freq { none = 0, each = 1, last = 2, packed = 3 };

/** one element of the SEQUENCE section (ParameterHandler.h:17-35) */
struct sequence_item {
  std::string name;
  std::string content;
  std::vector<double> duration;
  double dt;
  double chirp_min;
  double chirp_max;
  int no_of_chirps;
  int chirp_mode;
  int rabi_output_freq;
  int output_freq;
  int custom_freq;
  int comp;
  int compute_pn_freq;
  int analyze;
  int Nk;
  double time;
};

struct analyze_item {
  std::string name;
  bool content;
  bool separate;
};

struct zernike_polynomial {
  int n;
  int m;
  double alpha;
};

class ParameterHandler {
 public:
  // throws std::string when the file cannot be read or parsed
  explicit ParameterHandler(const std::string filename);
  ~ParameterHandler();

  void Get_Header(generic_header&, bool = true);

  // SIMULATION/<key> text; throws std::string if absent
  std::string Get_simulation(const std::string);
  // CONSTANTS/<key>; throws std::string if absent
  double Get_Constant(const std::string);
  // VCONSTANTS/<key>[p]; throws std::string if absent
  double Get_VConstant(const std::string, const int);
  std::vector<double> Get_VConstant(const std::string);
  bool Get_MomentumStates(std::vector<double>&, const int);

  double Get_t();
  double Get_dt();
  double Get_epsilon();
  double Get_stepsize();
  double Get_xMin();
  double Get_xMax();
  double Get_yMin();
  double Get_yMax();
  double Get_zMin();
  double Get_zMax();
  int Get_NX();
  int Get_NY();
  int Get_NZ();
  int Get_NA();
  int Get_NK();
  int Get_MaxIter();

  // constants of the CONSTANTS section as expression constants, plus the
  // functions Heaviside, rect and sign (ParameterHandler.cpp:298-312)
  void Setup_muParser(atus_host::Expr&);

  std::vector<sequence_item> m_sequence;
  std::vector<analyze_item> m_analyze;
  std::vector<zernike_polynomial> m_zernike;

 protected:
  void populate_constants();
  void populate_vconstants();
  void populate_algorithm();
  void populate_simulation();
  void populate_sequence();
  void populate_zernike();
  void populate_analyze();

  std::unique_ptr<atus_host::XmlNode> m_xml_doc;
  std::map<std::string, int> m_map_freq;
  std::map<std::string, double> m_map_constants;
  std::map<std::string, std::vector<double>> m_map_vconstants;
  std::map<std::string, std::string> m_map_algorithm;
  std::map<std::string, std::string> m_map_simulation;
};

#endif
