// Reader for the reference's input_file.prm (deal.II ParameterHandler text
// format: `subsection <name>` ... `set <key> = <value>` ... `end`, `#`
// comments) and the mapping of its entries onto the C-ABI structs.
//
// Entries and defaults: ParameterReader.cpp:29-208 (declare_parameters).
// Meaning after scaling: Parameters.hpp:117-287 (parse_and_scale_parameters).
// Constants of the initial equilibria: InitialConditions.cpp:5-81.
#pragma once
#include <algorithm>
#include <cctype>
#include <fstream>
#include <map>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/pecs_abi.h"

namespace pecs_host {

class ParameterFile {
 public:
  // Declared entries with the reference's defaults (ParameterReader.cpp:29-208).
  ParameterFile() {
    def("computational", "global refinements", "4");
    def("computational", "local refinements", "0");
    def("computational", "time step size", "0.05");
    def("computational", "end time", "10");
    def("computational", "end time 2", "20");
    def("computational", "time stamps", "50");
    def("computational", "restart status", "false");
    def("physical", "applied bias", "0.0");
    def("physical", "built in bias", "0.41");
    def("physical", "schottky bias", "0.0");
    def("physical", "illumination status", "false");
    def("physical", "insulated", "false");
    def("physical", "schottky status", "false");
    def("physical", "characteristic length", "1.0e-4");
    def("physical", "characteristic density", "1.0e16");
    def("physical", "characteristic time", "1.0e-12");
    def("physical", "intrinsic density", "2.564e9");
    def("physical", "semiconductor permittivity", "11.9");
    def("physical", "electrolyte permittivity", "100");
    def("physical", "photon flux", "1.2e17");
    def("physical", "absorption coefficient", "1.74974e5");
    def("mesh", "mesh length", "1.0");
    def("mesh", "mesh height", "1");
    def("mesh", "radius one", "0.5");
    def("mesh", "radius two", "0.5");
    def("mesh", "boundary layer", "0.1");
    def("electrons", "mobility", "1350.0");
    def("electrons", "transfer rate", "1e-19");
    def("electrons", "recombination time", "5e-5");
    def("electrons", "recombination velocity", "3e5");
    def("holes", "mobility", "480.0");
    def("holes", "transfer rate", "1e-16");
    def("holes", "recombination time", "5e-5");
    def("holes", "recombination velocity", "2.9e-2");
    def("reductants", "mobility", "1.0");
    def("oxidants", "mobility", "1.0");
  }

  // Parse a file; unknown subsections or keys are errors, as in ParameterHandler.
  void read(const std::string& path) {
    std::ifstream in(path);
    if (!in) throw std::runtime_error("cannot open parameter file " + path);
    std::string line, section;
    int lineno = 0;
    while (std::getline(in, line)) {
      ++lineno;
      const size_t hash = line.find('#');
      if (hash != std::string::npos) line.resize(hash);
      line = trim(line);
      if (line.empty()) continue;
      const std::string where = path + ":" + std::to_string(lineno);
      if (starts(line, "subsection ")) {
        if (!section.empty()) throw std::runtime_error(where + ": nested subsection");
        section = trim(line.substr(11));
        bool known = false;
        for (auto& kv : val_) known |= kv.first.first == section;
        if (!known) throw std::runtime_error(where + ": unknown subsection '" + section + "'");
      } else if (line == "end") {
        if (section.empty()) throw std::runtime_error(where + ": 'end' outside a subsection");
        section.clear();
      } else if (starts(line, "set ")) {
        const size_t eq = line.find('=');
        if (eq == std::string::npos) throw std::runtime_error(where + ": missing '='");
        const std::string key = trim(line.substr(4, eq - 4)), value = trim(line.substr(eq + 1));
        auto it = val_.find({section, key});
        if (it == val_.end()) throw std::runtime_error(where + ": unknown entry '" + key + "' in '" + section + "'");
        it->second = value;
      } else {
        throw std::runtime_error(where + ": cannot parse '" + line + "'");
      }
    }
    if (!section.empty()) throw std::runtime_error(path + ": subsection '" + section + "' not closed");
  }

  const std::string& get(const std::string& s, const std::string& k) const {
    auto it = val_.find({s, k});
    if (it == val_.end()) throw std::runtime_error("undeclared entry " + s + "/" + k);
    return it->second;
  }
  double get_double(const std::string& s, const std::string& k) const {
    const std::string& v = get(s, k);
    size_t pos = 0;
    const double d = std::stod(v, &pos);
    if (trim(v.substr(pos)).size()) throw std::runtime_error(s + "/" + k + ": not a number: " + v);
    return d;
  }
  int get_integer(const std::string& s, const std::string& k) const {
    const double d = get_double(s, k);
    if (d != (double)(int)d) throw std::runtime_error(s + "/" + k + ": not an integer");
    return (int)d;
  }
  bool get_bool(const std::string& s, const std::string& k) const {
    std::string v = get(s, k);
    std::transform(v.begin(), v.end(), v.begin(), [](unsigned char c) { return (char)std::tolower(c); });
    if (v == "true" || v == "1" || v == "yes") return true;
    if (v == "false" || v == "0" || v == "no") return false;
    throw std::runtime_error(s + "/" + k + ": not a bool: " + v);
  }

 private:
  void def(const char* s, const char* k, const char* v) { val_[{s, k}] = v; }
  static bool starts(const std::string& a, const char* b) { return a.rfind(b, 0) == 0; }
  static std::string trim(const std::string& s) {
    size_t a = 0, b = s.size();
    while (a < b && std::isspace((unsigned char)s[a])) ++a;
    while (b > a && std::isspace((unsigned char)s[b - 1])) --b;
    return s.substr(a, b - a);
  }
  std::map<std::pair<std::string, std::string>, std::string> val_;
};

struct RunSettings {
  pecs_mesh mesh{};
  pecs_params params{};
  pecs_scaled scaled{};
  double t_end = 0, t_end_2 = 0;
  double char_length = 1, char_density = 1, char_time = 1;
  int time_stamps = 0, local_refinements = 0;
  bool restart = false;
};

// Parameters::parse_and_scale_parameters + the SolarCellProblem constructor
// (SolarCell.cpp:15-92) + the fixed choices of run_full_system
// (penalty 1 at SolarCell.cpp:1944-1945, implicit Euler on the LDG part,
// equilibria of InitialConditions.cpp, R = 0 as SolarCell.hpp:86-98).
inline RunSettings settings_from(const ParameterFile& f, int degree) {
  RunSettings r;
  pecs_physical ph{};
  ph.char_length = f.get_double("physical", "characteristic length");
  ph.char_time = f.get_double("physical", "characteristic time");
  ph.char_density = f.get_double("physical", "characteristic density");
  r.char_length = ph.char_length;
  r.char_time = ph.char_time;
  r.char_density = ph.char_density;
  ph.electron_mobility = f.get_double("electrons", "mobility");
  ph.hole_mobility = f.get_double("holes", "mobility");
  ph.reductant_mobility = f.get_double("reductants", "mobility");
  ph.oxidant_mobility = f.get_double("oxidants", "mobility");
  ph.electron_recombo_t = f.get_double("electrons", "recombination time");
  ph.hole_recombo_t = f.get_double("holes", "recombination time");
  ph.electron_recombo_v = f.get_double("electrons", "recombination velocity");
  ph.hole_recombo_v = f.get_double("holes", "recombination velocity");
  ph.k_et = f.get_double("electrons", "transfer rate");
  ph.k_ht = f.get_double("holes", "transfer rate");
  ph.intrinsic_density = f.get_double("physical", "intrinsic density");
  ph.photon_flux = f.get_double("physical", "photon flux");
  ph.absorption_coeff = f.get_double("physical", "absorption coefficient");
  ph.applied_bias = f.get_double("physical", "applied bias");
  ph.built_in_bias = f.get_double("physical", "built in bias");
  ph.schottky_bias = f.get_double("physical", "schottky bias");
  if (pecs_scale_parameters(&ph, &r.scaled) != 0) throw std::runtime_error(pecs_global_error());

  r.mesh.r1 = f.get_double("mesh", "radius one");
  r.mesh.r2 = f.get_double("mesh", "radius two");
  r.mesh.H = f.get_double("mesh", "mesh height");
  r.mesh.L = f.get_double("mesh", "mesh length");
  r.mesh.n_refine = f.get_integer("computational", "global refinements");
  r.mesh.insulated = f.get_bool("physical", "insulated");
  r.mesh.schottky = f.get_bool("physical", "schottky status");
  r.local_refinements = f.get_integer("computational", "local refinements");
  r.mesh.boundary_layer = f.get_double("mesh", "boundary layer");
  r.mesh.local_refine = r.local_refinements > 0 ? r.local_refinements : 0;
  r.t_end = f.get_double("computational", "end time");
  r.t_end_2 = f.get_double("computational", "end time 2");
  r.time_stamps = f.get_integer("computational", "time stamps");
  r.restart = f.get_bool("computational", "restart status");

  pecs_params& p = r.params;
  p.degree = degree;
  p.dt = f.get_double("computational", "time step size");
  p.theta = 1.0;
  p.penalty = 1.0;
  for (int i = 0; i < 4; ++i) p.mu[i] = r.scaled.mobility[i];
  p.eps_s = f.get_double("physical", "semiconductor permittivity");
  p.eps_e = f.get_double("physical", "electrolyte permittivity");
  p.lambda2 = r.scaled.debye_length;
  p.k_et = r.scaled.k_et;
  p.k_ht = r.scaled.k_ht;
  p.v_rec_n = r.scaled.recombo_v[0];
  p.v_rec_p = r.scaled.recombo_v[1];
  p.n_e = 2.0;   // InitialConditions.cpp:20
  p.p_e = 0.0;   // InitialConditions.cpp:43
  p.r_inf = 30.0;  // InitialConditions.cpp:61
  p.o_inf = 29.0;  // InitialConditions.cpp:79
  p.phi_bi = r.scaled.built_in_bias;
  p.phi_app = r.scaled.applied_bias;
  p.phi_sch = r.scaled.schottky_bias;
  p.gen_mode = f.get_bool("physical", "illumination status") ? 0 : 2;
  p.G0 = r.scaled.photon_flux;
  p.alpha = r.scaled.absorption_coeff;
  p.srh = 0;
  p.n_i = r.scaled.intrinsic_density;
  p.tau_n = r.scaled.recombo_t[0];
  p.tau_p = r.scaled.recombo_t[1];
  p.leaf_cells = 0;
  return r;
}

// Steps taken per output interval by the reference loop (SolarCell.cpp:1978-2082).
inline std::vector<long> reference_step_counts(bool restart, double t_end, double t_end_2, double dt, int stamps) {
  std::vector<long> n(stamps, 0);
  double time = restart ? t_end : 0.0;
  for (int k = 0; k < stamps; ++k) {
    const double stamp = restart ? t_end + ((k + 1) * (t_end_2 - t_end) / stamps) : (k + 1) * t_end / stamps;
    while (time < stamp) {
      time += dt;
      ++n[k];
    }
  }
  return n;
}

}  // namespace pecs_host
