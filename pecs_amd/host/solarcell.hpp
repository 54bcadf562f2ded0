// C++ host classes in the reference's roles, on top of the C-ABI (header only).
//
//   ParameterSpace::Parameters      -> pecs_host::Parameters (scaled values)
//   ChargeCarrierSpace::Carrier     -> pecs_host::Carrier      (Carrier.hpp:38-101)
//   ChargeCarrierSpace::CarrierPair -> pecs_host::CarrierPair  (CarrierPair.hpp:30-110)
//   Poisson::PoissonData            -> pecs_host::PoissonData  (Poisson.hpp:26-84)
//   SOLARCELL::SolarCellProblem     -> pecs_host::SolarCellProblem (SolarCell.hpp:105-757)
//
// The member functions a SolarCellProblem calls per step keep their names and
// meaning (assemble_semiconductor_rhs, assemble_electrolyte_rhs,
// solve_full_system, assemble_Poisson_rhs, solve_Poisson); each forwards to the
// GPU through the C-ABI.  Vectors are std::vector<double> in the reference's
// component-wise layouts, so deal.II Vector<double>::begin() data can be
// copied in and out unchanged.  Errors are thrown as std::runtime_error, the
// way the reference lets deal.II exceptions reach main().
#pragma once
#include <chrono>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/pecs_abi.h"

namespace pecs_host {

inline void check(int rc, pecs_ctx* ctx) {
  if (rc != 0) throw std::runtime_error(std::string("pecs: ") + (ctx ? pecs_last_error(ctx) : pecs_global_error()));
}

// deal.II Vector<double>::block_write / block_read stream format: the size in
// decimal, a newline and '[', the raw values, then ']'.
inline void block_write(std::ostream& out, const std::vector<double>& v) {
  const std::string head = std::to_string(v.size()) + "\n[";
  out.write(head.data(), (std::streamsize)head.size());
  out.write(reinterpret_cast<const char*>(v.data()), (std::streamsize)(v.size() * sizeof(double)));
  const char outro = ']';
  out.write(&outro, 1);
}

inline void block_read(std::istream& in, std::vector<double>& v) {
  size_t sz = 0;
  in >> sz;
  char c = 0;
  in.read(&c, 1);
  if (c != '\n') throw std::runtime_error("block_read: expected newline after the size");
  in.read(&c, 1);
  if (c != '[') throw std::runtime_error("block_read: expected '['");
  v.resize(sz);
  in.read(reinterpret_cast<char*>(v.data()), (std::streamsize)(sz * sizeof(double)));
  in.read(&c, 1);
  if (!in || c != ']') throw std::runtime_error("block_read: truncated vector");
}

struct Parameters {  // scaled values, Parameters.hpp:25-80 after scaling
  pecs_mesh mesh{0.5, 0.5, 1.0, 1.0, 4, 0, 0};
  pecs_params p{};
  int time_steps = 200;
  // characteristic scales (Parameters.hpp:65-67), used only to rescale output
  double char_length = 1.0, char_density = 1.0, char_time = 1.0;
  Parameters() {
    p.degree = 1;
    p.dt = 1e-3;
    p.theta = 1.0;
    p.penalty = 1.0;
    for (double& m : p.mu) m = 1.0;
    p.eps_s = p.eps_e = p.lambda2 = 1.0;
    p.k_et = p.k_ht = 1e-2;
    p.n_e = 1.0;
    p.r_inf = p.o_inf = 1.0;
    p.G0 = p.alpha = 1.0;
    p.n_i = 0.1;
    p.tau_n = p.tau_p = 1.0;
  }
};

class SolarCellProblem;

struct Carrier {
  std::string name;
  double charge_number = 0.0;
  double scaled_mobility = 1.0;
  std::vector<double> solution, system_rhs;
  int which = 0;
  pecs_ctx* ctx = nullptr;
  void set_name(const std::string& n) { name = n; }
  // Carrier::solve (Carrier.cpp:34-40): solution = A^{-1} system_rhs
  void solve() { check(pecs_solve(ctx, which, system_rhs.data(), solution.data()), ctx); }
};

struct CarrierPair {
  Carrier carrier_1, carrier_2;
  double penalty = 1.0;
  double material_permittivity = 1.0;
  std::string material_name;
  void set_name(const std::string& n) { material_name = n; }
  // print_dofs / read_dofs (CarrierPair.cpp:89-121): <Name>.dofs per carrier
  void print_dofs(const std::string& dir = ".") const {
    for (const Carrier* c : {&carrier_1, &carrier_2}) {
      std::ofstream out(dir + "/" + c->name + ".dofs", std::ios::binary);
      if (!out) throw std::runtime_error("cannot write " + c->name + ".dofs");
      block_write(out, c->solution);
    }
  }
  void read_dofs(const std::string& dir = ".") {
    for (Carrier* c : {&carrier_1, &carrier_2}) {
      std::ifstream in(dir + "/" + c->name + ".dofs", std::ios::binary);
      if (!in) throw std::runtime_error("cannot read " + c->name + ".dofs");
      std::vector<double> v;
      block_read(in, v);
      if (v.size() != c->solution.size()) throw std::runtime_error(c->name + ".dofs: size mismatch");
      c->solution = v;
    }
  }
};

struct PoissonData {
  std::vector<double> solution, system_rhs;
  pecs_ctx* ctx = nullptr;
  // PoissonData::solve (Poisson.cpp:98-105), Neumann D.n DoFs come back as 0
  void solve() { check(pecs_solve(ctx, PECS_POISSON, system_rhs.data(), solution.data()), ctx); }
};

// Wall-clock sections in the manner of deal.II TimerOutput (SolarCell.cpp:1905-2087).
class SectionTimer {
 public:
  void enter(const std::string& s) { t0_[s] = std::chrono::steady_clock::now(); }
  void leave(const std::string& s) {
    acc_[s] += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0_[s]).count();
    cnt_[s]++;
  }
  void print(FILE* f) const {
    std::fprintf(f, "+---------------------------------+-----------+------------+\n");
    std::fprintf(f, "| Section                         | no. calls |  wall time |\n");
    for (auto& kv : acc_)
      std::fprintf(f, "| %-31s | %9d | %9.4fs |\n", kv.first.c_str(), cnt_.at(kv.first), kv.second);
    std::fprintf(f, "+---------------------------------+-----------+------------+\n");
  }

 private:
  std::map<std::string, std::chrono::steady_clock::time_point> t0_;
  std::map<std::string, double> acc_;
  std::map<std::string, int> cnt_;
};

class SolarCellProblem {
 public:
  SolarCellProblem(const Parameters& prm) : prm_(prm) {
    timer.enter("Make Grids + Assemble + Factor");
    check(pecs_create(&prm_.mesh, &prm_.p, &ctx_), nullptr);
    timer.leave("Make Grids + Assemble + Factor");
    int64_t ns = 0, ne = 0, np = 0;
    check(pecs_sizes(ctx_, &ns, &ne, &np), ctx_);
    // names, charge numbers and mobilities: SolarCell.cpp:53-85
    setup_carrier(electron_hole_pair.carrier_1, "Electrons", -1.0, prm.p.mu[0], PECS_ELECTRONS, ns);
    setup_carrier(electron_hole_pair.carrier_2, "Holes", 1.0, prm.p.mu[1], PECS_HOLES, ns);
    setup_carrier(redox_pair.carrier_1, "Reductants", -1.0, prm.p.mu[2], PECS_REDUCTANTS, ne);
    setup_carrier(redox_pair.carrier_2, "Oxidants", 1.0, prm.p.mu[3], PECS_OXIDANTS, ne);
    electron_hole_pair.set_name("Semiconductor-");
    redox_pair.set_name("Electrolyte-");
    electron_hole_pair.material_permittivity = prm.p.eps_s;
    redox_pair.material_permittivity = prm.p.eps_e;
    Poisson_object.solution.assign(np, 0.0);
    Poisson_object.system_rhs.assign(np, 0.0);
    Poisson_object.ctx = ctx_;
  }
  ~SolarCellProblem() {
    if (ctx_) pecs_destroy(ctx_);
  }
  SolarCellProblem(const SolarCellProblem&) = delete;
  SolarCellProblem& operator=(const SolarCellProblem&) = delete;

  // --- the per-step interface of the reference (host vectors) -------------
  void assemble_semiconductor_rhs() {  // SolarCell.cpp:1037-1071
    check(pecs_rhs_semiconductor(ctx_, electron_hole_pair.carrier_1.solution.data(),
                                 electron_hole_pair.carrier_2.solution.data(), Poisson_object.solution.data(),
                                 redox_pair.carrier_1.solution.data(), redox_pair.carrier_2.solution.data(),
                                 electron_hole_pair.carrier_1.system_rhs.data(),
                                 electron_hole_pair.carrier_2.system_rhs.data()),
          ctx_);
  }
  void assemble_electrolyte_rhs() {  // SolarCell.cpp:1417-1449
    check(pecs_rhs_electrolyte(ctx_, redox_pair.carrier_1.solution.data(), redox_pair.carrier_2.solution.data(),
                               Poisson_object.solution.data(), electron_hole_pair.carrier_1.solution.data(),
                               electron_hole_pair.carrier_2.solution.data(), redox_pair.carrier_1.system_rhs.data(),
                               redox_pair.carrier_2.system_rhs.data()),
          ctx_);
  }
  void solve_full_system() {  // SolarCell.cpp:1758-1782
    electron_hole_pair.carrier_1.solve();
    electron_hole_pair.carrier_2.solve();
    redox_pair.carrier_1.solve();
    redox_pair.carrier_2.solve();
  }
  void assemble_Poisson_rhs() {  // SolarCell.cpp:430-485
    check(pecs_rhs_poisson(ctx_, electron_hole_pair.carrier_1.solution.data(),
                           electron_hole_pair.carrier_2.solution.data(), redox_pair.carrier_1.solution.data(),
                           redox_pair.carrier_2.solution.data(), Poisson_object.system_rhs.data()),
          ctx_);
  }
  void solve_Poisson() { Poisson_object.solve(); }  // SolarCell.cpp:1750-1756

  // One loop body of run_full_system (SolarCell.cpp:2057-2075) through the
  // host-vector interface above.
  void step_host() {
    timer.enter("Assemble semiconductor rhs");
    assemble_semiconductor_rhs();
    timer.leave("Assemble semiconductor rhs");
    timer.enter("Assemble electrolyte rhs");
    assemble_electrolyte_rhs();
    timer.leave("Assemble electrolyte rhs");
    timer.enter("Solve LDG Systems");
    solve_full_system();
    timer.leave("Solve LDG Systems");
    timer.enter("Assemble Poisson rhs");
    assemble_Poisson_rhs();
    timer.leave("Assemble Poisson rhs");
    timer.enter("Solve Poisson system");
    solve_Poisson();
    timer.leave("Solve Poisson system");
  }

  // --- device-resident run (the fast path) --------------------------------
  void initial_state() {  // SolarCell.cpp:1996-2034
    check(pecs_initial_state(ctx_), ctx_);
    pull_state();
  }
  void set_state_from_members() {
    check(pecs_set_state(ctx_, electron_hole_pair.carrier_1.solution.data(),
                         electron_hole_pair.carrier_2.solution.data(), redox_pair.carrier_1.solution.data(),
                         redox_pair.carrier_2.solution.data(), nullptr),
          ctx_);
    pull_state();
  }
  void run_full_system(int nsteps) {  // SolarCell.cpp:2050-2088 loop, device resident
    timer.enter("Time steps (device)");
    check(pecs_step(ctx_, nsteps), ctx_);
    timer.leave("Time steps (device)");
    pull_state();
  }
  double steady_residual() const {
    double r = -1;
    check(pecs_last_residual(ctx_, &r), ctx_);
    return r;
  }
  void pull_state() {
    check(pecs_get_state(ctx_, electron_hole_pair.carrier_1.solution.data(),
                         electron_hole_pair.carrier_2.solution.data(), redox_pair.carrier_1.solution.data(),
                         redox_pair.carrier_2.solution.data(), Poisson_object.solution.data()),
          ctx_);
  }

  // print_results (SolarCell.cpp:1826-1858): Semiconductor-NNN.vtu,
  // Electrolyte-NNN.vtu and Poisson-NNN.vtu in dir from the members' solution
  // vectors (call pull_state first after device-resident steps).
  void print_results(unsigned time_step_number, const std::string& dir = ".") {
    timer.enter("Printing");
    check(pecs_write_vtu(&prm_.mesh, &prm_.p, electron_hole_pair.carrier_1.solution.data(),
                         electron_hole_pair.carrier_2.solution.data(), redox_pair.carrier_1.solution.data(),
                         redox_pair.carrier_2.solution.data(), Poisson_object.solution.data(), (int)time_step_number,
                         dir.c_str(), prm_.char_length, prm_.char_density, prm_.char_time),
          nullptr);
    timer.leave("Printing");
  }

  pecs_ctx* ctx() const { return ctx_; }
  CarrierPair electron_hole_pair, redox_pair;
  PoissonData Poisson_object;
  SectionTimer timer;

 private:
  void setup_carrier(Carrier& c, const char* name, double z, double mu, int which, int64_t n) {
    c.set_name(name);
    c.charge_number = z;
    c.scaled_mobility = mu;
    c.which = which;
    c.ctx = ctx_;
    c.solution.assign(n, 0.0);
    c.system_rhs.assign(n, 0.0);
  }
  Parameters prm_;
  pecs_ctx* ctx_ = nullptr;
};

}  // namespace pecs_host
