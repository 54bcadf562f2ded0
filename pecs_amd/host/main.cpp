// pecs_run: command-line driver in the role of source/main/main.cpp, running
// SolarCellProblem::run_full_system (SolarCell.cpp:1898-2095) on the GPU.
//
//   pecs_run [input_file.prm] [--degree K] [--steps N] [--steady-tol TOL]
//            [--check-every M] [--out DIR] [--ignore-local-refinement] [--host-loop]
//            [--no-vtu]
//
// Without --steps the step count is the reference's: `time < timeStamps[k]`
// with `time += delta_t` in double arithmetic, replayed on the host so the
// device runs exactly as many steps as the reference loop would.  With
// --steady-tol the run also stops once the relative steady residual
// ||x^{n+1}-x^n|| / (||x^n|| dt) over the five state vectors falls below TOL
// (checked every M steps).  At the end the carrier states are written as
// <Name>.dofs (CarrierPair::print_dofs, CarrierPair.cpp:89-104); with
// `restart status = true` they are read back first (read_dofs, 106-121) and
// the run continues from t_end to t_end_2.  Errors end the program with the
// reference's message frame and exit status 1 (main.cpp:28-51).  --host-loop
// runs the loop body through the per-call host-vector interface
// (assemble_*_rhs / solve_full_system / solve_Poisson, one C-ABI call each)
// instead of the device-resident pecs_step.  As print_results does, the
// initial state and the state at every output time are written as
// Semiconductor-NNN.vtu / Electrolyte-NNN.vtu / Poisson-NNN.vtu in DIR
// (NNN counting from 0, or from the number of outputs on a restart,
// SolarCell.cpp:1991-2038); --no-vtu skips them.
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <string>
#include <vector>

#include "prm.hpp"
#include "solarcell.hpp"

using namespace pecs_host;

namespace {

struct Cli {
  std::string prm = "input_file.prm";
  int degree = 1;  // main.cpp:11
  long steps = -1;
  double steady_tol = 0;
  int check_every = 10;
  std::string out = ".";
  bool ignore_local = false;
  bool host_loop = false;
  bool vtu = true;
};

Cli parse(int argc, char** argv) {
  Cli c;
  for (int i = 1; i < argc; ++i) {
    const std::string a = argv[i];
    auto next = [&]() -> std::string {
      if (i + 1 >= argc) throw std::runtime_error("missing value after " + a);
      return argv[++i];
    };
    if (a == "--degree") c.degree = std::stoi(next());
    else if (a == "--steps") c.steps = std::stol(next());
    else if (a == "--steady-tol") c.steady_tol = std::stod(next());
    else if (a == "--check-every") c.check_every = std::stoi(next());
    else if (a == "--out") c.out = next();
    else if (a == "--ignore-local-refinement") c.ignore_local = true;
    else if (a == "--host-loop") c.host_loop = true;
    else if (a == "--no-vtu") c.vtu = false;
    else if (a == "-h" || a == "--help") {
      std::printf("usage: pecs_run [input_file.prm] [--degree K] [--steps N] [--steady-tol TOL] "
                  "[--check-every M] [--out DIR] [--ignore-local-refinement] [--host-loop] [--no-vtu]\n");
      std::exit(0);
    } else if (!a.empty() && a[0] == '-') throw std::runtime_error("unknown option " + a);
    else c.prm = a;
  }
  if (c.check_every < 1) throw std::runtime_error("--check-every must be >= 1");
  return c;
}

int run(const Cli& cli) {
  ParameterFile file;
  file.read(cli.prm);
  RunSettings rs = settings_from(file, cli.degree);
  // one level of boundary-layer refinement runs as the reference's Grid.cpp:72-111;
  // deeper local refinement is refused
  if (rs.local_refinements > 1 && !cli.ignore_local)
    throw std::runtime_error("local refinements > 1 are not supported (one level of boundary-layer refinement is); "
                             "pass --ignore-local-refinement to run on the globally refined mesh");
  Parameters prm;
  prm.mesh = rs.mesh;
  if (cli.ignore_local) prm.mesh.local_refine = 0;
  prm.p = rs.params;
  prm.char_length = rs.char_length;
  prm.char_density = rs.char_density;
  prm.char_time = rs.char_time;
  SolarCellProblem sim(prm);
  double stats[8] = {0};
  pecs_stats(sim.ctx(), stats, 8);
  int64_t ns = 0, ne = 0, np = 0;
  pecs_sizes(sim.ctx(), &ns, &ne, &np);
  std::printf("degree %d, refinement %d: %lld semiconductor + %lld electrolyte carrier DoFs, %lld Poisson DoFs\n",
              prm.p.degree, prm.mesh.n_refine, (long long)ns, (long long)ne, (long long)np);
  std::printf("delta_t = %g\n", prm.p.dt);

  std::vector<long> counts;
  if (rs.restart) {
    sim.electron_hole_pair.read_dofs(cli.out);
    sim.redox_pair.read_dofs(cli.out);
    sim.set_state_from_members();
    counts = reference_step_counts(true, rs.t_end, rs.t_end_2, prm.p.dt, rs.time_stamps);
    std::printf("running until t = %g\n", rs.t_end_2);
  } else {
    sim.initial_state();
    counts = reference_step_counts(false, rs.t_end, rs.t_end_2, prm.p.dt, rs.time_stamps);
    std::printf("running until t = %g\n", rs.t_end);
  }
  if (cli.steps >= 0) counts.assign(1, cli.steps);
  unsigned time_step_number = rs.restart ? (unsigned)rs.time_stamps : 0u;
  double print_s = 0;  // output time, kept out of the step rate
  auto print = [&](bool pull) {
    if (!cli.vtu) return;
    const auto tp = std::chrono::steady_clock::now();
    if (pull) sim.pull_state();
    sim.print_results(time_step_number++, cli.out);
    print_s += std::chrono::duration<double>(std::chrono::steady_clock::now() - tp).count();
  };
  print(false);

  long done = 0;
  bool steady = false;
  double res = -1;
  print_s = 0;
  const auto t0 = std::chrono::steady_clock::now();
  for (size_t k = 0; k < counts.size() && !steady; ++k) {
    long left = counts[k];
    if (cli.host_loop) {
      for (; left > 0; --left, ++done) sim.step_host();
      std::printf("output %zu: step %ld\n", k + 1, done);
      print(false);
      continue;
    }
    while (left > 0 && !steady) {
      const long n = cli.steady_tol > 0 ? std::min<long>(left, cli.check_every) : left;
      sim.timer.enter("Time steps (device)");
      check(pecs_step(sim.ctx(), (int)n), sim.ctx());
      sim.timer.leave("Time steps (device)");
      left -= n;
      done += n;
      res = sim.steady_residual();
      steady = cli.steady_tol > 0 && res <= cli.steady_tol;
    }
    std::printf("output %zu: step %ld, steady residual %.3e\n", k + 1, done, res);
    print(true);
  }
  const double wall = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count() - print_s;
  if (!cli.host_loop) sim.pull_state();
  sim.electron_hole_pair.print_dofs(cli.out);
  sim.redox_pair.print_dofs(cli.out);
  const double dofs = 4.0 * (double)ns + (double)np;
  std::printf("%ld steps in %.3f s (%.3f ms/step, %.3e DOF-updates/s)%s\n", done, wall, 1e3 * wall / std::max(1L, done),
              dofs * done / std::max(wall, 1e-30), steady ? ", steady state reached" : "");
  sim.timer.print(stdout);
  return 0;
}

}  // namespace

int main(int argc, char** argv) {
  try {
    return run(parse(argc, argv));
  } catch (std::exception& exc) {
    std::cerr << "\n\n----------------------------------------------------\n"
              << "Exception on processing: \n"
              << exc.what() << "\nAborting!\n"
              << "----------------------------------------------------\n";
    return 1;
  }
}
