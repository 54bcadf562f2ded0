// CPU checks of the host layer that need no GPU: the .prm reader, the mapping
// onto the C-ABI structs, the reference step count and the .dofs format.
//
//   pecs_host_test prm <file> <degree>   print the settings as key=value lines
//   pecs_host_test steps <restart 0|1> <t_end> <t_end_2> <dt> <stamps>
//   pecs_host_test dofs <file> <n>       write v[i] = i + 0.25 as a .dofs file, read it back
//   pecs_host_test readdofs <file>       print the values of a .dofs file
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <string>

#include "prm.hpp"
#include "solarcell.hpp"

using namespace pecs_host;

int main(int argc, char** argv) {
  try {
    if (argc < 2) throw std::runtime_error("usage: pecs_host_test prm|steps|dofs|readdofs ...");
    const std::string mode = argv[1];
    if (mode == "prm" && argc == 4) {
      ParameterFile f;
      f.read(argv[2]);
      const RunSettings r = settings_from(f, std::atoi(argv[3]));
      const pecs_params& p = r.params;
      std::printf("r1=%.17g\nr2=%.17g\nH=%.17g\nL=%.17g\nn_refine=%d\ninsulated=%d\nschottky=%d\nbl=%.17g\n",
                  r.mesh.r1, r.mesh.r2, r.mesh.H, r.mesh.L, r.mesh.n_refine, r.mesh.insulated, r.mesh.schottky,
                  r.mesh.boundary_layer);
      std::printf("degree=%d\ndt=%.17g\nt_end=%.17g\nt_end_2=%.17g\ntime_stamps=%d\nrestart=%d\nlocal=%d\n", p.degree,
                  p.dt, r.t_end, r.t_end_2, r.time_stamps, (int)r.restart, r.local_refinements);
      for (int i = 0; i < 4; ++i) std::printf("mu%d=%.17g\n", i, p.mu[i]);
      std::printf("eps_s=%.17g\neps_e=%.17g\nlambda2=%.17g\nk_et=%.17g\nk_ht=%.17g\n", p.eps_s, p.eps_e, p.lambda2,
                  p.k_et, p.k_ht);
      std::printf("v_rec_n=%.17g\nv_rec_p=%.17g\nphi_bi=%.17g\nphi_app=%.17g\nphi_sch=%.17g\n", p.v_rec_n, p.v_rec_p,
                  p.phi_bi, p.phi_app, p.phi_sch);
      std::printf("gen_mode=%d\nG0=%.17g\nalpha=%.17g\nn_e=%.17g\np_e=%.17g\nr_inf=%.17g\no_inf=%.17g\n", p.gen_mode,
                  p.G0, p.alpha, p.n_e, p.p_e, p.r_inf, p.o_inf);
      return 0;
    }
    if (mode == "steps" && argc == 7) {
      const auto n = reference_step_counts(std::atoi(argv[2]) != 0, std::atof(argv[3]), std::atof(argv[4]),
                                           std::atof(argv[5]), std::atoi(argv[6]));
      for (long c : n) std::printf("%ld\n", c);
      return 0;
    }
    if (mode == "dofs" && argc == 4) {
      const size_t n = std::strtoul(argv[3], nullptr, 10);
      std::vector<double> v(n), w;
      for (size_t i = 0; i < n; ++i) v[i] = (double)i + 0.25;
      {
        std::ofstream out(argv[2], std::ios::binary);
        block_write(out, v);
      }
      std::ifstream in(argv[2], std::ios::binary);
      block_read(in, w);
      if (w != v) throw std::runtime_error("dofs round trip mismatch");
      std::printf("ok %zu\n", w.size());
      return 0;
    }
    if (mode == "readdofs" && argc == 3) {
      std::ifstream in(argv[2], std::ios::binary);
      if (!in) throw std::runtime_error(std::string("cannot open ") + argv[2]);
      std::vector<double> w;
      block_read(in, w);
      for (double x : w) std::printf("%.17g\n", x);
      return 0;
    }
    throw std::runtime_error("bad arguments");
  } catch (std::exception& e) {
    std::fprintf(stderr, "error: %s\n", e.what());
    return 1;
  }
}
