"""CPU oracle for the PECS per-time-step IMEX path -- TEST INFRASTRUCTURE ONLY.

This package is a plain numpy/scipy restatement of the reference deal.II
algorithm (mdh266/PECS).  It exists to check the HIP product path; it is
never the thing measured or shipped.  Only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import it, and only as the checker / the timed CPU baseline.

What it restates (reference file:line):
  * fe.py        -- QGauss(k+2) rules, FE_DGQ(k) Lagrange tabulation,
                    RT_{k-1} x DGQ_{k-1} tabulation, MappingQ1 geometry
                    (Assembly.cpp:9-68, SolarCell.cpp:23-31).
  * mesh.py      -- Grid_Maker meshes + boundary ids (Grid.cpp:44-461),
                    interface / region maps (SolarCell.cpp:156-372),
                    component-wise DoF layout (CarrierPair.cpp:23-64,
                    Poisson.cpp:18-69).
  * ldg.py       -- LDG mass, cell+bc and interior-flux matrices
                    (LDG.cpp:40-678), carrier right-hand sides
                    (SolarCell.cpp:1037-1726).
  * mixed.py     -- mixed RT matrix (MixedFEM.cpp:29-164), Neumann D.n
                    constraints (Poisson.cpp:32-49), Poisson right-hand side
                    (SolarCell.cpp:430-815).
  * solarcell.py -- parameters / scaling (Parameters.hpp:117-287), physics
                    point functions (Generation.cpp, InitialConditions.cpp,
                    BiasValues.cpp, SolarCell.hpp:86-98), the time loop
                    (SolarCell.cpp:1898-2095) with sparse direct solves in
                    the role of UMFPACK (Carrier.cpp:26-40, Poisson.cpp:90-105).
  * dealii_rt.py  -- deal.II's FE_RaviartThomas(k-1) node functionals, dual
                    basis and distribute_dofs / component_wise numbering
                    (SolarCell.cpp:23-24, Poisson.cpp:18-30), to check the
                    library's conversion to the reference's Poisson layout.
  * manufactured.py -- the four convergence studies
                    (test_functions.cpp, SolarCell.cpp:2105-3278), on the
                    reference's unit-square grids and on the solar-cell grid
                    the GPU path runs them on (tests/test_gpu_manufactured.py
                    checks the GPU against the latter).

Parity pins (what fixes each function without the reference, which cannot be
built here: it needs deal.II 8.x + UMFPACK + TBB):
  * quadrature, Lagrange and RT tabulations: closed forms and numpy's
    Gauss-Legendre routine (tests/test_oracle_fe.py);
  * element matrices: sympy-exact integrals on 2 x 2-cell rectangles
    (tests/golden/ldg_*.json, mixed_*.json, made by
    tests/golden/make_fixtures.py);
  * LDG / mixed matrices: the reference's manufactured solutions at
    textbook rates, orientation invariance of the interior flux
    (tests/test_oracle_manufactured.py, tests/test_oracle_invariants.py);
  * carrier and Poisson right-hand sides (ldg.carrier_rhs,
    mixed.poisson_rhs): 40-digit golden vectors on 2 x 2-cell rectangles and
    trapezoids with every term switched on (tests/golden/rhs_*.json, made by
    tests/golden/make_rhs_fixtures.py), with planted mistakes that must be
    caught (tests/test_oracle_rhs_fixtures.py); insulated-boundary mass
    conservation, the Boltzmann zero-current steady state at rate k+1 (drift
    signs and the permittivity), generation moments and Poisson functionals
    against closed-form integrals (tests/test_oracle_physics_pins.py);
    interface antisymmetry and the neutral fixed point
    (tests/test_oracle_invariants.py);
  * solves: residual checks of the sparse LU.
"""
