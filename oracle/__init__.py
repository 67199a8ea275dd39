"""CPU oracle for the Simulon MD step — TEST INFRASTRUCTURE ONLY.

This package restates, in numpy (FP32 and FP64, with every floating-point
operation in an explicit, fixed order), the algorithms of the reference's
PyTorch CPU path for one MD step:

  box.py         core/box.py:109-148                 minimum image, wrap, scale
  neighbor.py    core/neighbor_search/gpu_kdtree.py:90-156,
                 io_utils/reader.py:136-163, 364-376, 477-484
                                                     brute-force Verlet half list,
                                                     skin/2 trigger
  lj.py          core/force/lennard_jones_force.py:15-95
  eam.py         io_utils/eam_parser.py:23-210,
                 core/force/eam_force_cu.py:334-579  tables + two-pass EAM
  integrate.py   core/integrator/integrator.py:45-124,
                 io_utils/reader.py:319-339,
                 core/md_model.py:83-166,            BAOB Langevin velocity Verlet,
                 core/barostat.py:36-68              Berendsen scaling factor
  mechanics.py   core/barostat.py:78-208,
                 core/mechanics/loading.py:39-78     anisotropic barostat,
                                                     uniaxial tensile loading
  mlp.py         radial-descriptor MLP potential (no reference counterpart:
                 the reference's learned potential is an external CHGNet
                 wrapper; this restates the library's own network in FP64)

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
import this package, and only as the checker (or the timed CPU baseline).
The product path (simulon_amd) never imports it.

How each function is pinned is listed in its docstring ("pinned by: ...");
the checks live in tests/test_oracle_*.py.
"""
