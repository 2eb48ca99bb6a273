"""CPU oracle for the PETITE shower-transport hot path.

TEST INFRASTRUCTURE ONLY.  This package is a float64 numpy restatement of the
reference algorithm (kjkellyphys/PETITE, src/PETITE/shower.py,
dark_shower.py, all_processes.py, kinematics.py, moliere.py, particle.py,
physical_constants.py, radiative_return.py, atomic_annihilation.py,
atomic_compton.py); each function cites the reference lines it follows.
Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline
leg may import it, and only as the checker.  The product (``petite_amd``)
never imports, links or runs anything from here.

Random numbers: the reference draws from three unseeded/independent
generators (numpy global, Python ``random``, vegas), so its showers are not
reproducible.  The oracle instead consumes the same counter-based
Philox4x32-10 streams, in the same order, as the HIP kernels
(``oracle.philox``), which makes candidate-level comparisons possible.

Parity pins (what checks each function; see tests/):
* philox: Random123 known-answer vectors (test_oracle_philox.py).
* integrands + map transform: the sigma-table integral identity against the
  shipped ``sm_xsec.pkl``/``dark_xsec.pkl`` values (test_oracle_tables.py);
  closed-form total cross sections (Moller, Bhabha, mu-e) and Klein-Nishina
  (test_oracle_closed_forms.py).
* draw (accept/reject): the accepted points' law -- 1-D processes against f
  normalised on a fine grid, Brem/PairProd marginals against the VEGAS
  J f-weighted estimate, at max_F above sup J f (test_oracle_draw_law.py).
* kinematics: vertex invariants (test_oracle_physics.py), and the invariants
  recomputed from the four-vectors (momentum transfer q^2, Mandelstam t,
  the annihilation cut, the dark vector's angle) against the variables the
  integrands assign to the same map point (test_oracle_kinematics_pins.py).
* MCS: the Lynch-Dahl width against Highland's formula
  (test_oracle_moliere_pin.py), the Rayleigh law of the space angle and |p|
  preservation (test_oracle_physics.py).
* DarkAnn: its integral against the reference's luminosity table
  lumi_integral_list (test_oracle_dark_ann_lumi.py).
* transport: exponential free-path law (KS), per-shower energy budget.
* dark weights/drate: the stored ``dark_weights.pkl``/``dark_drate.pkl``
  caches and scipy quad (test_dark_tables.py, test_oracle_dark_yield.py); the
  panel rule of dark_panels.py against the same quad integrals and caches,
  and its running integrals against closed-form trapezoids
  (test_dark_tables.py).
Functions without such a check say "parity unpinned" in their docstring.
"""
