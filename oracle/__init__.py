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
* philox: Random123 known-answer vectors.
* integrands + map transform: sigma-table integral identity against the
  shipped ``sm_xsec.pkl``/``dark_xsec.pkl`` values; closed-form total cross
  sections (Moller, Bhabha, mu-e) and Klein-Nishina.
* draw (accept/reject): brute-force histogram of accepted samples against the
  normalised integrand on a fine grid.
* kinematics: energy conservation / on-shell invariants per vertex.
* MCS: Rayleigh law of the space angle (KS), |p| preserved.
* transport: exponential free-path law (KS), per-shower energy budget.
* dark weights/drate: stored ``dark_weights.pkl``/``dark_drate.pkl`` caches.
Functions without such a check say "parity unpinned" in their docstring.
"""
