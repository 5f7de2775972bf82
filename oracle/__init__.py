"""FP64 CPU oracle for the finite-element residual hot path.

TEST INFRASTRUCTURE ONLY.  Nothing in the product package (`folax_amd`) may
import, call, link or execute anything under this directory.  Only `tests/`,
`__graft_entry__.smoke()` and the `cpu_baseline` leg of `bench.py` use it, and
only as the checker (or as the timed CPU baseline), never as the measured or
shipped path.

`oracle.fe` is a from-scratch NumPy FP64 restatement of the reference
algorithm (folax `fol/geometries`, `fol/loss_functions`,
`fol/constitutive_material_models`, `fol/solvers`); every function cites the
reference file:line it follows.  `oracle/cpu_ref.c` is a plain-C/OpenMP
restatement of the same per-element arithmetic used as the timed CPU baseline.
"""
