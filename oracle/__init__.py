"""CPU FP64 oracle for the coupled drift-diffusion Newton solve.

TEST INFRASTRUCTURE ONLY.  Nothing in the product package may import,
call or execute anything in this directory: only ``tests/``,
``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg use
it, and only as the checker / the timed CPU reference, never as the
thing shipped.

It restates, in NumPy/SciPy, the reference's structured 2D/3D
drift-diffusion path (Boltzmann statistics, single semiconductor,
ohmic Dirichlet contacts and Robin gates):

  physics.py    constants, silicon parameters, Bernoulli, recombination,
                mobility (Caughey-Thomas, Lombardi CVT), BGN, lifetimes,
                ohmic contact values, control-volume widths
  dd.py         Device2D / Device3D (dimension-generic core): equilibrium
                Poisson Newton, coupled bias Newton, residual/Jacobian
                assembly with the 5-/7-point stencils and normal-axis
                gates, terminal current
  linsolve.py   symmetric Dirichlet elimination, direct (SuperLU) and
                node-block-Jacobi GMRES/BiCGStab solves with the
                reference's acceptance contract
  transient.py  theta-scheme transient steps (storage term, line-search
                Newton, adaptive dt) and the voltage waveforms

Every function cites the reference file:line it follows.  How each
function is pinned (golden fixtures generated from the reference,
closed forms, invariants) is recorded in tests/ and in the header of
each module.
"""
