"""CPU oracle for the MeltPoolDG heat-operator hot path.

TEST INFRASTRUCTURE ONLY. Only ``tests/``, ``__graft_entry__.smoke()`` and
``bench.py``'s CPU-baseline leg import this package, and only as the checker
(or the timed CPU baseline). The product package ``meltpooldg_amd`` never
imports it.

Modules
-------
cpu    ctypes front end of ``heat_oracle.cpp`` (C++ restatement of the
       reference operator, diagonal, CG and GMRES).
rng    SplitMix64 seeded vectors keyed by (seed, lattice index).
kron   independent Kronecker-product closed form of the operator for
       constant / separable coefficients and the vortex velocity.
brute  independent dense assembly by explicit quadrature on tiny meshes.
tables independent numpy construction of the 1D FE tables.

Parity pins (what checks each function):
  tables          numpy leggauss, roots of P_p', Lagrange identities, exactness
  oracle_vmult    kron (configs with constant/separable k, vortex velocity),
                  brute (nodal heaviside), golden fixtures, symmetry, A*1 = c*M*1
  oracle_diagonal kron closed-form diagonal, brute diag(A_ff)
  oracle_material reference unit_tests/material.output known answers
  oracle_solve    manufactured solutions, textbook dense PCG/GMRES iteration
                  counts on the brute-force matrix
"""
