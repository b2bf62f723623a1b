"""CPU oracle for the interior-penalty DG diffusion solve.

TEST INFRASTRUCTURE ONLY.  This package is a plain NumPy/SciPy FP64 restatement
of the reference algorithm (linfeng810/dg_nn_public, z02-implicit/z06-DGtoCG and
the variable-diffusivity operator of z05-ns).  It is the checker that the HIP path
in ``dg_nn_public_amd`` is compared against.  Only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may import
it, and only as the checker or the timed CPU baseline -- never as the product path.

Each function cites the reference file:line it restates.  Reference paths are
relative to the reference repository root; ``z02/`` abbreviates
``z02-implicit/z06-DGtoCG/`` and ``z05/`` abbreviates ``z05-ns/``.

Modules
-------
tables   reference-element quadrature, P3 Lagrange basis, face tables, gi_align
dg       topology, geometry, SIPG residual, block diagonal, smoother, RHS, L2 error
mg       DG<->P1CG transfers, P1CG sparsity, colouring, RAR, SFC levels, V-cycle
krylov   left-preconditioned GMRES(m) (MGS), MG-PCG, stand-alone MG
"""
