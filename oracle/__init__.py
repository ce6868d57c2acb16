"""CPU FP64 restatement of the reference's residual path -- TEST INFRASTRUCTURE ONLY.

This package is the checker the GPU kernels are compared against.  Only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may import it, and only as
the oracle -- never as the measured or shipped path.  The product (``aquaflux_amd``) never
imports it and fails loudly when its HIP library is missing.

The restatement is torch-CPU float64 (vectorised gather / ``index_add_`` in place of
``jax.ops.segment_sum``), following the reference's functions one for one; each function cites the
reference ``file:line`` it restates.  torch autograd of this restatement stands in for ``jax.jvp``
/ ``jax.vjp`` of the reference, which cannot be imported here (no JAX in the image).

Pinning (what checks each part beyond retyping its formula):
  * integer connectivity: ``connectivity.py`` loop generators vs the vectorised product generator
    and the reference's structured-grid unit-test facts (face counts, patch sizes, owner < nb);
  * geometry: closed forms of the reference's ``tests/unit/test_mesh.py`` (unit square, cube,
    tetrahedron, prism, parallelogram, far-from-origin centroid);
  * gradients: linear-field exactness; fluxes: the single-face closed forms of
    ``tests/unit/test_diffusion.py`` / ``test_advection.py`` / ``test_rhie_chow.py`` /
    ``test_pressure_force.py``; conservation; limiter psi==1 on linear fields;
  * SST closures: the closed forms of ``tests/unit/test_sst.py`` / ``test_sst_sources.py``;
  * derivatives: central finite differences and adjoint dot-product identities.
Functions with no such pin are marked "parity unpinned" in their docstrings.
"""
