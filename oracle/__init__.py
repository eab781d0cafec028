"""CPU oracle for the ADFWI propagator hot path -- TEST INFRASTRUCTURE ONLY.

This package restates, in eager PyTorch on the CPU (plus plain numpy loops for
the damping profiles and tiny brute-force steps), the reference algorithm of
ADFWI's acoustic and elastic staggered-grid propagators:

* ``oracle.boundary``  -- damping profiles      (ADFWI/propagator/boundary_condition.py:15-120)
* ``oracle.acoustic``  -- acoustic forward_kernel and its autograd path
                          (ADFWI/propagator/acoustic_kernels.py:16-301)
* ``oracle.elastic``   -- elastic 4th-order PML forward_kernel and autograd path
                          (ADFWI/propagator/elastic_kernels.py:20-84,175-423,917-1037)
* ``oracle.model``     -- model parameterisations (ADFWI/model/parameters.py:47-213,
                          ADFWI/model/acoustic_model.py:111-134, elastic_model.py:198-234)
* ``oracle.loops``     -- pure-Python loop restatement of one time step, for tiny grids.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this package, and only as the checker or as the timed
CPU baseline.  The product package ``adfwi_amd`` never imports it; its compute
path is the HIP library and fails loudly when that library is missing.

Pinning (what each function is checked against, in ``tests/``):

* golden vectors written by ``tests/golden/make_golden.py`` from the reference's
  own ``forward_kernel`` / ``bc_pml`` / ``bc_pml_xz`` (records, energy maps and
  autograd gradients on tiny grids, FS on and off, several checkpoint segment
  counts): pins ``acoustic.forward_kernel``, ``elastic.forward_kernel``,
  ``boundary.*``;
* closed forms: FD coefficients 9/8 and -1/24, Ricker, Gardner, Lame,
  Thomsen -> Cij (``model.*``);
* brute force: ``loops.acoustic_step`` / ``loops.elastic_step`` against the
  vectorised step on tiny grids (window bounds and free-surface rows);
* invariants: linearity in the wavelet, shot-batch independence, FP64
  dot-product (adjoint) tests of the autograd path.
"""
