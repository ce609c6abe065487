"""CPU oracle for the ADC(2)-s sigma-vector path.

TEST INFRASTRUCTURE ONLY.  Nothing in the product path (the HIP library
``libadc2g.so`` and the ``adc_ottawa_amd`` package) imports, links or calls
anything in this directory.  Only ``tests/``, ``__graft_entry__.smoke()`` and
the ``cpu_baseline`` leg of ``bench.py`` use it, and only as the checker.

Modules (each function cites the reference file:line it restates):

* ``csf``      configuration tables (source/adclib/select_fano.f90)
* ``elements`` matrix elements (source/adclib/adc_ph.f90) and the dense/COO
               build (source/adclib/get_matrix.f90)
* ``dfsigma``  an independent density-fitted restatement of the same sigma
* ``davidson`` block Davidson (source/eigen/block_davidson.f90)
* ``sil``      short-iterative-Lanczos propagation
               (source/propagation/sillib.f90, fvecprop.f90)
* ``mlp``      the learned preconditioner's network, forward and backward
* ``schur``    exact eigenvalues from the Schur-complement (Lowdin
               partitioning) of the diagonal 2h2p block
* ``coo_sigma.c`` C/OpenMP restatement of the reference's stored-matrix CPU
               path (COO build + ``hxvec_incore`` SpMM), the CPU baseline.

How each part is pinned (beyond restating the reference) is listed in the
module headers.
"""
