"""CPU FP64 restatement of the reference propagator and inversion loops.

TEST INFRASTRUCTURE ONLY.  This package is the parity oracle: only the test
suite, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py``
may import it, and only as the checker (or the reported CPU baseline), never
as the thing measured or shipped.  The product path (``seismic_amd``) never
imports it and fails loudly when its HIP library is missing.

Modules
-------
fd2d    forward / tangent / adjoint propagator (core/seis_forward2.py)
invert  cost+gradient, Gauss-Newton (dense and CG), LBFGS + strong Wolfe
        (core/seis_invert.py, core/seis_numerics.py)
priors  RowTotalVariation, TotalVariation, SquaredExponential, L2 smoothness
        (core/seis_prior.py)
"""
