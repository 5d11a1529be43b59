"""CPU oracle for the hipnufft hot path -- TEST INFRASTRUCTURE ONLY.

Nothing in the product (the ``hipnufft`` package, ``libhipnufft.so``) imports,
links or calls anything under this directory.  Only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` use
it, and only as the checker (or, for the baseline, as the thing timed on the
host cores), never as the thing measured on the GPU.

Contents
--------
ndft        exact FP64 non-uniform DFT (explicit matrix and implicit loops),
            restating ``src/mrinufft/operators/interfaces/nudft_numpy.py:12-81``.
nufft       the exponential-of-semicircle (ES) NUFFT in numpy: kernel
            parameters, deconvolution factors, the float32 fold/bin rule, the
            stable bin sort, spreading, interpolation and the type-1 / type-2
            transforms.  It restates the published FINUFFT algorithm
            (Barnett, Magland & af Klinteberg, SIAM J. Sci. Comput. 41(5),
            2019) that the reference reaches through
            ``operators/interfaces/finufft.py:23-80`` and
            ``operators/interfaces/cufinufft.py:51-137``.
mri         the MRI operator layer (SENSE expand/combine, density, batching,
            normalisation, data consistency) restated from
            ``operators/base.py:949-1139`` and
            ``operators/interfaces/cufinufft.py:478-1178``.
solvers     power method, CG and Pipe density, restated from
            ``operators/base.py:1155-1221``, ``extras/optim.py:801-902`` and
            ``operators/interfaces/cufinufft.py:1281-1331``; LSMR as
            scipy's published implementation, which the reference's
            ``extras/optim.py:498-798`` adapts; the exact SENSE operator
            as a scipy LinearOperator.
smaps       low-frequency coil-map estimation (window, blur, root sum of
            squares, LSQR as scipy's published implementation) restated
            from ``extras/smaps.py:83-306``.
stacked     the stacked 2.5D operator (z DFT and plane selection) restated
            from ``operators/stacked.py:447-790``.
orc         the off-resonance operator and its interpolator factorizations
            restated from ``operators/off_resonance.py:230-302`` and
            ``extras/field_map.py:214-566``.

How each function is pinned (checked independently of its own formula) is
listed in the docstring of each module; the tests live in
``tests/test_oracle_*.py``.
"""
