"""CPU oracle for the cmtip hot path (TEST INFRASTRUCTURE ONLY).

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this package, and only as the checker or as the timed
CPU baseline.  The product package ``cmtip_amd`` never imports it.

Modules
-------
nufft   FP64 type-1/type-2 NUFFT: exact direct sums (C, OpenMP), a numpy
        direct sum, and an ES-kernel fast route (C gridding + numpy FFT).
geom    quaternion -> rotation matrix, rotated q-vectors, alignment errors.
align   orientation matching restated on sklearn's euclidean_distances.
ac      the autocorrelation normal operator and scipy CG solve, restated.
synth   is NOT here: seeded synthetic inputs live in cmtip_amd.synth because
        the product's benchmark needs them too.
"""
