"""CPU oracle of the NUFFT hot path: TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import or run anything here, and only as the checker (or as
the timed CPU baseline). The library (``finufft_amd``) never calls it; the
product path has no CPU fallback.

Contents and what pins each part (the pinning tests live in tests/):

* ``nudft``  FP64 direct NUDFT of types 1/2/3 (restates test/utils/dirft1d.hpp,
  dirft2d.hpp, dirft3d.hpp and norms.hpp). Pinned by closed forms (a single
  point / single mode give pure exponentials; Parseval-style adjointness
  <type1(c), f> = <c, type2(f)>), by numpy.fft on equispaced points, and by
  brute-force loops on tiny inputs.
* ``kernel`` ES kernel, width/beta rules, Horner coefficient fit, fine grid
  size, Gauss-Legendre nodes and the kernel's Fourier series. Pinned by the
  reference's own tables and known-answer vectors (Horner coefficients of
  include/cufinufft/contrib/*.inc, next235 table of test/testutils.cpp:39-44,
  the quadrature test of test/testutils.cpp:70-87) and by scipy quadrature.
* ``sort``   bin index and stable bin sort (restates the GPU bin index of
  include/cufinufft/spreadinterp.hpp:214-229 and the stable order of the CPU
  bin_sort_singlethread_impl, include/finufft/spread.hpp:459-584). Pinned by
  numpy lexsort brute force and the directed-rounding fold of ``cfold``.
* ``cfold``  the device fold_rescale (spreadinterp.hpp:61-89) in C with the
  rounding mode switched by fesetround (oracle/csrc/oracle.c), independent of
  the library's exact-error emulation. Pinned by exhaustive comparisons with
  float128-free closed cases (multiples of 2*pi, +-pi, far outside the box).
* ``cpu``    CPU NUFFT restatement used as the CPU baseline (spread/interp in
  C with OpenMP, FFT by scipy.fft) - see its module docstring.
"""
