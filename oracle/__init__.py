"""CPU oracle of the list-1 path -- TEST INFRASTRUCTURE ONLY.

A plain restatement of the reference algorithm (xywei/volumential) used to
check the GPU product.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import or run anything here, and
only as the checker or as the timed CPU baseline; the product package
(``volumential_amd``) never imports it.

Modules:

* ``orbit_ref``   -- gallery and arithmetic orbit reduction, following
  ``list1_gallery.py:118-198`` and ``orbit_arithmetic.py:29-790``
  statement by statement, plus a brute-force group-orbit oracle.
* ``list1_ref``   -- the ``near_field`` kernel semantics of
  ``list1.py:1057-1148`` (numpy, per-entry orbit arithmetic).
* ``duffy_ref``   -- Duffy node rules and per-entry quadrature
  (``singular_integral_2d.py:594-949``,
  ``nearfield_potential_table.py:917-946, 1099-1177, 2098-2201``),
  and closed forms used to pin them.
* ``list1_cpu.c`` -- C/OpenMP restatement of the ``near_field`` kernel,
  built to ``oracle/liboracle_list1.so``; the timed CPU baseline.

Parity pins (see the tests): the reference's own known-answer values
(7/33/139 cases; q=3 3D: 2510 canonical rows, 7290 dense layout rows,
1576 metadata bytes; the sorting-network row), brute-force orbit
enumeration over all 2^d d! signed permutations, closed-form box
potentials for the Laplace kernels, Gauss-Legendre weights for the
constant kernel, and linearity / completeness of the list-1 sum.
"""
