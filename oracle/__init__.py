"""CPU oracle of the multislice ptychography hot path — TEST INFRASTRUCTURE ONLY.

A numpy/scipy restatement of PyPty's algorithm (every function cites the
reference file:line it follows).  Only ``tests/``, ``__graft_entry__.smoke()``
and the ``cpu_baseline`` leg of ``bench.py`` may import it, and only as the
checker or the timed CPU baseline; the product package ``pypty_amd`` never
imports or calls it.

Pinning (see tests/test_oracle.py):
  * FFT wrappers      brute-force DFT (n = 6, 12, 16), Parseval, shift theorem
  * grids / masks     closed forms of fftfreq and the cutoff radius
  * multislice        O = 1 reduces to free propagation; constant phase slices;
                      H(d1) H(d2) = H(d1 + d2); norm conservation
  * loss              closed forms for a flat probe (mode mean, both losses)
  * adjoint / grads   dot-product test and central finite differences of the
                      loss for object, probe and sub-pixel positions
  * scatteradd        np.add.at brute force, conservation of the sum
  * L-BFGS            textbook two-loop recursion on a convex quadratic
"""
