"""Oracle for the fastlpr_amd engine: TEST INFRASTRUCTURE ONLY.

Nothing in the product (the ``fastlpr_amd`` package, its C-ABI library or its
HIP kernels) imports, links or calls anything in this directory.  Only
``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` use it, and only as the checker / the timed CPU baseline.

Contents
--------
``restate``  NumPy/SciPy restatement of the reference fastLPR dataflow for
             ``cv_fastlpr`` / ``cv_fastkde`` (Gaussian-gridding type-1 NUFFT,
             FFT convolution with the kernel-moment spectra, per-node solves,
             Hutchinson DOF, GCV/LCV and the 1-SE rule).  Two precisions:
             ``mode="ref"`` mirrors the reference's accuracy<=4 arithmetic
             (complex64 spread grid and spectrum products, float32 S/T/solve);
             ``mode="f64"`` runs the same discrete estimator in float64.
``direct``   Brute-force float64 definitions that pin ``restate`` where the
             reference's fixtures cannot (2-D / 3-D order 2): the spread as an
             explicit loop over points and footprint taps, the moments as
             explicit sums over grid cells with the full d-dimensional kernel,
             and Gauss-Jordan solves per node.

Parity status of each function is stated in its docstring ("pinned by ...").
"""
