"""FP64 CPU oracle for the spectral 2-D turbulence step.

TEST INFRASTRUCTURE ONLY. Nothing in the product path imports this package: only
``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may use it,
and only as the checker (or, for the CPU baseline, as the thing timed on the host cores).

Modules
-------
``spectral``   the pseudo-spectral scheme in FP64 NumPy, restated from the equations of the
               reference (reference/spectral2d.py, README.md "The scheme").
``fftmodel``   the radix plan and the in-place radix-4 DIT line transform
               (reference/fft_ref.py, src/gpu/fft.js, src/shaders/fft.wgsl).
``gpusem``     semantics that exist only in the reference's WGSL/JS GPU path: the integer
               hash, stochastic and pointer forcing, the CFL reduction, spectrum seeding,
               scenarios, the BFECC dye and the diagnostics conventions.

Parity status of each function is stated in its docstring: "pinned by ..." names the golden
vector, closed form or invariant that a test checks it against; "parity unpinned" marks what no
such check covers.
"""
