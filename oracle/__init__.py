"""Oracle: CPU restatements of photonfdtd's FDTD stepping path.

TEST INFRASTRUCTURE ONLY. Only ``tests/``, ``__graft_entry__.smoke()`` and
``bench.py``'s CPU-baseline leg may import or run anything here, and only as
the checker / CPU baseline -- never as the thing measured or shipped. The
``photonfdtd_amd`` package never imports this directory; its ``backend="hip"``
path fails loudly when the HIP library is missing.

Contents
--------
fdtd_numpy.py  float64 NumPy restatement of the reference's in-core stepper
               (``src/photonfdtd/simulation.py:860-1459``) plus its own copy of
               the setup tables (CPML ``boundaries.py:31-97``, ADE
               ``simulation.py:572-669``), so setup bugs in the package are
               caught too. Subpixel smoothing restated over the whole refined
               domain (``smoothing.py:119-172``).
cpu_fdtd.cpp   C++/OpenMP restatement of the reference's compiled CPU stepper
               (``rust/src/kernel.rs:153-372`` fused H/E passes with the CPML
               psi folded into the curl, ``rust/src/blocked.rs:134-350``
               ghost-zone temporal blocking) plus ADE and per-step DFT: the
               benchmark's compiled CPU baseline. ``build_cpu.py`` builds
               ``libpfcpu.so``; ``fdtd_cpu.py`` drives it.
surrogate_torch.py  PyTorch CPU restatement of the surrogate network.
fdtd_torch.py  differentiable float64 restatement for the adjoint checks.

How each function is pinned (besides the golden vectors in
``tests/golden/``, written by the reference's own NumPy path):

* CPML b/c:        closed form at hand-computed nodes + 1-D absorption >= 20 dB.
* dt / n_steps:    closed form (Courant).
* Yee update:      speed of light within 3 %, plane-DFT == full-DFT slice,
                   reciprocity and linearity in the source amplitude.
* ADE recursion:   Lorentz phase index within 2.5 % of analytic, Drude finite,
                   ``omega0 dt >= 2`` guard, lossless-pole energy bound.
* DFT:             equals a brute-force post-hoc DFT of an every-step
                   FieldMonitor series (rel 1e-10).
* Subpixel:        the reference-written subpixel fixtures, bit for bit.
* cpu_fdtd.cpp:    float64 DFTs within 1e-12 of the reference-written
                   fixtures (as the reference checks its Rust backend against
                   NumPy), temporal blocking bit-identical to plain stepping.
"""
