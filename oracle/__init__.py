"""CPU oracle for torchlpc_amd: TEST INFRASTRUCTURE ONLY.

The oracle restates the reference CPU path (DiffAPF/torchlpc 0.8) so that the
HIP kernels can be checked against it:

* ``lpc_oracle.c`` (built to ``liblpc_oracle.so``) follows the reference loops:
  forward ``torchlpc/csrc/scan_cpu.cpp:80-122,143-164``, first-order scan
  ``scan_cpu.cpp:62-77``, backward composition ``torchlpc/core.py:186-229``.
* ``oracle.py`` wraps it with numpy, adds FP64 closed forms and the seeded
  workload generators shared by tests and the benchmark.
* ``Makefile`` target ``ref`` compiles the reference's own CPU op into
  ``_ref/torchlpc_ref_cpu.so`` (used to pin the restatement bit for bit).

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this package, and only as the checker or as the timed
CPU baseline. The product package ``torchlpc_amd`` never imports it.

Parity pins (each function of the oracle is checked by at least one of them):

* forward ``lpc``: bit-exact against ``_ref/torchlpc_ref_cpu.so`` and the committed
  golden fixtures (``tests/golden``); biquad impulse response closed form; scipy
  ``lfilter`` for time-invariant coefficients; linearity; zi/zf chaining.
* ``scan``: bit-exact against the reference op; the ``0.9**t`` impulse response.
* ``lpc_backward``: golden fixtures; finite differences of the oracle forward in
  FP64; the transpose identity <M^-1 u, v> = <u, M^-T v>.
"""
