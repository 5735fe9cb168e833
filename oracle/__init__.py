"""FP64 CPU oracle for the chi(k) forward model, the windowed chi(R) transform and the
parallel-tempering RMC swap sampler of dluccioni/CHIMERA.

THIS PACKAGE IS TEST INFRASTRUCTURE. Only `tests/`, `__graft_entry__.smoke()` and the
`cpu_baseline` leg of `bench.py` may import it, and only as the checker (or as the
timed CPU baseline). The product package `chimera_amd` never imports it and has no
CPU fallback.

Every function restates the reference's algorithm in plain NumPy (float64) and cites
the reference file:line it follows. Modules:

  lattice    supercells, neighbour lists, shells, reverse index, Warren-Cowley
  tables     the synthetic single- and multiple-scattering tables of the workloads
  mspaths    multiple-scattering path classes, entries, per-supercell entry counts
  forward    pair-space chi(k) (the reference's chi_reference) and count space
  fourier    Hanning window, resample, FFT transform, direct DFT at any R
  rmc        splitmix64 proposals, affected pairs, profiled chi^2, Metropolis,
             replica exchange, and the plain sampler loop (the CPU baseline)

Parity status of each function (how it is pinned) is listed in the module headers;
a function with no independent check says "parity unpinned".
"""
