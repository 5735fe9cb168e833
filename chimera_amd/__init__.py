"""MI355X-native chi(k) forward model, windowed chi(R) transform and parallel-tempering
RMC swap sampler (the hot path of dluccioni/CHIMERA), on hand-written gfx950 HIP kernels
behind the C ABI in include/chimera_hip.h.

Drop-in modules (same class names, arguments and exceptions as the reference):
  chimera_amd.forward        ForwardModel          (exafs_gpu/forward.py)
  chimera_amd.fourier        FourierTransform,     (exafs_gpu/fourier.py)
                             r_grid, hanning_window
  chimera_amd.multift        MultiEdgeFT           (chimera/multift.py)
  chimera_amd.method_a_rmc   RMCSampler            (methods/method_a_rmc.py): the
                             count-space engine for swap-only chains on the lattice, and
  chimera_amd.pair_rmc       PairRMCSampler        for every other move set (displacements,
                             energy prior, |chi(R)| targets, off-lattice starts);
                             RMCSampler(...) returns whichever the arguments need
chimera_amd.dropin.install() registers these under the reference's module names.
Host setup: lattice, scattering, mspaths, workloads; analytics: ess.
"""
__version__ = "0.1.0"
