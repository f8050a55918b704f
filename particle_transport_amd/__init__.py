"""MI355X-native trajectory and lag-statistics engine for particle_transport.

The compute path is libpt_hip.so (hand-written HIP for gfx950, C ABI in
include/pt_hip.h).  The modules here mirror the reference's Julia entry
points on top of it:

  time_units  reference clock, integer cadence, lag grids (time_units.jl)
  fields      input loading, grid, synthetic turbulence generator
  injection   particle injection (phase_space_gpu_runner.jl:196-245)
  runner      run_gpu_ensemble / run_mu_cache_ensemble and the C-ABI handles
  dmumu       <dmu^2>, D_mumu(mu, tau) products (compute_delta_mu2_dmumu_full.jl)
  dpp         D_pp(tau) and energy snapshots (compute_dpp_full.jl)
  serving     continuous-batching campaign server
"""
__all__ = ["time_units", "fields", "injection", "runner", "dmumu", "dpp"]
