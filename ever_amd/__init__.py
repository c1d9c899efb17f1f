"""ever_amd: MI355X-native (gfx950 HIP) exact volumetric ellipsoid tracer with the reference's PyTorch API.

Drop-in module: ``ever_amd.splinetracers.fast_ellipsoid_splinetracer`` (trace_rays, SplineTracer), and
``ever_amd.eval_sh`` (EvalSH, eval_sh). The compute runs in ``libever_amd.so`` (C ABI: include/ever_amd.h).
"""
