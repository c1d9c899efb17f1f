"""Benchmark: Mrays/s of forward + backward of the exact ellipsoid tracer, 1M clustered ellipsoids @ 1920x1080.

Workload (one step): LBVH rebuilt from the primitives, SH degree 3 colour prepass fused into the build, forward
trace of 1920x1080 rays, backward with N(0,1) upstream gradients on all five output columns, through the
drop-in autograd path (trace_rays_sh + .backward()). Scene: 1,000,000 ellipsoids from a 64-component Gaussian
mixture (std 0.15) in [-2,2]^3, aspect ratios up to 100:1, densities logU[0.1,200]; camera at radius 4.

Prints one JSON line: value = rays / second of whole steps (Mrays/s), the roofline record of the dominant
kernel (algorithmic bytes / its event-timed duration vs 8 TB/s HBM) and the CPU baseline (the FP32 C++
restatement of the reference algorithm, OpenMP, timed on a bounded ray subsample of the same workload).
"""
import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Mrays/s fwd+bwd, 1M ellipsoids @1080p, plain vs proposal at ≤1e-4 color err"
HBM_PEAK_GBS = 8000.0
# proposal network: pinball quantile of the training loss and the factor on its depth; the best of quantiles
# 0.5 / 0.75 / 0.9 x margins 1.0 / 1.25 / 1.5 / 2.0 measured on this workload (a short proposal costs a second round)
PROP_TAU = 0.9
PROP_MARGIN = 2.0
P_BYTES = 56  # per ray-primitive interaction record: mean 12 + scale 12 + quat 16 + density 4 + colour 12
G_BYTES = 56  # gradient record of the same fields


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--prims", type=int, default=1_000_000)
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--kbuf", type=int, default=19)
    ap.add_argument("--max-iters", type=int, default=500)
    ap.add_argument("--cpu-sample", type=int, default=131072, help="rays traced by the CPU baseline (0: skip)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-proposal", action="store_true", help="skip the learned-proposal form")
    ap.add_argument("--no-inference", action="store_true", help="skip the inference-renderer frame timing")
    ap.add_argument("--mlp-epochs", type=int, default=12)
    ap.add_argument("--hbm-probe-only", action="store_true",
                    help="only the 1 GiB copy (a kernel of known traffic, for calibrating the PMC byte counters)")
    return ap.parse_args()


def algorithmic_bytes(out, nr, n, d, max_iters):
    """Algorithmic HBM bytes of one launch of the forward trace kernel and of the backward trace kernel.

    Per ray: origin + direction in (24 B) and colour, state, iteration count, start-inside density, FP64 state
    and processing order out (forward 44 + 68 B per ray as counted here), 4 B per stored event id; per distinct
    primitive a ray touches, one 56 B record read (and in the backward one 56 B gradient written). Distinct
    primitives per ray are bounded below by max(#entry events, #exit events), so the figures are lower bounds.
    The SH prepass (part of the scene build) is counted separately."""
    import torch
    iters = out["iters"].long()
    if out.get("event_layout") == "slot":  # column s of the event list belongs to ray ray_order[s]
        iters = iters[out["ray_order"].long()]
    tri = out["tri_collection"].view(max_iters, nr)
    E = int(iters.sum())
    ent = torch.zeros(nr, dtype=torch.long, device=iters.device)
    ext = torch.zeros(nr, dtype=torch.long, device=iters.device)
    kmax = int(iters.max()) if nr > 0 else 0
    for k0 in range(0, kmax, 64):
        k1 = min(kmax, k0 + 64)
        blk = tri[k0:k1]
        valid = torch.arange(k0, k1, device=iters.device).view(-1, 1) < iters.view(1, -1)
        face = (blk & 1).bool()
        ent += (valid & face).sum(0)
        ext += (valid & ~face).sum(0)
    U = int(torch.maximum(ent, ext).sum())
    fwd = nr * 44 + U * P_BYTES + nr * 68 + 4 * E
    bwd = nr * (44 + 68) + 4 * E + U * (P_BYTES + G_BYTES)
    return dict(fwd=fwd, bwd=bwd, events=E, distinct=U)


def measured_traffic(kernel):
    """HBM bytes per launch of a trace kernel from the committed rocprofv3 PMC summary (FETCH_SIZE and
    WRITE_SIZE, each scaled by the reading of a plain copy kernel of known size), or None."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            t = json.load(f)
        return int(t[kernel]["hbm_bytes_per_launch"])
    except (OSError, KeyError, ValueError):
        return None


def cpu_baseline(sc, o, d, eye, sample, max_iters, nr_full):
    """FP32 C++ restatement of the reference (oracle, OpenMP, CPU BVH) on a deterministic ray subsample;
    the per-ray time is extrapolated to the full image and the BVH build counted once."""
    import numpy as np

    from oracle import oracle
    from oracle import quadrature as Q
    stride = max(1, o.shape[0] // sample)
    idx = np.arange(0, o.shape[0], stride)[:sample]
    means = sc["means"].numpy()
    t0 = time.time()
    rgb = oracle.eval_sh(means, sc["features"].numpy(), 3, np.asarray(eye))  # per-primitive SH prepass
    rgb = np.logaddexp(0.0, 10.0 * rgb) / 10.0                                # softplus(beta=10)
    feats = Q.RGB2SH(rgb).reshape(-1, 1, 3).astype(np.float32)
    t_sh = time.time() - t0
    args = [means, sc["scales"].numpy(), sc["quats"].numpy(), sc["densities"].numpy(), feats]
    ro, rd = o.numpy()[idx], d.numpy()[idx]
    t0 = time.time()
    oracle.forward(*args, ro[:1], rd[:1], 0.0, 100.0, max_iters=max_iters, use_bvh=True, dtype=np.float32)
    t_build = time.time() - t0
    t0 = time.time()
    f = oracle.forward(*args, ro, rd, 0.0, 100.0, max_iters=max_iters, use_bvh=True, dtype=np.float32)
    t_fwd = time.time() - t0 - t_build
    g = np.random.default_rng(1).normal(size=(len(idx), 5)).astype(np.float32)
    t0 = time.time()
    oracle.backward(*args, ro, rd, 0.0, max_iters, g, f, dtype=np.float32)
    t_bwd = time.time() - t0
    per_ray = max(t_fwd, 0.0) / len(idx) + t_bwd / len(idx)
    total = t_sh + t_build + per_ray * nr_full
    # the OpenMP threads that ran, bounded by the cores this process may use
    cores = min(oracle.num_threads(), len(os.sched_getaffinity(0)))
    return dict(value=round(nr_full / total / 1e6, 6), unit="Mrays/s", cores=cores, kind="port",
                sample=(f"{len(idx)} of {nr_full} rays (every {stride}th pixel), FP32, CPU BVH build "
                        f"{t_build:.2f}s + SH prepass {t_sh:.2f}s counted once, per-ray fwd+bwd "
                        f"{1e3 * per_ray:.3f} ms extrapolated to the full image"))


def hbm_copy_rate(dev, gib=1.0, reps=5):
    """Box-measured HBM bandwidth: event-timed device-to-device copies of a 1 GiB buffer (larger than the L2 and
    the 256 MB last-level cache together), read + write bytes per second."""
    import torch
    n = int(gib * 2**30) // 4
    src = torch.ones(n, dtype=torch.float32, device=dev)
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        dst.copy_(src)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    del src, dst
    return {"gbs": round(2 * n * 4 / (ms * 1e-3) / 1e9, 1), "bytes_per_copy": 2 * n * 4, "ms": round(ms, 4)}


MLP_FLOPS_PER_RAY = 2 * (63 * 128 + 128 * 128 + 128 * 128 + 128)  # 81,920 (inference)
BF16_PEAK_TFLOPS = 2500.0


def proposal_form(a, params, og, dg, grad, eye, nr, plain_value):
    """The learned-proposal form: the MLP is trained on the GPU (outside the timed region) on (ray, saturation
    depth) pairs of the plain tracer from 64 random cameras at radius 3-5 (256x256 each, 4.2M pairs); each timed
    step then runs MLP inference + proposal-bounded forward + backward."""
    import torch

    from ever_amd import proposal as PR
    from ever_amd import tracer as T
    from ever_amd.splinetracers import fast_ellipsoid_splinetracer as fes
    dev = og.device
    t0 = time.perf_counter()
    scene = T.Scene(*[x.detach() for x in params], color_mode=1, sh_origin=eye, sh_degree=3, activation=1).build()
    ro, rd, tg = PR.make_pairs(scene, n_cameras=64, width=256, height=256)
    torch.cuda.synchronize()
    t_pairs = time.perf_counter() - t0
    net = PR.ProposalNet(dev, seed=0)
    t0 = time.perf_counter()
    loss = PR.fit(net, ro, rd, tg, epochs=a.mlp_epochs, batch=65536, lr=3e-3, tau=PROP_TAU)
    torch.cuda.synchronize()
    t_fit = time.perf_counter() - t0
    del ro, rd, tg

    def step():
        prop = net.proposal(og, dg, margin=PROP_MARGIN)
        out = fes.trace_rays_sh(*params, og, dg, eye, 3, 0.0, 100.0, max_iters=a.max_iters, proposal=prop)
        torch.autograd.backward(out, grad)
        for p in params:
            p.grad = None

    for _ in range(a.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        step()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    # inference alone, event-timed on the current stream
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(5):
        prop = net.proposal(og, dg, margin=PROP_MARGIN)
    e1.record()
    torch.cuda.synchronize()
    infer_ms = e0.elapsed_time(e1) / 5
    # rendering (forward only, no event list, no touch counts): plain walk vs network + proposal-bounded walk
    def render(use_prop):
        p_ = net.proposal(og, dg, margin=PROP_MARGIN) if use_prop else None
        return T.trace_forward(scene, og, dg, 0.0, 100.0, a.max_iters, store_events=False, count_touches=False,
                               proposal=p_)["color"]
    render_ms = {}
    for key, use_prop in (("plain", False), ("proposal", True)):
        render(use_prop)
        e0.record()
        for _ in range(5):
            render(use_prop)
        e1.record()
        torch.cuda.synchronize()
        render_ms[key] = e0.elapsed_time(e1) / 5
    with torch.no_grad():
        plain = fes.trace_rays_sh(*params, og, dg, eye, 3, 0.0, 100.0, max_iters=a.max_iters)
        bounded = fes.trace_rays_sh(*params, og, dg, eye, 3, 0.0, 100.0, max_iters=a.max_iters, proposal=prop)
    ideal_depth = PR.saturation_depth(T.trace_forward(scene, og, dg, 0.0, 100.0, a.max_iters, store_events=False))
    sat = ideal_depth < PR.BACKGROUND_DEPTH
    rel = ((prop[sat] - ideal_depth[sat]) / ideal_depth[sat]).abs()
    # the ceiling of any proposal: the plain trace's own saturation depths (x1.02) as proposals, rendering only
    exact = torch.where(sat, ideal_depth * 1.02, torch.full_like(ideal_depth, float("inf"))).contiguous()

    def render_exact():
        return T.trace_forward(scene, og, dg, 0.0, 100.0, a.max_iters, store_events=False, count_touches=False,
                               proposal=exact)["color"]
    render_exact()
    e0.record()
    for _ in range(5):
        render_exact()
    e1.record()
    torch.cuda.synchronize()
    exact_ms = e0.elapsed_time(e1) / 5
    value = nr * a.steps / dt / 1e6
    achieved = MLP_FLOPS_PER_RAY * nr / (infer_ms * 1e-3) / 1e12
    return {"value": round(value, 4), "unit": "Mrays/s", "ms_per_step": round(1e3 * dt / a.steps, 3),
            "vs_plain": round(value / plain_value, 4),
            "color_max_abs_diff_vs_plain": float((plain - bounded).abs().max()),
            "proposal_rel_err_median": round(float(rel.median()), 4),
            "proposal_short_frac": round(float((prop[sat] < ideal_depth[sat]).float().mean()), 4),
            "render": {"plain_ms": round(render_ms["plain"], 4), "proposal_ms": round(render_ms["proposal"], 4),
                       "speedup": round(render_ms["plain"] / render_ms["proposal"], 4),
                       "exact_depth_ms": round(exact_ms, 4),
                       "exact_depth_speedup": round(render_ms["plain"] / exact_ms, 4),
                       "note": "forward only (no event list), network inference included; exact_depth: the plain "
                               "trace's saturation depths x1.02 as proposals (no network), the ceiling of any "
                               "proposal"},
            "mlp": {"inference_ms": round(infer_ms, 4), "flops_per_ray": MLP_FLOPS_PER_RAY,
                    "achieved_tflops": round(achieved, 2), "peak_tflops": BF16_PEAK_TFLOPS,
                    "mfma_util": round(achieved / BF16_PEAK_TFLOPS, 5), "dtype": "bf16 x bf16 -> f32"},
            "training": {"pairs": 64 * 256 * 256, "pairs_s": round(t_pairs, 3), "fit_s": round(t_fit, 3),
                         "epochs": a.mlp_epochs, "final_loss": round(loss, 5), "tau": PROP_TAU,
                         "margin": PROP_MARGIN}}


def inference_frames(params, og, dg, eye, nr, frames=10):
    """The inference renderer's frame (integration_utils/gaussian_renderer/fast_renderer.py:110-131 on this
    library: ever_amd.render.FastRenderer) on the same scene and camera: the tree built once, then per frame the
    colours rewritten for the camera (fused SH + softplus pass) and a forward trace without an event list, 200
    events at most per ray as the reference's renderer asks."""
    import torch

    from ever_amd import tracer as T
    with torch.no_grad():
        t = [x.detach() for x in params]
        scene = T.Scene(*t, color_mode=1, sh_origin=eye, sh_degree=3, activation=1).build()

        def frame():
            scene.update_features(t[4], sh_origin=eye, sh_degree=3)
            return T.trace_forward(scene, og, dg, 0.0, 1e7, 200, store_events=False, count_touches=False)

        for _ in range(2):
            frame()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(frames):
            frame()
        torch.cuda.synchronize()
        ms = 1e3 * (time.perf_counter() - t0) / frames
    return {"ms_per_frame": round(ms, 3), "fps": round(1e3 / ms, 2), "Mrays_s": round(nr / ms / 1e3, 2),
            "frames": frames, "note": "tree built once; per frame: fused recolouring + forward, no event list"}


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {a.gpus}); several GPUs would each trace a share of "
              "the rays (image tiles or whole cameras) against a replica of the scene",
              file=sys.stderr)
        sys.exit(2)
    import torch

    from ever_amd import _capi as C
    from ever_amd import scenes
    from ever_amd import tracer as T
    from ever_amd.splinetracers import fast_ellipsoid_splinetracer as fes

    dev = torch.device("cuda:0")
    if a.hbm_probe_only:
        print(json.dumps({"hbm_copy": hbm_copy_rate(dev)}), flush=True)
        return
    T.set_kbuf(a.kbuf)
    sc, o, d = scenes.config3(width=a.width, height=a.height, n=a.prims)
    eye = (0.0, 0.0, -4.0)
    nr = o.shape[0]
    D = {k: v.to(dev).contiguous() for k, v in sc.items()}
    params = [D[k].requires_grad_(True) for k in ("means", "scales", "quats", "densities", "features")]
    og, dg = o.to(dev), d.to(dev)
    grad = scenes.upstream_grad(nr, seed=1).to(dev)

    def step():
        out = fes.trace_rays_sh(*params, og, dg, eye, 3, 0.0, 100.0, max_iters=a.max_iters)
        torch.autograd.backward(out, grad)
        for p in params:
            p.grad = None

    for _ in range(a.warmup):
        step()
    torch.cuda.synchronize()
    timer = T.StageTimer()
    T.stage_timer = timer
    L = C.lib()
    C.check(L.ever_profile_enable(1), "ever_profile_enable")  # HIP events around the two trace kernels
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        step()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    C.check(L.ever_profile_enable(0), "ever_profile_enable")
    T.stage_timer = None
    stages = timer.summary()
    kernel_ms = {}
    for name, which in (("forward", 0), ("backward", 1)):
        tot, cnt = ctypes.c_double(), ctypes.c_int64()
        C.check(L.ever_profile_read(which, ctypes.byref(tot), ctypes.byref(cnt)), "ever_profile_read")
        kernel_ms[name] = tot.value / max(cnt.value, 1)

    # one untimed forward for the algorithmic byte count of this workload, configured as the step's forward (events
    # by processing slot, no touch counts); the counts below do not depend on the event-list layout
    scene = T.Scene(*[x.detach() for x in params], color_mode=1, sh_origin=eye, sh_degree=3, activation=1).build()
    out = T.trace_forward(scene, og, dg, 0.0, 100.0, a.max_iters, count_touches=False, event_layout="slot")
    ab = algorithmic_bytes(out, nr, a.prims, 16, a.max_iters)
    fwd_ms, bwd_ms = kernel_ms["forward"], kernel_ms["backward"]
    dom, dom_ms, dom_bytes = ("forward", fwd_ms, ab["fwd"]) if fwd_ms >= bwd_ms else ("backward", bwd_ms, ab["bwd"])
    achieved = dom_bytes / (dom_ms * 1e-3) / 1e9
    traffic = measured_traffic(dom)
    wave = T.get_traversal() == "wave" or (T.get_traversal() == "auto" and nr >= 65536)
    fwd_kernel = f"k_trace_forward{'_wave' if wave else ''}<{a.kbuf}>"
    rec = {
        "metric": METRIC,
        "value": round(nr * a.steps / dt / 1e6, 4),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(1e3 * dt / a.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "1M clustered ellipsoids (64-GMM, aspect<=100, SH3) @1920x1080, fwd+bwd, LBVH "
                               "rebuilt per step, plain traversal", "prims": a.prims, "rays": nr, "kbuf": a.kbuf,
                   "max_iters": a.max_iters, "sh_degree": 3},
        "stages_ms": {k: round(v, 3) for k, v in stages.items()},
        "kernel_ms": {k: round(v, 3) for k, v in kernel_ms.items()},
        "events_per_ray": round(ab["events"] / nr, 3),
        "roofline": {"bound": "hbm", "kernel": fwd_kernel if dom == "forward" else "k_trace_backward",
                     "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": traffic,
                     "algorithmic_bytes": int(dom_bytes), "launch_ms": round(dom_ms, 4),
                     "peak_measured_copy": hbm_copy_rate(dev)},
    }
    if not a.no_inference:
        rec["inference"] = inference_frames(params, og, dg, eye, nr)
    if not a.no_proposal:
        rec["proposal"] = proposal_form(a, params, og, dg, grad, eye, nr, rec["value"])
    if not a.no_cpu and a.cpu_sample > 0:
        rec["cpu_baseline"] = cpu_baseline(sc, o, d, eye, a.cpu_sample, a.max_iters, nr)
    print(json.dumps(rec), flush=True)


if __name__ == "__main__":
    main()
