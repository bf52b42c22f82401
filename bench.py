#!/usr/bin/env python
"""Benchmark: lnlike+grad evaluations/s on configuration 4 (512 uneven
patches: sources 1+Poisson(8) capped at 30, bands U{1..10}, exposures per
band U{1..16}, square tiles of side U{32..256}, 3-6 PSF Gaussians per band),
batched (one persistent launch over all patches, parameters -> proposals ->
image Gaussians -> pixel pass -> FP64 reduction -> stacked gradient, all on
the device) versus the plain single-patch drop-in (one synchronous
``HIPProposer.evaluate_proposal`` per patch).

One step = one lnlike+gradient evaluation of every patch.  ``value`` is
patch-evaluations per second of the batched path (inputs resident in HBM).
Prints one JSON line on stdout; progress goes to stderr.
"""

import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

METRIC = ("lnlike+grad evals/sec & Gaussian-pixel evals/s, batched vs plain, "
          "chi2 err<=1e-6")
PEAK_FP32_TFLOPS = 157.3   # MI355X FP32 vector peak (published)
FLOP_PER_PAIR = 98         # algorithmic FLOPs per (pixel, image Gaussian) pair
FLOP_PER_CUT_PAIR = 11     # dx, dy, v, exp argument of a pair beyond the cutoff
TRAFFIC_SUMMARY = os.path.join(HERE, "profiles", "pixel_pass_config4_traffic.txt")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--patches", type=int, default=512)
    ap.add_argument("--plain-sweeps", type=int, default=1)
    ap.add_argument("--skip-plain", action="store_true")
    ap.add_argument("--skip-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=20.0,
                    help="target length of the CPU baseline sample")
    ap.add_argument("--launch", type=str, default=None,
                    help="K,W,persistent,groups_per_cu (default: autotuner cache)")
    ap.add_argument("--precision", default="mixed", choices=["fp32", "fp64_reduce", "mixed"])
    ap.add_argument("--tune", type=str, default=None, metavar="CACHE_JSON",
                    help="run the launch autotuner on this workload, write the "
                         "cache to CACHE_JSON and exit")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 1 GiB device copy (calibrates the HBM "
                         "byte counters FETCH_SIZE / WRITE_SIZE) and exit")
    return ap.parse_args()


def copy_calibration():
    """A plain device-to-device copy of 1 GiB, five times: every dispatch reads
    and writes exactly 2^30 bytes."""
    import torch
    src = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(5):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes": src.numel() * 4, "dispatches": 5}), flush=True)


def measured_traffic(launch):
    """HBM bytes per pixel-pass launch from the committed rocprofv3
    byte-counter summary, if it was taken on this kernel source and launch
    shape; else (None, reason)."""
    from forcepho_amd import profiling
    if not os.path.exists(TRAFFIC_SUMMARY):
        return None, "no byte-counter summary"
    kv = profiling.read_summary(TRAFFIC_SUMMARY)
    shape = ",".join(str(launch[k]) for k in ("pix_per_lane", "waves_per_group",
                                              "persistent", "groups_per_cu"))
    if kv.get("kernel_source_sha256_16") != profiling.kernel_source_digest():
        return None, "byte-counter summary taken on other kernel source"
    if kv.get("launch") != shape:
        return None, f"byte-counter summary taken at launch {kv.get('launch')}, not {shape}"
    return float(kv["traffic_bytes_per_launch"]), os.path.relpath(TRAFFIC_SUMMARY, HERE)


def algorithmic_counts(specs):
    """Dense pairs, pixels and FLOPs over unpadded pixels (per evaluation of
    every patch): F = 98 N_pairs + 7 N_pix n_src + 6 N_pix."""
    pairs = npix = flops = 0
    for s in specs:
        n_pix = s.npix_unpadded
        p = s.n_pairs()
        pairs += p
        npix += n_pix
        flops += FLOP_PER_PAIR * p + 7 * n_pix * s.n_sources + 6 * n_pix
    return pairs, npix, flops


def build_patches(specs, log_every=64):
    """Upload every patch with data = noise, render the truth on the device
    with negated fluxes (the model is linear in flux) so that the residual
    buffer holds noise + model, and make it the data (swap_on_device)."""
    import ctypes
    from forcepho_amd import _lib, synth
    from forcepho_amd.limits import source_struct_dtype
    out = []
    for i, spec in enumerate(specs):
        patch, scene, theta = synth.make_patch(spec, config=4, patch_id=i,
                                               return_residual=True)
        rng = synth.rng_for(4, i, 2)
        noise = np.zeros(patch.npix, dtype=np.float32)
        for e in range(len(patch.exposure_N)):
            b = int(np.searchsorted(np.cumsum(patch.band_N), e, side="right"))
            s0, n = patch.exposure_start[e], patch.original_sizes[e]
            noise[s0:s0 + n] = rng.standard_normal(n, dtype=np.float32) * \
                np.float32(patch.sigma_band[b])
        patch.data = noise
        patch.send_to_device()
        neg = scene.get_proposal().copy()
        neg["fluxes"] *= -1.0
        nb, ns = len(patch.band_N), patch.n_sources
        c = np.zeros(nb)
        d = np.zeros((nb, ns, 7), dtype=np.float32)
        neg = np.ascontiguousarray(neg, dtype=source_struct_dtype)
        _lib.check(_lib.lib().fpho_evaluate(patch.device_patch, ctypes.c_void_p(neg.ctypes.data),
                                            _lib.fptr(c, ctypes.c_double),
                                            _lib.fptr(d, ctypes.c_float)))
        patch.swap_on_device()
        patch.data = None  # host copy no longer describes the device data
        out.append((patch, scene, theta))
        if (i + 1) % log_every == 0:
            log(f"  built {i + 1}/{len(specs)} patches")
    return out


def cpu_baseline(patches, specs, thetas, seconds):
    """The reference's own CPU kernel (compiled from its sources into
    oracle/_ref), one thread, on a seeded random subset of the patches whose
    pair count is sized for about ``seconds`` of work."""
    import oracle
    if oracle.reference_module() is None:
        return None
    from forcepho_amd import synth
    rng = np.random.default_rng(1234)
    order = rng.permutation(len(patches))
    target_pairs = seconds * 6.0e7
    sample, pairs = [], 0
    for i in order:
        if pairs >= target_pairs:
            break
        sample.append(i)
        pairs += specs[i].n_pairs()
    t_total = 0.0
    evals = 0
    for i in sample:
        patch, scene, theta = patches[i]
        patch.data = patch.retrieve_array("data")
        scene.set_all_source_params(thetas[i])
        prop = scene.get_proposal().copy()
        ref = oracle.ReferenceKernel(patch)
        t0 = time.perf_counter()
        ref.evaluate(prop)
        t_total += time.perf_counter() - t0
        evals += 1
        patch.data = None
    pairs_per_s = pairs / t_total
    return dict(pairs_per_s=pairs_per_s, seconds=t_total, patches=len(sample),
                sample_pairs=pairs, evals=evals)


def main():
    args = parse()
    if args.copy_calibration:
        copy_calibration()
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU runs are not "
                          "part of it (patches would be sharded across GPUs, see "
                          "DESIGN.md)", "n_gpus": args.gpus}))
        sys.exit(2)

    import torch
    from forcepho_amd import _lib, synth
    from forcepho_amd.autotune import load_or_default
    from forcepho_amd.batch import DeviceSplines, PatchBatch

    _lib.init(0)
    torch.cuda.init()
    t_setup = time.time()
    specs = synth.config4_specs(args.patches)
    pairs, npix, flops = algorithmic_counts(specs)
    log(f"config 4: {len(specs)} patches, {npix:.3e} px, {pairs:.3e} pairs, "
        f"{flops:.3e} algorithmic FLOP per sweep")
    patches = build_patches(specs)
    log(f"setup {time.time() - t_setup:.1f}s")

    batch = PatchBatch([p for p, _, _ in patches], precision=args.precision)
    if args.launch:
        K, W, P, G = (int(x) for x in args.launch.split(","))
        launch = dict(pix_per_lane=K, waves_per_group=W, persistent=P, groups_per_cu=G)
    else:
        launch = load_or_default("batch")
    batch.set_launch(**launch)
    log(f"launch {launch}")
    splines = DeviceSplines()
    thetas = [synth.make_proposals(th, len(p.band_N), 1, config=4, patch_id=i)[0]
              for i, (p, sc, th) in enumerate(patches)]
    theta = torch.as_tensor(np.concatenate(thetas), dtype=torch.float64, device="cuda")

    if args.tune:
        from forcepho_amd import autotune
        autotune.CACHE = args.tune
        best, table = autotune.tune(batch, lambda: batch.lnlike_and_grad(theta, splines),
                                    reps=3, log=log)
        autotune.save("batch", best, table[0][0])
        log(f"best launch {best}: {table[0][0]:.3f} ms")
        print(json.dumps({"best": best, "ms": table[0][0],
                          "table": [[ms, c] for ms, c in table]}))
        return

    # ---- batched: warm-up, then exactly `steps` timed sweeps
    for _ in range(args.warmup):
        batch.lnlike_and_grad(theta, splines)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        batch.lnlike_and_grad(theta, splines)
    torch.cuda.synchronize()
    t_batched = (time.perf_counter() - t0) / args.steps
    evals_per_s = len(patches) / t_batched
    log(f"batched: {t_batched * 1e3:.2f} ms/sweep, {evals_per_s:.1f} evals/s, "
        f"{pairs / t_batched:.3e} pairs/s")

    # ---- dominant kernel (pixel pass) timed with HIP events on its stream
    batch.set_timing(True)
    stage = np.zeros(3)
    n_t = max(2, min(args.steps, 5))
    for _ in range(n_t):
        batch.lnlike_and_grad(theta, splines)
        stage += np.array(batch.last_timing())
    batch.set_timing(False)
    stage /= n_t
    eval_ms = stage[1]
    # The reference algorithm evaluates a pair fully only inside the cutoff
    # (exp argument <= 36, common_kernel.cc:73,108) and stops after the exp
    # argument otherwise: 98 FLOP per in-cutoff pair, 11 per other pair.
    n_in = batch.count_pairs_in_cutoff()
    flops_alg = flops - FLOP_PER_PAIR * pairs + FLOP_PER_PAIR * n_in + \
        FLOP_PER_CUT_PAIR * max(pairs - n_in, 0)
    achieved = flops_alg / (eval_ms * 1e-3) / 1e12
    log(f"pairs inside the cutoff: {n_in:.4e} of {pairs:.4e} ({n_in / pairs:.3f}); "
        f"algorithmic {flops_alg:.3e} FLOP per sweep")
    log(f"stages (ms): convert {stage[0]:.3f} pixel pass {stage[1]:.3f} reduce {stage[2]:.3f}")

    # ---- chi2 error on a sample of the smallest patches vs the FP64 oracle
    chi2_err = None
    try:
        import oracle
        from forcepho_amd.limits import source_struct_dtype
        props = batch.proposals.cpu().numpy().view(source_struct_dtype)
        chi2_dev = batch.chi2_host()
        small = np.argsort([s.n_pairs() for s in specs])[:3]
        errs = []
        for i in small:
            p, sc, th = patches[i]
            p.data = p.retrieve_array("data")
            prop = props[batch.src_off[i]:batch.src_off[i + 1]]
            c_ref, _ = oracle.evaluate(p, prop)
            errs.append(np.max(np.abs(chi2_dev[i] - c_ref) / np.abs(c_ref)))
            p.data = None
        chi2_err = float(max(errs))
    except ImportError:
        pass

    # ---- plain single-patch drop-in
    plain = None
    if not args.skip_plain:
        from forcepho_amd.proposal import HIPProposer
        props_host = []
        for i, (p, sc, th) in enumerate(patches):
            sc.set_all_source_params(thetas[i])
            props_host.append(sc.get_proposal().copy())
        prop_er = HIPProposer()
        for (p, sc, th), pr in zip(patches, props_host):    # warm-up (plans, staging)
            prop_er.evaluate_proposal(pr, patch=p)
        t0 = time.perf_counter()
        for _ in range(args.plain_sweeps):
            for (p, sc, th), pr in zip(patches, props_host):
                prop_er.evaluate_proposal(pr, patch=p)
        t_plain = (time.perf_counter() - t0) / args.plain_sweeps
        plain = len(patches) / t_plain
        log(f"plain: {t_plain * 1e3:.1f} ms/sweep, {plain:.1f} evals/s")

    cpu = None
    if not args.skip_cpu:
        c = cpu_baseline(patches, specs, thetas, args.cpu_seconds)
        if c is not None:
            t_sweep = pairs / c["pairs_per_s"]
            cpu = {"value": len(patches) / t_sweep, "unit": "evals/s", "cores": 1,
                   "kind": "reference",
                   "sample": (f"forcepho's compute_gaussians_kernel.cc compiled from source "
                              f"(g++ -O3), 1 thread, {c['patches']} randomly chosen patches "
                              f"of the workload ({c['sample_pairs']:.3e} pairs, "
                              f"{c['seconds']:.1f}s), scaled per pair to the full sweep"),
                   "pairs_per_s": c["pairs_per_s"]}
            log(f"cpu reference: {c['pairs_per_s']:.3e} pairs/s on {c['patches']} patches")

    traffic, traffic_src = measured_traffic(launch)
    out = {
        "metric": METRIC, "value": evals_per_s, "unit": "evals/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": t_batched * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"config 4: {len(patches)} uneven patches batched in one "
                               f"launch (lnlike+grad of every patch per step)",
                   "patches": len(patches), "npix": npix, "n_pairs": pairs,
                   "precision": args.precision, "launch": launch},
        "pairs_per_s": pairs / t_batched,
        "plain_evals_per_s": plain,
        "batched_over_plain": (evals_per_s / plain) if plain else None,
        "chi2_rel_err_sample": chi2_err,
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": PEAK_FP32_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / PEAK_FP32_TFLOPS, "traffic": traffic,
                     "traffic_source": traffic_src,
                     "kernel": "eval pixel pass (eval_kernel / eval_persistent_kernel)",
                     "algorithmic_flop_per_sweep": flops_alg,
                     "pairs_in_cutoff": n_in,
                     "note": ("compute-bound on the FP32 VALU (no dense contraction, so no "
                              "MFMA); peak = MI355X FP32 vector 157.3 TFLOP/s, of which "
                              "78.6 is issuable without packed FP32; algorithmic FLOP = "
                              "98 per pair inside the cutoff + 11 per other pair + "
                              "7 N_pix n_src + 6 N_pix")},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
