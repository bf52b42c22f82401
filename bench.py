#!/usr/bin/env python
"""Benchmark: lnlike+grad evaluations/s on configuration 4 (512 uneven
patches: sources 1+Poisson(8) capped at 30, bands U{1..10}, exposures per
band U{1..16}, square tiles of side U{32..256}, 3-6 PSF Gaussians per band),
batched (one launch of each stage over all patches, parameters -> proposals
-> image Gaussians -> pixel pass -> FP64 reduction -> stacked gradient, all
on the device, at the autotuned launch shape) versus the plain single-patch
drop-in (one synchronous ``HIPProposer.evaluate_proposal`` per patch).

One step = one lnlike+gradient evaluation of every patch.  ``value`` is
patch-evaluations per second of the batched path (inputs resident in HBM).
The JSON line also carries the roofline of the pixel pass (algorithmic FLOP
over its HIP-event time; HBM bytes per launch and GB/s from the committed
byte-counter summary when it was taken on this kernel source; the device's
measured exp-issue ceiling), the compiled reference CPU kernel on one core
and on all cores of the host, and the batched/plain ratio.
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
CPU_SHARE = 16             # host cores per GPU on the MI355X nodes this runs on


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--patches", type=int, default=None,
                    help="patches (default 512 for config 4, 4096 for config 5)")
    ap.add_argument("--plain-sweeps", type=int, default=1)
    ap.add_argument("--skip-plain", action="store_true")
    ap.add_argument("--skip-cpu", action="store_true")
    ap.add_argument("--skip-hmc", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=20.0,
                    help="target length of the CPU baseline sample")
    ap.add_argument("--launch", type=str, default=None,
                    help="K,W,persistent,groups_per_cu (default: autotuner cache)")
    ap.add_argument("--precision", default="mixed", choices=["fp32", "fp64_reduce", "mixed", "mixed_centre"])
    ap.add_argument("--tune", type=str, default=None, metavar="CACHE_JSON",
                    help="run the launch autotuner on this workload, write the "
                         "cache to CACHE_JSON and exit")
    ap.add_argument("--config", type=int, default=4, choices=[4, 5],
                    help="4: the benchmark line (default); 5: the mosaic-scale stress "
                         "run (4096 patches, precision ablation, HMC), one JSON line")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 1 GiB device copy (calibrates the HBM "
                         "byte counters FETCH_SIZE / WRITE_SIZE) and exit")
    ap.add_argument("--issue-calibration", action="store_true",
                    help="only measure the VALU issue rate of FP32 FMA (and FMA with "
                         "v_exp_f32) chains at 8 waves per SIMD, and exit")
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


def issue_calibration():
    """Wave64 VALU instructions per SIMD per cycle of independent FP32 FMA
    chains (8-byte and 4-byte encodings), of FMA chains interleaved with
    v_exp_f32, and of v_exp_f32 chains alone (measurement library; the
    shader clock is the device's nominal one)."""
    import measure
    from forcepho_amd import _lib
    _lib.init(0)
    cus = _lib.cu_count()
    out = {}
    for kind, name in ((0, "fma_f32"), (2, "fmac_f32"), (1, "fma_f32+exp_f32"),
                       (3, "exp_f32"), (4, "mul_f32"), (5, "add_f32"),
                       (6, "fmac_f32_vgpr_operands")):
        ms, wi, clk = measure.issue_calibration(kind)
        cycles = ms * 1e-3 * clk * 1e3
        out[name] = {"ms": ms, "wave_instructions": wi, "clock_mhz": clk / 1e3,
                     "per_simd_per_cycle": wi / (4 * cus * cycles)}
        if kind in (0, 2, 4, 5, 6):
            # FP32 FMA = 2 FLOP per lane, multiply or add 1
            flop = 2 if kind in (0, 2, 6) else 1
            out[name]["tflops"] = wi * 64 * flop / (ms * 1e-3) / 1e12
    print(json.dumps({"issue_calibration": out, "cus": cus}), flush=True)


def exp_issue_rate():
    """Lane v_exp_f32 evaluations per second of the device, from 16
    independent exp chains per lane at 8 waves per SIMD on every CU (the
    measurement library's calibration kernel, kind 3)."""
    import measure
    ms, wi, clk = measure.issue_calibration(3)
    return wi * 64 / (ms * 1e-3)


def host_cpu():
    """Model name and logical CPU count of this host, and the CPUs this
    process may run on."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"model": model, "nproc": os.cpu_count(),
            "affinity": len(os.sched_getaffinity(0))}


def measured_traffic(launch, precision="mixed"):
    """HBM bytes per pixel-pass launch from the committed rocprofv3
    byte-counter summary, if it was taken on this kernel source, launch
    shape and precision mode (the default, mixed); else (None, reason)."""
    import profiling
    if precision != "mixed":
        return None, f"byte-counter summary taken in the mixed mode, not {precision}"
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


def _noise(patch, config, i):
    """Seeded Gaussian noise of each band's sigma over the patch's real
    pixels (padding stays 0), drawn on the device (the host's generator
    would take minutes at the mosaic workload's 4.9e9 pixels) and copied to
    a host array for upload."""
    import torch
    seed = int(np.random.SeedSequence([config, i, 2]).generate_state(1)[0])
    gen = torch.Generator(device="cuda")
    gen.manual_seed(seed)
    noise = torch.randn(patch.npix, generator=gen, device="cuda", dtype=torch.float32)
    sigma = torch.zeros_like(noise)
    band_of = np.repeat(np.arange(len(patch.band_N)), patch.band_N)
    for e in range(len(patch.exposure_N)):
        s0, n = int(patch.exposure_start[e]), int(patch.original_sizes[e])
        sigma[s0:s0 + n] = float(np.float32(patch.sigma_band[band_of[e]]))
    return (noise * sigma).cpu().numpy()


def build_patches(specs, log_every=64, config=4, keep_host=None, prebuilt=None):
    """Upload every patch with data = noise, render the truth on the device
    with negated fluxes (the model is linear in flux) so that the residual
    buffer holds noise + model, and make it the data (swap_on_device);
    then stop residual writes, so that evaluations write nothing but their
    outputs.  Host pixel arrays are dropped after upload except for the
    patches in ``keep_host`` (None keeps all).  ``prebuilt`` maps patch
    indices to (patch, scene, theta) built by the caller with host data;
    those are uploaded as they are."""
    import ctypes
    from forcepho_amd import _lib, synth
    from forcepho_amd.limits import source_struct_dtype
    prebuilt = prebuilt or {}
    out = []
    for i, spec in enumerate(specs):
        if i in prebuilt:
            patch, scene, theta = prebuilt[i]
            patch.return_residual = False
            patch.send_to_device()
            out.append((patch, scene, theta))
            continue
        patch, scene, theta = synth.make_patch(spec, config=config, patch_id=i,
                                               return_residual=True)
        patch.data = _noise(patch, config, i)
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
        patch.set_residual_writes(False)
        patch.data = None  # host copy no longer describes the device data
        if keep_host is not None and i not in keep_host:
            patch.ierr = patch.xpix = patch.ypix = None
        out.append((patch, scene, theta))
        if (i + 1) % log_every == 0:
            log(f"  built {i + 1}/{len(specs)} patches")
    return out


def time_plain(patch_list, props_host, sweeps=1):
    """Seconds per sweep of synchronous ``HIPProposer.evaluate_proposal``
    calls, one per patch, as a fitting loop makes them (no residual
    download: the residual stays on the device)."""
    from forcepho_amd.proposal import HIPProposer
    prop_er = HIPProposer()
    for p, pr in zip(patch_list, props_host):    # warm-up (plans, staging)
        prop_er.evaluate_proposal(pr, patch=p)
    t0 = time.perf_counter()
    for _ in range(sweeps):
        for p, pr in zip(patch_list, props_host):
            prop_er.evaluate_proposal(pr, patch=p)
    return (time.perf_counter() - t0) / sweeps


def time_hmc(patches, batch, splines, thetas, n_traj=3, n_steps=10, eps=1e-3):
    """Graph-captured HMC over every patch of the batch: one trajectory of
    ``n_steps`` leapfrog steps (each a full lnlike+gradient of every patch
    through the sampling-space transform) recorded in a HIP graph and
    replayed ``n_traj`` times."""
    import torch
    from forcepho_amd import synth
    from forcepho_amd.hmc import HMC, DevicePosterior
    lo, hi = [], []
    for (p, sc, th) in patches:
        l, h = synth.theta_bounds(th, len(p.band_N))
        lo.append(l)
        hi.append(h)
    lo, hi = np.concatenate(lo), np.concatenate(hi)
    post = DevicePosterior(batch, splines, lo, hi)
    z0 = post.inverse_transform(np.clip(np.concatenate(thetas), lo + 1e-9, hi - 1e-9))
    hmc = HMC(post, z0, step_size=eps, n_steps=n_steps, seed=5)
    hmc.capture(warmup=1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    hmc.run(n_traj)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    return {"evals_per_s": len(patches) * n_steps * n_traj / dt,
            "ms_per_leapfrog_step": dt / (n_steps * n_traj) * 1e3,
            "trajectories": n_traj, "leapfrog_steps": n_steps, "step_size": eps,
            "acceptance": float(hmc.n_accept.mean().item()) / (n_traj + 1)}


# ------------------------------------------------------------------- CPU leg
# Everything in bench.py that reads oracle/ is in this leg: the reference's
# own CPU kernel (compiled into oracle/_ref) timed as the baseline, and the
# FP64 oracle as the checker of the sweep's results on a sample of patches.
# Nothing the benchmark times runs through it; --skip-cpu skips it.

def oracle_chi2_check(patches, specs, thetas, batch):
    """Relative chi2 error of the last sweep on the three smallest patches
    against the FP64 oracle fed the FP64 source records the sweep converts
    (scipy splines), or None without the oracle."""
    try:
        import oracle
    except ImportError:
        return None
    chi2_dev = batch.chi2_host()
    small = [int(i) for i in np.argsort([s.n_pairs() for s in specs])[:3]]
    errs = []
    for i in small:
        p = patches[i][0]
        p.data = p.retrieve_array("data")
        c_ref, _ = oracle.evaluate64(p, oracle.sources_from_theta(thetas[i], len(p.band_N)))
        errs.append(np.max(np.abs(chi2_dev[i] - c_ref) / np.abs(c_ref)))
        p.data = None
    return {"max_rel_err": float(max(errs)), "patches": small,
            "n_pairs": [specs[i].n_pairs() for i in small],
            "note": "the three smallest patches, FP64 oracle on the host; the largest, "
                    "median and most-sourced patches are checked against committed oracle "
                    "fixtures by the GPU tests"}


def oracle_reference(patches, thetas, sample):
    """FP64 oracle chi2 and dchi2 of the sample patches at their parameter
    vectors (FP64 source records, as the sweep converts), {index: (chi2,
    dchi2)}."""
    import oracle
    ref = {}
    for i in sample:
        p = patches[i][0]
        p.data = p.retrieve_array("data")
        t0 = time.time()
        ref[i] = oracle.evaluate64(p, oracle.sources_from_theta(thetas[i], len(p.band_N)))
        p.data = None
        log(f"oracle patch {i}: {time.time() - t0:.1f}s")
    return ref


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


# patch attributes the compiled reference kernel reads (oracle.ReferenceKernel)
_REF_FIELDS = ("data", "ierr", "xpix", "ypix", "band_start", "band_N", "exposure_start",
               "exposure_N", "rad2", "D", "crpix", "crval", "CW", "G", "n_psf_per_source",
               "psfgauss_start", "psfgauss", "n_sources", "n_radii")


def _ref_worker(units):
    """Evaluate (patch fields, proposal) units with the compiled reference
    kernel in this process; returns the compute seconds."""
    import types
    import oracle
    t = 0.0
    for fields, prop in units:
        ref = oracle.ReferenceKernel(types.SimpleNamespace(**fields))
        t0 = time.perf_counter()
        ref.evaluate(prop)
        t += time.perf_counter() - t0
    return t


def cpu_baseline_all_cores(patches, specs, thetas, seconds, workers):
    """The compiled reference kernel on ``workers`` processes (the pybind11
    kernel holds the GIL, so threads would not run concurrently): a seeded
    random sample of patches sized for about ``seconds`` of compute per
    worker, assigned longest-first to the workers; throughput = sample pairs
    / the slowest worker's compute time (start-up and data transfer
    excluded)."""
    import multiprocessing
    from concurrent.futures import ProcessPoolExecutor
    import oracle
    if oracle.reference_module() is None:
        return None
    rng = np.random.default_rng(4321)
    order = rng.permutation(len(patches))
    target = seconds * 1.6e8 * workers
    sample, pairs = [], 0
    for i in order:
        if pairs >= target or len(sample) >= 4 * workers:
            break
        sample.append(int(i))
        pairs += specs[i].n_pairs()
    sample.sort(key=lambda i: -specs[i].n_pairs())
    bins = [[] for _ in range(workers)]
    load = [0] * workers
    for i in sample:
        w = int(np.argmin(load))
        patch, scene, _ = patches[i]
        fields = {}
        for name in _REF_FIELDS:
            fields[name] = (patch.retrieve_array("data") if name == "data"
                            else getattr(patch, name))
        scene.set_all_source_params(thetas[i])
        bins[w].append((fields, scene.get_proposal().copy()))
        load[w] += specs[i].n_pairs()
    ctx = multiprocessing.get_context("spawn")
    with ProcessPoolExecutor(workers, mp_context=ctx) as ex:
        list(ex.map(_ref_worker, [[] for _ in range(workers)]))   # start every worker
        times = list(ex.map(_ref_worker, [b for b in bins if b]))
    return dict(pairs_per_s=pairs / max(times), seconds=max(times), patches=len(sample),
                sample_pairs=pairs, workers=workers)


def main():
    args = parse()
    if args.copy_calibration:
        copy_calibration()
        return
    if args.issue_calibration:
        issue_calibration()
        return
    if args.config == 5 and args.gpus == 1:
        run_config5(args)
        return
    if args.gpus != 1:
        print(json.dumps({"error": "runs on one GPU only: more GPUs would each take a "
                          "share of the patches (independent units, balanced by pair "
                          "count), which is not built", "n_gpus": args.gpus}))
        sys.exit(2)

    import torch
    from forcepho_amd import _lib, synth
    from forcepho_amd.autotune import load_or_default
    from forcepho_amd.batch import DeviceSplines, PatchBatch

    _lib.init(0)
    torch.cuda.init()
    t_setup = time.time()
    specs = synth.config4_specs(args.patches or 512)
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
        best_ms = next(ms for ms, c in table if c == best)
        autotune.save("batch", best, best_ms)
        log(f"best launch {best}: {best_ms:.3f} ms (fastest {table[0][0]:.3f} ms)")
        print(json.dumps({"best": best, "ms": best_ms,
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
    import measure
    n_in = measure.count_pairs_in_cutoff(batch)
    flops_alg = flops - FLOP_PER_PAIR * pairs + FLOP_PER_PAIR * n_in + \
        FLOP_PER_CUT_PAIR * max(pairs - n_in, 0)
    achieved = flops_alg / (eval_ms * 1e-3) / 1e12
    log(f"pairs inside the cutoff: {n_in:.4e} of {pairs:.4e} ({n_in / pairs:.3f}); "
        f"algorithmic {flops_alg:.3e} FLOP per sweep")
    log(f"stages (ms): convert {stage[0]:.3f} pixel pass {stage[1]:.3f} reduce {stage[2]:.3f}")

    # ---- graph-captured HMC leapfrog over every patch (10 steps/trajectory)
    hmc_rate = None
    if not args.skip_hmc:
        hmc_rate = time_hmc(patches, batch, splines, thetas, n_traj=3)
        log(f"graph-captured HMC: {hmc_rate['evals_per_s']:.1f} evals/s "
            f"({hmc_rate['ms_per_leapfrog_step']:.2f} ms per leapfrog step)")

    # ---- plain single-patch drop-in
    plain = None
    if not args.skip_plain:
        props_host = []
        for i, (p, sc, th) in enumerate(patches):
            sc.set_all_source_params(thetas[i])
            props_host.append(sc.get_proposal().copy())
        t_plain = time_plain([p for p, _, _ in patches], props_host, args.plain_sweeps)
        plain = len(patches) / t_plain
        log(f"plain: {t_plain * 1e3:.1f} ms/sweep, {plain:.1f} evals/s")

    cpu = cpu_all = chi2_err = None
    if not args.skip_cpu:
        batch.lnlike_and_grad(theta, splines)   # the sweep the checker reads back
        chi2_err = oracle_chi2_check(patches, specs, thetas, batch)
        c = cpu_baseline(patches, specs, thetas, args.cpu_seconds)
        if c is not None:
            t_sweep = pairs / c["pairs_per_s"]
            cpu = {"value": len(patches) / t_sweep, "unit": "evals/s", "cores": 1,
                   "kind": "reference", "host": host_cpu(),
                   "sample": (f"forcepho's compute_gaussians_kernel.cc compiled from source "
                              f"(g++ -O3), 1 thread, {c['patches']} randomly chosen patches "
                              f"of the workload ({c['sample_pairs']:.3e} pairs, "
                              f"{c['seconds']:.1f}s), scaled per pair to the full sweep"),
                   "pairs_per_s": c["pairs_per_s"]}
            log(f"cpu reference: {c['pairs_per_s']:.3e} pairs/s on {c['patches']} patches")
        workers = min(CPU_SHARE, len(os.sched_getaffinity(0)))
        c = cpu_baseline_all_cores(patches, specs, thetas, args.cpu_seconds / 4, workers)
        if c is not None:
            t_sweep = pairs / c["pairs_per_s"]
            cpu_all = {"value": len(patches) / t_sweep, "unit": "evals/s", "cores": workers,
                       "kind": "reference", "host": host_cpu(),
                       "cores_note": (f"{workers} worker processes: the host's CPU share "
                                      f"for one GPU is {CPU_SHARE} cores, whatever nproc "
                                      f"reports"),
                       "sample": (f"forcepho's compute_gaussians_kernel.cc compiled from "
                                  f"source (g++ -O3), {workers} processes, {c['patches']} "
                                  f"randomly chosen patches ({c['sample_pairs']:.3e} pairs) "
                                  f"assigned longest-first, slowest process "
                                  f"{c['seconds']:.1f}s, scaled per pair to the full sweep"),
                       "pairs_per_s": c["pairs_per_s"]}
            log(f"cpu reference, {workers} processes: {c['pairs_per_s']:.3e} pairs/s")

    traffic, traffic_src = measured_traffic(launch, args.precision)
    # exp-issue ceiling: one exp per in-cutoff pair (the algorithmic count)
    # at the device's measured exp throughput, as a time per sweep
    exp_rate = exp_issue_rate()
    exp_ceiling = {"exp_per_s": exp_rate, "ms_per_sweep": n_in / exp_rate * 1e3,
                   "frac_of_pixel_pass": n_in / exp_rate * 1e3 / eval_ms,
                   "measured_by": "16 v_exp_f32 chains per lane, 8 waves/SIMD, all CUs"}
    log(f"exp issue: {exp_rate:.3e} exp/s; one exp per in-cutoff pair takes "
        f"{exp_ceiling['ms_per_sweep']:.2f} ms of the {eval_ms:.2f} ms pixel pass")
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
        "hmc_graph": hmc_rate,
        "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK_FP32_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / PEAK_FP32_TFLOPS, "traffic": traffic,
                     "traffic_source": traffic_src,
                     "hbm_GBps": (traffic / (eval_ms * 1e-3) / 1e9) if traffic else None,
                     "exp_issue_ceiling": exp_ceiling,
                     "kernel": "eval pixel pass (eval_kernel / eval_persistent_kernel)",
                     "algorithmic_flop_per_sweep": flops_alg,
                     "pairs_in_cutoff": n_in,
                     "unpacked_fp32_peak": PEAK_FP32_TFLOPS / 2,
                     "frac_of_unpacked": achieved / (PEAK_FP32_TFLOPS / 2),
                     "note": ("compute-bound on the FP32 VALU (no dense contraction, so no "
                              "MFMA); peak = MI355X FP32 vector 157.3 TFLOP/s, which counts "
                              "packed FP32; the build disables packed FP32, so one FMA per "
                              "lane per 4 cycles issues at most half of it "
                              "(unpacked_fp32_peak); algorithmic FLOP = 98 per pair inside "
                              "the cutoff + 11 per other pair + 7 N_pix n_src + 6 N_pix")},
        "cpu_baseline": cpu,
        "cpu_baseline_all_cores": cpu_all,
    }
    print(json.dumps(out), flush=True)


def run_config5(args):
    """Mosaic-scale stress on one GPU: 4096 patches (5% crowded: 60 sources,
    12 bands x 2 exposures of 256^2) resident in HBM.  For each precision
    mode: time per sweep, chi2 and gradient error against the FP64 oracle on
    a sample (the 3 smallest patches and one crowded patch).  Then
    graph-captured HMC trajectories over all patches, and the plain drop-in
    timed on a random sample of patches.  Prints one JSON line."""
    import torch
    from forcepho_amd import _lib, synth
    from forcepho_amd.autotune import load_or_default
    from forcepho_amd.batch import DeviceSplines, PatchBatch

    _lib.init(0)
    torch.cuda.init()
    t_setup = time.time()
    specs = synth.config5_specs(args.patches or 4096)
    pairs, npix, flops = algorithmic_counts(specs)
    order = np.argsort([s.n_pairs() for s in specs])
    crowded = [i for i, s in enumerate(specs) if s.crowded]
    sample = [int(i) for i in order[:3]] + crowded[:1]
    log(f"config 5: {len(specs)} patches ({len(crowded)} crowded), {npix:.3e} px, "
        f"{pairs:.3e} pairs per sweep")
    patches = build_patches(specs, log_every=512, config=5, keep_host=set(sample))
    log(f"setup {time.time() - t_setup:.1f}s; device memory "
        f"{torch.cuda.memory_allocated() / 1e9:.1f} GB (torch) ")
    free, total = torch.cuda.mem_get_info()
    resident = (total - free) / 1e9
    batch = PatchBatch([p for p, _, _ in patches], precision="mixed")
    launch = load_or_default("batch")
    batch.set_launch(**launch)
    splines = DeviceSplines()
    thetas = [synth.make_proposals(th, len(p.band_N), 1, config=5, patch_id=i)[0]
              for i, (p, sc, th) in enumerate(patches)]
    theta = torch.as_tensor(np.concatenate(thetas), dtype=torch.float64, device="cuda")

    # the CPU leg's checker: FP64 oracle results of the sample
    ref = {} if args.skip_cpu else oracle_reference(patches, thetas, sample)

    modes = {}
    for mode in ("mixed", "fp64_reduce", "fp32"):
        batch.set_precision(mode)
        batch.lnlike_and_grad(theta, splines)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            batch.lnlike_and_grad(theta, splines)
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / args.steps
        chi2 = batch.chi2_host()
        dchi2 = batch.dchi2_host()
        cerr, gbad, gsize, gerr = 0.0, 0, 0, 0.0
        for i in ref:
            c_ref, d_ref = ref[i]
            cerr = max(cerr, float(np.max(np.abs(chi2[i] - c_ref) / np.abs(c_ref))))
            d = dchi2[i].astype(np.float64)
            scale = np.abs(d_ref).reshape(d_ref.shape[0], -1).max(axis=1)[:, None, None]
            err = np.abs(d - d_ref)
            bad = (err > 1e-4 * np.abs(d_ref)) & (err > 1e-6 * scale)
            gbad += int(bad.sum())
            gsize += int(bad.size)
            gerr = max(gerr, float(np.max(err / scale)))
        modes[mode] = {"ms_per_sweep": dt * 1e3, "evals_per_s": len(patches) / dt,
                       "chi2_rel_err_max": cerr, "grad_mismatches": gbad,
                       "grad_components": gsize, "grad_err_max_over_band_max": gerr}
        log(f"{mode}: {dt * 1e3:.1f} ms/sweep, chi2 err {cerr:.2e}, grad mismatches "
            f"{gbad}/{gsize}")
    batch.set_precision("mixed")

    # algorithmic roofline of the pixel pass (mixed)
    batch.set_timing(True)
    batch.lnlike_and_grad(theta, splines)
    stage = np.array(batch.last_timing())
    batch.set_timing(False)
    import measure
    n_in = measure.count_pairs_in_cutoff(batch)
    flops_alg = flops - FLOP_PER_PAIR * pairs + FLOP_PER_PAIR * n_in + \
        FLOP_PER_CUT_PAIR * max(pairs - n_in, 0)
    achieved = flops_alg / (stage[1] * 1e-3) / 1e12

    # HMC over every patch, one trajectory captured in a graph
    hmc = time_hmc(patches, batch, splines, thetas, n_traj=2)
    log(f"HMC: {hmc['trajectories']} trajectories x {hmc['leapfrog_steps']} steps, "
        f"{hmc['evals_per_s']:.0f} evals/s, mean acceptance {hmc['acceptance']:.2f}")

    # plain drop-in on a random sample, scaled per pair to the sweep
    rng = np.random.default_rng(55)
    pick = rng.choice(len(patches), min(256, len(patches)), replace=False)
    props_host = []
    for i in pick:
        sc = patches[i][1]
        sc.set_all_source_params(thetas[i])
        props_host.append(sc.get_proposal().copy())
    t_pl = time_plain([patches[i][0] for i in pick], props_host, 1)
    pairs_pick = sum(specs[i].n_pairs() for i in pick)
    plain = len(patches) / (t_pl * pairs / pairs_pick)
    log(f"plain: {len(pick)} patches in {t_pl:.2f}s -> {plain:.1f} evals/s over the sweep")

    out = {
        "metric": "config 5 mosaic stress: lnlike+grad evals/s per precision mode",
        "n_gpus": 1, "steps": args.steps, "data": "synthetic",
        "config": {"workload": f"config 5: {len(patches)} patches, {len(crowded)} crowded "
                               f"(60 src, 12 bands x 2 exp of 256^2)",
                   "patches": len(patches), "npix": npix, "n_pairs": pairs,
                   "launch": launch, "device_resident_gb": resident},
        "modes": modes,
        "pairs_in_cutoff": n_in,
        "roofline_mixed": {"bound": "valu", "achieved": achieved, "peak": PEAK_FP32_TFLOPS,
                           "unit": "TFLOP/s",
                           "frac": achieved / PEAK_FP32_TFLOPS,
                           "pixel_pass_ms": float(stage[1])},
        "hmc": hmc,
        "plain_evals_per_s": plain,
        "batched_over_plain": modes["mixed"]["evals_per_s"] / plain,
        "oracle_sample": {"patches": sample, "crowded_patch": crowded[:1]},
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
