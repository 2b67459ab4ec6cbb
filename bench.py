"""Benchmark: Fisher-vector encoding throughput at K=1024, D=80 on one MI355X.

Workload (the metric's configuration): a K=1024 diagonal GMM fitted on the GPU
to SIFT-like 80-d descriptors, then improved Fisher vectors (signed square root
+ L2) of 4,096 images x 3,000 descriptors (12.29 M descriptors, FV dimension
163,840), all inputs resident in HBM when the timed region starts.  A step is
one encode of the whole batch through the library's batched device entry point
(vlhip_fisher_encode_batch).  value = descriptors encoded per second by the
selected encoder (--mode; default: the learned-gated encoder, which falls back to
the compensated FP16 matrix-core encoder for every descriptor group its gate does
not serve).

Beside it, in the same run:
  * the plain FP32 GPU encoder (the reference-parity path) on the same batch,
    and the speed-up of the selected encoder over it;
  * the reference's own OpenMP CPU path (VLFeat built from source,
    oracle/_ref) on a bounded sample of the same images, and the parity of the
    GPU result against it;
  * the error of the selected encoder against the FP64 oracle on a few images.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--mode gated|f16x3|bf16x3|fp32]
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import re
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

K_COMP = 1024
DIM = 80
NUM_IMAGES = 4096
PER_IMAGE = 3000
FLAGS = 3  # VL_FISHER_FLAG_IMPROVED
PEAK_FP32_TFLOPS = 157.3
PEAK_F16_TFLOPS = 2500.0
PEAK_HBM_GBS = 8000.0

# Algorithmic flop per descriptor of each kernel (K components, D dimensions):
# the E-step contraction [x'^2 | x'] . [w2 | w1] is 4DK, the posterior-weighted
# accumulation Q^T [x' | x'^2] is 4DK.
KERNEL_FLOP_PER_DESC = {
    "mfma_fisher": 8 * DIM * K_COMP,   # logits + moments of its components
    "mfma_lse": 4 * DIM * K_COMP,      # logits for the row statistics
    "fisher_accum": 4 * DIM * K_COMP,  # FP32 accumulation
    "fisher_estep": 4 * DIM * K_COMP,  # FP32 E-step
    "f32_fisher": 8 * DIM * K_COMP,    # FP32 matrix cores: logits + moments of its components
    "f32_lse": 4 * DIM * K_COMP,       # FP32 matrix cores: logits for the row statistics
}
# Matrix-core flop actually issued per descriptor and component by the
# compensated kernels, relative to the algorithmic count:
#   mfma_fisher: logits, three products over the 2D features (3 x 2 x 2D flop),
#                moments, two products (hi, lo) for each of x' and x'^2
#                (2 x 2 x 2D), posterior mass, one 16x16x32 product with ones
#                per 16 components (32 flop); algorithmic 8D
#   mfma_lse:    logits only, 3 x 2 x 2D against the algorithmic 4D
#   f32_fisher:  plain FP32 logits (2 x 2D) and moments (2 x 2D), plus the
#                posterior mass, one 16x16x4 product with ones per 16
#                components and 4 rows (32 flop per descriptor and component)
#   f32_lse:     logits only, exactly the algorithmic count
EXECUTED_PER_ALGORITHMIC = {"mfma_fisher": (3 * 2 * 2 * DIM + 2 * 2 * 2 * DIM + 32) / (8 * DIM),
                            "mfma_lse": (3 * 2 * 2 * DIM) / (4 * DIM),
                            "f32_fisher": (8 * DIM + 32) / (8 * DIM), "f32_lse": 1.0}


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def host_description(R) -> str:
    """CPU model, logical CPUs visible to this process and the SIMD path the
    reference's distance kernels take."""
    model = "unknown CPU"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    simd = "AVX" if (R.vl_get_simd_enabled() and R.vl_cpu_has_avx()) else "SSE2/scalar"
    return f"{model}, {len(os.sched_getaffinity(0))} logical CPUs visible, {simd} distance kernels"


def workload_descriptors(N: int, dim: int, seed: int, device: str = "cuda", chunk: int = 1 << 20):
    """The SIFT-like workload's N descriptors from the library's seeded C
    generator (libvlsynth, include/vlsynth.h: a C caller of
    vlsynth_sift_like(out, N, dim, seed, 0) gets the same bits), generated on
    the host in chunks and copied to a device tensor."""
    import torch
    from vlfeat_amd import synth
    out = torch.empty((N, dim), dtype=torch.float32, device=device)
    buf = np.empty((min(chunk, max(N, 1)), dim), np.float32)
    for s0 in range(0, N, chunk):
        n = min(chunk, N - s0)
        synth.c_sift_like(n, dim, seed=seed, row_offset=s0, out=buf[:n])
        out[s0:s0 + n].copy_(torch.from_numpy(buf[:n]))
    return out


def fit_gmm(x_host: np.ndarray, K: int, iters: int, seed: int, cov_bound: float = 1e-6):
    """GMM for the workload: EM on the GPU from random data points."""
    import vlfeat_amd as va
    rng = np.random.default_rng(seed)
    means = x_host[rng.choice(x_host.shape[0], K, replace=False)].copy()
    covs = np.tile(np.var(x_host, 0), (K, 1)).astype(np.float32)
    pri = np.full(K, 1.0 / K, np.float32)
    m, c, p, ll, _ = va.vl_gmm(x_host, K, init="custom", means=means, covs=covs, priors=pri,
                               max_iter=iters, cov_bound=cov_bound, want_post=False)
    return m, c, p, ll


def cpu_legs(x_host, offsets, m, c, p, budget_s: float, gpu_enc, gpu_fp32_enc, n_oracle: int, gpu_gated_enc=None):
    """The reference VLFeat (oracle/_ref, OpenMP over all host threads) on whole
    images until the time budget is used; parity of the GPU FP32 encoder
    against it; error of the selected GPU encoder against the FP64 oracle."""
    from oracle import oracle, reference
    out = {}
    R = reference()
    if R is not None:
        R.set_num_threads(0)
        threads = int(R.vl_get_max_threads())
        t0 = time.perf_counter()
        done_desc = 0
        done_img = 0
        encs = []
        max_img = min(len(offsets) - 1, x_host.shape[0] // PER_IMAGE)
        while time.perf_counter() - t0 < budget_s and done_img < max_img:
            xs = x_host[offsets[done_img]:offsets[done_img + 1]]
            e, _ = R.fisher(xs, m, c, p, FLAGS)
            encs.append(e)
            done_desc += xs.shape[0]
            done_img += 1
        dt = time.perf_counter() - t0
        out["cpu_baseline"] = {
            "value": done_desc / dt, "unit": "descriptors/s", "cores": threads, "kind": "reference",
            "sample": f"{done_img} images x {PER_IMAGE} descriptors (K={K_COMP}, D={DIM}, improved FV), "
                      f"vl_fisher_encode per image, OpenMP inside the image, {dt:.1f} s; {host_description(R)}"}
        # image-parallel upper bound of the CPU path (as an application's parfor
        # over images): one host thread per image, no OpenMP inside a call
        import concurrent.futures as cf
        R.set_num_threads(1)
        t0 = time.perf_counter()
        with cf.ThreadPoolExecutor(threads) as ex:
            list(ex.map(lambda i: R.fisher(x_host[offsets[i]:offsets[i + 1]], m, c, p, FLAGS), range(done_img)))
        dt2 = time.perf_counter() - t0
        R.set_num_threads(0)
        out["cpu_image_parallel"] = {
            "value": done_desc / dt2, "unit": "descriptors/s", "cores": threads, "kind": "reference",
            "sample": f"the same {done_img} images, {threads} host threads each encoding whole images "
                      f"(vl_set_num_threads(1)), {dt2:.1f} s"}
        nc = min(len(encs), gpu_fp32_enc.shape[0])
        r = np.stack(encs[:nc]).astype(np.float64)
        g = gpu_fp32_enc[:nc].astype(np.float64)
        err = np.linalg.norm(g - r, axis=1) / np.linalg.norm(r, axis=1)
        out["parity_fp32_vs_reference"] = {"images": nc, "max_rel_l2": float(err.max())}
    O = oracle()
    errs, errs32, errref, errg = [], [], [], []
    for i in range(min(n_oracle, len(offsets) - 1)):
        xs = x_host[offsets[i]:offsets[i + 1]]
        eo, _ = O.fisher(xs, m, c, p, FLAGS)
        nrm = np.linalg.norm(eo)
        errs.append(np.linalg.norm(gpu_enc[i].astype(np.float64) - eo) / nrm)
        errs32.append(np.linalg.norm(gpu_fp32_enc[i].astype(np.float64) - eo) / nrm)
        if gpu_gated_enc is not None:
            errg.append(np.linalg.norm(gpu_gated_enc[i].astype(np.float64) - eo) / nrm)
        if R is not None and i < r.shape[0]:
            errref.append(np.linalg.norm(r[i] - eo) / nrm)
    out["error_vs_fp64_oracle"] = {
        "images": len(errs), "selected_max_rel_l2": float(max(errs)), "fp32_max_rel_l2": float(max(errs32)),
        "reference_float_max_rel_l2": float(max(errref)) if errref else None,
        "gated_max_rel_l2": float(max(errg)) if errg else None}
    return out


def projection_leg(L, x, N: int, steps: int) -> dict:
    """The pre-encoding projection (apps/recognition/encodeImage.m:91-97) at the
    workload's size: the same N SIFT-like descriptors before their projection
    (128-d, generated from the same seed), projected to 80-d by the workload's
    fixed orthonormal matrix with vlhip_project (center 0, no renormalisation),
    so the output must reproduce the encoder's input.  Reported per launch: time
    (hipEvents on the library stream), algorithmic bytes (128 x 4 read + 80 x 4
    written per descriptor) and flops (2 x 128 x 80 per descriptor) against the
    HBM and FP32 matrix peaks."""
    import torch
    from vlfeat_amd import synth
    raw = workload_descriptors(N, 128, seed=1)
    P = torch.as_tensor(synth.c_projection(DIM, 128, 1), dtype=torch.float32, device=raw.device)
    cen = torch.zeros(128, dtype=torch.float32, device=raw.device)
    out = torch.empty((N, DIM), dtype=torch.float32, device=raw.device)
    torch.cuda.synchronize()
    L.vlhip_project(out.data_ptr(), raw.data_ptr(), N, 128, DIM, P.data_ptr(), cen.data_ptr(), 0)
    L.vlhip_synchronize()
    L.vlhip_timing_reset()
    L.vlhip_timing_enable(1)
    for _ in range(steps):
        L.vlhip_project(out.data_ptr(), raw.data_ptr(), N, 128, DIM, P.data_ptr(), cen.data_ptr(), 0)
    L.vlhip_synchronize()
    L.vlhip_timing_enable(0)
    import vlfeat_amd as va
    ms, cnt = va.kernel_timings()["project"]
    avg = ms / max(cnt, 1)
    rel = float(torch.linalg.norm(out - x) / torch.linalg.norm(x))
    byts, flops = N * (128 + DIM) * 4, N * 2 * 128 * DIM
    del raw, out
    res = {"kernel": "project", "ms_per_launch": avg, "descriptors_per_launch": N,
           "GB_per_s": byts / avg / 1e6, "hbm_frac": byts / avg / 1e6 / PEAK_HBM_GBS,
           "TFLOP_per_s": flops / avg / 1e9, "fp32_frac": flops / avg / 1e9 / PEAK_FP32_TFLOPS,
           "rel_l2_vs_encoder_input": rel}
    log(f"projection load stage: {avg:.3f} ms per launch, {res['GB_per_s']:.0f} GB/s, "
        f"{res['TFLOP_per_s']:.1f} TF/s, rel. diff. to the encoder's input {rel:.2e}")
    return res


def traffic_per_descriptor(kernel: str):
    """HBM bytes per descriptor of `kernel` from the committed PMC summary
    (FETCH_SIZE and WRITE_SIZE passes scaled by a copy-kernel calibration), if any."""
    path = os.path.join(ROOT, "profiles", f"traffic_{kernel}_k{K_COMP}_d{DIM}.txt")
    if not os.path.exists(path):
        return None
    m = re.search(r"^hbm_bytes_per_descriptor:\s*([0-9.eE+]+)", open(path).read(), re.M)
    return float(m.group(1)) if m else None


def batched_config(args) -> int:
    """Batched uneven encoding: 4,096 images with log-normal descriptor counts
    (median 3,000, sigma 0.5, clamped to [200, 20,000]), D=80, a fixed K=512
    GMM; improved Fisher vectors (compensated FP16) and hard-assignment VLAD
    of every image in one graph-captured encode plan whose component tiling
    the library's autotuner picks.  Reports images/s over whole-batch replays
    and the per-image latency (single-image plans, graph replay) p50/p99."""
    import torch
    import vlfeat_amd as va
    from vlfeat_amd import synth

    K, D = 512, DIM
    L = va.lib()
    L.vlhip_init()
    counts = synth.c_lognormal_counts(args.images, median=3000, sigma=0.5, lo=200, hi=20000, seed=4)
    off = np.concatenate([[0], np.cumsum(counts)]).astype(np.uint64)
    N = int(off[-1])
    log(f"batched: {args.images} images, {N} descriptors (median {int(np.median(counts))}, max {counts.max()})")
    x = workload_descriptors(N, D, seed=2)
    torch.cuda.synchronize()
    m, c, p, ll = fit_gmm(x[:300_000].cpu().numpy(), K, iters=10, seed=0)
    dev = torch.device("cuda")
    md, cd, pd = (torch.as_tensor(a, device=dev) for a in (m, c, p))
    n = args.images
    enc = torch.empty((n, 2 * K * D), dtype=torch.float32, device=dev)
    vlad = torch.empty((n, K * D), dtype=torch.float32, device=dev)
    terms = torch.zeros(n, dtype=torch.int64, device=dev)
    plan = L.vlhip_encode_plan_new(D, K, off.ctypes.data, n, FLAGS, 0)
    t0 = time.perf_counter()
    L.vlhip_encode_plan_run(plan, x.data_ptr(), md.data_ptr(), cd.data_ptr(), pd.data_ptr(), enc.data_ptr(),
                            vlad.data_ptr(), terms.data_ptr())
    first_s = time.perf_counter() - t0
    variant = C.c_int()
    ms = (C.c_double * 2)()
    nb = C.c_uint64()
    L.vlhip_encode_plan_info(plan, C.byref(variant), ms, C.byref(nb))
    cands = []
    cw, wv, sc, cms = C.c_int(), C.c_int(), C.c_int(), C.c_double()
    for ci in range(L.vlhip_encode_plan_candidates(plan, -1, None, None, None, None)):
        L.vlhip_encode_plan_candidates(plan, ci, C.byref(cw), C.byref(wv), C.byref(sc), C.byref(cms))
        cands.append({"components_per_wave": cw.value, "waves": wv.value, "schedule": sc.value, "ms": cms.value})
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        L.vlhip_encode_plan_run(plan, x.data_ptr(), md.data_ptr(), cd.data_ptr(), pd.data_ptr(), enc.data_ptr(),
                                vlad.data_ptr(), terms.data_ptr())
    dt = (time.perf_counter() - t0) / args.steps
    L.vlhip_encode_plan_delete(plan)
    log(f"batched: {dt * 1000:.2f} ms per replay, tiling {variant.value}; candidates " +
        ", ".join(f"{c['components_per_wave']}x{c['waves']}/s{c['schedule']}: {c['ms']:.2f} ms" for c in cands) +
        f"; first run (autotune + capture) {first_s:.2f} s")
    # per-image latency: single-image plans over a sample of the images
    rng = np.random.default_rng(0)
    sample = np.sort(rng.choice(n, min(256, n), replace=False))
    lat = []
    for i in sample:
        o1 = np.array([0, counts[i]], np.uint64)
        pl = L.vlhip_encode_plan_new(D, K, o1.ctypes.data, 1, FLAGS, 0)
        xi = x[int(off[i]):int(off[i + 1])]
        args_i = (xi.data_ptr(), md.data_ptr(), cd.data_ptr(), pd.data_ptr(), enc.data_ptr(), vlad.data_ptr(),
                  terms.data_ptr())
        L.vlhip_encode_plan_run(pl, *args_i)
        t0 = time.perf_counter()
        L.vlhip_encode_plan_run(pl, *args_i)
        lat.append(time.perf_counter() - t0)
        L.vlhip_encode_plan_delete(pl)
    lat = np.array(lat) * 1000
    result = {
        "metric": "batched uneven Fisher+VLAD encoding (K=512, D=80, log-normal image sizes)",
        "value": n / dt, "unit": "images/s", "n_gpus": 1, "steps": args.steps, "warmup": 1,
        "ms_per_step": dt * 1000, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "fp16", "data": "synthetic (seeded C generator vlsynth_sift_like: SIFT-like projected to 80-d, seed 2; K=512 GMM fitted by GPU EM)",
        "config": {"workload": "4096 images, n_i ~ LogNormal(log 3000, 0.5) clamped to [200, 20000], "
                               "improved FV + hard VLAD (nearest mean), one encode plan",
                   "images": n, "descriptors": N, "K": K, "D": D},
        "descriptors_per_s": N / dt,
        "tiling": {"chosen": variant.value, "chosen_ms": ms[0], "candidates": cands, "batches": nb.value,
                   "search": "every component tiling (components per wave x waves) with schedule 0, then "
                             "schedules 1 and 2 with the fastest tiling; one warm-up and one timed run each"},
        "first_run_s": first_s,
        "per_image_latency_ms": {"sample": len(lat), "p50": float(np.percentile(lat, 50)),
                                 "p99": float(np.percentile(lat, 99)), "max": float(lat.max()),
                                 "how": "single-image encode plan, graph replay, wall clock of the synchronous call"},
    }
    print(json.dumps(result), flush=True)
    return 0


def _em_finalize(out: np.ndarray, m_old: np.ndarray, K: int, D: int, lb: float):
    """M-step finalisation of gmm.c:1169-1202 in FP64, applied alike to every
    path's moments [S0 | mass | S1 | S2] for the error report."""
    S0 = out[:K]
    S1 = out[2 * K:2 * K + K * D].reshape(K, D)
    S2 = out[2 * K + K * D:].reshape(K, D)
    upd = (S0 >= 1e-6 / K)[:, None]
    s0 = np.where(S0 > 0, S0, 1.0)[:, None]
    mu = np.where(upd, S1 / s0, S1)
    cv = np.maximum(np.where(upd, S2 / s0 - (mu - m_old) ** 2, S2), lb)
    pri = S0 / max(S0.sum(), 1e-12)
    return {"S0": S0, "S1": S1, "S2": S2, "means": mu, "covariances": cv, "priors": pri}


def _err(a, b) -> dict:
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    d = np.abs(a - b)
    nz = np.abs(b) > 0
    return {"max_abs": float(d.max()), "max_rel": float((d[nz] / np.abs(b[nz])).max()) if nz.any() else 0.0,
            "rel_l2": float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))}


def em_config(args) -> int:
    """Precision form of the EM iteration: N = 4,000,000 SIFT-like 128-d
    descriptors (rectified, L2-normalised, clipped at 0.2, renormalised; not
    projected), K = 1024.  A step is one EM iteration on device-resident data:
    the log-likelihood and the M-step moments of the current model
    (vlhip_gmm_em_step) in the compensated FP16 (value) and BF16 matrix-core
    modes, and the FP32 parity path (posterior matrix + moments) beside them.
    The model is a 5-iteration FP32 EM fit on 500k of the descriptors
    (variance floor 1e-4).  Errors per output (log-likelihood, moments,
    means / covariances / priors after the M-step) against FP64:
      * on all N descriptors against the library's FP64 path (double data);
      * on the first --em-oracle-rows descriptors against the FP64 CPU oracle,
        which also pins that FP64 GPU path.
    cpu_baseline: the reference's OpenMP EM (oracle/_ref), one iteration timed
    as the difference of 2- and 1-iteration runs on a bounded sample."""
    import threading

    import torch
    import vlfeat_amd as va
    from vlfeat_amd import synth

    N, D, K = args.em_n, 128, 1024
    lb = 1e-4
    L = va.lib()
    L.vlhip_init()
    dev = torch.device("cuda")
    log(f"em: generating {N} descriptors (D={D})")
    x = workload_descriptors(N, D, seed=1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    m, c, p, ll_fit = fit_gmm(x[:500_000].cpu().numpy(), K, iters=5, seed=0, cov_bound=lb)
    log(f"em: model fitted in {time.perf_counter() - t0:.1f} s (FP32 EM, 5 iterations, 500k descriptors)")
    md, cd, pd = (torch.as_tensor(a, device=dev) for a in (m, c, p))
    KD = K * D
    outs = {}

    def mfma_step(mode, xp, n):
        o = np.zeros(2 * K + 2 * KD)
        ll = L.vlhip_gmm_em_step(mode, o.ctypes.data, pd.data_ptr(), md.data_ptr(), cd.data_ptr(), D, K, xp, n)
        return ll, o

    q32 = torch.empty((N, K), dtype=torch.float32, device=dev)

    def fp32_step(xp, n):
        o = np.zeros(2 * K + 2 * KD)
        ll = L.vlhip_gmm_posteriors(1, q32.data_ptr(), K, n, pd.data_ptr(), md.data_ptr(), D, cd.data_ptr(), xp)
        L.vlhip_gmm_moments(1, o.ctypes.data, q32.data_ptr(), xp, md.data_ptr(), n, D, K, 0.01 / K)
        return ll, o

    steps = {"f16x3": lambda: mfma_step(va.MODE_F16X3, x.data_ptr(), N),
             "bf16x3": lambda: mfma_step(va.MODE_BF16X3, x.data_ptr(), N),
             "fp32": lambda: fp32_step(x.data_ptr(), N)}
    timing = {}
    for name, fn in steps.items():
        for _ in range(args.warmup):
            fn()
        L.vlhip_timing_reset()
        L.vlhip_timing_enable(1)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            ll, o = fn()
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / args.steps
        L.vlhip_timing_enable(0)
        kt = va.kernel_timings()
        timing[name] = {"ms_per_iteration": dt * 1000, "descriptors_per_s": N / dt,
                        "kernels_ms_per_iteration": {k: v[0] / args.steps for k, v in kt.items() if v[1]},
                        "kernel_launches": {k: v[1] for k, v in kt.items() if v[1]}}
        outs[name] = (ll, o)
        log(f"em: {name}: {dt * 1000:.1f} ms per EM iteration; " +
            ", ".join(f"{k}={v[0] / args.steps:.2f}ms" for k, v in kt.items() if v[1]))
    del q32

    # FP64 over all N descriptors (the library's double path)
    x64 = x.double()
    m64, c64, p64 = (t.double() for t in (md, cd, pd))
    q64 = torch.empty((N, K), dtype=torch.float64, device=dev)
    o64 = np.zeros(2 * K + 2 * KD)
    ll64 = L.vlhip_gmm_posteriors(2, q64.data_ptr(), K, N, p64.data_ptr(), m64.data_ptr(), D, c64.data_ptr(),
                                  x64.data_ptr())
    L.vlhip_gmm_moments(2, o64.ctypes.data, q64.data_ptr(), x64.data_ptr(), m64.data_ptr(), N, D, K, 0.01 / K)
    torch.cuda.synchronize()
    del q64
    f64 = _em_finalize(o64, m.astype(np.float64), K, D, lb)
    full = {}
    for name, (ll, o) in outs.items():
        fo = _em_finalize(o, m.astype(np.float64), K, D, lb)
        e = {"loglikelihood_rel": abs(ll - ll64) / abs(ll64), "mass": _err(o[K:2 * K], o64[K:2 * K])}
        e.update({k: _err(fo[k], f64[k]) for k in ("S0", "S1", "S2", "means", "covariances", "priors")})
        full[name] = e
        log(f"em: {name} vs FP64 (all N): LL {e['loglikelihood_rel']:.2e}, means {e['means']['rel_l2']:.2e}, "
            f"covs {e['covariances']['rel_l2']:.2e}, priors {e['priors']['rel_l2']:.2e}")

    # FP64 CPU oracle on a subsample (threads over row blocks)
    from oracle import oracle, reference
    O = oracle()
    ns = min(args.em_oracle_rows, N)
    xs = x[:ns].cpu().numpy()
    nth = 16
    parts = [None] * nth
    bounds = np.linspace(0, ns, nth + 1).astype(int)

    def work(i):
        parts[i] = O.posteriors(xs[bounds[i]:bounds[i + 1]], m, c, p)
    t0 = time.perf_counter()
    th = [threading.Thread(target=work, args=(i,)) for i in range(nth)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    qo = np.concatenate([q for q, _ in parts])
    llo = float(sum(l for _, l in parts))
    log(f"em: oracle posteriors of {ns} descriptors in {time.perf_counter() - t0:.1f} s ({nth} threads)")
    xo = xs.astype(np.float64)
    qt = np.where(qo < 0.01 / K, 0.0, qo)
    S0 = qt.sum(0)
    S1 = qt.T @ xo
    mo = m.astype(np.float64)
    S2 = qt.T @ (xo * xo) - 2 * mo * S1 + mo * mo * S0[:, None]
    oo = np.concatenate([S0, qo.sum(0), S1.ravel(), S2.ravel()])
    fo_ref = _em_finalize(oo, m.astype(np.float64), K, D, lb)
    sub_outs = {"f16x3": mfma_step(va.MODE_F16X3, x.data_ptr(), ns), "bf16x3": mfma_step(va.MODE_BF16X3, x.data_ptr(), ns)}
    q32 = torch.empty((ns, K), dtype=torch.float32, device=dev)
    sub_outs["fp32"] = fp32_step(x.data_ptr(), ns)
    del q32
    q64 = torch.empty((ns, K), dtype=torch.float64, device=dev)
    o64s = np.zeros(2 * K + 2 * KD)
    ll64s = L.vlhip_gmm_posteriors(2, q64.data_ptr(), K, ns, p64.data_ptr(), m64.data_ptr(), D, c64.data_ptr(),
                                   x64.data_ptr())
    L.vlhip_gmm_moments(2, o64s.ctypes.data, q64.data_ptr(), x64.data_ptr(), m64.data_ptr(), ns, D, K, 0.01 / K)
    sub_outs["fp64_gpu"] = (ll64s, o64s)
    del q64, x64
    sub = {}
    for name, (ll, o) in sub_outs.items():
        fo = _em_finalize(o, m.astype(np.float64), K, D, lb)
        e = {"loglikelihood_rel": abs(ll - llo) / abs(llo), "mass": _err(o[K:2 * K], oo[K:2 * K])}
        e.update({k: _err(fo[k], fo_ref[k]) for k in ("S0", "S1", "S2", "means", "covariances", "priors")})
        sub[name] = e
        log(f"em: {name} vs FP64 oracle ({ns} rows): LL {e['loglikelihood_rel']:.2e}, "
            f"means {e['means']['rel_l2']:.2e}, covs {e['covariances']['rel_l2']:.2e}")

    # roofline of the dominant matrix-core kernel of the F16X3 iteration
    kt = timing["f16x3"]["kernels_ms_per_iteration"]
    kl = timing["f16x3"]["kernel_launches"]
    flop_desc = {"mfma_em_moments": 8 * D * K, "mfma_lse": 4 * D * K}
    kname = max((k for k in kt if k in flop_desc), key=lambda k: kt[k])
    per_launch_ms = kt[kname] * args.steps / kl[kname]
    desc_per_launch = N / (kl[kname] / args.steps)
    achieved = flop_desc[kname] * desc_per_launch / (per_launch_ms / 1000) / 1e12
    roofline = {"bound": "mfma", "achieved": achieved, "peak": PEAK_F16_TFLOPS, "unit": "TFLOP/s",
                "frac": achieved / PEAK_F16_TFLOPS, "traffic": None, "kernel": kname, "kernel_avg_ms": per_launch_ms,
                "flop_per_descriptor": flop_desc[kname], "descriptors_per_launch": desc_per_launch}

    cpu = None
    R = reference()
    if R is not None and not args.no_cpu:
        R.set_num_threads(0)
        threads = int(R.vl_get_max_threads())
        nc = min(args.em_cpu_rows, N)
        xc = x[:nc].cpu().numpy()
        # vl_gmm_em with T iterations runs T + 1 E-steps and T M-steps: one
        # iteration (E + M) is half the difference of a 3- and a 1-iteration
        # run, from a start far enough from convergence that neither stops
        # early (random data points, data variance, uniform priors)
        rng = np.random.default_rng(3)
        m0 = xc[rng.choice(nc, K, replace=False)].copy()
        c0 = np.tile(np.var(xc, 0), (K, 1)).astype(np.float32)
        p0 = np.full(K, 1.0 / K, np.float32)

        def cpu_run(t):
            t0 = time.perf_counter()
            R.gmm(xc, K, init="custom", means=m0, covs=c0, priors=p0, max_iter=t, cov_bound=lb, want_post=False)
            return time.perf_counter() - t0
        cpu_run(1)  # warm-up (thread pool, first touch of the buffers)
        t1 = cpu_run(1)
        t3 = cpu_run(3)
        it_s = max((t3 - t1) / 2, 1e-9)
        cpu = {"value": nc / it_s, "unit": "descriptors/s", "cores": threads, "kind": "reference",
               "sample": f"{nc} descriptors, K={K}, D={D}: one EM iteration (E-step + M-step) of the reference's "
                         f"vl_gmm_em (OpenMP), timed as half the difference of a 3-iteration ({t3:.2f} s) and a "
                         f"1-iteration ({t1:.2f} s) vl_gmm_cluster run from random data points"}
        log(f"em: reference CPU {cpu['value']:.3g} descriptors/s per EM iteration ({threads} threads)")

    result = {
        "metric": "EM iteration descriptors/s at N=4M, D=128, K=1024 (precision form)",
        "value": timing["f16x3"]["descriptors_per_s"], "unit": "descriptors/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": timing["f16x3"]["ms_per_iteration"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "fp16",
        "data": "synthetic (seeded C generator vlsynth_sift_like: SIFT-like 128-d, not projected, seed 1; model: 5-iteration FP32 EM fit)",
        "config": {"workload": "one EM iteration (LL + M-step moments) over 4M descriptors, K=1024, D=128",
                   "N": N, "K": K, "D": D, "variance_floor": lb},
        "modes": timing,
        "speedup_f16x3_vs_fp32_gpu": timing["fp32"]["ms_per_iteration"] / timing["f16x3"]["ms_per_iteration"],
        "speedup_bf16x3_vs_fp32_gpu": timing["fp32"]["ms_per_iteration"] / timing["bf16x3"]["ms_per_iteration"],
        "errors_vs_fp64_gpu_all_rows": full,
        "errors_vs_fp64_oracle_subsample": {"rows": ns, "per_mode": sub},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(result), flush=True)
    return 0


def dropin_config(args) -> int:
    """Plain FP32 drop-in at SIFT scale, through VLFeat's host C API exactly as a
    caller of the reference uses it: vl_gmm_cluster on N = 1,000,000 SIFT-like
    descriptors projected to 80-d (seed 1), K = 256, 30 EM iterations from a
    seeded custom start (random data points, data variance, uniform priors);
    vl_kmeans_cluster (Lloyd, and Elkan, 10 iterations each) on the same data; then
    vl_fisher_encode (improved), the homogeneous kernel map of the encodings,
    a linear SVM (vl_svm_train, SDCA) on the encodings, and vl_vlad_encode
    (one-hot assignments from
    vl_kmeans_quantize against the GMM means) of 1,000 images x 3,000 descriptors
    (seed 2).  The same calls run on the reference (oracle/_ref, OpenMP over the
    host cores) on the same inputs: wall-clock of both, and the parity of every
    output at full size (means / covariances / priors / LL after 30 iterations;
    every Fisher vector, VLAD vector and hard assignment computed by both from
    the reference's GMM)."""
    import torch
    import vlfeat_amd as va
    from oracle import reference
    from vlfeat_amd import synth

    N, D, K, n_img, per = args.dropin_n, DIM, 256, args.dropin_images, PER_IMAGE
    L = va.lib()
    L.vlhip_init()
    log(f"dropin: generating {N} + {n_img * per} descriptors")
    x = synth.c_sift_like(N, D, seed=1)
    xe = synth.c_sift_like(n_img * per, D, seed=2)
    torch.cuda.synchronize()
    rng = np.random.default_rng(7)
    m0 = x[rng.choice(N, K, replace=False)].copy()
    c0 = np.tile(np.var(x, 0), (K, 1)).astype(np.float32)
    p0 = np.full(K, 1.0 / K, np.float32)

    def run(lib, model=None):
        """EM, then the encoders with `model` (the reference's GMM, so that the
        encoders' parity does not include the EM's) or this run's own GMM."""
        out = {}
        t0 = time.perf_counter()
        m, c, p, ll, _ = lib.gmm(x, K, init="custom", means=m0, covs=c0, priors=p0, max_iter=30, want_post=False)
        out["em_s"] = time.perf_counter() - t0
        em = (m, c, p, ll)
        if model is not None:
            m, c, p = model
        t0 = time.perf_counter()
        fv = np.stack([lib.fisher(xe[i * per:(i + 1) * per], m, c, p, FLAGS)[0] for i in range(n_img)])
        out["fisher_s"] = time.perf_counter() - t0
        km = lib.kmeans_new(np.float32)
        lib.vl_kmeans_set_centers(km, m.ctypes.data, D, K)
        t0 = time.perf_counter()
        a = lib.kmeans_quantize(km, xe)[0]
        out["quantize_s"] = time.perf_counter() - t0
        lib.vl_kmeans_delete(km)
        # vl_kmeans_cluster (Lloyd, random-selection seeding, 10 iterations) on
        # the EM data, as the GMM's k-means initialisation would run it
        lib.seed(0)
        km = lib.kmeans_new(np.float32, algorithm=va.VlKMeansLloyd, max_iter=10)
        t0 = time.perf_counter()
        ke = lib.vl_kmeans_cluster(km, x.ctypes.data, D, N, K)
        out["kmeans_s"] = time.perf_counter() - t0
        kc = lib.kmeans_centers(km, np.float32)
        lib.vl_kmeans_delete(km)
        # the same with Elkan's refinement (kmeans.c:1162-1745)
        lib.seed(0)
        km = lib.kmeans_new(np.float32, algorithm=va.VlKMeansElkan, max_iter=10)
        t0 = time.perf_counter()
        kee = lib.vl_kmeans_cluster(km, x.ctypes.data, D, N, K)
        out["kmeans_elkan_s"] = time.perf_counter() - t0
        kce = lib.kmeans_centers(km, np.float32)
        lib.vl_kmeans_delete(km)
        # one dense one-hot matrix, rewritten in place per image (3,000 entries
        # cleared, 3,000 set), so that the timing is that of the encoder calls
        # rather than of allocating and zeroing 3 MB per image
        oh = np.zeros((per, K), np.float32)
        rows = np.arange(per)
        prev = None
        t0 = time.perf_counter()
        vl = []
        for i in range(n_img):
            ai = a[i * per:(i + 1) * per]
            if prev is not None:
                oh[rows, prev] = 0
            oh[rows, ai] = 1
            prev = ai
            vl.append(lib.vlad(xe[i * per:(i + 1) * per], m, oh, 0))
        out["vlad_s"] = time.perf_counter() - t0
        return out, em + (fv, a, np.stack(vl), kc, ke, kce, kee)

    R = reference() if not args.no_cpu else None
    r_t = r = None
    if R is not None:
        R.set_num_threads(0)
        r_t, r = run(R)
        log(f"dropin: reference ({int(R.vl_get_max_threads())} threads): EM {r_t['em_s']:.2f} s, "
            f"k-means {r_t['kmeans_s']:.2f} s (Elkan {r_t['kmeans_elkan_s']:.2f} s), "
            f"Fisher {r_t['fisher_s']:.2f} s, quantize {r_t['quantize_s']:.2f} s, VLAD {r_t['vlad_s']:.2f} s")
    g_t, g = run(L, model=r[:3] if r is not None else None)
    log(f"dropin: GPU library: EM {g_t['em_s']:.2f} s, k-means {g_t['kmeans_s']:.2f} s "
        f"(Elkan {g_t['kmeans_elkan_s']:.2f} s), Fisher {g_t['fisher_s']:.2f} s, "
        f"quantize {g_t['quantize_s']:.2f} s, VLAD {g_t['vlad_s']:.2f} s")
    # homogeneous kernel map of the encodings, as the recognition application
    # applies it before its linear SVM (traintest.m:99: order 1, chi2, the
    # MEX front-end's rectangular window): the library's array form (host
    # arrays, PCIe included) vs the reference's scalar entry point looped over
    # the values in C (toolbox/misc/vl_homkermap.c:170-184, one thread)
    fvs = np.ascontiguousarray(g[4], np.float32)
    L.homkermap(fvs, 1, 1.0, 1, -1.0, 1, array=True)  # warm-up: buffers, table, first launch
    L.vlhip_timing_reset()
    L.vlhip_timing_enable(1)
    t0 = time.perf_counter()
    hk = L.homkermap(fvs, 1, 1.0, 1, -1.0, 1, array=True)
    hk_s = time.perf_counter() - t0
    L.vlhip_timing_enable(0)
    hk_ms, _ = va.kernel_timings().get("homkermap", (float("nan"), 0))
    n_hk = fvs.size
    homk = {"values": n_hk, "gpu_library_s": hk_s, "kernel_ms": hk_ms,
            "kernel_GB_per_s": n_hk * 16 / hk_ms / 1e6 if hk_ms == hk_ms else None}
    if R is not None:
        import ctypes as C
        f = R.lib.ref_homkermap_array_f
        f.restype, f.argtypes = None, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t]
        hmap = R.vl_homogeneouskernelmap_new(1, 1.0, 1, -1.0, 1)
        hr = np.empty_like(hk)
        t0 = time.perf_counter()
        f(hmap, hr.ctypes.data, fvs.ctypes.data, n_hk)
        homk["reference_s"] = time.perf_counter() - t0
        R.vl_homogeneouskernelmap_delete(hmap)
        homk["bit_identical"] = bool(np.array_equal(hk, hr))
    log("dropin: homogeneous kernel map of the encodings: " + ", ".join(f"{k}={v}" for k, v in homk.items()))
    del hk
    # linear SVM on the encodings, as the recognition application trains one
    # per class (traintest.m, vl_svmtrain): SDCA, hinge loss, lambda 1e-4, the
    # reference's defaults otherwise (epsilon 1e-2, diagnostics once per
    # epoch, at most max(n, 10 / lambda) iterations); one-vs-rest labels of a
    # 10-class split of the images
    svm = {}
    labels = np.where(np.arange(n_img) % 10 == 0, 1.0, -1.0)
    t0 = time.perf_counter()
    sg = L.svm_train(fvs, labels, 1e-4, solver=2, loss=0, seed=0)
    svm["gpu_library_s"] = time.perf_counter() - t0
    svm.update({"points": n_img, "dimension": int(fvs.shape[1]), "iterations": sg["statistics"]["iteration"] + 1,
                "status": sg["statistics"]["status"], "objective": sg["statistics"]["objective"],
                "duality_gap": sg["statistics"]["dualityGap"]})
    if R is not None:
        t0 = time.perf_counter()
        sr = R.svm_train(fvs, labels, 1e-4, solver=2, loss=0, seed=0)
        svm["reference_s"] = time.perf_counter() - t0
        svm["model_rel_diff"] = float(np.linalg.norm(sg["model"] - sr["model"]) / np.linalg.norm(sr["model"]))
        svm["same_iterations"] = sg["statistics"]["iteration"] == sr["statistics"]["iteration"]
    log("dropin: SVM (SDCA) on the encodings: " + ", ".join(f"{k}={v}" for k, v in svm.items()))

    result = {
        "metric": "FP32 drop-in at SIFT scale (config 2): host C API, GPU library vs reference",
        "value": n_img * per / (g_t["fisher_s"] + g_t["vlad_s"] + g_t["quantize_s"]),
        "unit": "descriptors/s (Fisher + quantize + VLAD through the host API, PCIe included)",
        "n_gpus": 1, "steps": 1, "warmup": 0, "ms_per_step": 1000 * (g_t["fisher_s"] + g_t["vlad_s"] + g_t["quantize_s"]),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "fp32",
        "data": "synthetic (seeded C generator vlsynth_sift_like: SIFT-like projected to 80-d; EM data seed 1, images seed 2)",
        "config": {"workload": "vl_gmm_cluster N=1M K=256 D=80 30 iterations; improved FV + VLAD of "
                               f"{n_img} images x {per}", "N": N, "K": K, "D": D, "images": n_img},
        "gpu_library_s": g_t,
        "homkermap": homk,
        "svm": svm,
    }
    if R is not None:
        rel = lambda a, b: float(np.linalg.norm(np.asarray(a, np.float64) - b) / np.linalg.norm(np.asarray(b, np.float64)))
        fv_err = np.linalg.norm(g[4].astype(np.float64) - r[4], axis=1) / np.linalg.norm(r[4].astype(np.float64), axis=1)
        vl_err = np.linalg.norm(g[6].astype(np.float64) - r[6], axis=1) / np.maximum(
            np.linalg.norm(r[6].astype(np.float64), axis=1), 1e-30)
        parity = {"means_rel_l2": rel(g[0], r[0]), "covariances_rel_l2": rel(g[1], r[1]),
                  "priors_rel_l2": rel(g[2], r[2]), "loglikelihood_rel": abs(g[3] - r[3]) / abs(r[3]),
                  "fisher_max_rel_l2": float(fv_err.max()), "fisher_median_rel_l2": float(np.median(fv_err)),
                  "assignments_equal": int((g[5] == r[5]).sum()), "assignments_total": int(g[5].size),
                  "vlad_max_rel_l2": float(vl_err.max()),
                  "kmeans_centers_equal": bool(np.array_equal(g[7], r[7])), "kmeans_energy_equal": g[8] == r[8],
                  "elkan_centers_equal": bool(np.array_equal(g[9], r[9])), "elkan_energy_equal": g[10] == r[10]}
        log("dropin: parity " + ", ".join(f"{k}={v:.3g}" for k, v in parity.items()))
        result["reference_s"] = r_t
        result["parity_vs_reference"] = parity
        result["speedup_vs_reference"] = {k: r_t[k] / g_t[k] for k in g_t}
        result["cpu_baseline"] = {
            "value": n_img * per / (r_t["fisher_s"] + r_t["vlad_s"] + r_t["quantize_s"]), "unit": result["unit"],
            "cores": int(R.vl_get_max_threads()), "kind": "reference",
            "sample": f"the same {n_img} images through the reference's host API (OpenMP inside each call)"}
    print(json.dumps(result), flush=True)
    return 0


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--mode", default="gated", choices=["fp32", "fp32mfma", "bf16x3", "f16x3", "gated"])
    ap.add_argument("--images", type=int, default=NUM_IMAGES)
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    ap.add_argument("--oracle-images", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp32", action="store_true", help="skip the plain FP32 comparison run")
    ap.add_argument("--no-gated", action="store_true", help="skip the learned-gate leg")
    ap.add_argument("--gate-train", type=int, default=2_000_000, help="descriptors the gate is trained on")
    ap.add_argument("--gate-need", type=float, default=0.999)
    ap.add_argument("--config", default="headline", choices=["headline", "batched", "em", "dropin"],
                    help="batched: the uneven-image Fisher+VLAD plan benchmark (K=512); em: the EM precision "
                         "form (N=4M, D=128, K=1024) instead of the headline")
    ap.add_argument("--em-n", type=int, default=4_000_000)
    ap.add_argument("--dropin-n", type=int, default=1_000_000)
    ap.add_argument("--dropin-images", type=int, default=1000)
    ap.add_argument("--em-oracle-rows", type=int, default=65536)
    ap.add_argument("--em-cpu-rows", type=int, default=131072)
    ap.add_argument("--no-project", action="store_true", help="skip the projection load-stage leg")
    ap.add_argument("--copy-probe", action="store_true",
                    help="also run a 1 GiB device copy (HBM counter calibration under rocprofv3 --pmc)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this library runs on one GPU; a multi-GPU layout (images split across GPUs, "
                          "one all-reduce of EM statistics per iteration) is not built"}))
        return 2
    if args.config == "batched":
        return batched_config(args)
    if args.config == "em":
        return em_config(args)
    if args.config == "dropin":
        return dropin_config(args)

    import torch
    import vlfeat_amd as va
    from vlfeat_amd import synth

    L = va.lib()
    L.vlhip_init()
    n_img = args.images
    N = n_img * PER_IMAGE
    log(f"generating {N} descriptors (D={DIM}) with the seeded C generator")
    x = workload_descriptors(N, DIM, seed=1)
    torch.cuda.synchronize()
    train = x[: 500_000].cpu().numpy()
    log("fitting the K=1024 GMM (EM on the GPU, 500k descriptors)")
    t0 = time.perf_counter()
    m, c, p, ll = fit_gmm(train, K_COMP, iters=10, seed=0)
    log(f"GMM fitted in {time.perf_counter() - t0:.1f} s, LL={ll:.4g}")

    dev = torch.device("cuda")
    md = torch.as_tensor(m, device=dev)
    cd = torch.as_tensor(c, device=dev)
    pd = torch.as_tensor(p, device=dev)
    offsets = np.arange(n_img + 1, dtype=np.uint64) * PER_IMAGE
    od = torch.as_tensor(offsets.astype(np.int64), device=dev)
    modes = {"fp32": va.MODE_FP32, "fp32mfma": va.MODE_FP32_MFMA, "bf16x3": va.MODE_BF16X3, "f16x3": va.MODE_F16X3,
             "gated": va.MODE_GATED}
    torch.cuda.synchronize()

    gate_info = None
    gate = None
    if args.mode == "gated" or not args.no_gated:
        # learned gate (80 -> 256 -> 256 -> 1024, GELU, BF16 products) trained on the
        # GPU against the exact posteriors of the fitted GMM, outside the timed region
        n_train = min(args.gate_train, N)
        batch, epochs = 4096, 2
        steps = epochs * (n_train // batch)
        gate = L.vlhip_gate_new(DIM, K_COMP, 256, 256, 1)
        L.vlhip_gate_set_standardization_from_data(gate, x.data_ptr(), n_train)
        trace = np.empty(steps, np.float32)
        t0 = time.perf_counter()
        L.vlhip_gate_train(gate, x.data_ptr(), n_train, md.data_ptr(), cd.data_ptr(), pd.data_ptr(), steps, batch,
                           C.c_float(1e-3), trace.ctypes.data)
        torch.cuda.synchronize()
        gate_info = {"train_descriptors": n_train, "steps": steps, "batch": batch,
                     "train_s": time.perf_counter() - t0, "loss_first": float(trace[:20].mean()),
                     "loss_last": float(trace[-20:].mean()), "candidates": 16, "need": args.gate_need}
        log(f"gate trained: {steps} steps in {gate_info['train_s']:.1f} s, loss "
            f"{gate_info['loss_first']:.3f} -> {gate_info['loss_last']:.3f}")
        L.vlhip_fisher_set_gate(gate, 16, C.c_float(args.gate_need))

    def run(mode_name: str, steps: int, warmup: int):
        enc = torch.empty((n_img, 2 * K_COMP * DIM), dtype=torch.float32, device=dev)
        terms = torch.zeros(n_img, dtype=torch.int64, device=dev)

        def step():
            L.vlhip_fisher_encode_batch(1, enc.data_ptr(), md.data_ptr(), cd.data_ptr(), pd.data_ptr(), DIM, K_COMP,
                                        x.data_ptr(), od.data_ptr(), n_img, FLAGS, modes[mode_name],
                                        terms.data_ptr())
        for i in range(warmup):
            step()
            torch.cuda.synchronize()
            log(f"{mode_name}: warmup step {i} done")
        L.vlhip_timing_reset()
        L.vlhip_timing_enable(1)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            step()
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        L.vlhip_timing_enable(0)
        timings = va.kernel_timings()
        log(f"{mode_name}: timed {steps} steps in {dt:.3f} s; kernels: " +
            ", ".join(f"{k}={v[0] / max(v[1], 1):.3f}ms x{v[1]}" for k, v in timings.items()))
        return dt * 1000 / steps, timings, enc, terms

    ms_per_step, timings, enc, terms = run(args.mode, args.steps, args.warmup)
    value = N / (ms_per_step / 1000)

    # roofline of the dominant kernel: algorithmic flop per launch / average launch time
    kname, (kms, kcnt) = max(timings.items(), key=lambda kv: kv[1][0])
    avg_ms = kms / max(kcnt, 1)
    launches_per_step = kcnt / args.steps
    desc_per_launch = N / launches_per_step
    flop_per_launch = KERNEL_FLOP_PER_DESC.get(kname, 8 * DIM * K_COMP) * desc_per_launch
    achieved = flop_per_launch / (avg_ms / 1000) / 1e12
    peak = PEAK_FP32_TFLOPS if args.mode in ("fp32", "fp32mfma") else PEAK_F16_TFLOPS
    bpd = traffic_per_descriptor(kname)
    roofline = {"bound": "mfma", "achieved": achieved, "peak": peak, "unit": "TFLOP/s", "frac": achieved / peak,
                "traffic": bpd * desc_per_launch if bpd else None, "kernel": kname, "kernel_avg_ms": avg_ms,
                "launches_per_step": launches_per_step, "descriptors_per_launch": desc_per_launch,
                "flop_per_descriptor": KERNEL_FLOP_PER_DESC.get(kname, 8 * DIM * K_COMP),
                "kernels_ms_per_step": {k: v[0] / args.steps for k, v in timings.items() if v[1]}}
    if kname in EXECUTED_PER_ALGORITHMIC:
        ex = achieved * EXECUTED_PER_ALGORITHMIC[kname]
        roofline["executed_matrix_tflops"] = ex
        roofline["executed_frac"] = ex / peak
    # every matrix-core kernel of the step, and the step as a whole (8DK
    # algorithmic flop per descriptor over the whole step's wall time)
    per_kernel = {}
    for k, (kt_ms, kt_n) in timings.items():
        if k not in KERNEL_FLOP_PER_DESC or not kt_n:
            continue
        k_avg = kt_ms / kt_n
        k_desc = N / (kt_n / args.steps)
        k_tf = KERNEL_FLOP_PER_DESC[k] * k_desc / (k_avg / 1000) / 1e12
        e = {"avg_ms": k_avg, "launches_per_step": kt_n / args.steps, "achieved_tflops": k_tf, "frac": k_tf / peak}
        if k in EXECUTED_PER_ALGORITHMIC:
            e["executed_frac"] = k_tf * EXECUTED_PER_ALGORITHMIC[k] / peak
        per_kernel[k] = e
    roofline["per_kernel"] = per_kernel
    step_tf = 8 * DIM * K_COMP * N / (ms_per_step / 1000) / 1e12
    roofline["step_achieved_tflops"] = step_tf
    roofline["step_frac"] = step_tf / peak

    t_host = terms.cpu().numpy()
    result = {
        "metric": "Fisher descriptors/s at K=1024,D=80: gated BF16 vs plain FP32 GPU, rel-err<=1e-3",
        "value": value, "unit": "descriptors/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        # gated: the gate network runs in bf16, the exact posteriors and moments of
        # every group it does not serve in compensated fp16 (all groups of this workload)
        "dtype": {"fp32": "fp32", "fp32mfma": "fp32", "bf16x3": "bf16", "f16x3": "fp16", "gated": "fp16"}[args.mode],
        "data": "synthetic (seeded C generator vlsynth_sift_like: SIFT-like rectified/clipped Gaussian projected to 80-d, seed 1; GMM fitted by GPU EM)",
        "config": {"workload": "improved Fisher vectors, 4096 images x 3000 descriptors, K=1024, D=80",
                   "K": K_COMP, "D": DIM, "images": n_img, "descriptors_per_image": PER_IMAGE,
                   "mode": args.mode, **({"gate": "80-256-256-1024 GELU perceptron (bf16), top-16 candidates, "
                                                   f"need {args.gate_need}"} if args.mode == "gated" else {})},
        "roofline": roofline,
        "terms_per_descriptor": float(t_host.sum()) / N,
    }
    n_check = max(args.oracle_images, 64 if not args.no_cpu else 0)
    enc_sel = enc[: min(n_img, n_check)].cpu().numpy()
    del enc
    # The plain FP32 GPU comparator: FP32 operands on the FP32 matrix cores,
    # posteriors kept in registers (fp32mfma); the reference-parity FP32 path
    # (VALU arithmetic of the reference, posterior matrix written) beside it.
    enc32 = None
    if args.mode not in ("fp32", "fp32mfma") and not args.no_fp32:
        ms32, t32, e32, _ = run("fp32mfma", 2, 1)
        result["fp32_gpu"] = {"value": N / (ms32 / 1000), "ms_per_step": ms32, "mode": "fp32mfma",
                              "kernels_ms_per_step": {k: v[0] / 2 for k, v in t32.items() if v[1]},
                              "fp32_matrix_ceiling_descriptors_per_s": PEAK_FP32_TFLOPS * 1e12 / (8 * DIM * K_COMP)}
        result["speedup_vs_fp32_gpu"] = ms32 / ms_per_step
        enc32 = e32[: min(n_img, n_check)].cpu().numpy()
        del e32
        msp, tp, _, _ = run("fp32", 1, 1)
        result["fp32_parity_gpu"] = {"value": N / (msp / 1000), "ms_per_step": msp, "mode": "fp32",
                                     "kernels_ms_per_step": {k: v[0] for k, v in tp.items() if v[1]}}
    elif args.mode in ("fp32", "fp32mfma"):
        enc32 = enc_sel
    if args.mode == "gated":
        # the dense compensated-fp16 encoder the gate falls back to, on the same batch
        msd, tsd, _, _ = run("f16x3", 1, 1)
        result["f16x3_gpu"] = {"value": N / (msd / 1000), "ms_per_step": msd,
                               "kernels_ms_per_step": {k: v[0] for k, v in tsd.items() if v[1]}}
    enc_gated = None
    if gate_info is not None:
        if args.mode == "gated":
            msg, enc_gated = ms_per_step, enc_sel
        else:
            va.fisher_gate_stats(reset=True)
            msg, tg, eg, _ = run("gated", 1, 1)
            enc_gated = eg[: min(n_img, n_check)].cpu().numpy()
            del eg
        groups, served = va.fisher_gate_stats(reset=True)
        gate_info.update({"value": N / (msg / 1000), "ms_per_step": msg,
                          "served_group_fraction": served / max(groups, 1),
                          "speedup_vs_fp32_gpu": (result["fp32_gpu"]["ms_per_step"] / msg) if "fp32_gpu" in result
                          else None})
        result["gated_gpu"] = gate_info

    if not args.no_project:
        result["projection_load_stage"] = projection_leg(L, x, N, args.steps)

    if not args.no_cpu and enc32 is not None:
        x_host = x[: min(N, 1024 * PER_IMAGE)].cpu().numpy()
        result.update(cpu_legs(x_host, offsets, m, c, p, args.cpu_budget, enc_sel, enc32, args.oracle_images,
                               enc_gated))
        result["cpu_baseline"] = result.get("cpu_baseline")

    if args.copy_probe:
        a = torch.empty(1 << 28, dtype=torch.float32, device=dev)
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        log("copy probe: 1 GiB read + 1 GiB written by one copy kernel")
    if gate is not None:
        L.vlhip_fisher_set_gate(None, 16, C.c_float(args.gate_need))
        L.vlhip_gate_delete(gate)
    print(json.dumps(result), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
