#!/usr/bin/env python3
"""Benchmark: the 768-ciphertext batched bootstrap of one BERT-base layer on one MI355X.

Workload: the bootstrapping segment of an encrypted BERT-base layer, 768 ciphertexts
(hidden 768, column packing) at the reference parameters of test_shared_a.cuh:1207-1274:
N = 2^16, logn = 15, Q = 1 x 51 + 20 x 46 + 14 x 51 bits, P = 5 x 51 bits (dnum 7), secret
Hamming weight 192, scale 2^46, inputs with 4 primes, slots U(-1, 1); S2C-first, three
CoeffToSlot/SlotToCoeff layers, EvalMod by a degree-63 Chebyshev cosine + 2 double angles.
The batched form is the shared-a bootstrap (n = 8 members per chunk, 96 chunks,
BatchBootstrap.cuh:719-820, each chunk's linear transforms on one shared-a batch and its
EvalMod batched over the 16 real and imaginary parts); inputs are resident in HBM before
any timed region.

One step = one chunk of 8 of the 768 ciphertexts; `steps` chunks are timed back to back
(cycling through the 768 inputs). `value` is the measured wall time of one complete pass
over all 768 ciphertexts (96 chunks in one call), end to end (no extrapolation);
`value_from_steps_s` is ms_per_step x 96 for comparison.

`bootstrap_768` carries every form over all 768 ciphertexts, each measured end to end:
the plain per-ciphertext form (768 sequential single-ciphertext calls, each waited for, the
reference's loop), ordinary groups of 8, shared-a n = 4 and shared-a n = 8, with decryption
errors on a sample of outputs. `roofline` is the dominant kernel of a
shared-a chunk, whichever NTT tile pass takes more time per chunk (the plain forward pass
or the fused last pass with the ModDown combine and rescale; the other one is
`roofline_other_kernel`), from HIP events around each of its launches in one eager chunk. `cpu_baseline` is the oracle's key-switch
primitives at the same parameters, weighted by the key-switch work the GPU path counts per
chunk, extrapolated to 768 (a lower bound: transforms outside key switching and the
elementwise work are not counted).

`matmul_192` and `full_width_768` keep the MOAI ct-pt matmul (192 x 192 and 768 x 768 at
N = 2^16, L = 24 x 50-bit primes, incl. rescale) with the INT8-MFMA contraction's roofline;
`ntt` and `keyswitch` give the HBM rate of the transform and of one top-level rotation;
`serving` drives the bootstrapping server with requests of uneven sizes.
"""
from __future__ import annotations

import argparse
import gc
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "amortized s/input per BERT layer + 768-ct batched bootstrap s; HBM GB/s % peak"
HBM_PEAK_GBS = 8000.0          # MI355X HBM3E
INT8_PEAK_TOPS = 5000.0        # MI355X dense INT8 MFMA (POP/s x 1e3)
LOG_N, L, ROW_W, COL_W, INPUTS = 16, 24, 192, 192, 256
BITS = [50] * L + [60]
SCALE = 2.0 ** 40
BOOT_BITS = [51] + [46] * 20 + [51] * 14 + [51] * 5
BOOT_SCALE = 2.0 ** 46
N_CTS = 768
REF_SHARED_A_N8_S = 193.3      # reference, 768 cts shared-a n = 8, RTX PRO 6000 (README.md:32-34)

TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "contract_int8_192x192_N65536_L24_pmc.txt")
BOOT_TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "ntt_tile_passes_shared_a_n8_c3_pmc.txt")
# the NTT tile passes as the profiles name them (lazy integer butterflies, 512-thread
# blocks): the plain forward pass and the fused rescale / ModDown pass
PLAIN_TILE_KERNEL = "k_ntt_blocks<12,true,0,true,512>"
RESCALE_TILE_KERNEL = "k_ntt_blocks<12,true,2,true,512>"
TILE_NOTE = ("12 stages of 64-bit integer Harvey butterflies through LDS per 4096-word tile (4 rounds of radix-8 "
             "units; the two wave-uniform rounds read their twiddles through scalar loads); bound by VALU issue of "
             "the 64-bit modular products, not by HBM (the same tile's load and store alone run near peak)")


def measured_traffic(kernel: str, path: str = TRAFFIC_PROFILE):
    """HBM bytes per launch of `kernel` from a committed rocprofv3 PMC summary
    (FETCH_SIZE and WRITE_SIZE runs scaled by a copy calibration), or None."""
    try:
        with open(path) as f:
            for line in f:
                if line.startswith(f"traffic_bytes_per_launch ({kernel})"):
                    return int(line.split(":")[1].split()[0])
    except OSError:
        pass
    return None


def host_cpu():
    model = platform.processor() or "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        affinity = len(os.sched_getaffinity(0))
    except AttributeError:
        affinity = os.cpu_count() or 1
    return {"model": model, "nproc": affinity, "os_cpu_count": os.cpu_count()}


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def make_inputs(ctx, shm, rng):
    N = ctx.N
    xs, hs = [], []
    for _ in range(ROW_W):
        a = np.empty((2, L, N), np.uint64)
        for k in range(L):
            a[:, k, :] = rng.integers(0, int(ctx.coeff_modulus[k]), size=(2, N), dtype=np.uint64)
        a = a.ravel()
        ct = shm.Ciphertext(ctx, 2, 1).load(a)
        ct.scale = SCALE
        xs.append(a)
        hs.append(ct)
    return xs, hs


def cpu_baseline_matmul(xs, W, threads: int):
    from oracle import refcpu
    refcpu.build()
    refcpu.set_threads(threads)
    ref = refcpu.Ref(LOG_N, BITS, 1)
    # output columns in chunks of 8 until about 10 s of wall time (or all of them)
    done, dt = 0, 0.0
    while done < COL_W and dt < 10.0:
        k = min(8, COL_W - done)
        t0 = time.perf_counter()
        ref.ctpt_matmul(xs, W, L, np.full(ROW_W, SCALE), done, done + k)
        dt += time.perf_counter() - t0
        done += k
    per_matmul = dt / done * COL_W
    return {"value": per_matmul / INPUTS, "unit": "s/input", "cores": threads, "kind": "port",
            "sample": f"{done} of {COL_W} output columns of the same 192x192 matmul at full shape "
                      f"(N=2^16, L=24, incl. rescale) on {threads} threads = {dt:.2f} s, "
                      f"extrapolated x{COL_W / done:.2f}",
            "seconds_per_full_matmul": per_matmul}


def ntt_bench(shm, ctx, iters: int = 10):
    """Forward + inverse NTT of one batch of 64 polynomials x 24 limbs (N = 2^16).

    Algorithmic bytes of one transform: 16 B per word (read + write once, the
    per-NTT figure 16 * limbs * N); the two-pass kernels move twice that."""
    # four zero-filled buffers of 16 polynomials x 24 limbs (one launch each)
    bufs = [shm.SharedABatch(ctx, 15, 1) for _ in range(4)]
    ptrs, per, npoly = [b.data_ptr() for b in bufs], 16, 64

    def once():
        for p in ptrs:
            ctx.ntt_forward(p, per, L, 0)
        for p in ptrs:
            ctx.ntt_inverse(p, per, L, 0)

    once()
    shm.synchronize()
    t0 = time.perf_counter()
    for _ in range(iters):
        once()
    shm.synchronize()
    ms = (time.perf_counter() - t0) / (2 * iters) * 1e3
    words = npoly * L * ctx.N
    alg_gbs = 16 * words / (ms * 1e-3) / 1e9
    log(f"bench: NTT {npoly}x{L} limbs {ms:.3f} ms per transform, {alg_gbs:.0f} GB/s algorithmic")
    return {"limbs": npoly * L, "ms_per_transform": round(ms, 4), "algorithmic_bytes": 16 * words,
            "achieved_gbs": round(alg_gbs, 1), "frac": round(alg_gbs / HBM_PEAK_GBS, 4),
            "moved_gbs_two_pass": round(2 * alg_gbs, 1)}


def keyswitch_bench(shm, ctx, sk, gk, step: int, enc, iters: int = 10):
    """Rotation (one hybrid key switch) of a top-level ciphertext of the
    bootstrapping parameters. Algorithmic bytes: 8N [Ql + 2 beta (Ql + P)
    + 2 Ql + 2 Ql] (input limbs, key, read c0/c1, write)."""
    ev = shm.Evaluator(ctx)
    ct = sk.encrypt_symmetric(enc.encode(np.zeros(ctx.N // 2), 2.0 ** 46, 1), seed=9)
    ev.rotate_vector_inplace(ct, step, gk)
    shm.synchronize()
    t0 = time.perf_counter()
    for _ in range(iters):
        ev.rotate_vector_inplace(ct, step, gk)
    shm.synchronize()
    ms = (time.perf_counter() - t0) / iters * 1e3
    Ql, P = ctx.size_q, ctx.size_p
    beta = -(-Ql // P)
    alg = 8 * ctx.N * (Ql + 2 * beta * (Ql + P) + 2 * Ql + 2 * Ql)
    gbs = alg / (ms * 1e-3) / 1e9
    log(f"bench: key switch (rotate, Ql={Ql}, dnum={beta}) {ms:.3f} ms, {gbs:.0f} GB/s algorithmic")
    return {"Ql": Ql, "dnum": beta, "ms": round(ms, 4), "algorithmic_bytes": alg, "achieved_gbs": round(gbs, 1),
            "frac": round(gbs / HBM_PEAK_GBS, 4)}


def serving_bench(shm, boot, gk, rk, cts, n_req: int = 12, seed: int = 9, workers: int = 1):
    """Continuous batching: n_req requests of 1..4 ciphertexts arrive at random
    (exponential gaps, mean 150 ms) from a client thread while earlier ones run;
    the server groups slots of different requests (up to 8 per group), admits by the
    device's free memory, and runs groups on `workers` streams."""
    import threading
    rng = np.random.default_rng(seed)
    sizes = [int(x) for x in rng.integers(1, 5, n_req)]
    gaps = rng.exponential(0.150, n_req)
    srv = shm.BootstrapServer(boot, gk, rk, mem_budget=0, max_group=8, workers=workers)
    ids = []

    def client():
        k = 0
        for m, g in zip(sizes, gaps):
            time.sleep(float(g))
            ids.append(srv.submit([cts[(k + i) % len(cts)] for i in range(m)]))
            k += m

    # one pass of the same sequence first: the group shapes it forms get autotuned there
    warm = shm.BootstrapServer(boot, gk, rk, mem_budget=0, max_group=8, workers=workers)
    wids, k = [], 0
    for m in sizes:
        wids.append(warm.submit([cts[(k + i) % len(cts)] for i in range(m)]))
        k += m
    for rid in wids:
        warm.wait(rid)
    warm.close()
    t0 = time.perf_counter()
    th = threading.Thread(target=client)
    th.start()
    th.join()
    for rid in ids:
        srv.wait(rid)
    wall = time.perf_counter() - t0
    st = srv.stats()
    srv.close()
    log(f"bench: serving {n_req} requests / {sum(sizes)} ciphertexts: p50 {st['p50_ms']:.0f} ms, "
        f"p99 {st['p99_ms']:.0f} ms, {st['groups']} groups")
    return {"requests": n_req, "ciphertexts": sum(sizes), "sizes": sizes, "mean_gap_ms": 150, "workers": workers,
            "p50_ms": round(st["p50_ms"], 2), "p99_ms": round(st["p99_ms"], 2), "groups": int(st["groups"]),
            "ciphertexts_per_s": round(sum(sizes) / wall, 2), "gpu_busy_frac": round(st["busy_s"] / wall, 3),
            "admit_cap_gib": round(st["admit_cap_bytes"] / 2 ** 30, 1)}


def serving_packed_bench(shm, ctx, sk, boot, gk_packed, rk, n_req: int = 24, rate: float = 4.0, cols: int = 8,
                         seed: int = 11, workers: int = 1):
    """Serving with shared slot packing, C5's arrival process on the bootstrapping
    segment: Poisson arrivals (mean `rate` requests/s), each request one input of
    U[16, 128] tokens whose `cols` column ciphertexts (batch_input layout, data in
    slots [0, tokens)) are packed with the requests that queue up behind the running
    work into shared ciphertexts at 128-slot-or-smaller aligned blocks."""
    import threading
    rng = np.random.default_rng(seed)
    n = ctx.N // 2
    toks = [int(x) for x in rng.integers(16, 129, n_req)]
    gaps = rng.exponential(1.0 / rate, n_req)
    enc = shm.Encoder(ctx)
    reqs = []
    for r, L in enumerate(toks):
        row = []
        for c in range(cols):
            z = np.zeros(n)
            z[:L] = rng.uniform(-1, 1, L)
            row.append(sk.encrypt_symmetric(enc.encode(z, BOOT_SCALE, ctx.chain(4)), seed=9000 + r * cols + c))
        reqs.append((L, z, row))
    shm.synchronize()
    srv = shm.BootstrapServer(boot, gk_packed, rk, mem_budget=0, max_group=8, workers=workers)
    ids = []

    def client():
        for (L, _, row), g in zip(reqs, gaps):
            time.sleep(float(g))
            ids.append(srv.submit_packed(row, L))

    t0 = time.perf_counter()
    th = threading.Thread(target=client)
    th.start()
    th.join()
    res = [srv.wait(rid) for rid in ids]
    wall = time.perf_counter() - t0
    st = srv.stats()
    srv.close()
    dec = shm.Decryptor(ctx, sk)
    err = 0.0
    for (L, z, _), (outs, _, off) in list(zip(reqs, res))[:4]:
        got = enc.decode(dec.decrypt(outs[cols - 1]))
        err = max(err, float(np.max(np.abs(got[off:off + L] - z[:L]))))
    packs = st["shared_cts"] // cols
    log(f"bench: packed serving {n_req} requests (Poisson {rate}/s, tokens 16..128, {cols} columns): "
        f"p50 {st['p50_ms']:.0f} ms, p99 {st['p99_ms']:.0f} ms, {packs} packs, max err {err:.2e}")
    return {"requests": n_req, "arrival": f"Poisson, mean {rate} requests/s", "tokens": "uniform 16..128",
            "columns_per_request": cols, "workers": workers, "packs": int(packs),
            "requests_per_pack": round(n_req / max(packs, 1), 2),
            "bootstraps_saved_frac": round(1.0 - st["shared_cts"] / (n_req * cols), 3),
            "p50_ms": round(st["p50_ms"], 2), "p99_ms": round(st["p99_ms"], 2),
            "requests_per_s": round(n_req / wall, 2), "gpu_busy_frac": round(st["busy_s"] / wall, 3),
            "max_abs_err_sampled": err}


def full_width_bench(shm, ctx, hs, steps: int = 2):
    """One 768 x 768 matmul (the BERT-base projection width; rows contracted in
    three chunks of 256) on 768 distinct device ciphertexts: the 192 inputs and
    three device copies of each (limb values do not change the integer work)."""
    R = 768
    X = list(hs) + [h.copy() for _ in range(3) for h in hs]
    W = np.random.default_rng(7).normal(0.0, 0.02, (R, R))
    out = [shm.Ciphertext(ctx, 2, 2) for _ in range(R)]
    shm.ct_pt_matrix_mul_wo_pre(X, W, R, R, R, ctx, out=out)
    log("bench: 768x768 warmup done")
    ks = []
    t0 = time.perf_counter()
    for _ in range(steps):
        shm.ct_pt_matrix_mul_wo_pre(X, W, R, R, R, ctx, out=out)
        ks.append(shm.last_matmul_kernel_ms())
    ms = (time.perf_counter() - t0) / steps * 1e3
    assert shm.last_matmul_variant() == shm.MATMUL_INT8
    k_ms = float(np.mean(ks))
    tops = 2 * 35 * R * R * 2 * L * ctx.N / (k_ms * 1e-3) / 1e12
    log(f"bench: 768x768 matmul {ms:.2f} ms/step, contraction {k_ms:.2f} ms")
    return {"workload": "768x768 ct-pt matmul, N=2^16, L=24, incl. rescale", "steps": steps,
            "ms_per_step": round(ms, 3), "s_per_input": ms * 1e-3 / INPUTS,
            "contraction_ms": round(k_ms, 3), "contraction_int8_tops": round(tops, 1),
            "contraction_frac_of_int8_peak": round(tops / INT8_PEAK_TOPS, 4)}


def masked_bench(shm, ctx, hs, cols: int = 16):
    """The masked form (ct_pt_matrix_mul_wo_pre_w_mask) on the 192 inputs and
    `cols` outputs: every term encodes W[j][i] in the slots of a token mask
    (a full polynomial per term), so the time per term is the figure; it is
    extrapolated to the FFN output projection of the reference's layer
    (3072 x 768 terms)."""
    rng = np.random.default_rng(8)
    mask = (rng.random(ctx.N // 2) < 0.7).astype(np.int32)
    W = rng.normal(0.0, 0.02, (ROW_W, cols))
    out = [shm.Ciphertext(ctx, 2, 2) for _ in range(cols)]
    shm.ct_pt_matrix_mul_wo_pre_w_mask(hs[:8], W[:8], mask, 8, cols, 8, ctx, out=out)
    t0 = time.perf_counter()
    shm.ct_pt_matrix_mul_wo_pre_w_mask(hs, W, mask, ROW_W, cols, ROW_W, ctx, out=out)
    dt = time.perf_counter() - t0
    per = dt / (ROW_W * cols)
    log(f"bench: masked matmul {ROW_W}x{cols}: {dt * 1e3:.1f} ms, {per * 1e6:.2f} us per term")
    return {"workload": f"masked ct-pt matmul {ROW_W}x{cols}, N=2^16, L=24, 70% of slots selected, incl. rescale",
            "ms": round(dt * 1e3, 2), "us_per_term": round(per * 1e6, 3),
            "ffn_3072x768_s_extrapolated": round(per * 3072 * 768, 2)}


LAYER_BITS = [51] + [46] * 23 + [51] * 11 + [51] * 5   # test_single_layer.cuh:473-535 (shared-a S2C-first)


def ctct_bench(shm, rows: int = 4):
    """The attention's Q K^T product in column packing (Ct_ct_matrix_mul.cuh:10-181, called
    with col_X = 64 and row_X = 128 in single_att_block.cuh:210) on the reference layer's
    prime chain (35 data primes, 5 special primes, dnum 7) at 24 primes, num_batch 256:
    output 0 has no rotation and every other output takes 64; one output alone and `rows`
    outputs are timed, giving t(output 0) + 127 t(rotated output) for the 128 outputs."""
    ctx = shm.Context(16, LAYER_BITS, 5)
    sk = shm.SecretKey(ctx, seed=3, hamming_weight=192)
    nb, Ql = 256, 24
    gk = sk.create_galois_keys_from_steps([i * nb for i in range(1, rows)], seed=4)
    rk = sk.gen_relinkey(seed=5)
    rng = np.random.default_rng(9)
    chain = ctx.chain(Ql)

    def rand_ct():
        a = np.concatenate([rng.integers(0, int(ctx.coeff_modulus[k]), size=ctx.N, dtype=np.uint64)
                            for k in range(Ql)] * 2)
        c = shm.Ciphertext(ctx, 2, chain).load(a)
        c.scale = 2.0 ** 46
        return c

    base = [rand_ct() for _ in range(8)]
    X = [base[i % 8].copy() for i in range(64)]
    W = [base[(i + 3) % 8].copy() for i in range(64)]
    out = [shm.Ciphertext(ctx, 2, chain) for _ in range(rows)]
    shm.ct_ct_matrix_mul_colpacking(X, W, gk, rk, ctx, 64, 2, 64, 2, nb, out=out[:2])
    t1, _ = timed(shm, lambda: shm.ct_ct_matrix_mul_colpacking(X, W, gk, rk, ctx, 64, 1, 64, 1, nb, out=out[:1]))
    tr, _ = timed(shm, lambda: shm.ct_ct_matrix_mul_colpacking(X, W, gk, rk, ctx, 64, rows, 64, rows, nb, out=out))
    t_rot = (tr - t1) / (rows - 1)
    log(f"bench: ct-ct QK^T colpacking: output 0 {t1 * 1e3:.1f} ms, rotated output {t_rot * 1e3:.1f} ms")
    return {"workload": "ct-ct column-packing product, col_X=64, N=2^16, 24 of the layer's primes "
                        "(P=5, dnum 7), num_batch=256",
            "output0_ms": round(t1 * 1e3, 2), "rotated_output_ms": round(t_rot * 1e3, 2),
            "outputs_timed": rows,
            "qk_128_outputs_s_estimated": round(t1 + 127 * t_rot, 2)}


def boot_setup(shm, log_t0):
    ctx = shm.Context(16, BOOT_BITS, 5)
    sk = shm.SecretKey(ctx, seed=1, hamming_weight=192)
    boot = shm.Bootstrapper(ctx)
    steps = boot.rotation_steps()
    gk = sk.create_galois_keys_from_steps(steps, seed=2)
    rk = sk.gen_relinkey(seed=3)
    boot.prepare()
    log(f"bench: bootstrap keys + diagonals ready ({time.perf_counter() - log_t0:.1f} s)")
    return ctx, sk, boot, steps, gk, rk


def make_boot_inputs(shm, ctx, sk, count):
    enc = shm.Encoder(ctx)
    rng = np.random.default_rng(7)
    zs = [rng.uniform(-1, 1, ctx.N // 2) for _ in range(count)]
    cts = [sk.encrypt_symmetric(enc.encode(z, BOOT_SCALE, ctx.chain(4)), seed=100 + i) for i, z in enumerate(zs)]
    shm.synchronize()
    return zs, cts


def max_err(shm, ctx, sk, outs, zs, sample):
    enc, dec = shm.Encoder(ctx), shm.Decryptor(ctx, sk)
    return max(float(np.max(np.abs(enc.decode(dec.decrypt(outs[i]), complex_out=True) - zs[i]))) for i in sample)


def timed(shm, fn):
    shm.synchronize()
    t0 = time.perf_counter()
    r = fn()
    shm.synchronize()
    return time.perf_counter() - t0, r


def cpu_baseline_c3(shm, counters, chunk_cts, threads):
    """The oracle's key-switch primitives at the bootstrap parameters (ModUp of one
    polynomial, an inner product with two outputs, ModDown of one polynomial), timed at
    three levels and interpolated in Ql, weighted by the key-switch work the GPU path
    counted over one chunk, extrapolated to 768 ciphertexts."""
    from oracle import refcpu
    refcpu.build()
    refcpu.set_threads(threads)
    ref = refcpu.Ref(16, BOOT_BITS, 5)
    rng = np.random.default_rng(11)
    N, nP = ref.N, ref.nP

    def limbs(idx):
        return np.concatenate([rng.integers(0, int(ref.primes[k]), N, dtype=np.uint64) for k in idx])

    key = limbs(list(range(ref.nQP)) * (2 * ref.dnum()))
    meas = {}
    t_all = time.perf_counter()
    for Ql in (4, 22, 35):
        c2 = limbs(range(Ql))
        t0 = time.perf_counter()
        d = ref.modup(c2, Ql)
        t_up = time.perf_counter() - t0
        t0 = time.perf_counter()
        cx = ref.ks_inner(d, key, Ql)
        t_in = time.perf_counter() - t0
        t0 = time.perf_counter()
        ref.moddown(cx[:(Ql + nP) * N], Ql)
        t_dn = time.perf_counter() - t0
        meas[Ql] = (t_up, t_in / 2, t_dn)
    sample_s = time.perf_counter() - t_all
    pts = sorted(meas)

    def interp(Ql, i):
        xs = np.array(pts, float)
        ys = np.array([meas[p][i] for p in pts])
        return float(np.interp(Ql, xs, ys))

    per_chunk = sum(up * interp(q, 0) + inner * interp(q, 1) + dn * interp(q, 2)
                    for q, (up, inner, dn) in counters.items())
    est = per_chunk * N_CTS / chunk_cts
    return {"value": round(est, 1), "unit": "s per 768-ct batched bootstrap (extrapolated)", "cores": threads,
            "kind": "port",
            "sample": f"oracle ModUp / inner product / ModDown at Ql = 4, 22, 35 (N=2^16, P=5) on {threads} "
                      f"threads, {sample_s:.1f} s; weighted by the GPU path's key-switch counts over one "
                      f"shared-a chunk of {chunk_cts} and extrapolated x{N_CTS // chunk_cts}; key switching "
                      "only (a lower bound)",
            "per_level_s": {str(q): [round(v, 4) for v in meas[q]] for q in pts}}


def c1_pipeline(shm, threads, iters: int = 20):
    """NTT -> rotate(1) -> multiply_plain -> rescale -> iNTT at N = 2^13, on the
    GPU (through the C ABI) and on the oracle (all threads), end to end."""
    from oracle import refcpu
    refcpu.set_threads(threads)
    bits = [50] * 8 + [60]
    ctx = shm.Context(13, bits, 1)
    ref = refcpu.Ref(13, bits, 1)
    rng = np.random.default_rng(3)
    N, Ql = ctx.N, 8
    x = np.concatenate([rng.integers(0, int(ref.primes[k]), size=N, dtype=np.uint64) for k in range(Ql)] * 2)
    sk = shm.SecretKey(ctx, seed=1, hamming_weight=64)
    gk = sk.create_galois_keys_from_steps([1], seed=9)
    s_ref = ref.sk_ntt(ref.sample_ternary(1, 64))
    elt = ref.elt(1)
    key = ref.gen_galois_key(s_ref, elt, ref.splitmix64(9, elt))
    ev = shm.Evaluator(ctx)
    pt = shm.Encoder(ctx).encode(0.37, 2.0 ** 50, 1)
    cts = [shm.Ciphertext(ctx, 2, 1).load(x) for _ in range(iters + 1)]

    def gpu(ct):
        ctx.ntt_forward(ct.data_ptr(), 2, Ql, 0)
        ev.rotate_vector_inplace(ct, 1, gk)
        ev.multiply_plain_inplace(ct, pt)
        ev.rescale_to_next_inplace(ct)
        ctx.ntt_inverse(ct.data_ptr(), 2, Ql - 1, 0)

    gpu(cts[0])
    dt, _ = timed(shm, lambda: [gpu(c) for c in cts[1:]])
    gpu_ms = dt / iters * 1e3
    pt_ref = ref.encode_scalar(0.37, Ql, 2.0 ** 50)
    t0 = time.perf_counter()
    y = ref.ntt(x, np.tile(ref.pidx(Ql), 2))
    y = ref.rotate(y, Ql, elt, key)
    y = ref.multiply_plain(y, pt_ref, Ql)
    y = ref.rescale(y, Ql)
    y = ref.intt(y, np.tile(ref.pidx(Ql - 1), 2))
    cpu_ms = (time.perf_counter() - t0) * 1e3
    assert np.array_equal(cts[1].numpy(), y), "C1 pipeline: GPU differs from the oracle"
    log(f"bench: C1 pipeline GPU {gpu_ms:.3f} ms, oracle {cpu_ms:.1f} ms on {threads} threads")
    return {"workload": "NTT -> rotate(1) -> multiply_plain -> rescale -> iNTT, N=2^13, 8 x 50-bit + 60-bit",
            "gpu_ms": round(gpu_ms, 4), "cpu_ms": round(cpu_ms, 2), "cpu_threads": threads, "bitexact": True}


def matmul_bench(shm, args, threads):
    """The MOAI ct-pt matmul at C2 (192 x 192) with the contraction and rescale rooflines."""
    variant = shm.MATMUL_INT8 if args.variant == "int8" else shm.MATMUL_U64
    ctx = shm.Context(LOG_N, BITS, 1)
    rng = np.random.default_rng(5)
    log("bench: generating 192 input ciphertexts (N=2^16, L=24) and uploading to HBM")
    xs, hs = make_inputs(ctx, shm, rng)
    W = np.random.default_rng(6).normal(0.0, 0.02, (ROW_W, COL_W))
    out = [shm.Ciphertext(ctx, 2, 2) for _ in range(COL_W)]
    for w in range(2):
        shm.ct_pt_matrix_mul_wo_pre(hs, W, ROW_W, COL_W, ROW_W, ctx, variant=variant, out=out)
    if args.copy_calibration:
        hs[0].copy()
        log(f"bench: calibration copy of {hs[0].words() * 8} bytes done")
    kernel_ms, tile_ms = [], []
    steps = 10
    # per-kernel times from the serial form (contraction, then rescale)...
    shm.set_tuning("matmul_overlap", 0)
    t0 = time.perf_counter()
    for s in range(steps):
        shm.ct_pt_matrix_mul_wo_pre(hs, W, ROW_W, COL_W, ROW_W, ctx, variant=variant, out=out)
        kernel_ms.append(shm.last_matmul_kernel_ms())
        tile_ms.append(shm.last_matmul_rescale_tile_ms())
    ms_per_step = (time.perf_counter() - t0) / steps * 1e3  # every call synchronises before returning
    ref_out = [o.numpy() for o in out[:2]]
    # ...and the overlapped form (contraction in limb windows with the rescale on a
    # second stream), which gives the same limbs
    shm.set_tuning("matmul_overlap", 1)
    shm.ct_pt_matrix_mul_wo_pre(hs, W, ROW_W, COL_W, ROW_W, ctx, variant=variant, out=out)
    assert all(np.array_equal(a, o.numpy()) for a, o in zip(ref_out, out[:2]))
    t0 = time.perf_counter()
    for s in range(steps):
        shm.ct_pt_matrix_mul_wo_pre(hs, W, ROW_W, COL_W, ROW_W, ctx, variant=variant, out=out)
    overlap_ms = (time.perf_counter() - t0) / steps * 1e3
    shm.set_tuning("matmul_overlap", 0)
    k_ms, t_ms = float(np.mean(kernel_ms)), float(np.mean(tile_ms))
    N = ctx.N
    alg_step = ROW_W * 2 * L * N * 8 + COL_W * 2 * (L - 1) * N * 8   # read X, write Y (9.46 GB)
    contraction_bytes = ROW_W * 2 * L * N * 8 + COL_W * 2 * L * N * 8  # what the contraction kernel moves
    macs = ROW_W * COL_W * 2 * L * N
    tops = 2 * 35 * macs / (k_ms * 1e-3) / 1e12
    traffic_c = measured_traffic("k_contract_int8")
    traffic_t = measured_traffic(RESCALE_TILE_KERNEL)
    res = {"workload": "quarter-width ct-pt matmul 192x192 (MOAI column packing), N=2^16, L=24 x 50-bit primes, "
                       "256 inputs x 128 tokens per ciphertext, incl. rescale",
           "variant": args.variant, "last_variant": shm.last_matmul_variant(), "steps": steps,
           "ms_per_step": round(ms_per_step, 4), "s_per_input": ms_per_step * 1e-3 / INPUTS,
           "ms_per_step_overlapped": round(overlap_ms, 4),
           "overlap_note": "the overlapped form runs the contraction in 3 limb windows with the rescale on a second "
                           "stream (same limbs); both kernels fill the GPU, so it is not faster and the serial "
                           "form is the default",
           "algorithmic_bytes_per_step": alg_step,
           "step_hbm_gbs_algorithmic": round(alg_step / (ms_per_step * 1e-3) / 1e9, 1),
           "step_frac_of_hbm_peak": round(alg_step / (ms_per_step * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
           "contraction": {"bound": "mfma", "kernel": "k_contract_int8", "kernel_ms": round(k_ms, 4),
                           "achieved": round(tops, 1), "peak": INT8_PEAK_TOPS, "unit": "TOP/s",
                           "frac": round(tops / INT8_PEAK_TOPS, 4), "traffic": traffic_c,
                           "kernel_bytes_per_launch": contraction_bytes},
           "rescale_tile_pass": {"kernel": RESCALE_TILE_KERNEL, "kernel_ms": round(t_ms, 4),
                                 "traffic": traffic_t}}
    if traffic_c and traffic_t:
        res["traffic_over_algorithmic_step"] = round((traffic_c + traffic_t) / alg_step, 3)
    log(f"bench: matmul 192x192 {ms_per_step:.3f} ms/step ({overlap_ms:.3f} overlapped), contraction "
        f"{k_ms:.3f} ms, rescale tile {t_ms:.3f} ms")
    if not args.quick:
        t1 = time.perf_counter()
        shm.ct_pt_matrix_mul_wo_pre(hs, W, ROW_W, COL_W, ROW_W, ctx, variant=shm.MATMUL_PER_CT, out=out)
        plain_ms = (time.perf_counter() - t1) * 1e3
        res["plain_per_ct_ms_per_step"] = round(plain_ms, 2)
        res["batched_speedup_vs_plain"] = round(plain_ms / ms_per_step, 2)
        log(f"bench: matmul plain per-ciphertext form {plain_ms:.1f} ms")
        full = full_width_bench(shm, ctx, hs)
        full["masked"] = masked_bench(shm, ctx, hs)
        full["ctct_qk"] = ctct_bench(shm)
        res["full_width_768"] = full
        res["ntt"] = ntt_bench(shm, ctx)
    if not args.no_cpu_baseline:
        res["cpu_baseline"] = cpu_baseline_matmul(xs, W, threads)
    return res


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--variant", choices=["int8", "u64"], default="int8")
    ap.add_argument("--quick", action="store_true",
                    help="headline only: the shared-a n=8 form (steps and one 768 pass)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-matmul", action="store_true", help="skip the ct-pt matmul objects")
    ap.add_argument("--matmul-only", action="store_true",
                    help="only the ct-pt matmul objects (profiling); prints their JSON, not the headline line")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run one device-to-device copy of a ciphertext (profiler byte-counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on exactly one GPU; --gpus {args.gpus} is not supported: the 768 ciphertexts "
              "would shard by chunk across GPUs with the keys replicated on each, and only one GPU is measured here",
              file=sys.stderr)
        return 2

    import shmoai_amd as shm
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or min(16, os.cpu_count() or 1)
    if args.matmul_only:
        print(json.dumps({"matmul_192": matmul_bench(shm, args, threads)}), flush=True)
        return 0
    cpu_info = host_cpu()
    t_setup = time.perf_counter()
    ctx, sk, boot, steps, gk, rk = boot_setup(shm, t_setup)
    keys8 = shm.SharedAKeys(ctx, sk, 8, 192, 12, ctx.chain(4))
    keys8.add_galois(steps, seed=13)
    log(f"bench: shared-a n=8 keys resident ({time.perf_counter() - t_setup:.1f} s)")
    zs, cts = make_boot_inputs(shm, ctx, sk, N_CTS)
    log(f"bench: {N_CTS} input ciphertexts encrypted and resident in HBM")
    outs = [shm.Ciphertext(ctx, 2, 1) for _ in range(N_CTS)]
    sample = [0, 7, 100, 383, 384, 500, 760, 767]

    # headline form: shared-a n = 8, 96 chunks through the chunked entry point
    nch = N_CTS // 8

    def step(k):
        c = k % nch
        shm.bootstrap_shared_a_s2c_first(boot, keys8, cts[8 * c:8 * (c + 1)], rk, out=outs[8 * c:8 * (c + 1)])

    for w in range(args.warmup):
        step(w)
        log(f"bench: warmup step {w + 1}/{args.warmup} done")
    shm.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        step(k)  # each call waits for its chunk
    shm.synchronize()
    dt = time.perf_counter() - t0
    ms_per_step = dt / args.steps * 1e3
    log(f"bench: {args.steps} steps (chunks of 8), {ms_per_step:.2f} ms/step")
    full_s, _ = timed(shm, lambda: shm.bootstrap_shared_a_s2c_first_chunks(boot, keys8, cts, rk, out=outs))
    err8 = max_err(shm, ctx, sk, outs, zs, sample)
    log(f"bench: shared-a n=8, all {N_CTS} in one chunked call: {full_s:.2f} s, max err {err8:.2e}")
    boot768 = {"params": "N=2^16, logn=15, Q=1x51+20x46+14x51, P=5x51 (dnum 7), HW 192, scale 2^46, S2C-first, "
                         "EvalMod deg-63 Chebyshev + 2 double angles, output 22 primes",
               "shared_a_n8_s": round(full_s, 3), "max_abs_err_shared_a_n8": err8}

    # roofline of the dominant kernel and the key-switch work of one eager chunk
    shm.keyswitch_counters_reset()
    shm.probe_fused_ntt_begin()
    shm.bootstrap_shared_a_s2c_first(boot, keys8, cts[:8], rk, out=outs[:8])
    pr = shm.probe_fused_ntt_end()
    counters = shm.keyswitch_counters()
    pp = shm.probe_plain_ntt_last()

    def tile_roofline(p, kernel, what, note):
        avg_ms = p["ms"] / max(p["launches"], 1)
        alg = p["bytes"] / max(p["launches"], 1)
        gbs = alg / (avg_ms * 1e-3) / 1e9
        traffic = measured_traffic(kernel, BOOT_TRAFFIC_PROFILE)
        r = {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
             "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": traffic, "kernel": f"{kernel} ({what})",
             "launches_per_chunk": p["launches"], "kernel_ms_avg": round(avg_ms, 5),
             "algorithmic_bytes_per_launch": int(alg), "kernel_ms_per_chunk": round(p["ms"], 2), "note": note}
        if traffic:
            r["traffic_over_algorithmic"] = round(traffic / alg, 3)
        return r

    fused = tile_roofline(pr, RESCALE_TILE_KERNEL, "fused last NTT pass: ModDown combine and rescale",
                          TILE_NOTE)
    plain = tile_roofline(pp, PLAIN_TILE_KERNEL, "forward NTT tile pass: ModUp digits and other forward transforms",
                          TILE_NOTE)
    # the dominant kernel of the chunk (most time) is the roofline object
    roofline, roofline_other = (plain, fused) if pp["ms"] >= pr["ms"] else (fused, plain)
    avg_ms, gbs = fused["kernel_ms_avg"], fused["achieved"]
    log(f"bench: fused NTT pass {pr['launches']} launches, {avg_ms * 1e3:.1f} us avg, {gbs:.0f} GB/s algorithmic; "
        f"plain forward tile pass {pp['launches']} launches, {plain['kernel_ms_avg'] * 1e3:.1f} us avg, "
        f"{plain['achieved']:.0f} GB/s algorithmic")

    if not args.quick:
        # shared-a n = 4
        keys4 = shm.SharedAKeys(ctx, sk, 4, 192, 8, ctx.chain(4))
        keys4.add_galois(steps, seed=9)
        shm.bootstrap_shared_a_s2c_first(boot, keys4, cts[:4], rk, out=outs[:4])  # warm-up
        s4, _ = timed(shm, lambda: shm.bootstrap_shared_a_s2c_first_chunks(boot, keys4, cts, rk, out=outs))
        e4 = max_err(shm, ctx, sk, outs, zs, sample)
        boot768.update({"shared_a_n4_s": round(s4, 3), "max_abs_err_shared_a_n4": e4})
        log(f"bench: shared-a n=4, all {N_CTS}: {s4:.2f} s, max err {e4:.2e}")
        del keys4
        gc.collect()
        # shared-a n = 16 (BatchBootstrapper takes any n; the reference publishes n = 4 and 8)
        keys16 = shm.SharedAKeys(ctx, sk, 16, 192, 56, ctx.chain(4))
        keys16.add_galois(steps, seed=17)
        shm.bootstrap_shared_a_s2c_first(boot, keys16, cts[:16], rk, out=outs[:16])  # warm-up
        s16, _ = timed(shm, lambda: shm.bootstrap_shared_a_s2c_first_chunks(boot, keys16, cts, rk, out=outs))
        e16 = max_err(shm, ctx, sk, outs, zs, sample)
        boot768.update({"shared_a_n16_s": round(s16, 3), "max_abs_err_shared_a_n16": e16})
        log(f"bench: shared-a n=16, all {N_CTS}: {s16:.2f} s, max err {e16:.2e}")
        del keys16
        gc.collect()
        # shared-a n = 8 with EvalMod as a degree-40 Chebyshev cosine + 3 double angles: the
        # same depth (output level) with 13 + 3 instead of 16 + 2 relinearisations per part
        boot40 = shm.Bootstrapper(ctx, deg=40, double_angles=3)
        boot40.prepare()
        shm.bootstrap_shared_a_s2c_first(boot40, keys8, cts[:8], rk, out=outs[:8])  # warm-up
        s40, _ = timed(shm, lambda: shm.bootstrap_shared_a_s2c_first_chunks(boot40, keys8, cts, rk, out=outs))
        e40 = max_err(shm, ctx, sk, outs, zs, sample)
        boot768.update({"shared_a_n8_evalmod_deg40_r3_s": round(s40, 3),
                        "max_abs_err_shared_a_n8_evalmod_deg40_r3": e40,
                        "evalmod_deg40_r3_note": "EvalMod variant (not the headline configuration): "
                                                 "degree-40 Chebyshev cosine + 3 double angles, same output "
                                                 "level; the headline keeps degree 63 + 2"})
        log(f"bench: shared-a n=8, EvalMod deg 40 + 3 double angles, all {N_CTS}: {s40:.2f} s, max err {e40:.2e}")
        del boot40
        gc.collect()
        # ordinary S2C-first, groups of 8 (one call over all 768)
        so, _ = timed(shm, lambda: boot.bootstrap_s2c_first(cts, gk, rk, out=outs))
        eo = max_err(shm, ctx, sk, outs, zs, sample)
        boot768.update({"ordinary_group8_s": round(so, 3), "max_abs_err_ordinary_group8": eo})
        log(f"bench: ordinary groups of 8, all {N_CTS}: {so:.2f} s, max err {eo:.2e}")

        # the plain per-ciphertext form: one call per ciphertext, each waited for
        def sequential():
            for i in range(N_CTS):
                boot.bootstrap_s2c_first([cts[i]], gk, rk, out=[outs[i]])
        ss, _ = timed(shm, sequential)
        es = max_err(shm, ctx, sk, outs, zs, sample)
        best = min(full_s, s4, s16, so)
        boot768.update({"plain_sequential_s": round(ss, 3), "max_abs_err_plain_sequential": es,
                        "best_batched_s": round(best, 3),
                        "batched_speedup_vs_plain": round(ss / best, 3)})
        log(f"bench: plain per-ciphertext form, {N_CTS} sequential calls: {ss:.2f} s, max err {es:.2e}")
        enc = shm.Encoder(ctx)
        boot768["keyswitch"] = keyswitch_bench(shm, ctx, sk, gk, min(s for s in steps if s > 0), enc)
        boot768["serving"] = serving_bench(shm, boot, gk, rk, cts[:64])
        # Galois keys for the packing rotations: signed powers of two from the 16-slot block up
        pk_steps = sorted(set(steps) | {sg * (1 << k) for k in range(4, 15) for sg in (1, -1)})
        gk_packed = sk.create_galois_keys_from_steps(pk_steps, seed=2)
        boot768["serving_packed"] = serving_packed_bench(shm, ctx, sk, boot, gk_packed, rk)
        boot768["serving_packed_overload"] = serving_packed_bench(shm, ctx, sk, boot, gk_packed, rk, n_req=48,
                                                                  rate=16.0, seed=12)
        del gk_packed
        gc.collect()
    boot768["keyswitch_counts_per_chunk_n8"] = {str(q): list(v) for q, v in counters.items()}
    tuned = [ln for ln in shm.autotune_report().splitlines() if "N=65536" in ln]
    picks = {}
    for ln in tuned:
        k = ln.split("-> slots ")[1].split(" ")[0]
        picks[k] = picks.get(k, 0) + 1
    boot768["autotune"] = {"shapes_tuned": len(tuned), "slots_chosen": picks,
                           "examples": [ln for ln in tuned if "-> slots 1 " not in ln][:4]}

    cpu = None
    if not args.no_cpu_baseline:
        log(f"bench: CPU baseline (oracle key-switch primitives) on {threads} threads")
        cpu = cpu_baseline_c3(shm, counters, 8, threads)
        cpu.update(cpu_info)
        cpu["threads_note"] = ("OMP_NUM_THREADS of the GPU box (its CPU share for one GPU)"
                               if os.environ.get("OMP_NUM_THREADS") else "min(16, cpu_count)")

    del outs, cts
    gc.collect()
    extra = {}
    if not args.no_matmul and not args.quick:
        extra["c1_pipeline"] = c1_pipeline(shm, threads)
        extra["matmul_192"] = matmul_bench(shm, args, threads)

    line = {
        "metric": METRIC,
        "value": round(full_s, 3),
        "unit": "s per 768-ct batched bootstrap",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": round(full_s / REF_SHARED_A_N8_S, 4),
        "vs_baseline_note": "value / 193.3 s, the reference's 768-ct shared-a n=8 figure on an RTX PRO 6000 "
                            "(README.md:32-34); below 1 is faster",
        "dtype": "u64 (exact RNS limbs; FP64 only in the encoder)",
        "data": "synthetic (slots U(-1,1), random-init keys)",
        "config": {"workload": "768-ciphertext shared-a S2C-first bootstrap (one BERT-base layer's bootstrapping "
                               "segment), C3: N=2^16, Q=1x51+20x46+14x51, P=5x51, HW 192, scale 2^46; "
                               "96 chunks of n=8",
                   "step": "one chunk of 8 of the 768 ciphertexts", "ciphertexts": N_CTS,
                   "value_from_steps_s": round(ms_per_step * 1e-3 * N_CTS / 8, 3)},
        "roofline": roofline,
        "roofline_other_kernel": roofline_other,
        "cpu_baseline": cpu,
        "bootstrap_768": boot768,
        **extra,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
