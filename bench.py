#!/usr/bin/env python3
"""Benchmark: the quarter-width MOAI ct-pt matmul on one MI355X.

Workload (one step): 192 column-packed input ciphertexts (N = 2^16, 24 x 50-bit
primes, 256 inputs x 128 tokens = 32768 slots each) times a 192 x 192 weight
matrix W ~ N(0, 0.02) encoded at scale 2^40, then rescale
(Ct_pt_matrix_mul.cuh:18-185 semantics). Inputs are resident in HBM before
the timed region; limbs are uniform random residues (synthetic data).

Prints one JSON line: value = amortised seconds per input (step time / 256
inputs), plus the roofline of the step's dominant kernel and of the other large
one (the INT8-MFMA contraction and the tile pass of the fused rescale NTT, each
timed with HIP events on the stream it runs on), and a CPU baseline (the oracle,
on a bounded sample of the same workload, extrapolated).

The line also carries `bootstrap_768`: S2C-first bootstrapping at the reference
parameters (N = 2^16, Q = 1 x 51 + 20 x 46 + 14 x 51 bits, P = 5 x 51 bits,
HW 192, inputs with 4 primes, slots U(-1, 1)), timed on a bounded number of
ciphertexts and extrapolated to 768: ordinary one ciphertext per call (the
sequential form), ordinary in one group of 8 (batched linear transforms and
EvalMod), and shared-a chunks of n = 4 and n = 8, with the decryption error of
the timed outputs, and `serving`: requests of uneven sizes arriving at random
through the bootstrapping server (continuous batching), with p50/p99 latency.

`full_width_768` times one 768 x 768 matmul of the same kind (the BERT-base
projection width, W rows contracted in chunks of 256), and in `masked` the
masked form (a token mask in every weight plaintext) per term.

Two more objects give the HBM rate of the transform and key-switch kernels:
`ntt` (forward + inverse NTT of 64 polynomials x 24 limbs at N = 2^16) and
`bootstrap_768.keyswitch` (one rotation of a top-level ciphertext of the
bootstrapping parameters), each as algorithmic GB/s and fraction of peak.
`traffic` is each kernel's HBM bytes per launch from the committed rocprofv3
PMC summary in profiles/.
"""
from __future__ import annotations

import argparse
import json
import os
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


TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "contract_int8_192x192_N65536_L24_pmc.txt")


def measured_traffic(kernel: str):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 PMC summary
    (FETCH_SIZE and WRITE_SIZE runs scaled by a copy calibration), or None when
    the summary is absent."""
    try:
        with open(TRAFFIC_PROFILE) as f:
            for line in f:
                if line.startswith(f"traffic_bytes_per_launch ({kernel})"):
                    return int(line.split(":")[1].split()[0])
    except OSError:
        pass
    return None


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


def cpu_baseline(xs, W, threads: int):
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


BOOT_BITS = [51] + [46] * 20 + [51] * 14 + [51] * 5


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


def bootstrap_bench(shm, n_seq: int = 3, n_group: int = 8, shared=((4, 2), (8, 1))):
    """Seconds per ciphertext of S2C-first bootstrapping at the reference
    parameters: ordinary one ciphertext per call (the reference's sequential
    form), ordinary in one group of n_group ciphertexts (batched linear
    transforms and EvalMod), and shared-a chunks of n members."""
    ctx = shm.Context(16, BOOT_BITS, 5)
    sk = shm.SecretKey(ctx, seed=1, hamming_weight=192)
    boot = shm.Bootstrapper(ctx)
    t0 = time.perf_counter()
    steps = boot.rotation_steps()
    gk = sk.create_galois_keys_from_steps(steps, seed=2)
    rk = sk.gen_relinkey(seed=3)
    boot.prepare()
    keysets = {}
    for n_sh, _ in shared:
        keysets[n_sh] = shm.SharedAKeys(ctx, sk, n_sh, 192, 4 + n_sh, ctx.chain(4))
        keysets[n_sh].add_galois(steps, seed=5 + n_sh)
    setup_s = time.perf_counter() - t0
    log(f"bench: bootstrap keys + diagonals ready in {setup_s:.1f} s")
    n = ctx.N // 2
    rng = np.random.default_rng(7)
    total = max([n_group, n_seq] + [n_sh * ch for n_sh, ch in shared])
    zs = [rng.uniform(-1, 1, n) for _ in range(total)]
    enc, dec = shm.Encoder(ctx), shm.Decryptor(ctx, sk)
    ks = keyswitch_bench(shm, ctx, sk, gk, min(s for s in steps if s > 0), enc)
    cts = [sk.encrypt_symmetric(enc.encode(z, 2.0 ** 46, ctx.chain(4)), seed=100 + i) for i, z in enumerate(zs)]

    def max_err(outs):
        return max(float(np.max(np.abs(enc.decode(dec.decrypt(o), complex_out=True) - z)))
                   for o, z in zip(outs[:2], zs[:2]))

    boot.bootstrap_s2c_first(cts[:1], gk, rk)  # warm-up
    t0 = time.perf_counter()
    outs = [boot.bootstrap_s2c_first([c], gk, rk)[0] for c in cts[:n_seq]]  # each call synchronises
    seq_s = (time.perf_counter() - t0) / n_seq
    err_seq = max_err(outs)
    log(f"bench: ordinary bootstrap, one ciphertext per call: {seq_s * 1e3:.1f} ms/ct, max err {err_seq:.2e}")
    t0 = time.perf_counter()
    outs = boot.bootstrap_s2c_first(cts[:n_group], gk, rk)
    grp_s = (time.perf_counter() - t0) / n_group
    err_grp = max_err(outs)
    log(f"bench: ordinary bootstrap, group of {n_group}: {grp_s * 1e3:.1f} ms/ct, max err {err_grp:.2e}")
    res = {"params": "N=2^16, logn=15, Q=1x51+20x46+14x51, P=5x51 (dnum 7), HW 192, scale 2^46, "
                     "S2C-first, EvalMod deg-63 Chebyshev + 2 double angles, output 22 primes",
           "ordinary_sequential_s_per_ct": round(seq_s, 5),
           "ordinary_sequential_768_s_extrapolated": round(seq_s * 768, 2),
           f"ordinary_group{n_group}_s_per_ct": round(grp_s, 5),
           f"ordinary_group{n_group}_768_s_extrapolated": round(grp_s * 768, 2),
           "max_abs_err_ordinary": max(err_seq, err_grp)}
    best = grp_s
    for n_sh, chunks in shared:
        keys = keysets[n_sh]
        shm.bootstrap_shared_a_s2c_first(boot, keys, cts[:n_sh], rk)  # warm-up
        t0 = time.perf_counter()
        outs = []
        for c in range(chunks):
            outs += shm.bootstrap_shared_a_s2c_first(boot, keys, cts[c * n_sh:(c + 1) * n_sh], rk)
        sh_s = (time.perf_counter() - t0) / (n_sh * chunks)
        err_sh = max_err(outs)
        log(f"bench: shared-a n={n_sh} bootstrap {sh_s * 1e3:.1f} ms/ct, max err {err_sh:.2e}")
        res[f"shared_a_n{n_sh}_s_per_ct"] = round(sh_s, 5)
        res[f"shared_a_n{n_sh}_768_s_extrapolated"] = round(sh_s * 768, 2)
        res[f"max_abs_err_shared_a_n{n_sh}"] = err_sh
        best = min(best, sh_s)
    res.update({"best_768_s_extrapolated": round(best * 768, 2),
                "batched_speedup_vs_sequential": round(seq_s / best, 3),
                "keys_resident_setup_s": round(setup_s, 2), "keyswitch": ks,
                "sample": f"{n_seq} single-ciphertext calls, one group of {n_group}, shared-a chunks "
                          + ", ".join(f"{ch} x n={n_sh}" for n_sh, ch in shared)
                          + "; wall clock incl. host overhead, extrapolated to 768"})
    res["serving"] = serving_bench(shm, boot, gk, rk, cts)
    return res


def serving_bench(shm, boot, gk, rk, cts, n_req: int = 12, seed: int = 9):
    """Continuous batching: n_req requests of 1..4 ciphertexts arrive at random
    (exponential gaps, mean 150 ms) from a client thread while earlier ones run;
    the server groups slots of different requests (up to 8 per group)."""
    import threading
    rng = np.random.default_rng(seed)
    sizes = [int(x) for x in rng.integers(1, 5, n_req)]
    gaps = rng.exponential(0.150, n_req)
    srv = shm.BootstrapServer(boot, gk, rk, mem_budget=64 << 30, max_group=8)
    ids = []

    def client():
        k = 0
        for m, g in zip(sizes, gaps):
            time.sleep(float(g))
            ids.append(srv.submit([cts[(k + i) % len(cts)] for i in range(m)]))
            k += m

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
    return {"requests": n_req, "ciphertexts": sum(sizes), "sizes": sizes, "mean_gap_ms": 150,
            "p50_ms": round(st["p50_ms"], 2), "p99_ms": round(st["p99_ms"], 2), "groups": int(st["groups"]),
            "ciphertexts_per_s": round(sum(sizes) / wall, 2), "gpu_busy_frac": round(st["busy_s"] / wall, 3)}


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


def ctct_bench(shm, ctx, hs, rows: int = 4):
    """The attention's Q K^T product in column packing (Ct_ct_matrix_mul.cuh:10-181,
    called with col_X = 64 and row_X = 128 in single_att_block.cuh:210) on 64 + 64 of
    the resident ciphertexts (N=2^16, L=24), num_batch 256: `rows` of the 128 outputs
    are timed (each is 64 rotations, 64 tensor products, one relinearisation and
    rescale) and extrapolated to 128."""
    sk = shm.SecretKey(ctx, seed=3, hamming_weight=192)
    nb = 256
    gk = sk.create_galois_keys_from_steps([i * nb for i in range(1, rows)], seed=4)
    rk = sk.gen_relinkey(seed=5)
    X, W = hs[:64], hs[64:128]
    out = [shm.Ciphertext(ctx, 2, 2) for _ in range(rows)]
    shm.ct_ct_matrix_mul_colpacking(X, W, gk, rk, ctx, 64, 2, 64, 2, nb, out=out[:2])
    shm.synchronize()
    t0 = time.perf_counter()
    shm.ct_ct_matrix_mul_colpacking(X, W, gk, rk, ctx, 64, rows, 64, rows, nb, out=out)
    shm.synchronize()
    dt = time.perf_counter() - t0
    log(f"bench: ct-ct QK^T colpacking, {rows} of 128 outputs: {dt * 1e3:.1f} ms")
    return {"workload": "ct-ct column-packing product, col_X=64, N=2^16, L=24, num_batch=256",
            "outputs_timed": rows, "ms": round(dt * 1e3, 2), "ms_per_output": round(dt * 1e3 / rows, 2),
            "qk_128_outputs_s_extrapolated": round(dt / rows * 128, 2)}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--variant", choices=["int8", "u64"], default="int8")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-plain", action="store_true", help="skip the one-off per-ciphertext form timing")
    ap.add_argument("--no-bootstrap", action="store_true", help="skip the bootstrap timing")
    ap.add_argument("--no-full-width", action="store_true", help="skip the 768x768 matmul timing")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run one device-to-device copy of a 24 MiB ciphertext (profiler byte-counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on exactly one GPU; --gpus {args.gpus} is not supported "
              "(the matmul shards by coefficient columns, see DESIGN.md)", file=sys.stderr)
        return 2

    import shmoai_amd as shm
    variant = shm.MATMUL_INT8 if args.variant == "int8" else shm.MATMUL_U64
    ctx = shm.Context(LOG_N, BITS, 1)
    rng = np.random.default_rng(5)
    log("bench: generating 192 input ciphertexts (N=2^16, L=24) and uploading to HBM")
    xs, hs = make_inputs(ctx, shm, rng)
    W = np.random.default_rng(6).normal(0.0, 0.02, (ROW_W, COL_W))
    out = [shm.Ciphertext(ctx, 2, 2) for _ in range(COL_W)]

    for w in range(args.warmup):
        shm.ct_pt_matrix_mul_wo_pre(hs, W, ROW_W, COL_W, ROW_W, ctx, variant=variant, out=out)
        log(f"bench: warmup step {w + 1}/{args.warmup} done")
    if args.copy_calibration:
        hs[0].copy()
        log(f"bench: calibration copy of {hs[0].words() * 8} bytes done")
    kernel_ms, tile_ms = [], []
    t0 = time.perf_counter()
    for s in range(args.steps):
        shm.ct_pt_matrix_mul_wo_pre(hs, W, ROW_W, COL_W, ROW_W, ctx, variant=variant, out=out)
        kernel_ms.append(shm.last_matmul_kernel_ms())
        tile_ms.append(shm.last_matmul_rescale_tile_ms())
        assert shm.last_matmul_variant() == variant, "matmul fell back to another variant"
    dt = time.perf_counter() - t0  # every call synchronises its stream before returning
    ms_per_step = dt / args.steps * 1e3
    k_ms = float(np.mean(kernel_ms))
    t_ms = float(np.mean(tile_ms))
    log(f"bench: {args.steps} steps, {ms_per_step:.3f} ms/step, contraction kernel {k_ms:.3f} ms, "
        f"rescale tile pass {t_ms:.3f} ms")

    N = ctx.N
    bytes_alg = ROW_W * 2 * L * N * 8 + COL_W * 2 * L * N * 8        # read X limbs + write accumulators
    macs = ROW_W * COL_W * 2 * L * N                                  # modular MACs
    hbm_gbs = bytes_alg / (k_ms * 1e-3) / 1e9
    digit_ops = 2 * 35 * macs                                         # 7 x 5 int8 digit products
    tops = digit_ops / (k_ms * 1e-3) / 1e12
    contraction = {"bound": "mfma" if variant == shm.MATMUL_INT8 else "hbm",
                   "achieved": round(tops if variant == shm.MATMUL_INT8 else hbm_gbs, 2),
                   "peak": INT8_PEAK_TOPS if variant == shm.MATMUL_INT8 else HBM_PEAK_GBS,
                   "unit": "TOP/s" if variant == shm.MATMUL_INT8 else "GB/s",
                   "traffic": measured_traffic("k_contract_int8") if variant == shm.MATMUL_INT8 else None,
                   "kernel": "k_contract_int8" if variant == shm.MATMUL_INT8 else "k_contract_u64",
                   "kernel_ms": round(k_ms, 4),
                   "hbm_achieved_gbs": round(hbm_gbs, 1), "hbm_frac": round(hbm_gbs / HBM_PEAK_GBS, 4),
                   "algorithmic_bytes_per_launch": bytes_alg, "modular_macs_per_launch": macs}
    contraction["frac"] = round(contraction["achieved"] / contraction["peak"], 4)
    # the fused rescale's tile pass: per output limb it reads the delta limb and the
    # accumulator limb and writes the output limb (24 N bytes), 2 x 23 limbs per output
    tile_bytes = 24 * N * COL_W * 2 * (L - 1)
    tile_gbs = tile_bytes / (t_ms * 1e-3) / 1e9
    tile = {"bound": "hbm", "achieved": round(tile_gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(tile_gbs / HBM_PEAK_GBS, 4),
            "traffic": measured_traffic("k_ntt_blocks<12,true,2,true>"),
            "kernel": "k_ntt_blocks<12,true,2,true> (fused rescale NTT, tile pass)", "kernel_ms": round(t_ms, 4),
            "algorithmic_bytes_per_launch": tile_bytes,
            "note": "VALU-bound on 64-bit modular butterflies; HBM is the roofline it is quoted against"}
    # the roofline object describes the step's dominant kernel; the other one rides along
    roofline, other = (tile, contraction) if t_ms > k_ms else (contraction, tile)

    extra = {}
    if not args.no_plain:
        t1 = time.perf_counter()
        shm.ct_pt_matrix_mul_wo_pre(hs, W, ROW_W, COL_W, ROW_W, ctx, variant=shm.MATMUL_PER_CT, out=out)
        plain_ms = (time.perf_counter() - t1) * 1e3
        extra = {"plain_per_ct_ms_per_step": round(plain_ms, 2),
                 "batched_speedup_vs_plain": round(plain_ms / ms_per_step, 2)}
        log(f"bench: plain per-ciphertext form {plain_ms:.1f} ms")

    full = None
    if not args.no_full_width and variant == shm.MATMUL_INT8:
        full = full_width_bench(shm, ctx, hs)
        full["masked"] = masked_bench(shm, ctx, hs)
        full["ctct_qk"] = ctct_bench(shm, ctx, hs)

    ntt = ntt_bench(shm, ctx)
    boot = None
    if not args.no_bootstrap:
        boot = bootstrap_bench(shm)

    cpu = None
    if not args.no_cpu_baseline:
        threads = min(16, os.cpu_count() or 1)
        log(f"bench: CPU baseline (oracle) on {threads} threads")
        cpu = cpu_baseline(xs, W, threads)

    line = {
        "metric": METRIC,
        "value": ms_per_step * 1e-3 / INPUTS,
        "unit": "s/input",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64 (int8 MFMA digits, exact)" if variant == shm.MATMUL_INT8 else "u64",
        "data": "synthetic (uniform random limbs, W ~ N(0,0.02))",
        "config": {"workload": "quarter-width ct-pt matmul 192x192 (MOAI column packing), N=2^16, "
                               "L=24 x 50-bit primes, 256 inputs x 128 tokens per ciphertext, incl. rescale",
                   "variant": args.variant, "inputs_per_step": INPUTS},
        "roofline": roofline,
        "roofline_other_kernel": other,
        "cpu_baseline": cpu,
        "full_width_768": full,
        "ntt": ntt,
        "bootstrap_768": boot,
        **extra,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
