"""Benchmark of the batched gate-bootstrapping workload on one MI355X.

Workload (one step): 4096 independent gate bootstraps on seeded random bit
pairs, 2048 NAND + 2048 XOR, n=636, N=1024, BSK 2^7 x 3, KS 2^2 x 8, FP64
negacyclic FFT, one persistent blind-rotation launch per gate batch followed by
one batched key switch.  Keys are generated on the GPU from seed 1; inputs are
resident in HBM before the timed region.

Prints one JSON line (see the README of the driver contract).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "gpu-circuitbootstrapping-aestransciphering_amd")
sys.path.insert(0, PKG)

METRIC = "AES-128 blocks/sec + bootstraps/sec at 1 MI355X, batched FP32-split vs plain FP64"
BATCH = 4096
N_LWE, N_RING, BK_L, M_FFT = 636, 1024, 3, 512
# algorithmic FLOPs of one external product: (k+1)(l+1) FFTs of M points (5 M log2 M each)
# plus the Fourier-domain MAC 8 (k+1)^2 l M, k = 1
F_EP = 2 * (BK_L + 1) * 5 * M_FFT * 9 + 8 * 4 * BK_L * M_FFT
F_PBS = N_LWE * F_EP  # 148.49 MFLOP per blind rotation
BSK_BYTES = N_LWE * 2 * BK_L * 2 * M_FFT * 16  # Fourier BSK streamed once per pass (62.5 MB)
PEAK_FP64_TFLOPS = 78.6
PEAK_LDS_TBS = 78.6  # 256 CUs x 128 B/clk x 2.4 GHz
PEAK_INT8_TOPS = 5000.0  # dense int8 matrix-core peak
# circuit-bootstrapping level: N = 2048 (M = 1024), PBS gadget 2^9 x 4 (8 digit rows)
N_CB, M_CB, CB_ROWS = 2048, 1024, 8
F_EP_CB = {1: (CB_ROWS + 2) * 5 * M_CB * 10 + 8 * 2 * CB_ROWS * M_CB,      # one product: 643,072
           2: (CB_ROWS + 4) * 5 * M_CB * 10 + 8 * 2 * 2 * CB_ROWS * M_CB}  # two limbs: 876,544


def cpu_baseline(sample=24, threads=None):
    """The oracle's gate bootstrap on the host cores (a bounded sample): one thread, then
    `threads` threads (the box's CPU share, at most 16) each running `sample` bootstraps
    (the C call releases the GIL, so the threads run on separate cores)."""
    from concurrent.futures import ThreadPoolExecutor
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    import oracle as O
    if threads is None:
        threads = max(1, min(16, os.cpu_count() or 1))
    keys = O.GateKeys(1)
    rng = np.random.default_rng(5)
    a = rng.integers(0, 2, sample)
    b = rng.integers(0, 2, sample)
    c1 = [keys.encrypt(int(a[k]), 2 * k) for k in range(sample)]
    c2 = [keys.encrypt(int(b[k]), 2 * k + 1) for k in range(sample)]

    def run():
        for k in range(sample):
            keys.bootstrap(O.NAND if k % 2 == 0 else O.XOR, c1[k], c2[k])

    t0 = time.perf_counter()
    run()
    one = sample / (time.perf_counter() - t0)
    t0 = time.perf_counter()
    with ThreadPoolExecutor(threads) as pool:
        for f in [pool.submit(run) for _ in range(threads)]:
            f.result()
    many = threads * sample / (time.perf_counter() - t0)
    return {"value": round(many, 3), "unit": "bootstraps/s", "cores": threads, "kind": "port",
            "single_thread": round(one, 3),
            "sample": f"{sample} gate bootstraps (NAND/XOR alternating) per thread on {threads} threads, then "
                      f"{sample} on one thread; same parameter set, oracle C restatement, FP64 radix-2 FFT"}


def one_product_error(batch=256):
    """Max coefficient error of the PBS's one-product form against the exact
    integer product (both on the GPU; the exact one is bit-exact by the tests):
    digit polynomials in [-256, 256) times uniform torus-64 polynomials, N = 2048."""
    import numpy as np
    import torch
    import tfhe_amd
    rng = np.random.default_rng(5)
    x = torch.from_numpy(rng.integers(-256, 256, (batch, 2048)).astype(np.int32)).to("cuda")
    y = torch.from_numpy(rng.integers(0, 2 ** 64, (batch, 2048), dtype=np.uint64).view(np.int64)).to("cuda")
    exact = torch.empty_like(y)
    approx = torch.empty_like(y)
    tfhe_amd.polymul2048_batch(x, y, exact)
    tfhe_amd.polymul2048_batch(x, y, approx, approx=True)
    torch.cuda.synchronize()
    d = (approx.cpu().numpy().view(np.uint64) - exact.cpu().numpy().view(np.uint64)).view(np.int64)
    worst = int(np.abs(d).max())
    return {"products": batch, "max_abs_error_log2_of_torus": round(float(np.log2(max(worst, 1))) - 64, 2),
            "rms_error_log2_of_torus": round(float(np.log2(np.sqrt(np.mean(d.astype(np.float64) ** 2)))) - 64, 2)}


def aes_kernel_rooflines(ctx, pbs_approx, nb, pfks_linear=0):
    """PBS and PFKS of one AES round's shape timed on their own (HIP events): the PBS
    against the FP64 peak by its algorithmic FLOPs; the lookup PFKS against the LDS
    bandwidth by the bytes its selects read (4 per output word per digit), the linear
    PFKS against the dense int8 matrix-core peak by its products (inputs x 16392
    digits x 2N columns x 5 key bytes per function, two operations each)."""
    import numpy as np
    import torch
    rng = np.random.default_rng(13)
    lwe = torch.from_numpy(rng.integers(0, 2 ** 32, (nb, N_LWE + 1), dtype=np.uint64).astype(np.uint32)
                           .view(np.int32)).to("cuda")
    pbs_out = torch.empty((nb, 2, N_CB + 1), dtype=torch.int64, device="cuda")
    rows = torch.empty((2 * nb, 2 * N_CB), dtype=torch.int32, device="cuda")

    def timed(fn):
        fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        fn()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1)

    t_pbs = timed(lambda: ctx.pbs_batch(lwe, pbs_out))
    t_pfks = timed(lambda: [ctx.pfks_batch(u, pbs_out, 2 * nb, 2 * N_CB, rows) for u in (0, 1)])
    limbs = 1 if pbs_approx else 2
    tflops = nb * N_LWE * F_EP_CB[limbs] / (t_pbs * 1e-3) / 1e12
    out = {"inputs": nb,
           "pbs": {"ms": round(t_pbs, 3), "bound": "fp64", "achieved": round(tflops, 2), "peak": PEAK_FP64_TFLOPS,
                   "unit": "TFLOP/s", "frac": round(tflops / PEAK_FP64_TFLOPS, 4)}}
    if pfks_linear:
        ops = 2 * 2 * (2 * nb) * (N_CB + 1) * 8 * (2 * N_CB) * 5  # 2 ops x functions x inputs x digits x cols x bytes
        tops = ops / (t_pfks * 1e-3) / 1e12
        out["pfks_both_functions"] = {"ms": round(t_pfks, 3), "form": "linear key, int8 GEMM", "bound": "mfma",
                                      "achieved": round(tops, 1), "peak": PEAK_INT8_TOPS, "unit": "TOP/s",
                                      "frac": round(tops / PEAK_INT8_TOPS, 4)}
    else:
        lds_bytes = 2 * (2 * nb) * (N_CB + 1) * 8 * (2 * N_CB) * 4  # functions x inputs x digits x words x 4 B
        lds_tbs = lds_bytes / (t_pfks * 1e-3) / 1e12
        out["pfks_both_functions"] = {"ms": round(t_pfks, 3), "form": "lookup key", "bound": "lds",
                                      "achieved": round(lds_tbs, 2), "peak": PEAK_LDS_TBS, "unit": "TB/s",
                                      "frac": round(lds_tbs / PEAK_LDS_TBS, 4)}
    return out


def aes_ctr_bench(batches, reps=1, pbs_approx=0, pbs_multibit=1, plain_blocks=0, tune=True, pfks_linear=0):
    """AES-128 CTR transciphering: the keystream (circuit bootstrap + S-box LUT per
    round, linear layers as LWE additions) plus the public AES-CTR ciphertext bits, so
    the outputs encrypt the plaintext bits; every block decrypted and compared with the
    seeded random plaintext.
    pbs_approx=1 runs the one-product FP64 PBS instead of the exact two-limb one;
    pbs_multibit=2 blind-rotates two key bits per step (one-product form only)."""
    import numpy as np
    import torch
    import tfhe_amd
    maxb = max(batches)
    ctx = tfhe_amd.CbsContext(max_blocks=maxb, params=tfhe_amd.cbs_params(pbs_approx, pbs_multibit, pfks_linear))
    sk = tfhe_amd.lwe_keygen(N_LWE, 3)
    t0 = time.perf_counter()
    ctx.keygen(sk, 3)
    keygen_s = time.perf_counter() - t0
    rng = np.random.default_rng(11)
    key = bytes(rng.integers(0, 256, 16, dtype=np.uint8))
    ctx.set_round_keys(tfhe_amd.aes_encrypt_key_bits(sk, 5, key))
    nonce = bytes(rng.integers(0, 256, 8, dtype=np.uint8))
    out = {}
    d_ks = torch.empty((maxb, 128, N_LWE + 1), dtype=torch.int32, device="cuda")
    s64 = sk.astype(np.uint64)
    def workload(blocks):  # counters, seeded random plaintexts and their public AES-CTR ciphertexts
        ctrs = [nonce + int(b).to_bytes(8, "big") for b in range(blocks)]
        plain = [bytes(rng.integers(0, 256, 16, dtype=np.uint8)) for _ in range(blocks)]
        cipher = [bytes(p ^ k for p, k in zip(pt, tfhe_amd.aes128_encrypt_plain(key, c)))
                  for pt, c in zip(plain, ctrs)]
        return ctrs, plain, cipher

    warm = workload(min(batches))
    ctx.aes_ctr_transcipher(warm[0], warm[2], d_ks)  # warm-up (module load, first launches)
    torch.cuda.synchronize()
    for blocks in batches:
        ctrs, plain, cipher = workload(blocks)
        tuned = None
        if tune:  # launch shapes for this batch, timed before the timed region
            t0 = time.perf_counter()
            tuned = ctx.autotune(blocks)
            tuned["tune_s"] = round(time.perf_counter() - t0, 3)
            ctx.aes_ctr_transcipher(ctrs, cipher, d_ks)  # first launch with the tuned shapes
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            ctx.aes_ctr_transcipher(ctrs, cipher, d_ks)
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / reps
        got = d_ks[:blocks].cpu().numpy().view(np.uint32)
        ph = (got[:, :, -1].astype(np.uint64) - (got[:, :, :-1].astype(np.uint64) @ s64)) % 2 ** 32
        bits = ((ph + (1 << 30)) >> 31) & 1
        bad = 0
        for b in range(blocks):
            v = bytes(int(sum(int(bits[b, 8 * k + j]) << j for j in range(8))) for k in range(16))
            bad += v != plain[b]
        out[str(blocks)] = {"blocks_per_s": round(blocks / dt, 3), "s_per_launch": round(dt, 4),
                            "circuit_bootstraps_per_s": round(blocks * 1280 / dt, 1), "blocks_wrong": int(bad)}
        if tuned is not None:
            out[str(blocks)]["tuned_shapes"] = tuned
        print(f"aes128 ctr transcipher ({'one-product' if pbs_approx else 'exact'} PBS, {pbs_multibit} key bit(s) "
              f"per step): "
              f"{blocks} blocks in {dt:.3f} s, "
              f"{int(bad)} wrong", file=sys.stderr, flush=True)
        if blocks <= 32:  # the same transcipher with the keystream replayed from its captured hipGraph
            eager = d_ks[:blocks].clone()
            ctx.aes_ctr_transcipher(ctrs, cipher, d_ks, mode="graph")  # capture + first replay
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            ctx.aes_ctr_transcipher(ctrs, cipher, d_ks, mode="graph")
            torch.cuda.synchronize()
            gdt = time.perf_counter() - t0
            out[str(blocks)]["graph_s_per_launch"] = round(gdt, 4)
            out[str(blocks)]["graph_equals_eager"] = bool(torch.equal(d_ks[:blocks], eager))
    if pbs_multibit in (1, 3):  # the round's two big kernels on their own, at 17 blocks' shape (or less)
        out["kernels"] = aes_kernel_rooflines(ctx, pbs_approx, min(17, maxb) * 128, pfks_linear)
    if plain_blocks:  # the per-bootstrap form: every bootstrap, LUT and key switch a launch of one
        ctrs, plain, cipher = workload(plain_blocks)
        ctx.aes_ctr_transcipher(ctrs, cipher, d_ks)
        torch.cuda.synchronize()
        eager = d_ks[:plain_blocks].clone()
        t0 = time.perf_counter()
        ctx.aes_ctr_transcipher(ctrs, cipher, d_ks, mode="plain")
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        out["plain_per_bootstrap"] = {"blocks": plain_blocks, "blocks_per_s": round(plain_blocks / dt, 4),
                                      "s_per_launch": round(dt, 3),
                                      "equals_batched": bool(torch.equal(d_ks[:plain_blocks], eager))}
        print(f"aes128 ctr plain per-bootstrap form: {plain_blocks} block(s) in {dt:.2f} s", file=sys.stderr,
              flush=True)
    ctx.close()
    pbs = ("one FP64 product per digit polynomial, rounding error ~2^-33" if pbs_approx
           else "2-limb exact") + (", two key bits per blind-rotation step" if pbs_multibit == 2 else
                                   ", two key bits per step unless the autotuner measured one bit faster" if pbs_multibit == 3 else "")
    pf = ("linear 40-bit key as an int8 GEMM on the matrix cores" if pfks_linear
          else "lookup key, LDS selects")
    return {"form": f"fp64 FFT: PBS on the 64-bit torus ({pbs}), PFKS ({pf}) and CMux LUT on the 32-bit torus "
                    "(1-limb exact)", "keygen_s": round(keygen_s, 2), "batches": out}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--plain-steps", type=int, default=3, help="plain per-bootstrap form samples")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run three 512 MiB device copies (the HBM-counter calibration kernel)")
    ap.add_argument("--campaign", type=int, default=1_000_000,
                    help="gate bootstraps in the decryption-failure campaign (0 to skip)")
    ap.add_argument("--valu-rates", action="store_true",
                    help="also measure the FP64 and (unpacked) FP32 FMA throughput of the library's probe kernel")
    ap.add_argument("--aes-blocks", default="1,17,256,1024",
                    help="comma-separated AES-128 CTR batch sizes to time ('' to skip)")
    ap.add_argument("--aes-plain-blocks", type=int, default=1,
                    help="AES blocks timed in the plain per-bootstrap form (exact PBS; 0 to skip)")
    ap.add_argument("--aes-pbs", default="exact,exact-lin,approx-auto-lin",
                    help="comma-separated PBS forms of the AES runs: exact (two-limb), approx (one FP64 "
                         "product), approx-mb2 (one product, two key bits per step); exact-lin / approx-lin "
                         "with the linear-key PFKS on the matrix cores; approx-mb2-lin two key bits per step; "
                         "approx-auto-lin both key forms, two bits per step unless tuned to one")
    ap.add_argument("--no-aes-tune", action="store_true",
                    help="keep the kernels' model launch shapes (no tfhe_cbs_autotune before each AES batch)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              f"(bootstraps are independent: a multi-GPU run would give each GPU a key replica and a slice "
              f"of the batch)")
        sys.exit(2)

    import numpy as np
    import torch
    import tfhe_amd

    torch.cuda.init()
    stream = torch.cuda.current_stream()
    sptr = stream.cuda_stream

    ctx = tfhe_amd.GateContext(max_batch=BATCH)
    sk = tfhe_amd.lwe_keygen(N_LWE, 1)
    ctx.bt_keygen(sk, 1)
    rng = np.random.default_rng(1)
    a = rng.integers(0, 2, BATCH)
    b = rng.integers(0, 2, BATCH)
    c1 = np.stack([tfhe_amd.lwe_encrypt(sk, 2, 2 * k, int(a[k])) for k in range(BATCH)])
    c2 = np.stack([tfhe_amd.lwe_encrypt(sk, 2, 2 * k + 1, int(b[k])) for k in range(BATCH)])
    d1 = torch.from_numpy(c1.view(np.int32).copy()).cuda()
    d2 = torch.from_numpy(c2.view(np.int32).copy()).cuda()
    d_out = torch.empty_like(d1)
    d_lweN = torch.empty((BATCH, N_RING + 1), dtype=torch.int32, device="cuda")
    half = BATCH // 2
    L = tfhe_amd.lib()

    def ptr(t):
        return tfhe_amd._ptr(t)

    def step(events=None):
        # NAND on the first half, XOR on the second half
        for g, lo in ((tfhe_amd.NAND, 0), (tfhe_amd.XOR, half)):
            sl = slice(lo, lo + half)
            if events is not None:
                events.append(torch.cuda.Event(enable_timing=True))
                events[-1].record(stream)
            tfhe_amd.check(L.tfhe_gate_blind_rotate_batch(ctx.h, g, ptr(d1[sl]), ptr(d2[sl]), half,
                                                          ptr(d_lweN[sl]), sptr))
            if events is not None:
                events.append(torch.cuda.Event(enable_timing=True))
                events[-1].record(stream)
            tfhe_amd.check(L.tfhe_gate_keyswitch_batch(ctx.h, ptr(d_lweN[sl]), half, ptr(d_out[sl]), sptr))
            if events is not None:
                events.append(torch.cuda.Event(enable_timing=True))
                events[-1].record(stream)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    events = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step(events)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0

    # correctness of the timed output (last step)
    got = d_out.cpu().numpy().view(np.uint32)
    want = np.concatenate([1 - (a[:half] & b[:half]), a[half:] ^ b[half:]])
    s64 = sk.astype(np.uint64)
    ph = (got[:, -1].astype(np.uint64) - (got[:, :-1].astype(np.uint64) @ s64)) % 2 ** 32
    dec = ((ph + (1 << 29)) >> 30) & 3
    failures = int(np.count_nonzero(dec != want))

    # decryption-failure campaign: batches chained through their own outputs (inputs of
    # batch r+1 are the outputs of batch r and a rotation of them), every output decrypted
    campaign = None
    if args.campaign > 0:
        cur, cur_bits = d_out.clone(), want.copy()
        d_new = torch.empty_like(d_out)
        done = fails = 0
        worst = 0
        while done < args.campaign:
            d1.copy_(cur)
            d2.copy_(torch.roll(cur, 1, 0))
            bits_a, bits_b = cur_bits, np.roll(cur_bits, 1)
            for g, lo in ((tfhe_amd.NAND, 0), (tfhe_amd.XOR, half)):
                sl = slice(lo, lo + half)
                tfhe_amd.check(L.tfhe_gate_blind_rotate_batch(ctx.h, g, ptr(d1[sl]), ptr(d2[sl]), half,
                                                              ptr(d_lweN[sl]), sptr))
                tfhe_amd.check(L.tfhe_gate_keyswitch_batch(ctx.h, ptr(d_lweN[sl]), half, ptr(d_new[sl]), sptr))
            exp = np.concatenate([1 - (bits_a[:half] & bits_b[:half]), bits_a[half:] ^ bits_b[half:]])
            g_np = d_new.cpu().numpy().view(np.uint32)
            ph_c = (g_np[:, -1].astype(np.uint64) - (g_np[:, :-1].astype(np.uint64) @ s64)) % 2 ** 32
            dec_c = ((ph_c + (1 << 29)) >> 30) & 3
            fails += int(np.count_nonzero(dec_c != exp))
            err = ((ph_c - exp.astype(np.uint64) * (1 << 30) + 2 ** 31) % 2 ** 32).astype(np.int64) - 2 ** 31
            worst = max(worst, int(np.abs(err).max()))
            cur, d_new = d_new, cur
            cur_bits = exp
            done += BATCH
        campaign = {"bootstraps": done, "failures": fails,
                    "max_abs_phase_error_log2": round(float(np.log2(max(worst, 1))), 2), "margin_log2": 29.0,
                    "inputs": "chained: each batch's inputs are the previous outputs and a rotation of them"}

    br_ms, ks_ms = [], []
    for k in range(0, len(events), 3):
        br_ms.append(events[k].elapsed_time(events[k + 1]))
        ks_ms.append(events[k + 1].elapsed_time(events[k + 2]))
    br_avg = sum(br_ms) / len(br_ms)
    ks_avg = sum(ks_ms) / len(ks_ms)
    achieved_tflops = F_PBS * half / (br_avg * 1e-3) / 1e12

    # plain per-bootstrap form (reference launch structure), a few single gates
    plain_rate = None
    if args.plain_steps > 0:
        one_out = torch.empty((1, N_LWE + 1), dtype=torch.int32, device="cuda")
        ctx.eval_plain(tfhe_amd.NAND, d1[0:1], d2[0:1], one_out, stream=sptr)
        torch.cuda.synchronize()
        tp = time.perf_counter()
        for k in range(args.plain_steps):
            ctx.eval_plain(tfhe_amd.NAND, d1[k:k + 1], d2[k:k + 1], one_out, stream=sptr)
        torch.cuda.synchronize()
        plain_rate = args.plain_steps / (time.perf_counter() - tp)

    if args.copy_calibration:
        src = torch.empty(512 << 20, dtype=torch.uint8, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()

    traffic = None
    prof = os.path.join(ROOT, "profiles", "gate_blind_rotate_4096_nand_xor_traffic.json")
    if os.path.exists(prof):
        traffic = json.load(open(prof)).get("hbm_bytes_per_launch")

    ms_per_step = elapsed * 1e3 / args.steps
    line = {
        "metric": METRIC,
        "value": round(BATCH * args.steps / elapsed, 1),
        "unit": "bootstraps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "batched gate bootstraps: 4096 = 2048 NAND + 2048 XOR per step",
                   "n": N_LWE, "N": N_RING, "bsk_gadget": "2^7 x 3", "ks_gadget": "2^2 x 8",
                   "torus_bits": 32, "fft": "fp64 negacyclic, wavefront-resident", "batch": BATCH},
        "roofline": {"bound": "valu", "kernel": "k_gate_blind_rotate",
                     "achieved": round(achieved_tflops, 3), "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(achieved_tflops / PEAK_FP64_TFLOPS, 4), "traffic": traffic,
                     "algorithmic_flop_per_launch": F_PBS * half,
                     "avg_launch_ms": round(br_avg, 4)},
        "breakdown_ms": {"blind_rotate_per_2048": round(br_avg, 4), "keyswitch_per_2048": round(ks_avg, 4)},
        "plain_fp64_per_bootstrap_rate": None if plain_rate is None else round(plain_rate, 2),
        "decrypt_failures": failures,
    }
    if campaign is not None:
        line["failure_campaign"] = campaign
    ctx.close()
    if args.valu_rates:
        line["valu_fma_tflops"] = {"fp64": round(tfhe_amd.probe_fma_tflops(1), 2),
                                   "fp32_unpacked": round(tfhe_amd.probe_fma_tflops(0), 2)}
    if args.aes_blocks:
        sizes = [int(b) for b in args.aes_blocks.split(",") if b]
        forms = {"exact": ("aes128_ctr", 0, 1, 0), "approx": ("aes128_ctr_one_product_pbs", 1, 1, 0),
                 "approx-mb2": ("aes128_ctr_one_product_pbs_2bit", 1, 2, 0),
                 "exact-lin": ("aes128_ctr_linear_pfks", 0, 1, 1),
                 "approx-lin": ("aes128_ctr_one_product_pbs_linear_pfks", 1, 1, 1),
                 "approx-mb2-lin": ("aes128_ctr_one_product_pbs_2bit_linear_pfks", 1, 2, 1),
                 "approx-auto-lin": ("aes128_ctr_one_product_pbs_auto_bits_linear_pfks", 1, 3, 1)}
        for f in [f for f in args.aes_pbs.split(",") if f]:
            if f not in forms:
                sys.exit(f"--aes-pbs: unknown form {f!r} ({', '.join(forms)})")
            key, approx, mb, lin = forms[f]
            line[key] = aes_ctr_bench(sizes, pbs_approx=approx, pbs_multibit=mb,
                                      plain_blocks=args.aes_plain_blocks if f == "exact" else 0,
                                      tune=not args.no_aes_tune, pfks_linear=lin)
            if approx:
                line[key]["one_product_vs_exact_product"] = one_product_error()
        # batched blocks/s over the plain FP64 per-bootstrap form's
        plain = line.get("aes128_ctr", {}).get("batches", {}).get("plain_per_bootstrap")
        if plain:
            best = max((v["blocks_per_s"] for k in forms.values() if k[0] in line
                        for b, v in line[k[0]]["batches"].items() if b.isdigit()), default=None)
            line["aes128_batched_over_plain"] = None if best is None else round(best / plain["blocks_per_s"], 1)
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
