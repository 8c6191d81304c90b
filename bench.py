"""Benchmark of AES-128 CTR transciphering on one MI355X, with batched gate
bootstraps as a secondary line.

Headline workload (one step): homomorphic AES-128 CTR transciphering of 256
blocks: the keystream of 256 counter blocks under an encrypted, seeded random
128-bit key (10 rounds; SubBytes by circuit bootstrap of every state bit, an
8-bit CMux vertical-packing S-box LUT and a key switch; ShiftRows, MixColumns and
AddRoundKey as LWE additions), plus the public AES-CTR ciphertext bits, so the
outputs are LWE encryptions of the plaintext bits.  Level-0 LWE n=636 (sigma
2^-15), circuit bootstrap N=2048: one-product FP64 PBS (2^9 x 4) with two key bits
per blind-rotation step, linear-key PFKS (2^4 x 8) as an int8 GEMM on the matrix
cores, CB 2^8 x 2, KS 2^2 x 8.  Keys are generated on the GPU from seed 3; the
counters and the public ciphertext (4 KB for 256 blocks) are the only host inputs
of a step.  Every block of the last timed step is decrypted and compared with the
seeded random plaintext.

Secondary (key "gate_bootstraps"): 4096 independent NAND/XOR gate bootstraps per
step (n=636, N=1024, BSK 2^7 x 3, KS 2^2 x 8, FP64 FFT), with a decryption-failure
campaign over 10^6 chained bootstraps.  Side line "gate_functions": the gate context's
programmable bootstrap, EvalFunc and EvalSign over Z_2^32 on 1024 ciphertexts.  Side line
"cbs_lut_256_bits": 256 LWE bits through circuit bootstrapping and 32 random-table byte LUTs.

Prints one JSON line.
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
# the PBS's algorithmic FLOPs per circuit bootstrap: one FFT external product per key bit at the
# one-product count, for every form.  The two-key-bit form executes less by the same convention
# (10 FFTs + 3 pattern MACs + 196,608 FLOP of pattern factors per key-bit pair = 550,912 per key
# bit, 0.857 of this count), so its rate by executed arithmetic is 0.857 x the one reported.
F_PBS_CB = N_LWE * F_EP_CB[1]  # 409.0 MFLOP
F_MB2_PER_BIT = (CB_ROWS + 2) * 5 * M_CB * 10 + 3 * 8 * 2 * CB_ROWS * M_CB + 196608  # 1,101,824 per pair
# HomTrace + SchemeSwitch per circuit bootstrap: 2 CB levels x (11 trace rounds + 1 scheme switch)
# key switches, each 5 forward + 4 inverse FFT1024s and a 5-digit, 2-output, 2-limb MAC
F_HTSS_KS = 9 * 5 * M_CB * 10 + 8 * 5 * 2 * 2 * M_CB  # 624,640
F_HTSS_CB = 2 * 12 * F_HTSS_KS  # 14.99 MFLOP


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


def split32_product_bench(batch=4096):
    """The split-precision FP32 product forms (tfhe_polymul2048_split32_batch) beside the
    FP64 ones, N = 2048: digit polynomials in [-256, 256] times uniform polynomials,
    `batch` products per launch, each timed with HIP events (mean of 3 launches after a
    warm-up).  Errors are against the one-FP64-product form, which the tests show to be
    the exact product bit for bit."""
    import numpy as np
    import torch
    import tfhe_amd
    rng = np.random.default_rng(7)
    x = torch.from_numpy(rng.integers(-256, 257, (batch, 2048)).astype(np.int32)).to("cuda")
    y32 = torch.from_numpy(rng.integers(0, 2 ** 32, (batch, 2048), dtype=np.uint64).astype(np.uint32)
                           .view(np.int32)).to("cuda")
    y64 = torch.from_numpy(rng.integers(0, 2 ** 64, (batch, 2048), dtype=np.uint64).view(np.int64)).to("cuda")
    o32 = {f: torch.empty_like(y32) for f in tfhe_amd.SPLIT32_FORMS}
    o64 = torch.empty_like(y64)
    stream = torch.cuda.current_stream()

    def timed(fn):
        fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(3):
            fn()
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / 3

    out = {"products_per_launch": batch}
    for f in tfhe_amd.SPLIT32_FORMS:
        ms = timed(lambda: tfhe_amd.polymul2048_split32_batch(x, y32, o32[f], f))
        out[f] = {"ms": round(ms, 4), "us_per_product": round(1e3 * ms / batch, 4)}
    ref = o32["fp64"].cpu().numpy().view(np.uint32).astype(np.int64)
    for f in ("fp32_3x11", "fp32_exact_top"):
        d = (o32[f].cpu().numpy().view(np.uint32).astype(np.int64) - ref) % 2 ** 32
        d = np.where(d >= 2 ** 31, d - 2 ** 32, d)
        worst = int(np.abs(d).max())
        out[f]["max_abs_error_log2_of_torus"] = round(float(np.log2(max(worst, 1))) - 32, 2)
        out[f]["rms_error_log2_of_torus"] = round(float(np.log2(max(np.sqrt(np.mean(d.astype(np.float64) ** 2)),
                                                                    1e-30))) - 32, 2)
        out[f]["coefficients_wrong"] = float(np.mean(d != 0))
    for name, approx in (("fp64_torus64_two_limbs", False), ("fp64_torus64_one_product", True)):
        ms = timed(lambda: tfhe_amd.polymul2048_batch(x, y64, o64, approx=approx))
        out[name] = {"ms": round(ms, 4), "us_per_product": round(1e3 * ms / batch, 4)}
    return out


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
    lwe32 = torch.from_numpy(rng.integers(0, 2 ** 32, (nb, N_CB + 1), dtype=np.uint64).astype(np.uint32)
                             .view(np.int32)).to("cuda")
    ks_out = torch.empty((nb, N_LWE + 1), dtype=torch.int32, device="cuda")
    t_ks = timed(lambda: ctx.keyswitch_batch(lwe32, nb, ks_out))
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
    if ctx.p.ks_linear:
        ops = 2 * nb * N_CB * 8 * (N_LWE + 1) * 4  # 2 ops x inputs x digits x columns x key bytes
        tops = ops / (t_ks * 1e-3) / 1e12
        out["keyswitch"] = {"ms": round(t_ks, 3), "form": "linear key, int8 GEMM", "bound": "mfma",
                            "achieved": round(tops, 1), "peak": PEAK_INT8_TOPS, "unit": "TOP/s",
                            "frac": round(tops / PEAK_INT8_TOPS, 4)}
    else:
        lds_bytes = nb * N_CB * 8 * (N_LWE + 1) * 4  # inputs x digits x columns x 4 B of selected key rows
        lds_tbs = lds_bytes / (t_ks * 1e-3) / 1e12
        out["keyswitch"] = {"ms": round(t_ks, 3), "form": "lookup key", "bound": "lds", "achieved": round(lds_tbs, 2),
                            "peak": PEAK_LDS_TBS, "unit": "TB/s", "frac": round(lds_tbs / PEAK_LDS_TBS, 4)}
    return out


def aes_ctr_bench(batches, reps=1, pbs_approx=0, pbs_multibit=1, plain_blocks=0, tune=True, pfks_linear=0,
                  lwe_to_rgsw=0, ks_linear=0):
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
    ctx = tfhe_amd.CbsContext(max_blocks=maxb, params=tfhe_amd.cbs_params(pbs_approx, pbs_multibit, pfks_linear,
                                                                          lwe_to_rgsw, ks_linear))
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
              f"per step{', HT+SS' if lwe_to_rgsw else ''}): "
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
    if lwe_to_rgsw:  # PBS and HT+SS of one round's circuit bootstraps, at 17 blocks' shape (or less)
        out["kernels"] = htss_kernel_rooflines(ctx, min(17, maxb) * 128)
    elif pbs_multibit in (1, 3):  # the round's two big kernels on their own, at 17 blocks' shape (or less)
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
    pbs, conv, ks = form_text(pbs_approx, pbs_multibit, pfks_linear, lwe_to_rgsw, ks_linear)
    return {"form": f"fp64 FFT: PBS on the 64-bit torus ({pbs}), LWE->RGSW by {conv}, CMux LUT on the 32-bit "
                    f"torus (1-limb exact), N -> n key switch ({ks})", "keygen_s": round(keygen_s, 2), "batches": out}


def aes_cpu_baseline(threads=None, cbs_sample=1):
    """The oracle's AES-128 pieces per state bit on the host cores, scaled to blocks/s:
    one block is 1280 circuit bootstraps (PBS + linear-key PFKS of both levels), 160
    S-box LUTs (8 CMux each) and 1280 N -> n key switches.  A bounded sample is timed
    on one thread (cbs_sample circuit bootstraps, 4 LUTs, 16 key switches), then
    `threads` threads (the box's CPU share, at most 16) each run one circuit bootstrap
    (the C calls release the GIL).  The oracle is the exact two-limb FP64 PBS with one
    key bit per step; its PFKS regenerates every key entry it uses from Philox (the
    8 GB key is never materialised), which dominates its time."""
    import ctypes
    from concurrent.futures import ThreadPoolExecutor
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    import oracle as O
    if threads is None:
        threads = max(1, min(16, os.cpu_count() or 1))
    K = O.CbsKeys(3)
    K.ksk2  # noqa: B018 (materialise the key-switching key before timing)
    cts = [K.encrypt_bit(k & 1, 100 + k) for k in range(max(cbs_sample, threads))]
    # the oracle's PFKS has OpenMP loops: every timing thread runs them on itself alone
    gomp = ctypes.CDLL("libgomp.so.1")
    gomp.omp_set_num_threads(1)
    t0 = time.perf_counter()
    for k in range(cbs_sample):
        rgsw, _ = K.cbs(cts[k], linear=True)
    t_cbs = (time.perf_counter() - t0) / cbs_sample
    rf = K.rgsw_fourier(rgsw)
    table = np.random.default_rng(2).integers(0, 2 ** 32, K.p.N, dtype=np.uint64).astype(np.uint32)
    t0 = time.perf_counter()
    for _ in range(4):
        acc, _ = K.vertical_packing([rf] * 8, table)
    t_lut = (time.perf_counter() - t0) / 4
    lwe = K.extract64(np.zeros(2 * K.p.N, np.uint64), 0)
    t0 = time.perf_counter()
    for _ in range(16):
        K.ks2(lwe)
    t_ks = (time.perf_counter() - t0) / 16
    per_block = 1280 * t_cbs + 160 * t_lut + 1280 * t_ks
    t0 = time.perf_counter()
    def one_cbs(c):
        gomp.omp_set_num_threads(1)
        return K.cbs(c, linear=True)

    with ThreadPoolExecutor(threads) as pool:
        list(pool.map(one_cbs, cts[:threads]))
    t_cbs_many = (time.perf_counter() - t0) / threads  # wall time per circuit bootstrap on `threads` threads
    per_block_many = 1280 * t_cbs_many + (160 * t_lut + 1280 * t_ks) / threads
    return {"value": round(1.0 / per_block_many, 8), "unit": "blocks/s", "cores": threads, "kind": "port",
            "single_thread": round(1.0 / per_block, 8),
            "per_bit_s": {"circuit_bootstrap": round(t_cbs, 4), "sbox_lut_per_byte": round(t_lut, 5),
                          "keyswitch": round(t_ks, 5), "circuit_bootstrap_on_all_threads": round(t_cbs_many, 4)},
            "sample": f"scaled, not a whole block: {cbs_sample} circuit bootstraps, 4 S-box LUTs and 16 key "
                      f"switches on one thread, then one circuit bootstrap per thread on {threads} threads; "
                      f"blocks/s = 1 / (1280 CBS + 160 LUT + 1280 KS); oracle C restatement (exact two-limb "
                      f"FP64 PBS, linear-key PFKS regenerating its key entries from Philox)"}


def gate_bench(args):
    """Batched gate bootstraps (secondary line): 4096 = 2048 NAND + 2048 XOR per step,
    one persistent blind-rotation launch and one key switch per gate half."""
    import numpy as np
    import torch
    import tfhe_amd
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
    ptr = tfhe_amd._ptr

    def step(events=None):
        for g, lo in ((tfhe_amd.NAND, 0), (tfhe_amd.XOR, half)):  # NAND on one half, XOR on the other
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

    steps = args.gate_steps
    for _ in range(2):
        step()
    torch.cuda.synchronize()
    events = []
    t0 = time.perf_counter()
    for _ in range(steps):
        step(events)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    got = d_out.cpu().numpy().view(np.uint32)
    want = np.concatenate([1 - (a[:half] & b[:half]), a[half:] ^ b[half:]])
    s64 = sk.astype(np.uint64)
    ph = (got[:, -1].astype(np.uint64) - (got[:, :-1].astype(np.uint64) @ s64)) % 2 ** 32
    failures = int(np.count_nonzero((((ph + (1 << 29)) >> 30) & 3) != want))

    # decryption-failure campaign: batches chained through their own outputs (inputs of
    # batch r+1 are the outputs of batch r and a rotation of them), every output decrypted
    campaign = None
    if args.campaign > 0:
        cur, cur_bits = d_out.clone(), want.copy()
        d_new = torch.empty_like(d_out)
        done = fails = worst = 0
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
            fails += int(np.count_nonzero((((ph_c + (1 << 29)) >> 30) & 3) != exp))
            err = ((ph_c - exp.astype(np.uint64) * (1 << 30) + 2 ** 31) % 2 ** 32).astype(np.int64) - 2 ** 31
            worst = max(worst, int(np.abs(err).max()))
            cur, d_new = d_new, cur
            cur_bits = exp
            done += BATCH
        campaign = {"bootstraps": done, "failures": fails,
                    "max_abs_phase_error_log2": round(float(np.log2(max(worst, 1))), 2), "margin_log2": 29.0,
                    "inputs": "chained: each batch's inputs are the previous outputs and a rotation of them"}

    br_ms = [events[k].elapsed_time(events[k + 1]) for k in range(0, len(events), 3)]
    ks_ms = [events[k + 1].elapsed_time(events[k + 2]) for k in range(0, len(events), 3)]
    br_avg, ks_avg = sum(br_ms) / len(br_ms), sum(ks_ms) / len(ks_ms)
    achieved = F_PBS * half / (br_avg * 1e-3) / 1e12

    plain_rate = None  # plain per-bootstrap form (the reference's launch structure), a few single gates
    if args.plain_steps > 0:
        one_out = torch.empty((1, N_LWE + 1), dtype=torch.int32, device="cuda")
        ctx.eval_plain(tfhe_amd.NAND, d1[0:1], d2[0:1], one_out, stream=sptr)
        torch.cuda.synchronize()
        tp = time.perf_counter()
        for k in range(args.plain_steps):
            ctx.eval_plain(tfhe_amd.NAND, d1[k:k + 1], d2[k:k + 1], one_out, stream=sptr)
        torch.cuda.synchronize()
        plain_rate = args.plain_steps / (time.perf_counter() - tp)
    ctx.close()
    traffic = None
    prof = os.path.join(ROOT, "profiles", "gate_blind_rotate_4096_nand_xor_traffic.json")
    if os.path.exists(prof):
        traffic = json.load(open(prof)).get("hbm_bytes_per_launch")
    out = {"value": round(BATCH * steps / elapsed, 1), "unit": "bootstraps/s", "steps": steps,
           "ms_per_step": round(elapsed * 1e3 / steps, 3),
           "config": {"workload": "batched gate bootstraps: 4096 = 2048 NAND + 2048 XOR per step",
                      "n": N_LWE, "N": N_RING, "bsk_gadget": "2^7 x 3", "ks_gadget": "2^2 x 8",
                      "torus_bits": 32, "fft": "fp64 negacyclic, wavefront-resident"},
           "roofline": {"bound": "valu", "kernel": "k_gate_blind_rotate", "achieved": round(achieved, 3),
                        "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s", "frac": round(achieved / PEAK_FP64_TFLOPS, 4),
                        "traffic": traffic, "algorithmic_flop_per_launch": F_PBS * half,
                        "avg_launch_ms": round(br_avg, 4)},
           "breakdown_ms": {"blind_rotate_per_2048": round(br_avg, 4), "keyswitch_per_2048": round(ks_avg, 4)},
           "plain_fp64_per_bootstrap_rate": None if plain_rate is None else round(plain_rate, 2),
           "decrypt_failures": failures}
    if campaign is not None:
        out["failure_campaign"] = campaign
    return out


def cbs_lut_bench(reps, nbits=256):
    """Circuit bootstrapping + LUT (the 256-bit workload): 32 seeded random bytes as 256 LWE
    bits -> 256 circuit bootstraps -> RGSW to the Fourier domain -> 32 CMux vertical-packing
    lookups in a seeded random byte table (8 output bits each) -> N -> n key switch, every
    output byte decrypted against the table; in the exact FP64 form (2-limb PBS, lookup
    PFKS) and in the headline form (one-product two-bit PBS, linear PFKS and key switch)."""
    import numpy as np
    import torch
    import tfhe_amd
    rng = np.random.default_rng(17)
    table = rng.integers(0, 256, 256)
    poly = np.zeros(N_CB, np.uint32)
    for x in range(256):
        for j in range(8):
            poly[8 * x + j] = np.uint32(((int(table[x]) >> j) & 1) << 31)
    xs = rng.integers(0, 256, nbits // 8)
    bits = [(int(x) >> i) & 1 for x in xs for i in range(8)]
    sk = tfhe_amd.lwe_keygen(N_LWE, 3)
    cts = np.stack([tfhe_amd.lwe_encrypt(sk, 5, k, b, p=2) for k, b in enumerate(bits)])
    s64 = sk.astype(np.uint64)
    out = {"bits": nbits, "luts": nbits // 8, "reps": reps}
    forms = {"exact_fp64": tfhe_amd.cbs_params(),
             "one_product_two_bit_linear": tfhe_amd.cbs_params(pbs_approx=1, pbs_multibit=3, pfks_linear=1,
                                                               ks_linear=1)}
    for name, params in forms.items():
        ctx = tfhe_amd.CbsContext(max_blocks=(nbits + 127) // 128, params=params)
        ctx.keygen(sk, 3)
        d_cts = torch.from_numpy(cts.view(np.int32).copy()).cuda()
        d_poly = torch.from_numpy(poly.view(np.int32).copy()).cuda()
        d_rgsw = torch.empty((nbits, 4, 2, N_CB), dtype=torch.int32, device="cuda")
        d_f = torch.empty(nbits * 8 * M_CB * 2, dtype=torch.float64, device="cuda")
        d_l = torch.empty((nbits // 8, 8, N_CB + 1), dtype=torch.int32, device="cuda")
        d_o = torch.empty((nbits, N_LWE + 1), dtype=torch.int32, device="cuda")

        def call():
            ctx.cbs_batch(d_cts, d_rgsw)
            ctx.rgsw_to_fourier(d_rgsw, nbits, d_f)
            ctx.lut8(d_f, nbits // 8, d_poly, 0, 8, d_l)
            ctx.keyswitch_batch(d_l, nbits, d_o)

        call()
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(reps):
            call()
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t) / reps
        t = time.perf_counter()
        for _ in range(reps):
            ctx.cbs_batch(d_cts, d_rgsw)
        torch.cuda.synchronize()
        dt_cbs = (time.perf_counter() - t) / reps
        g = d_o.cpu().numpy().view(np.uint32)
        ph = (g[:, -1].astype(np.uint64) - g[:, :-1].astype(np.uint64) @ s64) % 2 ** 32
        got_bits = ((ph + (1 << 30)) >> 31) & 1
        fails = sum(int(sum(int(got_bits[8 * k + j]) << j for j in range(8)) != int(table[x])) for k, x in enumerate(xs))
        ctx.close()
        out[name] = {"ms_per_call": round(dt * 1e3, 3), "circuit_bootstraps_per_s": round(nbits / dt_cbs, 1),
                     "cbs_ms": round(dt_cbs * 1e3, 3), "luts_per_s_end_to_end": round(nbits // 8 / dt, 1),
                     "wrong_bytes": fails}
    return out


def gate_functions_bench(reps, batch=1024):
    """Programmable bootstrap and the large-precision functions on the gate context (side
    line): BootstrapFunc over Z_1024, EvalFunc of an arbitrary table of Z_8 over Z_1024 (two
    bootstraps through Z_2048; every output decrypted) and EvalSign over Z_2^32 (eight floors
    and the sign bootstrap, 17 bootstraps per ciphertext), `batch` ciphertexts per call."""
    import numpy as np
    import torch
    import tfhe_amd
    Q, P = tfhe_amd.FUNC_Q, 8
    ctx = tfhe_amd.GateContext(max_batch=batch)
    sk = tfhe_amd.lwe_keygen(N_LWE, 5)
    ctx.bt_keygen(sk, 5)
    s64 = sk.astype(np.uint64)
    rng = np.random.default_rng(3)

    def enc(vals, mod, seed):
        sh = 32 - int(mod).bit_length() + 1
        cts = np.stack([tfhe_amd.lwe_encrypt(sk, seed, k, 0, p=2) for k in range(len(vals))]).astype(np.uint64)
        cts[:, -1] = (cts[:, -1] + (np.asarray(vals, np.uint64) << np.uint64(sh))) % 2 ** 32
        if sh:
            cts = ((cts + (1 << (sh - 1))) >> sh) % int(mod)
        return torch.from_numpy(cts.astype(np.uint32).view(np.int32)).cuda()

    def timed(fn):
        fn()
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(reps):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t) / reps

    F = [(3 * m + 1) % P for m in range(P)]
    lut = np.array([F[(x * P) // Q] * (Q // P) for x in range(Q)], np.uint32)
    msgs = rng.integers(0, P, batch)
    d_f = enc(msgs * (Q // P), Q, 11)
    d_out = torch.empty_like(d_f)
    t_bf = timed(lambda: ctx.bootstrap_func(d_f, Q, lut, Q, d_out))
    t_ef = timed(lambda: ctx.eval_func(d_f, Q, lut, d_out))
    got = d_out.cpu().numpy().view(np.uint32)
    ph = ((got[:, -1].astype(np.uint64) - got[:, :-1].astype(np.uint64) @ s64) % Q).astype(np.int64)
    ef_fail = int(np.count_nonzero(((ph + Q // (2 * P)) // (Q // P)) % P != np.array(F)[msgs]))
    d_s = enc(rng.integers(0, 2 ** 32, batch, dtype=np.uint64), 2 ** 32, 12)
    t_sg = timed(lambda: ctx.eval_sign(d_s, 2 ** 32, d_out))
    ctx.close()
    return {"batch": batch, "reps": reps,
            "bootstrap_func_per_s": round(batch / t_bf, 1),
            "eval_func_per_s": round(batch / t_ef, 1), "eval_func_decrypt_failures": ef_fail,
            "eval_sign_mod_2_32_per_s": round(batch / t_sg, 1),
            "config": {"workload": "gate-context programmable bootstrap (n 636, N 1024, BSK 2^7 x 3, KS 2^2 x 8)",
                       "eval_func": "arbitrary table of Z_8 over Z_1024 (two bootstraps)",
                       "eval_sign": "Z_2^32 to Z_1024: 8 floors + 1 bootstrap = 17 bootstraps"}}


# form -> (bench key, pbs_approx, pbs_multibit, pfks_linear, lwe_to_rgsw, ks_linear)
AES_FORMS = {"exact": ("aes128_ctr", 0, 1, 0, 0, 0), "approx": ("aes128_ctr_one_product_pbs", 1, 1, 0, 0, 0),
             "approx-mb2": ("aes128_ctr_one_product_pbs_2bit", 1, 2, 0, 0, 0),
             "exact-lin": ("aes128_ctr_linear_pfks", 0, 1, 1, 0, 0),
             "approx-lin": ("aes128_ctr_one_product_pbs_linear_pfks", 1, 1, 1, 0, 0),
             "approx-mb2-lin": ("aes128_ctr_one_product_pbs_2bit_linear_pfks", 1, 2, 1, 0, 0),
             "approx-auto-lin": ("aes128_ctr_one_product_pbs_auto_bits_linear_pfks", 1, 3, 1, 0, 0),
             "approx-auto-links": ("aes128_ctr_one_product_pbs_auto_bits_linear_pfks_linear_ks", 1, 3, 1, 0, 1),
             "htss": ("aes128_ctr_homtrace_schemeswitch", 0, 1, 0, 1, 0)}


def form_text(pbs_approx, pbs_multibit, pfks_linear, lwe_to_rgsw, ks_linear=0):
    """(PBS, LWE->RGSW conversion, N -> n key switch) of a form, in words."""
    ks = ("linear key, signed digits, int8 GEMM on the matrix cores" if ks_linear
          else "lookup key, LDS-staged rows")
    if lwe_to_rgsw:
        return ("2-limb exact, gadget 2^10 x 4, LUT accumulator (BSK sigma 2^-54)",
                "HomTrace (11 automorphism key switches) + SchemeSwitch, gadget 2^8 x 5, torus-64 keys", ks)
    pbs = ("one FP64 product per digit polynomial, rounding error ~2^-33" if pbs_approx else "2-limb exact") + \
        (", two key bits per blind-rotation step" if pbs_multibit == 2 else
         ", two key bits per step unless the autotuner measured one bit faster" if pbs_multibit == 3 else "")
    return pbs, ("PFKS, linear 40-bit key as an int8 GEMM on the matrix cores" if pfks_linear
                 else "PFKS, lookup key, LDS selects"), ks


def htss_kernel_rooflines(ctx, nb):
    """One circuit-bootstrap launch of nb inputs in the HomTrace + SchemeSwitch form: the PBS
    (events inside the library) and the HT+SS kernel (the rest of the launch) against the FP64
    peak by their algorithmic FLOPs (2-limb exact PBS count; 2 levels x 12 key switches)."""
    import numpy as np
    import torch
    rng = np.random.default_rng(13)
    lwe = torch.from_numpy(rng.integers(0, 2 ** 32, (nb, N_LWE + 1), dtype=np.uint64).astype(np.uint32)
                           .view(np.int32)).to("cuda")
    rgsw = torch.empty((nb, 4, 2, N_CB), dtype=torch.int32, device="cuda")
    ctx.cbs_batch(lwe, rgsw)
    torch.cuda.synchronize()
    ctx.timing_read()
    ctx.timing_enable(True)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    ctx.cbs_batch(lwe, rgsw)
    e1.record()
    torch.cuda.synchronize()
    total = e0.elapsed_time(e1)
    pbs_ms, _, _ = ctx.timing_read()
    ctx.timing_enable(False)
    ht_ms = max(total - pbs_ms, 1e-6)
    pbs_tf = nb * N_LWE * F_EP_CB[2] / (pbs_ms * 1e-3) / 1e12
    ht_tf = nb * F_HTSS_CB / (ht_ms * 1e-3) / 1e12
    return {"inputs": nb,
            "pbs": {"ms": round(pbs_ms, 3), "bound": "fp64", "achieved": round(pbs_tf, 2),
                    "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s", "frac": round(pbs_tf / PEAK_FP64_TFLOPS, 4),
                    "flop_per_input": N_LWE * F_EP_CB[2]},
            "homtrace_schemeswitch": {"ms": round(ht_ms, 3), "bound": "fp64", "achieved": round(ht_tf, 2),
                                      "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                                      "frac": round(ht_tf / PEAK_FP64_TFLOPS, 4), "flop_per_input": F_HTSS_CB}}


def aes_headline(args):
    """AES-128 CTR transciphering of args.blocks blocks per step in the form args.aes_form:
    W untimed steps, then K timed steps with every PBS launch bracketed by HIP events
    inside the library (tfhe_cbs_timing_*) on the stream it runs on; then one launch of
    each side batch size.  Returns (headline dict, side dict)."""
    import numpy as np
    import torch
    import tfhe_amd
    key_name, approx, mb, lin, l2r, ksl = AES_FORMS[args.aes_form]
    sides = [int(b) for b in args.side_blocks.split(",") if b]
    maxb = max([args.blocks] + sides)
    ctx = tfhe_amd.CbsContext(max_blocks=maxb, params=tfhe_amd.cbs_params(approx, mb, lin, l2r, ksl))
    sk = tfhe_amd.lwe_keygen(N_LWE, 3)
    t0 = time.perf_counter()
    ctx.keygen(sk, 3)
    torch.cuda.synchronize()
    keygen_s = time.perf_counter() - t0
    rng = np.random.default_rng(11)
    key = bytes(rng.integers(0, 256, 16, dtype=np.uint8))
    ctx.set_round_keys(tfhe_amd.aes_encrypt_key_bits(sk, 5, key))
    nonce = bytes(rng.integers(0, 256, 8, dtype=np.uint8))
    d_ks = torch.empty((maxb, 128, N_LWE + 1), dtype=torch.int32, device="cuda")
    s64 = sk.astype(np.uint64)

    def workload(blocks):  # counters, seeded random plaintexts and their public AES-CTR ciphertexts
        ctrs = [nonce + int(b).to_bytes(8, "big") for b in range(blocks)]
        plain = [bytes(rng.integers(0, 256, 16, dtype=np.uint8)) for _ in range(blocks)]
        cipher = [bytes(p ^ k for p, k in zip(pt, tfhe_amd.aes128_encrypt_plain(key, c)))
                  for pt, c in zip(plain, ctrs)]
        return ctrs, plain, cipher

    def wrong_blocks(blocks, plain):
        got = d_ks[:blocks].cpu().numpy().view(np.uint32)
        ph = (got[:, :, -1].astype(np.uint64) - (got[:, :, :-1].astype(np.uint64) @ s64)) % 2 ** 32
        bits = ((ph + (1 << 30)) >> 31) & 1
        bad = 0
        for b in range(blocks):
            v = bytes(int(sum(int(bits[b, 8 * k + j]) << j for j in range(8))) for k in range(16))
            bad += v != plain[b]
        return int(bad)

    blocks = args.blocks
    ctrs, plain, cipher = workload(blocks)
    tuned = None
    if not args.no_aes_tune:
        t0 = time.perf_counter()
        tuned = ctx.autotune(blocks)
        tuned["tune_s"] = round(time.perf_counter() - t0, 3)
    for _ in range(args.warmup):
        ctx.aes_ctr_transcipher(ctrs, cipher, d_ks)
    torch.cuda.synchronize()
    ctx.timing_enable(True)
    t0 = time.perf_counter()
    for k in range(args.steps):
        ctx.aes_ctr_transcipher(ctrs, cipher, d_ks)
        print(f"aes128 step {k + 1}/{args.steps} issued", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    pbs_ms, launches, inputs = ctx.timing_read()
    ctx.timing_enable(False)
    bad = wrong_blocks(blocks, plain)
    print(f"aes128 {blocks} blocks x {args.steps} steps: {elapsed:.2f} s, {bad} wrong blocks", file=sys.stderr,
          flush=True)
    per_launch_inputs = inputs / max(launches, 1)
    avg_ms = pbs_ms / max(launches, 1)
    f_unit = N_LWE * F_EP_CB[2] if l2r else F_PBS_CB  # the exact form's count for the HT+SS PBS
    tflops = per_launch_inputs * f_unit / (avg_ms * 1e-3) / 1e12
    traffic = None
    prof = os.path.join(ROOT, "profiles", "pbs_two_bit_aes_256_blocks_traffic.json")
    if os.path.exists(prof):
        traffic = json.load(open(prof)).get("hbm_bytes_per_launch")
    head = {"value": round(blocks * args.steps / elapsed, 3), "elapsed_s": elapsed, "blocks_wrong": bad,
            "keygen_s": round(keygen_s, 2), "tuned_shapes": tuned,
            "roofline": {"bound": "valu", "kernel": "k_pbs2_manylut (" + form_text(approx, mb, lin, l2r, ksl)[0] + ")",
                         "achieved": round(tflops, 3), "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                         "frac": round(tflops / PEAK_FP64_TFLOPS, 4),
                         "traffic": traffic if args.aes_form in ("approx-auto-lin", "approx-auto-links") else None,
                         "algorithmic_flop_per_launch": int(per_launch_inputs * f_unit),
                         "avg_launch_ms": round(avg_ms, 3), "launches_timed": launches,
                         "inputs_per_launch": int(per_launch_inputs),
                         "pbs_share_of_step": round(pbs_ms / (elapsed * 1e3), 4)}}
    side = {}
    for sb in sides:
        s_ctrs, s_plain, s_cipher = workload(sb)
        s_tuned = None
        if not args.no_aes_tune:
            s_tuned = ctx.autotune(sb)
        ctx.aes_ctr_transcipher(s_ctrs, s_cipher, d_ks)  # first launch of this size
        torch.cuda.synchronize()
        reps = 3 if sb <= 32 else 1
        t0 = time.perf_counter()
        for _ in range(reps):
            ctx.aes_ctr_transcipher(s_ctrs, s_cipher, d_ks)
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / reps
        entry = {"blocks_per_s": round(sb / dt, 3), "s_per_launch": round(dt, 4), "reps": reps,
                 "circuit_bootstraps_per_s": round(sb * 1280 / dt, 1), "blocks_wrong": wrong_blocks(sb, s_plain),
                 "tuned_shapes": s_tuned}
        if sb <= 32:  # the keystream replayed from its captured hipGraph
            eager = d_ks[:sb].clone()
            ctx.aes_ctr_transcipher(s_ctrs, s_cipher, d_ks, mode="graph")  # capture + first replay
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            ctx.aes_ctr_transcipher(s_ctrs, s_cipher, d_ks, mode="graph")
            torch.cuda.synchronize()
            entry["graph_s_per_launch"] = round(time.perf_counter() - t0, 4)
            entry["graph_equals_eager"] = bool(torch.equal(d_ks[:sb], eager))
        side[str(sb)] = entry
        print(f"aes128 side batch {sb} blocks: {sb / dt:.2f} blocks/s, {entry['blocks_wrong']} wrong",
              file=sys.stderr, flush=True)
    side["kernels_at_17_blocks"] = (htss_kernel_rooflines(ctx, min(17, maxb) * 128) if l2r else
                                    aes_kernel_rooflines(ctx, approx, min(17, maxb) * 128, lin))
    ctx.close()
    return head, side


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5, help="timed AES steps (one transcipher call of --blocks blocks)")
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--blocks", type=int, default=256, help="AES blocks per step (the headline batch)")
    ap.add_argument("--side-blocks", default="1,17,1024",
                    help="comma-separated AES batch sizes timed once each beside the headline ('' to skip)")
    ap.add_argument("--aes-form", default="approx-auto-links", choices=sorted(AES_FORMS),
                    help="PBS / PFKS / key-switch form of the headline: exact (two-limb) or approx (one FP64 "
                         "product) PBS, -mb2 two key bits per step, -auto both key forms (two bits unless tuned to "
                         "one), -lin the linear-key PFKS on the matrix cores, -links also the linear-key N -> n key "
                         "switch; htss: HomTrace + SchemeSwitch")
    ap.add_argument("--extra-forms", default="htss",
                    help="comma-separated further AES forms timed at --extra-blocks (one launch per size; "
                         "'' to skip)")
    ap.add_argument("--extra-blocks", default="1,17,256")
    ap.add_argument("--aes-plain-blocks", type=int, default=0,
                    help="AES blocks timed in the per-bootstrap form of the exact PBS (a launch-overhead figure: "
                         "every bootstrap, LUT and key switch is a launch of one; 0 to skip)")
    ap.add_argument("--no-aes-tune", action="store_true",
                    help="keep the kernels' model launch shapes (no tfhe_cbs_autotune before each AES batch)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--gate-steps", type=int, default=20, help="timed steps of the gate line (0 to skip)")
    ap.add_argument("--plain-steps", type=int, default=3, help="plain per-bootstrap gate samples")
    ap.add_argument("--cbs-lut-reps", type=int, default=5,
                    help="timed calls of the 256-bit circuit bootstrap + LUT line (0 to skip)")
    ap.add_argument("--func-reps", type=int, default=3,
                    help="timed calls of the gate-context function line (0 to skip)")
    ap.add_argument("--campaign", type=int, default=1_000_000,
                    help="gate bootstraps in the decryption-failure campaign (0 to skip)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run three 512 MiB device copies (the HBM-counter calibration kernel)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              f"(bootstraps and AES blocks are independent: a multi-GPU run would give each GPU a key replica "
              f"and a slice of the batch)")
        sys.exit(2)

    import torch
    import tfhe_amd
    torch.cuda.init()
    tfhe_amd.lib()

    head, side = aes_headline(args)
    elapsed = head.pop("elapsed_s")
    line = {
        "metric": METRIC,
        "value": head.pop("value"),
        "unit": "AES-128 blocks/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed * 1e3 / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"AES-128 CTR transciphering, {args.blocks} blocks per launch (uneven-batch config "
                               f"1/17/256/1024 blocks; 256 the headline, the others side keys)",
                   "blocks_per_step": args.blocks, "form": args.aes_form,
                   "n": N_LWE, "N": N_CB, "pbs_gadget": "2^10 x 4" if AES_FORMS[args.aes_form][4] else "2^9 x 4",
                   "cb_gadget": "2^8 x 2",
                   "conversion_gadget": "2^8 x 5" if AES_FORMS[args.aes_form][4] else "2^4 x 8",
                   "ks_gadget": "2^2 x 8",
                   "pbs": form_text(*AES_FORMS[args.aes_form][1:])[0],
                   "lwe_to_rgsw": form_text(*AES_FORMS[args.aes_form][1:])[1],
                   "keyswitch": form_text(*AES_FORMS[args.aes_form][1:])[2],
                   "circuit_bootstraps_per_step": 1280 * args.blocks},
        "roofline": head.pop("roofline"),
        "blocks_wrong": head.pop("blocks_wrong"),
        "aes128_side_batches": side,
        "aes128_keygen_s": head.pop("keygen_s"),
        "aes128_tuned_shapes": head.pop("tuned_shapes"),
    }
    if args.extra_forms:
        sizes = [int(b) for b in args.extra_blocks.split(",") if b]
        for f in [f for f in args.extra_forms.split(",") if f]:
            key, approx, mb, lin, l2r, ksl = AES_FORMS[f]
            line[key] = aes_ctr_bench(sizes, pbs_approx=approx, pbs_multibit=mb,
                                      plain_blocks=args.aes_plain_blocks if f == "exact" else 0,
                                      tune=not args.no_aes_tune, pfks_linear=lin, lwe_to_rgsw=l2r, ks_linear=ksl)
    line["one_product_vs_exact_product"] = one_product_error()
    line["split_fp32_product"] = split32_product_bench()
    if args.gate_steps > 0:
        line["gate_bootstraps"] = gate_bench(args)
    if args.cbs_lut_reps > 0:
        line["cbs_lut_256_bits"] = cbs_lut_bench(args.cbs_lut_reps)
    if args.func_reps > 0:
        line["gate_functions"] = gate_functions_bench(args.func_reps)
    if args.copy_calibration:
        src = torch.empty(512 << 20, dtype=torch.uint8, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = aes_cpu_baseline()
        if args.gate_steps > 0:
            line["gate_bootstraps"]["cpu_baseline"] = cpu_baseline()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
