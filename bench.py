"""Benchmark: Kyber-768 decapsulations per second at batch 65536 on one MI355X,
fused persistent kernel vs the plain per-op kernels, with the CPU oracle timed
on the host cores in the same run.

A step is one decapsulation pass over one batch of 65536 requests (each with
its own key pair) whose ciphertexts and secret keys are already resident in
HBM.  The inputs come from seeded coins, expanded by the engine's own key
generation and encapsulation kernels.  Prints one JSON line.

--workload frodo976 measures FrodoKEM-976 decapsulation of 4096 ciphertexts
under one device-resident key, with the shared-A int8-MFMA product timed on
its own; --workload newhope1024 measures 16384 NewHope-1024 CPA
keygen+encaps+decaps triples (and the CCA-KEM's rates beside them); --workload kyber_mixed measures 65536 mixed
Kyber-512/768/1024 requests (1:2:1) through the graph-captured fused kernels;
--workload serving replays a Poisson trace (2M requests/s, 90% Kyber 1:2:1,
10% FrodoKEM-976, 1024 device-resident key pairs per set) through the serving
loop, fused and plain.  All print one JSON line of the same shape.
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

BATCH = 65536
ALG = "kyber768"
K = 3
KECCAK_PER_DECAPS = 54         # G 1 + A^T 36 + PRF 7 + H(ct) 9 + KDF 1
OPS_PER_KECCAK = 7440          # 24 rounds x 155 64-bit ops = 7440 32-bit ops
NEWHOPE_KECCAK_PER_ENC = 64    # gen_a 16 SHAKE128 streams (one block each) + s', e', e'' 3 x 16 SHAKE256
BYTES_PER_DECAPS = 1088 + 2400 + 32  # compulsory: ct in, sk in, ss out
# 32-bit vector peak: 256 CUs x 4 SIMDs x 64 lanes per 2 cycles x 2.4 GHz = 78.6 T lane-ops/s, the
# lane rate of the published 157.3 TFLOP/s FP32 vector peak (one FMA = 2 flops).  valu_issue_check()
# measures the sustained issue rate on the box (pqc_valu_issue_rate).
VALU_PEAK_TOPS = 256 * 4 * 64 / 2 * 2.4e9 / 1e12
HBM_PEAK_GBPS = 8000.0
# Bytes between L2 and memory per fused decapsulation launch at batch 65536
# (k_fused_dec32<3>) from rocprofv3 FETCH_SIZE and WRITE_SIZE runs, each scaled
# by the same counter's reading of a 256 MiB device copy
# (profiles/kyber768_decaps_b65536.md): 786.8 MB read + 515.1 MB written.  Most of
# it is the A^T scratch of the 32-request tiles, written and read back per wave.
MEASURED_TRAFFIC_BYTES_B65536 = 1301.9e6
# HBM bytes per Frodo-976 decapsulation GEMM launch at batch 4096, same method
# (profiles/frodo976_decaps_b4096.md): 186.2 MB read + 0.07 MB written.
MEASURED_TRAFFIC_FRODO_GEMM_B4096 = 186.3e6
# HBM bytes per NewHope-1024 fused encryption launch at batch 16384, same method
# (profiles/newhope1024_cpa_triples_b16384.md): 35.86 MB read + 35.65 MB written.
MEASURED_TRAFFIC_NEWHOPE_ENC_B16384 = 71.5e6


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def usable_cores():
    return len(os.sched_getaffinity(0))


def cpu_threads(args):
    """Threads for the CPU baselines: every core this process may run on, unless
    --cpu-threads names a number."""
    return args.cpu_threads if args.cpu_threads > 0 else usable_cores()


def host_cpu_info():
    """The host the CPU baselines ran on: logical CPUs, the ones this process may
    use, and the model name."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.lower().startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"host_cpus": os.cpu_count(), "usable_cpus": usable_cores(), "cpu_model": model}


def timed_cpu(run, total, first, min_s=1.0):
    """Times the CPU oracle for at least `min_s` seconds.  run(m) processes items
    0..m-1 (and checks them); a first pass over `first` items sets the rate, then
    one pass over enough items (up to `total`), repeated while the total time is
    under min_s.  Returns (items processed, seconds) of the timed passes."""
    t0 = time.perf_counter()
    run(min(first, total))
    dt0 = max(time.perf_counter() - t0, 1e-6)
    m = int(min(total, max(first, first / dt0 * min_s * 1.1)))
    done, dt = 0, 0.0
    while dt < min_s:
        t0 = time.perf_counter()
        run(m)
        dt += time.perf_counter() - t0
        done += m
    return done, dt


def cpu_entry(value, unit, thr, sample, single=None):
    out = {"value": round(value, 1), "unit": unit, "cores": thr, "kind": "port", "sample": sample}
    out.update(host_cpu_info())
    if single is not None:
        out["single_core_value"] = round(single, 1)
    return out


# GPU time each side measurement (the legs beside a workload's own timed steps) runs
# before its timing starts
SECONDARY_WARM_MS = 30.0


def timed(fn, stream, steps, warmup, pqc, warm_ms=0.0):
    """`warmup` untimed launches; with warm_ms, more until about warm_ms of GPU time
    has run back to back: after a host-side phase leaves the GPU idle, its first
    tens of milliseconds of work run up to ~10% slower."""
    for _ in range(warmup):
        fn()
    if warm_ms > 0:
        a, b = pqc.Event(), pqc.Event()
        a.record(stream)
        fn()
        b.record(stream)
        b.synchronize()
        for _ in range(min(1000, int(warm_ms / max(a.elapsed_ms(b), 0.01)))):
            fn()
    stream.synchronize()
    evs = [(pqc.Event(), pqc.Event()) for _ in range(steps)]
    t0 = time.perf_counter()
    for a, b in evs:
        a.record(stream)
        fn()
        b.record(stream)
    stream.synchronize()
    wall = time.perf_counter() - t0
    per = [a.elapsed_ms(b) for a, b in evs]
    return wall, per


def calibrate_copy(args, pqc, stream):
    """One device-to-device copy of a known size, for scaling HBM byte counters."""
    if args.calibrate_copy_mb:
        nb = args.calibrate_copy_mb << 20
        src, dst = pqc.DeviceBuffer(nb), pqc.DeviceBuffer(nb)
        pqc.lib().pqc_memcpy_dtod(dst.ptr, src.ptr, nb, stream.handle)
        stream.synchronize()
        log("calibration copy: %d MiB device to device" % args.calibrate_copy_mb)
        del src, dst


def timed_run(fn, stream, launches, warmup, pqc):
    """Average duration of one launch over `launches` back-to-back launches between
    one pair of events on the stream (no per-launch event overhead)."""
    for _ in range(warmup):
        fn()
    a, b = pqc.Event(), pqc.Event()
    a.record(stream)
    for _ in range(launches):
        fn()
    b.record(stream)
    b.synchronize()
    return a.elapsed_ms(b) / launches


VALU_OPS = ("v_xor_b32", "v_add_u32", "v_alignbit_b32", "v_mul_lo_u32", "v_bitop3_b32", "v_lshlrev_b32",
            "v_mad_u32_u24", "v_sub_u32", "v_and_b32", "v_ashrrev_i32", "v_fma_f32", "v_mul_f32", "v_mul_u32_u24",
            "v_mad_i32_i24", "v_bfe_u32", "v_perm_b32")


def valu_issue_check(pqc):
    """Sustained cycles per wave64 instruction per SIMD of sixteen instruction
    kinds at 8, 2 and 1 waves per SIMD (pqc_valu_issue_rate), the simple-op rate
    the 8-wave figures imply against VALU_PEAK_TOPS, and the Keccak-f throughput
    of the three state layouts at 1 and 2 waves per SIMD (pqc_keccak_rate)."""
    out = {}
    for wps in (8, 2, 1):
        row = {}
        for op, name in enumerate(VALU_OPS):
            cyc = ctypes.c_double()
            rc = pqc.lib().pqc_valu_issue_rate(op, wps, ctypes.byref(cyc), None)
            if rc:
                raise pqc.PQCError(rc, "pqc_valu_issue_rate")
            row[name] = round(cyc.value, 2)
        out["cycles_%d_waves_per_simd" % wps] = row
    best = min(out["cycles_8_waves_per_simd"]["v_xor_b32"], out["cycles_8_waves_per_simd"]["v_add_u32"])
    out["simple_op_rate_T"] = round(256 * 4 * 64 / best * 2.4e9 / 1e12, 2)
    kr = {}
    for form, name in enumerate(("single_lane", "lane_pair", "lane_pair_interleaved")):
        for wps in (1, 2):
            r = ctypes.c_double()
            rc = pqc.lib().pqc_keccak_rate(form, wps, ctypes.byref(r), None)
            if rc:
                raise pqc.PQCError(rc, "pqc_keccak_rate")
            kr["%s_%d_waves_per_simd" % (name, wps)] = round(r.value / 1e9, 3)
    out["keccak_G_state_perms_per_s"] = kr
    log("VALU issue (cycles per wave64 instruction per SIMD): %s" % out)
    return out


def host_buffer_rates(pqc, ctx, n, ct, sk, ss_want, reps=3):
    """PCIe-inclusive decapsulation rate: ciphertexts and secret keys start in
    host memory and the shared secrets end there (PQC_F_HOST_PTRS), from
    pageable numpy arrays (staged through the context's pinned double buffers)
    and from pqc_host_alloc memory (copied directly).  Not the headline value,
    which has its inputs resident in HBM."""
    ct_h, sk_h = ct.download((n, 1088)), sk.download((n, 2400))
    res = {}
    pin = [pqc.PinnedBuffer(n * w) for w in (1088, 2400, 32)]
    p_ct, p_sk, p_ss = pin[0].view((n, 1088)), pin[1].view((n, 2400)), pin[2].view((n, 32))
    p_ct[:] = ct_h
    p_sk[:] = sk_h
    for name, args_ in (("pageable", (ct_h, sk_h, None)), ("pinned", (p_ct, p_sk, p_ss))):
        got = ctx.dec_host(ALG, args_[0], args_[1], out=args_[2])  # warm-up + check
        assert (got == ss_want).all(), "host-buffer decaps differ from the device-buffer run"
        t0 = time.perf_counter()
        for _ in range(reps):
            ctx.dec_host(ALG, args_[0], args_[1], out=args_[2])
        dt = (time.perf_counter() - t0) / reps
        res[name + "_per_s"] = round(n / dt, 1)
        res[name + "_ms"] = round(dt * 1e3, 3)
        res[name + "_GBps_h2d"] = round(n * (1088 + 2400) / dt / 1e9, 2)
        log("host buffers (%s): %.3f ms per %d decaps" % (name, dt * 1e3, n))
    for b in pin:
        b.free()
    res["note"] = "ct + sk host->device, decaps, ss device->host per call; bytes match the device run"
    return res


def small_batch_latency(pqc, ctx, stream, ct, sk, ss, ss_want, reps=50):
    """Kernel time of decapsulation batches of 1, 64 and 256 requests (HIP events,
    p50 / p99 over `reps` launches): the one-workgroup-per-request kernel that
    such batches take by default, and the 64-request tile kernel forced with
    lowlat_max = -1.  The outputs are checked equal to the full batch's."""
    from pqc_amd import tuner as T
    out = {}
    for m in (1, 64, 256):
        row = {}
        for name, lim in (("lowlat", 0), ("tiles", -1)):
            T.put_tuning(ctx, lowlat_max=lim)
            _, per = timed(lambda: ctx.dec(ALG, m, ss, ct, sk, stream=stream), stream, reps, 3, pqc)
            got = ss.download((m, 32))
            assert (got == ss_want[:m]).all(), "small-batch decaps differ from the full batch"
            row[name] = {"p50": round(float(np.percentile(per, 50)) * 1e3, 1),
                         "p99": round(float(np.percentile(per, 99)) * 1e3, 1)}
        out["n%d" % m] = row
        log("decaps n=%d: lowlat p50 %.1f us, tiles p50 %.1f us" % (m, row["lowlat"]["p50"], row["tiles"]["p50"]))
    T.put_tuning(ctx, lowlat_max=0)
    return out


def cpu_baseline(ct_h, sk_h, ss_gpu, threads):
    """The oracle's decapsulation of the same batch on `threads` host threads and
    on one, each timed for at least a second; outputs checked equal to the GPU's."""
    from tests import oracle_lib as O

    def run(thr):
        def f(m):
            ss = O.kyber_batch("dec", K, [np.ascontiguousarray(ct_h[:m]), np.ascontiguousarray(sk_h[:m])],
                               nthreads=thr)
            assert (ss == ss_gpu[:m]).all(), "CPU oracle and GPU disagree"
        return f

    O.kyber_batch("dec", K, [ct_h[:64], sk_h[:64]], nthreads=threads)  # warm
    n1, dt = timed_cpu(run(threads), len(ct_h), 256 * threads)
    n2, dt1 = timed_cpu(run(1), len(ct_h), 512)
    return n1 / dt, n1, n2 / dt1, n2


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--batch", type=int, default=BATCH)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="host threads for the CPU baselines (0 = every core this process may use)")
    ap.add_argument("--only", choices=["fused", "plain"], default=None,
                    help="time one path only (profiling runs)")
    ap.add_argument("--calibrate-copy-mb", type=int, default=0,
                    help="first run one device-to-device copy of this many MiB (counter calibration)")
    ap.add_argument("--workload", choices=["kyber768", "frodo976", "newhope1024", "kyber_mixed", "serving"],
                    default="kyber768")
    ap.add_argument("--rate", type=float, default=2e6, help="serving: Poisson arrival rate, requests/s")
    ap.add_argument("--duration", type=float, default=0.5, help="serving: trace length, seconds")
    ap.add_argument("--slots", type=int, default=4,
                    help="serving: server slots (with Kyber and Frodo in graph mode, half take the Frodo requests)")
    ap.add_argument("--no-secondary", action="store_true",
                    help="kyber768: skip the other configurations' figures (config.secondary)")
    ap.add_argument("--fixed-rate", type=float, default=4e6,
                    help="kyber768: Poisson arrival rate of the fixed-rate latency phase, decaps/s (0 = skip)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the engine shards by request with no collective, "
              "so multi-GPU runs are not measured here (use --gpus 1)", file=sys.stderr)
        sys.exit(2)
    # The serving loop drives up to 3 streams per slot; with HIP's default of 4
    # hardware queues per process, streams share queues and batches serialise
    # behind each other.  Must be set before the runtime initialises.
    if int(os.environ.get("GPU_MAX_HW_QUEUES", "4")) < 16:
        os.environ["GPU_MAX_HW_QUEUES"] = "16"
    if args.workload == "frodo976":
        return bench_frodo(args)
    if args.workload == "newhope1024":
        return bench_newhope(args)
    if args.workload == "kyber_mixed":
        return bench_mixed(args)
    if args.workload == "serving":
        return bench_serving(args)

    import pqc_amd as pqc
    n = args.batch
    ctx = pqc.Context([ALG], max_batch=n)
    stream = pqc.Stream()
    rng = np.random.default_rng(0)
    coins = pqc.DeviceBuffer.from_array(rng.integers(0, 256, (n, 64), dtype=np.uint8))
    ecoins = pqc.DeviceBuffer.from_array(rng.integers(0, 256, (n, 32), dtype=np.uint8))
    pk, sk = pqc.DeviceBuffer(n * 1184), pqc.DeviceBuffer(n * 2400)
    ct, ss_enc = pqc.DeviceBuffer(n * 1088), pqc.DeviceBuffer(n * 32)
    ss_f, ss_p = pqc.DeviceBuffer(n * 32), pqc.DeviceBuffer(n * 32)
    ctx.keypair_derand(ALG, n, pk, sk, coins, stream=stream)
    ctx.enc_derand(ALG, n, ct, ss_enc, pk, ecoins, stream=stream)
    stream.synchronize()
    log("inputs ready: %d Kyber-768 key pairs and ciphertexts in HBM" % n)
    calibrate_copy(args, pqc, stream)

    fused = lambda: ctx.dec(ALG, n, ss_f, ct, sk, stream=stream)
    plain = lambda: ctx.dec(ALG, n, ss_p, ct, sk, plain=True, stream=stream)
    if args.only == "plain":
        wall_p, per_p = timed(plain, stream, args.steps, args.warmup, pqc)
        log("plain: %.3f ms/step" % (wall_p * 1e3 / args.steps))
        return
    wall_f, per_f = timed(fused, stream, args.steps, args.warmup, pqc)
    log("fused: %.3f ms/step" % (wall_f * 1e3 / args.steps))
    if args.only == "fused":
        return
    wall_p, per_p = timed(plain, stream, max(3, args.steps // 4), 1, pqc, warm_ms=SECONDARY_WARM_MS)
    log("plain: %.3f ms/step" % (wall_p * 1e3 / max(3, args.steps // 4)))

    a, b, e = ss_f.download((n, 32)), ss_p.download((n, 32)), ss_enc.download((n, 32))
    bitexact = bool((a == b).all() and (a == e).all())

    # the rest of the triple on the same batch: key generation and encapsulation,
    # fused vs plain (outputs compared), beside the decapsulation above
    pk2, sk2, ct2, ss2 = (pqc.DeviceBuffer(n * b_) for b_ in (1184, 2400, 1088, 32))
    tsteps = max(3, args.steps // 4)
    tri = {}
    for name, f_fn, p_fn in (
            ("keypair", lambda: ctx.keypair_derand(ALG, n, pk2, sk2, coins, stream=stream),
             lambda: ctx.keypair_derand(ALG, n, pk2, sk2, coins, plain=True, stream=stream)),
            ("encaps", lambda: ctx.enc_derand(ALG, n, ct2, ss2, pk, ecoins, stream=stream),
             lambda: ctx.enc_derand(ALG, n, ct2, ss2, pk, ecoins, plain=True, stream=stream))):
        wf, _ = timed(f_fn, stream, tsteps, 1, pqc, warm_ms=SECONDARY_WARM_MS)
        out_f = (pk2 if name == "keypair" else ct2).download((n, 1184 if name == "keypair" else 1088))
        wp, _ = timed(p_fn, stream, tsteps, 1, pqc, warm_ms=SECONDARY_WARM_MS)
        out_p = (pk2 if name == "keypair" else ct2).download((n, 1184 if name == "keypair" else 1088))
        tri[name] = {"fused_per_s": round(n / (wf / tsteps), 1), "plain_per_s": round(n / (wp / tsteps), 1),
                     "fused_over_plain": round(wp / wf, 3), "bitexact": bool((out_f == out_p).all())}
        log("%s: fused %.3f ms, plain %.3f ms" % (name, wf * 1e3 / tsteps, wp * 1e3 / tsteps))

    host_io = host_buffer_rates(pqc, ctx, n, ct, sk, a)
    valu_check = valu_issue_check(pqc)
    small = small_batch_latency(pqc, ctx, stream, ct, sk, ss_f, a)

    ms_f = wall_f * 1e3 / args.steps
    ms_p = wall_p * 1e3 / max(3, args.steps // 4)
    kern_ms = float(np.mean(per_f))
    rate_f = n / (ms_f / 1e3)
    rate_p = n / (ms_p / 1e3)
    ops = n * KECCAK_PER_DECAPS * OPS_PER_KECCAK
    achieved = ops / (kern_ms / 1e3) / 1e12

    out = {
        "metric": "Kyber-768 decaps/s at batch 64K (fused vs plain); p99 latency at fixed rate",
        "value": round(rate_f, 1),
        "unit": "decaps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_f, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int16",
        "data": "synthetic: seeded coins, keys and ciphertexts generated on the GPU by the engine",
        "config": {
            "workload": "kyber768_decaps_fused",
            "alg": "Kyber-768 round 2 (k=3, eta=2, du=10, dv=4)",
            "batch": n,
            "plain_decaps_per_s": round(rate_p, 1),
            "plain_ms_per_step": round(ms_p, 4),
            "fused_over_plain": round(rate_f / rate_p, 3),
            "batch_latency_ms_p50": round(float(np.percentile(per_f, 50)), 4),
            "batch_latency_ms_p99": round(float(np.percentile(per_f, 99)), 4),
            "bitexact_fused_plain_encaps": bitexact,
            "keypair": tri["keypair"],
            "encaps": tri["encaps"],
            "host_buffers": host_io,
            "small_batch_latency_us": small,
        },
        "roofline": {
            "bound": "valu",
            "kernel": "k_fused_dec32<3> (fused Kyber-768 decaps, 32-request tiles, two waves per SIMD)",
            "achieved": round(achieved, 3),
            "peak": round(VALU_PEAK_TOPS, 2),
            "unit": "Tops/s (int32)",
            "frac": round(achieved / VALU_PEAK_TOPS, 4),
            "peak_note": "256 CUs x 4 SIMDs x 64 lanes / 2 cycles x 2.4 GHz (the FP32 vector peak's lane rate); "
                         "issue_check: sustained cycles per wave64 instruction measured on this GPU",
            "issue_check": valu_check,
            "work_per_launch": "%d decaps x 54 Keccak-f x 7440 int32 ops" % n,
            "kernel_ms": round(kern_ms, 4),
            "hbm_GBps_compulsory": round(n * BYTES_PER_DECAPS / (kern_ms / 1e3) / 1e9, 1),
            "hbm_frac_compulsory": round(n * BYTES_PER_DECAPS / (kern_ms / 1e3) / 1e9 / HBM_PEAK_GBPS, 4),
            "traffic": MEASURED_TRAFFIC_BYTES_B65536 if n == BATCH else None,
        },
        "cpu_baseline": None,
    }
    if not args.no_cpu:
        ct_h = ct.download((n, 1088))
        sk_h = sk.download((n, 2400))
        thr = cpu_threads(args)
        r_all, s_all, r_one, s_one = cpu_baseline(ct_h, sk_h, a, thr)
        out["cpu_baseline"] = cpu_entry(
            r_all, "decaps/s", thr,
            "%d Kyber-768 decaps of the same batch on %d threads (and %d on 1 thread: %.1f/s), each leg timed for "
            "at least 1 s; outputs checked equal to the GPU's" % (s_all, thr, s_one, r_one), single=r_one)
        out["config"]["gpu_over_cpu"] = round(rate_f / r_all, 1)
    ctx.close()
    if args.fixed_rate > 0:
        out["config"]["fixed_rate"] = fixed_rate_latency(args.fixed_rate, 0.25)
    if not args.no_secondary:
        out["config"]["secondary"] = secondary_workloads(args)
    print(json.dumps(out), flush=True)


def _compact(line, keys):
    """The figures of one workload's line that the headline line carries."""
    c, r, cpu = line["config"], line.get("roofline") or {}, line.get("cpu_baseline") or {}
    out = {"metric": line["metric"], "value": line["value"], "unit": line["unit"],
           "ms_per_step": line["ms_per_step"], "steps": line["steps"], "workload": c["workload"]}
    if r:
        out["roofline"] = {k: r[k] for k in ("kernel", "bound", "achieved", "peak", "unit", "frac", "kernel_ms",
                                             "kernel_timing", "traffic") if k in r}
    if cpu:
        out["cpu_baseline"] = {k: cpu[k] for k in ("value", "unit", "cores", "kind", "cpu_model") if k in cpu}
        out["gpu_over_cpu"] = c.get("gpu_over_cpu")
    for k in keys:
        out[k] = c.get(k)
    return out


def secondary_workloads(args):
    """The other configurations, each run as its own workload would run it (fewer
    steps), reduced to their main figures: FrodoKEM-976 shared-key decapsulation
    (with the int8-MFMA utilisation against its 40% target, and the pipelined
    rate), NewHope-1024 CPA triples fused vs plain, the graph-captured mixed
    Kyber batch, and a 0.1 s slice of the 2 M requests/s serving trace."""
    import copy
    sa = copy.copy(args)
    sa.steps, sa.warmup, sa.batch, sa.calibrate_copy_mb = min(args.steps, 10), min(args.warmup, 2), BATCH, 0
    sec = {}
    t0 = time.perf_counter()
    fr = run_frodo(sa)
    sec["frodo976"] = _compact(fr, ("batch", "decaps_equal_encaps", "pipelined", "batch_latency_ms_p50"))
    sec["frodo976"]["mfma_frac_target"] = 0.40
    log("secondary frodo976: %.1f s" % (time.perf_counter() - t0))
    t0 = time.perf_counter()
    nh = run_newhope(sa)
    sec["newhope1024"] = _compact(nh, ("batch", "plain_triples_per_s", "fused_over_plain", "fused_equals_plain",
                                       "messages_recovered"))
    log("secondary newhope1024: %.1f s" % (time.perf_counter() - t0))
    t0 = time.perf_counter()
    mx = run_mixed(sa)
    sec["kyber_mixed"] = _compact(mx, ("batch", "counts_512_768_1024", "tuning", "autotune", "fused_over_plain",
                                       "bitexact_fused_plain"))
    log("secondary kyber_mixed: %.1f s" % (time.perf_counter() - t0))
    t0 = time.perf_counter()
    sv = run_serving(sa, duration=0.1, with_plain=False)
    sec["serving"] = _compact(sv, ("offered_rate", "requests", "mode", "p50_us", "p99_us", "mean_batch", "errors"))
    log("secondary serving: %.1f s" % (time.perf_counter() - t0))
    return sec


def fixed_rate_latency(rate, duration, keys=1024):
    """p50 / p99 latency of Kyber-768 decapsulations arriving as a Poisson stream at
    `rate` per second, through the serving loop (continuous batching, 1024
    device-resident key pairs), fused kernels vs plain per-op kernels on the same
    trace; latency = arrival to completion, host ciphertext in to host secret out."""
    from pqc_amd.server import RequestSet, Server
    n = int(rate * duration)
    rng = np.random.default_rng(5)
    arrivals = np.cumsum(rng.exponential(1e6 / rate, size=n))
    algs = np.full(n, 2, np.int32)
    ops = np.full(n, 2, np.int32)
    kid = rng.integers(0, keys, n).astype(np.uint32)
    res = {"rate": rate, "requests": n, "slots": 2, "max_wait_us": 25}
    for plain, graph in ((False, False), (False, True), (True, False)):
        # two slots (one batch on the GPU while the next fills) and batches closed
        # after 25 us, which kept batches near 600 requests on every box measured
        # (50 us let some settle near 1000, doubling p50).  Batches of up to 8192
        # decapsulations run the one-workgroup-per-request kernel.
        srv = Server(["kyber768"], keys_per_alg=keys, key_seed=0, max_batch=65536, slots=2,
                     arena_bytes=128 << 20, max_wait_us=25, plain=plain, graph=graph)
        ws = RequestSet.random(algs, ops, kid, rng)
        srv.run_trace(arrivals, ws.reqs)  # warm-up: the same trace once
        rs = RequestSet.random(algs, ops, kid, rng)
        st, _ = srv.run_trace(arrivals, rs.reqs)
        srv.close()
        log("fixed rate %.1fM/s plain=%s graph=%s: %.0f decaps/s, p50 %.0f us, p99 %.0f us, mean batch %.0f" %
            (rate / 1e6, plain, graph, st.ops_per_s, st.p50_us, st.p99_us, st.mean_batch))
        res["plain" if plain else "fused_graph" if graph else "fused"] = {"completed_per_s": round(st.ops_per_s, 1), "p50_us": round(st.p50_us, 1),
                                              "p99_us": round(st.p99_us, 1), "mean_batch": round(st.mean_batch, 1),
                                              "errors": int(st.errors)}
    return res


# ---- FrodoKEM-976: 4096 decapsulations under one device-resident key -------------
FRODO_BATCH = 4096
FRODO_N, FRODO_NBAR = 976, 8
MFMA_I8_PEAK_TOPS = 5000.0  # dense int8 MFMA, MI355X
# The shared-A product timed alone: back-to-back launches after ~50 ms of the same
# launches.  Right after the host checks leave the GPU idle, a window of a few
# hundred launches reads up to ~10% slower than the ones that follow (68 us
# against 60-63 us at 4096 requests), so a short window right after warm-up
# measures the clocks' ramp as much as the kernel.
FRODO_GEMM_WARMUP, FRODO_GEMM_TIMED = 800, 200


def bench_frodo(args):
    print(json.dumps(run_frodo(args)), flush=True)


def run_frodo(args):
    import pqc_amd as pqc
    n = FRODO_BATCH if args.batch == BATCH else args.batch
    ctx = pqc.Context(["frodo976"], max_batch=n)
    stream = pqc.Stream()
    rng = np.random.default_rng(0)
    rnd = rng.integers(0, 256, (1, 64), dtype=np.uint8)
    pk, sk = ctx.frodo_keypair_host(rnd)
    mu = pqc.DeviceBuffer.from_array(rng.integers(0, 256, (n, 24), dtype=np.uint8))
    ct, ss_enc, ss = pqc.DeviceBuffer(n * 15744), pqc.DeviceBuffer(n * 24), pqc.DeviceBuffer(n * 24)
    ctx.frodo_load_key(pk[0], False, host=True)
    ctx.frodo_enc(n, ct, ss_enc, mu, stream=stream)
    ctx.frodo_load_key(sk[0], True, host=True)
    stream.synchronize()
    log("inputs ready: one FrodoKEM-976 key pair, %d ciphertexts in HBM" % n)
    calibrate_copy(args, pqc, stream)

    dec = lambda: ctx.frodo_dec(n, ss, ct, stream=stream)
    wall, per = timed(dec, stream, args.steps, args.warmup, pqc)
    ms = wall * 1e3 / args.steps
    log("frodo decaps: %.3f ms/step" % ms)
    ok = bool((ss.download((n, 24)) == ss_enc.download((n, 24))).all())
    pipe = frodo_pipelined(pqc, ctx, stream, pk[0], sk[0], n, args, rng)
    # the shared-A product alone (its decapsulation form, comparing against ct),
    # on the same stream and the same batch
    gemm = lambda: ctx.frodo_gemm(n, ct, stream=stream)
    gemm_ms = timed_run(gemm, stream, FRODO_GEMM_TIMED, FRODO_GEMM_WARMUP, pqc)
    rows = FRODO_NBAR * n
    int8_ops = 2 * 2 * rows * FRODO_N * FRODO_N          # hi and lo digit products, unpadded
    modq_ops = 2 * rows * FRODO_N * FRODO_N              # S'.A mod 2^16
    achieved = int8_ops / (gemm_ms / 1e3) / 1e12
    out = {
        "metric": "FrodoKEM-976 decaps/s at batch 4096 under one key (shared-A S'.A on int8 MFMA)",
        "value": round(n / (ms / 1e3), 1),
        "unit": "decaps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int8 MFMA (hi/lo digits of uint16), uint16 mod 2^16",
        "data": "synthetic: one seeded key pair and %d seeded encapsulations made on the GPU" % n,
        "config": {
            "workload": "frodo976_decaps_shared_a",
            "batch": n,
            "gemm": "(%dx976).(976x976) S'.A plus S'.B, mod 2^16" % rows,
            "batch_latency_ms_p50": round(float(np.percentile(per, 50)), 4),
            "batch_latency_ms_p99": round(float(np.percentile(per, 99)), 4),
            "decaps_equal_encaps": ok,
            "pipelined": pipe,
        },
        "roofline": {
            "bound": "mfma",
            "kernel": "k_gemm<decaps> (shared-A product + compare, v_mfma_i32_16x16x64_i8)",
            "achieved": round(achieved, 2),
            "peak": MFMA_I8_PEAK_TOPS,
            "unit": "TOPS (int8)",
            "frac": round(achieved / MFMA_I8_PEAK_TOPS, 4),
            "work_per_launch": "%d rows x 976 x 976 x 2 digits x 2 ops = %.2f G int8 ops "
                               "(%.2f G mod-2^16 ops)" % (rows, int8_ops / 1e9, modq_ops / 1e9),
            "kernel_ms": round(gemm_ms, 4),
            "kernel_timing": "%d back-to-back launches after %d warm-up launches" % (FRODO_GEMM_TIMED, FRODO_GEMM_WARMUP),
            "traffic": MEASURED_TRAFFIC_FRODO_GEMM_B4096 if n == FRODO_BATCH else None,
        },
        "cpu_baseline": None,
    }
    pipe["best_over_single_batch"] = round(pipe["best"]["decaps_per_s"] / out["value"], 3)
    if not args.no_cpu:
        from tests import oracle_lib as O
        ct_h = ct.download((n, 15744))
        ss_h = ss.download((n, 24))
        thr = cpu_threads(args)

        def run(m):
            ss_cpu = O.frodo_dec_batch(np.ascontiguousarray(ct_h[:m]), sk[0].tobytes(), nthreads=thr)
            assert (ss_cpu == ss_h[:m]).all(), "CPU oracle and GPU disagree"

        m, dt = timed_cpu(run, n, 2 * thr)
        out["cpu_baseline"] = cpu_entry(m / dt, "decaps/s", thr,
                                        "%d FrodoKEM-976 decaps of the same batch on %d threads, sharing A, timed for "
                                        "at least 1 s; outputs checked equal to the GPU's" % (m, thr))
        out["config"]["gpu_over_cpu"] = round(out["value"] / (m / dt), 1)
    ctx.close()
    return out


def frodo_pipelined(pqc, ctx, stream, pk, sk, n, args, rng):
    """Decapsulation of several consecutive n-ciphertext batches under the same key
    as one call, in two forms: one chain over the whole call (one k_shake launch
    over every request: 4 x 4096 requests put one stream wave on each SIMD), and
    chunks of `chunk` requests run on four streams (tuning.frodo_chunk).  Reports
    decaps/s of each against the single-batch rate, and checks that the first
    batch's shared secrets equal those of the same ciphertexts decapsulated as a
    single batch."""
    from pqc_amd import tuner as T
    rows = []
    for pn, chunk in ((4 * n, 0), (4 * n, n), (16 * n, 0), (16 * n, 4 * n)):
        big = pqc.Context(["frodo976"], max_batch=pn)
        try:
            if chunk:
                T.put_tuning(big, frodo_chunk=chunk)
            mu = pqc.DeviceBuffer.from_array(rng.integers(0, 256, (pn, 24), dtype=np.uint8))
            ct, ss_enc, ss = pqc.DeviceBuffer(pn * 15744), pqc.DeviceBuffer(pn * 24), pqc.DeviceBuffer(pn * 24)
            big.frodo_load_key(pk, False, host=True)
            big.frodo_enc(pn, ct, ss_enc, mu, stream=stream)
            big.frodo_load_key(sk, True, host=True)
            steps = max(3, args.steps // 2)
            wall, per = timed(lambda: big.frodo_dec(pn, ss, ct, stream=stream), stream, steps, 1, pqc,
                              warm_ms=SECONDARY_WARM_MS)
            ms = wall * 1e3 / steps
            ss_h = ss.download((pn, 24))
            ok = bool((ss_h == ss_enc.download((pn, 24))).all())
            one = pqc.DeviceBuffer(n * 24)
            ctx.frodo_dec(n, one, ct, stream=stream)  # the same first n ciphertexts as one batch
            stream.synchronize()
            ok = ok and bool((one.download((n, 24)) == ss_h[:n]).all())
        finally:
            big.close()
        log("frodo %d decaps per call, chunk %s: %.3f ms" % (pn, chunk or "none", ms))
        rows.append({"batch_per_call": pn, "chunk": chunk or None, "streams": min(4, pn // chunk) if chunk else 1,
                     "ms_per_call": round(ms, 4), "decaps_per_s": round(pn / (ms / 1e3), 1), "bitexact": ok})
    best = max(rows, key=lambda r: r["decaps_per_s"])
    return {"best": best, "runs": rows}


# ---- NewHope-1024 CPA: 16384 keygen + encaps + decaps triples ----------------------
NEWHOPE_BATCH = 16384


def bench_newhope(args):
    print(json.dumps(run_newhope(args)), flush=True)


def run_newhope(args):
    import pqc_amd as pqc
    n = NEWHOPE_BATCH if args.batch == BATCH else args.batch
    ctx = pqc.Context(["newhope1024cpa", "newhope1024", "newhope512cpa", "newhope512"], max_batch=n)
    stream = pqc.Stream()
    calibrate_copy(args, pqc, stream)
    rng = np.random.default_rng(0)
    seed, msg, coin = (pqc.DeviceBuffer.from_array(rng.integers(0, 256, (n, 32), dtype=np.uint8)) for _ in range(3))
    pk, sk, ct, m2 = pqc.DeviceBuffer(n * 1824), pqc.DeviceBuffer(n * 1792), pqc.DeviceBuffer(n * 2176), \
        pqc.DeviceBuffer(n * 32)

    def triple(plain=False):
        ctx.newhope_keypair(n, pk, sk, seed, stream=stream, plain=plain)
        ctx.newhope_enc(n, ct, msg, pk, coin, stream=stream, plain=plain)
        ctx.newhope_dec(n, m2, ct, sk, stream=stream, plain=plain)

    wall, per = timed(triple, stream, args.steps, args.warmup, pqc)
    ms = wall * 1e3 / args.steps
    log("newhope triples (fused): %.3f ms/step" % ms)
    ok = bool((m2.download((n, 32)) == msg.download((n, 32))).all())
    ct_fused = ct.download((n, 2176))
    pwall, _ = timed(lambda: triple(True), stream, args.steps, args.warmup, pqc)
    plain_ms = pwall * 1e3 / args.steps
    log("newhope triples (plain): %.3f ms/step" % plain_ms)
    same = bool((ct.download((n, 2176)) == ct_fused).all())
    # the three fused kernels, one launch each, timed alone on the same stream
    kms = {name: timed_run(fn, stream, 20, 2, pqc) for name, fn in (
        ("keypair", lambda: ctx.newhope_keypair(n, pk, sk, seed, stream=stream)),
        ("enc", lambda: ctx.newhope_enc(n, ct, msg, pk, coin, stream=stream)),
        ("dec", lambda: ctx.newhope_dec(n, m2, ct, sk, stream=stream)))}
    log("fused kernels: keypair %.1f us, enc %.1f us, dec %.1f us"
        % tuple(kms[k] * 1e3 for k in ("keypair", "enc", "dec")))
    cca = newhope_cca_rates(pqc, ctx, stream, n, rng, max(3, args.steps // 4))
    nh512 = newhope512_rates(pqc, ctx, stream, n, rng, max(3, args.steps // 4))

    # The per-operation path's largest kernels: poly_sample and the NTT, alone.
    polys = pqc.DeviceBuffer(n * 2048)
    L = pqc.lib()
    sample_ms = timed_run(lambda: L.pqc_newhope1024_poly_sample(n, polys.ptr, coin.ptr, 32, 0, stream.handle),
                          stream, 20, 2, pqc)
    ntt_ms = timed_run(lambda: L.pqc_newhope1024_poly_transform(0, n, polys.ptr, stream.handle), stream, 20, 2, pqc)
    ntt_bytes = n * 2048 * 2              # read + write one uint16 polynomial
    log("plain poly_sample %.1f us, poly_ntt %.1f us" % (sample_ms * 1e3, ntt_ms * 1e3))
    enc_ops = n * NEWHOPE_KECCAK_PER_ENC * OPS_PER_KECCAK
    enc_ms = kms["enc"]
    out = {
        "metric": "NewHope-1024 CPA keygen+encaps+decaps triples/s at batch 16384 (fused per-element kernels vs "
                  "plain batched kernels)",
        "value": round(n / (ms / 1e3), 1),
        "unit": "triples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "uint16",
        "data": "synthetic: seeded 32-byte seeds, messages and coins",
        "config": {"workload": "newhope1024_cpa_triples", "batch": n,
                   "batch_latency_ms_p50": round(float(np.percentile(per, 50)), 4),
                   "batch_latency_ms_p99": round(float(np.percentile(per, 99)), 4),
                   "messages_recovered": ok,
                   "plain_triples_per_s": round(n / (plain_ms / 1e3), 1),
                   "plain_ms_per_step": round(plain_ms, 4),
                   "fused_over_plain": round(plain_ms / ms, 3),
                   "fused_equals_plain": same,
                   "fused_kernel_us": {k: round(v * 1e3, 1) for k, v in kms.items()},
                   "cca_kem": cca,
                   "newhope512": nh512,
                   "plain_kernels": {"poly_sample_us": round(sample_ms * 1e3, 1), "poly_ntt_us": round(ntt_ms * 1e3, 1),
                                     "poly_ntt_hbm_frac": round(ntt_bytes / ntt_ms / 1e6 / HBM_PEAK_GBPS, 4)}},
        "roofline": {"bound": "valu", "kernel": "newhope::k_enc_fused<1024> (cpapke_enc, one wave per element)",
                     "achieved": round(enc_ops / enc_ms / 1e9, 2), "peak": VALU_PEAK_TOPS,
                     "unit": "T int32 ops/s", "frac": round(enc_ops / enc_ms / 1e9 / VALU_PEAK_TOPS, 4),
                     "work_per_launch": "%d encryptions x %d Keccak-f x 7440 two-input 32-bit ops (the sponge work "
                                        "only: NTT, pointwise and packing work is not counted, so frac is a lower "
                                        "bound)" % (n, NEWHOPE_KECCAK_PER_ENC),
                     "kernel_ms": round(enc_ms, 4),
                     "traffic": MEASURED_TRAFFIC_NEWHOPE_ENC_B16384 if n == NEWHOPE_BATCH else None,
                     "traffic_compulsory": n * (1824 + 32 + 32 + 2176)},
        "cpu_baseline": None,
    }
    if not args.no_cpu:
        from tests import oracle_lib as O
        thr = cpu_threads(args)
        s_h, m_h, c_h = seed.download((n, 32)), msg.download((n, 32)), coin.download((n, 32))
        ct_h = ct.download((n, 2176))

        def run(m):
            opk, osk, oct_, om2 = O.newhope_triples(s_h[:m], m_h[:m], c_h[:m], nthreads=thr)
            assert (oct_ == ct_h[:m]).all(), "CPU oracle and GPU disagree"

        m, dt = timed_cpu(run, n, 64 * thr)
        out["cpu_baseline"] = cpu_entry(m / dt, "triples/s", thr,
                                        "%d NewHope-1024 triples of the same batch on %d threads, timed for at least "
                                        "1 s; ciphertexts checked equal to the GPU's" % (m, thr))
        out["config"]["gpu_over_cpu"] = round(out["value"] / (m / dt), 1)
    ctx.close()
    return out


def newhope_cca_rates(pqc, ctx, stream, n, rng, steps, alg="newhope1024"):
    """NewHope CCA-KEM keygen, encaps and decaps over the same batch size
    (the specification's transform over the CPA kernels above)."""
    pkb, skb, ctb, _ = pqc.sizes(alg)
    kc = pqc.DeviceBuffer.from_array(rng.integers(0, 256, (n, 64), dtype=np.uint8))
    ec = pqc.DeviceBuffer.from_array(rng.integers(0, 256, (n, 32), dtype=np.uint8))
    pk, sk, ct = pqc.DeviceBuffer(n * pkb), pqc.DeviceBuffer(n * skb), pqc.DeviceBuffer(n * ctb)
    ss, ss2 = pqc.DeviceBuffer(n * 32), pqc.DeviceBuffer(n * 32)
    res = {}
    for name, fn in (("keypair", lambda: ctx.keypair_derand(alg, n, pk, sk, kc, stream=stream)),
                     ("encaps", lambda: ctx.enc_derand(alg, n, ct, ss, pk, ec, stream=stream)),
                     ("decaps", lambda: ctx.dec(alg, n, ss2, ct, sk, stream=stream))):
        wall, _ = timed(fn, stream, steps, 1, pqc, warm_ms=SECONDARY_WARM_MS)
        res[name + "_per_s"] = round(n / (wall / steps), 1)
        log("%s cca %s: %.3f ms" % (alg, name, wall * 1e3 / steps))
    res["shared_secrets_equal"] = bool((ss.download((n, 32)) == ss2.download((n, 32))).all())
    res["sizes"] = "pk %d, sk %d, ct %d, ss 32" % (pkb, skb, ctb)
    return res


def newhope512_rates(pqc, ctx, stream, n, rng, steps):
    """NewHope-512: CPA triples through the N = 512 kernels and the CCA-KEM."""
    seed, msg, coin = (pqc.DeviceBuffer.from_array(rng.integers(0, 256, (n, 32), dtype=np.uint8)) for _ in range(3))
    pk, sk, ct, m2 = (pqc.DeviceBuffer(n * b) for b in (928, 896, 1088, 32))

    def triple():
        ctx.newhope_keypair(n, pk, sk, seed, stream=stream, N=512)
        ctx.newhope_enc(n, ct, msg, pk, coin, stream=stream, N=512)
        ctx.newhope_dec(n, m2, ct, sk, stream=stream, N=512)

    wall, _ = timed(triple, stream, steps, 1, pqc, warm_ms=SECONDARY_WARM_MS)
    log("newhope512 triples: %.3f ms" % (wall * 1e3 / steps))
    return {"cpa_triples_per_s": round(n / (wall / steps), 1),
            "messages_recovered": bool((m2.download((n, 32)) == msg.download((n, 32))).all()),
            "cca_kem": newhope_cca_rates(pqc, ctx, stream, n, rng, steps, "newhope512")}


# ---- mixed Kyber-512/768/1024 (1:2:1), 65536 encaps + decaps, fused in a graph ---------
# Keccak-f per decapsulation: G 1 + A^T 4 K^2 + PRF 2K+1 + H(ct) ceil((ct+1)/136) + KDF 1
MIXED_KECCAK_PER_DECAPS = {1: 1 + 16 + 5 + 6 + 1, 2: 1 + 36 + 7 + 9 + 1, 3: 1 + 64 + 9 + 12 + 1}
def bench_mixed(args):
    print(json.dumps(run_mixed(args)), flush=True)


def run_mixed(args):
    import pqc_amd as pqc
    from pqc_amd import tuner as T
    n = args.batch
    ctx = pqc.Context(["kyber512", "kyber768", "kyber1024"], max_batch=n)
    stream = pqc.Stream()
    base = np.array([1, 2, 2, 3] * ((n + 3) // 4), dtype=np.int32)[:n]
    algs = np.random.default_rng(0).permutation(base)
    lay = pqc.MixedLayout(algs, np.random.default_rng(1))
    ar = pqc.DeviceArenas(lay)
    rng = np.random.default_rng(2)
    ar.buf["kcoins"].upload(rng.integers(0, 256, (lay.nbytes["kcoins"],), dtype=np.uint8))
    ar.buf["coins"].upload(rng.integers(0, 256, (lay.nbytes["coins"],), dtype=np.uint8))
    for a in (1, 2, 3):
        ctx.keypair_derand(a, lay.count[a], ar.region_ptr("pk", a), ar.region_ptr("sk", a),
                           ar.region_ptr("kcoins", a), stream=stream)
    ss_dec = pqc.DeviceBuffer(lay.nbytes["ss"])
    stream.synchronize()
    log("inputs ready: %d requests (Kyber-512/768/1024 = %d/%d/%d), one key pair each" %
        (n, lay.count[1], lay.count[2], lay.count[3]))

    def enc_dec_n(m, s):
        ctx.kem_enc_mixed(m, ar.enc_reqs, ar.buf["ct"], ar.buf["pk"], ar.buf["ss"], ar.buf["coins"], stream=s)
        ctx.kem_dec_mixed(m, ar.dec_reqs, ar.buf["ct"], ar.buf["sk"], ss_dec, stream=s)

    def enc_dec(s):
        enc_dec_n(n, s)

    tuned, autotune = learned_autotune(ctx, T, n, enc_dec_n, stream)
    log("tuned (model-guided): %s %.3f ms; %s" % (tuned, autotune["model_pick_ms"], autotune))
    graph = pqc.Graph.capture(stream, lambda: enc_dec(stream))
    wall, per = timed(lambda: graph.launch(stream), stream, args.steps, args.warmup, pqc)
    ms = wall * 1e3 / args.steps
    ss_e, ss_d = ar.download("ss"), ss_dec.download((lay.nbytes["ss"],))
    ct_f = ar.download("ct")
    ok_fused = bool((ss_e == ss_d).all())

    def plain_step():
        for a in (1, 2, 3):
            ctx.enc_derand(a, lay.count[a], ar.region_ptr("ct", a), ar.region_ptr("ss", a), ar.region_ptr("pk", a),
                           ar.region_ptr("coins", a), plain=True, stream=stream)
        for a in (1, 2, 3):
            ctx.dec(a, lay.count[a], ss_region(a), ar.region_ptr("ct", a), ar.region_ptr("sk", a), plain=True,
                    stream=stream)

    def ss_region(a):
        from pqc_amd.mixed import _Slice
        off, cnt = lay.region["ss"][a]
        return _Slice(ss_dec, off, cnt * 32)

    # the dominant kernel, k_fused_mixed_dec32, timed alone on the stream (with the
    # descriptor classification it needs: two small kernels of ~40 us together)
    dec_ms = timed_run(lambda: ctx.kem_dec_mixed(n, ar.dec_reqs, ar.buf["ct"], ar.buf["sk"], ss_dec, stream=stream),
                       stream, 10, 2, pqc)
    dec_ops = sum(lay.count[a] * MIXED_KECCAK_PER_DECAPS[a] for a in (1, 2, 3)) * OPS_PER_KECCAK
    log("mixed decaps alone: %.3f ms" % dec_ms)

    psteps = max(3, args.steps // 4)
    wall_p, _ = timed(plain_step, stream, psteps, 1, pqc, warm_ms=SECONDARY_WARM_MS)
    ms_p = wall_p * 1e3 / psteps
    bitexact = bool(ok_fused and (ar.download("ct") == ct_f).all() and (ar.download("ss") == ss_e).all() and
                    (ss_dec.download((lay.nbytes["ss"],)) == ss_d).all())
    graph.destroy()
    out = {
        "metric": "mixed Kyber-512/768/1024 (1:2:1) encaps+decaps requests/s at batch 64K, fused graph vs plain",
        "value": round(n / (ms / 1e3), 1),
        "unit": "requests/s (encaps + decaps each)",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int16",
        "data": "synthetic: seeded key pairs (one per request) and coins, made on the GPU",
        "config": {"workload": "kyber_mixed_enc_dec_graph", "batch": n,
                   "counts_512_768_1024": [lay.count[1], lay.count[2], lay.count[3]],
                   "tuning": tuned,
                   "autotune": autotune,
                   "plain_ms_per_step": round(ms_p, 4),
                   "fused_over_plain": round(ms_p / ms, 3),
                   "batch_latency_ms_p50": round(float(np.percentile(per, 50)), 4),
                   "batch_latency_ms_p99": round(float(np.percentile(per, 99)), 4),
                   "bitexact_fused_plain": bitexact},
        "roofline": {"bound": "valu", "kernel": "k_fused_mixed_dec32 (mixed decaps, 32-request tiles of all three parameter sets, two waves per SIMD)",
                     "achieved": round(dec_ops / dec_ms / 1e9, 3), "peak": VALU_PEAK_TOPS, "unit": "T int32 ops/s",
                     "frac": round(dec_ops / dec_ms / 1e9 / VALU_PEAK_TOPS, 4),
                     "work_per_launch": "Keccak-f per decaps 29 / 54 / 87 (Kyber-512 / 768 / 1024) x 7440 ops, "
                                        "%d / %d / %d requests" % (lay.count[1], lay.count[2], lay.count[3]),
                     "kernel_ms": round(dec_ms, 4), "traffic": None},
        "cpu_baseline": None,
    }
    if not args.no_cpu:
        # the oracle on the host cores: encaps + decaps of a sample of each parameter set
        from tests import oracle_lib as O
        thr = cpu_threads(args)
        pk_h, sk_h, co_h = ar.download("pk"), ar.download("sk"), ar.download("coins")
        ct_h, ssd_h = ar.download("ct"), ss_dec.download((lay.nbytes["ss"],))
        views = {a: [np.ascontiguousarray(lay.view(x, kd, a)) for x, kd in
                     ((pk_h, "pk"), (co_h, "coins"), (sk_h, "sk"), (ct_h, "ct"), (ssd_h, "ss"))] for a in (1, 2, 3)}

        def run(m):
            # m requests in the batch's 1:2:1 proportions: m/4 Kyber-512, m/2 Kyber-768, m/4 Kyber-1024
            for a, share in ((1, 4), (2, 2), (3, 4)):
                q = max(1, m // share)
                pk_a, co_a, sk_a, ct_a, ss_a = (v[:q] for v in views[a])
                oct_, oss = O.kyber_batch("enc", a + 1, [pk_a, co_a], nthreads=thr)
                oss2 = O.kyber_batch("dec", a + 1, [ct_a, sk_a], nthreads=thr)
                assert (oct_ == ct_a).all() and (oss2 == ss_a).all(), "CPU oracle and GPU disagree"

        done, dt = timed_cpu(run, min(lay.count[1], lay.count[3]) * 4, 64 * thr)
        out["cpu_baseline"] = cpu_entry(done / dt, "requests/s (encaps + decaps each)", thr,
                                        "%d requests in the batch's 1:2:1 proportions encapsulated and decapsulated "
                                        "on %d threads, timed for at least 1 s; outputs checked equal to the GPU's"
                                        % (done, thr))
        out["config"]["gpu_over_cpu"] = round(out["value"] / (done / dt), 1)
    ctx.close()
    return out


def learned_autotune(ctx, T, n, run_n, stream, budget=4):
    """Grid waves x tile order for the mixed batch of n, chosen by the learned cost
    model.  The model (trained on the GPU by the library's own kernels) learns from
    exhaustive sweeps of the first n/4, n/2 and 3n/4 requests of the same batch;
    then pqc_autotune ranks the 16 candidates for n with it and measures only the
    best-ranked `budget`, installing the fastest of those.  An exhaustive sweep at
    n follows as the reference: the pick's regret is its time in that sweep over
    the sweep's best.  The model's pick is what the timed graph runs."""
    train = [n // 4, n // 2, 3 * n // 4]
    default = T.get_tuning(ctx).as_dict()
    # medians of several runs: each run times eager launches from Python, and one
    # slow launch in two runs made a worse configuration look better
    for m in train:
        T.autotune(ctx, T.request("mixed_kyber", m, k=3, op=2, budget=0, reps=3), lambda s, m=m: run_n(m, s), stream)
    x, y = T.samples(ctx)
    model = T.CostModel.of_context(ctx)
    loss = model.fit(x, y, steps=3000, lr=3e-3)
    T.put_tuning(ctx, **default)
    req = T.request("mixed_kyber", n, k=3, op=2, budget=budget, use_model=True, reps=5)
    guided = T.autotune(ctx, req, lambda s: run_n(n, s), stream)
    pick = guided.best.as_dict()
    before = len(T.samples(ctx)[1])
    sweep = T.autotune(ctx, T.request("mixed_kyber", n, k=3, op=2, budget=0, reps=5), lambda s: run_n(n, s), stream)
    sx, sy = T.samples(ctx)
    sx, sy = sx[before:], np.exp(sy[before:])
    def in_sweep(tuning_dict):
        t = T.Tuning()
        for k, v in tuning_dict.items():
            setattr(t, k, v)
        f = T.features(req, t)
        row = [i for i in range(len(sy)) if np.allclose(sx[i], f)]
        return float(sy[row[0]]) if row else None

    pick_in_sweep = in_sweep(pick) or float(guided.best_ms)
    T.put_tuning(ctx, **pick)
    return pick, {"train_batches": train, "train_samples": int(len(y)), "model_loss": round(float(loss), 5),
                  "model_trials": int(guided.measured), "sweep_trials": int(sweep.measured),
                  "trials_saved": int(sweep.measured - guided.measured),
                  "model_pick": pick, "model_pick_ms": round(float(guided.best_ms), 4),
                  "model_pick_ms_in_sweep": round(pick_in_sweep, 4),
                  "sweep_best": sweep.best.as_dict(), "sweep_best_ms": round(float(sweep.best_ms), 4),
                  "regret": round(pick_in_sweep / float(sweep.best_ms), 4),
                  "default": default, "default_ms_in_sweep": in_sweep(default)}


# ---- serving loop: Poisson arrivals, mixed Kyber / Frodo, 1024 keys per set ------------
def serving_trace(n, rate, seed, keys):
    """Arrival times (us), algorithm ids (90 % Kyber at 512:768:1024 = 1:2:1,
    10 % FrodoKEM-976), operations (encaps / decaps 50 / 50, an assumption), keys."""
    rng = np.random.default_rng(seed)
    arrivals = np.cumsum(rng.exponential(1e6 / rate, size=n))
    kyber = rng.choice(np.array([1, 2, 2, 3], np.int32), size=n)
    algs = np.where(rng.random(n) < 0.9, kyber, 5).astype(np.int32)
    ops = rng.choice(np.array([1, 2], np.int32), size=n)
    kid = rng.integers(0, keys, size=n).astype(np.uint32)
    return arrivals, algs, ops, kid, rng


def bench_serving(args):
    print(json.dumps(run_serving(args)), flush=True)


def run_serving(args, duration=None, with_plain=True):
    from pqc_amd.server import RequestSet, Server
    n = int(args.rate * (duration or args.duration))
    keys = 1024
    arrivals, algs, ops, kid, rng = serving_trace(n, args.rate, 0, keys)
    algs_names = ["kyber512", "kyber768", "kyber1024", "frodo976"]
    results = {}
    slots = args.slots
    for plain, graph in ((False, False), (False, True)) + (((True, False),) if with_plain else ()):
        # the plain path cannot hold the offered rate: replay a prefix of the trace
        m = n if not plain else min(n, int(0.2 * n))
        t0 = time.perf_counter()
        srv = Server(algs_names, keys_per_alg=keys, key_seed=0, max_batch=16384, slots=slots,
                     arena_bytes=96 << 20, max_wait_us=200, plain=plain, graph=graph)
        log("server (plain=%s) up in %.1f s: %d key pairs per set" % (plain, time.perf_counter() - t0, keys))
        # untimed warm-up: a short trace at the same rate (first launches, pinned pages)
        wm = min(m, 50000)
        ws = RequestSet.random(algs[:wm], ops[:wm], kid[:wm], rng)
        srv.run_trace(arrivals[:wm], ws.reqs)
        rs = RequestSet.random(algs[:m], ops[:m], kid[:m], np.random.default_rng(1234))  # same inputs every mode
        st, lat = srv.run_trace(arrivals[:m], rs.reqs)
        results[(plain, graph)] = (st, m, rs)
        log("plain=%s graph=%s: %d requests, %.0f ops/s, p50 %.0f us, p99 %.0f us, mean batch %.0f, errors %d" %
            (plain, graph, m, st.ops_per_s, st.p50_us, st.p99_us, st.mean_batch, st.errors))
        if not plain and not graph and not args.no_cpu:
            cpu = serving_cpu_baseline(srv, rs, algs, ops, kid, cpu_threads(args))
        srv.close()
    st, m, rs_e = results[(False, False)]
    sg, mg, rs_g = results[(False, True)]
    sp, mp, _ = results.get((True, False), (None, 0, None))
    graph_same = bool((rs_e.outbuf == rs_g.outbuf).all())
    out = {
        "metric": "KEM serving loop at Poisson 2M req/s (90% Kyber 1:2:1 / 10% Frodo-976, 1024 keys per set): "
                  "ops/s and p50/p99 latency, fused vs plain",
        "value": round(sg.ops_per_s, 1),
        "unit": "requests/s",
        "n_gpus": 1,
        "steps": 1,
        "warmup": 0,
        "ms_per_step": round(sg.duration_s * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int16 / uint16",
        "data": "synthetic: seeded Poisson trace, random coins and ciphertexts, seeded key pairs",
        # the primary figures are graph mode (PQC_SERVER_GRAPH): with Kyber and Frodo
        # in one server the slots form two pools, and the Kyber batches replay as
        # graphs while the Frodo batches stay eager; "eager" is every batch eager
        "config": {"workload": "serving_poisson", "offered_rate": args.rate, "requests": m, "slots": slots,
                   "mode": "graph (Kyber pool captured, Frodo pool eager)",
                   "hw_queues": int(os.environ.get("GPU_MAX_HW_QUEUES", "4")),
                   "p50_us": round(sg.p50_us, 1), "p90_us": round(sg.p90_us, 1), "p99_us": round(sg.p99_us, 1),
                   "p999_us": round(sg.p999_us, 1), "mean_batch": round(sg.mean_batch, 1), "errors": int(sg.errors),
                   "eager": {"ops_per_s": round(st.ops_per_s, 1), "p50_us": round(st.p50_us, 1),
                             "p99_us": round(st.p99_us, 1), "mean_batch": round(st.mean_batch, 1),
                             "errors": int(st.errors), "bytes_equal_graph": graph_same},
                   "plain": None if sp is None else {
                       "requests": mp, "ops_per_s": round(sp.ops_per_s, 1), "p50_us": round(sp.p50_us, 1),
                       "p99_us": round(sp.p99_us, 1), "mean_batch": round(sp.mean_batch, 1),
                       "errors": int(sp.errors)}},
        "roofline": None,
        "cpu_baseline": None if args.no_cpu else cpu,
    }
    return out


def serving_cpu_baseline(srv, rs, algs, ops, kid, thr, min_s=1.0):
    """The oracle on the host cores over the first requests of the same trace
    (same mix, same keys and inputs), one request per task on a thread pool,
    enough of them for at least `min_s` seconds; every output is checked equal to
    what the server returned."""
    from concurrent.futures import ThreadPoolExecutor
    from tests import oracle_lib as O
    m = min(rs.n, 50000)  # enough for well over a second on the host, without copying the whole trace
    keys = {}
    for i in range(m):
        keys.setdefault((int(algs[i]), int(kid[i])), None)
    for a, k in keys:
        keys[(a, k)] = srv.key(a, k)
    ins = [rs.inbuf[rs.in_off[i]:rs.in_off[i + 1]].tobytes() for i in range(m)]

    def one(i):
        a, op = int(algs[i]), int(ops[i])
        pk, sk = keys[(a, int(kid[i]))]
        if a == 5:
            return b"".join(O.frodo_enc(pk, ins[i])) if op == 1 else O.frodo_dec(ins[i], sk)
        return b"".join(O.kyber_enc(a + 1, pk, ins[i])) if op == 1 else O.kyber_dec(a + 1, ins[i], sk)

    with ThreadPoolExecutor(thr) as ex:
        def run(k):
            got = list(ex.map(one, range(k)))
            assert all(got[i] == rs.out(i) for i in range(k)), "CPU oracle and server disagree"

        run(min(m, 2 * thr))  # warm
        done, dt = timed_cpu(run, m, min(m, 64 * thr), min_s)
    return cpu_entry(done / dt, "requests/s", thr,
                     "the first requests of the trace (same mix, keys and inputs), %d in all, one request per task on "
                     "%d threads, no arrival pacing, timed for at least %.0f s; outputs checked equal to the server's"
                     % (done, thr, min_s))


if __name__ == "__main__":
    main()
