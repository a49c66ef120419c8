"""Benchmark: 30-qubit FP64 random circuit on one MI355X.

One step = reset to |0...0> and execute the whole 40-layer circuit (1800
gates: per layer one gate per qubit from {H,X,Y,Z,RX,RY,RZ,P,U} and 15 CX/CZ
on a random perfect matching) on a 16 GiB state resident in HBM.

Prints ONE JSON line:
  value      gates/s over the timed steps (whole job, one GPU)
  roofline   the dominant kernel class: algorithmic HBM bytes / device time,
             measured with HIP events on the library stream over the timed
             region, against the 8 TB/s HBM peak
  adjoint    24-qubit, 480-parameter adjoint gradient: grads/s and the ratio
             against parameter shift on the GPU
  fusion     60-layer random circuits at 32 q FP64 and 34 q FP32 through the
             tile executor and the dense 5-qubit MFMA blocks, with fidelity,
             norm and <Z0 Z_n-1> errors against an unfused FP64 run at 28 q
  batch      1024 independent small-circuit adjoint gradients (n in 10..20,
             depth 2..12): p50/p99 latency and grads/s against a loop of
             single-circuit calls
  qft        20-qubit QFT of a Gaussian state plus 4096 samples
  per_stride plain-kernel GB/s per target qubit for the gate classes
  cpu_baseline  the reference's per-kind rayon loop shapes restated in C
             (OpenMP, vectorised; oracle/spz_oracle.c:orc_rayon_*) timed on
             this host on a bounded sample of the workload, at the host's
             thread count and on 1 thread;
             the same leg uses the oracle as the checker of the GPU results
             (26-qubit replica, QFT amplitudes and samples, adjoint gradient
             against parameter shift)
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12
MFMA_PEAK_F64 = 78.6e12   # dense FP64 matrix peak
MFMA_PEAK_F32 = 157.3e12  # dense FP32 matrix peak


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=3)
    p.add_argument("--warmup", type=int, default=1)
    p.add_argument("--qubits", type=int, default=30)
    p.add_argument("--layers", type=int, default=40)
    p.add_argument("--seed", type=int, default=2)
    p.add_argument("--mode", choices=["auto", "plain", "strict"], default="auto")
    p.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    p.add_argument("--no-replica", action="store_true", help="skip the 26q parity replica")
    p.add_argument("--no-strides", action="store_true", help="skip the per-stride sweep")
    p.add_argument("--cpu-threads", type=int, default=0)
    p.add_argument("--cpu-budget-1t", type=float, default=8.0,
                   help="seconds of the 1-thread CPU baseline sample")
    p.add_argument("--replica-layers", type=int, default=4)
    p.add_argument("--no-adjoint", action="store_true", help="skip the 24q adjoint leg")
    p.add_argument("--adjoint-qubits", type=int, default=24)
    p.add_argument("--adjoint-layers", type=int, default=10)
    p.add_argument("--calibrate", action="store_true",
                   help="first apply one plain X on the top qubit: a pass that reads and "
                        "writes every byte of the state once (counter calibration)")
    p.add_argument("--ps-subset", type=int, default=24,
                   help="parameters timed with GPU parameter shift (extrapolated to all)")
    p.add_argument("--cpu-ps-subset", type=int, default=4,
                   help="24q adjoint gradients checked against oracle parameter shift")
    p.add_argument("--no-fusion", action="store_true", help="skip the 32q/34q fusion leg")
    p.add_argument("--fusion-layers", type=int, default=60)
    p.add_argument("--fusion-qubits", type=str, default="32,34",
                   help="FP64 and FP32 qubit counts of the fusion leg")
    p.add_argument("--precision-qubits", type=int, default=28)
    p.add_argument("--dense-layers", type=int, default=6,
                   help="layers timed through the dense 5-qubit blocks at 32q/34q")
    p.add_argument("--no-batch", action="store_true", help="skip the batched adjoint leg")
    p.add_argument("--batch-count", type=int, default=1024)
    p.add_argument("--batch-group-qubits", type=int, default=0,
                   help="circuits above this many qubits run on groups of workgroups "
                        "(0: the library default, < 0: never)")
    p.add_argument("--batch-wg-max", type=int, default=0,
                   help="persistent-kernel switch point (0: autotuned)")
    p.add_argument("--batch-wg-per-cu", type=int, default=0,
                   help="resident persistent-kernel workgroups per CU (0: the library "
                        "default, 2; < 0: as many as fit)")
    p.add_argument("--no-qft", action="store_true", help="skip the 20q QFT leg")
    return p.parse_args()


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def per_stride(spz, N, n):
    """GB/s of the plain kernels per target qubit (one launch each, event-timed)."""
    from spinoza_amd.workloads import gate_bytes
    s = spz.State(n)
    classes = {
        "H": lambda t: spz.apply(spz.Gate.H, s, t),
        "RZ": lambda t: spz.apply(spz.Gate.RZ(0.3), s, t),
        "P": lambda t: spz.apply(spz.Gate.P(0.3), s, t),
        "CX(c=t+1 mod n)": lambda t: spz.c_apply(spz.Gate.X, s, (t + 1) % n, t),
    }
    byts = {
        "H": gate_bytes({"kind": N.H}, n),
        "RZ": gate_bytes({"kind": N.RZ}, n),
        "P": gate_bytes({"kind": N.P}, n),
        "CX(c=t+1 mod n)": gate_bytes({"kind": N.X, "ones": 1}, n),
    }
    out = {}
    N.check(N.lib().spz_profile_enable(1))
    for name, fn in classes.items():
        row = []
        for t in range(n):
            fn(t)  # warm
            N.check(N.lib().spz_sync())
            N.check(N.lib().spz_profile_reset())
            reps = 3
            for _ in range(reps):
                fn(t)
            tot = 0.0
            import ctypes
            for cls in (N.K_PAIR, N.K_DIAG):
                ms = ctypes.c_double()
                N.check(N.lib().spz_profile_get(cls, None, ctypes.byref(ms), None))
                tot += ms.value
            row.append(round(byts[name] * reps / (tot * 1e-3) / 1e9, 1) if tot > 0 else None)
        out[name] = row
    N.check(N.lib().spz_profile_enable(0))
    del s
    return out


def adjoint_leg(spz, N, n, layers, ps_subset, seed=4):
    """24-qubit hardware-efficient ansatz (RY, RZ on every qubit + CX ladder,
    10 layers, 480 parameters), Ising energy with hx = 1: adjoint gradient
    timed over a few repetitions, GPU parameter shift timed on a subset of
    parameters and extrapolated, max |adjoint - parameter shift| on that
    subset."""
    import numpy as np
    from spinoza_amd.autograd import adjoint_gradient, param_shift_gradient
    from spinoza_amd.workloads import hea
    ops, thetas, where = hea(n, layers, seed=seed)
    s = spz.State(n)
    adjoint_gradient(s, ops, 1.0)  # warm
    N.check(N.lib().spz_sync())
    reps = 3
    t0 = time.perf_counter()
    for _ in range(reps):
        s.set_basis(0)
        e, g, drift = adjoint_gradient(s, ops, 1.0, ckpt_every=len(ops) // 5)
    N.check(N.lib().spz_sync())
    t_adj = (time.perf_counter() - t0) / reps
    sub = list(range(0, len(where), max(1, len(where) // ps_subset)))[:ps_subset]
    param_shift_gradient(n, ops, where, 1.0, subset=sub[:1])  # warm
    t0 = time.perf_counter()
    ps = param_shift_gradient(n, ops, where, 1.0, subset=sub)
    N.check(N.lib().spz_sync())
    t_ps_per_param = (time.perf_counter() - t0) / len(sub)
    err = max(abs(g[where[k]] - v) for k, v in ps.items())
    del s
    return {"qubits": n, "layers": layers, "params": len(where), "ops": len(ops),
            "energy": e, "adjoint_s": round(t_adj, 4),
            "grads_per_s": round(len(where) / t_adj, 1),
            "param_shift_s_per_param": round(t_ps_per_param, 5),
            "param_shift_s_full_extrapolated": round(t_ps_per_param * len(where), 3),
            "speedup_vs_param_shift": round(t_ps_per_param * len(where) / t_adj, 1),
            "param_shift_subset": len(sub), "max_abs_grad_err_vs_param_shift": err,
            "checkpoint_drift": drift}, (n, ops, where, g)


def timed(fn, sync):
    sync()
    t0 = time.perf_counter()
    out = fn()
    sync()
    return time.perf_counter() - t0, out


def fusion_leg(spz, N, layers, n64, n32, n_prec, dense_layers, seed=3):
    """Random circuits (the 30q generator) with `layers` layers.

    Precision at n_prec qubits: FP64 and FP32 storage through the tile
    executor and the dense 5-qubit blocks against the unfused FP64 plain run:
    fidelity |<psi|psi_ref>|^2, |norm - norm_ref| and |<Z0 Z_n-1> - ref|.
    Throughput at n64 qubits FP64 and n32 qubits FP32 for both executors, with
    the dense-block kernel's HBM and MFMA fractions; the dense-block rate is
    taken on the first `dense_layers` layers of the same circuit (one block
    pass per 5-qubit run makes the whole circuit take about a minute)."""
    import ctypes
    from spinoza_amd.executor import encode_ops, run_ops
    from spinoza_amd.workloads import random_circuit
    sync = lambda: N.check(N.lib().spz_sync())
    modes = {"tile": N.EXEC_AUTO, "dense5": N.EXEC_DENSE5}
    out = {"layers": layers, "precision": {"qubits": n_prec}}
    ops = random_circuit(n_prec, layers, seed)
    ref = spz.State(n_prec, "f64")
    run_ops(ref, ops, mode=N.EXEC_PLAIN)
    ref_norm = ref.norm()
    ref_zz = spz.expect_zz(ref, 0, n_prec - 1)
    for prec in ("f64", "f32"):
        for name, mode in modes.items():
            s = spz.State(n_prec, prec)
            run_ops(s, ops, mode=mode)
            nrm = s.norm()
            # state fidelity of the normalised states; the norm error is reported apart
            fid = abs(spz.inner(s, ref)) ** 2 / (nrm * ref_norm)
            out["precision"][f"{name}_{prec}"] = {
                "fidelity": fid, "infidelity": 1.0 - fid,
                "norm_err": abs(nrm - ref_norm),
                "zz_err": abs(spz.expect_zz(s, 0, n_prec - 1) - ref_zz)}
            del s
    del ref
    for n, prec in ((n64, "f64"), (n32, "f32")):
        full = random_circuit(n, layers, seed)
        s = spz.State(n, prec)
        row = {"gates": len(full), "state_bytes": (16 if prec == "f64" else 8) << n}
        for name, mode in modes.items():
            ops = full if name == "tile" else random_circuit(n, min(layers, dense_layers), seed)
            enc = encode_ops(ops)
            s.set_basis(0)
            N.check(N.lib().spz_profile_enable(1))
            N.check(N.lib().spz_profile_reset())
            dt, _ = timed(lambda: run_ops(s, ops, mode=mode, encoded=enc), sync)
            cnt, ms, by = ctypes.c_uint64(), ctypes.c_double(), ctypes.c_double()
            cls = N.K_TILE if name == "tile" else N.K_DENSE
            N.check(N.lib().spz_profile_get(cls, ctypes.byref(cnt), ctypes.byref(ms),
                                            ctypes.byref(by)))
            N.check(N.lib().spz_profile_enable(0))
            r = {"layers": len(ops) // (n + n // 2), "gates": len(ops), "s": round(dt, 3),
                 "gates_per_s": round(len(ops) / dt, 1),
                 "launches": cnt.value, "kernel_ms": round(ms.value, 1),
                 "hbm_frac": round(by.value / (ms.value * 1e-3) / HBM_PEAK, 4) if ms.value else None,
                 "norm_err": abs(s.norm() - 1.0),
                 "zz": spz.expect_zz(s, 0, n - 1)}
            if name == "dense5" and ms.value:
                # 32x32 complex block: 8 * 32 real flops per amplitude
                flops = cnt.value * 256.0 * (1 << n)
                peak = MFMA_PEAK_F64 if prec == "f64" else MFMA_PEAK_F32
                r["mfma_tflops"] = round(flops / (ms.value * 1e-3) / 1e12, 2)
                r["mfma_frac"] = round(flops / (ms.value * 1e-3) / peak, 4)
            row[name] = r
        out[f"{n}q_{prec}"] = row
        del s
    return out


def batch_leg(spz, N, count, seed=5, group_qubits=0, wg_max=0, wg_per_cu=0):
    """`count` hardware-efficient-ansatz circuits, n ~ U{10..20}, depth ~
    U{2..12} (SplitMix64 seed 5), Ising energy (hx = 1): one batched call
    after autotuning the per-workgroup switch point, against looping the
    single-circuit adjoint."""
    import numpy as np
    from spinoza_amd.autograd import adjoint_gradient, batch_adjoint, batch_autotune
    from spinoza_amd.rng import SplitMix64
    from spinoza_amd.workloads import hea
    sync = lambda: N.check(N.lib().spz_sync())
    rng = SplitMix64(seed)
    circuits, nparams = [], 0
    for i in range(count):
        n = 10 + rng.next_u64() % 11
        depth = 2 + rng.next_u64() % 11
        ops, _, where = hea(n, depth, seed=seed * 100003 + i)
        circuits.append((n, ops))
        nparams += len(where)
    wg = wg_max or batch_autotune(circuits)
    run = lambda: batch_adjoint(circuits, 1.0, wg_max_qubits=wg, group_qubits=group_qubits,
                                wg_per_cu=wg_per_cu)
    run()  # warm (graph capture, buffers)
    dt, (energies, grads, st) = timed(run, sync)
    states = {}
    for n, _ in circuits:
        if n not in states:
            states[n] = spz.State(n)
    adjoint_gradient(states[circuits[0][0]], circuits[0][1], 1.0)  # warm
    err = 0.0

    def loop():
        res = []
        for n, ops in circuits:
            states[n].set_basis(0)
            res.append(adjoint_gradient(states[n], ops, 1.0))
        return res

    dt_loop, res = timed(loop, sync)
    for (e1, g1, _), e, g in zip(res, energies, grads):
        err = max(err, abs(e - e1), float(np.max(np.abs(g - g1))))
    del states
    return {"circuits": count, "params": nparams, "wg_max_qubits": wg,
            "wg_per_cu": wg_per_cu if wg_per_cu else 2,  # (0: the library default)
            "n_small": st["n_small"], "n_large": st["n_large"], "n_group": st["n_group"],
            "total_ms": round(dt * 1e3, 2), "p50_ms": round(st["p50_ms"], 4),
            "p99_ms": round(st["p99_ms"], 4), "grads_per_s": round(nparams / dt, 1),
            "loop_single_ms": round(dt_loop * 1e3, 2),
            "loop_grads_per_s": round(nparams / dt_loop, 1),
            "speedup_vs_loop": round(dt_loop / dt, 2), "max_abs_diff_vs_single": err}


def qft_leg(spz, N, n=20, seed=1):
    """QFT (H + controlled phases, MSB first, then the swap layer) of a
    normalised Gaussian state (SplitMix64 seed 1, Box-Muller) and 4096
    inverse-CDF samples from SplitMix64 seed 2 uniforms.  Returns the GPU
    timings and the device state for the checker."""
    from spinoza_amd.executor import encode_ops, run_ops
    from spinoza_amd.rng import gaussian_state, uniforms
    from spinoza_amd.workloads import qft
    sync = lambda: N.check(N.lib().spz_sync())
    ops = qft(n)
    enc = encode_ops(ops)
    re, im = gaussian_state(seed, n)
    src = spz.State.from_arrays(re, im)
    s = spz.State(n)
    u = uniforms(seed + 1, 4096)
    s.copy_from(src)
    run_ops(s, ops, encoded=enc)  # warm
    reps = 20

    def run():
        for _ in range(reps):
            s.copy_from(src)
            run_ops(s, ops, encoded=enc)

    dt, _ = timed(run, sync)
    dts, smp = timed(lambda: spz.sample(s, u), sync)
    return {"qubits": n, "gates": len(ops), "ms_per_qft": round(dt * 1e3 / reps, 4),
            "ms_per_4096_samples": round(dts * 1e3, 4)}, (s, smp, u, re, im, ops)


def cpu_baseline(ops, n, threads, budget_s=20.0):
    """Time the CPU baseline (the reference's per-kind parallel loop shapes,
    oracle/spz_oracle.c:orc_rayon_*) on the first gates of the same workload
    until about `budget_s` of work; returns (gates done, seconds)."""
    from oracle import pyoracle
    pyoracle.set_threads(threads)
    st = pyoracle.OracleState(n)
    done = 0
    t0 = time.perf_counter()
    for o in ops:
        pyoracle.rayon_execute(st, [o])
        done += 1
        if time.perf_counter() - t0 > budget_s:
            break
    dt = time.perf_counter() - t0
    del st
    return done, dt


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def main():
    a = parse()
    if a.gpus != 1:
        print(json.dumps({"error": "this build runs on exactly one GPU; --gpus must be 1 "
                          "(multi-GPU sharding is a design note, not built)"}))
        sys.exit(2)
    import ctypes
    import spinoza_amd as spz
    from spinoza_amd import _native as N
    from spinoza_amd.executor import encode_ops, run_ops
    from spinoza_amd.workloads import gate_bytes, random_circuit

    n, L = a.qubits, a.layers
    mode = {"auto": N.EXEC_AUTO, "plain": N.EXEC_PLAIN, "strict": N.EXEC_STRICT}[a.mode]
    ops = random_circuit(n, L, a.seed)
    enc = encode_ops(ops)
    alg_bytes_circuit = sum(gate_bytes(o, n) for o in ops)
    log(f"workload: {n}q, {L} layers, {len(ops)} gates, mode={a.mode}")

    s = spz.State(n)
    if a.calibrate:
        spz.apply(spz.Gate.X, s, n - 1)
        N.check(N.lib().spz_sync())
        log(f"calibration pass: X on qubit {n - 1}, {32 * (1 << n)} bytes moved")
    for w in range(a.warmup):
        s.set_basis(0)
        run_ops(s, ops, mode=mode, encoded=enc)
        N.check(N.lib().spz_sync())
        log(f"warmup {w} done")

    N.check(N.lib().spz_profile_enable(1))
    N.check(N.lib().spz_profile_reset())
    N.check(N.lib().spz_sync())
    t0 = time.perf_counter()
    for k in range(a.steps):
        s.set_basis(0)
        run_ops(s, ops, mode=mode, encoded=enc)
    N.check(N.lib().spz_sync())
    t1 = time.perf_counter()
    log(f"timed {a.steps} steps in {t1 - t0:.3f}s")

    prof = {}
    for cls, name in enumerate(N.K_NAMES):
        cnt, ms, by = ctypes.c_uint64(), ctypes.c_double(), ctypes.c_double()
        N.check(N.lib().spz_profile_get(cls, ctypes.byref(cnt), ctypes.byref(ms), ctypes.byref(by)))
        prof[name] = (cnt.value, ms.value, by.value)
    N.check(N.lib().spz_profile_enable(0))
    dom = max((k for k in prof if k != "other"), key=lambda k: prof[k][1])
    cnt, ms, by = prof[dom]
    achieved = by / (ms * 1e-3) if ms > 0 else 0.0

    ms_step = (t1 - t0) * 1e3 / a.steps
    value = len(ops) * a.steps / (t1 - t0)
    # HBM bytes per launch from the committed rocprofv3 FETCH_SIZE/WRITE_SIZE
    # summary of this kernel on this workload; kept only while the profiled
    # launch time is within 15% of the live one (a changed kernel makes the
    # summary stale, and the line then says so)
    traffic, traffic_src = None, None
    tpath = os.path.join(ROOT, "profiles", f"traffic_{dom}_{n}q.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            tj = json.load(f)
        live_us = ms / max(cnt, 1) * 1e3
        prof_us = tj.get("avg_launch_us") or 0.0
        fresh = prof_us > 0 and abs(prof_us - live_us) <= 0.15 * live_us
        traffic = tj.get("hbm_bytes_per_launch") if fresh else None
        traffic_src = {"file": f"profiles/traffic_{dom}_{n}q.json", "kernel": tj.get("kernel"),
                       "profiled_launch_us": prof_us, "live_launch_us": round(live_us, 1),
                       "used": bool(fresh)}

    result = {
        "metric": "30q f64 gates/s + % HBM roofline; 24q adjoint grads/s vs param-shift",
        "value": round(value, 2),
        "unit": "gates/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(ms_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"{n}-qubit random circuit, FP64 SoA state, {L} layers",
                   "qubits": n, "layers": L, "gates_per_step": len(ops), "seed": a.seed,
                   "executor": a.mode, "state_bytes": 16 * (1 << n)},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": round(achieved / 1e9, 1),
                     "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK, 4), "traffic": traffic,
                     "traffic_source": traffic_src,
                     "launches": cnt, "avg_launch_ms": round(ms / max(cnt, 1), 4),
                     "alg_bytes_per_launch": by / max(cnt, 1)},
        "kernel_time_ms": {k: round(v[1], 2) for k, v in prof.items() if v[0]},
        "alg_gb_per_step": round(alg_bytes_circuit / 1e9, 2),
        "effective_gb_s": round(alg_bytes_circuit * a.steps / (t1 - t0) / 1e9, 1),
    }
    del s

    adj_ctx = None
    if not a.no_adjoint:
        log(f"adjoint leg: {a.adjoint_qubits}q, {a.adjoint_layers} layers")
        result["adjoint"], adj_ctx = adjoint_leg(spz, N, a.adjoint_qubits, a.adjoint_layers,
                                                 a.ps_subset)

    if not a.no_fusion:
        n64, n32 = (int(x) for x in a.fusion_qubits.split(","))
        log(f"fusion leg: {n64}q FP64, {n32}q FP32, {a.fusion_layers} layers; "
            f"precision at {a.precision_qubits}q")
        result["fusion"] = fusion_leg(spz, N, a.fusion_layers, n64, n32, a.precision_qubits,
                                      a.dense_layers)

    if not a.no_batch:
        log(f"batched adjoint leg: {a.batch_count} circuits")
        result["batch"] = batch_leg(spz, N, a.batch_count, group_qubits=a.batch_group_qubits,
                                    wg_max=a.batch_wg_max, wg_per_cu=a.batch_wg_per_cu)

    qft_ctx = None
    if not a.no_qft:
        log("qft leg: 20q")
        result["qft"], qft_ctx = qft_leg(spz, N)

    if not a.no_strides:
        log("per-stride sweep")
        result["per_stride_gb_s"] = per_stride(spz, N, n)

    if not a.no_cpu:
        # cpu_baseline leg: the oracle as the timed CPU baseline and as the
        # checker of the GPU results above
        import numpy as np
        from oracle import pyoracle
        threads = a.cpu_threads or min(16, os.cpu_count() or 1)
        log(f"cpu baseline: reference loop shapes, {threads} threads, first gates of the "
            f"{n}q workload")
        done, dt = cpu_baseline(ops, n, threads)
        log("cpu baseline: 1 thread")
        done1, dt1 = cpu_baseline(ops, n, 1, budget_s=a.cpu_budget_1t)
        pyoracle.set_threads(threads)
        result["cpu_baseline"] = {
            "value": round(done / dt, 4), "unit": "gates/s", "cores": threads,
            "kind": "port",
            "sample": f"first {done} of the {len(ops)} gates of the same {n}q circuit "
                      f"({dt:.1f} s) on a 16 GiB host state: the reference's per-kind "
                      f"rayon loop shapes (chunk loops, insert-bit controlled loops, "
                      f"CZ as c_apply(P(pi))) in C with OpenMP, vectorised",
            "value_1_thread": round(done1 / dt1, 4),
            "sample_1_thread": f"first {done1} gates ({dt1:.1f} s)",
            "nproc": os.cpu_count(), "cpu_model": cpu_model()}
        result["speedup_vs_cpu"] = round(value / (done / dt), 1)
        checks = {}
        if not a.no_replica:
            rn = 26
            rops = random_circuit(rn, a.replica_layers, a.seed)
            log(f"check: 26q replica, {len(rops)} gates on the oracle")
            ref = pyoracle.OracleState(rn)
            pyoracle.execute(ref, rops)
            g = spz.State(rn)
            run_ops(g, rops, mode=mode)
            re, im = g.download()
            checks["replica_26q"] = {
                "layers": a.replica_layers, "gates": len(rops),
                "max_abs_err": float(max(np.max(np.abs(re - ref.re)), np.max(np.abs(im - ref.im)))),
                "bit_identical": bool(np.array_equal(re, ref.re) and np.array_equal(im, ref.im))}
            del g, ref
        if qft_ctx is not None:
            s20, smp, u, re0, im0, qops = qft_ctx
            log("check: 20q QFT and samples on the oracle")
            ref = pyoracle.OracleState(20)
            ref.re[:] = re0
            ref.im[:] = im0
            t0 = time.perf_counter()
            pyoracle.execute(ref, qops)
            t_cpu = time.perf_counter() - t0
            re, im = s20.download()
            checks["qft_20q"] = {
                "max_abs_err": float(max(np.max(np.abs(re - ref.re)), np.max(np.abs(im - ref.im)))),
                "samples_bit_identical": bool(np.array_equal(smp, ref.sample(u))),
                "oracle_ms": round(t_cpu * 1e3, 1)}
            del ref
        if adj_ctx is not None and a.cpu_ps_subset > 0:
            an, aops, where, g = adj_ctx
            sub = where[::max(1, len(where) // a.cpu_ps_subset)][:a.cpu_ps_subset]
            log(f"check: {an}q adjoint against oracle parameter shift ({len(sub)} parameters)")
            err = 0.0
            for i in sub:
                e = []
                for sh in (math.pi / 2, -math.pi / 2):
                    sops = [dict(o) for o in aops]
                    sops[i]["theta"] += sh
                    st = pyoracle.OracleState(an)
                    pyoracle.execute(st, sops)
                    e.append(st.ising_energy(1.0))
                    del st
                err = max(err, abs(g[i] - 0.5 * (e[0] - e[1])))
            checks["adjoint_vs_cpu_param_shift"] = {"qubits": an, "params": len(sub),
                                                   "max_abs_err": err}
        result["checks"] = checks

    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
