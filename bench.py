#!/usr/bin/env python
"""Plain SSB throughput on one MI355X: resident uint16 I[256,256,128,128] -> float32 phase.

One step is one end-to-end reconstruction at the R1 boundary: exact uint64
mean diffraction pattern over the full 2 GiB stream, on-device bright-field
selection, BF gather + scan FFT fused with the trotter reduction, inverse
FFT and angle.  The input is the seeded synthetic weak-phase dataset of the
benchmark configuration (300 kV, 20 mrad, 1.25 mrad/px, defocus C10 = 200 A,
5e4 e-/probe), generated on the GPU before timing.

Prints one JSON line (see the repository contract): value = recon/s,
roofline of the dominant kernel (the mean-DP stream) from hipEvents recorded
inside the timed region around that kernel in its first ten steps (an event
pair costs ~6 us of queue time per reconstruction, so the other steps run
without one), the whole-reconstruction roofline, GPU latencies at
the R1 (raw frames -> phase) and R2 (resident Fourier stack -> phase)
boundaries, the batched search legs (phase-image variance and exact
objective, direct and hipGraph replay, uneven scans in one launch), the fit
leg (learned start vs plain search), and the CPU baseline: the NumPy/SciPy
restatement timed on this host after one warm-up run, p50/p95/max of at
least five runs, R1 and R2, float32 beside float64, on the 64²x64² (padded
to 128²) CPU-path workload and on this 256²x128² workload.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time
from pathlib import Path

REPO = Path(__file__).resolve().parent
sys.path.insert(0, str(REPO))

METRIC = "SSB recon/s 256²×128² u16; aberration-fit s, learned vs search, same phase err"
HBM_PEAK_GBS = 8000.0
SCAN, DET = 256, 128
C10 = 200.0                      # 20 nm defocus in the engine's numeric (Angstrom) convention
TRAFFIC_FILE = REPO / "profiles" / "mean_dp_stream_traffic.txt"


def _abi_lib():
    from quantem.gpu_amd.ssb.compute.hip import _abi
    return _abi.lib()


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def read_traffic() -> float | None:
    """HBM bytes per launch of the mean-DP stream, from the committed rocprofv3 counter summary."""
    if not TRAFFIC_FILE.exists():
        return None
    for line in TRAFFIC_FILE.read_text().splitlines():
        if line.startswith("traffic_bytes_per_launch:"):
            try:
                return float(line.split(":", 1)[1].strip())
            except ValueError:
                return None
    return None


def percentiles(ms: list[float]) -> dict:
    """p50 / p95 / max by nearest rank (rank = ceil(p/100 * n)), in ms."""
    s = sorted(ms)
    n = len(s)

    def rank(p):
        return s[max(0, min(n - 1, math.ceil(p / 100.0 * n) - 1))]

    return {"p50_ms": rank(50), "p95_ms": rank(95), "max_ms": s[-1], "runs": n}


def _time_runs(fn, runs: int) -> dict:
    fn()                                            # one untimed warm-up
    ts = []
    for _ in range(runs):
        t0 = time.perf_counter()
        fn()
        ts.append((time.perf_counter() - t0) * 1e3)
    return percentiles(ts)


def cpu_baseline(cases, runs: int = 5) -> dict:
    """The oracle's NumPy/SciPy restatement of the reconstruction, timed on
    this host's cores: FP32 (the reference algorithm as shipped, "port") and
    the FP64 oracle beside it, at R1 (resident u16 frames -> phase: scan
    normalisation, mean DP, BF selection, gather, scan FFT, corrected Fourier
    sum, inverse FFT, angle) and R2 (prepared Fourier stack -> phase), one
    untimed warm-up then ``runs`` timed runs each, p50/p95/max.  ``cases``
    maps a configuration name to (host u16 frames, SyntheticConfig, C10);
    ``value`` is FP32 R1 of the configuration measured on the GPU."""
    import numpy as np
    from oracle import ssb_oracle as O

    cores = O._workers()
    out = {}
    for name, (host, cfg, c10) in cases.items():
        row = {}
        for label, dt in (("fp32", np.float32), ("fp64", np.float64)):
            def r1(dt=dt):
                p = O.prepare(host, voltage_kV=300, semiangle_mrad=cfg.semiangle_mrad, scan_sampling=cfg.step_A,
                              det_sampling=cfg.det_sampling_mrad, dtype=dt)
                return np.angle(O.object_wave(p.G, p.dc, p.geom, c10, 0.0, 0.0))
            row[f"R1_{label}"] = _time_runs(r1, runs)
            p = O.prepare(host, voltage_kV=300, semiangle_mrad=cfg.semiangle_mrad, scan_sampling=cfg.step_A,
                          det_sampling=cfg.det_sampling_mrad, dtype=dt)
            row[f"R2_{label}"] = _time_runs(lambda p=p: np.angle(O.object_wave(p.G, p.dc, p.geom, c10, 0.0, 0.0)),
                                            runs)
            row["n_bf"] = int(p.bf.rows.size)
            del p
            log(f"CPU {name} {label}: R1 p50 {row[f'R1_{label}']['p50_ms']:.1f} ms, "
                f"R2 p50 {row[f'R2_{label}']['p50_ms']:.1f} ms")
        out[name] = row
    main = out["config2"]["R1_fp32"]
    return {"value": 1e3 / main["p50_ms"], "unit": "recon/s", "cores": cores, "kind": "port",
            "sample": f"scan 256x256 x detector 128x128 u16 (the GPU workload) and scan 64x64 x detector 64x64 "
                      f"(padded to 128x128); R1 and R2, FP32 restatement and FP64 oracle, 1 warm-up + {runs} "
                      f"timed runs each, scipy.fft and BF chunks on {cores} threads; value = 1/p50 of FP32 R1 "
                      f"at 256x256",
            "runs": out}


def _phase_of(ssb, state):
    """Mean per-BF phase (the image the exact objective scores) at a search state."""
    import numpy as np
    ssb._set_search_state(state)
    mags = np.zeros(14)
    mags[0], mags[4] = state["C10_nm"], state["C30"]
    return ssb.reconstruct_full(mags, np.zeros(14)).cpu().numpy().astype(np.float64)


def _rms_wrapped(a, b) -> float:
    import numpy as np
    d = np.angle(np.exp(1j * (a - b)))
    return float(np.sqrt(np.mean(d * d)))


def fit_leg(dev, n_sets: int, train_steps: int, search_trials: int, learned_trials: int, refine: str,
            rounds: int | None, seed: int = 2027) -> dict:
    """Aberration fit at scan 256x256, detector 128x128: learned start + a
    local batched search + refine, against the plain batched search (global
    Sobol) + the same refine.

    Per held-out dataset (C10 U(+-500 A), C30 U(+-2e5 A), rotation U(+-5 deg),
    5e4 e-/probe): both fits run from zero aberrations.  The reference optimum
    of the exact objective is the minimum of a least-squares quadratic model
    through 2 x 64 points around a fit's result (untimed), taken from the fit
    with the lower loss; the same model from the other fit's result gives the
    reference's own spread.  (The exact loss is rough at the ~3e-8 level near
    its optimum: Nelder-Mead polishes from two nearby starts land up to
    ~1.2e-4 rad apart in phase, the model minima within ~4e-5.)  Each fit's
    phase error is the RMS of the wrapped phase difference to the reference.
    Training time is reported on its own."""
    import numpy as np
    import torch
    from quantem.gpu_amd.ssb.learned import LearnedInit, FeatureConfig
    from quantem.gpu_amd.ssb.synthetic import SyntheticConfig, generate
    from quantem.gpu_amd.ssb.compute.hip.backend import HipSSBBackend

    base = SyntheticConfig(scan=SCAN, det=DET)
    fcfg = FeatureConfig.for_scan(SCAN, base.step_A, base.semiangle_mrad * 1e-3, base.det_sampling_mrad * 1e-3,
                                  base.wavelength)
    net = LearnedInit(device=dev, seed=1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    hist = net.train(fcfg, steps=train_steps, batch=256, lr=2e-3, seed=2, log_every=max(1, train_steps // 4))
    torch.cuda.synchronize()
    train_s = time.perf_counter() - t0
    log(f"learned initializer trained: {train_steps} steps in {train_s:.1f} s, loss {hist[0][1]:.4f} -> {hist[-1][1]:.4f}")
    rng = np.random.default_rng(seed)
    zero = {"C10_nm": 0.0, "C12_nm": 0.0, "phi12_deg": 0.0, "C30": 0.0, "rotation_angle_deg": 0.0}
    keys = ("C10_nm", "C30", "rotation_angle_deg")
    rows = []
    for i in range(n_sets + 1):                     # set 0 warms both paths up and is not reported
        truth = {"C10": float(rng.uniform(-500, 500)), "C30": float(rng.uniform(-2e5, 2e5))}
        rot = float(rng.uniform(-5, 5))
        sc = SyntheticConfig(scan=SCAN, det=DET, aberrations=truth, rotation_deg=rot, dose=5e4, seed=9000 + i)
        data = generate(sc, device="cuda")
        ssb = HipSSBBackend(data, semiangle=sc.semiangle_mrad, scan_sampling=sc.step_A,
                            det_sampling=sc.det_sampling_mrad, voltage_kV=sc.voltage_kV)
        out = {}
        for name in ("search", "learned"):
            ssb._set_search_state(zero)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            if name == "search":
                ssb.fit_search(trials=search_trials, refine=refine, rounds=rounds)
            else:
                ssb.fit_learned(net, fcfg, trials=learned_trials, refine=refine, rounds=rounds)
            torch.cuda.synchronize()
            out[name] = (time.perf_counter() - t0, ssb._search_state(), float(ssb._best_loss), ssb._refine_nfev)
        refs = {}
        for name, (_, state, _, _) in out.items():
            ssb._set_search_state(state)
            ssb.refine_polish(rounds=2, batch=64, seed=0)
            refs[name] = (ssb._search_state(), float(ssb._best_loss))
        best_name = min(out, key=lambda n: out[n][2])
        opt_state, opt_loss = refs[best_name]
        ph_opt = _phase_of(ssb, opt_state)
        spread = _rms_wrapped(_phase_of(ssb, refs["search"][0]), _phase_of(ssb, refs["learned"][0]))
        row = {"truth": {"C10": truth["C10"], "C30": truth["C30"], "rotation_deg": rot},
               "optimum": {k: opt_state[k] for k in keys}, "optimum_loss": opt_loss,
               "reference_spread_rad": spread}
        for name, (secs, state, loss, nfev) in out.items():
            row[name] = {"s": secs, "loss": loss, "phase_err_rad": _rms_wrapped(_phase_of(ssb, state), ph_opt),
                         "params": {k: state[k] for k in keys}, "refine_evals": nfev}
        if i > 0:
            rows.append(row)
            log(f"fit set {i}: search {row['search']['s']:.3f} s err {row['search']['phase_err_rad']:.2e}, "
                f"learned {row['learned']['s']:.3f} s err {row['learned']['phase_err_rad']:.2e}")
        ssb.close()
        del data
    ts = float(np.median([r["search"]["s"] for r in rows]))
    tl = float(np.median([r["learned"]["s"] for r in rows]))
    return {
        "workload": "aberration fit (C10, C30, rotation), scan 256x256, detector 128x128, u16, exact objective",
        "search_s": ts, "learned_s": tl, "speedup": ts / tl,
        "search_trials": search_trials, "learned_trials": learned_trials, "refine": refine,
        "refine_evals": {"search": float(np.median([r["search"]["refine_evals"] for r in rows])),
                         "learned": float(np.median([r["learned"]["refine_evals"] for r in rows]))},
        "max_phase_err_rad": {"search": max(r["search"]["phase_err_rad"] for r in rows),
                              "learned": max(r["learned"]["phase_err_rad"] for r in rows)},
        "phase_err_bar_rad": 1e-4,
        "reference": "least-squares quadratic model minimum (2 x 64 points) around the lower-loss fit's result",
        "reference_spread_max_rad": max(r["reference_spread_rad"] for r in rows),
        "sets_within_bar": {name: sum(r[name]["phase_err_rad"] <= 1e-4 for r in rows) for name in ("search", "learned")},
        "n_sets": len(rows),
        "training_s": train_s, "training_steps": train_steps, "training_loss": [hist[0][1], hist[-1][1]],
        "sets": rows,
    }


def series_leg(data, spec, dev, knobs: dict | None, in_flight: int = 3, recons: int = 30) -> dict:
    """A series of reconstructions with ``in_flight`` plans on as many
    streams (default priority), each plan's raw-data stream pass ordered
    after the previous plan's (``ssb_recon_u16_after``), so one stream pass
    runs while the other plans finish their BF stages.  Reported beside
    ``value`` (one reconstruction at a time), never as it; every plan's
    phase must equal the one-at-a-time result bit for bit."""
    import torch
    from quantem.gpu_amd.ssb.compute.hip.engine import SSBPlan, HipSSBEngine
    from quantem.gpu_amd.ssb.compute.hip import tuner
    engs = [HipSSBEngine(SSBPlan(spec, device=dev)) for _ in range(in_flight)]
    for e in engs:
        for knob, v in (knobs or {}).items():
            e.tune_set(tuner.KNOBS[knob], int(v))
    phases = [torch.empty((SCAN, SCAN), dtype=torch.float32, device=dev) for _ in range(in_flight)]
    main = torch.cuda.current_stream(dev)
    streams = [main] + [torch.cuda.Stream(device=dev) for _ in range(in_flight - 1)]

    def run(n):
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        e0.record(main)
        for st in streams[1:]:
            st.wait_event(e0)
        for i in range(n):
            k = i % in_flight
            with torch.cuda.stream(streams[k]):
                engs[k].recon_u16(data, C10, 0.0, 0.0, 0.0, 0.0, None, phase_out=phases[k],
                                  after=engs[(k - 1) % in_flight])
        for st in streams[1:]:
            ev = torch.cuda.Event()
            ev.record(st)
            main.wait_event(ev)
        e1.record(main)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / n

    run(2 * in_flight)
    ms = run(recons)
    single = torch.empty_like(phases[0])
    engs[0].recon_u16(data, C10, 0.0, 0.0, 0.0, 0.0, None, phase_out=single)
    torch.cuda.synchronize()
    identical = all(torch.equal(single, ph) for ph in phases)
    for e in engs:
        e.plan.close()
    return {"in_flight": in_flight, "recons": recons, "ms_per_recon": ms, "recon_per_s": 1e3 / ms,
            "phase_identical_to_single": bool(identical)}


def decode_leg(data, reps: int = 5, distinct: int = 64) -> dict:
    """GPU decode of bitshuffle-LZ4 chunks (one frame each) into the resident
    uint16 array: ``distinct`` frames of the benchmark input compressed on
    the host by the library's encoder, tiled to every frame of the scan
    (65,536 chunks, 2 GiB decoded), decoded ``reps`` times; checks the
    decoded frames bit for bit."""
    import numpy as np
    import torch
    from quantem.gpu_amd.ssb.compute.hip.bslz4 import Bslz4Chunks, encode_u16

    frames = data.reshape(-1, data.shape[-2], data.shape[-1])
    n = frames.shape[0]
    uniq = frames[:distinct].cpu().numpy()
    chunks = [np.frombuffer(encode_u16(f), dtype=np.uint8) for f in uniq]
    period = np.concatenate(chunks)
    sizes = np.tile(np.array([c.size for c in chunks], dtype=np.int64), n // distinct)
    buf = np.tile(period, n // distinct)
    dec = Bslz4Chunks((buf, sizes), (1,) + tuple(frames.shape[1:]))
    out = torch.empty_like(frames)
    dec.decode(out)                                  # warm, and validates every chunk
    stream = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(reps):
        dec.decode(out, sync=False)
    e1.record(stream)
    torch.cuda.synchronize()
    dec.check()
    ms = e0.elapsed_time(e1) / reps
    ok = bool(torch.equal(out.view(n // distinct, distinct, *frames.shape[1:]),
                          torch.from_numpy(uniq).to(out.device).unsqueeze(0).expand(n // distinct, -1, -1, -1)))
    out_bytes = 2.0 * out.numel()
    alg = out_bytes + dec.compressed_bytes
    # detector binning by 2 of the decoded frames (uint32 sums, the binned
    # load's default), timed the same way
    from quantem.gpu_amd.ssb.compute.hip.binning import bin_frames
    binned = bin_frames(out, 2)
    e0.record(stream)
    for _ in range(reps):
        bin_frames(out, 2, out=binned)
    e1.record(stream)
    torch.cuda.synchronize()
    bin_ms = e0.elapsed_time(e1) / reps
    bin_bytes = out_bytes + 4.0 * binned.numel()
    bin_ok = bool(torch.equal(binned.reshape(n, -1).to(torch.int64).sum(dim=1),
                              out.reshape(n, -1).to(torch.int64).sum(dim=1)))
    binning = {"workload": f"det_bin 2 of {n} frames of {frames.shape[1]}x{frames.shape[2]} u16 -> u32",
               "ms": bin_ms, "totals_preserved": bin_ok,
               "roofline": {"bound": "hbm", "achieved": bin_bytes / (bin_ms * 1e-3) / 1e9, "peak": HBM_PEAK_GBS,
                            "unit": "GB/s", "frac": bin_bytes / (bin_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                            "algorithmic_bytes": bin_bytes}}
    del binned
    return {"workload": f"bitshuffle-LZ4 decode, {n} chunks of one {frames.shape[1]}x{frames.shape[2]} u16 frame",
            "ms": ms, "decoded_GBps": out_bytes / (ms * 1e-3) / 1e9,
            "compression_ratio": out_bytes / dec.compressed_bytes,
            "roofline": {"bound": "hbm", "achieved": alg / (ms * 1e-3) / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": alg / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                         "algorithmic_bytes": alg},
            "bit_exact": ok, "binning": binning}


def _event_times(fn, n: int, stream) -> list[float]:
    import torch
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(n)]
    for a, b in evs:
        a.record(stream)
        fn()
        b.record(stream)
    torch.cuda.synchronize()
    return [a.elapsed_time(b) for a, b in evs]


def gpu_latency(eng, data, step, n: int) -> dict:
    """Per-call hipEvent latency of R1 (resident u16 frames -> phase, the
    ``value`` workload) and R2 (prepared resident Fourier stack -> phase, the
    reference's published boundary), p50/p95/max over ``n`` calls each."""
    import torch
    stream = torch.cuda.current_stream()
    r1 = percentiles(_event_times(step, n, stream))
    eng.mean_dp(data)
    eng.select_bf(0.0, None)
    eng.prepare_g(data)
    eng.cache_rotation(0.0)
    phase = torch.empty((SCAN, SCAN), dtype=torch.float32, device=data.device)
    warm = lambda: eng.reconstruct_object(C10, 0.0, 0.0, phase_out=phase)
    for _ in range(3):
        warm()
    r2 = percentiles(_event_times(warm, n, stream))
    return {"R1": r1, "R2": r2,
            "R2_bytes": 8.0 * eng.num_bf * SCAN * (SCAN // 2 + 1) + 4.0 * SCAN * SCAN}


def gpu_latency_small(data, cfg, dev, n: int) -> dict:
    """R1 latency of the scan 64 x 64, detector 64 x 64 configuration (padded
    to 128 x 128 on the device first, as the backend does)."""
    import torch
    from quantem.gpu_amd.ssb.compute.hip.engine import HipSSBEngine, PlanSpec, SSBPlan, normalize_scan
    padded = normalize_scan(data)
    spec = PlanSpec(128, 64, 64, 4096, cfg.wavelength, cfg.semiangle_mrad * 1e-3, cfg.det_sampling_mrad * 1e-3,
                    cfg.det_sampling_mrad * 1e-3, cfg.step_A, cfg.step_A)
    e = HipSSBEngine(SSBPlan(spec, device=dev))
    phase = torch.empty((128, 128), dtype=torch.float32, device=dev)
    fn = lambda: e.recon_u16(normalize_scan(data), 0.0, 0.0, 0.0, 0.0, 0.0, None, phase_out=phase)
    for _ in range(3):
        fn()
    out = {"workload": "scan 64x64 x detector 64x64 u16, zero aberrations, padded to 128x128 (pad included)",
           "R1": percentiles(_event_times(fn, n, torch.cuda.current_stream()))}
    e.plan.close()
    del padded
    return out


def trials_leg(dev, data, cfg, n_trials: int = 512, reps: int = 5, seed: int = 3) -> dict:
    """Batched aberration search (scan 256 x 256, detector 128 x 128): 512
    trials with C10 U(+-500 A), C30 U(+-2e5 A), rotation U(+-5 deg), each
    evaluated with the phase-image-variance objective in one persistent launch
    replayed from a hipGraph; beside it the same trials with the exact
    objective (the per-trial kernels, eager and graph-replayed), and the
    uneven-scan mix (scans 128, 192 -> 256 and 256 at detector 128, plus the
    benchmark dataset) in one launch.  A search unit is one (trial, BF pixel,
    Hermitian-half Q) evaluation."""
    import numpy as np
    import torch
    from quantem.gpu_amd.ssb.compute.hip.backend import HipSSBBackend
    from quantem.gpu_amd.ssb.compute.hip.trials import TrialBatch, params_to_vectors, sobol_design
    from quantem.gpu_amd.ssb.synthetic import SyntheticConfig, generate

    def backend(d, c):
        return HipSSBBackend(d, semiangle=c.semiangle_mrad, scan_sampling=c.step_A, det_sampling=c.det_sampling_mrad,
                             voltage_kV=c.voltage_kV)

    ranges = {"C10_nm": (-500.0, 500.0), "C30": (-2.0e5, 2.0e5), "rotation_angle_deg": (-5.0, 5.0)}
    params = sobol_design(ranges, n_trials, seed)
    mags, angs, rot = params_to_vectors(params)
    main = backend(data, cfg)
    eng = main._engine
    stream = torch.cuda.current_stream()

    def timed(fn):
        fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(reps):
            fn()
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps

    tb = TrialBatch([eng], n_trials)
    tb.set(mags, angs, rot)
    ms_graph = timed(lambda: tb.phase_variance(use_graph=True))
    ms_direct = timed(lambda: tb.phase_variance(use_graph=False))
    var = tb.phase_variance().cpu().numpy()[0]
    tb.close()
    nbf = eng.num_bf
    units = float(n_trials) * nbf * SCAN * (SCAN // 2 + 1)
    out_exact = torch.empty(n_trials, dtype=torch.float32, device=data.device)
    ms_exact = timed(lambda: eng.batch_phase_loss(mags, angs, rot, out=out_exact))
    ms_exact_graph = timed(lambda: eng.batch_phase_loss(mags, angs, rot, out=out_exact, graph=True))
    best = params[int(np.argmax(var))]
    # uneven scans in one launch: 128, 192 (padded to 256) and 256 at detector 128
    extra = []
    for scan, s_ in ((128, 41), (192, 42)):
        c = SyntheticConfig(scan=scan, det=DET, aberrations={"C10": 120.0, "C30": -5.0e4}, rotation_deg=1.0,
                            dose=5e4, seed=s_)
        extra.append(backend(generate(c, device="cuda"), c))
    engines = [b._engine for b in extra] + [eng]
    mix = TrialBatch(engines, n_trials)
    mix.set(mags, angs, rot)
    ms_mix = timed(lambda: mix.phase_variance(use_graph=True))
    mix_units = float(n_trials) * sum(e.num_bf * e.n * (e.n // 2 + 1) for e in engines)
    mix.close()
    for b in extra:
        b.close()
    main.close()
    return {
        "workload": f"{n_trials} trials (C10 U(+-500 A), C30 U(+-2e5 A), rotation U(+-5 deg)), scan 256x256, "
                    f"detector 128x128, N_bf {nbf}",
        "phase_variance": {"objective": "phase-image variance var_R angle(O_t) (maximised)",
                           "ms_per_batch_graph": ms_graph, "ms_per_batch_direct": ms_direct,
                           "trials_per_s": n_trials / (ms_graph * 1e-3), "units": units,
                           "units_per_s": units / (ms_graph * 1e-3), "best": best},
        "exact": {"objective": "mean_R var_b phi_b (the reference's default fit objective)",
                  "ms_per_batch_eager": ms_exact, "ms_per_batch_graph": ms_exact_graph,
                  "trials_per_s": n_trials / (ms_exact_graph * 1e-3)},
        "uneven_mix": {"scans": [128, 256, 256], "source_scans": [128, 192, 256], "ms_per_batch_graph": ms_mix,
                       "units_per_s": mix_units / (ms_mix * 1e-3)},
        "roofline": search_unit_roofline(units / (ms_graph * 1e-3)),
    }


SEARCH_PROFILE = REPO / "profiles" / "trials_search_512x256x128.md"
VALU_LANES_PER_S = 1024 * 16 * 2.4e9   # 256 CUs x 4 SIMDs, one wave64 VALU instruction per 4 cycles, 2.4 GHz


def read_profile_figure(path: Path, key: str) -> float | None:
    """A ``key: value`` line of a committed rocprofv3 summary, or None."""
    if not path.exists():
        return None
    for line in path.read_text().splitlines():
        if line.strip().startswith(key + ":"):
            try:
                return float(line.split(":", 1)[1].strip())
            except ValueError:
                return None
    return None


def search_unit_roofline(units_per_s: float) -> dict:
    """The batched search against its VALU issue bound: VALU (and
    transcendental) lane-instructions per search unit from the committed
    counter summary of the same launch, the bound = FP32 VALU lane rate
    (packed FP32 unused) / instructions per unit, frac = achieved / bound."""
    per = read_profile_figure(SEARCH_PROFILE, "valu_lane_instructions_per_unit")
    trans = read_profile_figure(SEARCH_PROFILE, "trans_f32_lane_instructions_per_unit")
    if not per:
        return {"bound": "valu", "achieved": units_per_s, "unit": "units/s", "peak": None, "frac": None}
    peak = VALU_LANES_PER_S / per
    return {"bound": "valu", "achieved": units_per_s, "peak": peak, "unit": "units/s", "frac": units_per_s / peak,
            "valu_lane_instructions_per_unit": per, "trans_f32_lane_instructions_per_unit": trans,
            "counters": str(SEARCH_PROFILE.relative_to(REPO))}


def copy_probe() -> int:
    """Five launches of an elementwise copy reading and writing 1 GiB each: the
    reference kernel of known traffic for the FETCH_SIZE/WRITE_SIZE counters."""
    import torch
    n = 1 << 28
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(5):
        torch.mul(src, 1.0, out=dst)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes_read": 4 * n, "copy_probe_bytes_written": 4 * n}), flush=True)
    return 0


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu-baseline", action="store_true",
                    help="skip the CPU baseline (used for profiler runs)")
    ap.add_argument("--cpu-runs", type=int, default=5, help="timed CPU-baseline runs after one warm-up")
    ap.add_argument("--fit-sets", type=int, default=5,
                    help="held-out datasets of the aberration-fit comparison (0 skips it)")
    ap.add_argument("--fit-train-steps", type=int, default=4000)
    ap.add_argument("--fit-search-trials", type=int, default=512)
    ap.add_argument("--fit-trials", type=int, default=16, help="trials of the local search around the learned start")
    ap.add_argument("--fit-refine", default="newton", choices=["newton", "nelder-mead", "quadratic"])
    ap.add_argument("--fit-rounds", type=int, default=None,
                    help="batches of the newton refine (default 8) or rounds of the quadratic refine (default 4)")
    ap.add_argument("--no-series", action="store_true",
                    help="skip the series measurement (3 reconstructions in flight)")
    ap.add_argument("--no-decode", action="store_true",
                    help="skip the bitshuffle-LZ4 decode measurement")
    ap.add_argument("--no-tune", action="store_true",
                    help="keep the default launch shapes instead of autotuning them before timing")
    ap.add_argument("--no-trials", action="store_true", help="skip the batched-search leg")
    ap.add_argument("--trials-only", action="store_true",
                    help="only the batched-search leg (used for profiler runs)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run a plain 1 GiB float32 copy kernel (calibrates HBM byte counters)")
    args = ap.parse_args()
    if args.copy_probe:
        return copy_probe()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one MI355X: the SSB path runs on one GPU and this library has no "
              "multi-GPU mode (the path would shard by BF pixels, one all-reduce of the Fourier-sum plane, or by "
              "trials).", file=sys.stderr)
        return 2

    import torch
    from quantem.gpu_amd.ssb.synthetic import SyntheticConfig, generate
    from quantem.gpu_amd.ssb.compute.hip.engine import SSBPlan, PlanSpec, HipSSBEngine
    from quantem.gpu_amd.optics import electron_wavelength_angstrom

    if not torch.cuda.is_available():
        print("bench.py needs a ROCm GPU", file=sys.stderr)
        return 1
    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    cfg = SyntheticConfig(scan=SCAN, det=DET, aberrations={"C10": C10}, dose=5e4, seed=2026)
    log("generating synthetic 256x256x128x128 u16 input on the GPU")
    data = generate(cfg, device="cuda")
    torch.cuda.synchronize()

    lam = electron_wavelength_angstrom(300e3)
    spec = PlanSpec(SCAN, DET, DET, 4096, lam, cfg.semiangle_mrad * 1e-3,
                    cfg.det_sampling_mrad * 1e-3, cfg.det_sampling_mrad * 1e-3, cfg.step_A, cfg.step_A)
    if args.trials_only:
        print(json.dumps({"trials": trials_leg(dev, data, cfg)}), flush=True)
        return 0
    eng = HipSSBEngine(SSBPlan(spec, device=dev))
    phase = torch.empty((SCAN, SCAN), dtype=torch.float32, device=dev)

    def step():
        eng.recon_u16(data, C10, 0.0, 0.0, 0.0, 0.0, None, phase_out=phase)

    for _ in range(max(1, args.warmup)):
        step()
    torch.cuda.synchronize()
    tuning = None
    if not args.no_tune:
        from quantem.gpu_amd.ssb.compute.hip import tuner
        tuning = tuner.autotune(eng, data, c10=C10, repeats=5)
        log(f"autotuned launch shapes in {tuning['seconds']:.1f} s: {tuning['config']}")
        for _ in range(max(1, args.warmup)):
            step()
        torch.cuda.synchronize()
    eng.mean_dp(data)
    nbf = eng.select_bf(0.0, None).rows.size
    log(f"warmup done; N_bf = {nbf}")

    stream = torch.cuda.current_stream(dev)
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    # timed region: hipEvents only around the dominant kernel (the mean-DP
    # stream pass: the boundaries before and after its stage), on the plan's
    # stream, in the first n_dom steps (each event pair adds ~6 us of queue
    # time to its reconstruction; the kernel's duration does not vary by step)
    lib = _abi_lib()
    names = [lib.ssb_stage_name(i).decode() for i in range(int(lib.ssb_stage_count()))]
    k_stream = names.index("mean_dp_stream")
    n_dom = min(10, args.steps)
    eng.timing_begin(n_dom, boundaries=(k_stream, k_stream + 1))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    e0.record(stream)
    for _ in range(args.steps):
        step()
    e1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    dom = eng.timing_end()
    gpu_ms = e0.elapsed_time(e1)
    ms_per_step = gpu_ms / args.steps
    value = args.steps / (gpu_ms / 1e3)
    log(f"{args.steps} steps: {ms_per_step:.4f} ms/recon (wall {wall * 1e3 / args.steps:.4f} ms)")
    # per-stage breakdown from a separate, untimed pass with events at every stage boundary
    n_diag = min(10, args.steps)
    eng.timing_begin(n_diag)
    for _ in range(n_diag):
        step()
    stages = eng.timing_end()

    n_r, n_k = SCAN * SCAN, DET * DET
    dom_bytes = 2.0 * n_r * n_k                       # raw u16 read once by the stream
    dom_ms = dom["mean_dp_stream"] / dom["calls"]
    achieved = dom_bytes / (dom_ms * 1e-3) / 1e9
    recon_bytes = 2.0 * n_r * n_k + 2.0 * n_r * nbf + 4.0 * n_r
    recon_achieved = recon_bytes / (ms_per_step * 1e-3) / 1e9
    traffic = read_traffic()

    result = {
        "metric": METRIC,
        "value": value,
        "unit": "recon/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded weak-phase forward model, Poisson, u16)",
        "config": {"workload": "plain SSB, R1: resident u16 I[256,256,128,128] -> f32 phase[256,256]",
                   "scan": [SCAN, SCAN], "detector": [DET, DET], "n_bf": int(nbf),
                   "aberrations_A": {"C10": C10}, "voltage_kV": 300, "semiangle_mrad": 20.0,
                   "det_sampling_mrad": 1.25, "dose": 5e4},
        "roofline": {"bound": "hbm", "kernel": "k_dp_partial (mean_dp_stream)",
                     "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "algorithmic_bytes_per_launch": dom_bytes, "avg_launch_ms": dom_ms,
                     "timed_launches": int(dom["calls"])},
        "recon_roofline": {"bound": "hbm", "achieved": recon_achieved, "peak": HBM_PEAK_GBS,
                           "unit": "GB/s", "frac": recon_achieved / HBM_PEAK_GBS,
                           "algorithmic_bytes_per_recon": recon_bytes},
        "stages_ms_per_recon": {k: v / stages["calls"] for k, v in stages.items() if k != "calls"},
        "tuning": tuning,
    }
    log("per-call latency: R1 and R2 (warm resident Fourier stack)")
    result["latency"] = gpu_latency(eng, data, step, args.steps)
    cfg1 = SyntheticConfig(scan=64, det=64, dose=1e4, seed=64)
    data1 = generate(cfg1, device="cuda")
    result["latency"]["config1"] = gpu_latency_small(data1, cfg1, dev, args.steps)
    if not args.no_cpu_baseline:
        log("timing the CPU baseline (FP32 NumPy/SciPy restatement and FP64 oracle, R1 and R2)")
        cases = {"config2": (data.cpu().numpy(), cfg, C10), "config1": (data1.cpu().numpy(), cfg1, 0.0)}
        result["cpu_baseline"] = cpu_baseline(cases, runs=args.cpu_runs)
        del cases
    else:
        result["cpu_baseline"] = None
    del data1
    eng.plan.close()
    if not args.no_series:
        log("series of reconstructions, 3 in flight")
        result["series"] = series_leg(data, spec, dev, (tuning or {}).get("config"))
        log(f"series: {result['series']['ms_per_recon']:.4f} ms/recon, identical "
            f"{result['series']['phase_identical_to_single']}")
    if not args.no_decode:
        log("bitshuffle-LZ4 decode of the same frames")
        result["decode"] = decode_leg(data)
    if not args.no_trials:
        log("batched aberration search: 512 trials in one launch")
        result["trials"] = trials_leg(dev, data, cfg)
        log(f"trials: phase variance {result['trials']['phase_variance']['ms_per_batch_graph']:.2f} ms / 512, "
            f"exact {result['trials']['exact']['ms_per_batch_graph']:.2f} ms / 512")
    del data
    if args.fit_sets > 0:
        log("aberration fit: learned start + batched refine vs plain batched search")
        result["fit"] = fit_leg(dev, args.fit_sets, args.fit_train_steps, args.fit_search_trials,
                                args.fit_trials, args.fit_refine, args.fit_rounds)
    print(json.dumps(result), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
