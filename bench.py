"""Throughput benchmark of the photon-transport hot path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--histories H]

Workload (config 2): CLARITAS small cell (bore 25 mm, water 50 mm), 40-bin
lognormal PSD (mass median 20 um, sigma_g 1.8, m = 1.46 + 1e-4 i), Rouse
profile on a 64x128 FP32 r-z grid, Rankine free surface (core 3 mm, depth
5 mm), 860 nm collimated 1 mm beam, 1800-angle Mie table, plain-Mie event
law, concentration sweep {10, 31.6, 100, 316, 1000, 3162, 5000} mg/L.
One step = the whole sweep with H histories per concentration (default
1e8), traced in one persistent batched launch from HBM-resident tables.
Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

FP32_PEAK_TFLOPS = 157.3  # MI355X published FP32 peak
TRAFFIC_PROFILE = ROOT / 'profiles' / 'transport_kernel_config2_hbm_traffic.txt'


def traffic_per_history():
    """HBM bytes per history of the transport kernel on this workload, from
    the committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary (each counter
    scaled by its reading on a plain copy of known size); None if absent."""
    try:
        for line in TRAFFIC_PROFILE.read_text().splitlines():
            if line.startswith('hbm_bytes_per_history:'):
                return float(line.split(':', 1)[1].split()[0])
    except (OSError, ValueError):
        pass
    return None


def copy_calibration(mib: int) -> None:
    """A plain device-to-device copy of `mib` MiB (3 times), the known-size
    reference for scaling the FETCH_SIZE / WRITE_SIZE counters."""
    import torch
    n = mib * (1 << 20) // 4
    a = torch.ones(n, dtype=torch.float32, device='cuda')
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({'copy_calibration_bytes': n * 4, 'repeats': 3}))


def cpu_baseline(histories_per_conc: int, seed: int):
    """The CPU oracle (restatement of the reference algorithm) on a bounded
    sample of the same workload, all host cores, timed on this host."""
    import numpy as np
    from oracle import pyoracle
    from tardiis_and_claritas_amd import ClaritasForwardModel, workloads
    t_total, n_total = 0.0, 0
    for i, c in enumerate(workloads.CONFIG2_CONCENTRATIONS_MG_L):
        kw, mat, conc, _, _ = workloads.config2(c, 'plain_mie')
        m = ClaritasForwardModel(**kw)
        om = pyoracle.OracleModel(m, mat, conc)
        t0 = time.perf_counter()
        om.run(histories_per_conc, seed + i, 0, threads=0)
        t_total += time.perf_counter() - t0
        n_total += histories_per_conc
    cores = os.cpu_count() or 1
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        pass
    omp = os.environ.get('OMP_NUM_THREADS')
    threads = int(omp) if omp and omp.isdigit() else cores
    model = ''
    try:
        for line in open('/proc/cpuinfo'):
            if line.startswith('model name'):
                model = line.split(':', 1)[1].strip()
                break
    except OSError:
        pass
    return {'value': n_total / t_total, 'unit': 'photon histories/s', 'cores': threads, 'kind': 'port',
            'sample': f'{histories_per_conc} histories at each of the 7 config-2 concentrations '
                      f'({n_total} total, plain-Mie law), CPU oracle in C with OpenMP; host CPU: {model}',
            'seconds': t_total}


SURROGATE_DIR = ROOT / 'tardiis_and_claritas_amd'
SURROGATE_WEIGHTS = SURROGATE_DIR / 'surrogate_weights.npz'
SURROGATE_SPOT = SURROGATE_DIR / 'surrogate_spotcheck.npz'
SURROGATE_HELDOUT = SURROGATE_DIR / 'surrogate_heldout.npz'
SURROGATE_EVAL = SURROGATE_DIR / 'surrogate_eval.json'


def surrogate_line(seed: int) -> dict:
    """Second half of the metric, measured in this run: the trained 5-member
    BF16 ensemble (committed weights) answers the committed spot-check
    queries: 1,024 config-3 configurations with distinct optical parameters,
    whose Monte Carlo labels carry ~20,000 effective counts on S90/S0 (0.7 %
    relative error), and the earlier set of 1,024 (64 optical configurations
    x 16 transport variants, ~1 % label error), scored separately.
    queries/s is the inference kernel on 65536 device-resident queries; the
    batched Monte Carlo rate is measured on the spot-check configurations;
    the speed-up is Monte Carlo seconds per query for 1 % relative error on
    S90/S0 times queries/s.  Queries whose ensemble spread exceeds the
    committed fallback threshold are answered by Monte Carlo run now, at 1 %
    relative error."""
    import numpy as np
    from tardiis_and_claritas_amd import surrogate as sg
    from tardiis_and_claritas_amd.batched import BatchedRunner
    ev = json.loads(SURROGATE_EVAL.read_text())
    runner = BatchedRunner(0)
    sur = sg.Surrogate.load(runner.ctx, SURROGATE_WEIGHTS)
    P, S_ref, n_ref = sg.load_spot(SURROGATE_SPOT)
    pred, spread = sur.predict(sg.features(P))
    m = sg.set_metrics(pred, S_ref, n_ref)
    P2, S2, n2 = sg.load_spot(SURROGATE_HELDOUT)
    m2 = sg.set_metrics(sur.predict(sg.features(P2))[0], S2, n2)
    q_s = sur.ens.time_infer(65536, 50) * 1e-3 / 65536
    # live batched Monte Carlo rate on 256 spot-check configurations (one launch)
    sub = {k: v[:256] for k, v in P.items()}
    n_mc = np.full(256, 2_000_000, dtype=np.int64)
    sg.monte_carlo(sub, n_mc, seed + 900, runner=runner)
    sg.monte_carlo(sub, n_mc, seed + 901, runner=runner)   # one batch: its launch is timed by HIP events
    rate = float(n_mc.sum()) / (runner.last_ms * 1e-3)
    speed = sg.matched_speedup(pred, m['surrogate_rms_est'], q_s, rate)
    thr = float(ev['bench_fallback_threshold'])
    fb = sg.fallback_eval(P, pred, spread, S_ref, n_ref, {'threshold': thr}, q_s, rate, runner, seed + 902)['threshold']
    print(f'bench: surrogate {1 / q_s:.3g} queries/s; spot-check median {m["ntu_error_median"]:.4f} '
          f'p90 {m["ntu_error_p90"]:.4f} rms est {m["surrogate_rms_est"]:.4f}; Monte Carlo {rate:.3g} histories/s',
          file=sys.stderr, flush=True)
    return {
        'queries_per_s': 1.0 / q_s,
        'mc_histories_per_s': rate,
        'mc_seconds_per_query_at_1pct': speed['mc_query_s_median'],
        'speedup_vs_batched_mc': speed['speedup_at_1pct'],
        'speedup_vs_batched_mc_at_matched_rms': speed.get('speedup_at_matched_rms'),
        'meets_1pct': speed['meets_1pct'],
        'ntu_error_median': m['ntu_error_median'],
        'ntu_error_p90': m['ntu_error_p90'],
        'surrogate_rms_est': m['surrogate_rms_est'],
        'rms_vs_labels': m['rms_vs_labels'],
        'label_rms': m['label_rms'],
        'label_se_median': m['label_se_median'],
        'held_out_configs': m['n'],
        'earlier_spot_check': {k: m2[k] for k in ('n', 'ntu_error_median', 'ntu_error_p90', 'rms_vs_labels',
                                                  'label_rms', 'surrogate_rms_est')},
        'training_configs': ev['n_train'],
        'fallback': fb,
        'evaluation': 'live: committed weights (' + SURROGATE_WEIGHTS.name + ') on the committed spot-check sets ('
                      + SURROGATE_SPOT.name + ', ' + SURROGATE_HELDOUT.name + '); surrogate_rms_est = '
                      'sqrt(mean squared log error of S90/S0 against the labels - mean label variance); '
                      'speed-up = live Monte Carlo seconds for 1 % relative error on S90/S0 (median over the '
                      'spot-check queries) x live queries/s; the fallback answers queries above the spread '
                      'threshold by Monte Carlo now',
    }


def hybrid_line(h: int, seed: int) -> dict:
    """The same sweep under the reference's production event law
    (hybrid_mie_exact_rate, CLARITAS.py:9941): h histories per concentration in
    one launch, one warm-up launch and two timed ones (HIP events)."""
    import numpy as np
    from tardiis_and_claritas_amd import workloads
    from tardiis_and_claritas_amd.batched import BatchedRunner, flops_per_history_model, total_counters
    specs = [workloads.config2(c, 'hybrid_mie_exact_rate')[:3] for c in workloads.CONFIG2_CONCENTRATIONS_MG_L]
    r = BatchedRunner(0)
    r.prepare(specs)
    budgets = [h] * len(specs)
    r.run_prepared(budgets, seed + 700)
    ms, fl = [], []
    for k in range(2):
        T = r.run_prepared(budgets, seed + 701 + k)
        ms.append(r.last_ms)
        fl.append(flops_per_history_model(total_counters(T)))
    n = h * len(specs)
    t = float(np.mean(ms)) * 1e-3
    ach = float(np.mean(fl)) / t / 1e12
    print(f'bench: hybrid-law sweep {n:.3g} histories: {", ".join("%.1f" % m for m in ms)} ms', file=sys.stderr,
          flush=True)
    return {
        'value': n / t, 'unit': 'photon histories/s', 'histories_per_launch': n, 'launch_ms': t * 1e3,
        'workload': 'the config-2 sweep with the reference production law hybrid_mie_exact_rate (smooth '
                    'spheres; conditional exact-rate closure for absorbing bins), GENERIC kernel variant',
        'roofline': {'bound': 'valu', 'achieved': ach, 'peak': FP32_PEAK_TFLOPS, 'unit': 'TFLOP/s',
                     'frac': ach / FP32_PEAK_TFLOPS, 'flops_per_history': float(np.mean(fl)) / n,
                     'flop_model': 'the plain-law model above; the FP64 ray tracing of the conditional '
                                   'closure trials is not counted, so frac is a lower bound'},
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=3)
    ap.add_argument('--warmup', type=int, default=1)
    ap.add_argument('--histories', type=float, default=1e8, help='histories per concentration per step')
    ap.add_argument('--cpu-histories', type=int, default=0, help='CPU sample per concentration (0 = auto)')
    ap.add_argument('--no-cpu', action='store_true')
    ap.add_argument('--block', type=int, default=0)
    ap.add_argument('--blocks-per-cu', type=int, default=0)
    ap.add_argument('--chunk', type=int, default=0)
    ap.add_argument('--depth', type=int, default=2, help='launches in flight (separate HIP streams)')
    ap.add_argument('--queue-order', type=int, default=0, help='0 as given, 1 most expensive configuration first')
    ap.add_argument('--roofline-launches', type=int, default=2,
                    help='launches timed alone (one in flight) for the roofline figure')
    ap.add_argument('--hybrid-histories', type=float, default=1e7,
                    help='histories per concentration of the hybrid-law sweep (0 skips it)')
    ap.add_argument('--no-surrogate', action='store_true')
    ap.add_argument('--copy-calibration', type=int, default=0, metavar='MIB',
                    help='only run a plain device copy of MIB MiB (counter calibration) and exit')
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.gpus != 1:
        print(json.dumps({'error': 'the transport runs on one GPU: --gpus must be 1 (sharding history ranges '
                                   'over several GPUs is designed, not implemented)'}))
        sys.exit(2)
    import numpy as np
    from tardiis_and_claritas_amd import workloads
    from tardiis_and_claritas_amd.batched import Pipeline, flops_per_history_model, total_counters

    H = int(args.histories)
    seed = workloads.SEED
    launch = {'block': args.block, 'blocks_per_cu': args.blocks_per_cu, 'chunk': args.chunk,
              'queue_order': args.queue_order}
    specs = []
    for c in workloads.CONFIG2_CONCENTRATIONS_MG_L:
        kw, mat, conc, _, _ = workloads.config2(c, 'plain_mie')
        specs.append((kw, mat, conc))
    pipe = Pipeline(specs, depth=args.depth, launch=launch if any(launch.values()) else None)
    budgets = [H] * len(specs)
    print(f'bench: config-2 sweep, {len(specs)} concentrations x {H:.3g} histories per step, '
          f'{args.depth} launch(es) in flight', file=sys.stderr, flush=True)
    if args.warmup:
        pipe.run([budgets] * args.warmup, [seed + 1000 * (w + 1) for w in range(args.warmup)])
        print(f'bench: {args.warmup} warmup step(s) done', file=sys.stderr, flush=True)
    try:
        import torch
        sync = torch.cuda.synchronize if torch.cuda.is_available() else (lambda: None)
    except Exception:  # torch is only used for the device-wide barrier
        sync = lambda: None
    sync()
    t0 = time.perf_counter()
    results = pipe.run([budgets] * args.steps, [seed + k for k in range(args.steps)])
    sync()
    elapsed = time.perf_counter() - t0
    kernel_ms = list(pipe.last_ms)
    flops = 0.0
    for k, T in enumerate(results):
        flops += flops_per_history_model(total_counters(T))
        print(f'bench: step {k + 1}/{args.steps}: launch {kernel_ms[k]:.1f} ms', file=sys.stderr, flush=True)
    n_hist = args.steps * H * len(specs)
    value = n_hist / elapsed
    step_ms = elapsed / args.steps * 1e3
    tph = traffic_per_history()
    # Roofline of the dominant kernel: the same launch timed alone (one launch
    # in flight, so its HIP-event interval is its own duration), algorithmic
    # FLOP of that launch from its event counters over that duration.
    # (--roofline-launches 0 with --depth 1: the timed steps are such launches)
    stage = pipe.stages[0]
    r_ms, r_flops = [], []
    for k in range(args.roofline_launches):
        T = stage.run_prepared(budgets, seed + 500 + k)
        r_ms.append(stage.last_ms)
        r_flops.append(flops_per_history_model(total_counters(T)))
    if not r_ms:
        if args.depth != 1:
            raise SystemExit('--roofline-launches 0 needs --depth 1')
        r_ms = kernel_ms
        r_flops = [flops_per_history_model(total_counters(T)) for T in results]
    launch_ms = float(np.mean(r_ms))
    achieved = float(np.mean(r_flops)) / (launch_ms * 1e-3) / 1e12
    print(f'bench: roofline launches alone: {", ".join("%.1f" % m for m in r_ms)} ms', file=sys.stderr, flush=True)
    out = {
        'metric': 'photon histories/s (1 MI355X) + surrogate-vs-MC speedup at <=1% NTU error',
        'value': value,
        'unit': 'photon histories/s',
        'n_gpus': 1,
        'steps': args.steps,
        'warmup': args.warmup,
        'ms_per_step': step_ms,
        'higher_is_better': True,
        'scaling': 'weak',
        'vs_baseline': None,
        'dtype': 'fp32',
        'data': 'synthetic (seeded Philox photon histories; analytic Rouse field; lognormal PSD)',
        'config': {
            'workload': 'config2: lognormal PSD 40 bins (median 20 um, sigma_g 1.8), Rouse r-z FP32 64x128, '
                        'Rankine surface (core 3 mm, depth 5 mm), 860 nm, plain-Mie event law, '
                        '7 concentrations 10-5000 mg/L',
            'histories_per_concentration_per_step': H,
            'histories_per_step': H * len(specs),
            'launch': 'one persistent batched launch per step (hipGraph), %d launches in flight on separate '
                      'HIP streams' % args.depth,
        },
        'roofline': {
            'bound': 'valu',
            'achieved': achieved,
            'peak': FP32_PEAK_TFLOPS,
            'unit': 'TFLOP/s',
            'frac': achieved / FP32_PEAK_TFLOPS,
            'traffic': (tph * H * len(specs)) if tph else None,
            'traffic_source': 'bytes per launch = HBM bytes per history x histories per launch; bytes per '
                              'history from the committed rocprofv3 FETCH_SIZE/WRITE_SIZE runs of this '
                              'workload, copy-calibrated (' + TRAFFIC_PROFILE.name + ')',
            'kernel': 'transport_kernel<1> (plain-Mie persistent kernel)',
            'launch_ms': launch_ms,
            'launches_timed': len(r_ms),
            'time_basis': 'HIP-event duration of one launch of the step workload run alone (one launch in '
                          'flight), on the launch stream; the timed steps above keep two in flight',
            'flop_model': 'F = 16 + 30 N_flight + 16 N_null + 60 N_scat + 40 N_iface + 10 N_abs + 45 N_exit '
                          'per history (FP32; FMA=2), event counts from the launch counters; peak is the '
                          'FP32 figure, issued on the VALU',
            'flops_per_history': float(np.mean(r_flops)) / (H * len(specs)),
        },
    }
    if args.hybrid_histories > 0:
        out['hybrid_law'] = hybrid_line(int(args.hybrid_histories), seed)
    if not args.no_surrogate:
        out['surrogate'] = surrogate_line(seed)
    if not args.no_cpu:
        per = args.cpu_histories or int(2e7)
        out['cpu_baseline'] = cpu_baseline(per, seed)
    print(json.dumps(out))


if __name__ == '__main__':
    main()
