#!/usr/bin/env python3
"""Benchmark of the D3 energy + gradient + virial path on one MI355X.

Workload (default ``--config c1``): periodic triclinic cell of 2,000 atoms
(28 A cubic with +-5 % off-diagonal strain, H/C/N/O/Si/S), PBE0-D3(BJ) with
the ATM three-body term, default cutoffs (CN 40, two-body 60, three-body 40
bohr).  One step is one complete evaluation of energy, gradient and virial
(CN, weights, C6, two-body, ATM, dCN/dR chain rule) with the positions
resident in device memory.  Both arithmetic paths are timed: plain FP64 and
mixed (FP32 pair/triple arithmetic, FP64 accumulation).

Prints one JSON line.  ``value`` is the mixed path's E+grad pair-terms/s
(two-body terms of the reference's loops per second of complete
evaluations); the FP64 path, triple-terms/s, cells/s, the roofline of the
dominant (ATM) kernel and a CPU baseline (the C oracle on the host cores, on
a bounded sample of the same workload) are reported beside it.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "simple-dftd3_amd"))

METRIC = "E+grad pair-terms/s & molecules/s: mixed vs FP64 on GPU, |dE|<=1e-8 Eh/atom"
# Algorithmic operation counts per work unit (mul and add counted separately,
# an FMA as two; div, rcp, sqrt, exp and pow as one each).
OPS_PER_TRIPLE = 233    # ATM E+grad+sigma (atm.f90:733-806)
OPS_PER_PAIR_BJ = 63    # two-body BJ E+grad+sigma (rational.f90:462-497)
OPS_PER_TRIPLE_E = 60   # ATM energy only (atm.f90:500-621), the arithmetic of a pairwise triple term
KERNEL_GROUPS = ("pack", "order", "cn", "weights", "c6", "disp2", "atm_list", "atm", "cn_chain", "reduce")
PEAK_TFLOPS = {"fp64": 78.6, "mixed": 157.3}  # MI355X vector FP64 / FP32
# HBM bytes per launch of the mixed ATM kernel on the 2,000-atom cell,
# measured by two rocprofv3 --pmc runs of this script (FETCH_SIZE, then
# WRITE_SIZE, KiB per dispatch) scaled by a 256 MiB device copy (FETCH x2.00,
# WRITE x1.00); commands and figures in
# profiles/atm_kernel__config1_periodic_2000.md.  "kernel_sha" is the SHA-256
# (first 16 hex digits) of csrc/d3_kernels.hip the counters were read on: a
# report on other kernel sources gives traffic = null.
# The mixed-precision three-body kernel of a large call without a plan: one
# wavefront per block (the fourth template argument; 16 persistent blocks per
# CU) for lists with lattice images, two (8 per CU) otherwise.
def atm_kernel_name(periodic):
    return f"atm_kernel<float, true, false, {1 if periodic else 2}, false>"


ATM_TRAFFIC_C1 = {"kernel": atm_kernel_name(True), "fetch_kib": 12_889_626, "write_kib": 4_879_442,
                  "fetch_scale": 2.0, "write_scale": 1.0, "kernel_sha": "a5ee85cc7b04f681"}
KERNEL_SOURCE = os.path.join(ROOT, "simple-dftd3_amd", "csrc", "d3_kernels.hip")


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def workload(config):
    from dftd3 import synthetic
    if config == "c1":
        num, pos, lat = synthetic.periodic_c1()
        sysd = dict(numbers=num, positions=pos, lattice=lat, periodic=[True] * 3, kind="rational",
                    param=dict(synthetic.PBE0_BJ))
        return [sysd], dict(workload="config1-periodic-2000-atoms-pbe0-d3bj-atm", natoms=len(num), cells=1,
                            lattice="28 A cubic, +-5% off-diagonal strain", elements="H/C/N/O/Si/S")
    if config == "c0":
        num, pos = synthetic.molecule_c0()
        sysd = dict(numbers=num, positions=pos, kind="rational", param=dict(synthetic.PBE0_BJ))
        return [sysd], dict(workload="config0-molecule-200-atoms-pbe0-d3bj-atm", natoms=len(num), cells=1)
    if config == "c2":
        num, pos = synthetic.ellipsoid_c2()
        sysd = dict(numbers=num, positions=pos, kind="rational", param=dict(synthetic.PBE0_BJ))
        return [sysd], dict(workload="config2-ellipsoid-20000-atoms-pbe0-d3bj-atm", natoms=len(num), cells=1,
                            elements="H/C/N/O/S protein-like")
    if config == "c4":
        systems = []
        for num, pos, lat in synthetic.cells_c4():
            systems.append(dict(numbers=num, positions=pos, lattice=lat, periodic=[True] * 3, kind="rational",
                                param=dict(synthetic.PBE0_BJ)))
        return systems, dict(workload="config4-autograd-256-cells-64-512-atoms-pbe0-d3bj-atm",
                             natoms=int(sum(len(s["numbers"]) for s in systems)), cells=len(systems),
                             elements="H/C/N/O/Si")
    if config == "c3":
        systems = []
        for num, pos, param, kind in synthetic.batch_c3():
            systems.append(dict(numbers=num, positions=pos, kind="rational" if kind == 0 else "zero", param=param))
        return systems, dict(workload="config3-batch-4096-molecules-10-300-atoms-bj-zero-atm",
                             natoms=int(sum(len(s["numbers"]) for s in systems)), cells=len(systems),
                             elements="H/B/C/N/O/F/P/S/Cl")
    raise SystemExit(f"unknown config {config}")


def timed_steps(run, steps, sync):
    sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        run()
    sync()
    return (time.perf_counter() - t0) / steps


def host_cpu():
    """Model name, logical CPUs and physical cores of the host, and the CPUs
    this process may run on."""
    model, logical, cores = "unknown", 0, set()
    phys = core = None
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                key, _, val = line.partition(":")
                key, val = key.strip(), val.strip()
                if key == "model name":
                    model = val
                    logical += 1
                elif key == "physical id":
                    phys = val
                elif key == "core id":
                    core = val
                elif not key and phys is not None:
                    cores.add((phys, core))
                    phys = core = None
    except OSError:
        pass
    return dict(model=model, logical_cpus=logical or os.cpu_count(), physical_cores=len(cores) or None,
                usable_cpus=len(os.sched_getaffinity(0)))


def cpu_baseline(systems, counts, stride, threads, reps=3):
    """Oracle (FP64 C, OpenMP) on every stride-th outer atom of each loop:
    the median of `reps` timings on `threads` threads, and one timing on one
    thread of a ten times sparser sample."""
    os.environ["OMP_NUM_THREADS"] = str(threads)
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    s = systems[0]
    args_ = (s["numbers"], s["positions"], s.get("lattice"), s.get("periodic"), oracle.RATIONAL, s["param"])
    if len(s["numbers"]) <= 500:  # small systems: whole evaluations (a sample would time a few atoms)
        stride = 1
    stride1 = stride * 10 if stride > 1 else 1
    oracle.set_num_threads(threads)
    times = sorted(oracle.time_sampled_dispersion(stride, *args_) for _ in range(reps))
    t = times[len(times) // 2]
    t_full = t * stride
    cores = oracle.num_threads()
    oracle.set_num_threads(1)
    t1 = oracle.time_sampled_dispersion(stride1, *args_) * stride1
    oracle.set_num_threads(cores)
    return dict(value=float(counts[1] / t_full), unit="pair-terms/s", cores=cores, kind="port",
                sample=(f"oracle E+grad+sigma with the outer atom loops restricted to every {stride}-th atom, "
                        f"median of {reps} runs ({', '.join(f'{x:.2f}' for x in times)} s), scaled by {stride} "
                        f"to one full evaluation ({t_full:.1f} s)") if stride > 1 else
                       (f"oracle E+grad+sigma, median of {reps} full evaluations "
                        f"({', '.join(f'{x:.3f}' for x in times)} s)"),
                seconds_per_eval_estimate=t_full, host=host_cpu(),
                one_thread=dict(value=float(counts[1] / t1), seconds_per_eval_estimate=t1,
                                sample=f"every {stride1}-th outer atom, scaled" if stride1 > 1
                                else "one full evaluation"))


def source_sha(path):
    import hashlib
    with open(path, "rb") as fh:
        return hashlib.sha256(fh.read()).hexdigest()[:16]


def bench_batched(args, systems, config, spec, positions, counts):
    """Config 3: the whole batch as one planned evaluation (HIP graph of the
    kernel sequence, three-body launch shape autotuned), FP64 and mixed,
    against one FP64 evaluation per molecule."""
    import torch
    from dftd3 import library, torch_api
    nmol = len(systems)
    res, outs = {}, {}
    for prec in ("fp64", "mixed"):
        plan = torch_api.BatchPlan(spec, positions, precision=prec)
        best, times = plan.autotune(reps=2)
        log(f"{prec}: three-body shape {best} chosen; " +
            ", ".join(f"{w}w/{b}bpc {t:.2f} ms" for (w, b), t in times.items()))
        plan.capture()
        for _ in range(args.warmup):
            plan.replay()
        t = timed_steps(plan.replay, args.steps, torch.cuda.synchronize)
        if plan.overflowed():
            raise RuntimeError("three-body lists exceeded the planned capacity")
        outs[prec] = {k: v.clone() for k, v in plan.out.items()}
        res[prec] = dict(ms_per_step=t * 1e3, molecules_per_s=nmol / t, pair_terms_per_s=counts[1] / t,
                         triple_terms_per_s=counts[2] / t, shape=list(best),
                         autotune_ms={f"{w}x{b}": round(v, 4) for (w, b), v in times.items()})
        log(f"{prec}: batched graph replay {t * 1e3:.2f} ms ({nmol / t:.0f} molecules/s)")
        del plan
    # kernel breakdown of one batched evaluation (dynamic entry point, HIP events)
    ws = torch_api.workspace(positions.device)
    kernel_ms = {}
    for prec in ("fp64", "mixed"):
        out = torch_api.run_batch(spec, positions, precision=prec)
        torch.cuda.synchronize()
        library.set_kernel_timing(ws, True)
        torch_api.run_batch(spec, positions, precision=prec, out=out)
        torch.cuda.synchronize()
        kernel_ms[prec] = {k: library.get_kernel_time(ws, k)[0] for k in KERNEL_GROUPS}
        library.set_kernel_timing(ws, False)
    # one FP64 evaluation per molecule (each call plans, launches and reads
    # back its list size)
    singles = []
    for s in systems:
        sp, p = torch_api.batch_spec([s])
        pt = torch.as_tensor(p, dtype=torch.float64, device=positions.device).contiguous()
        singles.append((sp, pt, torch_api.run_batch(sp, pt, precision="fp64")))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for sp, pt, out in singles:
        torch_api.run_batch(sp, pt, precision="fp64", out=out)
    torch.cuda.synchronize()
    t_single = time.perf_counter() - t0
    log(f"fp64 per-molecule launches: {t_single * 1e3:.1f} ms for {nmol} molecules")
    e_single = torch.cat([o["energy"] for _, _, o in singles])
    natoms = spec.natoms.astype(float)
    de = ((outs["mixed"]["energy"] - outs["fp64"]["energy"]).abs().cpu().numpy() / natoms).max()
    dg = (outs["mixed"]["gradient"] - outs["fp64"]["gradient"]).abs().max().item()
    ds = (outs["mixed"]["sigma"] - outs["fp64"]["sigma"]).abs().max().item()
    d_single = ((outs["fp64"]["energy"] - e_single).abs().cpu().numpy() / natoms).max()
    mixed, fp64 = res["mixed"], res["fp64"]
    line = dict(
        metric=METRIC, value=mixed["pair_terms_per_s"], unit="pair-terms/s", n_gpus=1, steps=args.steps,
        warmup=args.warmup, ms_per_step=mixed["ms_per_step"], higher_is_better=True, scaling="weak",
        vs_baseline=None, dtype="f32", data="synthetic",
        config=dict(config, pair_terms=int(counts[1]), triple_terms=int(counts[2]), cn_terms=int(counts[0]),
                    parallelism="single-gpu", accumulation="f64", launch="hip-graph replay of a planned batch"),
        molecules_per_s=mixed["molecules_per_s"], triple_terms_per_s=mixed["triple_terms_per_s"],
        fp64_batched=fp64,
        fp64_per_molecule=dict(molecules_per_s=nmol / t_single, ms_total=t_single * 1e3,
                               pair_terms_per_s=counts[1] / t_single),
        mixed_batched_over_fp64_per_molecule=(nmol / mixed["ms_per_step"] * 1e3) / (nmol / t_single),
        mixed_over_fp64=fp64["ms_per_step"] / mixed["ms_per_step"], mixed_shape=mixed["shape"],
        mixed_autotune_ms=mixed["autotune_ms"], kernel_ms=kernel_ms,
        mixed_vs_fp64_error=dict(energy_per_atom=float(de), max_gradient=dg, max_sigma=ds),
        fp64_batched_vs_per_molecule_energy_per_atom=float(d_single),
        roofline=None,
    )
    return line


def bench_autograd(args, systems, config, spec, positions, counts):
    """Config 4: energies of 256 cells from the autograd forward pass and
    position + lattice adjoints from its backward pass (float32 inputs, mixed
    kernels), against the same with float64 inputs (FP64 kernels)."""
    import torch
    from dftd3.autograd import dispersion_energy
    lat = torch.as_tensor(np.stack([s["lattice"] for s in systems]), device=positions.device)
    res, grads = {}, {}
    # both runs see the float32-rounded coordinates, so that the difference is
    # that of the arithmetic
    pos32, lat32 = positions.to(torch.float32), lat.to(torch.float32)
    for prec, dtype in (("mixed", torch.float32), ("fp64", torch.float64)):
        pos = pos32.to(dtype).detach().requires_grad_(True)
        lt = lat32.to(dtype).detach().requires_grad_(True)

        def step():
            e = dispersion_energy(spec, pos, lt)
            gp, gl = torch.autograd.grad(e.sum(), (pos, lt))
            return e, gp, gl

        for _ in range(args.warmup):
            step()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            e, gp, gl = step()
        torch.cuda.synchronize()
        t = (time.perf_counter() - t0) / args.steps
        grads[prec] = (e.double(), gp.double(), gl.double())
        res[prec] = dict(ms_per_step=t * 1e3, cells_per_s=len(systems) / t, pair_terms_per_s=counts[1] / t,
                         triple_terms_per_s=counts[2] / t)
        log(f"{prec}: forward + backward {t * 1e3:.2f} ms ({len(systems) / t:.0f} cells/s)")
    natoms = torch.as_tensor(spec.natoms, dtype=torch.float64, device=positions.device)
    de = ((grads["mixed"][0] - grads["fp64"][0]).abs() / natoms).max().item()
    dg = (grads["mixed"][1] - grads["fp64"][1]).abs().max().item()
    dl = (grads["mixed"][2] - grads["fp64"][2]).abs().max().item()
    mixed, fp64 = res["mixed"], res["fp64"]
    line = dict(
        metric=METRIC, value=mixed["pair_terms_per_s"], unit="pair-terms/s", n_gpus=1, steps=args.steps,
        warmup=args.warmup, ms_per_step=mixed["ms_per_step"], higher_is_better=True, scaling="weak",
        vs_baseline=None, dtype="f32", data="synthetic",
        config=dict(config, pair_terms=int(counts[1]), triple_terms=int(counts[2]), cn_terms=int(counts[0]),
                    parallelism="single-gpu", accumulation="f64",
                    step="autograd forward (one E+grad+sigma evaluation) + backward (adjoints from the saved gradient and virial)"),
        molecules_per_s=mixed["cells_per_s"], triple_terms_per_s=mixed["triple_terms_per_s"], fp64=fp64,
        mixed_over_fp64=fp64["ms_per_step"] / mixed["ms_per_step"],
        mixed_vs_fp64_error=dict(energy_per_atom=de, max_position_adjoint=dg, max_lattice_adjoint=dl),
        roofline=None,
    )
    print(json.dumps(line), flush=True)
    return 0


def bench_pairwise(args, systems, config, spec, positions, counts):
    """Pairwise-resolved energies (get_pairwise_dispersion, disp.f90:403-457)
    of the workload: the two-body and three-body pair energy matrices, FP64
    and mixed, timed per kernel with HIP events, against the oracle's
    pairwise loops on the host cores (sampled outer atoms, scaled)."""
    import torch
    from dftd3 import library, torch_api
    ws = torch_api.workspace(positions.device)
    res, outs = {}, {}
    for prec in ("fp64", "mixed"):
        run = lambda: torch_api.run_pairwise(spec, positions, precision=prec, ws=ws)  # noqa: E731
        for _ in range(args.warmup):
            run()
        torch.cuda.synchronize()
        library.set_kernel_timing(ws, True)
        t = timed_steps(run, args.steps, torch.cuda.synchronize)
        all_ms = {k: library.get_kernel_time(ws, k) for k in KERNEL_GROUPS}
        library.set_kernel_timing(ws, False)
        t_plain = timed_steps(run, args.steps, torch.cuda.synchronize)
        outs[prec] = run()
        torch.cuda.synchronize()
        km = {k: v[0] / max(v[1], 1) for k, v in all_ms.items() if v[1]}
        atm_s = km.get("atm", 0.0) / 1e3
        res[prec] = dict(ms_per_step=t_plain * 1e3, ms_per_step_with_kernel_events=t * 1e3,
                         pair_terms_per_s=counts[1] / t_plain, triple_terms_per_s=counts[2] / t_plain,
                         atm_pairs_ms=atm_s * 1e3,
                         atm_pairs_tflops=counts[2] * OPS_PER_TRIPLE_E / atm_s / 1e12 if atm_s > 0 else 0.0,
                         kernel_ms=km)
        log(f"{prec}: pairwise {t_plain * 1e3:.2f} ms/step, three-body pair kernel + finalize {atm_s * 1e3:.2f} ms")
    natoms = config["natoms"]
    d2 = (outs["mixed"]["pair2"] - outs["fp64"]["pair2"]).abs().max().item()
    d3 = (outs["mixed"]["pair3"] - outs["fp64"]["pair3"]).abs().max().item()
    rows = [torch_api.blocks(outs[p], "pair2")[0].sum(dim=1) + torch_api.blocks(outs[p], "pair3")[0].sum(dim=1)
            for p in ("mixed", "fp64")]
    drow = (rows[0] - rows[1]).abs().max().item()
    mixed, fp64 = res["mixed"], res["fp64"]
    line = dict(
        metric=METRIC, value=mixed["pair_terms_per_s"], unit="pair-terms/s", n_gpus=1, steps=args.steps,
        warmup=args.warmup, ms_per_step=mixed["ms_per_step"], higher_is_better=True, scaling="weak",
        vs_baseline=None, dtype="f32", data="synthetic",
        config=dict(config, pair_terms=int(counts[1]), triple_terms=int(counts[2]), cn_terms=int(counts[0]),
                    parallelism="single-gpu", accumulation="f64",
                    step="pairwise-resolved energies: n x n two-body and three-body pair energy matrices"),
        molecules_per_s=config["cells"] / (mixed["ms_per_step"] / 1e3),
        triple_terms_per_s=mixed["triple_terms_per_s"],
        fp64=dict(value=fp64["pair_terms_per_s"], unit="pair-terms/s", ms_per_step=fp64["ms_per_step"],
                  triple_terms_per_s=fp64["triple_terms_per_s"], atm_pairs_ms=fp64["atm_pairs_ms"],
                  atm_pairs_tflops=fp64["atm_pairs_tflops"]),
        mixed_over_fp64=fp64["ms_per_step"] / mixed["ms_per_step"],
        mixed_vs_fp64_error=dict(max_pair2=d2, max_pair3=d3, max_row_sum_per_atom=drow),
        kernel_ms=dict(fp64=fp64["kernel_ms"], mixed=mixed["kernel_ms"]),
        roofline=dict(kernel="atm_pairs_kernel<float, false> + pair3_finalize_kernel", bound="valu",
                      achieved=mixed["atm_pairs_tflops"], peak=PEAK_TFLOPS["mixed"], unit="TFLOP/s",
                      frac=mixed["atm_pairs_tflops"] / PEAK_TFLOPS["mixed"], traffic=None,
                      ops_per_triple=OPS_PER_TRIPLE_E, triples_per_launch=int(counts[2]),
                      launch_ms=mixed["atm_pairs_ms"]),
    )
    if not args.no_cpu_baseline:
        log(f"cpu baseline: oracle pairwise loops on every {args.cpu_stride}-th atom, {args.cpu_threads} threads")
        os.environ["OMP_NUM_THREADS"] = str(args.cpu_threads)
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle
        s = systems[0]
        t = oracle.time_sampled_pairwise(args.cpu_stride, s["numbers"], s["positions"], s.get("lattice"),
                                         s.get("periodic"), oracle.RATIONAL, s["param"])
        t_full = t * args.cpu_stride
        line["cpu_baseline"] = dict(
            value=float(counts[1] / t_full), unit="pair-terms/s", cores=oracle.num_threads(), kind="port",
            sample=f"oracle pairwise loops with the outer atom loops restricted to every {args.cpu_stride}-th "
                   f"atom ({t:.2f} s), scaled by {args.cpu_stride} to one full evaluation ({t_full:.1f} s)",
            seconds_per_eval_estimate=t_full)
    print(json.dumps(line), flush=True)
    return 0


def bench_hessian(args, systems, config, spec, positions, counts):
    """Analytic Hessian (get_dispersion_hessian, disp.f90:236-308, FP64) of the
    workload, timed per kernel with HIP events, against the oracle's
    restatement on the host (single-threaded loops, one evaluation)."""
    import torch
    from dftd3 import library, torch_api
    ws = torch_api.workspace(positions.device)
    run = lambda: torch_api.run_hessian(spec, positions, ws=ws)  # noqa: E731
    for _ in range(args.warmup):
        run()
    torch.cuda.synchronize()
    library.set_kernel_timing(ws, True)
    t_ev = timed_steps(run, args.steps, torch.cuda.synchronize)
    km = {k: library.get_kernel_time(ws, k) for k in KERNEL_GROUPS + ("hessian",)}
    library.set_kernel_timing(ws, False)
    t = timed_steps(run, args.steps, torch.cuda.synchronize)
    out = run()
    torch.cuda.synchronize()
    km = {k: v[0] / max(v[1], 1) for k, v in km.items() if v[1]}
    log(f"fp64: energy + Hessian {t * 1e3:.2f} ms/step (Hessian kernels {km.get('hessian', 0.0):.2f} ms)")
    n = config["natoms"]
    line = dict(
        metric=METRIC, value=counts[1] / t, unit="pair-terms/s", n_gpus=1, steps=args.steps, warmup=args.warmup,
        ms_per_step=t * 1e3, higher_is_better=True, scaling="weak", vs_baseline=None, dtype="f64",
        data="synthetic",
        config=dict(config, pair_terms=int(counts[1]), triple_terms=int(counts[2]), cn_terms=int(counts[0]),
                    parallelism="single-gpu", hessian_dim=3 * n,
                    step="energy + analytic (3n x 3n) Hessian, FP64"),
        molecules_per_s=config["cells"] / t, triple_terms_per_s=counts[2] / t,
        ms_per_step_with_kernel_events=t_ev * 1e3, kernel_ms=km, roofline=None,
    )
    if not args.no_cpu_baseline:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle
        s = systems[0]
        log("cpu baseline: oracle Hessian, one evaluation")
        t0 = time.perf_counter()
        ref = oracle.hessian(s["numbers"], s["positions"], s.get("lattice"), s.get("periodic"), oracle.RATIONAL,
                             s["param"])
        t_cpu = time.perf_counter() - t0
        h = torch_api.hessian_blocks(out)[0].cpu().numpy()
        line["max_abs_diff_vs_oracle"] = float(np.abs(h - ref["hessian"]).max())
        line["cpu_baseline"] = dict(value=float(counts[1] / t_cpu), unit="pair-terms/s", cores=1, kind="port",
                                    sample=f"oracle Hessian, one full evaluation ({t_cpu:.2f} s)",
                                    seconds_per_eval=t_cpu)
    print(json.dumps(line), flush=True)
    return 0


def bench_partition(args, systems, config, spec, positions, counts):
    """Work partition (set_work_partition, partition.f90): E+grad+sigma of the
    workload's system as P parts, each part evaluated in turn on the one GPU
    through the model API (host positions in, results out), mixed precision;
    reports the time per part and checks that the parts sum to the whole."""
    from dftd3.interface import DispersionModel, RationalDampingParam
    s = systems[0]
    p = s["param"]
    param = RationalDampingParam(s6=p["s6"], s8=p["s8"], s9=p["s9"], a1=p["a1"], a2=p["a2"], alp=p["alp"])

    def model(part, nparts):
        m = DispersionModel(s["numbers"], s["positions"], s.get("lattice"),
                            None if s.get("periodic") is None else np.asarray(s["periodic"]))
        m.set_precision("mixed")
        if nparts > 1:
            m.set_work_partition(part, nparts)
        return m

    def time_model(m):
        for _ in range(args.warmup):
            m.get_dispersion(param, grad=True)
        t = timed_steps(lambda: m.get_dispersion(param, grad=True), args.steps, lambda: None)
        return t, m.get_dispersion(param, grad=True)

    t_whole, whole = time_model(model(0, 1))
    nparts = args.partition
    per_part, e_sum, g_sum = [], 0.0, 0.0
    for part in range(nparts):
        t, res = time_model(model(part, nparts))
        per_part.append(t * 1e3)
        e_sum += res["energy"]
        g_sum = g_sum + res["gradient"]
    de = abs(e_sum - whole["energy"])
    dg = float(np.abs(g_sum - whole["gradient"]).max())
    log(f"mixed: whole {t_whole * 1e3:.2f} ms, {nparts} parts: " + " ".join(f"{t:.2f}" for t in per_part)
        + f" ms; sum of parts vs whole: dE {de:.1e}, max dg {dg:.1e}")
    line = dict(
        metric=METRIC, value=counts[1] / (max(per_part) * 1e-3), unit="pair-terms/s", n_gpus=1, steps=args.steps,
        warmup=args.warmup, ms_per_step=max(per_part), higher_is_better=True, scaling="strong", vs_baseline=None,
        dtype="f32", data="synthetic",
        config=dict(config, pair_terms=int(counts[1]), triple_terms=int(counts[2]), parallelism=f"parts{nparts}",
                    step="one part of a work partition through the model API (host buffers), the slowest part"),
        whole_ms=t_whole * 1e3, part_ms=per_part, sum_of_parts_error=dict(energy=de, max_gradient=dg),
        roofline=None, cpu_baseline=None,
    )
    print(json.dumps(line), flush=True)


def bench_ewald(args, systems, config, spec, positions, counts):
    """Reciprocal-space two-body term (set_ewald_summation: low-rank C6 with
    the default tolerance, cutoff from the damping radii, direct sum over the
    reciprocal lattice) of one 3D-periodic cell, E+grad+sigma, FP64, through
    the model API, timed per kernel with HIP events; CPU baseline: the
    oracle's restatement (OpenMP over the vectors) on a bounded sample."""
    import torch
    from dftd3 import library
    from dftd3.interface import DispersionModel, RationalDampingParam
    s = systems[0]
    if not all(s.get("periodic") or [False]):
        raise SystemExit("--ewald needs a 3D-periodic workload (c1)")
    p = s["param"]
    param = RationalDampingParam(s6=p["s6"], s8=p["s8"], s9=p["s9"], a1=p["a1"], a2=p["a2"], alp=p["alp"])
    model = DispersionModel(s["numbers"], s["positions"], s["lattice"], np.asarray(s["periodic"]))
    model.set_ewald_summation(mesh=args.ewald_mesh)
    spme = args.ewald_mesh >= 0
    run = lambda: model.get_dispersion(param, grad=True)  # noqa: E731
    for _ in range(args.warmup):
        out = run()
    groups = KERNEL_GROUPS + ("lowrank", "ewald_sf", "ewald_f", "ewald_atom", "spme_setup", "spme_spread",
                              "spme_fft", "spme_conv", "spme_gather")
    library.set_kernel_timing(model._disp, True)
    t_ev = timed_steps(run, args.steps, lambda: None)
    km = {k: library.get_kernel_time(model._disp, k) for k in groups}
    library.set_kernel_timing(model._disp, False)
    t = timed_steps(run, args.steps, lambda: None)
    km = {k: v[0] / max(v[1], 1) for k, v in km.items() if v[1]}
    rank, lr_err, kcut, nvec = library.get_model_lowrank(model._disp, last_evaluation=True)
    n = config["natoms"]
    if spme:
        return report_spme(args, config, km, t, t_ev, rank, lr_err, kcut, nvec, counts)
    units = n * nvec  # (atom, vector) pairs of one pass, one vector of each +-k pair
    ops_sf = 4 * rank + 12         # phase (two complex products), 2 rank FMAs per (atom, vector)
    ops_atom = 8 * rank + 32       # 4 rank FMAs (A, B), phase, energy / dE/dCN / gradient sums
    t_sf, t_atom = km.get("ewald_sf", 0.0) / 1e3, km.get("ewald_atom", 0.0) / 1e3
    ach_atom = units * ops_atom / t_atom / 1e12 if t_atom > 0 else 0.0
    ach_sf = units * ops_sf / t_sf / 1e12 if t_sf > 0 else 0.0
    log(f"ewald: rank {rank}, kcut {kcut:.3f}, {nvec} vectors, {t * 1e3:.2f} ms/step; "
        f"sf {t_sf * 1e3:.3f} ms ({ach_sf:.2f} TFLOP/s), atom {t_atom * 1e3:.3f} ms ({ach_atom:.2f} TFLOP/s)")
    line = dict(
        metric=METRIC, value=units / t, unit="atom-vector-terms/s", n_gpus=1, steps=args.steps,
        warmup=args.warmup, ms_per_step=t * 1e3, higher_is_better=True, scaling="weak", vs_baseline=None,
        dtype="f64", data="synthetic",
        config=dict(config, parallelism="single-gpu", step="E+grad+sigma with the two-body term in reciprocal space",
                    lowrank_rank=rank, lowrank_max_error=lr_err, kcut=kcut, vectors=int(nvec),
                    atom_vector_terms=int(units), triple_terms=int(counts[2])),
        molecules_per_s=config["cells"] / t, ms_per_step_with_kernel_events=t_ev * 1e3, kernel_ms=km,
        host_path="positions uploaded and results downloaded every step (model API, 48 KB each way)",
        roofline=dict(kernel="ewald_atom_kernel<true>", bound="valu", achieved=ach_atom, peak=PEAK_TFLOPS["fp64"],
                      unit="TFLOP/s", frac=ach_atom / PEAK_TFLOPS["fp64"], traffic=None, ops_per_term=ops_atom,
                      terms_per_launch=int(units), launch_ms=t_atom * 1e3,
                      sf_kernel=dict(achieved=ach_sf, ops_per_term=ops_sf, launch_ms=t_sf * 1e3)),
    )
    if not args.no_cpu_baseline:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        os.environ["OMP_NUM_THREADS"] = str(args.cpu_threads)
        import oracle
        stride = args.cpu_stride * 10
        log(f"cpu baseline: oracle on every {stride}-th vector and atom, {args.cpu_threads} threads")
        oracle.set_sample_stride(stride)
        try:
            t0 = time.perf_counter()
            oracle.ewald(s["numbers"], s["positions"], s["lattice"], s["periodic"], oracle.RATIONAL, s["param"])
            t_cpu = (time.perf_counter() - t0) * stride
        finally:
            oracle.set_sample_stride(1)
        line["cpu_baseline"] = dict(value=float(units / t_cpu), unit="atom-vector-terms/s",
                                    cores=oracle.num_threads(), kind="port",
                                    sample=f"oracle E+grad+sigma with every loop over vectors and outer atoms "
                                           f"restricted to every {stride}-th entry, time scaled by {stride} "
                                           f"({t_cpu:.1f} s per full evaluation)", seconds_per_eval_estimate=t_cpu)
    print(json.dumps(line), flush=True)
    return 0


def report_spme(args, config, km, t, t_ev, rank, lr_err, kcut, npoints, counts):
    """The particle-mesh variant of the --ewald line: per-kernel times, no
    roofline (the transforms are hipFFT's) and no CPU baseline."""
    line = dict(
        metric=METRIC, value=config["cells"] / t, unit="cells/s", n_gpus=1, steps=args.steps, warmup=args.warmup,
        ms_per_step=t * 1e3, higher_is_better=True, scaling="weak", vs_baseline=None, dtype="f64",
        data="synthetic",
        config=dict(config, parallelism="single-gpu",
                    step="E+grad+sigma with the two-body term on a particle mesh (order-6 B-splines)",
                    lowrank_rank=rank, lowrank_max_error=lr_err, kcut=kcut, mesh_points=int(npoints),
                    triple_terms=int(counts[2])),
        ms_per_step_with_kernel_events=t_ev * 1e3, kernel_ms=km, roofline=None,
        host_path="positions uploaded and results downloaded every step (model API, 48 KB each way)")
    log(f"spme: rank {rank}, kcut {kcut:.3f}, {npoints} mesh points, {t * 1e3:.2f} ms/step")
    print(json.dumps(line), flush=True)
    return 0


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", default="c1")
    ap.add_argument("--cpu-stride", type=int, default=10)
    ap.add_argument("--calibration-copy-mib", type=int, default=0,
                    help="run one device copy of this size first (reference kernel for FETCH/WRITE_SIZE counters)")
    ap.add_argument("--cpu-threads", type=int, default=len(os.sched_getaffinity(0)),
                    help="threads of the CPU baseline (default: every CPU this process may use)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-oracle-check", action="store_true",
                    help="skip the errors of both paths against one full FP64 oracle evaluation (default "
                         "workload only)")
    ap.add_argument("--no-extra", action="store_true",
                    help="report only the 2,000-atom cell (no 20,000-atom ellipsoid and 4,096-molecule batch keys)")
    ap.add_argument("--pairwise", action="store_true",
                    help="time the pairwise-resolved energies (pair energy matrices) of the workload instead")
    ap.add_argument("--hessian", action="store_true",
                    help="time the energy and analytic Hessian of the workload instead (FP64)")
    ap.add_argument("--ewald", action="store_true",
                    help="time the workload with the two-body term summed in reciprocal space instead (FP64)")
    ap.add_argument("--partition", type=int, default=1,
                    help="time the P parts of the reference's work partition, one after another (P > 1)")
    ap.add_argument("--ewald-mesh", type=int, default=-1,
                    help="with --ewald: < 0 sums the reciprocal vectors directly, >= 0 uses a particle mesh "
                         "(0: derived from the cutoff)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU; --gpus {args.gpus} is not supported (each box has a single MI355X; "
              "the sharding of this path over several GPUs is described in DESIGN.md)", file=sys.stderr)
        return 2

    import torch
    from dftd3 import library, torch_api

    dev = torch.device("cuda:0")
    if args.calibration_copy_mib > 0:
        n = args.calibration_copy_mib * (1 << 20) // 8
        src = torch.ones(n, dtype=torch.float64, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        log(f"calibration copy: {args.calibration_copy_mib} MiB read + {args.calibration_copy_mib} MiB written")
        del src, dst
    systems, config = workload(args.config)
    spec, pos = torch_api.batch_spec(systems)
    positions = torch.as_tensor(pos, dtype=torch.float64, device=dev).contiguous()
    counts = torch_api.count_terms(spec, positions).sum(axis=0)
    log(f"{config['workload']}: CN terms {counts[0]}, pair terms {counts[1]}, triple terms {counts[2]}")
    if args.config == "c4":
        return bench_autograd(args, systems, config, spec, positions, counts)
    if args.pairwise:
        return bench_pairwise(args, systems, config, spec, positions, counts)
    if args.hessian:
        return bench_hessian(args, systems, config, spec, positions, counts)
    if args.ewald:
        return bench_ewald(args, systems, config, spec, positions, counts)
    if args.partition > 1:
        return bench_partition(args, systems, config, spec, positions, counts)
    if args.config == "c3":
        print(json.dumps(bench_batched(args, systems, config, spec, positions, counts)), flush=True)
        return 0
    line = bench_single(args, systems, config, spec, positions, counts,
                        cpu=not args.no_cpu_baseline, check=args.config == "c1" and not args.no_oracle_check)
    if args.config == "c1" and not args.no_extra:
        line["config2"] = extra_config(args, "c2")
        line["config3"] = extra_config(args, "c3")
    print(json.dumps(line), flush=True)
    return 0


def extra_config(args, name):
    """The 20,000-atom ellipsoid (single system, both paths, ATM roofline) or
    the 4,096-molecule batch (planned, graph-replayed) as a summary for the
    default report line."""
    import torch
    from dftd3 import torch_api
    systems, config = workload(name)
    spec, pos = torch_api.batch_spec(systems)
    positions = torch.as_tensor(pos, dtype=torch.float64, device="cuda:0").contiguous()
    counts = torch_api.count_terms(spec, positions).sum(axis=0)
    log(f"{config['workload']}: CN terms {counts[0]}, pair terms {counts[1]}, triple terms {counts[2]}")
    if name == "c3":
        b = bench_batched(args, systems, config, spec, positions, counts)
        keep = ("ms_per_step", "molecules_per_s", "triple_terms_per_s", "fp64_batched", "fp64_per_molecule",
                "mixed_batched_over_fp64_per_molecule", "mixed_over_fp64", "mixed_shape", "mixed_vs_fp64_error",
                "fp64_batched_vs_per_molecule_energy_per_atom")
        out = dict(workload=config["workload"], unit="molecules/s", value=b["molecules_per_s"],
                   pair_terms_per_s=b["value"], **{k: b[k] for k in keep})
        out["config"] = b["config"]
        return out
    b = bench_single(args, systems, config, spec, positions, counts, cpu=False, check=False)
    keep = ("ms_per_step", "molecules_per_s", "triple_terms_per_s", "fp64", "mixed_over_fp64", "mixed_vs_fp64_error",
            "roofline")
    out = dict(workload=config["workload"], unit="pair-terms/s", value=b["value"], **{k: b[k] for k in keep})
    out["config"] = b["config"]
    return out


def bench_single(args, systems, config, spec, positions, counts, cpu, check):
    """One system (the 2,000-atom cell or the 20,000-atom ellipsoid): both
    paths timed per evaluation and per kernel (HIP events on the launch
    stream), the ATM roofline, and optionally the CPU baseline and the
    errors against one full oracle evaluation."""
    import torch
    from dftd3 import library, torch_api
    ws = torch_api.workspace(positions.device)
    outs = {}
    res = {}
    for prec in ("fp64", "mixed"):
        out = torch_api.run_batch(spec, positions, precision=prec)
        run = lambda: torch_api.run_batch(spec, positions, precision=prec, out=out)  # noqa: E731
        for _ in range(args.warmup):
            run()
        torch.cuda.synchronize()
        library.set_kernel_timing(ws, True)
        t = timed_steps(run, args.steps, torch.cuda.synchronize)
        atm_ms, atm_calls = library.get_kernel_time(ws, "atm")
        all_ms = {k: library.get_kernel_time(ws, k) for k in KERNEL_GROUPS}
        library.set_kernel_timing(ws, False)
        t_plain = timed_steps(run, args.steps, torch.cuda.synchronize)  # without event records
        outs[prec] = {k: v.clone() for k, v in out.items()}
        atm_avg = atm_ms / max(atm_calls, 1) / 1e3
        achieved = counts[2] * OPS_PER_TRIPLE / atm_avg / 1e12 if atm_avg > 0 else 0.0
        res[prec] = dict(ms_per_step=t_plain * 1e3, ms_per_step_with_kernel_events=t * 1e3,
                         pair_terms_per_s=counts[1] / t_plain, triple_terms_per_s=counts[2] / t_plain,
                         cells_per_s=config["cells"] / t_plain, atm_kernel_ms=atm_avg * 1e3,
                         atm_tflops=achieved, kernel_ms={k: v[0] / max(v[1], 1) for k, v in all_ms.items()})
        log(f"{prec}: {t_plain * 1e3:.2f} ms/step, ATM kernel {atm_avg * 1e3:.2f} ms ({achieved:.2f} TFLOP/s)")
    natoms = config["natoms"]
    de = (outs["mixed"]["energy"] - outs["fp64"]["energy"]).abs().max().item()
    dg = (outs["mixed"]["gradient"] - outs["fp64"]["gradient"]).abs().max().item()
    ds = (outs["mixed"]["sigma"] - outs["fp64"]["sigma"]).abs().max().item()
    mixed, fp64 = res["mixed"], res["fp64"]
    traffic, traffic_note = None, None
    if config["workload"].startswith("config1"):
        t = ATM_TRAFFIC_C1
        if t["kernel_sha"] == source_sha(KERNEL_SOURCE):
            traffic = (t["fetch_kib"] * t["fetch_scale"] + t["write_kib"] * t["write_scale"]) * 1024.0
            traffic_note = f"rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE on d3_kernels.hip sha256 {t['kernel_sha']}"
        else:
            traffic_note = "not measured on the current kernel source"
    line = dict(
        metric=METRIC, value=mixed["pair_terms_per_s"], unit="pair-terms/s", n_gpus=1, steps=args.steps,
        warmup=args.warmup, ms_per_step=mixed["ms_per_step"], higher_is_better=True, scaling="weak",
        vs_baseline=None, dtype="f32", data="synthetic",
        config=dict(config, pair_terms=int(counts[1]), triple_terms=int(counts[2]), cn_terms=int(counts[0]),
                    parallelism="single-gpu", accumulation="f64"),
        molecules_per_s=mixed["cells_per_s"], triple_terms_per_s=mixed["triple_terms_per_s"],
        fp64=dict(value=fp64["pair_terms_per_s"], unit="pair-terms/s", ms_per_step=fp64["ms_per_step"],
                  molecules_per_s=fp64["cells_per_s"], triple_terms_per_s=fp64["triple_terms_per_s"],
                  atm_kernel_ms=fp64["atm_kernel_ms"], atm_tflops=fp64["atm_tflops"],
                  atm_frac_of_fp64_peak=fp64["atm_tflops"] / PEAK_TFLOPS["fp64"]),
        mixed_over_fp64=fp64["ms_per_step"] / mixed["ms_per_step"],
        mixed_vs_fp64_error=dict(energy_per_atom=de / natoms, max_gradient=dg, max_sigma=ds),
        kernel_ms=dict(fp64=fp64["kernel_ms"], mixed=mixed["kernel_ms"]),
        roofline=dict(kernel=atm_kernel_name(systems[0].get("lattice") is not None), bound="valu", achieved=mixed["atm_tflops"],
                      peak=PEAK_TFLOPS["mixed"], unit="TFLOP/s", frac=mixed["atm_tflops"] / PEAK_TFLOPS["mixed"],
                      traffic=traffic, ops_per_triple=OPS_PER_TRIPLE, triples_per_launch=int(counts[2]),
                      launch_ms=mixed["atm_kernel_ms"], traffic_unit="bytes per launch (HBM read + write)",
                      traffic_source=traffic_note, kernel_sha=source_sha(KERNEL_SOURCE)),
    )
    if cpu:
        log(f"cpu baseline: oracle on every {args.cpu_stride}-th atom, {args.cpu_threads} threads, median of 3")
        line["cpu_baseline"] = cpu_baseline(systems, counts, args.cpu_stride, args.cpu_threads)
    if check:
        line["error_vs_oracle"] = oracle_errors(systems, outs, args.cpu_threads)
    return line




def oracle_errors(systems, outs, threads):
    """Energy (per atom), gradient and virial of both paths against one full
    FP64 oracle evaluation of the (single-system) workload."""
    os.environ["OMP_NUM_THREADS"] = str(threads)
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    s = systems[0]
    log(f"oracle: one full evaluation on {threads} threads")
    t0 = time.perf_counter()
    ref = oracle.dispersion(s["numbers"], s["positions"], s.get("lattice"), s.get("periodic"), oracle.RATIONAL,
                            s["param"])
    t_ref = time.perf_counter() - t0
    n = len(s["numbers"])
    res = dict(oracle_seconds=t_ref)
    for prec, out in outs.items():
        e = float(out["energy"].reshape(-1)[0].item())
        g = out["gradient"].reshape(-1, 3).cpu().numpy()
        sg = out["sigma"].reshape(3, 3).cpu().numpy()
        res[prec] = dict(energy_per_atom=abs(e - ref["energy"]) / n, energy_rel=abs(e - ref["energy"]) / abs(ref["energy"]),
                         max_gradient=float(np.abs(g - ref["gradient"]).max()),
                         max_sigma=float(np.abs(sg - ref["sigma"]).max()))
    return res


if __name__ == "__main__":
    sys.exit(main())
