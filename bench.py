"""Elemental-generation benchmark on one MI355X.

Workload (default): 32^3 x 64 lattice, Nev = 200, 33 momenta (|n|^2 <= 4),
13 derivative combinations (num_nabla = 2), Haar-random SU(3) links,
orthonormal complex64 eigenvectors -- the shape on which the mixed form is
required to reach >= 5x the FP64 form at <= 1e-6 normwise error.

One step = elementals of every timeslice (calc_all over all Lt timeslices)
with links and eigenvectors already resident in HBM and the output written
to HBM.  value = Lt * P * D * Ne^2 / seconds per step of the mixed form; the
FP64 form is measured on the same inputs and reported beside it, together
with the mixed form's error against the FP64 form.

Prints one JSON line.  Only one GPU is used; --gpus above 1 is refused.
"""

from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "elementals/s (t x p x d x Nev^2) mixed vs FP64 on 1 GPU at <=1e-6 rel err"
PEAK = {"fp64": 78.6, "mixed": 157.3}  # TFLOP/s dense MFMA: FP64, FP32
PEAK_FP16 = 2500.0  # TFLOP/s dense FP16 MFMA
HBM_PEAK_GBS = 8000.0
# HBM bytes per contraction launch read from rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE passes of
# "--workload C --steps 1 --warmup 0 --calibrate-copy 4096", scaled by the factors the 4 GiB copy
# in the same run gives (2.00 x FETCH_SIZE, 1.00 x WRITE_SIZE); the summary and the commands are in
# profiles/elemental_32c64_nev200_hbm_traffic.md.  Keyed by (workload, form, mixed kernel id
# (ED_Q_MIXED_KERNEL; None for FP64), strip height, item order) -- the tile order changes the bytes.
TRAFFIC_MEASURED = {
    ("C", "mixed", 3, 4, 0): (127743801.0 * 2.0 + 3239083.4) * 1024,  # contract_mixed_ws_kernel<3>, 8 timeslices
    ("C", "fp64", None, 2, 0): (434679524.2 * 2.0 + 2751779.8) * 1024,  # contract_kernel<double, 3>, 8 timeslices
    # contract_mixed16_kernel<2> (the tuned mixed kernel), 8 timeslices:
    # profiles/elemental_32c64_nev200_tuned_kernel_trace.md
    ("C", "mixed", 5, 2, 1): (180556618.2 * 2.0 + 3430244.9) * 1024,
}
TRAFFIC_SOURCE = {
    ("C", "mixed", 3, 4, 0): "profiles/elemental_32c64_nev200_hbm_traffic.md",
    ("C", "fp64", None, 2, 0): "profiles/elemental_32c64_nev200_hbm_traffic.md",
    ("C", "mixed", 5, 2, 1): "profiles/elemental_32c64_nev200_tuned_kernel_trace.md",
}
# the library's ED_Q_MIXED_KERNEL answer -> the kernel the mixed form runs
MIXED_KERNEL = {
    0: "contract_kernel<double, 3>",
    1: "contract_kernel<float, 3>",
    2: "contract_mixed_f16_kernel<3>",
    3: "contract_mixed_ws_kernel<3>",
    4: "contract_mixed_ws_kernel<3> (64-bit workspace addresses)",
    5: "contract_mixed16_kernel<2>",
}

WORKLOADS = {
    # name: (latt [Lx,Ly,Lz,Lt], Ne, num_nabla, max |n|^2, links)
    "C": ([32, 32, 32, 64], 200, 2, 4, "haar"),
    "B": ([16, 16, 16, 128], 100, 2, 4, "smooth"),
    "small": ([16, 16, 16, 8], 64, 2, 4, "haar"),
    # 7.2 GB of FP32 workspace per timeslice: the mixed contraction's 64-bit-address form
    "L": ([48, 48, 48, 8], 200, 2, 4, "smooth"),
}
# batched uneven form: three lattices in one graph launch, num_nabla 1 (4
# derivative combinations), 19 momenta (|n|^2 <= 2), smooth links eps = 0.3
BATCHED = {
    "D": [([16, 16, 16, 32], 64), ([24, 24, 24, 48], 128), ([32, 32, 32, 64], 256)],
    "D-small": [([8, 8, 8, 4], 16), ([12, 12, 12, 6], 32), ([16, 16, 16, 8], 64)],
}
# differentiation form: lattice, Nev, num_nabla (4 derivative combinations), the 7 momenta |n|^2 <= 1
DIFFERENTIATION = {"E": ([24, 24, 24, 64], 128, 1)}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def mixed_precision(kernel):
    """What runs in which precision in the mixed contraction kernels 2-5 (ED_Q_MIXED_KERNEL)."""
    h16 = kernel == 5
    return {
        "stencil": "FP64 arithmetic, " + ("FP16 hi/lo pair storage of W 2^k (22 bits)" if h16 else "complex64 storage"),
        "split_sum": "FP16 MFMA on the hi/lo pairs of both operands, FP32 sums" if h16 else "FP32 MFMA",
        "phase_contraction": "FP16 MFMA on a three-term hi/lo split of phase and Q, FP32 partial sums per 32 sites",
        "reduction_and_phase_table": "FP64",
    }


def blockwise_error(E, Eref, tables=False):
    """Error of E against Eref (torch tensors [Lt, D, P, Ne, Ne]) per (d, p) block.

    Returns (max, mean) over blocks of the normwise max error
    max_{t,e,f}|dE| / max_{t,e,f}|Eref|; with ``tables`` also the [D][P]
    tables of that error and of the Frobenius error ||dE||_F / ||Eref||_F.
    """
    import torch

    nd, npm = E.shape[1], E.shape[2]
    dmax = torch.zeros((nd, npm), dtype=torch.float64, device=E.device)
    rmax = torch.zeros_like(dmax)
    dfro = torch.zeros_like(dmax)
    rfro = torch.zeros_like(dmax)
    for d in range(nd):  # one derivative at a time bounds the temporary memory
        diff = (E[:, d] - Eref[:, d]).abs().permute(1, 0, 2, 3).reshape(npm, -1)
        ref = Eref[:, d].abs().permute(1, 0, 2, 3).reshape(npm, -1)
        dmax[d], rmax[d] = diff.amax(-1), ref.amax(-1)
        dfro[d], rfro[d] = diff.square().sum(-1), ref.square().sum(-1)
    row = rmax.amax(1, keepdim=True)
    den = torch.where(rmax >= 1e-8 * row, rmax, row.expand_as(rmax))
    err = dmax / den
    if not tables:
        return float(err.max()), float(err.mean())
    fro = (dfro / torch.where(rmax >= 1e-8 * row, rfro, rfro.amax(1, keepdim=True).expand_as(rfro))).sqrt()

    def table(x):
        return [[float(f"{v:.3g}") for v in r] for r in x.cpu().tolist()]

    return float(err.max()), float(err.mean()), table(err), table(fro)


def host_description():
    """CPU model, thread count and the BLAS numpy runs on (threadpoolctl's view)."""
    model = platform.processor() or platform.machine()
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    blas = []
    try:
        from threadpoolctl import threadpool_info

        blas = [f"{i.get('internal_api')} {i.get('version')} ({i.get('threading_layer', '')})"
                for i in threadpool_info() if i.get("user_api") == "blas"]
    except Exception:
        pass
    return model, cores, blas


def cpu_baseline(latt, Ne, num_nabla, moms, seconds_target=20.0):
    """Reference loop nest (oracle/reference_port.py) timed on the host cores.

    A full timeslice takes tens of minutes on a CPU, so the sample times a few
    hop applications and a few (split, momentum) contractions of timeslice 0
    and extrapolates with the exact operation counts of elemental.py:290-338:
    per timeslice sum_d 2^{|d|} splits, each recomputing both displaced
    vectors (|d| hops in total) and contracting once per momentum.
    """
    import numpy as np

    from threadpoolctl import threadpool_limits

    from easydistillation_amd import synthetic as S
    from oracle.elemental_oracle import derivative_list
    from oracle.reference_port import ReferencePortGenerator

    model, avail, _ = host_description()
    # every host core this process may use (the box's CPU share), set explicitly for the BLAS
    threads = int(os.environ.get("OMP_NUM_THREADS") or avail)
    with threadpool_limits(limits=threads):
        return _cpu_sample(latt, Ne, num_nabla, moms, seconds_target, threads, model, np, S, derivative_list,
                           ReferencePortGenerator)


def _cpu_sample(latt, Ne, num_nabla, moms, seconds_target, threads, model, np, S, derivative_list,
                ReferencePortGenerator):
    Lx, Ly, Lz, _ = latt
    one = [Lx, Ly, Lz, 1]
    rng = np.random.default_rng(0)
    U = S.gauge_np(rng, one, "haar")
    V = ((rng.standard_normal((1, Ne, Lz, Ly, Lx, 3)) + 1j * rng.standard_normal((1, Ne, Lz, Ly, Lx, 3))) / np.sqrt(6 * Lx * Ly * Lz)).astype(np.complex64)
    gen = ReferencePortGenerator(one, U, V, num_nabla, moms)
    Vt = V[0].astype(np.complex128)
    dl = derivative_list(num_nabla)
    n_splits = sum(2 ** len(d) for d in dl)
    n_hops = sum(len(d) * 2 ** len(d) for d in dl)  # each split applies |d| hops in total
    ph = gen._phase[tuple(moms[1 % len(moms)])]
    # hops
    t_hop, n_h, W = 0.0, 0, Vt
    t0 = time.perf_counter()
    while True:
        W = gen._nD(Vt, U[:, 0], (n_h % 3,))
        n_h += 1
        t_hop = time.perf_counter() - t0
        if n_h >= 2 and t_hop > seconds_target / 2:
            break
        if n_h >= 16:
            break
    # contractions
    t_con, n_c = 0.0, 0
    t0 = time.perf_counter()
    while True:
        np.einsum("zyx,ezyxc,fzyxc->ef", ph, W.conj(), Vt, optimize=True)
        n_c += 1
        t_con = time.perf_counter() - t0
        if n_c >= 2 and t_con > seconds_target / 2:
            break
        if n_c >= 16:
            break
    per_t = n_hops * (t_hop / n_h) + n_splits * len(moms) * (t_con / n_c)
    elementals_t = len(moms) * len(dl) * Ne * Ne
    _, _, blas = host_description()
    return {
        "value": elementals_t / per_t,
        "unit": "elementals/s",
        "cores": threads,
        "kind": "port",
        "cpu": model,
        "blas": blas,
        "numpy": np.__version__,
        "threads_set": f"threadpoolctl limit {threads} (BLAS and OpenMP pools)",
        "sample": (
            f"timeslice of {Lx}x{Ly}x{Lz}, Ne={Ne}: {n_h} hop applications ({t_hop / n_h:.2f} s each) and "
            f"{n_c} split contractions ({t_con / n_c:.2f} s each) of the reference loop nest, extrapolated "
            f"to its {n_hops} hops + {n_splits}x{len(moms)} contractions per timeslice ({per_t:.0f} s/t)"
        ),
    }


def bench_differentiation(args):
    """Differentiation form: L = sum_{t,d,p} |sum_k w_k E[t,d,p,k,k]|^2 and dL/dU, dL/dV, dL/dw.

    One step = the forward FP64 form over every timeslice plus the backward
    through the hand-written adjoint (ed_backward, which recomputes the
    displaced vectors) and the profile-trace adjoint.  value = Lt P D Ne^2 /
    seconds per step.  The roofline object covers the contraction adjoint
    (adj_contract_kernel, FP64 vector arithmetic): per (t, site, row, column)
    a phase sum of 2 P complex products (16 P flops) for each of the D
    derivatives and, per split term and colour, two complex products and
    their scaled accumulation (20 flops): V3 Ne^2 (16 P D + 60 S) flops per
    timeslice, S = sum over derivatives of 2^|d| split terms.  The reference
    has no reverse pass, so there is no CPU baseline.
    """
    import torch

    from easydistillation_amd import _native
    from easydistillation_amd import synthetic as S
    from easydistillation_amd.autograd import ElementalEngine, profiled_loss
    from easydistillation_amd.insertion import derivative, num_derivative

    dev = torch.device("cuda:0")
    latt, Ne, num_nabla = DIFFERENTIATION[args.workload]
    Lx, Ly, Lz, Lt = latt
    moms = S.unit_momenta()
    P, D = len(moms), num_derivative(num_nabla)
    nsplit = sum(2 ** len(derivative(d)) for d in range(D))
    t0 = time.time()
    U = S.gauge_torch(latt, "haar", seed=0, device=dev)
    V = S.eigvecs_torch(latt, Ne, seed=1, device=dev, log=log).to(torch.complex128)
    gen = torch.Generator(device=dev).manual_seed(2)
    w = torch.rand(Ne, dtype=torch.float64, device=dev, generator=gen) + 0.5
    log(f"inputs in {time.time() - t0:.1f} s")
    eng = ElementalEngine(latt, Ne, num_nabla, moms)
    eng.ctx.set_option(_native.ED_OPT_HOP_LDS, args.hop_lds)
    Ug, Vg, wg = (x.clone().requires_grad_() for x in (U, V, w))
    del U, V

    def step():
        loss = profiled_loss(Ug, Vg, wg, eng)
        return torch.autograd.grad(loss, [Ug, Vg, wg])

    for i in range(args.warmup):
        step()
    torch.cuda.synchronize()
    eng.ctx.set_option(_native.ED_OPT_PROFILE, 1)
    eng.ctx.profile_reset()
    t0 = time.perf_counter()
    for i in range(args.steps):
        step()
        torch.cuda.synchronize()
        log(f"differentiation step {i} done at {time.perf_counter() - t0:.2f} s")
    dt = (time.perf_counter() - t0) / args.steps
    eng.ctx.set_option(_native.ED_OPT_PROFILE, 0)
    a_ms, a_n = eng.ctx.profile_read(_native.ED_KERNEL_ADJOINT)
    c_ms, c_n = eng.ctx.profile_read(_native.ED_KERNEL_CONTRACT)
    h_ms, _ = eng.ctx.profile_read(_native.ED_KERNEL_STENCIL)
    adj_flops = float(Lx * Ly * Lz) * Ne * Ne * (16 * P * D + 60 * nsplit) * Lt * args.steps / max(a_n, 1)
    a_avg = a_ms / max(a_n, 1)
    achieved = adj_flops / (a_avg * 1e-3) / 1e12
    line = {
        "metric": "elementals/s through forward + backward (dL/dU, dL/dV, dL/dw), FP64, 1 GPU",
        "value": Lt * P * D * Ne * Ne / dt,
        "unit": "elementals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded Haar SU(3) links, orthonormal eigenvectors, weights U(0.5, 1.5))",
        "config": {
            "workload": f"{Lx}^3x{Lt}" if Lx == Ly == Lz else f"{Lx}x{Ly}x{Lz}x{Lt}",
            "Nev": Ne,
            "momenta": P,
            "derivatives": D,
            "form": "differentiation",
            "elementals_per_step": Lt * P * D * Ne * Ne,
        },
        "roofline": {
            "bound": "valu",
            "achieved": achieved,
            "peak": PEAK["fp64"],
            "unit": "TFLOP/s",
            "frac": achieved / PEAK["fp64"],
            "traffic": None,
            "kernel": "adj_contract_kernel",
            "avg_launch_ms": a_avg,
            "flops_per_launch": adj_flops,
            "flops_basis": "V3 Ne^2 (16 P D + 60 S) FP64 flops per timeslice (phase sums and split-operand gradients)",
        },
        "forward_contract_ms_per_step": c_ms / args.steps,
        "stencil_ms_per_step": h_ms / args.steps,
        "adjoint_contract_ms_per_step": a_ms / args.steps,
        "cpu_baseline": None,
        "cpu_baseline_note": "the reference has no reverse pass",
    }
    print(json.dumps(line), flush=True)


def bench_batched(args):
    """Batched uneven form: every timeslice of three lattices per graph launch (mixed form).

    One step = one replay of the batch's hipGraph (staging and hops of each
    lattice, then one persistent grouped contraction).  value = sum over
    lattices of Lt P D Ne^2 / seconds per step.  Each lattice is also timed
    alone (its own tuned batch) for the per-lattice figures, and every batch
    output is compared with the FP64 form of the same lattice.
    """
    import torch

    import easydistillation_amd as ed
    from easydistillation_amd import _native
    from easydistillation_amd import synthetic as S
    from easydistillation_amd.batch import ElementalBatch

    dev = torch.device("cuda:0")
    moms = S.momenta_by_norm(2)
    num_nabla = 1
    D = ed.num_derivative(num_nabla)
    P = len(moms)
    gens = []
    for i, (latt, Ne) in enumerate(BATCHED[args.workload]):
        t0 = time.time()
        U = S.gauge_torch(latt, "smooth", 0.3, seed=21 + i, device=dev)
        torch.cuda.synchronize()
        log(f"lattice {latt}: links in {time.time() - t0:.1f} s")
        V = S.eigvecs_torch(latt, Ne, seed=31 + i, device=dev, log=log)
        g = ed.ElementalGenerator(latt, None, None, num_nabla, moms, Ne=Ne, mode="mixed")
        g.load_arrays(U, V)
        del U, V
        gens.append(g)
        log(f"lattice {latt} Ne={Ne}: inputs in {time.time() - t0:.1f} s")
    torch.cuda.synchronize()

    batch = ElementalBatch(gens)
    t0 = time.time()
    tuned_ms = batch.tune(1)
    shape = {
        "grid_per_cu": batch.option(_native.ED_BOPT_GRID_PER_CU),
        "order": batch.option(_native.ED_BOPT_ORDER),
        "strip": batch.option(_native.ED_BOPT_STRIP),
        "fp16_split_sum": batch.option(_native.ED_BOPT_H16),
    }
    log(f"autotuned in {time.time() - t0:.1f} s: {tuned_ms:.1f} ms per run, {shape}")
    if shape["fp16_split_sum"] == 1:
        # the contexts' flop queries then describe the FP16 split-sum kernel the batch runs
        for g in gens:
            g.set_option(_native.ED_OPT_MIXED_WS, 2)
    for i in range(args.warmup):
        batch.run()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        batch.batch.run()
        batch.batch.synchronize()
        log(f"batched step {i} done at {time.perf_counter() - t0:.2f} s")
    dt = (time.perf_counter() - t0) / args.steps
    elems = batch.elementals()
    total = sum(elems)

    per = []
    f32 = f16 = 0
    for g, e in zip(gens, elems):
        nt = g.latt_size[3]
        f32_t = g.context.query(_native.ED_Q_MIXED_F32_FLOPS_PER_T)
        f16_t = g.context.query(_native.ED_Q_MIXED_F16_FLOPS_PER_T)
        f32 += f32_t * nt
        f16 += f16_t * nt
        single = ElementalBatch([g])
        single.tune(1)
        ms = single.time(max(1, args.steps))
        per.append(
            {
                "lattice": "x".join(str(v) for v in g.latt_size),
                "Nev": g.Ne,
                "elementals_per_run": e,
                "ms_alone": ms,
                "elementals_per_s_alone": e / (ms * 1e-3),
            }
        )
        del single
        log(f"alone {per[-1]}")
    equiv = f32 + f16 * PEAK["mixed"] / PEAK_FP16

    # error of each batch output against the FP64 form of the same lattice
    outs = batch.run()
    errs = []
    for g, o in zip(gens, outs):
        ref = g.calc_all(mode="fp64", out=torch.empty_like(o))
        errs.append(blockwise_error(o, ref)[0])
        del ref
    log(f"mixed batch vs fp64 blockwise error per lattice: {errs}")

    line = {
        "metric": "elementals/s (t x p x d x Nev^2), batched uneven lattices, mixed form, 1 GPU",
        "value": total / dt,
        "unit": "elementals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32+fp16x3",
        "precision": mixed_precision(5 if shape["fp16_split_sum"] == 1 else 3),
        "data": "synthetic (seeded smooth SU(3) links eps=0.3, orthonormal complex64 eigenvectors)",
        "config": {
            "workload": " + ".join(f"{p['lattice']}/Nev{p['Nev']}" for p in per),
            "momenta": P,
            "derivatives": D,
            "form": "mixed",
            "launch": "one hipGraph replay per step",
            "tuned_launch_shape": shape,
        },
        "roofline": {
            "bound": "mfma",
            "achieved": equiv / dt / 1e12,
            "peak": PEAK["mixed"],
            "unit": "TFLOP/s",
            "frac": equiv / dt / 1e12 / PEAK["mixed"],
            "traffic": None,
            "kernel": "whole graph (stencils + "
            + ("contract_mixed16_item_kernel" if shape["fp16_split_sum"] == 1 else "contract_mixed_group_kernel")
            + ")",
            "flops_basis": "executed FP32 MFMA flops plus FP16 MFMA flops at the 157.3:2500 peak ratio, over the whole step",
        },
        "per_lattice": per,
        "mixed_vs_fp64_error_max": max(errs),
        "mixed_vs_fp64_error_per_lattice": errs,
    }
    print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="C", choices=sorted(WORKLOADS) + sorted(BATCHED) + sorted(DIFFERENTIATION))
    ap.add_argument("--forms", default="mixed,fp64")
    ap.add_argument("--mixed-ws", type=int, default=1, choices=(0, 1, 2),
                    help="mixed form's FP16 stage-2 kernel: 1 warp-specialized with FP32 stage 1 (default), "
                    "2 warp-specialized with FP16 stage 1, 0 uniform waves")
    ap.add_argument("--tbatch", type=int, default=0, help="timeslices per contraction launch (0: the library's choice)")
    ap.add_argument("--hop-lds", type=int, default=-1, choices=(-1, 0, 1),
                    help="derivative hops: -1 LDS-staged for 16-byte workspace elements (library default), "
                    "1 LDS-staged for every form, 0 one wave per site")
    ap.add_argument("--no-tune", action="store_true",
                    help="skip the autotuner (easydistillation_amd.tune) and run the library defaults / --mixed-ws")
    ap.add_argument("--tune-cache", default=None, help="JSON table of tuned shapes (default: the tuner's own)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--ref-loop-t", type=int, default=16,
                    help="timeslices of the reference-style 'for t: data[t] = calc(t)' loop to time (0: skip)")
    ap.add_argument(
        "--calibrate-copy",
        type=int,
        default=0,
        metavar="MIB",
        help="first run one device-to-device copy of this many MiB (known HBM bytes for PMC byte-counter scaling)",
    )
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(
            "bench.py measures one GPU only: elemental generation shards over timeslices with no "
            "communication, so the one-GPU rate times the GPU count is the job's rate; run with --gpus 1",
            file=sys.stderr,
        )
        sys.exit(2)

    if args.workload in DIFFERENTIATION:
        bench_differentiation(args)
        return
    if args.workload in BATCHED:
        bench_batched(args)
        return

    import torch

    import easydistillation_amd as ed
    from easydistillation_amd import _native
    from easydistillation_amd import synthetic as S

    latt, Ne, num_nabla, n2, kind = WORKLOADS[args.workload]
    Lx, Ly, Lz, Lt = latt
    moms = S.momenta_by_norm(n2)
    D = ed.num_derivative(num_nabla)
    P = len(moms)
    dev = torch.device("cuda:0")
    log(f"workload {args.workload}: latt={latt} Ne={Ne} D={D} P={P} links={kind}")

    t0 = time.time()
    U = S.gauge_torch(latt, kind, 0.3, seed=11, device=dev)
    torch.cuda.synchronize()
    log(f"links built in {time.time() - t0:.1f} s")
    V = S.eigvecs_torch(latt, Ne, seed=12, device=dev, log=log)
    torch.cuda.synchronize()
    log(f"inputs built in {time.time() - t0:.1f} s")

    if args.calibrate_copy:
        n = args.calibrate_copy << 20
        src = torch.ones(n // 4, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        log(f"calibration copy: {n} bytes read + {n} bytes written")
        del src, dst

    forms = [f for f in args.forms.split(",") if f]
    results = {}
    outs = {}
    for form in forms:
        g = ed.ElementalGenerator(latt, None, None, num_nabla, moms, Ne=Ne, mode=form)
        g.set_option(_native.ED_OPT_MIXED_WS, args.mixed_ws)
        g.set_option(_native.ED_OPT_HOP_LDS, args.hop_lds)
        if args.tbatch > 0:
            g.set_option(_native.ED_OPT_TBATCH, args.tbatch)
        g.load_arrays(U, V)
        tuned = None
        if not args.no_tune:
            from easydistillation_amd.tune import autotune

            t1 = time.perf_counter()
            tuned = autotune(g, cache=args.tune_cache)
            log(f"{form}: tuned contraction {tuned} in {time.perf_counter() - t1:.1f} s")
        out = torch.empty((Lt, D, P, Ne, Ne), dtype=torch.complex128, device=dev)
        for i in range(args.warmup):
            t1 = time.perf_counter()
            g.calc_range(0, Lt, out=out)
            log(f"{form}: warmup {i} {time.perf_counter() - t1:.2f} s")
        torch.cuda.synchronize()
        g.context.set_option(_native.ED_OPT_PROFILE, 1)
        g.context.profile_reset()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for i in range(args.steps):
            g.calc_range(0, Lt, out=out)
            log(f"{form}: step {i} done at {time.perf_counter() - t0:.2f} s")
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / args.steps
        g.context.set_option(_native.ED_OPT_PROFILE, 0)
        c_ms, c_n = g.context.profile_read(_native.ED_KERNEL_CONTRACT)
        h_ms, h_n = g.context.profile_read(_native.ED_KERNEL_STENCIL)
        s_ms, s_n = g.context.profile_read(_native.ED_KERNEL_STAGE)
        ex_flops_t = g.context.query(_native.ED_Q_CONTRACT_FLOPS_PER_T)
        f16_flops_t = 0
        mixed_kernel = g.context.query(_native.ED_Q_MIXED_KERNEL) if form == "mixed" else None
        if mixed_kernel in (2, 3, 4, 5):
            # the mixed kernel runs stage 1 on FP32 MFMA and stage 2 on FP16 MFMA; its
            # roofline time is fp32_flops / FP32 peak + fp16_flops / FP16 peak, quoted
            # as FP32-equivalent flops
            ex_flops_t = g.context.query(_native.ED_Q_MIXED_F32_FLOPS_PER_T)
            f16_flops_t = g.context.query(_native.ED_Q_MIXED_F16_FLOPS_PER_T)
        equiv_flops_t = ex_flops_t + f16_flops_t * PEAK["mixed"] / PEAK_FP16
        algo_flops_t = g.context.query(_native.ED_Q_ALGO_FLOPS_PER_T)
        hops_t = g.context.query(_native.ED_Q_HOPS_PER_T)
        s = 16 if form == "fp64" else 8
        hop_bytes_t = hops_t * Lx * Ly * Lz * (2 * Ne * 3 * s + 144)  # links are complex128 in both forms
        results[form] = {
            "ms_per_step": dt * 1e3,
            "elementals_per_s": Lt * P * D * Ne * Ne / dt,
            "contract_ms_per_launch": c_ms / max(c_n, 1),
            "contract_launches_per_step": c_n // args.steps,
            "contract_tflops_executed": equiv_flops_t * Lt / (c_ms / args.steps * 1e-3) / 1e12,
            "contract_fp16_tflops": f16_flops_t * Lt / (c_ms / args.steps * 1e-3) / 1e12,
            "contract_tflops_algorithmic": algo_flops_t * Lt / (c_ms / args.steps * 1e-3) / 1e12,
            "executed_flops_per_launch": equiv_flops_t * Lt * args.steps / max(c_n, 1),
            "executed_fp16_flops_per_launch": f16_flops_t * Lt * args.steps / max(c_n, 1),
            "stencil_gbs": hop_bytes_t * Lt / (h_ms / args.steps * 1e-3) / 1e9 if h_ms > 0 else None,
            "stencil_ms_per_step": h_ms / args.steps,
            "stage_ms_per_step": s_ms / args.steps,
            "mixed_kernel": mixed_kernel,
            "strip": g.context.get_option(_native.ED_OPT_STRIP),
            "item_order": g.context.get_option(_native.ED_OPT_ITEM_ORDER),
            "hop_lds": g.context.get_option(_native.ED_OPT_HOP_LDS),
            "tuned": tuned,
        }
        log(f"{form}: {json.dumps(results[form])}")
        outs[form] = out
        del g

    err = None
    if "mixed" in outs and "fp64" in outs:
        err = blockwise_error(outs["mixed"], outs["fp64"], tables=True)
        log(f"mixed vs fp64 blockwise error: max {err[0]:.3e} mean {err[1]:.3e}")
    # V^dag V = 1 sanity check on the reference-numerics form
    chk = outs.get("fp64", outs.get("mixed"))
    eye_err = float((chk[:, 0, moms.index((0, 0, 0))] - torch.eye(Ne, dtype=chk.dtype, device=dev)).abs().max())
    del outs, chk
    torch.cuda.empty_cache()

    head = "mixed" if "mixed" in results else forms[0]
    r = results[head]

    ref_loop = None
    if args.ref_loop_t > 0:
        # the reference's driver: data[t] = calc(t) into a preallocated host array
        # (tests/test_elemental.py:37-45), here over the first ref_loop_t timeslices
        import numpy as np

        nt = min(Lt, args.ref_loop_t)
        g = ed.ElementalGenerator(latt, None, None, num_nabla, moms, Ne=Ne, mode=head)
        g.load_arrays(U, V)
        if not args.no_tune:
            from easydistillation_amd.tune import autotune

            autotune(g, cache=args.tune_cache)  # the shape's table entry, as for the timed forms
        data = np.zeros((nt, D, P, Ne, Ne), dtype=np.complex128)
        g.calc(0)  # warm-up: kernels, page-locked staging buffer
        g.load_arrays(U, V)  # drops the computed window
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        for t in range(nt):
            data[t] = g.calc(t)
        dt_loop = (time.perf_counter() - t1) / nt
        chk = float(np.abs(data[0, 0, moms.index((0, 0, 0))] - np.eye(Ne)).max())
        ref_loop = {
            "form": head,
            "timeslices": nt,
            "ms_per_timeslice": dt_loop * 1e3,
            "elementals_per_s": P * D * Ne * Ne / dt_loop,
            "identity_check_max_abs": chk,
            "note": (
                f"'for t: data[t] = calc(t)' into a preallocated numpy array; calc(t) computes aligned "
                f"windows of {g.WINDOW} timeslices per launch into HBM (the next window computing while the "
                f"host takes this one) and copies each timeslice out through page-locked buffers (the next "
                f"timeslice's copy overlapping the host copy; {D * P * Ne * Ne * 16 / 1e6:.0f} MB per "
                f"timeslice, PCIe included)"
            ),
        }
        # the same loop with the array in device memory, as the reference's driver
        # runs it on its cupy backend (backend.zeros on the GPU)
        from easydistillation_amd.backend import get_backend, set_backend

        before = get_backend()
        set_backend("torch")
        try:
            g.load_arrays(U, V)  # drops the computed window
            ddev = torch.zeros((nt, D, P, Ne, Ne), dtype=torch.complex128, device=dev)
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            for t in range(nt):
                ddev[t] = g.calc(t)
            torch.cuda.synchronize()
            dt_dev = (time.perf_counter() - t1) / nt
        finally:
            set_backend(before)
        ref_loop["device_array"] = {
            "ms_per_timeslice": dt_dev * 1e3,
            "elementals_per_s": P * D * Ne * Ne / dt_dev,
            "note": "'for t: data[t] = calc(t)' with data a device tensor (the reference driver on its cupy backend)",
        }
        log(f"reference-style loop, device array: {json.dumps(ref_loop['device_array'])}")
        del g, data, ddev
    # the measured bytes hold for the library's own timeslice batch (8 per launch at this workload)
    strip_eff = r["strip"] or (4 if head == "mixed" and r["mixed_kernel"] else 2)
    all_f16 = head == "mixed" and r["mixed_kernel"] == 5
    traffic_key = (args.workload, head, r["mixed_kernel"], strip_eff, r["item_order"]) if args.tbatch == 0 else None
    line = {
        "metric": METRIC,
        "value": r["elementals_per_s"],
        "unit": "elementals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32+fp16x3" if head == "mixed" and r["mixed_kernel"] in (2, 3, 4, 5) else (
            "fp32" if head == "mixed" else "fp64"),
        "precision": mixed_precision(r["mixed_kernel"]) if head == "mixed" and r["mixed_kernel"] in (2, 3, 4, 5) else None,
        "data": "synthetic (seeded Haar SU(3) links, orthonormal complex64 eigenvectors)",
        "config": {
            "workload": f"{Lx}^3x{Lt}" if Lx == Ly == Lz else f"{Lx}x{Ly}x{Lz}x{Lt}",
            "Nev": Ne,
            "momenta": P,
            "derivatives": D,
            "links": kind,
            "form": head,
            "elementals_per_step": Lt * P * D * Ne * Ne,
        },
        "roofline": {
            "bound": "mfma",
            # the FP16 stage-1 kernel executes only FP16 MFMAs: quoted against the FP16 peak
            "achieved": r["contract_fp16_tflops"] if all_f16 else r["contract_tflops_executed"],
            "peak": PEAK_FP16 if all_f16 else PEAK[head],
            "unit": "TFLOP/s",
            "frac": (r["contract_fp16_tflops"] / PEAK_FP16) if all_f16 else r["contract_tflops_executed"] / PEAK[head],
            "traffic": TRAFFIC_MEASURED.get(traffic_key),
            "traffic_source": TRAFFIC_SOURCE.get(traffic_key),
            "kernel": MIXED_KERNEL[r["mixed_kernel"]] if head == "mixed" else "contract_kernel<double, 3>",
            "flops_basis": (
                "executed FP16 MFMA flops of the factorized contraction per launch (split sum and phase "
                "contraction both on FP16 MFMA with hi/lo operand splits)"
                if all_f16
                else "executed MFMA flops of the factorized contraction per launch; the mixed kernel's FP16 "
                "MFMA flops are converted to FP32-equivalent at the FP32:FP16 peak ratio 157.3:2500"
                if head == "mixed"
                else "executed FP64 MFMA flops of the factorized contraction per launch"
            ),
            "executed_flops_per_launch": r["executed_flops_per_launch"],
            "avg_launch_ms": r["contract_ms_per_launch"],
            "algorithmic_equivalent_tflops": r["contract_tflops_algorithmic"],
        },
        "forms": results,
        "mixed_vs_fp64_error": None if err is None else {
            "max": err[0],
            "mean": err[1],
            "metric": "per (d,p) max|dE|/max|E_fp64|",
            "per_dp_normwise_max": err[2],
            "per_dp_frobenius": err[3],
            "tables": (f"[d][p]: d in derivative index order (derivative.py:23-33), p over the {P} momenta with "
                       f"|n|^2 <= {n2} sorted by |n|^2, then lexicographically"),
        },
        "identity_check_max_abs": eye_err,
    }
    if "fp64" in results and "mixed" in results:
        line["speedup_mixed_vs_fp64"] = results["fp64"]["ms_per_step"] / results["mixed"]["ms_per_step"]
    if "fp64" in results and results["fp64"]["stencil_gbs"]:
        line["stencil_roofline_fp64"] = {
            "bound": "hbm",
            "achieved": results["fp64"]["stencil_gbs"],
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": results["fp64"]["stencil_gbs"] / HBM_PEAK_GBS,
        }
    if ref_loop is not None:
        line["reference_loop"] = ref_loop
    if not args.no_cpu_baseline:
        log("cpu baseline sample ...")
        line["cpu_baseline"] = cpu_baseline(latt, Ne, num_nabla, moms, args.cpu_seconds)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
