#!/usr/bin/env python
"""J+K Fock-build throughput on one MI355X.

One step = one full J+K build (all 36 Coulomb and 54 exchange classes, the
device list preparation for the density, and AO assembly) of the lattice64
workload: 64 atoms on a 4x4x4 lattice (2.5 bohr, jitter 0.3), per atom 3 s
(3 prims), 2 p (2 prims), 1 spherical d (1 prim), nbf 896 (960 Cartesian),
D = C C^T with C (896 x 224) ~ N(0, 0.1^2), Schwarz threshold 1e-12.  Inputs
(pair lists, Schwarz factors, density) are resident in HBM before timing.

Prints one JSON line (see the keys below).  ``--gpus N`` with N > 1 exits: this
build runs on one GPU.
"""

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (ROOT, os.path.join(ROOT, "vlx-twoe-ints_amd")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

PEAK_FP64_VALU = 78.6e12     # MI355X FP64 vector, FLOP/s
PEAK_HBM = 8.0e12            # bytes/s


def cpu_baseline(eng, D_cart, nq, seed=0):
    """Time the CPU oracle (FP64 C restatement, OpenMP) on a bounded sample of
    the same workload and scale to one full J+K build.

    Per class: a random sample of bra blocks (J) or (i,k) blocks restricted to
    their first 8 partner blocks (K) is evaluated; the oracle reports how many
    primitive quartets it evaluated, giving seconds per quartet for the class,
    which is multiplied by the class's surviving-quartet count N_q."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import _util as U
    from vlxmi import pairs as P
    fl = eng.fl
    rng = np.random.default_rng(seed)
    cores = int(os.environ.get("OMP_NUM_THREADS", "0")) or min(16, os.cpu_count() or 1)
    os.environ["OMP_NUM_THREADS"] = str(cores)
    est = spent = 0.0
    nsamp = 0
    for ab, cd in P.j_classes():
        name = "computeCoulombFock" + (ab + cd).upper()
        if name not in nq or nq[name][0] == 0:
            continue
        n = fl.jlists[ab].nblk
        sel = np.sort(rng.choice(n, size=min(n, cores), replace=False)).astype(np.uint32)
        ne = np.zeros(1, dtype=np.uint64)
        tm = []
        U.oracle_matJ(fl, D_cart, "component", bra_blocks=sel, classes=[(ab, cd)], n_eval=ne, timer=tm)
        dt = sum(tm)
        spent += dt
        nsamp += int(ne[0])
        if ne[0]:
            est += dt / float(ne[0]) * nq[name][0]
    for cls in P.k_classes():
        name = "computeExchangeFock" + cls.upper()
        if name not in nq or nq[name][0] == 0:
            continue
        n = fl.iklists[cls[0] + cls[2]].nblk
        sel = np.sort(rng.choice(n, size=min(n, cores), replace=False)).astype(np.uint32)
        ne = np.zeros(1, dtype=np.uint64)
        tm = []
        U.oracle_matK(fl, D_cart, "component", ik_blocks=sel, classes=[cls], max_j=8, n_eval=ne, timer=tm)
        dt = sum(tm)
        spent += dt
        nsamp += int(ne[0])
        if ne[0]:
            est += dt / float(ne[0]) * nq[name][0]
    return {"value": 1.0 / est, "unit": "builds/s", "cores": cores, "kind": "port",
            "sample": "FP64 C oracle (OpenMP, %d threads): per class, %d random bra blocks (J) or (i,k) blocks "
                      "capped at 8 partner blocks (K); %d primitive quartets evaluated in %.1f s, seconds per "
                      "quartet x the class's surviving-quartet count -> %.1f s per full J+K build"
                      % (cores, cores, nsamp, spent, est),
            "seconds_per_build": est}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="lattice64")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--per-class-out", default=None, help="write per-class times/quartets/FLOPs (JSON)")
    ap.add_argument("--calib-copy-mib", type=int, default=0,
                    help="run one device copy of this many MiB first (reference kernel for HBM counters)")
    ap.add_argument("--precision", choices=("both", "fp64", "mixed"), default="both")
    a = ap.parse_args()
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this build runs the J/K path on one GPU only; --gpus %d is not supported "
              "(the multi-GPU sharding is described in DESIGN.md, not built)." % a.gpus, file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    from vlxmi import configs, flops
    from vlxmi import pairs as P
    from vlxmi.fock import FockEngine

    if a.calib_copy_mib:
        n = a.calib_copy_mib * (1 << 20) // 8
        src = torch.ones(n, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    mol, D = configs.make_workload(a.workload)
    t0 = time.time()
    eng = FockEngine(mol)
    torch.cuda.synchronize()
    setup_s = time.time() - t0
    Dd = eng.dev.tensor(D)

    def timed(precision):
        eng.precision = precision
        for _ in range(a.warmup):
            eng.build(Dd)
        torch.cuda.synchronize()
        eng.timer = []
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(a.steps):
            J, K = eng.build(Dd)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        timer, eng.timer = eng.timer, None
        per = {}
        for name, e0, e1 in timer:
            per.setdefault(name, []).append(e0.elapsed_time(e1) * 1e-3)
        return J, K, dt, per

    if a.precision == "mixed":
        J64, K64, dt64, per64 = timed("mixed")
    else:
        J64, K64, dt64, per64 = timed("fp64")
    if a.precision == "fp64":
        J, K, dt, per = J64, K64, dt64, per64
    else:
        J, K, dt, per = timed("mixed")
    kern_total = sum(sum(v) for v in per.values()) / a.steps
    EJ64, EK64 = eng.energies(Dd, J64, K64)
    EJ, EK = eng.energies(Dd, J, K)

    # algorithmic FLOPs of every class launch
    model = flops.load_model()
    Dc_np = eng.cart_density(Dd).cpu().numpy()
    maxD = eng.max_D(eng.cart_density(Dd))
    W = {}
    for ab, cd in P.j_classes():
        if ab in eng.fl.jlists and cd in eng.fl.jlists:
            nq = flops.count_j(eng.fl, Dc_np, ab, cd)
            W["computeCoulombFock" + (ab + cd).upper()] = (nq, nq * model["J"][ab + cd]["total"])
    for cls in P.k_classes():
        nq = eng.count_k_quartets(cls, maxD.get(cls[1] + cls[3], 0.0))
        W["computeExchangeFock" + cls.upper()] = (nq, nq * model["K"][cls]["total"])
    for k in list(W):
        W[k + "_mixed"] = W[k]
    total_flop = sum(v[1] for k, v in W.items() if not k.endswith("_mixed"))
    dom = max(per64, key=lambda k: sum(per64[k]))
    dom_t = sum(per64[dom]) / len(per64[dom])
    achieved = W[dom][1] / dom_t
    dom_m = max(per, key=lambda k: sum(per[k]))
    dom_mt = sum(per[dom_m]) / len(per[dom_m])
    top = sorted(per, key=lambda k: -sum(per[k]))[:6]
    dJ = float((J - J64).abs().max().item())
    dK = float((K - K64).abs().max().item())

    if a.per_class_out:
        rows = {k: {"avg_s": sum(v) / len(v), "quartets": W.get(k, (0, 0))[0], "flop": W.get(k, (0, 0))[1],
                    "tflops": W.get(k, (0, 0))[1] / (sum(v) / len(v)) / 1e12}
                for pp in (per64, per) for k, v in pp.items()}
        with open(a.per_class_out, "w") as f:
            json.dump(rows, f, indent=1, sort_keys=True)

    cpu = None
    if not a.no_cpu:
        cpu = cpu_baseline(eng, Dc_np, W)

    ms = 1e3 * dt / a.steps
    out = {
        "metric": "J+K Fock builds/s, mixed precision (FP32 VRR/HRR below Schwarz x |D| 1e-6) vs plain FP64",
        "value": a.steps / dt,
        "unit": "builds/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64+f32",
        "data": "synthetic",
        "config": {"workload": "lattice64: 64 atoms, 4x4x4 lattice 2.5 bohr +-0.3 jitter, per atom 3s(3)+2p(2)+1 sph d(1), "
                               "nbf 896 (960 Cartesian), D=CC^T C(896x224)~N(0,0.01), thr 1e-12, classes ss..dd (36 J + 54 K)",
                   "precision": "mixed (fp32 threshold 1e-6)", "nbf": mol.nbf, "naos_cart": mol.naos_cart},
        "roofline": {"bound": "valu", "kernel": dom_m, "achieved": W[dom_m][1] / dom_mt / 1e12,
                     "peak": PEAK_FP64_VALU / 1e12, "unit": "TFLOP/s", "frac": W[dom_m][1] / dom_mt / PEAK_FP64_VALU,
                     "traffic": None, "algorithmic_flop_per_launch": W[dom_m][1],
                     "quartets_per_launch": W[dom_m][0], "avg_launch_s": dom_mt,
                     "fp64_path": {"kernel": dom, "achieved": achieved / 1e12, "frac": achieved / PEAK_FP64_VALU,
                                   "avg_launch_s": dom_t}},
        "cpu_baseline": cpu,
        "fp64_builds_per_s": a.steps / dt64,
        "mixed_builds_per_s": a.steps / dt,
        "mixed_speedup": dt64 / dt,
        "mixed_error_vs_fp64": {"max_abs_dJ": dJ, "max_abs_dK": dK, "dE_J": EJ - EJ64, "dE_K": EK - EK64},
        "whole_build": {"algorithmic_flop": total_flop, "kernel_s": kern_total,
                        "tflops_over_kernels": total_flop / kern_total / 1e12,
                        "tflops_over_wall": total_flop / (dt / a.steps) / 1e12},
        "E_J": EJ64, "E_K": EK64, "setup_s": setup_s,
        "top_kernels_ms": {k: 1e3 * sum(per[k]) / len(per[k]) for k in top},
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
