#!/usr/bin/env python3
"""H_eff benchmark on one MI355X.

Headline (the metric's configuration): the two-site H_eff matvec on the
middle bond of the Hubbard chain at D = 2048 (L = 64, t = 1, U = 4, half
filling, U1 x U1 charges, d = 4, w = 6): the bond's charge sectors come from
a DMRG ramp (tests/golden/hubbard_L64_D2048_bonds.json), the environment
blocks and the 32 unit input vectors are seeded Gaussians.  One step = 32
grouped block-sparse FP64 matvecs, inputs resident in HBM; the value is
TFLOP/s with the planner's flop count (sum of 2 m n k over the sector
products, plus the block mode stage).  Beside it: the roofline of the
grouped-GEMM kernel, the mixed precisions and the plain per-sector loop on
the same inputs, the CPU per-sector restatement (baseline and full-size
parity), the D = 2048 DMRG sweep seconds per precision (ramp from the
charge-density-wave state, mixed sweep + FP64 polish schedule, plain
per-sector sweep), the dense D = 512 microbenchmark, J1-J2 gradients
(L = 40, D = 256), the TFIM TDVP2 quench (L = 100, 40 steps) and the XXZ
L = 20 D = 64 DMRG2 against exact diagonalisation.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("2-site H_eff matvec TFLOP/s + DMRG sweep s at D=2048; "
          "mixed vs FP64 at ΔE/L≤1e-9")
FP64_PEAK_TFLOPS = 78.6
NVEC = 32


def make_inputs(kind, D, seed=0):
    from finitemptools_amd import models
    rng = np.random.default_rng(seed)
    W = models.heisenberg_xxz_mpo(4, J=1.0, Jzz=1.0)[1]
    w, d = W.shape[0], W.shape[1]
    std = 1.0 / np.sqrt(D * w)
    L = rng.standard_normal((D, w, D)) * std
    R = rng.standard_normal((D, w, D)) * std
    shape = (D, d, d, D) if kind == 2 else (D, d, D)
    xs = []
    for i in range(NVEC):
        x = np.random.default_rng(1000 + i).standard_normal(shape)
        xs.append(x / np.linalg.norm(x))
    return W, L, R, xs


def flops(kind, D, w=5, d=2):
    if kind == 2:
        return 4.0 * D ** 3 * w * d ** 2 + 4.0 * D ** 2 * w ** 2 * d ** 3
    return 4.0 * D ** 3 * w * d + 2.0 * D ** 2 * w ** 2 * d ** 2


def gpu_run(kind, D, steps, warmup, prec="f64", profile=False):
    import torch
    from finitemptools_amd import _lib
    from finitemptools_amd.derivatives import HeffOperator
    from finitemptools_amd.engine import colmajor_empty, to_device
    W, L, R, xs = make_inputs(kind, D)
    Wd, Ld, Rd = to_device(W), to_device(L), to_device(R)
    op = HeffOperator(kind, Ld, Rd, Wd, Wd if kind == 2 else None, prec=prec)
    xd = [to_device(x) for x in xs]
    yd = [colmajor_empty(op.shape) for _ in xs]
    eng = op.engine
    for _ in range(warmup):
        for x, y in zip(xd, yd):
            op.apply(x, out=y)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        for x, y in zip(xd, yd):
            op.apply(x, out=y)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    prof = None
    if profile:
        # a second pass of the same steps with HIP events around every launch
        # (kept out of the timed pass above: the events add gaps)
        eng.profile_reset()
        eng.profile_enable(True)
        for _ in range(steps):
            for x, y in zip(xd, yd):
                op.apply(x, out=y)
        torch.cuda.synchronize()
        eng.profile_enable(False)
        prof = {k: eng.profile_read(c) for k, c in
                (("gemm", _lib.MH_K_GEMM), ("wstage", _lib.MH_K_WSTAGE))}
    sec = (t1 - t0) / (steps * NVEC)
    return {"sec_per_matvec": sec, "tflops": flops(kind, D) / sec / 1e12,
            "y0": yd[0].cpu().numpy(), "inputs": (W, L, R, xs), "prof": prof}


BOND_FIXTURE = os.path.join(ROOT, "tests", "golden", "hubbard_L64_D2048_bonds.json")
HUBBARD_L, MID_SITE = 64, 31


def middle_bond_case(seed=0):
    """The two-site H_eff of the Hubbard chain's middle bond at D = 2048
    (L = 64, t = 1, U = 4, half filling, U1 x U1): charge sectors from the
    committed fixture (a DMRG ramp to D = 2048), seeded iid N(0, 1/(D w))
    environment blocks and NVEC seeded unit vectors."""
    from finitemptools_amd import blockmps as bm
    from finitemptools_amd import blocksparse as bs
    from finitemptools_amd import models
    with open(BOND_FIXTURE) as f:
        b = json.load(f)["bonds"][str(MID_SITE)]

    def sp(lst):
        return bs.Space([(tuple(q), int(n)) for q, n in lst])

    phys = bm.hubbard_charges()
    xl = bs.BlockLayout(sp(b["x_l"]), phys, 2, sp(b["x_r"]))
    Lay = bs.EnvLayout(sp(b["L_space"]), [tuple(q) for q in b["L_qw"]])
    Ray = bs.EnvLayout(sp(b["R_space"]), [tuple(q) for q in b["R_qw"]])
    mpo = models.hubbard_mpo(HUBBARD_L, 1.0, 4.0)
    W1, W2 = mpo[MID_SITE], mpo[MID_SITE + 1]
    rng = np.random.default_rng(seed)
    std = 1.0 / np.sqrt(2048 * W1.shape[0])
    Le = rng.standard_normal(Lay.size) * std
    Re = rng.standard_normal(Ray.size) * std
    xs = []
    for i in range(NVEC):
        x = np.random.default_rng(1000 + i).standard_normal(xl.size)
        xs.append(x / np.linalg.norm(x))
    return phys, xl, Lay, Ray, W1, W2, Le, Re, xs


def _oracle_blocks(phys, xl, Lay, Ray, W1, W2, Le, Re, x):
    """the CPU per-sector restatement on the same blocks"""
    from finitemptools_amd import blocksparse as bs
    from oracle import blocksparse as OB

    def env_entries(lay):
        return [((q, w), off, lay.space.dim[q], lay.space.dim[bs.qsub(q, lay.qw[w])])
                for (q, w), off in lay.off.items()]

    xe = [((b.qa, b.p[0], b.p[1], b.qb), b.off, b.rows, b.cols) for b in xl.blocks]
    args = (OB.blocks_from_flat(x, xe), W1, W2, OB.blocks_from_flat(Le, env_entries(Lay)),
            OB.blocks_from_flat(Re, env_entries(Ray)), phys, Lay.qw, Ray.qw,
            dict(xl.lspace.sectors), dict(xl.rspace.sectors))
    return (lambda: OB.flat_from_blocks(OB.ac2_prime(*args), xe, xl.size))


def blocksparse_headline(steps, warmup, cpu=True, extras=True):
    """One step = NVEC FP64 two-site H_eff matvecs on the D = 2048 middle
    bond, each a grouped plan (two grouped-GEMM launches and one block mode
    product), inputs resident in HBM.  Flops: the sum of 2 m n k of the
    reference's graded contraction over every charge-allowed block, plus the
    mode stage's (the plan executes fewer: it skips the MPO's structural
    zeros)."""
    import torch
    from finitemptools_amd import _lib
    from finitemptools_amd import blockmps as bm
    from finitemptools_amd import blocksparse as bs
    from finitemptools_amd.engine import engine
    phys, xl, Lay, Ray, W1, W2, Le, Re, xs = middle_bond_case()
    Ld = torch.as_tensor(Le, device="cuda")
    Rd = torch.as_tensor(Re, device="cuda")
    xd = [torch.as_tensor(x, device="cuda") for x in xs]
    yd = [torch.empty_like(x) for x in xd]
    plan = bs.make_plan("heff2", xl, Lay, Ray, W1, W2, "f64")
    # the rate is quoted on the flops of the reference's graded contraction
    # (every charge-allowed block, the MPO's zero entries included); the plan
    # executes fewer (it skips the MPO's structural zeros)
    F = bs.heff2_reference_flops(xl, Lay, Ray, W1, W2)
    F_exec = plan.flops()

    def step(p):
        for x, y in zip(xd, yd):
            p.apply(x, Ld, Rd, y)

    for _ in range(max(1, warmup)):
        step(plan)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        step(plan)
    torch.cuda.synchronize()
    sec = (time.perf_counter() - t0) / (steps * NVEC)
    # a second pass with HIP events around every launch on the engine's stream
    eng = engine(0)
    eng.profile_reset()
    eng.profile_enable(True)
    step(plan)
    torch.cuda.synchronize()
    eng.profile_enable(False)
    gms, gn, gfl = eng.profile_read(_lib.MH_K_GEMM)
    wms, wn, wfl = eng.profile_read(_lib.MH_K_WSTAGE)
    y0 = yd[0].cpu().numpy()
    out = {"sec_per_matvec": sec, "flops": F, "executed_flops": F_exec, "tflops": F / sec / 1e12,
           "gemm": {"ms": gms, "launches": gn, "flops": gfl},
           "mode": {"ms": wms, "launches": wn, "flops": wfl},
           "sectors": len(xl.lspace.sectors), "max_sector": max(d for _, d in xl.lspace.sectors),
           "vector_len": xl.size}
    if not extras:
        return out
    # mixed precisions and the plain per-sector FP64 loop on the same inputs
    mixed = {}
    for prec in ("f32", "bf16x3"):
        pm = bs.make_plan("heff2", xl, Lay, Ray, W1, W2, prec)
        step(pm)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            step(pm)
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) / (steps * NVEC)
        ym = yd[0].cpu().numpy()
        mixed[prec] = {"ms": round(ms * 1e3, 4), "tflops": round(F / ms / 1e12, 3),
                       "speedup_vs_f64": round(sec / ms, 3),
                       "rel_err_vs_f64": float(np.linalg.norm(ym - y0) / np.linalg.norm(y0))}
    out["mixed"] = mixed
    with bm.plain_path():
        pp = bs.make_plan("heff2", xl, Lay, Ray, W1, W2, "f64")
        step(pp)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        step(pp)
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) / NVEC
    out["f64_plain_per_sector"] = {"ms": round(ms * 1e3, 4), "tflops": round(F / ms / 1e12, 3),
                                   "grouped_speedup": round(ms / sec, 2)}
    if cpu:
        out["cpu"] = cpu_baseline_blocksparse(phys, xl, Lay, Ray, W1, W2, Le, Re, xs, F, y0)
    return out


def cpu_baseline_blocksparse(phys, xl, Lay, Ray, W1, W2, Le, Re, xs, F, y0, runs=5):
    """The per-sector CPU restatement (oracle/blocksparse.py: one host BLAS
    GEMM per coupled sector and step, as @tensor on graded tensors does) on
    the same middle-bond blocks: median of `runs` matvecs after a warm-up on
    all BLAS threads, and of 3 on one thread.  Also the full-size parity of
    the FP64 device result against it."""
    from threadpoolctl import threadpool_info, threadpool_limits
    f = _oracle_blocks(phys, xl, Lay, Ray, W1, W2, Le, Re, xs[0])
    yc = f()   # warm-up
    parity = float(np.linalg.norm(y0 - yc) / np.linalg.norm(yc))
    ts = []
    for _ in range(runs):
        t0 = time.perf_counter()
        f()
        ts.append(time.perf_counter() - t0)
    threads = max((i.get("num_threads", 1) for i in threadpool_info()), default=1)
    t1 = []
    with threadpool_limits(1):
        for _ in range(3):
            t0 = time.perf_counter()
            f()
            t1.append(time.perf_counter() - t0)
    med, med1 = float(np.median(ts)), float(np.median(t1))
    return {"value": round(F / med / 1e12, 5), "unit": "TFLOP/s", "cores": threads,
            "os_cpu_count": os.cpu_count(), "kind": "port",
            "one_thread": {"value": round(F / med1 / 1e12, 5), "s_per_matvec": round(med1, 4)},
            "s_per_matvec": round(med, 4),
            "sample": (f"the D=2048 middle-bond two-site H_eff matvec (same blocks as the GPU "
                       f"line), median of {runs} after a warm-up on {threads} BLAS threads "
                       f"and of 3 on one thread; numpy per-sector restatement "
                       f"(oracle/blocksparse.py)"),
            "gpu_f64_rel_err_vs_cpu": parity}


def complex_tfim_matvec(D=512, nvec=8, steps=5, precs=("f64", "f32", "bf16x3")):
    """The unit of work of the TDVP quench at its bond-dimension cap: the
    two-site H_eff of the TFIM chain (w = 3, d = 2) on ComplexF64 data at
    D = 512, 4 x F2 = 26.1 GFLOP of real arithmetic per matvec; seeded complex
    Gaussian environments (std 1/sqrt(D w)) and unit vectors; one vector
    checked against the oracle.  "tflops" counts 8 real flops per complex
    multiply-add; the FP64 GEMMs execute 6 of them on MFMA (three-product
    form), which "mfma_tflops_executed" reports."""
    import torch
    from finitemptools_amd import models
    from finitemptools_amd.derivatives import HeffOperator
    from finitemptools_amd.engine import colmajor_empty, to_device
    from oracle import contractions as C
    W = models.tfim_mpo(4, 1.0, 1.0, dtype=np.complex128)[1]
    w, d = W.shape[0], W.shape[1]
    rng = np.random.default_rng(3)
    std = 1.0 / np.sqrt(2 * D * w)
    L = (rng.standard_normal((D, w, D)) + 1j * rng.standard_normal((D, w, D))) * std
    R = (rng.standard_normal((D, w, D)) + 1j * rng.standard_normal((D, w, D))) * std
    xs = []
    for i in range(nvec):
        r = np.random.default_rng(2000 + i)
        x = r.standard_normal((D, d, d, D)) + 1j * r.standard_normal((D, d, d, D))
        xs.append(x / np.linalg.norm(x))
    F = 4.0 * flops(2, D, w=w, d=d)
    Wd, Ld, Rd = to_device(W), to_device(L), to_device(R)
    xd = [to_device(x) for x in xs]
    out = {"workload": f"ac2_prime two-site H_eff, TFIM w={w}, d={d}, D={D}, ComplexF64 storage",
           "gflop_per_matvec": round(F / 1e9, 2)}
    yref = None
    for prec in precs:
        op = HeffOperator(2, Ld, Rd, Wd, Wd, prec=prec)
        yd = [colmajor_empty(op.shape, dtype=torch.complex128) for _ in xs]
        for x, y in zip(xd, yd):
            op.apply(x, out=y)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            for x, y in zip(xd, yd):
                op.apply(x, out=y)
        torch.cuda.synchronize()
        sec = (time.perf_counter() - t0) / (steps * nvec)
        ent = {"ms": round(sec * 1e3, 4), "tflops": round(F / sec / 1e12, 3)}
        if prec == "f64":
            ent["mfma_tflops_executed"] = round(0.75 * F / sec / 1e12, 3)
        y0 = yd[0].cpu().numpy()
        if yref is None:
            yref = C.ac2_prime(xs[0], W, W, L, R)
            ent["rel_err_vs_oracle"] = float(np.linalg.norm(y0 - yref) / np.linalg.norm(yref))
        else:
            ent["rel_err_vs_oracle"] = float(np.linalg.norm(y0 - yref) / np.linalg.norm(yref))
            ent["speedup_vs_f64"] = round(out["f64"]["ms"] / ent["ms"], 3)
        out[prec] = ent
    return out


def stored_traffic(name):
    """HBM bytes per launch from a committed counter summary under profiles/
    (its line "traffic_bytes_per_launch: N"), or None if absent."""
    path = os.path.join(ROOT, "profiles", name)
    try:
        for line in open(path):
            if line.startswith("traffic_bytes_per_launch:"):
                return float(line.split(":")[1])
    except OSError:
        pass
    return None


SWEEP_REPS = 3   # timed D = 2048 sweeps per precision (the median is reported)


def blocksparse_hubbard(L=64, ramp=(256, 512, 1024, 2048), plain=True):
    """Charge-conserving two-site DMRG of the Hubbard chain (t = 1, U = 4,
    half filling, U1 x U1): ramp from the charge-density-wave product state
    through D = 256 .. 2048 (one sweep each), then SWEEP_REPS sweeps per
    precision from the same D = 2048 state (the median reported, the last
    followed by an FP64 polish sweep), and the same sweep as the plain
    per-sector FP64 loop."""
    import torch
    from finitemptools_amd import blockmps as bm
    from finitemptools_amd import blocksparse as bs
    from finitemptools_amd import models
    from finitemptools_amd.dmrg import DMRG2
    from finitemptools_amd.truncation import MPSTruncation
    phys = bm.hubbard_charges()
    psi = bm.prodmps(phys, [phys.index((0, 0)), phys.index((1, 1))] * (L // 2))
    mpo = bm.BlockMPO(models.hubbard_mpo(L, 1.0, 4.0), phys)
    env = bm.BlockExpectationCache(mpo, psi)
    t0 = time.perf_counter()
    ramp_log = []
    for D in ramp:
        alg = DMRG2(trunc=MPSTruncation(D=D, eps=1e-10))
        ts = time.perf_counter()
        e, _ = bm.sweep(env, alg)
        ramp_log.append({"D": D, "s": round(time.perf_counter() - ts, 2), "E_per_site": e[-1] / L})
        print(f"  ramp D={D}: {ramp_log[-1]['s']} s", file=sys.stderr, flush=True)
    ramp_s = time.perf_counter() - t0
    # planning cost of the middle bond's H_eff (native C planner vs Python)
    site = L // 2 - 1
    x = bm.two_site(psi[site], psi[site + 1])
    Le, Re = env.hstorage[site], env.hstorage[site + 2]
    args = (x.layout, Le.layout, Re.layout, mpo[site], mpo[site + 1])
    ts = time.perf_counter()
    bs.plan_heff2(*args)
    py_plan_ms = 1e3 * (time.perf_counter() - ts)
    ts = time.perf_counter()
    bs.host_native_tables("heff2", *args, "f64")
    native_plan_ms = 1e3 * (time.perf_counter() - ts)
    plan_ms = {"python": round(py_plan_ms, 2), "native": round(native_plan_ms, 2)}
    # (b) one sweep per precision from the same state
    base_sites = [bm.BlockTensor(t.layout, t.data.clone()) for t in psi.sites]
    base_env = [bm.BlockEnv(q.layout, q.data.clone()) for q in env.hstorage]
    sweeps, schedule = {}, {}
    # the grouped-GEMM tile tuner times its variants on the first stage of
    # each shape class and precision: the FP64 classes are known from the
    # ramp; one untimed sweep per mixed precision from the same state lets
    # those precisions meet the same warm tables
    for prec in ("f32", "bf16x3"):
        psi.sites = [bm.BlockTensor(t.layout, t.data.clone()) for t in base_sites]
        env.hstorage = [bm.BlockEnv(q.layout, q.data.clone()) for q in base_env]
        env.prec = prec
        bm.sweep(env, DMRG2(trunc=MPSTruncation(D=ramp[-1], eps=1e-10), prec=prec))
    for prec in ("f64", "f32", "bf16x3"):
        # the median of SWEEP_REPS sweeps from the same state (each gives the
        # same energies bit for bit; the time moves with the host's speed)
        runs = []
        for _ in range(SWEEP_REPS):
            psi.sites = [bm.BlockTensor(t.layout, t.data.clone()) for t in base_sites]
            env.hstorage = [bm.BlockEnv(q.layout, q.data.clone()) for q in base_env]
            env.prec = prec
            alg = DMRG2(trunc=MPSTruncation(D=ramp[-1], eps=1e-10), prec=prec)
            torch.cuda.synchronize()
            ts = time.perf_counter()
            e, delta = bm.sweep(env, alg)
            torch.cuda.synchronize()
            runs.append((time.perf_counter() - ts, e, alg.stats))
            print(f"  sweep D=2048 {prec}: {runs[-1][0]:.2f} s", file=sys.stderr, flush=True)
        dt, e, st = sorted(runs, key=lambda r: r[0])[len(runs) // 2]
        sweeps[prec] = {"s": round(dt, 3), "s_runs": [round(r[0], 3) for r in runs],
                        "E_per_site": e[-1] / L,
                        "lanczos_s": round(st["lanczos_s"], 3), "split_s": round(st.get("split_s", 0), 3),
                        "plan_s": round(st.get("plan_s", 0), 3), "env_s": round(st["env_s"], 3),
                        "matvec_flops": st["matvec_flops"],
                        "lanczos_tflops": round(st["matvec_flops"] / max(st["lanczos_s"], 1e-9) / 1e12, 3),
                        "flop_count": "executed (the plans skip the MPO's structural zeros)"}
        # schedule: this sweep, then one FP64 polish sweep from its state
        env.prec = "f64"
        alg = DMRG2(trunc=MPSTruncation(D=ramp[-1], eps=1e-10), prec="f64")
        torch.cuda.synchronize()
        ts = time.perf_counter()
        e2, _ = bm.sweep(env, alg)
        torch.cuda.synchronize()
        schedule[f"{prec}+f64"] = {"s": round(sweeps[prec]["s"] + time.perf_counter() - ts, 3),
                                   "E_per_site": e2[-1] / L}
    # the same sweep as the plain per-sector FP64 loop: one GEMM launch per
    # sector product in every plan, one eigensolve per sector in sequence
    if not plain:
        for prec in ("f32", "bf16x3"):
            sweeps[prec]["dE_per_site_vs_f64"] = sweeps[prec]["E_per_site"] - sweeps["f64"]["E_per_site"]
            sweeps[prec]["speedup_vs_f64"] = round(sweeps["f64"]["s"] / sweeps[prec]["s"], 3)
        return {"workload": f"Hubbard L={L} t=1 U=4 half filling U1xU1, DMRG2 ramp {list(ramp)}",
                "ramp_s": round(ramp_s, 2), "ramp": ramp_log, "plan_ms_middle_bond": plan_ms,
                "sweep_D2048": sweeps, "schedule_mixed_then_f64": schedule}
    psi.sites = [bm.BlockTensor(t.layout, t.data.clone()) for t in base_sites]
    env.hstorage = [bm.BlockEnv(q.layout, q.data.clone()) for q in base_env]
    env.prec = "f64"
    alg = DMRG2(trunc=MPSTruncation(D=ramp[-1], eps=1e-10), prec="f64")
    with bm.plain_path():
        torch.cuda.synchronize()
        ts = time.perf_counter()
        e, delta = bm.sweep(env, alg)
        torch.cuda.synchronize()
    st = alg.stats
    sweeps["f64_plain_per_sector"] = {
        "s": round(time.perf_counter() - ts, 3), "E_per_site": e[-1] / L,
        "lanczos_s": round(st["lanczos_s"], 3), "split_s": round(st.get("split_s", 0), 3),
        "env_s": round(st["env_s"], 3),
        "dE_per_site_vs_f64": e[-1] / L - sweeps["f64"]["E_per_site"]}
    print(f"  block sweep plain f64: {sweeps['f64_plain_per_sector']['s']} s", file=sys.stderr, flush=True)
    for prec in ("f32", "bf16x3"):
        sweeps[prec]["dE_per_site_vs_f64"] = sweeps[prec]["E_per_site"] - sweeps["f64"]["E_per_site"]
        sweeps[prec]["speedup_vs_f64"] = round(sweeps["f64"]["s"] / sweeps[prec]["s"], 3)
    for prec in ("f64", "f32", "bf16x3"):
        sweeps[prec]["speedup_vs_plain_f64"] = round(sweeps["f64_plain_per_sector"]["s"] / sweeps[prec]["s"], 3)
    # mixed sweep + FP64 polish against two FP64 sweeps from the same state:
    # the energy criterion is |dE/L| <= 1e-9
    ref = schedule["f64+f64"]
    for prec in ("f32", "bf16x3"):
        ent = schedule[f"{prec}+f64"]
        ent["dE_per_site_vs_f64+f64"] = ent["E_per_site"] - ref["E_per_site"]
        ent["meets_1e-9"] = bool(abs(ent["dE_per_site_vs_f64+f64"]) <= 1e-9)
        ent["speedup_vs_f64+f64"] = round(ref["s"] / ent["s"], 3)
    return {"workload": f"Hubbard L={L} t=1 U=4 half filling U1xU1, DMRG2 ramp {list(ramp)}",
            "ramp_s": round(ramp_s, 2), "ramp": ramp_log, "plan_ms_middle_bond": plan_ms,
            "sweep_D2048": sweeps, "schedule_mixed_then_f64": schedule}


def xxz_dmrg_L20_D64(cpu=True):
    """The reference's CPU run configuration: XXZ (J = Delta = 1), L = 20,
    two-site DMRG to D = 64, 4 sweeps from a seeded random right-canonical
    MPS; energy and per-sweep times on the device, and the oracle's numpy
    sweep of the same configuration on the host cores."""
    import json as _json
    import torch
    from finitemptools_amd import dmrg, models
    from finitemptools_amd.envs import environments, mpo_to_device
    from finitemptools_amd.mps import randommps
    from finitemptools_amd.truncation import MPSTruncation
    L, D = 20, 64
    ed = _json.load(open(os.path.join(ROOT, "tests", "golden", "ed_energies.json")))["xxz_L20_J1_D1"]
    psi = randommps(L, 2, D, seed=0).right_canonicalize()
    env = environments(mpo_to_device(models.heisenberg_xxz_mpo(L)), psi)
    alg = dmrg.DMRG2(trunc=MPSTruncation(D=D, eps=1e-8))
    sweeps = []
    for _ in range(4):
        torch.cuda.synchronize()
        ts = time.perf_counter()
        e, delta = dmrg.sweep(env, alg)
        torch.cuda.synchronize()
        sweeps.append(round(time.perf_counter() - ts, 4))
    out = {"workload": "XXZ L=20 J=Delta=1, DMRG2 D=64, 4 sweeps, seed 0",
           "E": e[-1], "E_minus_ED": e[-1] - ed, "sweep_s": sweeps,
           "lanczos_s": round(alg.stats["lanczos_s"], 4), "split_s": round(alg.stats["svd_s"], 4),
           "env_s": round(alg.stats["env_s"], 4)}
    if cpu:
        from oracle import dmrg as OD
        psi_h = OD.right_canonicalize(OD.randommps(L, 2, D, seed=0))
        envh = OD.ExpectationCache(models.heisenberg_xxz_mpo(L), psi_h)
        algh = OD.DMRG2(trunc=OD.MPSTruncation(D=D, eps=1e-8))
        ts = time.perf_counter()
        eh, _ = OD.compute(envh, algh, nsweeps=1)
        out["cpu_oracle_first_sweep_s"] = round(time.perf_counter() - ts, 3)
        out["cpu_oracle_first_sweep_E"] = eh[-1]
    return out


def gradients_converged_j1j2(L=40, D=256, J2=0.25, nsweeps=5, h=1e-5):
    """dE/dJ2 at a converged DMRG2 ground state of the J1-J2 chain: adjoint
    kernels vs Hellmann-Feynman vs the central finite difference of ground
    energies (two more DMRG runs at J2 +- h)."""
    import torch
    from finitemptools_amd import dmrg, models
    from finitemptools_amd import gradients as GR
    from finitemptools_amd.envs import environments, mpo_to_device
    from finitemptools_amd.mps import randommps
    from finitemptools_amd.truncation import MPSTruncation

    def ground(j2):
        psi = randommps(L, 2, 32, seed=1).right_canonicalize()
        mpo = mpo_to_device(models.j1j2_mpo(L, 1.0, j2, keep_j2=True))
        alg = dmrg.DMRG2(trunc=MPSTruncation(D=D, eps=1e-12))
        e, _ = dmrg.compute(environments(mpo, psi), alg, nsweeps)
        return e[-1], psi, mpo

    ts = time.perf_counter()
    E0, psi, mpo = ground(J2)
    dmrg_s = time.perf_counter() - ts
    sites = [psi[i] for i in range(L)]
    W1 = mpo_to_device(models.j1j2_mpo(L, 1.0, 1.0, keep_j2=True))
    W0 = mpo_to_device(models.j1j2_mpo(L, 1.0, 0.0, keep_j2=True))
    dW = [a - b for a, b in zip(W1, W0)]
    torch.cuda.synchronize()
    ts = time.perf_counter()
    res = GR.energy_gradients(mpo, sites, want_dA=True)
    torch.cuda.synchronize()
    grad_s = time.perf_counter() - ts
    g = GR.dE_dtheta(dW, res)
    hf = GR.hellmann_feynman(mpo_to_device(models.j1j2_dJ2_mpo(L)), sites)
    fd = (ground(J2 + h)[0] - ground(J2 - h)[0]) / (2 * h)
    return {"workload": f"J1-J2 L={L} J2={J2}, DMRG2 D={D}, {nsweeps} sweeps",
            "E0": E0, "dmrg_s": round(dmrg_s, 2), "gradient_s": round(grad_s, 4),
            "dE_dJ2_adjoint": g, "hellmann_feynman": hf, "central_fd_h1e-5": fd,
            "adjoint_minus_hf": g - hf, "adjoint_minus_fd": g - fd,
            "dE_dA_norm": float(sum(torch.sum(a * a).item() for a in res.dA) ** 0.5)}


def tdvp_tfim_quench(L=100, nsteps=40, dt=0.05, D=512, precs=("f64", "f32")):
    """TDVP2 quench of the transverse-field Ising chain (J = 1, h = 0 -> 1)
    from |+>^L, ComplexF64 storage, truncation D <= 512 at eps = 1e-8, Lanczos
    exponentiate (dimension <= 12, FP64 re-orthogonalisation in the mixed
    forms): <sz_i(t)> at every step against the free-fermion closed form and,
    for the mixed precision, against the FP64 trajectory.  (Split-BF16 is left
    out: its ~1e-5 relative noise lies above the 1e-8 truncation threshold, so
    the bond dimension runs to the cap and the run takes minutes.)"""
    import torch
    from finitemptools_amd import models, tdvp
    from finitemptools_amd.envs import environments, local_expectations, mpo_to_device
    from finitemptools_amd.mps import prodmps
    from finitemptools_amd.truncation import MPSTruncation
    from oracle import ed
    from oracle.freefermion import tfim_quench_sz
    mpo = mpo_to_device(models.tfim_mpo(L, 1.0, 1.0, dtype=np.complex128), torch.complex128)
    exact = tfim_quench_sz(L, 1.0, 1.0, [dt * (k + 1) for k in range(nsteps)])
    res, traj = {"workload": f"TFIM L={L} J=1 quench h=0->1 from |+>^L, TDVP2 dt={dt}, {nsteps} steps, "
                             f"C64, D<={D} eps=1e-8"}, {}
    for prec in precs:
        psi = prodmps([np.ones(2) / np.sqrt(2)] * L, dtype=torch.complex128)
        alg = tdvp.TDVP2(stepsize=-1j * dt, trunc=MPSTruncation(D=D, eps=1e-8), prec=prec,
                         ishermitian=True)
        env = environments(mpo, psi, prec)
        rows, sec = [], 0.0
        for _ in range(nsteps):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            tdvp.sweep(env, alg)
            torch.cuda.synchronize()
            sec += time.perf_counter() - t0
            rows.append(local_expectations(psi, ed.SZ))
        traj[prec] = np.array(rows)
        ent = {"s": round(sec, 2), "s_per_step": round(sec / nsteps, 3),
               "D_max": max(t.shape[2] for t in psi.data),
               "max_abs_err_vs_closed_form": float(np.abs(traj[prec] - exact).max())}
        if prec != "f64" and "f64" in traj:
            ent["max_abs_diff_vs_f64"] = float(np.abs(traj[prec] - traj["f64"]).max())
            ent["speedup_vs_f64"] = round(res["f64"]["s"] / sec, 3)
        res[prec] = ent
        print(f"  tdvp {prec}: {sec:.1f} s", file=sys.stderr, flush=True)
    # D = 512 is the truncation cap of the configuration; at eps = 1e-8 the
    # 40-step quench's bond dimension stays far below it, so the D = 512
    # unit of work is measured on its own (ac2_prime_D512_c64_tfim)
    res["D_cap"] = D
    res["D_cap_reached"] = any(res[p]["D_max"] >= D for p in precs)
    return res


def gradients_j1j2(L=40, D=256, J2=0.25):
    """dE/dJ2 and dE/dA of E = <psi|H|psi>/<psi|psi> for the J1-J2 chain on a
    random D = 256 MPS (the cost does not depend on convergence), with stored
    and with recomputed environments."""
    import torch
    from finitemptools_amd import gradients as GR
    from finitemptools_amd import models
    from finitemptools_amd.envs import mpo_to_device
    from finitemptools_amd.mps import randommps
    psi = randommps(L, 2, D, seed=0).right_canonicalize()
    sites = [psi[i] for i in range(L)]
    mpo = mpo_to_device(models.j1j2_mpo(L, 1.0, J2, keep_j2=True))
    W1 = mpo_to_device(models.j1j2_mpo(L, 1.0, 1.0, keep_j2=True))
    W0 = mpo_to_device(models.j1j2_mpo(L, 1.0, 0.0, keep_j2=True))
    dW = [a - b for a, b in zip(W1, W0)]
    out = {"workload": f"J1-J2 L={L} D={D} J2={J2}, random right-canonical MPS"}
    GR.energy_gradients(mpo, sites, mode="store")  # warm-up
    for mode in ("store", "recompute"):
        torch.cuda.synchronize()
        ts = time.perf_counter()
        res = GR.energy_gradients(mpo, sites, mode=mode)
        torch.cuda.synchronize()
        g = GR.dE_dtheta(dW, res)
        out[mode] = {"s": round(time.perf_counter() - ts, 4), "dE_dJ2": g,
                     "resident_env_MB": round(res.stats["resident_env_bytes"] / 2 ** 20, 2),
                     "right_env_updates": res.stats["right_updates"]}
    hf = GR.hellmann_feynman(mpo_to_device(models.j1j2_dJ2_mpo(L)), sites)
    out["hellmann_feynman"] = hf
    out["adjoint_minus_hf"] = out["store"]["dE_dJ2"] - hf
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-blocksparse", action="store_true")
    ap.add_argument("--no-gradients", action="store_true")
    ap.add_argument("--no-plain", action="store_true",
                    help="leave out the plain per-sector FP64 sweep (about a minute)")
    ap.add_argument("--no-tdvp", action="store_true")
    ap.add_argument("--only", choices=("all", "headline", "blocksparse", "complex", "dense", "tdvp"),
                    default="all",
                    help="run one section only (profiling): the D=2048 middle-bond matvec, "
                         "the block-sparse ramp and sweeps, the complex or the dense D=512 matvec")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MiB",
                    help="first copy MiB of device memory three times (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: the H_eff engine has no multi-GPU path "
              f"(requested --gpus {args.gpus}).", file=sys.stderr)
        sys.exit(2)
    import torch
    if not torch.cuda.is_available():
        print("no HIP device", file=sys.stderr)
        sys.exit(1)
    from oracle import contractions as C

    if args.copy_calibration:
        # a device copy of known size: HBM counters of this kernel calibrate
        # the counter readings of the others (profiling runs only)
        n = args.copy_calibration * 2 ** 20 // 8
        a = torch.ones(n, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            torch.mul(a, 1.0, out=b)   # elementwise: reads and writes MiB each
        torch.cuda.synchronize()
        del a, b
    if args.only == "headline":
        hb = blocksparse_headline(args.steps, args.warmup, cpu=False, extras=False)
        print(json.dumps(hb))
        return
    if args.only == "blocksparse":
        print(json.dumps({"blocksparse_hubbard_L64_D2048": blocksparse_hubbard(plain=not args.no_plain)}))
        return
    if args.only == "complex":
        print(json.dumps({"ac2_prime_D512_c64_tfim": complex_tfim_matvec()}))
        return
    if args.only == "tdvp":
        print(json.dumps({"tdvp_tfim_L100_quench": tdvp_tfim_quench(precs=("f64",))}))
        return

    import ctypes
    from finitemptools_amd.engine import engine
    eng = engine(0)
    peak_meas = ctypes.c_double()
    eng.check(eng.lib.mh_probe_mfma_f64(eng.bind_stream(), ctypes.byref(peak_meas)))

    # ---- headline: the D = 2048 middle-bond block-sparse matvec
    print("block-sparse D=2048 middle-bond matvec ...", file=sys.stderr, flush=True)
    hb = blocksparse_headline(args.steps, args.warmup, cpu=not args.no_cpu)
    print(f"  {hb['tflops']:.2f} TFLOP/s", file=sys.stderr, flush=True)
    gms, gn, gfl = hb["gemm"]["ms"], hb["gemm"]["launches"], hb["gemm"]["flops"]
    achieved = gfl / (gms * 1e-3) / 1e12
    # algorithmic bytes of a grouped-GEMM launch: its two operand buffers read
    # once and its output written once, averaged over the plan's two stages
    # (step 1: L, x -> T1; step 3: T2, R -> y)
    from finitemptools_amd import blocksparse as bs
    _, xl, Lay, Ray, W1, W2, _, _, _ = middle_bond_case()
    sz = bs.make_plan("heff2", xl, Lay, Ray, W1, W2, "f64").sizes()
    alg_bytes = 8.0 * (sz[0] + sz[2] + sz[4] + sz[5] + sz[3] + sz[1]) / 2.0
    traffic = stored_traffic("blocksparse_hubbard_L64_D2048_pmc.md")

    if args.only == "dense":
        print(json.dumps({"ac2_prime_D512_f64": dense_d512(args, C, peak_meas.value)}))
        return
    out = {
        "metric": METRIC,
        "value": round(hb["tflops"], 3),
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(hb["sec_per_matvec"] * NVEC * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": ("synthetic: charge sectors of the Hubbard L=64 D=2048 middle bond "
                 "(tests/golden/hubbard_L64_D2048_bonds.json, from a DMRG ramp), seeded "
                 "Gaussian environment blocks and unit input vectors"),
        "config": {"workload": ("Hubbard L=64 D=2048 (U1xU1, d=4, w=6) middle-bond two-site "
                                "H_eff, grouped block-sparse FP64, 32 vectors per step"),
                   "L": HUBBARD_L, "D": 2048, "site": MID_SITE, "sectors": hb["sectors"],
                   "max_sector": hb["max_sector"], "vector_len": hb["vector_len"],
                   "gflop_per_matvec": round(hb["flops"] / 1e9, 4),
                   "executed_gflop_per_matvec": round(hb["executed_flops"] / 1e9, 4),
                   "flop_count": ("gflop_per_matvec: sum of 2mnk of the reference's graded "
                                  "contraction over every charge-allowed block (what value is "
                                  "quoted on); executed: the plan's products, which leave out "
                                  "the MPO's structural zeros (roofline.achieved uses these)"),
                   "vectors_per_step": NVEC},
        "roofline": {"bound": "mfma",
                     "kernel": "k_ggemm: the two grouped-GEMM launches of the block plan",
                     "achieved": round(achieved, 3), "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(achieved / FP64_PEAK_TFLOPS, 4),
                     "traffic": traffic,
                     "traffic_source": ("stored: profiles/blocksparse_hubbard_L64_D2048_pmc.md "
                                        "(rocprofv3 FETCH_SIZE + WRITE_SIZE of the same launches, "
                                        "copy-kernel calibrated)"),
                     "algorithmic_bytes": alg_bytes,
                     "measured_peak": round(peak_meas.value, 2),
                     "frac_of_measured_peak": round(achieved / peak_meas.value, 4),
                     "avg_launch_ms": round(gms / max(gn, 1), 5), "launches": gn,
                     "gflop_per_launch": round(gfl / max(gn, 1) / 1e9, 4),
                     "mode_stage": {"kernel": "k_bmode",
                                    "avg_launch_ms": round(hb["mode"]["ms"] / max(hb["mode"]["launches"], 1), 5),
                                    "launches": hb["mode"]["launches"]}},
        "blocksparse_matvec_D2048": {"f64_ms": round(hb["sec_per_matvec"] * 1e3, 4),
                                     "mixed": hb["mixed"],
                                     "f64_plain_per_sector": hb["f64_plain_per_sector"]},
    }
    if "cpu" in hb:
        out["cpu_baseline"] = hb["cpu"]
    print("dense D=512 matvecs ...", file=sys.stderr, flush=True)
    out["ac2_prime_D512_f64"] = dense_d512(args, C, peak_meas.value)
    if not args.no_blocksparse:
        print("block-sparse Hubbard ramp ...", file=sys.stderr, flush=True)
        bsh = blocksparse_hubbard(plain=not args.no_plain)
        out["blocksparse_hubbard_L64_D2048"] = bsh
        sw = bsh["sweep_D2048"]
        out["sweep_D2048_s"] = {k: sw[k]["s"] for k in sw}
        out["sweep_D2048_split_s"] = {k: sw[k].get("split_s") for k in sw}
        if "cpu_baseline" in out:
            # the CPU restatement's matvec rate applied to the FP64 sweep's
            # Lanczos matvec flops (splits and environment updates left out);
            # the sweep counts the plans' executed flops, scaled here to the
            # graded contraction's count by the middle bond's ratio
            rate = out["cpu_baseline"]["value"] * 1e12
            ratio = hb["flops"] / hb["executed_flops"]
            out["cpu_baseline"]["sweep_D2048_lanczos_s_extrapolated"] = round(
                sw["f64"]["matvec_flops"] * ratio / rate, 1)
    if not args.no_gradients:
        print("gradients ...", file=sys.stderr, flush=True)
        out["gradients_j1j2_L40_D256"] = gradients_j1j2()
        conv = {}
        for j2 in (0.0, 0.25, 0.5):
            conv[f"J2={j2}"] = gradients_converged_j1j2(J2=j2)
            print(f"  converged gradient J2={j2} done", file=sys.stderr, flush=True)
        # Majumdar-Ghosh point: the open even chain's ground state is the
        # nearest-neighbour dimer product, so dE/dJ2 = sum <S_i . S_i+2> = 0
        conv["majumdar_ghosh_abs_dE_dJ2"] = abs(conv["J2=0.5"]["dE_dJ2_adjoint"])
        out["gradients_j1j2_L40_D256_converged"] = conv
    if not args.no_tdvp:
        print("tdvp TFIM L=100 ...", file=sys.stderr, flush=True)
        out["tdvp_tfim_L100_quench"] = tdvp_tfim_quench()
        print("complex TFIM matvec D=512 ...", file=sys.stderr, flush=True)
        out["ac2_prime_D512_c64_tfim"] = complex_tfim_matvec()
    print("xxz L=20 D=64 ...", file=sys.stderr, flush=True)
    out["dmrg_xxz_L20_D64"] = xxz_dmrg_L20_D64(cpu=not args.no_cpu)
    print(json.dumps(out))


def dense_d512(args, C, peak_meas):
    """The dense microbenchmark: 32 two-site matvecs at D = 512 (XXZ w = 5,
    d = 2) and the one-site matvec at D = 1024, FP64 with its own launch
    roofline, and the mixed precisions against FP64."""
    D2, D1 = 512, 1024
    r2 = gpu_run(2, D2, args.steps, args.warmup, profile=True)
    W, L, R, xs = r2["inputs"]
    ref = C.ac2_prime(xs[0], W, W, L, R)
    err2 = float(np.linalg.norm(r2["y0"] - ref) / np.linalg.norm(ref))
    r1 = gpu_run(1, D1, max(1, args.steps // 2), args.warmup)
    mixed = {}
    # matrix-core arithmetic per GEMM flop and the published dense peak of
    # that format: FP32 MFMA 157.3 TFLOP/s; split-BF16 runs three BF16 MFMA
    # products per product, against the 2500 TFLOP/s BF16 peak
    fmt = {"f32": (1.0, 157.3, "FP32"), "bf16x3": (3.0, 2500.0, "BF16")}
    for prec in ("f32", "bf16x3"):
        rm = gpu_run(2, D2, args.steps, args.warmup, prec=prec)
        mult, peak, unit = fmt[prec]
        mixed[prec] = {"tflops_f64_equivalent": round(rm["tflops"], 3),
                       "mfma_tflops_in_format": round(mult * rm["tflops"], 1),
                       "format_peak_tflops": peak, "format": unit,
                       "frac_of_format_peak": round(mult * rm["tflops"] / peak, 4),
                       "ms": round(rm["sec_per_matvec"] * 1e3, 4),
                       "speedup_vs_f64": round(r2["sec_per_matvec"] / rm["sec_per_matvec"], 3),
                       "rel_err_vs_f64_oracle": float(np.linalg.norm(rm["y0"] - ref) / np.linalg.norm(ref))}
    gms, gn, gfl = r2["prof"]["gemm"]
    avg_ms = gms / max(gn, 1)
    achieved = gfl / max(gn, 1) / (avg_ms * 1e-3) / 1e12
    return {"workload": "ac2_prime two-site H_eff matvec, D=512, d=2, w=5 (XXZ), 32 vectors per step",
            "tflops": round(r2["tflops"], 3), "ms": round(r2["sec_per_matvec"] * 1e3, 4),
            "rel_err_vs_oracle": err2,
            "roofline": {"kernel": "k_heff_step1<5,4,5> and k_gemm (GEMM 3)",
                         "achieved": round(achieved, 3), "peak": FP64_PEAK_TFLOPS,
                         "frac": round(achieved / FP64_PEAK_TFLOPS, 4),
                         "frac_of_measured_peak": round(achieved / peak_meas, 4),
                         "avg_launch_ms": round(avg_ms, 5), "launches": gn,
                         "traffic": stored_traffic("heff_f64_ac2_D512_pmc_traffic.md"),
                         "traffic_source": "stored: profiles/heff_f64_ac2_D512_pmc_traffic.md"},
            "ac_prime_D1024_f64": {"tflops": round(r1["tflops"], 3),
                                   "ms": round(r1["sec_per_matvec"] * 1e3, 4)},
            "mixed": mixed}


if __name__ == "__main__":
    main()
