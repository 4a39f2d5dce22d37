"""Throughput of the MI355X Maxwell step on the 32^3-element N=7 FP64 box.

Workload (one "step" = one 5-stage RK step of cem_maxwell_op_rk over the
whole mesh): 32x32x32 periodic box on [0,2pi]^3, N=7 (16,777,216 GLL points
x 6 fields), per-element eps ~ U[1,4] (seed 1), mu = 1, upwind flux, initial
state a superposition of 8 plane waves with seeded wavevectors |k| <= 4
(seed 2), dt = 0.25*dxmin.  Synthetic data, generated in-process.

Metric: G(GLL-pt.field.stage)/s = npts*6*5*steps / seconds / 1e9, with the
state resident in HBM for the whole timed region.

Also reported:
  roofline      the fused stage kernel (5 launches per step): algorithmic
                bytes per launch = B_ALG x 6 x npts, B_ALG = 32 B per
                GLL-pt.field.stage (the 2N-storage floor: read u, read k,
                write k, write u), over its average launch duration from HIP
                events recorded on the library's stream.  The kernel skips
                the k read of stage 1 and the k write of stage 5 (a_1 = 0),
                so it moves 28.8 B per unit on average: achieved_moved and
                frac_moved price the same launches at that count.  traffic
                is the HBM bytes per launch that rocprofv3 --pmc FETCH_SIZE
                and WRITE_SIZE counted for this workload (the summary named
                in traffic_source);
  cpu_baseline  the CPU oracle (a loop-for-loop C restatement of the
                reference path, OpenMP) on the same workload for a bounded
                number of steps on this host's cores, and on one core;
  parity        relative L2 (bm1-weighted) between the GPU and the CPU
                oracle after those steps.
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# Algorithmic HBM bytes per GLL-pt.field.stage: the 2N-storage floor, read
# u, read k, write k, write u.
B_ALG = 32.0
# What the kernel moves of them per stage: the first stage need not read k
# (a_1 = 0) and the last need not write it (the next step's first stage
# ignores it), 24 B there.  Mean over a step: 28.8 B.
B_STAGE = (24.0, 32.0, 32.0, 32.0, 24.0)
B_MOVED = sum(B_STAGE) / len(B_STAGE)
FP64_PEAK_TFLOPS = 78.6


def f_alg(N):
    """Algorithmic FP64 flops per GLL-pt.field.stage: the contractions
    (6 (N+1)), curl combine, inverse mass and RK update (19.5), flux and lift
    per face node amortised over the element (80 / (N+1)); 77.5 at N = 7."""
    return 6.0 * (N + 1) + 19.5 + 80.0 / (N + 1)


def _hip():
    for name in ("libamdhip64.so", "/opt/rocm/lib/libamdhip64.so"):
        try:
            return C.CDLL(name)
        except OSError:
            continue
    raise RuntimeError("libamdhip64.so not found")


class StreamTimer:
    """hipEvent pair on a given stream."""

    def __init__(self, stream):
        self.hip = _hip()
        self.stream = C.c_void_p(stream)
        self.a, self.b = C.c_void_p(), C.c_void_p()
        assert self.hip.hipEventCreate(C.byref(self.a)) == 0
        assert self.hip.hipEventCreate(C.byref(self.b)) == 0

    def start(self):
        assert self.hip.hipEventRecord(self.a, self.stream) == 0

    def stop(self):
        assert self.hip.hipEventRecord(self.b, self.stream) == 0

    def ms(self):
        assert self.hip.hipEventSynchronize(self.b) == 0
        t = C.c_float()
        assert self.hip.hipEventElapsedTime(C.byref(t), self.a, self.b) == 0
        return float(t.value)


def copy_calibration(nbytes):
    """One device-to-device hipMemcpy of nbytes (the runtime's plain copy
    kernel): under rocprofv3 --pmc its FETCH_SIZE / WRITE_SIZE per dispatch
    give the counter's scale against a known byte count."""
    hip = _hip()
    a, b = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(a), C.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(C.byref(b), C.c_size_t(nbytes)) == 0
    assert hip.hipMemset(a, 1, C.c_size_t(nbytes)) == 0
    assert hip.hipDeviceSynchronize() == 0
    for _ in range(3):
        assert hip.hipMemcpy(b, a, C.c_size_t(nbytes), 3) == 0   # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)
    print(json.dumps({"copy_calibration_bytes": nbytes, "copies": 3}), flush=True)


def workload(nel=32, N=7):
    from nekcem_amd.mesh import BoxMesh
    from nekcem_amd import usr as cases
    mesh = BoxMesh((nel, nel, nel), N)
    n3 = mesh.n1 ** 3
    eps = np.repeat(cases.element_eps(mesh.nelt, seed=1), n3)
    x, y, z = mesh.coords()
    H, E = cases.plane_waves(x, y, z, seed=2)
    del x, y, z
    return mesh, eps, H, E


def batched(ncav=256, steps=50, warmup=3, seed=3, log=sys.stderr):
    """Batched-cavity speedup: ncav independent PEC cavities (n^3 elements,
    n ~ U{2..16}, N ~ U{5..9}, per-element eps ~ U[1,4], Gaussian-pulse Ez,
    each with its own CFL-0.25 dt) advanced together by one batch vs. one
    plain handle per cavity run one after another; same steps, same kernels,
    bit-identical states (checked on every cavity)."""
    from nekcem_amd import usr
    from nekcem_amd.maxwell import BatchSolver, MaxwellSolver
    from nekcem_amd.mesh import box_coords, box_dxmin
    cavs = usr.cavity_batch(ncav, seed)
    dts = np.array([0.25 * box_dxmin(c["nel"], c["N"], c["lo"], c["hi"]) for c in cavs])
    npts = sum(int(np.prod(c["nel"])) * (c["N"] + 1) ** 3 for c in cavs)

    def init(c):
        cv = cavs[c]
        x, y, z = box_coords(cv["nel"], cv["N"], cv["lo"], cv["hi"])
        return usr.gaussian_ez(x, y, z)

    b = BatchSolver(cavs)
    for c in range(ncav):
        b.set_state(c, *init(c))
    b.step(dts, warmup)
    b.sync()
    t0 = time.perf_counter()
    b.step(dts, steps)
    b.sync()
    t_batch = time.perf_counter() - t0
    print(f"[bench] batched: {ncav} cavities, {npts} points, {b.info()['norders']} orders, "
          f"{steps} steps in {t_batch:.3f} s", file=log, flush=True)
    t_seq = 0.0
    identical = True
    for c in range(ncav):
        cv = cavs[c]
        s = MaxwellSolver.from_box(cv["nel"], cv["N"], cv["lo"], cv["hi"], cv["bc"], eps_elem=cv["eps_elem"])
        s.set_state(*init(c))
        s.step(dts[c], warmup)
        s.sync()
        t0 = time.perf_counter()
        s.step(dts[c], steps)
        s.sync()
        t_seq += time.perf_counter() - t0
        gb = b.get_state(c)
        gs = s.get_state()
        identical &= all(np.array_equal(x, y) for x, y in zip(gb, gs))
        s.close()
    b.close()
    print(f"[bench] sequential: {t_seq:.3f} s, speedup {t_seq / t_batch:.2f}", file=log, flush=True)
    return {"metric": "batched speedup", "value": t_seq / t_batch, "unit": "x",
            "t_batched_s": t_batch, "t_sequential_s": t_seq, "cavities": ncav, "npts": npts,
            "steps": steps, "batched_G_per_s": npts * 30 * steps / t_batch / 1e9,
            "sequential_G_per_s": npts * 30 * steps / t_seq / 1e9,
            # both sides stream at least B_ALG bytes per unit: no schedule beats
            # HBM_PEAK_GBS / B_ALG G/s, so the batch is at most this much faster
            "hbm_bound_x": (HBM_PEAK_GBS / B_ALG) / (npts * 30 * steps / t_seq / 1e9),
            "bitwise_identical_to_sequential": bool(identical),
            "workload": f"{ncav} PEC cavities on [0,1]^3, n^3 elements n~U{{2..16}}, N~U{{5..9}}, "
                        f"per-element eps U[1,4], Gaussian Ez (width 0.2), own CFL-0.25 dt, seed {seed}"}


def adjoint(nel=16, N=7, nsteps=500, every=50, ndir=32, nprobe=64, seed=4, cap_steps=10, log=sys.stderr):
    """Checkpointed adjoint of the probe loss on the dielectric box: time
    relative to the forward, and agreement with central differences
    (h = 1e-6) of the HIP forward along ndir seeded random directions."""
    from nekcem_amd import usr
    from nekcem_amd.maxwell import MaxwellSolver
    from nekcem_amd.mesh import BoxMesh
    rng = np.random.default_rng(seed)
    lo, hi = (0.0, 0.0, 0.0), (1.0, 1.0, 1.0)
    mesh = BoxMesh((nel, nel, nel), N, lo, hi, ("PEC",) * 3)
    n = mesh.npts
    x, y, z = mesh.coords()
    eps = 2.0 + 0.3 * usr.smooth_field(x, y, z, seed)
    eps_t = 2.0 + 0.3 * usr.smooth_field(x, y, z, seed + 1)
    amp = usr.gaussian_source(x, y, z, mesh.bmn)
    probes = rng.choice(n, nprobe, replace=False)
    dt = mesh.dt(0.25)
    T = nsteps * dt
    src = dict(direction=(0.0, 0.0, 1.0), t0=0.15 * T, tw=0.05 * T)
    H0, E0 = np.zeros((3, n)), np.zeros((3, n))
    s = MaxwellSolver(mesh, eps=eps)
    s.set_source(amp, **src)

    def forward(e):
        s.set_materials(e)
        s.time = 0.0
        s.set_state(H0, E0)
        s.step(dt, nsteps)
        return s.get_state()[1][2, probes]

    target = forward(eps_t)
    # forward timing (plain steps) and the checkpointed adjoint
    s.set_materials(eps)
    s.time = 0.0
    s.set_state(H0, E0)
    s.sync()
    t0 = time.perf_counter()
    s.step(dt, nsteps)
    s.sync()
    t_fwd = time.perf_counter() - t0
    s.time = 0.0
    s.set_state(H0, E0)
    s.sync()
    t0 = time.perf_counter()
    s.adjoint_forward(dt, nsteps, every)
    s.sync()
    t_fwd_ckpt = time.perf_counter() - t0
    info = s.adjoint_info()
    Ez = s.get_state()[1][2, probes]
    diff = Ez - target
    loss = float(np.sum(diff * diff))
    lamE = np.zeros((3, n))
    np.add.at(lamE[2], probes, 2.0 * diff)
    t0 = time.perf_counter()
    ge, _ = s.adjoint_backward(np.zeros((3, n)), lamE)
    t_adj = time.perf_counter() - t0
    print(f"[bench] adjoint: forward {t_fwd:.3f} s, forward+checkpoints {t_fwd_ckpt:.3f} s, "
          f"backward {t_adj:.3f} s, loss {loss:.4e}, store {info}", file=log, flush=True)
    # the same sweep with the stage store capped at `cap_steps` steps: one
    # window held at a time, every window start recomputed
    s.adjoint_set_store(cap_steps)
    s.time = 0.0
    s.set_state(H0, E0)
    s.sync()
    t0 = time.perf_counter()
    s.adjoint_forward(dt, nsteps, every)
    s.sync()
    t_fwd_cap = time.perf_counter() - t0
    info_cap = s.adjoint_info()
    t0 = time.perf_counter()
    ge_cap, _ = s.adjoint_backward(np.zeros((3, n)), lamE)
    t_adj_cap = time.perf_counter() - t0
    s.adjoint_set_store(0)
    s.adjoint_release()
    capped = {"store_steps": cap_steps, "t_forward_checkpointed_s": t_fwd_cap, "t_adjoint_s": t_adj_cap,
              "adjoint_over_forward": t_adj_cap / t_fwd, "adjoint_store": info_cap,
              "gradient_bitwise_equal_to_default": bool(np.array_equal(ge_cap, ge))}
    print(f"[bench] adjoint, store capped at {cap_steps} steps: backward {t_adj_cap:.3f} s, "
          f"store {info_cap}", file=log, flush=True)
    h = 1e-6
    errs = []
    for _ in range(ndir):
        d = rng.standard_normal(n)
        lp = float(np.sum((forward(eps + h * d) - target) ** 2))
        lm = float(np.sum((forward(eps - h * d) - target) ** 2))
        fd = (lp - lm) / (2 * h)
        errs.append(abs(fd - float(np.dot(ge, d))) / abs(fd))
    s.close()
    err = max(errs) if errs else None
    if errs:
        print(f"[bench] adjoint vs FD over {ndir} directions: max rel err {err:.2e}", file=log, flush=True)
    return {"metric": "adjoint/fwd time", "value": t_adj / t_fwd, "unit": "x",
            "t_forward_s": t_fwd, "t_forward_checkpointed_s": t_fwd_ckpt, "t_adjoint_s": t_adj,
            "gradient_over_forward": (t_fwd_ckpt + t_adj) / t_fwd,
            "adjoint_store": info, "capped_store": capped,
            "fd_directions": ndir, "fd_h": h, "fd_max_rel_err": err, "loss": loss,
            "workload": f"PEC box {nel}^3 elements N={N} ({n} points), per-node eps = 2 + 0.3 s(x), "
                        f"Gaussian-pulse Ez source, {nprobe} probe nodes, {nsteps} steps, "
                        f"checkpoint every {every}"}


def drude_region(mesh, half=0.25):
    """Elements whose centre lies in the central cube of half side (of the
    unit box): (listed nodes, owning Drude element per node, element ids)."""
    n3 = mesh.n1 ** 3
    inside = np.ones(mesh.nelt, bool)
    for v in mesh.coords():
        inside &= np.abs(v.reshape(-1, n3).mean(axis=1) - 0.5) < half
    elems = np.nonzero(inside)[0]
    dindex = (elems[:, None] * n3 + np.arange(n3)[None, :]).ravel()
    owner = np.repeat(np.arange(len(elems)), n3)
    return dindex, owner, elems


def dispersive(nel=24, N=9, nsteps=2000, every=100, ndir=2, nprobe=64, seed=5, log=sys.stderr):
    """Dispersive box: a Drude medium (omega_p = 1, gamma = 0.1 per element)
    in the central cube of half side, vacuum elsewhere, PEC walls, upwind,
    Gaussian-pulse Ez source.  Forward throughput over nsteps, then the
    checkpointed adjoint of the probe loss with respect to per-element
    gamma and omega_p (target from a forward with seeded perturbed
    parameters), checked against central differences (h = 1e-6) of the
    HIP forward along ndir seeded directions."""
    from nekcem_amd import usr
    from nekcem_amd.maxwell import MaxwellSolver
    from nekcem_amd.mesh import BoxMesh
    rng = np.random.default_rng(seed)
    mesh = BoxMesh((nel, nel, nel), N, (0.0, 0.0, 0.0), (1.0, 1.0, 1.0), ("PEC",) * 3)
    n = mesh.npts
    x, y, z = mesh.coords()
    dindex, owner, elems = drude_region(mesh)
    ne = len(elems)
    gam, wp = np.full(ne, 0.1), np.full(ne, 1.0)
    gam_t = gam * rng.uniform(0.8, 1.2, ne)
    wp_t = wp * rng.uniform(0.9, 1.1, ne)
    probes = rng.choice(n, nprobe, replace=False)
    dt = mesh.dt(0.25)
    T = nsteps * dt
    H0, E0 = np.zeros((3, n)), np.zeros((3, n))
    s = MaxwellSolver.from_box((nel, nel, nel), N, (0.0, 0.0, 0.0), (1.0, 1.0, 1.0), ("PEC",) * 3)
    s.set_source(usr.gaussian_source(x, y, z, mesh.bmn), direction=(0.0, 0.0, 1.0), t0=0.15 * T, tw=0.05 * T)

    def forward(g, w, timed=False):
        s.set_drude(dindex, g[owner], (w * w)[owner])
        s.time = 0.0
        s.set_state(H0, E0)
        s.sync()
        t0 = time.perf_counter()
        s.step(dt, nsteps)
        s.sync()
        el = time.perf_counter() - t0
        Ez = s.get_state()[1][2, probes]
        return (Ez, el) if timed else Ez

    target = forward(gam_t, wp_t)
    _, t_fwd = forward(gam, wp, timed=True)
    s.adjoint_set_grads(False)      # the gradient asked for is w.r.t. gamma and omega_p only
    s.set_drude(dindex, gam[owner], (wp * wp)[owner])
    s.time = 0.0
    s.set_state(H0, E0)
    s.sync()
    t0 = time.perf_counter()
    s.adjoint_forward(dt, nsteps, every)
    s.sync()
    t_fwd_ckpt = time.perf_counter() - t0
    info = s.adjoint_info()
    diff = s.get_state()[1][2, probes] - target
    loss = float(np.sum(diff * diff))
    lamE = np.zeros((3, n))
    np.add.at(lamE[2], probes, 2.0 * diff)
    t0 = time.perf_counter()
    s.adjoint_backward(np.zeros((3, n)), lamE)
    ga, gb = s.adjoint_drude_grad()
    t_adj = time.perf_counter() - t0
    g_gam = np.bincount(owner, weights=ga, minlength=ne)
    g_wp = 2.0 * wp * np.bincount(owner, weights=gb, minlength=ne)
    lay = s.layout()
    print(f"[bench] dispersive: stage layout {lay}", file=log, flush=True)
    print(f"[bench] dispersive: forward {t_fwd:.3f} s ({nsteps} steps), forward+checkpoints "
          f"{t_fwd_ckpt:.3f} s, backward {t_adj:.3f} s, loss {loss:.4e}, adjoint {info}",
          file=log, flush=True)
    h = 1e-6
    errs = []
    for _ in range(ndir):
        dg, dw = rng.standard_normal(ne), rng.standard_normal(ne)
        lp = float(np.sum((forward(gam + h * dg, wp + h * dw) - target) ** 2))
        lm = float(np.sum((forward(gam - h * dg, wp - h * dw) - target) ** 2))
        fd = (lp - lm) / (2 * h)
        errs.append(abs(fd - float(np.dot(g_gam, dg) + np.dot(g_wp, dw))) / abs(fd))
    s.close()
    err = max(errs) if errs else None
    if errs:
        print(f"[bench] dispersive adjoint vs FD over {ndir} directions: max rel err {err:.2e}",
              file=log, flush=True)
    return {"metric": "G(GLL-pt*field*stage)/s FP64 forward; adjoint/fwd time",
            "value": n * 30 * nsteps / t_fwd / 1e9, "unit": "G(GLL-pt*field*stage)/s",
            "adjoint_over_forward": t_adj / t_fwd,
            "t_forward_s": t_fwd, "t_forward_checkpointed_s": t_fwd_ckpt, "t_adjoint_s": t_adj,
            "adjoint_window_steps": info["window"], "adjoint_device_bytes": info["bytes"],
            "stage_layout": lay,
            "fd_directions": ndir, "fd_h": h, "fd_max_rel_err": err, "loss": loss,
            "workload": f"PEC box {nel}^3 elements N={N} ({n} points), Drude (omega_p=1, gamma=0.1) in "
                        f"{ne} central elements, vacuum elsewhere, Gaussian-pulse Ez source, "
                        f"{nprobe} probe nodes, {nsteps} steps, checkpoint every {every}; "
                        f"gradient w.r.t. per-element gamma and omega_p"}


def load_traffic(nel, N, npts):
    """HBM bytes per stage launch of this workload from its rocprofv3 --pmc
    summary under profiles/ (FETCH_SIZE and WRITE_SIZE per dispatch, scaled
    by the counters' reading for a device copy of known size): (bytes,
    summary file) or (None, None) when no summary covers the workload."""
    name = f"stage_kernel_box{nel}_N{N}_traffic"
    p = os.path.join(ROOT, "profiles", name + ".json")
    if not os.path.exists(p):
        return None, None
    d = json.load(open(p))
    if d.get("npts") != npts:
        return None, None
    return d.get("hbm_bytes_per_launch"), f"profiles/{name}.txt"


def cpu_facts():
    """nproc, the cores this process may run on, and the CPU model."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count()
    return {"nproc": os.cpu_count(), "usable_cores": usable, "cpu_model": model}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--nel", type=int, default=32)
    ap.add_argument("--N", type=int, default=7)
    ap.add_argument("--cpu-steps", type=int, default=16,
                    help="steps of the CPU oracle for cpu_baseline (0 = skip)")
    ap.add_argument("--cpu-steps-1core", type=int, default=1,
                    help="steps of the one-thread CPU oracle sample (0 = skip)")
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--layout", type=int, default=-1,
                    help="stage-kernel layout of the main workload (-1: the autotuner picks)")
    ap.add_argument("--batched-steps", type=int, default=50,
                    help="steps of the batched-cavity speedup measurement (0 = skip)")
    ap.add_argument("--batched-only", action="store_true")
    ap.add_argument("--adjoint-dirs", type=int, default=32,
                    help="FD directions of the adjoint check (0 = skip the adjoint measurement)")
    ap.add_argument("--adjoint-only", action="store_true")
    ap.add_argument("--adjoint-steps", type=int, default=500)
    ap.add_argument("--adjoint-every", type=int, default=50)
    ap.add_argument("--dispersive-dirs", type=int, default=8,
                    help="FD directions of the dispersive (Drude) adjoint check (-1 = skip the measurement)")
    ap.add_argument("--dispersive-steps", type=int, default=2000)
    ap.add_argument("--dispersive-only", action="store_true")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="only run a device-to-device copy of BYTES (counter calibration)")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.batched_only:
        print(json.dumps(batched(steps=args.batched_steps)), flush=True)
        return
    if args.adjoint_only:
        print(json.dumps(adjoint(nsteps=args.adjoint_steps, every=args.adjoint_every, ndir=args.adjoint_dirs)),
              flush=True)
        return
    if args.dispersive_only:
        print(json.dumps(dispersive(nsteps=args.dispersive_steps, ndir=max(args.dispersive_dirs, 0))),
              flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}): the step is "
              f"built for one GPU's element set; sharding elements in blocks across GPUs "
              f"with a per-stage exchange of partition-face traces is not built",
              file=sys.stderr)
        sys.exit(2)

    from nekcem_amd.maxwell import MaxwellSolver

    t0 = time.time()
    mesh, eps, H, E = workload(args.nel, args.N)
    npts = mesh.npts
    solver = MaxwellSolver(mesh, eps=eps, graph=not args.no_graph)
    dt = mesh.dt(0.25)
    solver.set_state(H, E)
    setup_s = time.time() - t0
    print(f"[bench] setup {setup_s:.1f}s npts={npts} nelt={mesh.nelt} dt={dt:.3e} "
          f"info={solver.info()}", file=sys.stderr, flush=True)

    if args.layout >= 0:
        solver.set_layout(args.layout)
    solver.step(dt, args.warmup)
    solver.sync()
    layout = solver.layout()
    print(f"[bench] stage layout (autotuned) {layout}", file=sys.stderr, flush=True)
    timer = StreamTimer(solver.stream())
    solver.sync()
    w0 = time.perf_counter()
    timer.start()
    solver.step(dt, args.steps)
    timer.stop()
    solver.sync()
    wall = time.perf_counter() - w0
    ev_ms = timer.ms()
    solver.check_finite()
    print(f"[bench] timed {args.steps} steps: wall {wall*1e3:.2f} ms, events {ev_ms:.2f} ms",
          file=sys.stderr, flush=True)

    units = npts * 6 * 5 * args.steps
    value = units / wall / 1e9
    launches = 5 * args.steps
    avg_launch_s = ev_ms * 1e-3 / launches
    alg_bytes = B_ALG * 6 * npts
    achieved = alg_bytes / avg_launch_s / 1e9
    moved = B_MOVED * 6 * npts / avg_launch_s / 1e9
    traffic, traffic_src = load_traffic(args.nel, args.N, npts)

    # CPU baseline + parity on the same workload (bounded sample)
    cpu = None
    parity = None
    if args.cpu_steps > 0:
        from oracle import oracle as O
        ref = O.BoxCase(O.BoxSpec(nel=mesh.nel, N=args.N), eps=eps)
        cores = O.num_threads()
        c0 = time.perf_counter()
        out = ref.run(H, E, dt, args.cpu_steps)
        cpu_s = time.perf_counter() - c0
        cpu = {"value": npts * 30 * args.cpu_steps / cpu_s / 1e9, "unit": "G(GLL-pt*field*stage)/s",
               "cores": cores, "kind": "port",
               "sample": f"{args.cpu_steps} RK5 step(s) of the same {args.nel}^3-element N={args.N} "
                         f"workload, C oracle with OpenMP on {cores} threads, {cpu_s:.1f} s"}
        cpu.update(cpu_facts())
        if args.cpu_steps_1core > 0:
            O.set_num_threads(1)
            c0 = time.perf_counter()
            ref.run(H, E, dt, args.cpu_steps_1core)
            one_s = time.perf_counter() - c0
            O.set_num_threads(cores)
            cpu["value_1core"] = npts * 30 * args.cpu_steps_1core / one_s / 1e9
            cpu["sample_1core"] = (f"{args.cpu_steps_1core} RK5 step(s) of the same workload, "
                                   f"C oracle on 1 thread, {one_s:.1f} s")
        g = MaxwellSolver(mesh, eps=eps)
        g.set_state(H, E)
        g.step(dt, args.cpu_steps)
        gH, gE = g.get_state()
        g.close()
        bm = mesh.bmn
        worst = 0.0
        for a, b in ((gH, out["H"]), (gE, out["E"])):
            for c in range(3):
                den = np.sqrt(np.sum(bm * b[c] ** 2))
                if den > 0:
                    worst = max(worst, float(np.sqrt(np.sum(bm * (a[c] - b[c]) ** 2)) / den))
        parity = {"rel_l2_max_over_fields": worst, "steps": args.cpu_steps,
                  "norm": "per field, sqrt(sum bm1 (gpu-cpu)^2)/sqrt(sum bm1 cpu^2)"}

    bat = batched(steps=args.batched_steps) if args.batched_steps > 0 else None
    adj = (adjoint(nsteps=args.adjoint_steps, every=args.adjoint_every, ndir=args.adjoint_dirs)
           if args.adjoint_dirs > 0 else None)
    disp = (dispersive(nsteps=args.dispersive_steps, ndir=args.dispersive_dirs)
            if args.dispersive_dirs >= 0 else None)

    line = {
        "metric": f"G(GLL-pt*field*stage)/s FP64, {args.nel}^3el N={args.N}",
        "value": value,
        "unit": "G(GLL-pt*field*stage)/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"periodic box {args.nel}^3 hexahedra, N={args.N}, "
                               f"{npts} GLL points x 6 fields, per-element eps U[1,4], "
                               f"8 plane waves, upwind, CFL 0.25, RK5 (5 stages/step)",
                   "nelt": mesh.nelt, "N": args.N, "npts": npts,
                   "graph": not args.no_graph, "stage_layout": layout},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "traffic_source": traffic_src,
                     "kernel": "nkc::stage_kernel (fused stage)",
                     "avg_launch_us": avg_launch_s * 1e6,
                     "alg_bytes_per_launch": alg_bytes, "alg_bytes_per_unit": B_ALG,
                     "moved_bytes_per_unit": B_MOVED, "achieved_moved": moved,
                     "frac_moved": moved / HBM_PEAK_GBS,
                     "fp64": {"alg_flops_per_unit": f_alg(args.N),
                              "achieved_tflops": f_alg(args.N) * units / wall / 1e12,
                              "peak_tflops": FP64_PEAK_TFLOPS,
                              "frac": f_alg(args.N) * units / wall / 1e12 / FP64_PEAK_TFLOPS}},
        "cpu_baseline": cpu,
        "parity": parity,
        "batched": bat,
        "adjoint": adj,
        "dispersive": disp,
    }
    solver.close()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
