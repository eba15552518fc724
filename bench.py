"""Benchmark of the plane-current PEEC solve on one MI355X.

Headline workload (the timed steps): the two-layer board, 256 x 256 cells per layer at 0.25 mm, a 0.5 mm slot in
the lower plane, 20 stitching vias, one Manhattan route, solved at 10 frequencies log-spaced 1 MHz - 1 GHz, plain
FP64, GMRES(50) to a relative residual of 1e-10. One step = the 10 solves (one route at every frequency) with the
board and the right-hand sides resident, run as one batch (each solve follows the reference's GMRES and is
bit-identical to its solve alone). value = solves per second over the timed steps. The same batch is then timed in
mixed precision (the reference's restart cycles in FP32, every restart an FP64 refinement step) and with
FP32 Krylov bases, each reported with its solutions' relative errors against an FP64 solve of the same system at tol
1e-13.

Also reported: the L_p product's (the FFT matvec's) roofline. Every L_p launch of the timed steps is bracketed by
HIP events on the library's stream (gt_ctx_lp_timing); achieved = the algorithmic bytes of a launch (per right-hand
side: strip currents in and out, one read and one write of the padded spectra per transform; once per launch: the
kernel quadrants) / the average launch. traffic = the HBM bytes of an L_p launch from the FETCH_SIZE / WRITE_SIZE runs
committed under profiles/, scaled to the timed launches' average right-hand sides. Beside it: a single-right-hand-side
product over 200 launches; the served what-if stream (serve_stream: p50/p99 latency and solves/s of the continuous-
batching server beside the same stream solved one solve at a time), the 512^2 what-if latency (serve_whatif_512) and
the served backlog against the plain FP64 solver one solve at a time (serve_backlog); the oracle's CPU solve of one
frequency (best of 3, with the host's CPU model, affinity and library versions); and the six-layer 2048^2 board
(large: FP64, FP32-basis and mixed solves to 1e-8 with their solves/s and errors, its L_p product's GB/s against peak,
and the same product on the host); and the four-layer 512^2 board's 256 right-hand sides (batched: 64 nets x 4
frequencies in blocks of 16, mixed to 1e-8, FP64 on two of the blocks beside it, with a sample's errors); and the
copper rasteriser (raster: build_grid of a 200 x 200 mm pour with 9801 antipads at 0.1 mm, sub-grid 4, against the
oracle's rate on a 10 mm cut).

Progress lines go to stderr; stdout carries the one JSON line.

python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline] [--no-serve] [--no-large] [--no-batched]
python bench.py --lp-only [--workload W] [--lp-reps R] [--lp-batch N]   (the counter runs: L_p launches of N
                                                         right-hand sides, then a 64 MiB device copy as the scale)
python bench.py --large      (the six-layer 2048^2 board alone)
python bench.py --batched    (the four-layer 512^2 board's 256 right-hand sides alone)
python bench.py --raster     (the copper rasteriser's 200 x 200 mm board alone)
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time
from typing import Optional

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "PEEC solves/sec + matvec HBM GB/s; p50/p99 what-if latency vs plain FP64 at 1e-8"
HBM_PEAK_GBS = 8000.0


def lp_bytes(st) -> tuple:
    """The algorithmic bytes of one L_p product, split as (per right-hand side, per launch): the strip currents in
    and out once and one HBM read and write of the padded spectra per transform, 2c n_o + 4c S P0 P1, for every
    right-hand side; the kernel quadrants, r S^2 (P0/2+1)(P1/2+1), once per launch whatever its right-hand sides
    (c = 16, r = 8: FP64), summed over the orientations."""
    c, r, S = 16, 8, st.n_sheets
    vec = ker = 0
    for P, n_o in ((st.Px, st.nxb), (st.Py, st.nyb)):
        if n_o:
            vec += 2 * c * n_o + 4 * c * S * P[0] * P[1]
            ker += r * S * S * (P[0] // 2 + 1) * (P[1] // 2 + 1)
    return vec, ker


def lp_algorithmic_bytes(st, nr: int = 1) -> int:
    """B of one launch of nr right-hand sides (lp_bytes): nr times the vector bytes, the kernel bytes once."""
    vec, ker = lp_bytes(st)
    return nr * vec + ker


def log(msg: str):
    """A progress line on stderr (the JSON result alone goes to stdout)."""
    print(f"[bench {time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)


def time_lp(st, reps: int = 200, nr: int = 1):
    """Average duration of one L_p launch of nr right-hand sides on the library's stream (HIP events on that
    stream)."""
    import ctypes as C

    import torch

    from groundtruth_amd import _lib
    dev = f"cuda:{st.device}"
    rng = np.random.default_rng(4)
    I = [torch.from_numpy(rng.standard_normal(st.ns) + 1j * rng.standard_normal(st.ns)).to(dev) for _ in range(nr)]
    out = [torch.empty_like(v) for v in I]
    L = _lib.lib()
    h = st.handle
    X = (C.c_void_p * nr)(*[v.data_ptr() for v in I])
    Y = (C.c_void_p * nr)(*[v.data_ptr() for v in out])
    stream = torch.cuda.ExternalStream(L.gt_ctx_stream(st.ctx.handle), device=dev)
    torch.cuda.synchronize()
    for _ in range(5):
        _lib.check(L.gt_lp_matvec_batch(h, nr, X, Y))
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(reps):
        L.gt_lp_matvec_batch(h, nr, X, Y)
    e1.record(stream)
    e1.synchronize()
    return e0.elapsed_time(e1) / reps * 1e-3


def measured_traffic(workload: str, name: Optional[str] = None):
    """(HBM bytes per L_p launch, right-hand sides of that launch) from the committed FETCH_SIZE / WRITE_SIZE runs
    of this workload (lines `traffic_bytes_per_launch: N` and `traffic_rhs_per_launch: R` in
    profiles/lp_traffic_<workload>.md, or profiles/<name>), or None."""
    path = os.path.join(ROOT, "profiles", name or f"lp_traffic_{workload}.md")
    if not os.path.exists(path):
        return None
    vals = {}
    with open(path) as fh:
        for line in fh:
            for k in ("traffic_bytes_per_launch", "traffic_rhs_per_launch"):
                if line.startswith(k + ":"):
                    vals[k] = int(line.split(":")[1])
    if "traffic_bytes_per_launch" not in vals:
        return None
    return vals["traffic_bytes_per_launch"], vals.get("traffic_rhs_per_launch", 1)


def lp_only(args):
    """Set-up, a few L_p products, then a plain device copy of 64 MiB (the scale for the memory counters)."""
    import torch

    from groundtruth_amd import boards, peec
    board = boards.WORKLOADS[args.workload]()
    st = peec.PlaneStack(board.layers, board.branches, device=0)
    st.handle
    t_lp = time_lp(st, reps=args.lp_reps, nr=args.lp_batch)
    n = 64 << 20
    a = torch.empty(n, dtype=torch.uint8, device="cuda:0").fill_(1)
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"workload": args.workload, "rhs_per_launch": args.lp_batch, "lp_us": t_lp * 1e6,
                      "bytes_per_launch": lp_algorithmic_bytes(st, args.lp_batch), "copy_bytes": n}))


def steps_only(args):
    """Set-up, one warm-up step and --steps timed steps of the workload's batch in one precision (--steps-only
    fp64 | fp32-basis | mixed), then three plain device copies of 64 MiB (the scale for the memory counters): the
    kernels of the headline steps alone, for a per-kernel profile."""
    import torch

    from groundtruth_amd import boards, peec
    board = boards.WORKLOADS[args.workload]()
    route = board.routes[0]
    st = peec.PlaneStack(board.layers, board.branches, device=0)
    ex = peec.RouteExcitation(st, route.trace, route.driver, route.receiver)
    ops = [st.operator(f) for f in board.freqs]
    rhs = []
    for f in board.freqs:
        top, J = ex.rhs(f)
        rhs.append(torch.from_numpy(np.concatenate([top, J[st.keep].astype(complex)])).to("cuda:0"))
    x = [torch.zeros_like(r) for r in rhs]
    prec = args.steps_only

    def step():
        res = st.gmres_batch_dev(ops, rhs, x, board.tol, restart=board.restart, precision=prec)
        if not all(r.converged for r in res):
            raise RuntimeError(f"{prec} GMRES did not converge")
        return [int(r.iterations) for r in res]

    step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        its = step()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    n = 64 << 20
    a = torch.empty(n, dtype=torch.uint8, device="cuda:0").fill_(1)
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"workload": args.workload, "precision": prec, "steps": args.steps,
                      "ms_per_step": dt / args.steps * 1e3, "iterations": its, "copy_bytes": n}))


def cpu_lp_large(st, I: np.ndarray, y_dev: np.ndarray) -> dict:
    """The six-layer board's L_p product on the host by the oracle's numpy / scipy path (the reference's FFT matvec,
    peec.py:2054-2135: fft2 / ifft2 with workers=-1 and the numpy multiply-add over sheet pairs), with the device
    stack's own kernel quadrants: timed once (one product takes tens of seconds on the host), and its result held to
    the device product of the same vector (a full-size check)."""
    from oracle import peec as O

    from groundtruth_amd import peec
    o = O.PlaneStack.__new__(O.PlaneStack)
    o.nx, o.ny, o.n_sheets = st.nx, st.ny, st.n_sheets
    o.xa, o.ya, o.nxb, o.nyb, o.ns = st.xa, st.ya, st.nxb, st.nyb, st.ns
    o.Px, o.Py = tuple(st.Px), tuple(st.Py)
    o.dL = o.M_near = None              # every strip of this board a pitch wide
    # the device keeps its quadrants [qj][qi][t][s] pre-scaled by 1 / (P0 P1); the oracle's are [qi][qj][t][s], unscaled
    o._kq = {"x": np.ascontiguousarray(peec.kernel_quadrant(st, 0).transpose(1, 0, 2, 3)) * (o.Px[0] * o.Px[1]),
             "y": np.ascontiguousarray(peec.kernel_quadrant(st, 1).transpose(1, 0, 2, 3)) * (o.Py[0] * o.Py[1])}
    t0 = time.perf_counter()
    y = o.Lp_matvec(I)
    dt = time.perf_counter() - t0
    err = float(np.max(np.abs(y - y_dev)) / np.max(np.abs(y)))
    return {"seconds": dt, "cores": len(os.sched_getaffinity(0)), "kind": "port",
            "sample": "one L_p product of the six-layer 2048^2 board (FFT 4096 x 4096, 6 sheets, both orientations), "
                      "oracle's numpy / scipy path, timed once", "rel_diff_vs_device": err}


def large(cpu: bool = True) -> dict:
    """The six-layer 2048^2 board (full inter-layer mutual blocks, two planes split into Voronoi regions, 200
    stitching vias, 1 A at 1 GHz): set-up, one plain FP64 solve to 1e-8 (its solves/s), one FP64 GMRES over an FP32
    basis and one mixed-precision solve (FP32 inner GMRES under FP64 refinement), each with its true residual and its
    distance from the FP64 solve, the L_p product's achieved HBM GB/s against peak (10 one-right-hand-side launches
    timed by HIP events on the library's stream), and the same product on the host."""
    import torch

    from groundtruth_amd import boards, peec
    board = boards.large()
    route = board.routes[0]
    f = float(board.freqs[0])
    import gc
    gc.collect()                    # the earlier sections' stacks and vectors, before the footprint check
    free, total = torch.cuda.mem_get_info(0)
    log(f"six-layer 2048^2 board: set-up ({free / 1e9:.0f} of {total / 1e9:.0f} GB free)")
    t0 = time.perf_counter()
    st = peec.PlaneStack(board.layers, board.branches, device=0)
    ex = peec.RouteExcitation(st, route.trace, route.driver, route.receiver)
    st.handle
    setup_s = time.perf_counter() - t0
    rhs_top, J = ex.rhs(f)
    rhs = np.concatenate([rhs_top, np.asarray(J, float)[st.keep].astype(complex)])
    op = st.operator(f)
    out = {"workload": "large: six layers of 2048x2048 cells at 0.05 mm, two planes split into Voronoi regions, 200 "
                       "stitching vias (999 branches), 1 A at 1 GHz, GMRES(100) to 1e-8",
           "unknowns": int(st.nb + st.n_keep), "setup_s": round(setup_s, 2), "woodbury": st.woodbury_form,
           "schur": st.schur_stats(), "device_gb": round(st.device_bytes() / 1e9, 2)}
    sols = {}
    for prec in ("fp64", "fp32-basis", "mixed"):
        log(f"six-layer board: one {prec} solve")
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        sol = ex.solve_many([f], tol=board.tol, restart=board.restart, precision=prec)[0]
        dt = time.perf_counter() - t1
        res = float(np.linalg.norm(rhs - st.matvec(op, sol.x)) / np.linalg.norm(rhs))
        sols[prec] = sol
        out[prec] = {"seconds": round(dt, 3), "solves_per_s": 1.0 / dt, "iterations": sol.iterations,
                     "true_residual": res, "Z_loop": [sol.Z_loop.real, sol.Z_loop.imag]}
    for prec in ("mixed", "fp32-basis"):
        out[prec]["x_vs_fp64"] = float(np.linalg.norm(sols[prec].x - sols["fp64"].x) / np.linalg.norm(sols["fp64"].x))
        out[prec]["Z_vs_fp64"] = abs(sols[prec].Z_loop - sols["fp64"].Z_loop) / abs(sols["fp64"].Z_loop)
        out[prec]["speedup_vs_fp64"] = out["fp64"]["seconds"] / out[prec]["seconds"]
    del sols
    t_lp = time_lp(st, reps=10, nr=1)
    B = lp_algorithmic_bytes(st, 1)
    out["lp"] = {"us_per_launch": t_lp * 1e6, "bytes_per_launch": B, "achieved": B / t_lp / 1e9,
                 "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": B / t_lp / 1e9 / HBM_PEAK_GBS}
    mt = measured_traffic("large", "lp_traffic_six_layer_2048.md")
    out["lp"]["traffic"] = mt[0] if mt else None
    if cpu:
        log("six-layer board: the L_p product on the host")
        rng = np.random.default_rng(4)
        I = rng.standard_normal(st.ns) + 1j * rng.standard_normal(st.ns)
        y = st.Lp_matvec(I)
        out["cpu_lp"] = cpu_lp_large(st, I, y)
        out["cpu_lp"]["host_over_device"] = out["cpu_lp"]["seconds"] / t_lp
    del st, ex, op
    torch.cuda.empty_cache()
    return out


SERVE_TOL = 1e-8                # the served stream's tolerance (plain FP64 GMRES)
SERVE_RATE = 4.0                # served stream: Poisson arrivals per second
SERVE_SECONDS = 15.0            # served stream: the window the arrivals fall in


def serve_boards():
    """The served stream's resident boards (name, Board, stack): 1, 4 and 2 layers at 128^2, 128^2 and 256^2 cells."""
    from groundtruth_amd import boards, peec
    out = []
    for name, b in (("single_plane_128", boards.scaled("single_plane", 128)),
                    ("four_layer_128", boards.scaled("four_layer", 128, n_nets=4)),
                    ("two_layer_256", boards.two_layer())):
        out.append((name, b, peec.PlaneStack(b.layers, b.branches, device=0)))
    return out


def serve_requests(bds, rate: float, seconds: float, seed: int = 1):
    """A seeded Poisson stream: arrival times at `rate` per second over `seconds`; each request a board drawn
    uniformly, one of its nets and 1-4 of its frequencies; half of the requests are what-if edits of the resident
    board: on boards of two or more layers half of those add 1-2 stitching vias at copper cells (a what-if stack on
    the board's factor), the rest (and every edit of a one-layer board) move the board's slot (a topology of its
    own on the board's kernel spectra: its own numbering, factor and session)."""
    from groundtruth_amd import boards
    from groundtruth_amd.serve import Request
    rng = np.random.default_rng(seed)
    t, out = 0.0, []
    while True:
        t += rng.exponential(1.0 / rate)
        if t >= seconds:
            return out
        name, b, st = bds[int(rng.integers(len(bds)))]
        route = b.routes[int(rng.integers(len(b.routes)))]
        fs = [float(f) for f in rng.choice(b.freqs, size=int(rng.integers(1, min(4, len(b.freqs)) + 1)),
                                           replace=False)]
        extra, layers = [], None
        if rng.random() < 0.5:
            if len(b.layers) > 1 and rng.random() < 0.5:
                while len(extra) < int(rng.integers(1, 3)):
                    i, j = (int(v) for v in rng.integers(1, st.nx - 1, 2))
                    extra += boards._stitch(b.layers, i, j, f"whatif{len(out)}_{len(extra)}")
            else:
                layers, _ = boards.moved_slot(b, rng, protect=[route.driver, route.receiver])
        out.append((t, Request(name, route, fs, extra=extra, layers=layers, tol=SERVE_TOL)))


def serve_stream(rate: float, seconds: float, batched: bool = True) -> dict:
    """The served what-if stream in real time: requests are submitted at their arrival times and latency runs from
    the arrival to the last solve's end. batched: the continuous-batching server; else the same requests one solve
    at a time in arrival order (RouteExcitation.solve, plain FP64)."""
    import torch

    from groundtruth_amd import peec
    from groundtruth_amd.serve import Server
    t_set = time.perf_counter()
    bds = serve_boards()
    srv = Server(capacity=16)
    for name, b, st in bds:
        srv.add_board(name, st, routes=b.routes)
    setup_s = time.perf_counter() - t_set
    stream = serve_requests(bds, rate, seconds)
    stacks = {name: st for name, b, st in bds}
    # the boards' analysis (untimed): every net at every frequency, the solutions the what-ifs start from
    t_pr = time.perf_counter()
    for name, b, st in bds:
        srv.prime(name, b.freqs, tol=SERVE_TOL)
    prime_s = time.perf_counter() - t_pr
    # one what-if of each kind per board (untimed): the first use of the warm start's and the moved slot's routines
    from groundtruth_amd import boards as B
    from groundtruth_amd.serve import Request
    rng = np.random.default_rng(99)
    for name, b, st in bds:
        r = b.routes[0]
        if len(b.layers) > 1:
            srv.submit(Request(name, r, [b.freqs[0]], extra=B._stitch(b.layers, 3, 3, "warm-up"), tol=SERVE_TOL))
        srv.submit(Request(name, r, [b.freqs[0]], tol=SERVE_TOL,
                           layers=B.moved_slot(b, rng, protect=[r.driver, r.receiver])[0]))
    srv.drain()
    srv.reset_stats()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    lat = []
    solves = 0
    if batched:
        i = 0
        while i < len(stream) or srv.busy():
            now = time.perf_counter() - t0
            while i < len(stream) and stream[i][0] <= now:
                srv.submit(stream[i][1], now=t0 + stream[i][0])
                i += 1
            if srv.busy():
                srv.turn()
            elif i < len(stream):
                time.sleep(max(0.0, stream[i][0] - (time.perf_counter() - t0)))
        st_ = srv.stats()
        failed = st_["failed"]
        lat = [r.latency for r in srv.completed if r.error is None]
        solves = st_["solves"]
        extra = {"peak_sessions": st_["peak_sessions"], "peak_bytes": st_["peak_bytes"],
                 "budget_bytes": st_["budget_bytes"], "warm_starts": st_["warm_starts"]}
    else:
        failed = 0
        exs = {}
        for t_arr, req in stream:
            wait = t_arr - (time.perf_counter() - t0)
            if wait > 0:
                time.sleep(wait)
            st = stacks[req.board]
            key = (req.board, id(req.route))
            if key not in exs:
                exs[key] = srv.excitation(req.board, req.route)
            ex = exs[key]
            try:
                if req.layers is not None:      # a topology of its own: a new stack on the board's spectra, cold
                    d = st.with_layers(req.layers, st.branches + list(req.extra))
                    ex = peec.RouteExcitation(d, req.route.trace, req.route.driver, req.route.receiver)
                else:
                    d = st.with_branches(req.extra) if req.extra else st
                    if req.extra:
                        ex = ex.on(d)
                for f in req.fs:
                    xb = (srv.board_solution(req.board, req.route, f)
                          if req.extra and req.layers is None else None)
                    ex.solve(f, tol=SERVE_TOL, guess=[d.map_state(xb, st)] if xb is not None else None)
                    solves += 1
                lat.append(time.perf_counter() - (t0 + t_arr))
                if d is not st:
                    d.release()
            except peec.SolverError:
                failed += 1
        extra = {}
    span = time.perf_counter() - t0
    lat_ms = np.array(lat) * 1e3
    srv.close()
    out = {"requests": len(stream), "failed": failed, "solves": solves,
           "p50_ms": float(np.percentile(lat_ms, 50)), "p99_ms": float(np.percentile(lat_ms, 99)),
           "max_ms": float(lat_ms.max()), "solves_per_s": solves / span, "span_s": span,
           "setup_s": round(setup_s, 2), "prime_s": round(prime_s, 2)}
    out.update(extra)
    return out


BACKLOG_REQUESTS = 40          # the throughput case: what-if requests on the workload board, submitted at once


def serve_backlog(n_req: int = BACKLOG_REQUESTS, seed: int = 7, prepare_thread: bool = True) -> dict:
    """The throughput comparison against the plain solver on the workload board (two_layer, 2 x 256^2 cells): a backlog of
    what-if requests (1-2 added stitching vias at random copper cells, 1-2 of the board's frequencies each, GMRES to
    the workload's 1e-10) served by the continuous-batching server -- the board resident and primed, the what-ifs
    warm-started from its solutions and batched together in its session -- with the Krylov bases in FP32 and in
    FP64, against the same solves by the plain FP64 GPU solver one at a time from zero
    (RouteExcitation.on(what-if).solve(f)). Each run is timed from the first submission to the last result; a sample
    of the served solutions is compared with FP64 solves of the same systems at 1e-13."""
    import torch

    from groundtruth_amd import boards, peec
    from groundtruth_amd.serve import Request, Server
    b = boards.two_layer()
    st = peec.PlaneStack(b.layers, b.branches, device=0)
    rng = np.random.default_rng(seed)
    specs = []
    for k in range(n_req):
        extra = []
        for e in range(int(rng.integers(1, 3))):
            i, j = (int(v) for v in rng.integers(1, st.nx - 1, 2))
            extra += boards._stitch(b.layers, i, j, f"backlog{k}_{e}")
        fs = [float(f) for f in rng.choice(b.freqs, size=int(rng.integers(1, 3)), replace=False)]
        specs.append((extra, fs))
    n_solves = sum(len(fs) for _, fs in specs)
    out = {"requests": n_req, "solves": n_solves, "tol": b.tol, "unknowns": st.nb + st.n_keep}
    served_x = {}
    for precision in ("fp32-basis", "fp64"):
        srv = Server(capacity=16, precision=precision, restart=b.restart, prepare_thread=prepare_thread)
        srv.add_board("b", st, routes=b.routes)
        srv.prime("b", b.freqs, tol=b.tol)
        srv.submit(Request("b", b.routes[0], [b.freqs[0]], extra=boards._stitch(b.layers, 3, 3, "warm-up"),
                           tol=b.tol))
        srv.drain()
        srv.reset_stats()
        reqs = [Request("b", b.routes[0], fs, extra=extra, tol=b.tol) for extra, fs in specs]
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for q in reqs:
            srv.submit(q)
        srv.drain()
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        sstats = srv.stats()
        its = [s.iterations for q in reqs if q.error is None for s in q.solutions]
        out[precision] = {"solves_per_s": n_solves / dt, "span_s": dt, "failed": sstats["failed"],
                          "iterations_mean": float(np.mean(its)) if its else None,
                          "warm_starts": sstats["warm_starts"], "p50_ms": sstats["p50_ms"],
                          "p99_ms": sstats["p99_ms"]}
        if precision == "fp32-basis":
            served_x = {(k, f): s.x for k, q in enumerate(reqs[:3]) for f, s in zip(q.fs, q.solutions)}
        srv.close()
    # the plain FP64 GPU solver, one solve at a time from zero
    ex = peec.RouteExcitation(st, b.routes[0].trace, b.routes[0].driver, b.routes[0].receiver)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    its = []
    for extra, fs in specs:
        e = ex.on(st.with_branches(extra))
        for f in fs:
            its.append(e.solve(f, tol=b.tol, restart=b.restart).iterations)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    out["plain_fp64_one_at_a_time"] = {"solves_per_s": n_solves / dt, "span_s": dt,
                                       "iterations_mean": float(np.mean(its))}
    for precision in ("fp32-basis", "fp64"):
        out[precision]["speedup_vs_plain"] = out[precision]["solves_per_s"] / (n_solves / dt)
    errs = []
    for (k, f), x in served_x.items():
        ref = ex.on(st.with_branches(specs[k][0])).solve(f, tol=1e-13, restart=b.restart)
        errs.append(float(np.linalg.norm(x - ref.x) / np.linalg.norm(ref.x)))
    out["fp32-basis"]["rel_error_sample"] = errs
    out["error_reference"] = "FP64 solves of the same systems at tol 1e-13"
    out["target_speedup"] = 10.0
    return out


def batched(fp64_blocks: Optional[list] = None) -> dict:
    """The batched multi-right-hand-side case at its own size: the four-layer board of 512 x 512 cells per layer at
    0.2 mm (seeded voids, 64 nets of 5-40 segments with 0-4 layer changes), each net at the board's 4 frequencies in
    10-500 MHz: 256 uneven right-hand sides, GMRES(50) to 1e-8 in mixed precision (FP32 cycles, FP64 refinement), in
    blocks of 16 of one frequency (similar iteration counts), each block's launches shared by its 16 solves. The plain
    FP64 solve of the blocks `fp64_blocks` (all when None) beside it, and the error of a sample against FP64 solves
    at 1e-13."""
    import torch

    from groundtruth_amd import boards, peec
    board = boards.four_layer()
    log("four-layer 512^2 board, 64 nets x 4 frequencies: set-up")
    t0 = time.perf_counter()
    st = peec.PlaneStack(board.layers, board.branches, device=0)
    st.handle
    exs = [peec.RouteExcitation(st, r.trace, r.driver, r.receiver) for r in board.routes]
    setup_s = time.perf_counter() - t0
    jobs = [(ex, float(f)) for f in board.freqs for ex in exs]       # by frequency: a block's solves alike
    ops = [st.operator(f) for _, f in jobs]
    rhs = []
    for ex, f in jobs:
        top, J = ex.rhs(f)
        rhs.append(torch.from_numpy(np.concatenate([top, np.asarray(J, float)[st.keep].astype(complex)])).to("cuda:0"))
    torch.cuda.synchronize()
    B = 16
    blocks = [list(range(k, min(k + B, len(jobs)))) for k in range(0, len(jobs), B)]
    fp64_blocks = list(range(len(blocks))) if fp64_blocks is None else fp64_blocks
    out = {"workload": "four_layer: 4 x 512x512 cells at 0.2 mm, 64 nets x 4 frequencies (10-500 MHz) = 256 "
                       "right-hand sides, GMRES(50) to 1e-8",
           "unknowns": int(st.nb + st.n_keep), "branches": int(st.nbr), "setup_s": round(setup_s, 2),
           "right_hand_sides": len(jobs), "block": B}
    xs, tb = {}, {}
    for prec, which in (("mixed", range(len(blocks))), ("fp64", fp64_blocks)):
        x = {}
        its, res_max, conv, tb[prec] = [], 0.0, 0, {}
        log(f"four-layer board: {sum(len(blocks[k]) for k in which)} right-hand sides, {prec}")
        for k in which:
            blk = blocks[k]
            for i in blk:
                x[i] = torch.zeros_like(rhs[i])
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            res = st.gmres_batch_dev([ops[i] for i in blk], [rhs[i] for i in blk], [x[i] for i in blk], board.tol,
                                     restart=board.restart, precision=prec)
            torch.cuda.synchronize()
            tb[prec][k] = time.perf_counter() - t1
            its += [int(r.iterations) for r in res]
            conv += sum(bool(r.converged) for r in res)
            res_max = max(res_max, max(float(r.residual) for r in res))
        n = sum(len(blocks[k]) for k in which)
        dt = sum(tb[prec].values())
        out[prec] = {"right_hand_sides": n, "seconds": round(dt, 3), "solves_per_s": n / dt, "converged": conv,
                     "iterations": {"min": min(its), "median": float(np.median(its)), "max": max(its)},
                     "residual_max": res_max}
        xs[prec] = x
    same = sorted(tb["fp64"])
    out["mixed"]["speedup_vs_fp64"] = sum(tb["fp64"][k] for k in same) / sum(tb["mixed"][k] for k in same)
    out["mixed"]["speedup_blocks"] = same
    sample = [i for k in same[:2] for i in (blocks[k][0], blocks[k][-1])]
    xr = [torch.zeros_like(rhs[i]) for i in sample]
    st.gmres_batch_dev([ops[i] for i in sample], [rhs[i] for i in sample], xr, 1e-13, restart=board.restart)
    torch.cuda.synchronize()
    for prec, x in xs.items():
        out[prec]["rel_error_sample"] = [float((x[i] - r).norm() / r.norm()) for i, r in zip(sample, xr)]
    out["error_reference"] = f"FP64 solves of right-hand sides {sample} at tol 1e-13"
    del rhs, xs, xr, st, exs, ops
    torch.cuda.empty_cache()
    return out


WHATIF_RATE = 4.0              # the 512^2 what-if stream: Poisson arrivals per second
WHATIF_SECONDS = 12.0


def serve_whatif_512(rate: float = WHATIF_RATE, seconds: float = WHATIF_SECONDS, seed: int = 3) -> dict:
    """The latency case (p99 under 200 ms): what-if requests on a resident two-layer board of 512 x 512 cells per layer
    (1.56 M unknowns; 0.125 mm pitch, the 256^2 board's size), each one added stitching via at a random copper cell
    and one of the board's frequencies, Poisson arrivals, FP64 GMRES to 1e-8 from the board's solution (primed before
    the stream, untimed)."""
    from groundtruth_amd import boards, peec
    from groundtruth_amd.serve import Request, Server
    b = boards.two_layer(n=512, pitch_mm=0.125)
    t0 = time.perf_counter()
    st = peec.PlaneStack(b.layers, b.branches, device=0)
    srv = Server(capacity=16)
    srv.add_board("two_layer_512", st, routes=b.routes)
    setup_s = time.perf_counter() - t0
    srv.prime("two_layer_512", b.freqs, tol=SERVE_TOL)
    rng = np.random.default_rng(seed)

    def whatif(k):
        i, j = (int(v) for v in rng.integers(1, 511, 2))
        return Request("two_layer_512", b.routes[0], [float(b.freqs[int(rng.integers(len(b.freqs)))])],
                       extra=boards._stitch(b.layers, i, j, f"whatif{k}"), tol=SERVE_TOL)
    srv.submit(whatif(-1))            # the warm start's first use, untimed
    srv.drain()
    srv.reset_stats()
    arr, t = [], 0.0
    while True:
        t += rng.exponential(1.0 / rate)
        if t >= seconds:
            break
        arr.append(t)
    reqs = [whatif(k) for k in range(len(arr))]
    t0 = time.perf_counter()
    i = 0
    while i < len(arr) or srv.busy():
        now = time.perf_counter() - t0
        while i < len(arr) and arr[i] <= now:
            srv.submit(reqs[i], now=t0 + arr[i])
            i += 1
        if srv.busy():
            srv.turn()
        elif i < len(arr):
            time.sleep(max(0.0, arr[i] - (time.perf_counter() - t0)))
    s = srv.stats()
    its = [q.solutions[0].iterations for q in srv.completed if q.error is None]
    srv.close()
    return {"requests": s["requests"], "failed": s["failed"], "p50_ms": s["p50_ms"], "p99_ms": s["p99_ms"],
            "max_ms": s["max_ms"], "solves_per_s": s["solves_per_s"], "rate_per_s": rate,
            "unknowns": st.nb + st.n_keep, "setup_s": round(setup_s, 2), "warm_starts": s["warm_starts"],
            "iterations_median": float(np.median(its)) if its else None, "tol": SERVE_TOL,
            "target_p99_ms": 200.0}


RASTER_MM = 200.0               # the rasteriser's board: 200 x 200 mm, antipads every 2 mm
RASTER_D = 0.1                 # pitch (mm)
RASTER_SUB = 4                 # sub-grid factor
RASTER_SAMPLE_MM = 10.0        # the oracle's sample of the same pattern


def raster(reps: int = 3, cpu: bool = True) -> dict:
    """The copper rasteriser (build_grid) on a 200 x 200 mm ground pour with 9801 antipads at 0.1 mm and a sub-grid
    of 4 (2001^2 nodes, 8004^2 fine samples): seconds per grid (best of `reps`, from the polygons to every Grid field),
    fine samples per second, and the oracle's rate on a 10 x 10 mm cut of the same pattern (its loops take minutes
    at 20 mm)."""
    from groundtruth_amd import boards
    from groundtruth_amd import raster as RS
    geom = boards.antipad_pour(RASTER_MM)
    bbox = (0.0, 0.0, RASTER_MM, RASTER_MM)
    RS.build_grid(boards.antipad_pour(20.0), (0.0, 0.0, 20.0, 20.0), RASTER_D, sub=RASTER_SUB)      # warm-up
    times = []
    for _ in range(reps):
        t0 = time.perf_counter()
        g = RS.build_grid(geom, bbox, RASTER_D, sub=RASTER_SUB)
        times.append(time.perf_counter() - t0)
    fine = (g.nx * RASTER_SUB) * (g.ny * RASTER_SUB)
    out = {"workload": f"{RASTER_MM:.0f} x {RASTER_MM:.0f} mm ground pour, {len(geom['GND'].rings) - 1} antipads "
                       f"(64-gons), d {RASTER_D} mm, sub-grid {RASTER_SUB}: {g.nx}^2 nodes, {g.nx * RASTER_SUB}^2 "
                       "fine samples",
           "seconds_per_grid": min(times), "fine_samples_per_s": fine / min(times), "nodes": int(g.node.sum()),
           "strips": int(g.xb.sum() + g.yb.sum())}
    if cpu:
        from oracle import raster as OR
        small = boards.antipad_pour(RASTER_SAMPLE_MM)
        t0 = time.perf_counter()
        o = OR.build_grid(small, (0.0, 0.0, RASTER_SAMPLE_MM, RASTER_SAMPLE_MM), RASTER_D, sub=RASTER_SUB)
        dt = time.perf_counter() - t0
        n_small = (o.nx * RASTER_SUB) * (o.ny * RASTER_SUB)
        out["cpu_oracle"] = {"fine_samples_per_s": n_small / dt, "seconds": dt, "cores": 1, "kind": "port",
                             "sample": f"the oracle's build_grid of a {RASTER_SAMPLE_MM:.0f} x {RASTER_SAMPLE_MM:.0f} "
                                       f"mm cut of the same pattern ({o.nx}^2 nodes, {n_small} fine samples)"}
        out["speedup_vs_cpu_oracle"] = out["fine_samples_per_s"] / out["cpu_oracle"]["fine_samples_per_s"]
    return out


def host_info() -> dict:
    """The host the CPU numbers ran on: cores, the affinity the process got, the CPU model, numpy and scipy."""
    import scipy
    model = platform.processor() or platform.machine()
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"cpu_count": os.cpu_count(), "affinity": len(os.sched_getaffinity(0)), "cpu_model": model,
            "numpy": np.__version__, "scipy": scipy.__version__}


def cpu_baseline(board, route, f: float) -> dict:
    """The oracle (the numpy / scipy restatement of the reference's path, with the reference's settings: scipy FFTs on
    every core, the numpy multiply-add, a complex128 basis with zgemv Gram-Schmidt, SuperLU on one thread) solving one
    frequency of the workload on this host: after a warm-up solve, the best of 3, each its own SuperLU factorisation
    (timed apart) plus its GMRES. The GPU's value excludes the board's set-up (a resident board); the host's includes
    the factorisation, which the reference redoes per frequency."""
    from oracle import peec as O
    st = O.PlaneStack(board.layers, board.branches)
    ex = O.RouteExcitation(st, route.trace, route.driver, route.receiver)
    ex.solve(f, tol=board.tol, restart=board.restart)          # warm-up (and the kernels, cached)
    best, fac, its = None, None, 0
    for _ in range(3):
        st._ops.clear()
        t0 = time.perf_counter()
        st.operator(f)
        t1 = time.perf_counter()
        sol = ex.solve(f, tol=board.tol, restart=board.restart)
        t2 = time.perf_counter()
        if best is None or t2 - t0 < best:
            best, fac, its = t2 - t0, t1 - t0, sol.iterations
    info = host_info()
    return {"value": 1.0 / best, "unit": "solves/s", "cores": info["affinity"], "kind": "port",
            "sample": f"one of the workload's 10 solves (f = {f:.4g} Hz, GMRES({board.restart}) to {board.tol:g}, "
                      f"{its} iterations), best of 3 after a warm-up: {best:.2f} s, of which the SuperLU "
                      f"factorisation {fac:.2f} s; scipy FFTs on all {info['affinity']} cores, SuperLU on 1",
            "host": info}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="two_layer")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--lp-only", action="store_true")
    ap.add_argument("--lp-reps", type=int, default=20)
    ap.add_argument("--lp-batch", type=int, default=8, help="--lp-only: right-hand sides per L_p launch")
    ap.add_argument("--serve-rate", type=float, default=SERVE_RATE, help="served stream: requests per second")
    ap.add_argument("--serve-seconds", type=float, default=SERVE_SECONDS, help="served stream: arrival window")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--serve-only", action="store_true")
    ap.add_argument("--large", action="store_true", help="the six-layer 2048^2 board alone (a few minutes)")
    ap.add_argument("--batched", action="store_true", help="the four-layer 512^2 board's 256 right-hand sides alone")
    ap.add_argument("--no-batched", action="store_true", help="skip the four-layer 256-right-hand-side case")
    ap.add_argument("--no-large", action="store_true", help="skip the six-layer board (about three minutes)")
    ap.add_argument("--raster", action="store_true", help="the copper rasteriser's board alone")
    ap.add_argument("--no-raster", action="store_true", help="skip the copper rasteriser's board")
    ap.add_argument("--steps-only", choices=["fp64", "fp32-basis", "mixed"],
                    help="only the workload's batched steps in this precision, then 64 MiB copies (for profiles)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU: --gpus {args.gpus} is not supported (independent boards and frequencies "
              "would shard across GPUs without collectives, one process per GPU)",
              file=sys.stderr)
        sys.exit(2)
    if args.lp_only:
        lp_only(args)
        return
    if args.steps_only:
        steps_only(args)
        return
    if args.large:
        print(json.dumps({"large": large(cpu=not args.no_cpu_baseline)}))
        return
    if args.batched:
        print(json.dumps({"batched": batched()}))
        return
    if args.raster:
        print(json.dumps({"raster": raster(cpu=not args.no_cpu_baseline)}))
        return
    if args.serve_only:
        print(json.dumps({"served_backlog": serve_backlog(),
                          "served": serve_stream(args.serve_rate, args.serve_seconds),
                          "served_one_at_a_time": serve_stream(args.serve_rate, args.serve_seconds, batched=False)}))
        return
    import torch

    from groundtruth_amd import boards, peec
    board = boards.WORKLOADS[args.workload]()
    route = board.routes[0]
    log(f"workload {args.workload}: set-up")
    t_setup = time.perf_counter()
    st = peec.PlaneStack(board.layers, board.branches, device=0)
    ex = peec.RouteExcitation(st, route.trace, route.driver, route.receiver)
    st.handle
    ops = [st.operator(f) for f in board.freqs]
    setup_s = time.perf_counter() - t_setup

    # right-hand sides resident in HBM before the timed region (one per frequency)
    dev = "cuda:0"
    rhs = []
    for f in board.freqs:
        rhs_top, J = ex.rhs(f)
        rhs.append(np.concatenate([rhs_top, J[st.keep].astype(complex)]))
    rhs_dev = [torch.from_numpy(r).to(dev) for r in rhs]
    x_dev = [torch.zeros_like(r) for r in rhs_dev]
    torch.cuda.synchronize()

    def step():
        res = st.gmres_batch_dev(ops, rhs_dev, x_dev, board.tol, restart=board.restart)
        for op, r in zip(ops, res):
            if not r.converged:
                raise RuntimeError(f"GMRES did not converge at {op.f:.4g} Hz")
        return [int(r.iterations) for r in res]

    import ctypes as C

    from groundtruth_amd import _lib
    L = _lib.lib()
    log(f"{args.warmup} warm-up and {args.steps} timed steps (plain FP64)")
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    # every L_p launch of the timed steps bracketed by HIP events on the library's stream
    _lib.check(L.gt_ctx_lp_timing(st.ctx.handle, 1))
    t0 = time.perf_counter()
    its = []
    for _ in range(args.steps):
        its = step()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    n_lp, n_rhs, lp_ms = C.c_int64(), C.c_int64(), C.c_double()
    _lib.check(L.gt_ctx_lp_time(st.ctx.handle, C.byref(n_lp), C.byref(n_rhs), C.byref(lp_ms)))
    _lib.check(L.gt_ctx_lp_timing(st.ctx.handle, 0))

    # the same batch in mixed precision: FP32 FFT / Krylov inner solves, FP64 Arnoldi, Schur and refinement to 1e-8
    x_mix = [torch.zeros_like(r) for r in rhs_dev]

    def step_mixed():
        res = st.gmres_batch_dev(ops, rhs_dev, x_mix, board.tol, restart=board.restart, precision="mixed")
        for op, r in zip(ops, res):
            if not r.converged:
                raise RuntimeError(f"mixed-precision GMRES did not reach {board.tol:g} at {op.f:.4g} Hz")
        return [int(r.iterations) for r in res], [float(r.residual) for r in res]

    log("the same steps in mixed precision")
    for _ in range(args.warmup):
        step_mixed()
    torch.cuda.synchronize()
    t0m = time.perf_counter()
    for _ in range(args.steps):
        its_mix, res_mix = step_mixed()
    torch.cuda.synchronize()
    dt_mix = time.perf_counter() - t0m
    # the same batch with the Krylov basis stored in FP32 (one FP64 GMRES, the plain run's tolerance)
    x_f32b = [torch.zeros_like(r) for r in rhs_dev]

    def step_f32b():
        res = st.gmres_batch_dev(ops, rhs_dev, x_f32b, board.tol, restart=board.restart, precision="fp32-basis")
        for op, r in zip(ops, res):
            if not r.converged:
                raise RuntimeError(f"FP32-basis GMRES did not converge at {op.f:.4g} Hz")
        return [int(r.iterations) for r in res], [float(r.residual) for r in res]

    log("the same steps with FP32 Krylov bases")
    for _ in range(args.warmup):
        step_f32b()
    torch.cuda.synchronize()
    t0b = time.perf_counter()
    for _ in range(args.steps):
        its_f32b, res_f32b = step_f32b()
    torch.cuda.synchronize()
    dt_f32b = time.perf_counter() - t0b
    # an FP64 reference solution at tol 1e-13 (untimed) to state every run's errors against
    x_ref = [torch.zeros_like(r) for r in rhs_dev]
    st.gmres_batch_dev(ops, rhs_dev, x_ref, 1e-13, restart=board.restart)
    torch.cuda.synchronize()
    err_mix = [float((xm - xr).norm() / xr.norm()) for xm, xr in zip(x_mix, x_ref)]
    err_fp64 = [float((xp - xr).norm() / xr.norm()) for xp, xr in zip(x_dev, x_ref)]
    err_f32b = [float((xb - xr).norm() / xr.norm()) for xb, xr in zip(x_f32b, x_ref)]
    lp_us = lp_ms.value / max(1, n_lp.value) * 1e3           # average L_p launch in the timed steps
    lp_rhs = n_rhs.value / max(1, n_lp.value)                # right-hand sides per launch
    solves = args.steps * len(board.freqs)
    # the drop-in path (host vectors in and out, Solution unpacked) for the PCIe-inclusive rate
    t1 = time.perf_counter()
    for f in board.freqs:
        ex.solve(f, tol=board.tol, restart=board.restart)
    dt_host = time.perf_counter() - t1
    t_lp = time_lp(st)
    B = lp_algorithmic_bytes(st)
    B_launch = lp_algorithmic_bytes(st, 1) + (lp_rhs - 1) * lp_bytes(st)[0]    # the kernel term once per launch
    achieved = B_launch / (lp_us * 1e-6) / 1e9
    traffic = measured_traffic(args.workload)
    out = {
        "metric": METRIC,
        "value": solves / dt,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c128",
        "data": "synthetic (seeded board generator, default_rng(1002))",
        "config": {"workload": f"{args.workload}: {len(board.layers)} x {st.nx}x{st.ny} cells, "
                               f"{len(board.freqs)} frequencies 1 MHz-1 GHz, GMRES({board.restart}) to "
                               f"{board.tol:g}, plain FP64",
                   "unknowns": st.nb + st.n_keep, "strips": st.ns, "vias": st.nbr,
                   "iterations": its, "setup_s": round(setup_s, 3), "fft": [list(st.Px), list(st.Py)],
                   "lp_plan": getattr(st, "lp_plan", None)},
        "matvec_GBps": achieved,
        "value_pcie_inclusive": len(board.freqs) / dt_host,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": round(traffic[0] * lp_rhs / traffic[1]) if traffic else None,
                     "traffic_measured": {"bytes": traffic[0], "rhs_per_launch": traffic[1]} if traffic else None,
                     "kernel": "L_p product (lp_rows_fwd + lp_cols_mac + lp_rows_inv), as launched by the timed "
                               "batched solves",
                     "launches": n_lp.value, "rhs_per_launch": lp_rhs, "us_per_launch": lp_us,
                     "bytes_per_launch": round(B_launch),
                     "bytes_formula": "rhs_per_launch x (2c n_o + 4c S P0 P1) + r S^2 (P0/2+1)(P1/2+1) summed over "
                                      "the orientations, c = 16, r = 8: the kernel quadrants once per launch",
                     "single_rhs": {"us_per_launch": t_lp * 1e6, "achieved": B / t_lp / 1e9,
                                    "bytes_per_launch": B}},
    }
    out["mixed"] = {"value": solves / dt_mix, "unit": "solves/s", "ms_per_step": dt_mix / args.steps * 1e3,
                    "speedup_vs_fp64": (solves / dt_mix) / (solves / dt), "tol": board.tol,
                    "iterations": its_mix, "residual": res_mix,
                    "rel_error": err_mix, "rel_error_fp64_run": err_fp64,
                    "error_reference": "the plain FP64 solve of the same system at tol 1e-13",
                    "precision": "the reference's restarted GMRES with FP32 cycles (L_p FFT convolution, incidence "
                                 "rows, Krylov vectors and basis) and every restart an FP64 refinement step (r = b - K x,"
                                 " x += M^-1 y); FP64 Arnoldi dot products, Givens and Schur preconditioner"}
    out["fp32_basis"] = {"value": solves / dt_f32b, "unit": "solves/s", "ms_per_step": dt_f32b / args.steps * 1e3,
                         "speedup_vs_fp64": (solves / dt_f32b) / (solves / dt), "tol": board.tol,
                         "iterations": its_f32b, "residual": res_f32b, "rel_error": err_f32b,
                         "rel_error_fp64_run": err_fp64,
                         "error_reference": "the plain FP64 solve of the same system at tol 1e-13",
                         "precision": "Krylov basis stored in FP32; operator products, Arnoldi sums, Givens, Schur "
                                      "preconditioner, iterate and true-residual checks in FP64"}
    if not args.no_serve:
        log("served what-if stream, batched and one at a time; 512^2 what-ifs; the backlog")
        served = serve_stream(args.serve_rate, args.serve_seconds)
        plain = serve_stream(args.serve_rate, args.serve_seconds, batched=False)
        out["served"] = dict(served, rate_per_s=args.serve_rate, tol=SERVE_TOL,
                             p99_vs_one_at_a_time=plain["p99_ms"] / served["p99_ms"],
                             one_at_a_time=plain,
                             stream="Poisson arrivals; boards single_plane 128^2, four_layer 4x128^2 (4 nets), "
                                    "two_layer 2x256^2 drawn uniformly; 1 net x 1-4 frequencies per request; half "
                                    "the requests are what-ifs: on multi-layer boards half of those add 1-2 "
                                    "stitching vias (a what-if stack on the resident factor, warm from the board "
                                    "solution at f), the rest move the board's slot (a new topology on the board's "
                                    "kernel spectra, cold); plain FP64 GMRES to 1e-8; one_at_a_time: the same "
                                    "stream solved in arrival order one solve at a time (the boards primed before "
                                    "the stream, untimed)")
        out["whatif_512"] = serve_whatif_512()
        out["served_backlog"] = serve_backlog()
    if not args.no_cpu_baseline:
        log("the oracle's solve on the host (best of 3)")
        out["cpu_baseline"] = cpu_baseline(board, route, board.freqs[0])
    del st, ex, ops, rhs_dev, x_dev, x_mix, x_f32b, x_ref
    torch.cuda.empty_cache()
    if not args.no_batched:
        try:
            out["batched"] = batched(fp64_blocks=[0, 9])     # FP64 beside mixed on two blocks (f0, f2)
        except Exception as e:      # noqa: BLE001 - the headline line above stands on its own
            out["batched"] = {"error": f"{type(e).__name__}: {e}"}
    if not args.no_large:
        try:
            out["large"] = large(cpu=not args.no_cpu_baseline)
        except Exception as e:      # noqa: BLE001 - the headline line above stands on its own
            out["large"] = {"error": f"{type(e).__name__}: {e}"}
    if not args.no_raster:
        try:
            log("the copper rasteriser: 200 x 200 mm board")
            out["raster"] = raster(cpu=not args.no_cpu_baseline)
        except Exception as e:      # noqa: BLE001 - the headline line above stands on its own
            out["raster"] = {"error": f"{type(e).__name__}: {e}"}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
