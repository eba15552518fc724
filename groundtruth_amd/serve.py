"""Serving what-if solves on one device: boards stay resident, requests are batched continuously, admission is by
device-memory footprint, and every request's latency is recorded.

The reference solves a board's (model, frequency) pairs as independent tasks: `run_parallel` (simulate.py:3964-4013)
runs them on a thread pool, and `Admission` (simulate.py:3930-3961) lets a task start only while the memory it
plans to use fits the budget. Here the tasks are GMRES solves on the device, and they share launches instead of
threads: a `Session` (gt_session) runs up to 16 solves as one batch, a new solve joins it between two Arnoldi
steps, and a finished one leaves without waiting for the rest. A board's session also takes the solves of its
what-if stacks (peec.py:1871-1909 `with_branches`: the board plus a few candidate vias), each over its own unknowns,
so what-ifs of different sizes share launches with each other and with the board's own solves. `Server` keeps one
session per board (and one for a what-if that cannot join its board's, e.g. a via joining two islands), turns them
round-robin, admits sessions and what-if stacks by their exact device bytes (gt_session_open_bytes, the footprint
gpu.py:593-605 `device_bytes` estimates), and reports p50/p99 latency and solves per second. The Krylov bases are
FP64, or FP32 with everything else FP64 (precision "fp32-basis").

A served solve is the same solve as `RouteExcitation.solve(f)`: the same right-hand side, norm and GMRES state
machine, so its solution is bit-identical alone and in any mix (the library's batches keep each right-hand side's
arithmetic to itself).
"""
from __future__ import annotations

import ctypes as C
import math
import time
from collections import OrderedDict, deque
from concurrent.futures import FIRST_COMPLETED, Future, ThreadPoolExecutor, wait
from dataclasses import dataclass, field
from typing import Optional

import numpy as np

from . import _lib
from .peec import BATCH, GMRES_CHECK_MISSES, GMRES_EXTRA_CYCLES, PlaneStack, RouteExcitation
from .problem import SolverError


_BASIS = {"fp64": 0, "fp32-basis": 1}


def _basis(precision: str) -> int:
    if precision not in _BASIS:
        raise ValueError("precision is 'fp64' or 'fp32-basis'")
    return _BASIS[precision]


def session_plan_bytes(stack: PlaneStack, capacity: int = BATCH, restart: Optional[int] = None,
                       precision: str = "fp64", max_extra: int = 0) -> int:
    """Device bytes a session of `capacity` slots on `stack` adds: its GMRES workspace and the growth of the stack's
    product scratch to that batch (for what-ifs of up to max_extra added branches)."""
    b = int(_lib.lib().gt_session_open_bytes(stack.handle, int(capacity), int(restart or 0), _basis(precision),
                                             int(max_extra)))
    if b < 0:
        raise ValueError("1 to 16 slots and a non-negative restart and max_extra")
    return b


class Session:
    """Continuous batching of solves (gt_session) on `stack` and on what-if stacks derived from it with up to
    max_extra added branches: `submit` puts a solve into a free slot, `step` advances the running ones, `poll`
    collects a finished one. precision "fp32-basis" stores the Krylov bases in FP32."""

    def __init__(self, stack: PlaneStack, capacity: int = BATCH, restart: Optional[int] = None,
                 precision: str = "fp64", max_extra: int = 0):
        self.stack = stack
        self.capacity = int(capacity)
        self.max_extra = int(max_extra)
        h = C.c_void_p()
        _lib.check(_lib.lib().gt_session_open(stack.handle, self.capacity, int(restart or 0), _basis(precision),
                                              self.max_extra, C.byref(h)), "gt_session_open")
        self._h = h

    def submit(self, op, b_dev, x_dev, tol: float, bnorm: float = 0.0, max_cycles: int = 8,
               extra_cycles: int = GMRES_EXTRA_CYCLES, max_its: int = 0) -> int:
        """The slot the solve x = M^-1 y, (K M^-1) y = b took, or -1 when every slot is taken. op: an operator of the
        session's stack or of a what-if stack derived from it. b_dev, x_dev: torch complex128 device vectors of the
        op's stack's nb + n_keep values, b's work finished; b stays unchanged until the poll."""
        p = _lib.GmresParams(float(tol), float(bnorm), 0, int(max_cycles), int(extra_cycles), int(max_its),
                             GMRES_CHECK_MISSES)
        slot = C.c_int32()
        _lib.check(_lib.lib().gt_session_submit(self._h, op.handle, C.c_void_p(b_dev.data_ptr()),
                                                C.c_void_p(x_dev.data_ptr()), C.byref(p), C.byref(slot)),
                   "gt_session_submit")
        return int(slot.value)

    def step(self, max_steps: int = 1):
        """Set up the new solves, advance the running ones by up to max_steps Arnoldi steps, finish the stopped
        ones. (running, ready)."""
        run, rd = C.c_int32(), C.c_int32()
        _lib.check(_lib.lib().gt_session_step(self._h, int(max_steps), C.byref(run), C.byref(rd)),
                   "gt_session_step")
        return int(run.value), int(rd.value)

    def poll(self, slot: int):
        """The finished solve's GmresResult (its slot then free), or None while it runs."""
        res, done = _lib.GmresResult(), C.c_int32()
        _lib.check(_lib.lib().gt_session_poll(self._h, int(slot), C.byref(res), C.byref(done)), "gt_session_poll")
        return res if done.value else None

    def device_bytes(self) -> int:
        return int(_lib.lib().gt_session_device_bytes(self._h))

    def close(self):
        if self._h:
            _lib.lib().gt_session_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            if getattr(self, "_h", None) and _lib._LIB is not None:
                _lib._LIB.gt_session_destroy(self._h)
        except Exception:       # noqa: BLE001 - interpreter shutdown
            pass


class Admission:
    """Device-memory admission after simulate.py:3930-3961 `Admission`: a claim is granted while the bytes held
    plus its own fit the budget, and always when nothing is held (a claim is never left waiting for memory that
    nothing will free). Non-blocking: the serving loop asks again on its next turn. `reserve` takes bytes that stay
    held for good (the resident boards) and count against the budget without counting as a claim."""

    def __init__(self, budget: float):
        self.budget = float(budget)
        self.used = 0.0
        self.used_peak = 0.0
        self.held = 0
        self.peak = 0

    def reserve(self, nbytes: float) -> bool:
        if self.used + nbytes > self.budget:
            return False
        self.used += nbytes
        self.used_peak = max(self.used_peak, self.used)
        return True

    def try_hold(self, need: float) -> bool:
        if self.held and self.used + need > self.budget:
            return False
        self.used += need
        self.used_peak = max(self.used_peak, self.used)
        self.held += 1
        self.peak = max(self.peak, self.held)
        return True

    def release(self, need: float):
        self.used -= need
        self.held -= 1

    def try_grow(self, need: float) -> bool:
        """More bytes for a claim already held (a what-if stack running in its board's session): granted while
        they fit, and always when that claim is the only one."""
        if self.held > 1 and self.used + need > self.budget:
            return False
        self.used += need
        self.used_peak = max(self.used_peak, self.used)
        return True

    def shrink(self, need: float):
        self.used -= need

    def true_up(self, planned: float, actual: float):
        """A claim admitted by its plan, now of its exact bytes (a what-if stack once made)."""
        self.used += actual - planned
        self.used_peak = max(self.used_peak, self.used)


@dataclass
class Request:
    """One what-if request: a net (route) of a resident board, the frequencies to solve it at and, for a what-if,
    the branches (candidate vias) added to the board. `solutions` (one Solution per frequency, in order) and
    `latency` (s, from submission to the last solve's end) are filled when it completes."""
    board: str
    route: object                    # .trace, .driver, .receiver (a boards.Route), as RouteExcitation takes them
    fs: list
    extra: list = field(default_factory=list)
    layers: Optional[list] = None    # the board's layers with new copper (a moved slot): a topology of its own
    tol: float = 1e-8
    max_cycles: int = 8
    warm: bool = True                # a what-if starts from the board's solution at f when the server holds one
    tag: object = None
    # filled by the server
    arrival: float = 0.0
    done_at: float = 0.0
    solutions: list = field(default_factory=list)
    error: Optional[str] = None
    _hold: float = 0.0               # device bytes its admission holds: planned before its stack is made, then the
                                     # what-if stack's own while it runs in its board's session
    _stack: object = None            # its what-if stack, planned (not yet on the device) at admission
    _grow: bool = False              # _hold is a growth of its board's session claim (else a claim of its own)
    _attached: bool = False          # admitted into its board's session, its solves not yet queued there

    @property
    def latency(self) -> float:
        return self.done_at - self.arrival


class _Solve:
    """One (request, frequency) solve in flight on `stack` (the board or a what-if stack): GMRES on b = rhs - K x0
    from the warm start x0 (None: cold)."""
    __slots__ = ("req", "k", "f", "ex", "op", "b", "x", "bnorm", "x0", "slot", "stack")

    def __init__(self, req, k, f, ex, op, b, x, bnorm, x0, stack):
        self.req, self.k, self.f, self.ex, self.op, self.b, self.x, self.bnorm = req, k, f, ex, op, b, x, bnorm
        self.x0 = x0
        self.stack = stack
        self.slot = -1


class _Served:
    """A stack the server runs solves on, its session and the solves waiting for or holding its slots."""

    def __init__(self, stack: PlaneStack, need: float):
        self.stack = stack
        self.need = need
        self.session: Optional[Session] = None
        self.whatif = False
        self.attached = 0                      # what-ifs admitted into this session, not yet queued in it
        self.waiting: deque = deque()
        self.slots: dict = {}
        self.last_used = 0.0


class Server:
    """The serving loop of one device. `add_board` makes a board resident (its spectra and Schur factor stay on the
    device); `submit` queues a request; `turn` admits what fits, steps every session once and collects the finished
    solves; `drain` turns until everything submitted has completed.

    Per turn, each session with work gets `steps_per_turn` Arnoldi steps; a new solve enters its board's session as
    soon as a slot is free, so solves of different requests, frequencies and what-ifs share launches while they
    overlap. A what-if request gets the stack derived from its board with the added branches; with up to
    `max_extra` of them it joins the board's session, else (or when the added vias join islands, so the stack has a
    factor of its own) it gets a session of its own. Sessions and what-if stacks are admitted by their planned
    device bytes against `budget` (default: 90 % of the device memory free when the server starts); idle sessions
    are closed, oldest first, when a new one needs their room. precision: the sessions' Krylov bases, "fp64" or
    "fp32-basis"."""

    def __init__(self, device: int = 0, budget: Optional[float] = None, capacity: int = BATCH,
                 steps_per_turn: int = 1, restart: Optional[int] = None, precision: str = "fp64",
                 max_extra: int = 8, prepare_thread: bool = True):
        import torch
        _basis(precision)
        self.device = device
        self.dev = f"cuda:{device}"
        self.capacity = int(capacity)
        self.steps_per_turn = int(steps_per_turn)
        self.restart = restart
        self.precision = precision
        self.max_extra = int(max_extra)
        if budget is None:
            free, _ = torch.cuda.mem_get_info(device)
            budget = 0.9 * free
        self.admission = Admission(budget)
        self.boards: dict = {}
        self._excitations: dict = {}
        self._served: OrderedDict = OrderedDict()   # id(stack) -> _Served
        self._pending: deque = deque()              # requests not yet split into solves (their stack not admitted)
        self._open: dict = {}                       # id(request) -> solves left
        self.completed: list = []
        self.solves_done = 0
        self.warm_starts = 0
        self.t0: Optional[float] = None
        self._board_x: dict = {}                    # (board, route id, f key) -> the board's latest x (device)
        # a request's set-up (its what-if stack, right-hand sides, operators, warm starts) runs on a worker thread
        # while the sessions step: the library's calls release the interpreter, and its own lock orders them
        self._worker = ThreadPoolExecutor(max_workers=1, thread_name_prefix="gt-serve-prep") if prepare_thread else None
        self._prepping: deque = deque()             # (request, future of its prepared solves)
        self._prepared: deque = deque()             # (request, stack, solves, warm starts) waiting for admission

    # ---------------------------------------------------------------- boards and requests
    def add_board(self, name: str, stack: PlaneStack, routes=()):
        """Make `stack` resident and servable as `name` (its device bytes are held for the server's life), with the
        excitations of `routes` (their coupling to the board's strips) computed ahead of the requests."""
        stack.handle
        if not self.admission.reserve(stack.device_bytes()):
            raise MemoryError(f"board {name} does not fit the device budget")
        self.boards[name] = stack
        for r in routes:
            self.excitation(name, r)

    def prime(self, name: str, freqs, routes=None, tol: float = 1e-8):
        """Solve the board's nets at `freqs` (all its registered routes by default) and keep the solutions: the
        analysis of the board a what-if edits, whose solutions the what-ifs then start from."""
        rts = routes if routes is not None else [r for (b, _), (r, _) in self._excitations.items() if b == name]
        for r in rts:
            self.submit(Request(name, r, [float(f) for f in freqs], tol=tol))
        self.drain()

    def board_solution(self, name: str, route, f: float) -> Optional[np.ndarray]:
        """The board's latest solution x of the route at f (None when it has not been solved here)."""
        x = self._board_x.get((name, id(route), PlaneStack.freq_key(f)))
        return None if x is None else x.cpu().numpy()

    def reset_stats(self):
        """Forget the completed requests and the peaks (after a warm-up)."""
        self.completed = []
        self.solves_done = 0
        self.warm_starts = 0
        self.t0 = None
        self.admission.used_peak = self.admission.used
        self.admission.peak = self.admission.held

    def submit(self, req: Request, now: Optional[float] = None):
        if req.board not in self.boards:
            raise ValueError(f"unknown board {req.board!r}")
        req.arrival = time.perf_counter() if now is None else now
        if self.t0 is None:
            self.t0 = req.arrival
        self._pending.append(req)

    def busy(self) -> bool:
        return bool(self._pending or self._prepping or self._prepared or self._open)

    # ---------------------------------------------------------------- the loop
    def _stack_for(self, req: Request) -> PlaneStack:
        """The request's stack, not yet on the device: its board; a what-if derived from it (added branches); or,
        for new copper (a moved slot), a stack of its own on the board's grid sharing its kernel spectra."""
        base = self.boards[req.board]
        if req.layers is not None:
            return base.with_layers(req.layers, base.branches + list(req.extra))
        return base.with_branches(req.extra) if req.extra else base

    def _plan(self, req: Request) -> tuple:
        """The device bytes to admit a request by before anything of it is made, as (a claim of its own, a growth
        of its board's session claim): nothing for the board itself; for a what-if derived from the board that joins
        the board's session, the exact set-up peak of its stack (gt_stack_derive_bytes) as a growth; one with a
        session of its own, that peak and the session as a claim; for a stack of its own (new copper, or vias
        joining islands) the board's resident bytes and a session of the request's size, its planned peer."""
        board = self.boards[req.board]
        session = lambda: session_plan_bytes(board, min(self.capacity, max(1, len(req.fs))), self.restart,  # noqa: E731
                                             self.precision, 0)
        if req.layers is None and not req.extra:
            req._stack = board
            return 0.0, 0.0
        if req.layers is None:
            st = board.with_branches(req.extra)
            req._stack = st
            if st is board:
                return 0.0, 0.0
            if st._base is board:
                _, peak = st.plan_bytes()
                if self._host(req, st) is board:
                    return 0.0, float(peak)
                return float(peak) + session(), 0.0
        else:
            req._stack = None                   # new copper: the stack is built on the worker thread
        return float(board.device_bytes() + session()), 0.0

    def _admit_request(self, req: Request) -> bool:
        """Admit a request by its plan (nothing of it is made before): a claim of its own, or a growth of its
        board's session claim (the session opened first, and kept open while the request is attached)."""
        claim, grow = self._plan(req)
        if claim > 0:
            if not self._hold(claim):
                return False
            req._hold, req._grow = claim, False
        if grow > 0:
            sv = self._admit_served(self.boards[req.board], len(req.fs))
            if sv is None or not self.admission.try_grow(grow):
                return False
            sv.attached += 1
            req._hold, req._grow, req._attached = grow, True, True
        return True

    def excitation(self, board: str, route) -> RouteExcitation:
        """The route's excitation on the board, its strip coupling computed once per board and route."""
        key = (board, id(route))
        hit = self._excitations.get(key)
        if hit is None:
            hit = self._excitations[key] = (route, RouteExcitation(self.boards[board], route.trace, route.driver,
                                                                    route.receiver))
        return hit[1]

    def _excitation(self, req: Request, stack: PlaneStack) -> RouteExcitation:
        """The request's excitation on its stack: the board's; moved onto a what-if derived from it by
        RouteExcitation.on (the strips' coupling reused); computed afresh on a stack of other copper."""
        board = self.boards[req.board]
        ex = self.excitation(req.board, req.route)
        if stack is board:
            return ex
        if getattr(stack, "_base", None) is board:
            return ex.on(stack)
        r = req.route
        return RouteExcitation(stack, r.trace, r.driver, r.receiver)

    def _host(self, req: Request, stack: PlaneStack) -> PlaneStack:
        """The stack whose session runs the request's solves: its board's, for the board itself and for a what-if
        derived from it with up to max_extra added branches; else the what-if stack's own."""
        board = self.boards[req.board]
        if stack is board:
            return board
        if getattr(stack, "_base", None) is board and stack.nbr - board.nbr <= self.max_extra:
            return board
        return stack

    def _hold(self, need: float) -> bool:
        """Admission of `need` bytes, closing idle sessions (oldest first) to make room."""
        while not self.admission.try_hold(need):
            idle = [k for k, s in self._served.items() if not s.slots and not s.waiting and not s.attached]
            if not idle:
                return False
            self._close(idle[0])
        return True

    def _admit_served(self, stack: PlaneStack, n_solves: int, claim: float = 0.0) -> Optional[_Served]:
        """The stack's session, made when its bytes are admitted (None: not now). A board's session has every
        slot and room for what-ifs; a what-if stack's own, made for one request, as many slots as that request's
        solves: its bytes are the request's claim (admitted as planned before the stack was made), trued up to the
        session's and the stack's exact bytes."""
        sv = self._served.get(id(stack))
        if sv is not None:
            return sv
        whatif = stack is not None and all(stack is not b for b in self.boards.values())
        cap = min(self.capacity, max(1, n_solves)) if whatif else self.capacity
        extra = 0 if whatif else self.max_extra
        need = session_plan_bytes(stack, cap, self.restart, self.precision, extra)
        if whatif:
            need += stack.device_bytes()           # the what-if stack's own layout, incidence, factor, columns
            self.admission.true_up(claim, need)    # the request's claim becomes the session's
        elif not self._hold(need):
            return None
        sv = _Served(stack, need)
        sv.whatif = whatif
        sv.session = Session(stack, cap, self.restart, self.precision, extra)
        self._served[id(stack)] = sv
        return sv

    def _close(self, key):
        sv = self._served.pop(key)
        if sv.session is not None:
            sv.session.close()
        if sv.whatif:
            sv.stack.release()                 # a what-if stack of its own: its device copy goes with its session
        self.admission.release(sv.need)

    def _prepare(self, req: Request):
        """The request's solves (run on the worker thread, after its admission): its stack (a what-if derived from
        the board, or a topology of its own), the right-hand sides of RouteExcitation.solve, the operators, and for
        a what-if derived from the board the warm start of PlaneStack.solve from the board's solution at f
        (peec.py:2568-2576), its products in the board's scratch. (stack, solves, warm starts taken)."""
        import torch
        board = self.boards[req.board]
        stack = req._stack if req._stack is not None else self._stack_for(req)
        req._stack = stack
        stack.handle
        whatif = stack is not board
        ex = self._excitation(req, stack)
        solves, warm = [], 0
        for k, f in enumerate(req.fs):
            rhs_top, J = ex.rhs(f)
            rhs = np.concatenate([np.asarray(rhs_top, complex), np.asarray(J, float)[stack.keep].astype(complex)])
            op = stack.operator(f)
            b = torch.from_numpy(rhs).to(self.dev)
            x0 = None
            if whatif and req.warm and getattr(stack, "_base", None) is board:
                xb = self._board_x.get((req.board, id(req.route), PlaneStack.freq_key(f)))
                g = stack.map_state_dev(xb, board)
                if g is not None:
                    torch.cuda.synchronize(self.device)
                    x0, b = stack.warm_start_dev(op, b, [g], host=board)
                    warm += x0 is not None
            x = torch.empty_like(b)
            solves.append(_Solve(req, k, f, ex, op, b, x, float(np.linalg.norm(rhs)), x0, stack))
        torch.cuda.synchronize(self.device)         # the right-hand sides are on the device
        return stack, solves, warm

    def _release_request(self, req: Request):
        if req._attached:                      # a failed what-if lets its board's session close again
            sv = self._served.get(id(self.boards[req.board]))
            if sv is not None:
                sv.attached -= 1
            req._attached = False
        if req._hold:
            if req._grow:
                self.admission.shrink(req._hold)
            else:
                self.admission.release(req._hold)
            req._hold = 0.0

    def _fail(self, req: Request, e: Exception):
        req.error = str(e)
        req.done_at = time.perf_counter()
        self.completed.append(req)

    def _admit_prepared(self):
        """Queue the prepared requests whose session (and what-if stack bytes) the admission grants."""
        still = deque()
        while self._prepared:
            req, stack, solves, warm = item = self._prepared.popleft()
            try:
                host = self._host(req, stack)
                if host is stack and stack is not self.boards[req.board]:
                    sv = self._admit_served(host, len(req.fs), req._hold)   # the claim moves to its session
                    req._hold = 0.0
                else:
                    sv = self._admit_served(host, len(req.fs))
                    if sv is not None and host is not stack and req._hold:
                        # a what-if in its board's session: its growth trued up to the stack's own bytes (incidence
                        # and Woodbury columns), held while it runs
                        own = float(stack.device_bytes())
                        self.admission.true_up(req._hold, own)
                        req._hold = own
                        sv.attached -= 1
                        req._attached = False
            except (SolverError, ValueError, NotImplementedError) as e:
                self._release_request(req)
                self._fail(req, e)
                continue
            if sv is None:
                still.append(item)
                continue
            req.solutions = [None] * len(req.fs)
            sv.waiting.extend(solves)
            self.warm_starts += warm
            self._open[id(req)] = len(req.fs)
        self._prepared = still

    def turn(self) -> list:
        """Hand new requests to the set-up worker, admit the prepared ones, step every session with work once,
        collect. The requests completed in this turn."""
        refused = deque()
        while self._pending:
            req = self._pending.popleft()
            # admitted by its planned bytes before any of its device memory is made; refused ones wait holding none
            try:
                ok = self._admit_request(req)
            except (SolverError, ValueError, NotImplementedError) as e:
                self._release_request(req)
                self._fail(req, e)
                continue
            if not ok:
                self._release_request(req)
                refused.append(req)
                continue
            if self._worker is not None:
                self._prepping.append((req, self._worker.submit(self._prepare, req)))
            else:
                fut = Future()
                try:
                    fut.set_result(self._prepare(req))
                except Exception as e:          # noqa: BLE001 - handed on as the future's exception
                    fut.set_exception(e)
                self._prepping.append((req, fut))
        self._pending = refused
        while self._prepping and self._prepping[0][1].done():
            req, fut = self._prepping.popleft()
            e = fut.exception()
            if e is not None:
                self._release_request(req)
                if not isinstance(e, (SolverError, ValueError, NotImplementedError)):
                    raise e
                self._fail(req, e)
                continue
            self._prepared.append((req,) + fut.result())
        self._admit_prepared()
        finished = []
        stepped = False
        for key in list(self._served):
            sv = self._served[key]
            if not sv.waiting and not sv.slots:
                continue
            while sv.waiting:
                s = sv.waiting[0]
                slot = sv.session.submit(s.op, s.b, s.x, s.req.tol, bnorm=s.bnorm, max_cycles=s.req.max_cycles)
                if slot < 0:
                    break
                s.slot = slot
                sv.slots[slot] = sv.waiting.popleft()
            _, ready = sv.session.step(self.steps_per_turn)
            stepped = True
            if ready:
                finished += self._collect(sv)
            sv.last_used = time.perf_counter()
            self._served.move_to_end(key)
            if sv.whatif and not sv.slots and not sv.waiting:
                self._close(key)                       # a what-if stack serves its one request
        if not stepped and self._prepping:
            wait([self._prepping[0][1]], timeout=0.05, return_when=FIRST_COMPLETED)
        return finished

    def _collect(self, sv: _Served) -> list:
        import torch
        done = []
        for slot in list(sv.slots):
            res = sv.session.poll(slot)
            if res is None:
                continue
            s = sv.slots.pop(slot)
            self.solves_done += 1
            req = s.req
            if not res.converged:
                req.error = f"GMRES did not converge at {s.f:.4g} Hz ({res.iterations} iterations)"
            else:
                st = s.stack
                xd = s.x
                if s.x0 is not None:
                    with torch.cuda.stream(st.torch_stream()):
                        xd = s.x + s.x0
                if st is self.boards[req.board]:
                    self._board_x[(req.board, id(req.route), PlaneStack.freq_key(s.f))] = xd
                # Z_loop, the moment and the branch currents now from the device; the vector copied to the host
                # behind the running work, the grids unpacked when first read
                sol = s.ex.unpack_dev(s.f, xd, int(res.iterations))
                sol.residual = float(res.residual)
                req.solutions[s.k] = sol
            self._open[id(req)] -= 1
            if self._open[id(req)] == 0:
                del self._open[id(req)]
                self._release_request(req)
                st = req._stack
                if st is not None and st is not self.boards[req.board] and self._served.get(id(st)) is None:
                    st.release()                # a what-if derived from its board: its incidence and columns
                req.done_at = time.perf_counter()
                self.completed.append(req)
                done.append(req)
        return done

    def drain(self, limit_s: float = 600.0):
        t_end = time.perf_counter() + limit_s
        while self.busy():
            self.turn()
            if time.perf_counter() > t_end:
                raise TimeoutError("requests still open after the time limit")

    def close(self):
        for _, fut in self._prepping:
            fut.cancel()
        if self._worker is not None:
            self._worker.shutdown(wait=True)
        self._prepping.clear()
        for key in list(self._served):
            self._close(key)

    # ---------------------------------------------------------------- statistics
    def stats(self) -> dict:
        """p50 / p99 latency (ms) of the completed requests, solves per second since the first submission, and the
        admission's peaks."""
        lat = np.array([r.latency for r in self.completed if r.error is None]) * 1e3
        span = (max(r.done_at for r in self.completed) - self.t0) if self.completed else 0.0
        return {
            "requests": len(self.completed),
            "failed": sum(r.error is not None for r in self.completed),
            "solves": self.solves_done,
            "warm_starts": self.warm_starts,
            "p50_ms": float(np.percentile(lat, 50)) if lat.size else math.nan,
            "p99_ms": float(np.percentile(lat, 99)) if lat.size else math.nan,
            "max_ms": float(lat.max()) if lat.size else math.nan,
            "solves_per_s": self.solves_done / span if span > 0 else math.nan,
            "peak_sessions": self.admission.peak,
            "peak_bytes": self.admission.used_peak,
            "budget_bytes": self.admission.budget,
        }
