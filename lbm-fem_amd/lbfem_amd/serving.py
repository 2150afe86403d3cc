"""Continuous-batching service of flow solves on one GPU.

Requests (a flow, a mesh, a degree, Re, U0, and either a number of time steps
or a steady-state tolerance with a step cap) arrive over time. The server keeps
every admitted request's populations resident and advances all of them
together through one running lbfem_batch (one launch per (degree, scheme,
precision) group per step, steps replayed from captured HIP graphs). Between
ticks it admits requests that have arrived, as slots and the HBM budget allow,
by adding them to the running batch, and retires requests by removing them:
  - a fixed-step request after its steps;
  - a steady request when max |u - u_prev| / U0 over a window of `window`
    steps falls below its tolerance (the steady test of the plain-method
    configuration; the reference's own test is the mean rate,
    apps/cgdbe/cgdbe.cc:385-393), or at its step cap ("not steady");
  - any request whose populations are no longer finite ("diverged", as the
    reference's run ends with an error, apps/cgdbe/cgdbe.cc:583-590).
The velocity change and the finiteness check are one device reduction over the
members concerned (lbfem_batch_velocity_change); the batch is never rebuilt
and its graphs are kept while members come and go. A request's scheme (mesh,
initial populations: host work of up to a second for the largest meshes) is
prepared on worker threads as soon as it arrives and the budget allows, while
the device steps the others; admission takes prepared requests in arrival
order, at the end of each tick (at most tick_steps steps). A request's result
is bitwise the one it gets stepped alone under the same retirement rules.
Reported per request: arrival, admission and completion times, steps, status;
overall: time-to-solution percentiles and node.population updates per second.
"""
import concurrent.futures
import time

import numpy as np

from . import workloads


def request_bytes(inst, precision_bytes=4):
    """HBM held by a request: two population buffers of 9 values per node, the monitor's previous velocity
    (2 FP64 values per node), plus geometry (small)."""
    return workloads.nodes(inst) * (9 * 2 * precision_bytes + 16) + (1 << 16)


class Server:
    def __init__(self, A, precision=None, budget_bytes=240e9, max_slots=96, tick_steps=200, window=1000,
                 graph_steps=8, grid=2048, scheme=None, cfl=0.5, keep_results=False, workers=2):
        self.A = A
        self.precision = A.MIXED if precision is None else precision
        self.budget = budget_bytes
        self.max_slots = max_slots
        self.tick_steps = tick_steps
        self.window = window
        self.graph_steps = graph_steps
        self.grid = grid  # persistent blocks per group launch: membership changes keep the captured graphs
        self.scheme = scheme
        self.cfl = cfl
        self.keep_results = keep_results
        self.workers = workers

    def _next_event(self, a):
        """steps until this member's next retirement check"""
        if a["req"].get("mode") == "steady":
            return min(self.window - a["since"], a["cap"] - a["done"])
        return a["cap"] - a["done"]

    def run(self, requests, clock=time.perf_counter):
        """Serve `requests` (dicts with arrival [s], case, N, p, Re, U, mach, steps, id, and optionally
        mode="steady" with tol) in open loop: a request is visible from its arrival time on. Returns
        (per-request records, summary)."""
        A = self.A
        pending = sorted(requests, key=lambda r: r["arrival"])
        waiting, active, done = [], [], []
        preparing = []  # (request, future of its scheme), in arrival order; their memory is reserved
        used = 0
        pool = concurrent.futures.ThreadPoolExecutor(max_workers=self.workers)
        pb = 4 if self.precision == A.MIXED else 8
        batch = A.Batch([], graph_steps=self.graph_steps)
        batch.set_shape(-1, self.grid)
        t0 = clock()
        ticks = 0

        def retire(a, status, t_now):
            nonlocal used
            batch.remove(a["scheme"])
            r = a["req"]
            rec = {"id": r["id"], "arrival": r["arrival"], "admit": a["admit"], "done": t_now,
                   "steps": a["done"], "nodes": workloads.nodes(r), "status": status, "tts": t_now - r["arrival"]}
            if self.keep_results:
                rec["f"] = a["scheme"].f
            done.append(rec)
            used -= request_bytes(r, pb)

        while pending or waiting or preparing or active:
            now = clock() - t0
            while pending and pending[0]["arrival"] <= now:
                waiting.append(pending.pop(0))
            # preparation: FIFO while slots and memory allow (the memory is reserved from here)
            while (waiting and len(active) + len(preparing) < self.max_slots
                   and used + request_bytes(waiting[0], pb) <= self.budget):
                r = waiting.pop(0)
                used += request_bytes(r, pb)
                preparing.append((r, pool.submit(workloads.make_scheme, A, r, self.precision, self.scheme,
                                                 self.cfl)))
            # admission: prepared requests in arrival order join the running batch; with nothing to step,
            # wait for the next one
            if preparing and not active:
                preparing[0][1].result()
            fresh = []
            while preparing and preparing[0][1].done():
                r, fut = preparing.pop(0)
                s = fut.result()
                a = {"req": r, "scheme": s, "done": 0, "since": 0, "cap": r["steps"], "admit": clock() - t0}
                batch.add(s)
                active.append(a)
                fresh.append(a)
            steady_new = [a["scheme"] for a in fresh if a["req"].get("mode") == "steady"]
            if steady_new:  # the start of their first window
                batch.velocity_change(steady_new)
            if not active:
                if pending:  # idle until the next arrival
                    time.sleep(max(0.0, pending[0]["arrival"] - (clock() - t0)))
                continue
            k = min(self.tick_steps, min(self._next_event(a) for a in active))
            batch.step(k)
            ticks += 1
            for a in active:
                a["done"] += k
                a["since"] += k
            # the members due a check: fixed-step members every tick (finiteness), steady members at the end of
            # a window (and then their window restarts)
            due = [a for a in active if a["req"].get("mode") != "steady" or a["since"] >= self.window
                   or a["done"] >= a["cap"]]
            mx, _, bad = batch.velocity_change([a["scheme"] for a in due])
            t_now = clock() - t0
            idx = {id(s): i for i, s in enumerate(batch.schemes)}
            still = []
            for a in active:
                i = idx[id(a["scheme"])]
                status = None
                if bad[i] > 0:
                    status = "diverged"
                elif a["req"].get("mode") == "steady":
                    if a["since"] >= self.window:
                        a["since"] = 0
                        if mx[i] / a["req"]["U"] < a["req"].get("tol", 1e-6):
                            status = "steady"
                    if status is None and a["done"] >= a["cap"]:
                        status = "not steady"
                elif a["done"] >= a["cap"]:
                    status = "steps"
                if status is None:
                    still.append(a)
                else:
                    retire(a, status, t_now)
            active = still
        batch.close()
        pool.shutdown()
        t_end = clock() - t0
        tts = np.array([d["tts"] for d in done])
        work = sum(d["nodes"] * d["steps"] for d in done)
        first = min(d["arrival"] for d in done) if done else 0.0
        statuses = {}
        for d in done:
            statuses[d["status"]] = statuses.get(d["status"], 0) + 1
        summary = {
            "requests": len(done), "ticks": ticks, "wall_s": t_end, "statuses": statuses,
            "tts_p50_s": float(np.percentile(tts, 50)) if len(tts) else None,
            "tts_p99_s": float(np.percentile(tts, 99)) if len(tts) else None,
            "tts_max_s": float(tts.max()) if len(tts) else None,
            "MLUPS": 9 * work / max(1e-12, t_end - first) / 1e6,
            "node_steps": work,
        }
        return done, summary


def run_alone(A, request, precision, window=1000, scheme=None, cfl=0.5):
    """A request stepped by itself with the server's retirement rules (the same checks at the same steps):
    returns (scheme, steps, status). The served result of the request is bitwise this scheme's populations."""
    s = workloads.make_scheme(A, request, precision, scheme, cfl)
    cap = request["steps"]
    steady = request.get("mode") == "steady"
    if steady:
        s.velocity_change()
    done, since = 0, 0
    while True:
        k = min(window - since, cap - done) if steady else cap - done
        s.step(k)
        done += k
        since += k
        if not steady or since >= window or done >= cap:
            mx, _, bad = s.velocity_change()
            if bad > 0:
                return s, done, "diverged"
            if steady:
                if since >= window:
                    since = 0
                    if mx / request["U"] < request.get("tol", 1e-6):
                        return s, done, "steady"
                if done >= cap:
                    return s, done, "not steady"
            else:
                return s, done, "steps"
