"""Continuous-batching service of flow solves on one GPU.

Requests (a flow, a mesh, a degree, Re, U0 and a number of time steps) arrive
over time. The server keeps every admitted request's populations resident and
advances all of them together through one lbfem_batch (one launch per
(degree, scheme, precision) group per step, steps replayed from captured HIP
graphs). Between ticks it admits requests that have arrived, as slots and the
HBM budget allow, and retires those that have taken all their steps; a
request's result is bitwise the one it gets stepped alone. Reported per request:
arrival, admission and completion times; overall: time-to-solution percentiles
and node.population updates per second.
"""
import time

import numpy as np

from . import workloads


def request_bytes(inst, precision_bytes=4):
    """HBM held by a request: two population buffers of 9 values per node, plus geometry (small)."""
    return workloads.nodes(inst) * 9 * 2 * precision_bytes + (1 << 16)


class Server:
    def __init__(self, A, precision=None, budget_bytes=240e9, max_slots=96, tick_steps=400, graph_steps=8,
                 scheme=None, cfl=0.5, keep_results=False):
        self.A = A
        self.precision = A.MIXED if precision is None else precision
        self.budget = budget_bytes
        self.max_slots = max_slots
        self.tick_steps = tick_steps
        self.graph_steps = graph_steps
        self.scheme = scheme
        self.cfl = cfl
        self.keep_results = keep_results

    def run(self, requests, clock=time.perf_counter):
        """Serve `requests` (dicts with arrival [s], case, N, p, Re, mach, steps, id) in open loop: a request
        is visible from its arrival time on. Returns (per-request records, summary)."""
        A = self.A
        pending = sorted(requests, key=lambda r: r["arrival"])
        waiting, active, done = [], [], []
        used = 0
        batch = None
        members_changed = False
        t0 = clock()
        ticks = 0
        pb = 4 if self.precision == A.MIXED else 8
        while pending or waiting or active:
            now = clock() - t0
            while pending and pending[0]["arrival"] <= now:
                waiting.append(pending.pop(0))
            # admission: FIFO while slots and memory allow
            while waiting and len(active) < self.max_slots and used + request_bytes(waiting[0], pb) <= self.budget:
                r = waiting.pop(0)
                s = workloads.make_scheme(A, r, self.precision, self.scheme, self.cfl)
                active.append({"req": r, "scheme": s, "left": r["steps"], "admit": clock() - t0})
                used += request_bytes(r, pb)
                members_changed = True
            if not active:
                if pending:  # idle until the next arrival
                    time.sleep(max(0.0, pending[0]["arrival"] - (clock() - t0)))
                continue
            if members_changed:
                if batch is not None:
                    batch.close()
                batch = A.Batch([a["scheme"] for a in active], graph_steps=self.graph_steps)
                members_changed = False
            k = min(self.tick_steps, min(a["left"] for a in active))
            batch.step(k)
            batch.synchronize()
            ticks += 1
            t_now = clock() - t0
            still = []
            for a in active:
                a["left"] -= k
                if a["left"] > 0:
                    still.append(a)
                    continue
                r = a["req"]
                rec = {"id": r["id"], "arrival": r["arrival"], "admit": a["admit"], "done": t_now,
                       "steps": r["steps"], "nodes": workloads.nodes(r), "tts": t_now - r["arrival"]}
                if self.keep_results:
                    rec["f"] = a["scheme"].f
                done.append(rec)
                used -= request_bytes(r, pb)
                members_changed = True
            if members_changed:
                batch.close()
                batch = None
                active = still
        if batch is not None:
            batch.close()
        t_end = clock() - t0
        tts = np.array([d["tts"] for d in done])
        work = sum(d["nodes"] * d["steps"] for d in done)
        first = min(d["arrival"] for d in done) if done else 0.0
        summary = {
            "requests": len(done), "ticks": ticks, "wall_s": t_end,
            "tts_p50_s": float(np.percentile(tts, 50)) if len(tts) else None,
            "tts_p99_s": float(np.percentile(tts, 99)) if len(tts) else None,
            "tts_max_s": float(tts.max()) if len(tts) else None,
            "MLUPS": 9 * work / max(1e-12, t_end - first) / 1e6,
            "node_steps": work,
        }
        return done, summary
