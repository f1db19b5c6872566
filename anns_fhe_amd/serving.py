"""Serving loop for encrypted IVF-PQ queries on one GPU.

No reference counterpart: the reference serves one query at a time (a CPU
thread pool over OpenFHE, software_golden_model/fhe_core/main.cpp:259-305; its
FPGA path is strictly sequential, cpp_host_model/src/interactive_query_manager.cpp).

* Parameter sets.  Each set (for example N=2^14/L=8 and N=2^16/L=24) owns one
  engine context (its own HIP stream and HBM arenas), its encoded database, its
  evaluation keys (from the secure generator unless a seed is given) and a number
  of query slots (fixed, or as many as a share of the free HBM holds).  Sets of one
  `group` are replicas (same parameters and an explicit seed, so the same keys and
  database words): a
  request names the group and is admitted to the replica with the most free slots,
  which gives the group one more worker and stream on the device.
* Admission.  A request is admitted when its set has a free slot and the HBM
  budget (static state of every set + the slots of every admitted request)
  allows it; otherwise it waits, first come first served.
* Continuous batching.  One worker thread per set loops: every request admitted
  since its last launch goes through stage 1 together (the rotations of all of
  their queries are one batched key-switch submission), every request whose
  client round trip has returned goes through stage 3 together.  Requests that
  arrive while a batch is on the GPU join the next launch.  Both stages are
  launched before either's results are read back; the reads run on the context's
  copy stream, so stage 1's results reach their clients while stage 3 still runs,
  and stage 3's parts are read while the next part's sums run.  The two sets' work
  overlaps on the device (separate streams), and the client side (decrypt,
  cluster choice, top-k) runs on a thread pool, overlapping the GPU.
* Latency of a request = completion of its top-k minus its arrival time.
"""
import collections
import concurrent.futures as cf
import threading
import time
from dataclasses import dataclass, field

import numpy as np

from . import ivfpq as IV
from .client import CkksClient
from .driver import FheContext


@dataclass
class Request:
    rid: int
    pset: str          # parameter set, or group of replica sets
    query: np.ndarray
    n_probe: int
    arrival: float
    ct: np.ndarray = None
    slot: int = -1
    probes: np.ndarray = None
    ids: np.ndarray = None
    dist: np.ndarray = None
    t_admit: float = 0.0
    t_done: float = 0.0
    events: dict = field(default_factory=dict)
    replica: str = ""  # the set it was admitted to


class ParamSet:
    """One parameter set: context, client keys, encoded database, slots."""

    def __init__(self, name, ctx, ck, index, max_slots, max_g, lt=2, log_scale=40, seed=None, threads=0, log=None,
                 group=None):
        self.name = name
        self.group = group or name
        self.ctx, self.ck = ctx, ck
        self.max_g = max_g
        self.client = IV.FheIvfPqClient(ck, dim=index.dim, log_scale=log_scale, lt=lt, seed=seed)
        self.server = IV.FheIvfPqServer(ctx, ck, index, lt=lt, log_scale=log_scale, max_queries=max_slots,
                                        max_blocks_per_query=max_g, threads=threads, log=log)
        rel, rots = self.client.evaluation_keys()
        self.server.load_keys(rel, rots)
        self.max_slots = max_slots
        ksz = rel.size
        self.static_bytes = 8 * (self.server.pt_words + index.dim * ksz)
        # a slot's own words and its share of the output area
        self.slot_bytes = 8 * (self.server.slot_words + max_g * 2 * lt * ctx.N)

    @staticmethod
    def sizes(N, L, K, dnum, index, max_slots, lt=2, max_probe=8):
        """(poly_words, key_words, max_g) for a set serving `index` with up to max_probe probes."""
        nb = N // 2
        n_blocks = 1 + sum(-(-len(l) // nb) for l in index.lists)
        per = sorted((-(-len(l) // nb) for l in index.lists), reverse=True)
        max_g = max(1, sum(per[:max_probe]))
        words = IV.FheIvfPqServer.arena_words(N, L, lt, index.dim, n_blocks, max_slots, max_g)
        return words, index.dim * 2 * dnum * (lt + K) * N + 8, max_g

    @classmethod
    def create(cls, name, N, L, K, dnum, index, max_slots=None, lt=2, log_scale=40, max_probe=8, seed=None, threads=0,
               log=None, group=None, hbm_share=0.5, device=0, slot_cap=256):
        """max_slots None: as many query slots as `hbm_share` of the device's free HBM holds
        beside the set's database, keys and pipeline workspace (at most slot_cap)."""
        if max_slots is None:
            max_slots = cls.slots_for_memory(N, L, K, dnum, index, lt, max_probe, hbm_share, device, slot_cap)
        words, kw, max_g = cls.sizes(N, L, K, dnum, index, max_slots, lt, max_probe)
        ctx = FheContext(N, L, K, dnum, poly_words=words, key_words=kw, device=device)
        ck = CkksClient(0, 0, 0, 0, ctx=ctx)
        return cls(name, ctx, ck, index, max_slots, max_g, lt, log_scale, seed, threads, log, group)

    @classmethod
    def slots_for_memory(cls, N, L, K, dnum, index, lt=2, max_probe=8, hbm_share=0.5, device=0, slot_cap=256):
        """Query slots that fit in hbm_share of the free device memory (hipMemGetInfo): the
        arena grows by a fixed number of words per slot; one slot's worth of key-switch
        workspace per slot is reserved on top (the stage-1 pipeline's chunks)."""
        from . import _lib
        free, total = _lib.mem_info(device)
        w1, kw, _ = cls.sizes(N, L, K, dnum, index, 1, lt, max_probe)
        w2, _, _ = cls.sizes(N, L, K, dnum, index, 2, lt, max_probe)
        per_slot = 8 * (w2 - w1)
        static = 8 * (w1 - (w2 - w1) + kw)
        avail = int(free * hbm_share) - static
        per_slot_total = 2 * per_slot  # the slot's words and its share of the workspace
        if avail < per_slot_total:
            raise MemoryError(f"{name_of(N, L)}: {free} free bytes hold no query slot")
        return max(1, min(slot_cap, avail // per_slot_total))

    def encrypt(self, q, seed=None):
        """The query under the set's key: with a set seed (reproducible keys) `seed` makes
        the encryption reproducible too; without one it is drawn from the secure generator."""
        if self.client.seed0 is None:
            seed = None
        return self.ck.encrypt_values(self.client.pk, np.asarray(q, np.float64), self.ck.L, self.client.scale, seed)

    def close(self):
        close = getattr(self.ctx, "close", None)
        if close:
            close()


def name_of(N, L):
    return f"N=2^{N.bit_length() - 1}/L={L}"


def poisson_arrivals(rate, n, seed):
    rng = np.random.default_rng(seed)
    return np.cumsum(rng.exponential(1.0 / rate, n))


class ServingLoop:
    def __init__(self, psets, hbm_budget_bytes, max_batch=32, client_workers=8):
        self.ps = {p.name: p for p in psets}
        self.groups = {}
        for p in psets:
            self.groups.setdefault(p.group, []).append(p.name)
        for names in self.groups.values():  # replicas answer each other's clients
            pk0 = self.ps[names[0]].client.pk
            if any(not np.array_equal(self.ps[m].client.pk, pk0) for m in names[1:]):
                raise ValueError("replica sets of a group must share their keys (same parameters and seed)")
        self.budget = hbm_budget_bytes
        self.static = sum(p.static_bytes for p in psets)
        if self.static > self.budget:
            raise ValueError("HBM budget below the resident database and keys")
        self.max_batch = max_batch
        self.client_workers = client_workers

    def run(self, requests, encrypt=True):
        """Serves `requests` (sorted by arrival, seconds from start) and returns them completed."""
        if encrypt:
            for r in requests:
                if r.ct is None:
                    r.ct = self.ps[self.groups[r.pset][0]].encrypt(r.query, 7919 * (r.rid + 1))
        lock = threading.Condition()
        waiting = collections.deque()
        free = {n: list(range(p.max_slots)) for n, p in self.ps.items()}
        used = [self.static]
        st1 = {n: [] for n in self.ps}   # admitted, waiting for stage 1
        st3 = {n: [] for n in self.ps}   # client stage 2 done, waiting for stage 3
        done = []
        pool = cf.ThreadPoolExecutor(self.client_workers)
        t0 = time.perf_counter()
        n_total = len(requests)
        stop = threading.Event()
        errors = []

        def now():
            return time.perf_counter() - t0

        def admit():
            # FCFS: the head of the queue blocks the ones behind it
            while waiting:
                r = waiting[0]
                name = max(self.groups[r.pset], key=lambda m: len(free[m]))  # first of the freest
                p = self.ps[name]
                if not free[name] or used[0] + p.slot_bytes > self.budget:
                    break
                waiting.popleft()
                r.replica = name
                r.slot = free[name].pop()
                used[0] += p.slot_bytes
                r.t_admit = now()
                st1[name].append(r)

        def release(r):
            p = self.ps[r.replica]
            free[r.replica].append(r.slot)
            used[0] -= p.slot_bytes

        def client_stage2(r, coarse):
            p = self.ps[r.replica]
            srv = p.server
            scale = p.client.result_scale(p.ck.q[srv.lt - 1])
            d = p.client.decrypt_distances(coarse, srv.lt - 1, scale, srv.S)[:srv.ix.n_list]
            r.probes = np.argsort(d, kind="stable")[:r.n_probe]
            r.events["stage2"] = now()
            with lock:
                st3[r.replica].append(r)
                lock.notify_all()

        def client_stage4(r, blocks, fine):
            p = self.ps[r.replica]
            srv = p.server
            scale = p.client.result_scale(p.ck.q[srv.lt - 1])
            ids, dist = [], []
            dec = p.client.decrypt_distances_batch(fine, srv.lt - 1, scale, srv.S) if blocks else None
            for g, b in enumerate(blocks):
                cand = srv.blocks[b][1]
                ids.append(cand)
                dist.append(dec[g][:len(cand)])
            ids = np.concatenate(ids)
            dist = np.concatenate(dist)
            order = IV.topk_stable(dist, 10)
            r.ids, r.dist = ids[order], dist[order]
            r.t_done = now()
            with lock:
                done.append(r)
                lock.notify_all()

        def worker(name):
            p = self.ps[name]
            srv = p.server
            try:
                while not stop.is_set():
                    with lock:
                        admit()
                        b1 = st1[name][:self.max_batch]
                        del st1[name][:len(b1)]
                        b3 = st3[name][:self.max_batch]
                        del st3[name][:len(b3)]
                        if not b1 and not b3:
                            lock.wait(0.002)
                            continue
                    # launch both stages, their reads queued on the copy stream; stage 1's
                    # results come back (and go to the clients) while stage 3 runs
                    pend1 = pend3 = None
                    if b1:
                        for r in b1:
                            srv.upload(r.slot, r.ct)
                        srv.stage1_coarse_centroid_distance([r.slot for r in b1])
                        pend1 = srv.read_distances_submit([(r.slot, 1) for r in b1])
                    if b3:
                        pend3 = srv.stage3_submit([(r.slot, r.probes) for r in b3], chunks=2)
                    if pend1 is not None:
                        cmap = srv.collect(pend1)
                        for r in b1:
                            coarse = cmap[r.slot][0]
                            r.events["stage1"] = now()
                            pool.submit(client_stage2, r, coarse)
                    if pend3 is not None:
                        fmap = srv.collect(pend3)
                        outs = [pend3["blocks"][r.slot] for r in b3]
                        for r, blocks in zip(b3, outs):
                            fine = fmap.get(r.slot)
                            r.events["stage3"] = now()
                            with lock:
                                release(r)
                                lock.notify_all()
                            pool.submit(client_stage4, r, blocks, fine)
            except Exception as e:  # surfaced by run()
                errors.append(e)
                stop.set()
                with lock:
                    lock.notify_all()

        threads = [threading.Thread(target=worker, args=(n,), daemon=True) for n in self.ps]
        for th in threads:
            th.start()
        k = 0
        try:
            while len(done) < n_total and not errors:
                with lock:
                    while k < n_total and requests[k].arrival <= now():
                        waiting.append(requests[k])
                        k += 1
                    lock.notify_all()
                    if len(done) >= n_total:
                        break
                    nxt = requests[k].arrival - now() if k < n_total else 0.01
                    lock.wait(max(0.0, min(nxt, 0.01)))
        finally:
            stop.set()
            with lock:
                lock.notify_all()
            for th in threads:
                th.join()
            pool.shutdown(wait=True)
        if errors:
            raise errors[0]
        self.wall = now()
        return requests


def latency_stats(requests):
    lat = np.array([r.t_done - r.arrival for r in requests]) * 1e3
    span = max(r.t_done for r in requests) - min(r.arrival for r in requests)
    return {"p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
            "mean_ms": float(lat.mean()), "queries_per_s": len(requests) / span, "n": len(requests)}
