"""Schedule autotuner for the engine on the device it runs on.

Picks, by measurement, the knobs of fhea_ctx_set_tuning that change speed but
not results.  tune() times a caller's workload (for example the HMult batch being
served) over the knobs below, one knob at a time, and checks that its result words
do not change; tune_ntt() and tune_chunk() time the transform / a submit alone.
The knobs:
  * the split of a two-pass FP64 NTT (stages in the first pass) for forward and
    inverse transforms, among every split whose passes have 6..12 stages;
  * the number of resident NTT workgroups per CU;
  * the number of batch items per key-switch pipeline chunk, on a caller-given
    list of work items (for example the HMult batch being served);
  * one to four streams for a batched HMult (its parts overlapping on the device).
Every NTT candidate's output (first and last polynomial of the batch) is checked
against the default schedule's word for word before it can be chosen.  Choices can be cached in a JSON file keyed by the
parameter set.
"""
import json
import os

import numpy as np

from . import _lib

KNOBS = {"ntt_split_fwd": _lib.TUNE_NTT_SPLIT_FWD, "ntt_split_inv": _lib.TUNE_NTT_SPLIT_INV,
         "ntt_wg_per_cu": _lib.TUNE_NTT_WG_PER_CU, "chunk": _lib.TUNE_CHUNK, "streams": _lib.TUNE_STREAMS}


def _time(ctx, fn, reps, rounds=2):
    """Milliseconds per run of fn: the best of `rounds` timings of `reps` runs each
    (the minimum is the estimate least disturbed by other activity on the device)."""
    fn()
    ctx.sync()
    best = None
    for _ in range(rounds):
        ctx.timer_start()
        for _ in range(reps):
            fn()
        ms = ctx.timer_stop() / reps
        best = ms if best is None else min(best, ms)
    return best


def ntt_splits(logN):
    """First-pass lengths of the two-pass NTT; at N = 2^13, 2^14 also logN (one pass
    over the whole limb)."""
    two = [k for k in range(6, 13) if 6 <= logN - k <= 12] if logN > 12 else []
    return two + ([logN] if logN in (13, 14) else [])


def tune_ntt(ctx, scratch_off, polys, limbs, reps=5, log=None, src=None):
    """Chooses the forward and inverse splits and the workgroups per CU on
    `polys` x `limbs` limbs: the residues already at `src`, or (src None) uniform
    data written at scratch_off.  The transforms write to scratch_off (+ polys *
    limbs * N when the data is written there too).  Returns the timings."""
    N = ctx.N
    logN = N.bit_length() - 1
    W = polys * limbs * N
    if src is None:
        rng = np.random.default_rng(12345)
        data = np.concatenate([np.stack([rng.integers(0, q, N, dtype=np.uint64) for q in ctx.q[:limbs]]).reshape(-1)
                               for _ in range(polys)])
        src, dst = scratch_off, scratch_off + W
        ctx.write_poly(src, data)
    else:
        dst = scratch_off
    timings = {}
    P = limbs * N

    def sample():  # first and last polynomial of the output
        return np.concatenate([ctx.read_poly(dst, P), ctx.read_poly(dst + W - P, P)])

    for inverse, knob in ((False, "ntt_split_fwd"), (True, "ntt_split_inv")):
        ctx.set_tuning(KNOBS[knob], 0)
        ctx.ntt_batch(src, dst, polys, limbs, inverse)
        ref = sample()
        best = (None, 0)
        for k in [0] + ntt_splits(logN):
            ctx.set_tuning(KNOBS[knob], k)
            ctx.ntt_batch(src, dst, polys, limbs, inverse)
            if not np.array_equal(sample(), ref):
                raise RuntimeError(f"NTT split {k} ({knob}) changed the result")
            ms = _time(ctx, lambda: ctx.ntt_batch(src, dst, polys, limbs, inverse), reps)
            timings[f"{knob}={k}"] = round(ms, 4)
            if best[0] is None or ms < best[0]:
                best = (ms, k)
        ctx.set_tuning(KNOBS[knob], best[1])
        if log:
            log(f"autotune {knob}: {best[1]} ({best[0]:.3f} ms)")
    best = (None, 0)
    for w in (0, 4, 3, 2):
        ctx.set_tuning(KNOBS["ntt_wg_per_cu"], w)
        ms = _time(ctx, lambda: (ctx.ntt_batch(src, dst, polys, limbs, False),
                                 ctx.ntt_batch(dst, dst, polys, limbs, True)), reps)
        timings[f"ntt_wg_per_cu={w}"] = round(ms, 4)
        if best[0] is None or ms < best[0]:
            best = (ms, w)
    ctx.set_tuning(KNOBS["ntt_wg_per_cu"], best[1])
    return timings


def candidates(ctx, knob):
    if knob in ("ntt_split_fwd", "ntt_split_inv"):
        return [0] + ntt_splits(ctx.N.bit_length() - 1)
    if knob == "ntt_wg_per_cu":
        return [0, 4, 3, 2]
    if knob == "streams":
        return [1, 2]
    return [4, 8, 16, 32, 64]


def tune(ctx, run, check=None, knobs=("streams", "ntt_split_fwd", "ntt_split_inv", "ntt_wg_per_cu", "chunk"),
         reps=3, log=None):
    """Coordinate search over `knobs`, timing `run()` (the workload itself, launched
    on the context's stream) for every candidate of one knob with the others fixed
    at their best so far.  check(): a result digest of the workload (for example its
    output words); every candidate must reproduce the digest of the starting
    schedule, else RuntimeError.  Returns {knob=value: ms}."""
    timings = {}
    ref = None
    if check is not None:
        run()
        ctx.sync()
        ref = check()
    for knob in knobs:
        best = (None, ctx.get_tuning(KNOBS[knob]))
        for v in candidates(ctx, knob):
            ctx.set_tuning(KNOBS[knob], v)
            ms = _time(ctx, run, reps)
            if check is not None and not np.array_equal(check(), ref):
                raise RuntimeError(f"schedule {knob}={v} changed the workload's result")
            timings[f"{knob}={v}"] = round(ms, 4)
            if best[0] is None or ms < best[0]:
                best = (ms, v)
        ctx.set_tuning(KNOBS[knob], best[1])
        if log:
            log(f"autotune {knob}: {best[1]} ({best[0]:.3f} ms)")
    return timings


def tune_chunk(ctx, works, candidates=(8, 16, 32), reps=3, log=None):
    """Chooses the key-switch chunk size on a representative list of work items."""
    timings = {}
    best = (None, ctx.get_tuning(KNOBS["chunk"]))
    for c in candidates:
        ctx.set_tuning(KNOBS["chunk"], c)
        ms = _time(ctx, lambda: ctx.submit(works), reps)
        timings[f"chunk={c}"] = round(ms, 4)
        if best[0] is None or ms < best[0]:
            best = (ms, c)
    ctx.set_tuning(KNOBS["chunk"], best[1])
    if log:
        log(f"autotune chunk: {best[1]} ({best[0]:.3f} ms)")
    return timings


def current(ctx):
    return {k: ctx.get_tuning(v) for k, v in KNOBS.items()}


def apply(ctx, choices):
    for k, v in choices.items():
        ctx.set_tuning(KNOBS[k], int(v))


def cache_key(ctx):
    return f"N{ctx.N}_L{ctx.L}_K{ctx.K}_dnum{ctx.dnum}_q{ctx.q[0].bit_length()}"


def load_cached(ctx, path):
    """Applies cached choices for this parameter set; returns them or None."""
    if path and os.path.exists(path):
        with open(path) as f:
            entry = json.load(f).get(cache_key(ctx))
        if entry:
            apply(ctx, entry)
            return entry
    return None


def save_cached(ctx, path):
    data = {}
    if os.path.exists(path):
        with open(path) as f:
            data = json.load(f)
    data[cache_key(ctx)] = current(ctx)
    tmp = path + ".tmp"
    with open(tmp, "w") as f:
        json.dump(data, f, indent=1, sort_keys=True)
    os.replace(tmp, path)
