"""CPU model of the step-3 LDS ring: the order in which each wave of a GEMM workgroup issues direct
global->LDS loads (DMA) into the two ring slots, reads fragments back, waits (vmcnt / lgkmcnt) and
meets the workgroup barrier -- transcribed from chain_start / mainloop_chain
(ozablas_amd/csrc/igemm_core.hpp) and the persistent loops that drive them (k_igemm_s2,
k_igemm_s1g, k_igemm_s1 in k_igemm.hip; k_group_gemm in k_group.hip) -- checked against the ring
protocol for chains of work items of every length mix.

Hardware semantics assumed (and nothing more):
  * vmcnt(n) returns when at most n of the wave's vector-memory operations (direct loads, stores,
    atomics) are outstanding, and the counter drains in issue order;
  * lgkmcnt(0) returns when the wave's LDS reads are done;
  * s_barrier orders everything a wave did before it before everything any wave does after it.
NOT assumed: that two direct loads of one wave into the same LDS bytes land in issue order (the
counter drains in order, but the data writes into LDS are not promised to), nor that a wave's loads
finish before it exits.  Rules checked (the same R1-R4 as the kernel source states):
  R1 a read of a stage sees exactly that stage: for every wave, the last of its loads into the slot
     that is complete before the read (wait + barrier) carries the stage, and no load into the slot
     is in flight around the read;
  R2 no load into a slot is issued while any wave's read of that slot may be pending;
  R3 a wave never has two loads into one slot in flight (a vmcnt wait covers the first before the
     second is issued);
  R4 no load is in flight when a wave ends.
The previous schedule (refetch past the last stage, next prologue issued right after the epilogue's
barrier) is transcribed too, and the checker must reject it: it relied on in-order landing.
"""
import itertools

import pytest

# ------------------------------------------------------------------------------------ checker


def _barrier_index(ops):
    """bidx[i] = number of barriers strictly before position i."""
    out, n = [], 0
    for op in ops:
        out.append(n)
        if op[0] == "bar":
            n += 1
    out.append(n)
    return out


def check(waves):
    """waves: list of per-wave op lists.  Returns a list of violation strings (empty when valid)."""
    bidx = [_barrier_index(w) for w in waves]
    nbar = {b[-1] for b in bidx}
    if len(nbar) != 1:
        return [f"waves meet different numbers of barriers: {sorted(nbar)}"]

    def hb(w1, i1, w2, i2):
        """Event (w1, i1) happens before (w2, i2)."""
        if w1 == w2:
            return i1 < i2
        return bidx[w2][i2] > bidx[w1][i1]  # a barrier after i1 in w1 precedes i2 in w2

    def vm_done_at(w, i):
        """Position of the first wait in wave w that covers the VM op at i, or None."""
        ops = waves[w]
        later = 0
        for j in range(i + 1, len(ops)):
            if ops[j][0] == "vmw" and later >= ops[j][1]:
                return j
            if ops[j][0] in ("dma", "vm"):
                later += 1
        return None

    def lg_done_at(w, i):
        ops = waves[w]
        for j in range(i + 1, len(ops)):
            if ops[j][0] == "lgw":
                return j
        return None

    bad = []
    dmas = [[(i, op) for i, op in enumerate(w) if op[0] == "dma"] for w in waves]
    done = [{i: vm_done_at(w, i) for i, _ in dmas[w]} for w in range(len(waves))]
    # R4
    for w in range(len(waves)):
        for i, op in dmas[w]:
            if done[w][i] is None:
                bad.append(f"R4 wave {w} op {i}: load {op[1:]} still in flight at exit")
    # R3
    for w in range(len(waves)):
        last = {}
        for i, op in dmas[w]:
            s = op[1]
            if s in last:
                j = last[s]
                if done[w][j] is None or done[w][j] > i:
                    bad.append(f"R3 wave {w}: load {waves[w][j][1:]} (op {j}) still in flight when "
                               f"load {op[1:]} (op {i}) targets the same slot")
            last[s] = i
    # R1 / R2
    for w, ops in enumerate(waves):
        for i, op in enumerate(ops):
            if op[0] != "rd":
                continue
            slot, want = op[1], op[2]
            rdone = lg_done_at(w, i)
            for v in range(len(waves)):
                content = None
                for j, d in dmas[v]:
                    if d[1] != slot:
                        continue
                    c = done[v][j]
                    before = c is not None and hb(v, c, w, i)
                    after = rdone is not None and hb(w, rdone, v, j)
                    if before:
                        content = d[2]
                    elif not after:
                        bad.append(f"R2 wave {v} op {j}: load {d[1:]} may overlap wave {w}'s read "
                                   f"(op {i}) of slot {slot}")
                if want is not None and content != want:
                    bad.append(f"R1 wave {w} op {i}: read of slot {slot} expects {want}, wave {v}'s "
                               f"part holds {content}")
    return bad


# ------------------------------------------------------------------ transcription: current code

WAVES = 3  # wave 0 (thread 0's atomics and LDS writes) and two others; all waves run the same loop
STORES = 8  # epilogue stores per wave (any count: every wait after them is vmcnt(0))


def chain_start(ops, x, nit):
    """chain_start<C>: stage 0, and stage 1 when nit >= 2, into slots base, base + 1 (base = 0)."""
    ops.append(("dma", 0, (x, 0)))
    if nit >= 2:
        ops.append(("dma", 1, (x, 1)))


def mainloop_chain(ops, x, nit, base, nxt, ksteps):
    """mainloop_chain<C>.  nxt = (item, nit) of the next item, or None (tail = refetch)."""

    def slot(i):
        return (base + i) & 1

    def src(i):
        if i < nit:
            return (x, i)
        if nxt is not None and i - nit < nxt[1]:
            return (nxt[0], i - nit)
        return (x, nit - 1)  # refetch of this item's last stage

    ops += [("vmw", 0), ("lgw",), ("bar",)]  # wait_vmcnt<0>, lds_barrier
    if nit == 1:
        ops.append(("dma", slot(1), src(1)))
    ops.append(("rd", slot(0), (x, 0)))  # load_frags kk = 0
    for it in range(nit - 1):
        ops += [("rd", slot(it), (x, it))] * (ksteps - 1)  # k-steps 1..KSTEPS-1
        ops += [("vmw", 0), ("lgw",), ("bar",)]
        ops.append(("rd", slot(it + 1), (x, it + 1)))
        ops.append(("dma", slot(it + 2), src(it + 2)))
    ops += [("rd", slot(nit - 1), (x, nit - 1))] * (ksteps - 1)
    ops += [("lgw",), ("bar",)]  # last iteration: lds_barrier, no vmcnt wait
    ops.append(("dma", slot(nit + 1), src(nit + 1)))


def persistent(items, ksteps, dyn, wave):
    """k_igemm_s2 / k_igemm_s1g (dyn: dynamic order with claims) and k_group_gemm (dyn=False):
    one workgroup's chain of items, each given by its stage count."""
    ops = []
    if dyn:  # first_claims: thread 0's atomic, the LDS write, __syncthreads
        if wave == 0:
            ops += [("vm", "atomic"), ("vmw", 0)]
        ops += [("vmw", 0), ("lgw",), ("bar",)]
    chain_start(ops, 0, items[0])
    base = 0
    for x, nit in enumerate(items):
        more = x + 1 < len(items)
        mainloop_chain(ops, x, nit, base, (x + 1, items[x + 1]) if more else None, ksteps)
        if dyn and more and wave == 0:
            ops.append(("vm", "atomic"))  # the ticket for the item after next
        ops += [("vm", "store")] * STORES
        if not more:
            break
        if dyn:  # publish_claim: thread 0 waits for its ticket, writes it, lgkmcnt(0), barrier
            if wave == 0:
                ops.append(("vmw", 0))
            ops += [("lgw",), ("bar",)]
        base = (base + nit) & 1
    ops.append(("vmw", 0))  # final wait_vmcnt<0>
    return ops


def planar_s1(S, KT, ksteps, wave):
    """k_igemm_s1: one tile, items = groups g = 2..S+1 ((g-1)*KT stages each) in ascending order."""
    ops = []
    nits = [(g - 1) * KT for g in range(2, S + 2)]
    chain_start(ops, 0, nits[0])
    base = 0
    for x, nit in enumerate(nits):
        nxt = (x + 1, nits[x + 1]) if x + 1 < len(nits) else None
        mainloop_chain(ops, x, nit, base, nxt, ksteps)
        ops += [("vm", "store")] * STORES
        base = (base + nit) & 1
    ops.append(("vmw", 0))
    return ops


# ----------------------------------------------------------------- transcription: old schedule


def old_persistent(items_nit, n_items, ksteps, dyn, wave):
    """The earlier k_igemm_s2: issue_prologue / mainloop<C, false> per tile; the last iteration's
    load was unconditional (a refetch of the last stage into slot it % 2) and the next tile's
    prologue was issued after one barrier with no vmcnt wait."""
    nit = items_nit
    ops = []

    def prologue(x):
        for s in range(2):
            ops.append(("dma", s, (x, min(s, nit - 1))))

    prologue(0)
    for x in range(n_items):
        more = x + 1 < n_items
        if dyn and wave == 0:
            ops.append(("vm", "atomic"))
        ops += [("vmw", 0), ("lgw",), ("bar",)]  # stage_ready(0)
        ops.append(("rd", 0, (x, 0)))
        for it in range(nit):
            ops += [("rd", it % 2, (x, it))] * (ksteps - 1)
            ops += [("vmw", 0), ("lgw",), ("bar",)]  # stage_ready(it + 1)
            ops.append(("rd", (it + 1) % 2, (x, it + 1) if it + 1 < nit else None))  # stale at the end
            ops.append(("dma", it % 2, (x, min(it + 2, nit - 1))))
        if dyn:
            ops.append(("lgw",))  # thread 0's s_next write, lgkmcnt(0)
        ops.append(("bar",))  # block_sync_lds
        if more:
            prologue(x + 1)
        ops += [("vm", "store")] * STORES
    return ops


# ------------------------------------------------------------------------------------ tests


def chains(max_len=4, nits=(1, 2, 3, 4)):
    for n in range(1, max_len + 1):
        yield from itertools.product(nits, repeat=n)


@pytest.mark.parametrize("ksteps", [2, 4])  # BK = 64 (128-edge grouped tiles) and BK = 128
@pytest.mark.parametrize("dyn", [False, True])
def test_chained_items_keep_the_ring_protocol(ksteps, dyn):
    for items in chains():
        waves = [persistent(list(items), ksteps, dyn, w) for w in range(WAVES)]
        bad = check(waves)
        assert not bad, f"items {items}: {bad[:3]}"


@pytest.mark.parametrize("S", [1, 2, 3, 4])
@pytest.mark.parametrize("KT", [1, 2, 3])
def test_planar_scheme1_groups_keep_the_ring_protocol(S, KT):
    waves = [planar_s1(S, KT, 2, w) for w in range(WAVES)]
    assert not check(waves)


def test_checker_rejects_the_earlier_schedule():
    """The refetch past the last stage stays in flight when the next tile's prologue loads the
    same slot (R3), and the workgroup exits with it in flight (R4); in the static order the stale
    read of the next slot can still be pending when the next prologue loads it (R2)."""
    for dyn in (True, False):
        for nit in (1, 2, 3):
            waves = [old_persistent(nit, 3, 4, dyn, w) for w in range(WAVES)]
            bad = check(waves)
            rules = {b[:2] for b in bad}
            assert "R3" in rules and "R4" in rules, (dyn, nit, bad[:3])
            if not dyn:
                assert "R2" in rules, (nit, bad[:3])


def test_checker_catches_planted_mistakes():
    """Each rule fires on a one-line change of the current schedule."""
    items = [3, 1, 2]

    def variant(mutate):
        return [mutate(persistent(items, 4, True, w)) for w in range(WAVES)]

    def drop_first(kind):
        def f(ops):
            ops = list(ops)
            k = next(i for i, op in enumerate(ops) if op == kind and i > 12)
            del ops[k]
            return ops
        return f

    def drop_last_wait(ops):
        ops = list(ops)
        assert ops[-1] == ("vmw", 0)
        return ops[:-1]

    # a stage wait dropped: a read can see a load in flight (R1/R2)
    assert {b[:2] for b in check(variant(drop_first(("vmw", 0))))} & {"R1", "R2"}
    # an lgkmcnt wait dropped before a barrier: a load may overwrite a pending read (R2)
    assert "R2" in {b[:2] for b in check(variant(drop_first(("lgw",))))}
    # the final drain dropped: a load in flight at exit (R4)
    assert "R4" in {b[:2] for b in check(variant(drop_last_wait))}
