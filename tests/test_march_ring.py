"""LDS ring schedule of the marching stage kernel (march_body in
navierstokes_amd/csrc/ns_kernels.hip), checked on the CPU.

The kernel keeps window rows (U and T) and level-1 rows in LDS rings indexed
by row modulo the ring depth.  Waves of a block run the same phases between
barriers, but within one barrier epoch a fast wave may be in a later phase
than a slow one.  The model below lists, per phase, the rows each phase reads
and writes, with the barriers where the kernel has them; within every epoch
no slot may be written that another access of the same epoch reads or
writes for a different row, and no two rows written in one epoch may share a
slot.  The kernel's ring (2 RB + 9 window rows, RB + 6 level-1 rows) passes;
a window ring of only the RB + 9 rows live in one step fails unless barriers
follow the outputs and the prologue, which the kernel does not have (a
6-row-per-step kernel with that ring and those barriers was tried and
dropped: its register limit for 3 waves per SIMD made it spill).
"""
import pytest


def ring_depths(rb, tight):
    return (rb + 9 if tight else 2 * rb + 9), rb + 6


def schedule(rb, y0, nsteps, tight, barriers=True):
    """Epochs of the march: lists of phases, each a list of accesses
    (array, row, is_write), in the kernel's order."""
    epochs = [[]]

    def phase(acc):
        epochs[-1].append(acc)

    def barrier():
        epochs.append([])

    wy0 = y0 - 6
    # prologue: window rows y0-6 .. y0+5, barrier, level-1 rows y0-3 .. y0+2
    phase([("w", r, True) for r in range(wy0, y0 + 6)])
    barrier()
    phase([("w", r, False) for r in range(y0 - 6, y0 + 6)] + [("l", r, True) for r in range(y0 - 3, y0 + 3)])
    if tight and barriers:
        barrier()
    for k in range(nsteps):
        base = y0 + k * rb
        # (a) rows base+6 .. base+rb+5 stored
        phase([("w", r, True) for r in range(base + 6, base + rb + 6)])
        barrier()
        # (b) level-1 rows base+3 .. base+rb+2 from window rows base .. base+rb+5
        phase([("w", r, False) for r in range(base, base + rb + 6)] +
              [("l", r, True) for r in range(base + 3, base + rb + 3)])
        barrier()
        # (c) outputs base .. base+rb-1: window and level-1 rows base-3 .. base+rb+2
        phase([("w", r, False) for r in range(base - 3, base + rb + 3)] +
              [("l", r, False) for r in range(base - 3, base + rb + 3)])
        if tight and barriers:
            barrier()
    return epochs


def hazards(rb, tight, barriers=True, nsteps=12, y0=30):
    wr, lr = ring_depths(rb, tight)
    depth = {"w": wr, "l": lr}
    origin = {"w": y0 - 6, "l": y0 - 3}
    bad = []
    for e, epoch in enumerate(schedule(rb, y0, nsteps, tight, barriers)):
        acc = [a for ph in epoch for a in ph]
        slot_rows = {}
        for arr, row, wr_ in acc:
            s = (arr, (row - origin[arr]) % depth[arr])
            slot_rows.setdefault(s, set()).add((row, wr_))
        for s, rows in slot_rows.items():
            written = {r for r, w in rows if w}
            touched = {r for r, _ in rows}
            if written and len(touched) > 1:
                bad.append((e, s, sorted(rows)))
    return bad


@pytest.mark.parametrize("rb", [4, 8])
def test_kernel_ring_is_hazard_free(rb):
    assert hazards(rb, tight=False) == []


@pytest.mark.parametrize("rb", [4, 6, 8])
def test_live_rows_only_ring_needs_more_barriers(rb):
    # the model detects the overlap of (a) with the previous step's (c) and
    # with the prologue's level-1 rows in a ring of the live rows only
    assert hazards(rb, tight=True, barriers=False) != []
    assert hazards(rb, tight=True, barriers=True) == []


def test_fp64_ring_fits_two_blocks_per_cu():
    # FP64, nk=4 viscous, 4 rows per step: U (4 planes) + T in the window
    # ring, 5 level-1 fields: two blocks (8 waves) per CU's 160 KB
    wr, lr = ring_depths(4, False)
    assert 2 * 8 * (5 * wr * 70 + 5 * lr * 64) <= 160 * 1024
