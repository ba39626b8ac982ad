"""The in-place correlation's block order, modelled on the host: every wait ends.

k_fxcorr<..., ORD = true> (fxcorr_kernels.hpp): each workgroup takes a ticket from a device
counter when it starts, maps it to an item with xcd_map (fft_kernels.hpp), raises its item's read
flag without waiting for anyone, then waits for the flag of item - 1 before storing.  The model
starts workgroups in ticket order whenever one of `slots` resident slots is free (the hardware may
keep fewer workgroups resident than the grid holds) and checks that the launch always drains.
"""
import pytest


def xcd_map(t, g):
    """fft_kernels.hpp xcd_map: the XCD-aware spread of an index over a grid of g workgroups."""
    return t if g & 7 else (t & 7) * (g >> 3) + (t >> 3)


def drains(g, slots):
    flag = [False] * g
    running = []
    nxt = done = 0
    while done < g:
        while len(running) < slots and nxt < g:
            w = xcd_map(nxt, g)
            nxt += 1
            flag[w] = True  # loads returned, flag raised: no wait before it
            running.append(w)
        ready = [w for w in running if w == 0 or flag[w - 1]]
        if not ready:
            return False
        for w in ready:
            running.remove(w)
            done += 1
    return True


@pytest.mark.parametrize("g", [8, 16, 24, 64, 1368, 8000])
@pytest.mark.parametrize("slots", [8, 9, 16, 512])
def test_every_wait_ends(g, slots):
    assert drains(g, slots)


@pytest.mark.parametrize("g", [8, 16, 1368, 75264])
def test_items_are_a_permutation(g):
    assert sorted(xcd_map(t, g) for t in range(g)) == list(range(g))


def test_backward_waits_are_the_range_starts():
    """Only the first item of each XCD range waits for an item with a later ticket (7 of them)."""
    g = 1368
    ticket_of = {xcd_map(t, g): t for t in range(g)}
    late = [w for w in range(1, g) if ticket_of[w - 1] > ticket_of[w]]
    assert late == [x * (g // 8) for x in range(1, 8)]
