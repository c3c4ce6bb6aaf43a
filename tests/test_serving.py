"""Host-side admission policy of the reward server (no GPU needed)."""

import pytest

from panoptic_evaluator.serving import Admission, Ticket


def ticket(nbytes):
    return Ticket(request=None, images=1, footprint=nbytes, arrival=0.0)


def test_fifo_within_budget():
    adm = Admission(100)
    a, b, c = ticket(40), ticket(50), ticket(30)
    for t in (a, b, c):
        adm.offer(t)
    assert adm.admit() == [a, b]          # c would exceed 100
    assert adm.in_use == 90 and list(adm.waiting) == [c]
    adm.release(a)
    assert adm.admit() == [c]             # memory freed -> admitted
    assert adm.in_use == 80


def test_head_of_line_blocks_later_small_requests():
    adm = Admission(100)
    big, small = ticket(80), ticket(10)
    adm.offer(ticket(30))
    first = adm.admit()
    adm.offer(big)
    adm.offer(small)
    assert adm.admit() == []              # FIFO: small waits behind big
    adm.release(first[0])
    assert adm.admit() == [big, small]


def test_oversized_request_runs_alone():
    adm = Admission(100)
    huge, next_one = ticket(250), ticket(10)
    adm.offer(huge)
    adm.offer(next_one)
    assert adm.admit() == [huge]          # nothing else in use: admitted alone
    assert adm.admit() == []
    adm.release(huge)
    assert adm.admit() == [next_one] and adm.in_use == 10


def test_budget_must_be_positive():
    with pytest.raises(ValueError):
        Admission(0)
