"""The schedule of the AV sweeps run as a pipeline over z slabs (the x-y
walk and the z walk on two streams; miso_amd/csrc/tail_plan.hpp), checked
on the CPU by tests/cpp/test_tail_plan.cpp over a sweep of grid heights and
chunkings: every x-y slab and z slab that touch the same plane of the same
buffer, one of them writing, are ordered by the plan's waits (x-y first, as
the sequential walks), and the slabs write every plane exactly once."""
import ctypes
import os

import pytest

LIB = os.path.join(os.path.dirname(__file__), "cpp", "_build", "libtail_plan_check.so")


@pytest.mark.skipif(not os.path.exists(LIB), reason="tests/cpp not built (run __graft_entry__.build())")
def test_tail_plan_orders_every_conflict():
    lib = ctypes.CDLL(LIB)
    msg = ctypes.create_string_buffer(512)
    plans = ctypes.c_int(0)
    rc = lib.tail_plan_check(msg, 512, ctypes.byref(plans))
    assert rc == 0, msg.value.decode()
    assert plans.value > 300, plans.value  # most shapes of the sweep have a pipeline
