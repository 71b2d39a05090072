"""Failure handling of the serving loop, forced from the host (no kernel is
launched): admission rejection against a tiny HBM budget, dimension errors,
cancellation of a queued request, wait time-outs, and requests failing
cleanly when a worker has no usable device.  Every server here is created on
device index 99, which does not exist on any machine, so its workers never
obtain a context and no request ever reaches a kernel (the descriptors are
fake)."""
import ctypes as C

import pytest

import mex_linear_algebra_amd as mela


NO_DEVICE = 99
FAILED = ("DEVICE", "INVALID")


def fake(rows, cols):
    # a descriptor only: the host-side paths exercised here never touch data
    return mela.BitMatrix(C.c_void_p(0x1000), rows, cols, (cols + 63) // 64)


def test_admission_rejects_oversize_request():
    s = mela.Server(device=NO_DEVICE, hbm_budget=1 << 20, workers=1)
    with pytest.raises(mela.MelaError) as ei:
        s.submit(mela.OP_NULL, fake(4096, 4096))
    assert ei.value.status == "ADMISSION"
    assert s.stats()["rejected"] == 1
    s.close()


def test_matmul_dimension_error():
    s = mela.Server(device=NO_DEVICE, hbm_budget=1 << 30, workers=1)
    with pytest.raises(mela.MelaError, match="Inner dimensions must agree."):
        s.submit(mela.OP_MATMUL, fake(10, 20), fake(21, 5))
    s.close()


def test_queue_cancel_and_timeout():
    # with no device every request fails, which frees its memory again
    s = mela.Server(device=NO_DEVICE, hbm_budget=1 << 30, workers=1)
    tickets = [s.submit(mela.OP_RANK, fake(64, 64)) for _ in range(3)]
    for t in tickets:
        try:
            r = s.wait(t, 30.0)
            ok = r.status == 0
        except mela.MelaError as e:
            ok = e.status in FAILED + ("CANCELLED",)
        assert ok
    st = s.stats()
    assert st["completed"] == 0
    assert st["submitted"] == 3
    assert st["completed"] + st["failed"] + st["cancelled"] == 3
    s.close()


def test_cancel_pending_before_admission():
    # a budget so small that a second request must queue behind the first
    # (a 64 x 64 rank job reserves about 82 KiB: working and pivot rows,
    # metadata and the active list)
    one = 120 << 10
    s = mela.Server(device=NO_DEVICE, hbm_budget=one, workers=1)
    t1 = s.submit(mela.OP_RANK, fake(64, 64))
    t2 = s.submit(mela.OP_RANK, fake(64, 64))
    s.cancel(t2)
    with pytest.raises(mela.MelaError) as ei:
        s.wait(t2, 5.0)
    # t1 fails at once for want of a device and frees its memory, so t2 can
    # be admitted (and fail the same way) before the cancel lands
    assert ei.value.status in ("CANCELLED",) + FAILED
    st = s.stats()
    assert st["cancelled"] == (1 if ei.value.status == "CANCELLED" else 0)
    with pytest.raises(mela.MelaError) as ei:
        s.wait(t1, 30.0)
    assert ei.value.status in FAILED
    s.close()


def test_wait_timeout_on_blocked_request():
    one = 120 << 10
    s = mela.Server(device=NO_DEVICE, hbm_budget=one, workers=1)
    t1 = s.submit(mela.OP_RANK, fake(64, 64))
    with pytest.raises(mela.MelaError):
        s.wait(t1, 30.0)  # fails (no device) and frees its memory
    t2 = s.submit(mela.OP_RANK, fake(64, 64))
    with pytest.raises(mela.MelaError) as ei:
        s.wait(t2, 30.0)
    assert ei.value.status in FAILED
    with pytest.raises(mela.MelaError) as ei:
        s.wait(123456, 0.01)
    assert "unknown ticket" in str(ei.value)
    s.close()


def test_wait_zero_timeout():
    s = mela.Server(device=NO_DEVICE, hbm_budget=1 << 30, workers=1)
    t = s.submit(mela.OP_NULL, fake(128, 256))
    with pytest.raises(mela.MelaError) as ei:
        s.wait(t, 0.0)
    # not finished yet (TIMEOUT) or already failed for want of a device
    assert ei.value.status in ("TIMEOUT",) + FAILED
    s.close()


def test_routing_by_batch_member_cost():
    # the scheduler sends a rank / null request to the device-resident loop
    # only while one workgroup finishes it about as fast as the single-matrix
    # path; the device never exists here, so every request fails, but the
    # routing is counted all the same
    s = mela.Server(device=NO_DEVICE, hbm_budget=1 << 34, workers=1)
    cases = [
        (mela.OP_RANK, 256, 256, True),     # small: batched
        (mela.OP_RANK, 128, 8192, True),    # short and wide rank: batched
        (mela.OP_RANK, 4096, 8192, False),  # one workgroup would take ~28 ms
        (mela.OP_NULL, 128, 8192, False),   # 8192 x 8064 null basis from one workgroup
        (mela.OP_RANK, 5000, 100, False),   # beyond the batch kernel's 4096 rows
    ]
    for op, m, n, _ in cases:
        t = s.submit(op, fake(m, n))
        with pytest.raises(mela.MelaError):
            s.wait(t, 30.0)
    st = s.stats()
    assert st["batched_requests"] == sum(1 for c in cases if c[3])
    assert st["failed"] == len(cases)
    s.close()
