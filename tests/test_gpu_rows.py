"""GPU tests of the device-resident row seam (rtdl_amd.rows): paged row
streams over batch rows, and DLPack hand-off of rows to torch without a
host copy."""
from __future__ import annotations

import os
import subprocess
import sys

import numpy as np
import pytest

from rtdl_amd import batch as B
from rtdl_amd import rows as R
from rtdl_amd.device import DeviceBuffer
from tests import cases

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _say(msg):
    print(msg, flush=True)


def _batch(rng):
    pts = cases.points(rng, 20000, 3, kind="mixture")
    s3 = np.stack([pts["x"], pts["y"], pts["z"]], 1).astype(np.float32)
    scene = B.prepare_point_instances(DeviceBuffer.from_numpy(s3), [0, len(s3)])
    q = DeviceBuffer.from_numpy(rng.random((5000, 3)).astype(np.float32))
    jobs = [B.BatchJob(scene, B.JOB_FIXED_RADIUS, q, radius=0.03, k=32), B.BatchJob(scene, B.JOB_KNN, q, k=8)]
    res, _ = B.run_batch(jobs)
    return scene, jobs, res


def test_row_pages_stream_a_job_in_row_order(gpu_ready):
    _say("case: paged row stream over batch rows")
    scene, jobs, res = _batch(np.random.default_rng(91))
    for job, r in zip(jobs, res):
        whole = B.job_rows(job, r)
        pages = list(R.row_pages(job, r, 777, stream_id=f"job{job.kind}"))
        assert sum(p.valid_count for p in pages) == len(whole) == r["rows"]
        assert [p.row_offset for p in pages] == list(range(0, len(whole), 777))
        assert pages[-1].next_token is None and all(p.next_token for p in pages[:-1])
        assert not any(p.overflowed for p in pages)
        got = np.concatenate([p.rows for p in pages])
        for name in whole.dtype.names:
            assert np.array_equal(got[name], whole[name])
        # resuming from a page's token gives the rest of the stream
        rest = list(R.row_pages(job, r, 777, stream_id=f"job{job.kind}", start_token=pages[1].next_token))
        assert [p.row_offset for p in rest] == [p.row_offset for p in pages[2:]]
        with pytest.raises(ValueError, match="does not continue"):
            next(R.row_pages(job, r, 777, stream_id="other", start_token=pages[1].next_token))


_TORCH_SCRIPT = r"""
import sys
try:
    import torch
except Exception as exc:
    print("no torch:", exc); sys.exit(3)
if not torch.cuda.is_available():
    print("torch sees no GPU"); sys.exit(3)
torch.zeros(1, device="cuda")  # torch's HIP runtime first; the library shares it
sys.path.insert(0, ROOT)
import numpy as np
from rtdl_amd import batch as B, rows as R
from rtdl_amd.device import DeviceBuffer
rng = np.random.default_rng(92)
s3 = rng.random((30000, 3)).astype(np.float32)
scene = B.prepare_point_instances(DeviceBuffer.from_numpy(s3), [0, len(s3)])
q = DeviceBuffer.from_numpy(rng.random((4000, 3)).astype(np.float32))
job = B.BatchJob(scene, B.JOB_KNN, q, k=8)
(res,), _ = B.run_batch([job])
view = R.device_rows(job, res)
t = torch.from_dlpack(view)
assert t.is_cuda and t.dtype == torch.uint8 and tuple(t.shape) == (res["rows"], 24), (t.device, t.dtype, t.shape)
assert t.data_ptr() == view.data_ptr(), "not zero-copy"
host = view.download()
assert np.array_equal(t.cpu().numpy().reshape(-1), host)
ids = t.view(torch.int32)[:, 1].cpu().numpy()  # neighbor_id: the row's second u32
assert np.array_equal(ids, B.job_rows(job, res)["neighbor_id"].astype(np.int32))
del t
print("ok", res["rows"])
"""


def test_dlpack_rows_reach_torch_in_place(gpu_ready):
    """In a fresh process: torch wraps a batch job's device rows through
    DLPack at the same address (no copy), with the same bytes as a host
    download.  torch is initialised first so that the library runs on the
    HIP runtime torch loads (one runtime per process)."""
    _say("case: DLPack rows to torch.from_dlpack (fresh process)")
    script = "ROOT = " + repr(ROOT) + "\n" + _TORCH_SCRIPT
    proc = subprocess.run([sys.executable, "-c", script], capture_output=True, text=True, timeout=240)
    if proc.returncode == 3:
        pytest.skip(proc.stdout.strip())
    assert proc.returncode == 0, proc.stdout[-2000:] + proc.stderr[-3000:]
    assert proc.stdout.strip().startswith("ok")


def test_pinned_per_call_rows_survive_later_calls_and_are_reused(gpu_ready):
    """Per-call row sets of 1 MiB or more come back in pinned buffers that
    free_rows keeps for later calls: rows the caller still holds must not
    change when later calls run (and reuse released buffers), and every
    call's rows equal the oracle's."""
    import gc

    from oracle import native as oracle
    from rtdl_amd import mi355x_runtime as m

    _say("case: pinned per-call rows held across later calls, released buffers reused")
    rng = np.random.default_rng(123)
    s = rng.random((30000, 3)).astype(np.float32).astype(np.float64)
    search = {"id": np.arange(len(s), dtype=np.uint32), "x": s[:, 0], "y": s[:, 1], "z": s[:, 2]}

    def queries(n, seed):
        q = np.random.default_rng(seed).random((n, 3)).astype(np.float32).astype(np.float64)
        return {"id": np.arange(n, dtype=np.uint32), "x": q[:, 0], "y": q[:, 1], "z": q[:, 2]}

    qa, qb = queries(6000, 1), queries(5000, 2)  # 6000 x 16 x 24 B = 2.3 MB of rows
    want_a = oracle.neighbors("knn", qa, search, k=16)
    want_b = oracle.neighbors("knn", qb, search, k=16)

    def check(got, want):
        assert len(got) == len(want)
        assert np.array_equal(got["query_id"], want["qid"])
        assert np.array_equal(got["neighbor_id"], want["nid"])
        assert np.array_equal(got["distance"], want["dist"])

    a1 = m.knn_rows_mi355x(qa, search, k=16)
    assert a1.nbytes >= 1 << 20
    keep = a1.copy()
    b1 = m.knn_rows_mi355x(qb, search, k=16)  # a1 is still held: b1 gets another buffer
    check(a1, want_a)
    check(b1, want_b)
    assert np.array_equal(a1, keep)
    del b1
    gc.collect()
    for _ in range(3):  # released buffers are handed out again
        b2 = m.knn_rows_mi355x(qb, search, k=16)
        check(b2, want_b)
        assert np.array_equal(a1, keep)
        del b2
        gc.collect()
    del a1
    gc.collect()
    a2 = m.knn_rows_mi355x(qa, search, k=16)
    check(a2, want_a)
