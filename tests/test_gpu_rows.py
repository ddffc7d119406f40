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
