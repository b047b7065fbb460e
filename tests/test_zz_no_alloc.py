"""The no-allocation 2D critical-point judge (cpsz_critical_points_2d_noalloc_dev)
against the allocating one (cpsz_critical_points_2d_dev), called through the C
ABI twice in a row, the second call on a smaller field: the scratch array keeps
the first call's contents, so a counter or list left uncleared would show."""
import ctypes as C

import numpy as np
import pytest
import torch

import cpsz_amd as cz
from cpsz_amd.synth import field_2d

pytestmark = pytest.mark.gpu


def _noalloc(dU, dV, r1, r2, cap):
    ids = np.zeros(cap, np.int64)
    types = np.zeros(cap, np.int32)
    cnt = C.c_size_t()
    rc = cz.lib().cpsz_critical_points_2d_noalloc_dev(0, None, dU.data_ptr(), dV.data_ptr(), r1, r2,
                                                      ids.ctypes.data, types.ctypes.data, cap, C.byref(cnt))
    return rc, cnt.value, ids, types


@pytest.mark.parametrize("shapes", [((700, 900), (130, 170)), ((300, 400), (40, 41))])
def test_noalloc_matches_allocating_judge_twice(shapes):
    for k, (r1, r2) in enumerate(shapes):
        U, V = field_2d(r1, r2, 300 + k)
        dU, dV = torch.from_numpy(U).cuda(), torch.from_numpy(V).cuda()
        ref = cz.critical_points_2d_dev(dU.data_ptr(), dV.data_ptr(), r1, r2)
        rc, n, ids, types = _noalloc(dU, dV, r1, r2, 1 << 16)
        print(f"[noalloc] {r1}x{r2}: {n} critical points (allocating judge: {len(ref)})", flush=True)
        assert rc == 0, cz.lib().cpsz_last_error()
        assert n == len(ref) and n > 0
        assert np.all(np.diff(ids[:n]) > 0)  # sorted by id
        assert dict(zip(ids[:n].tolist(), types[:n].tolist())) == ref


def test_noalloc_rejects_short_output_arrays():
    U, V = field_2d(300, 400, 301)
    dU, dV = torch.from_numpy(U).cuda(), torch.from_numpy(V).cuda()
    ref = cz.critical_points_2d_dev(dU.data_ptr(), dV.data_ptr(), 300, 400)
    assert len(ref) > 1
    rc, n, _, _ = _noalloc(dU, dV, 300, 400, len(ref) - 1)
    assert rc == -1 and n == len(ref)
    rc, n, ids, types = _noalloc(dU, dV, 300, 400, len(ref))
    assert rc == 0 and dict(zip(ids[:n].tolist(), types[:n].tolist())) == ref
