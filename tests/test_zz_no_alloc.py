"""GPU: mskg_frame_device, the frame compaction that allocates nothing
(its block counts and offsets live in the library's fixed per-GPU scratch),
called through the C ABI with device buffers.  Its records and totals equal
the allocating mskg_frame's, on two calls in a row -- the second on a
smaller lattice, so the scratch holds the first call's larger contents --
and a list cut at a capacity below its length keeps its first records and
reports the full total."""
import ctypes as C

import numpy as np
import pytest
import torch

from mulskips_amd import _lib
from mulskips_amd.engine import Engine
from tests.test_frames import frame_case

pytestmark = pytest.mark.gpu


def device_frame(eng, caps=None):
    m, v, ne, nat, ncov = eng.frame()
    want = (m, v, ne)
    caps = caps or [len(a) for a in want]
    outs = [torch.full((max(c, 1), 4), -7, dtype=torch.int32, device="cuda") for c in caps]
    counts = torch.zeros(5, dtype=torch.int64, device="cuda")
    torch.cuda.synchronize()
    L = _lib.lib()
    rc = L.mskg_frame_device(eng.h, C.c_void_p(outs[0].data_ptr()), caps[0], C.c_void_p(outs[1].data_ptr()), caps[1],
                             C.c_void_p(outs[2].data_ptr()), caps[2], C.c_void_p(counts.data_ptr()))
    assert rc == 0, L.mskg_last_error(eng.h)
    return want, (nat, ncov), [o.cpu().numpy() for o in outs], counts.cpu().numpy(), caps


def check(eng):
    want, hdr, got, counts, caps = device_frame(eng)
    assert list(counts) == [len(want[0]), len(want[1]), len(want[2]), hdr[0], hdr[1]]
    for w, g, c in zip(want, got, caps):
        np.testing.assert_array_equal(g[:c], w)
    return len(want[0])


def test_frame_without_allocation(tmp_path):
    o, d, box, lat, void = frame_case(tmp_path, "cube")
    from tests.helpers import oracle_state
    _, mc, bulk, _ = oracle_state(o)
    big = Engine(box, d.tables())
    big.load_state(lat, mc, bulk, void)
    n_big = check(big)
    o2, d2, box2, lat2, void2 = frame_case(tmp_path, "surface")
    _, mc2, bulk2, _ = oracle_state(o2)
    small = Engine(box2, d2.tables())
    small.load_state(lat2, mc2, bulk2, void2)
    assert small.box[0] * small.box[1] * small.box[2] < big.box[0] * big.box[1] * big.box[2]
    n_small = check(small)
    assert n_big > 0 and n_small > 0
    # a list cut at its capacity: the first records, the full total
    want, hdr, got, counts, caps = device_frame(small, caps=[n_small // 2, 0, 0])
    assert counts[0] == n_small
    np.testing.assert_array_equal(got[0][:n_small // 2], want[0][:n_small // 2])
    check(big)  # the larger lattice again, after the smaller one
    big.close()
    small.close()
    o.close()
    o2.close()
