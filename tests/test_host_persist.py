"""Host-side check of the persistent step's ticket table (no GPU needed).

The persistent step runs one workgroup per ticket; a ticket waits for every item of the
earlier phases of its replica.  stz_persist_probe builds the table the plan builds and
decodes every ticket with the kernel's own decoder: each item of each phase (rate blocks,
group sums, chunk totals, the decision, athermal chunks, local-update row groups, the
bookkeeping) must be covered exactly once with indices in range, and every ticket of an
earlier phase of a replica must precede every ticket of a later one (the waits only look
back, so the step cannot deadlock).  Item counts must equal the kernel sequence's grids.
"""
import ctypes as C

import numpy as np
import pytest

import metallic_glasses_amd as mg

L = C.CDLL(mg.stz.LIB_PATH)
L.stz_persist_probe.argtypes = [C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_longlong)]

ENS64 = [((128, 128, 128), 16), ((96, 96, 96), 16), ((64, 64, 128), 16), ((32, 32, 32), 16)]


def _probe(classes, M=20, prec=mg.MIXED):
    shape = []
    for n, nrep in classes:
        dim = 2 if len(n) == 2 else 3
        nn = list(n) + [1] * (3 - len(n))
        shape += [dim, nn[0], nn[1], nn[2], M, prec, nrep]
    arr = (C.c_int * len(shape))(*shape)
    out = (C.c_longlong * 3)()
    rc = L.stz_persist_probe(len(classes), arr, out)
    return rc, list(out)


def _expected_items(n, nrep, M, prec):
    N = int(np.prod(n))
    nblk = -(-N // 256)
    ngrp = -(-M * nblk // 256)
    dim = len(n)
    g = (n[0], 1, n[1]) if dim == 2 else n
    h2 = g[2] // 2 + 1
    rowv = 1024 if prec == mg.MIXED else 512
    ipw = 1 if g[0] >= rowv else rowv // g[0]
    items = g[1] * h2
    nlb = -(-items // ipw)
    per = nblk + -(-ngrp // 4) + max(-(-ngrp // 256), -(-nblk // 256)) + 1 + -(-nblk // 256) + nlb + 1
    return per * nrep


@pytest.mark.parametrize("classes,prec", [
    (ENS64, mg.MIXED), (ENS64, mg.FP64),
    ([((512, 512, 256), 1)], mg.MIXED),
    ([((256, 256, 256), 1)], mg.FP64),
    ([((512, 512), 1)], mg.FP64),
    ([((16, 16, 16), 3), ((8, 16, 32), 2), ((24, 24, 24), 1)], mg.MIXED),
    ([((7, 5, 3), 1)], mg.FP64),
], ids=["ens64_mixed", "ens64_fp64", "512x512x256_mixed", "256cube_fp64", "2d_512sq", "small", "tiny"])
def test_ticket_table_covers_every_item_once_in_dependency_order(classes, prec):
    rc, (tickets, items, bad) = _probe(classes, prec=prec)
    assert rc == 0
    assert bad == 0
    assert items == sum(_expected_items(n, k, 20, prec) for n, k in classes)
    assert 0 < tickets <= items


def test_ticket_table_refuses_too_many_classes():
    rc, _ = _probe([((8, 8, 8), 1)] * 9)
    assert rc == -3      # STZ_E_ARG
