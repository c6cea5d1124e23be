"""swpc_lab_grad_noalloc, the entry point that allocates no device memory,
against swpc_lab_grad (the same kernels with the trainer's buffers),
called through the C ABI: equal loss and gradient, bit for bit, on two
calls in a row whose second has fewer records (the fixed scratch keeps the
first call's contents), against the FP64 oracle of the gradient, and a
clear error for more records than the scratch holds."""
import ctypes as C

import numpy as np
import pytest

from openswpc_amd import lab
from openswpc_amd._lib import SwpcError, lib

pytestmark = pytest.mark.gpu


def _noalloc(P, nout, rec):
    rec = np.ascontiguousarray(rec, np.float32)
    g = np.zeros(lab.NP, np.float32)
    loss = C.c_double(0)
    rc = lib().swpc_lab_grad_noalloc(P.ctypes.data_as(C.POINTER(C.c_float)), nout, len(rec),
                                     rec.ctypes.data_as(C.POINTER(C.c_float)), C.byref(loss),
                                     g.ctypes.data_as(C.POINTER(C.c_float)))
    return rc, loss.value, g


def _records(rng, n):
    rec = np.zeros((n, lab.REC), np.float32)
    rec[:, :61] = rng.standard_normal((n, 61))
    rec[:, 61:66] = rng.standard_normal((n, 5))
    return rec


def test_noalloc_grad_equals_trainer_grad():
    from oracle import lab as olab
    rng = np.random.default_rng(11)
    P = (rng.standard_normal(lab.NP) * 0.2).astype(np.float32)
    limit = 16384
    tr = lab.Trainer(1024)
    try:
        for n, nout in ((limit, 5), (1000, 4)):    # the second call is smaller
            print(f"noalloc gradient: {n} records, nout {nout}", flush=True)
            rec = _records(rng, n)
            rc, loss, g = _noalloc(P, nout, rec)
            assert rc == 0, lib().swpc_last_error().decode()
            lt, gt = tr.grad(P, nout, rec)
            assert loss == lt
            assert np.array_equal(g, gt)
            lref, gref = olab.loss_grad(P.astype(np.float64), rec[:, :61], rec[:, 61:66], nout)
            assert abs(loss - lref) <= 2e-6 * lref
            assert np.linalg.norm(g - gref) <= 1e-5 * np.linalg.norm(gref)
    finally:
        tr.close()


def test_noalloc_grad_rejects_more_records_than_its_scratch():
    rng = np.random.default_rng(12)
    P = (rng.standard_normal(lab.NP) * 0.2).astype(np.float32)
    rec = _records(rng, 16384 + 128)
    rc, _, _ = _noalloc(P, 5, rec)
    assert rc != 0
    assert "exceed" in lib().swpc_last_error().decode()
    with pytest.raises(SwpcError):
        from openswpc_amd._lib import check
        check(rc, "swpc_lab_grad_noalloc")
