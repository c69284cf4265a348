"""Argument checks of the C ABI on a live case.

Every rejected call returns a nonzero status before it launches anything and
leaves its message for ih_last_error.  The reference's launchers print a CUDA
error and carry on (src/grad_host.f90:7-35); the replacement reports instead.
After the rejected calls the same case still computes the operator bit for bit
against the oracle, so no check leaves the case half-changed.  The grid checks of
ih_case_create run before it looks for a device, so those tests need no GPU.
"""
import re

import pytest

import impact_hip as ih
from test_gpu_parity import pair, rnd, same

SPEC = ((17, 13, 11), (1, 1, 1, 1, 1, 1), (2.0, 1.0, 1.0), (0.0, 1.5, 1.5))


def rejects(fn, *args, match):
    with pytest.raises(ih.IHError, match=re.escape(match)):
        fn(*args)
    assert match in ih.last_error()


@pytest.mark.gpu
def test_rejected_calls_leave_the_case_usable():
    oc, gc = pair(SPEC)
    lo, hi = list(oc.Sp), list(oc.Np)
    # BiCGstab runs on the pressure range only, with a known preconditioner
    rejects(gc.bicgstab, 1e-8, 100, [lo[0] + 1, lo[1], lo[2]], hi, match="pressure range")
    rejects(gc.bicgstab_precond, 1e-8, 100, lo, hi, 2, match="preconditioner 2")
    rejects(gc.bicgstab_mixed, 1e-8, 100, lo, hi, 0, match="at least one sweep")
    # multigrid tables: level range and table names; the cycle's configuration
    G = len(gc.mg_levels())
    rejects(gc.mg_coeff, 0, "cdg1", match=f"level 0 out of 1..{G}")
    rejects(gc.mg_coeff, G + 1, "cdg1", match=f"level {G + 1} out of 1..{G}")
    rejects(gc.mg_coeff, 1, "cdg9", match="unknown multigrid table")
    rejects(gc.mg_coeff, 1, "cQ1", match="unknown multigrid table")
    rejects(gc.mg_configure, 0, 4, 4, 4, match="n_grids_limit")
    # a box reaching past the ghost layers, and an unknown benchmark kernel
    rejects(gc.get_box, ih.PRE, [-4, 1, 1], [5, 5, 5], match="outside the field's index range")
    rejects(gc.bench_kernel, "no_such_kernel", 1, match="unknown kernel 'no_such_kernel'")
    # the case still computes: one operator application, bit-exact
    p = rnd(oc.shape, 60)
    lap = rnd(oc.shape, 61)
    gc.upload(ih.TMP1, p)
    gc.upload(ih.TMP2, lap)
    oc.product_div_grad(p, lap, oc.zeros())
    gc.product_div_grad(ih.TMP1, ih.TMP2)
    same(gc.download(ih.TMP2), lap, "div grad after the rejected calls")
    gc.close()


@pytest.mark.parametrize("N,bc,match", [
    ((7, 13, 11), (1, 1, 1, 1, 1, 1), "N[0] = 7 out of range"),
    ((17, 13, 11), (-1, 1, 1, 1, 1, 1), "periodic must be set on both sides"),
    ((17, 13, 11), (1, 1, 0, 0, 1, 1), "only -1 (periodic) and 1 (wall)"),
], ids=["too-small", "one-sided-periodic", "neighbour-block"])
def test_case_create_rejects_bad_grids(N, bc, match):
    with pytest.raises(ih.IHError, match=re.escape(match)):
        ih.Case(N, bc, (1.0, 1.0, 1.0), (0.0, 0.0, 0.0))


@pytest.mark.gpu
def test_case_create_rejects_a_missing_device():
    with pytest.raises(ih.IHError, match=re.escape("device 99 out of range")):
        ih.Case((17, 13, 11), (1, 1, 1, 1, 1, 1), device=99)
