"""The allocation-free image statistics entry point (cmtip_image_stats_noalloc) against
cmtip_image_stats, through the C ABI, on two calls in a row (the second smaller), and
its size limit.  Runs last: a failure here must not hide other tests."""

import pytest
import torch

from cmtip_amd._lib import check, lib, ptr, stream
from tests.gpu_util import dev, say

pytestmark = pytest.mark.gpu

def _stats(fn_name, X, rows, P, with_workspace):
    d = X.device
    c = torch.empty(P, dtype=torch.float32, device=d)
    mom = torch.empty(2, dtype=torch.float64, device=d)
    if with_workspace:
        colsum = torch.empty(2 * P, dtype=torch.float64, device=d)
        check(lib().cmtip_image_stats(ptr(X), rows, P, ptr(colsum), ptr(c), ptr(mom), stream()), fn_name)
    else:
        check(lib().cmtip_image_stats_noalloc(ptr(X), rows, P, ptr(c), ptr(mom), stream()), fn_name)
    torch.cuda.synchronize()
    return c.cpu(), mom.cpu()

def test_noalloc_matches_workspace_form():
    d = dev()
    g = torch.Generator(device=d).manual_seed(11)
    # a large call, then a smaller one: the second must not see the first's column sums
    for rows, P in ((700, 16384), (33, 1000)):
        X = torch.poisson(torch.rand((rows, P), device=d, generator=g) * 20, generator=g)
        say(f"image_stats_noalloc rows={rows} P={P}")
        c0, m0 = _stats("cmtip_image_stats", X, rows, P, True)
        c1, m1 = _stats("cmtip_image_stats_noalloc", X, rows, P, False)
        assert torch.equal(c0, c1)
        assert torch.equal(m0, m1)
        # and the plain definition (column means, total sum and sum of squares)
        Xd = X.double().cpu()
        assert torch.allclose(c1.double(), Xd.mean(0), rtol=1e-6)
        assert torch.allclose(m1, torch.stack([Xd.sum(), (Xd * Xd).sum()]), rtol=1e-12)

def test_noalloc_rejects_oversized_rows():
    d = dev()
    X = torch.zeros((1, 8), device=d)
    c = torch.empty(8, device=d)
    mom = torch.empty(2, dtype=torch.float64, device=d)
    say("image_stats_noalloc size limit")
    rc = lib().cmtip_image_stats_noalloc(ptr(X), 1, 262145, ptr(c), ptr(mom), stream())
    assert rc == 1  # CMTIP_ERR_ARG, before any launch
    assert b"exceeds" in lib().cmtip_last_error()
