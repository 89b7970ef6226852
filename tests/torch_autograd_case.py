"""Runs weighted_pagerank through torch autograd and compares the gradients with
the C ABI's backward on the same inputs.  Started as its own process by
tests/test_gpu_wpr.py: torch is imported before the library so that both use
torch's HIP runtime."""
import sys
from pathlib import Path

import torch  # first: the library then binds to the HIP runtime torch loaded

sys.path.insert(0, str(Path(__file__).resolve().parent.parent))
import numpy as np  # noqa: E402

import oracle  # noqa: E402
import suminc_amd  # noqa: E402
from suminc_amd.wpr import weighted_pagerank  # noqa: E402


def main():
    if not torch.cuda.is_available():
        print("SKIP torch sees no GPU")
        return 0
    g = oracle.Graph.rmat(12, 16, seed=4, weights="f32")
    off, col, w = g.csr()
    rng = np.random.default_rng(4)
    p = np.zeros(g.n, np.float32)
    idx = rng.choice(g.n, 64, replace=False)
    p[idx] = rng.dirichlet(np.ones(64)).astype(np.float32)
    c = rng.standard_normal(g.n).astype(np.float32)
    with suminc_amd.Context(pr_precision=suminc_amd.PR_FP32_KAHAN, termcheck_threshold=1e-9) as ctx:
        ctx.upload_csr(off, col, w)
        wt = torch.tensor(w, device="cuda", requires_grad=True)
        pt = torch.tensor(p, device="cuda", requires_grad=True)
        ct = torch.tensor(c, device="cuda")
        r = weighted_pagerank(wt, pt, ctx)
        loss = (r * ct).sum()
        loss.backward()
        gw_t = wt.grad.cpu().numpy()
        gp_t = pt.grad.cpu().numpy()
        r_t = r.detach().cpu().numpy()
        ctx.wpr_set_weights(w)
        r_np, _ = ctx.wpr_forward(p)
        gp, gw, _ = ctx.wpr_backward(c)
    assert np.array_equal(r_t, r_np), "forward differs"
    assert np.array_equal(gw_t, gw), "weight gradient differs"
    assert np.array_equal(gp_t, gp), "personalization gradient differs"
    # the FP64 oracle agrees with the autograd result
    r0, _ = g.wpr_forward(w, p.astype(np.float64), float(np.float32(0.85)))
    _, gp0, gw0 = g.wpr_adjoint(w, r0, c.astype(np.float64), float(np.float32(0.85)))
    assert np.abs(gw_t - gw0).max() <= 1e-4 * np.abs(gw0).max()

    # two forwards with different personalization vectors, then one backward of a
    # loss over both: each forward's gradient must come from its own rank
    p2 = np.zeros(g.n, np.float32)
    idx2 = rng.choice(g.n, 32, replace=False)
    p2[idx2] = rng.dirichlet(np.ones(32)).astype(np.float32)
    c2 = rng.standard_normal(g.n).astype(np.float32)
    with suminc_amd.Context(pr_precision=suminc_amd.PR_FP32_KAHAN, termcheck_threshold=1e-9) as ctx:
        ctx.upload_csr(off, col, w)
        wt = torch.tensor(w, device="cuda", requires_grad=True)
        pa = torch.tensor(p, device="cuda", requires_grad=True)
        pb = torch.tensor(p2, device="cuda", requires_grad=True)
        ra = weighted_pagerank(wt, pa, ctx)
        rb = weighted_pagerank(wt, pb, ctx)
        loss = (ra * torch.tensor(c, device="cuda")).sum() + (rb * torch.tensor(c2, device="cuda")).sum()
        loss.backward()
        gw2 = wt.grad.cpu().numpy().astype(np.float64)
        gpa, gpb = pa.grad.cpu().numpy(), pb.grad.cpu().numpy()
    ra0, _ = g.wpr_forward(w, p.astype(np.float64), float(np.float32(0.85)))
    rb0, _ = g.wpr_forward(w, p2.astype(np.float64), float(np.float32(0.85)))
    _, gpa0, gwa0 = g.wpr_adjoint(w, ra0, c.astype(np.float64), float(np.float32(0.85)))
    _, gpb0, gwb0 = g.wpr_adjoint(w, rb0, c2.astype(np.float64), float(np.float32(0.85)))
    gw20 = gwa0 + gwb0
    assert np.abs(gw2 - gw20).max() <= 1e-4 * np.abs(gw20).max(), "two-forward weight gradient differs"
    assert np.abs(gpa - gpa0).max() <= 1e-4 * np.abs(gpa0).max()
    assert np.abs(gpb - gpb0).max() <= 1e-4 * np.abs(gpb0).max()
    print("OK")
    return 0


if __name__ == "__main__":
    sys.exit(main())
