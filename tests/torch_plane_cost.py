"""PyTorch-ROCm harness of the plane-cost backward kernel: a torch.autograd.Function
whose forward and backward call hpm_cost_grad_device on the tensors' device
memory.  Run as a script (test_gpu_grad.py starts it in a fresh process):
torch's own HIP runtime must initialise the GPU before libhpm_mi355x.so does.
Prints one JSON line with the checks' results; exit status 0 when all pass."""
import json
import os
import sys

import numpy as np
import torch

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
sys.path.insert(0, os.path.join(os.path.dirname(HERE), "hpm-mvs_plusplus_amd"))
sys.path.insert(0, os.path.join(os.path.dirname(HERE), "oracle"))


def main():
    if not torch.cuda.is_available():
        print(json.dumps({"skipped": "torch without a GPU"}))
        return 0
    dev = torch.device("cuda:0")
    torch.ones(1, device=dev).sum().item()  # torch's runtime first
    from test_gpu_grad import _problem, _ctx
    sc, ids, cams, imgs, params, weights, gt = _problem(noise_seed=2)
    ctx = _ctx(cams, imgs)

    class PlaneCost(torch.autograd.Function):
        """Per-pixel aggregated NCC cost of (depth, theta, phi); the backward
        is the hand-written HIP kernel's gradient."""

        @staticmethod
        def forward(fctx, p4, wts):
            cost = torch.empty(p4.shape[:2], dtype=torch.float32, device=p4.device)
            grad = torch.empty(p4.shape[:2] + (3,), dtype=torch.float32, device=p4.device)
            torch.cuda.synchronize()
            ctx.cost_grad_device(p4.data_ptr(), wts.data_ptr(), cost.data_ptr(), grad.data_ptr())
            ctx.synchronize()
            fctx.save_for_backward(grad)
            return cost

        @staticmethod
        def backward(fctx, gout):
            (grad,) = fctx.saved_tensors
            g4 = torch.zeros(grad.shape[:2] + (4,), dtype=torch.float32, device=grad.device)
            g4[..., :3] = gout[..., None] * grad
            return g4, None

    p4 = torch.zeros(params.shape[:2] + (4,), dtype=torch.float32)
    p4[..., :3] = torch.from_numpy(params)
    p4 = p4.to(dev).contiguous().requires_grad_(True)
    wts = torch.from_numpy(weights).to(dev).contiguous()
    cost = PlaneCost.apply(p4, wts)
    cost.sum().backward()
    ref_cost, ref_grad = ctx.cost_grad(params, weights)
    res = {"cost_equal": bool(np.array_equal(cost.detach().cpu().numpy(), ref_cost)),
           "grad_equal": bool(np.array_equal(p4.grad[..., :3].cpu().numpy(), ref_grad))}
    # a few steps of plain gradient descent on the depth through autograd lower the cost
    live = ((cost > 0.02) & (cost < 1.5)).detach()
    x = p4.detach().clone()
    start = float(cost.detach()[live].mean())
    for _ in range(3):
        x.requires_grad_(True)
        c = PlaneCost.apply(x, wts)
        c.sum().backward()
        with torch.no_grad():
            g = x.grad[..., 0]
            x = x.detach()
            x[..., 0] -= 2e-4 * torch.sign(g) * x[..., 0] * live
    end = float(PlaneCost.apply(x, wts).detach()[live].mean())
    res.update({"mean_cost_start": start, "mean_cost_after_3_steps": end, "descends": end < start})
    print(json.dumps(res), flush=True)
    return 0 if res["cost_equal"] and res["grad_equal"] and res["descends"] else 1


if __name__ == "__main__":
    sys.exit(main())
