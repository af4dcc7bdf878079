"""torch.autograd view of the differentiable window (bn_window_* of include/buryn_amd.h).

``WindowLoss.apply(window, w_ee, w_ei, w_ie, w_ii, g_leak, tau_decay_e)`` returns the window
loss as a 0-d float64 tensor; its backward calls the library's hand-written adjoint (checkpoint
recompute, surrogate spike derivative) and returns the gradients of the four weight matrices
(float32, same device as the inputs), g_leak and tau_decay_e.  The weights may live on the GPU
(copied device to device, no host round trip) or on the host.  PyTorch is only the gradient
interface here: every step, loss and adjoint runs in the HIP library.
"""
import os

import torch

from . import engine as _engine
from .engine import Window


def _check_one_hip_runtime() -> None:
    # The PyTorch wheel bundles its own HIP runtime.  Loaded after torch, the engine library
    # binds to that runtime (same soname); loaded before, it brings a second runtime into the
    # process and PyTorch can no longer open the GPU.
    if _engine._lib is None:
        return
    torch_lib = os.path.realpath(os.path.join(os.path.dirname(torch.__file__), "lib"))
    with open("/proc/self/maps") as f:
        hip = {line.split()[-1] for line in f if "libamdhip64" in line}
    if any(not os.path.realpath(p).startswith(torch_lib) for p in hip):
        raise ImportError("the engine library was loaded before torch: import torch (or this module) "
                          "before the first engine call to share one HIP runtime with PyTorch")


_check_one_hip_runtime()


class WindowLoss(torch.autograd.Function):
    @staticmethod
    def forward(ctx, window: Window, w_ee, w_ei, w_ie, w_ii, g_leak, tau_decay_e):
        ws = [w.detach().to(torch.float32).contiguous() for w in (w_ee, w_ei, w_ie, w_ii)]
        for w, shape in zip(ws, window.shapes):
            if tuple(w.shape) != shape:
                raise ValueError(f"weight shape {tuple(w.shape)} != {shape}")
        if ws[0].is_cuda:
            torch.cuda.current_stream(ws[0].device).synchronize()
            window.set_params(ws, float(g_leak), float(tau_decay_e))
        else:
            window.set_params([w.numpy() for w in ws], float(g_leak), float(tau_decay_e))
        out = window.forward(outputs=False)
        window._generation = getattr(window, "_generation", 0) + 1
        ctx.window = window
        ctx.generation = window._generation
        ctx.device = w_ee.device
        ctx.scalar_like = (g_leak, tau_decay_e)
        return torch.tensor(out["loss"], dtype=torch.float64, device=w_ee.device)

    @staticmethod
    def backward(ctx, grad_loss):
        window = ctx.window
        if window._generation != ctx.generation:
            raise RuntimeError("the window ran another forward since this loss was computed; "
                               "call backward before the next forward")
        if ctx.device.type == "cuda":
            dW = [torch.empty(s, dtype=torch.float32, device=ctx.device) for s in window.shapes]
            r = window.backward(dW)
        else:
            r = window.backward()
            dW = [torch.from_numpy(d) for d in r["dW"]]
        g = float(grad_loss)
        grads = [d * g for d in dW]

        def scalar(value, like):
            if isinstance(like, torch.Tensor):
                return torch.tensor(value * g, dtype=like.dtype, device=like.device)
            return None

        return (None, *grads, scalar(r["dg_leak"], ctx.scalar_like[0]), scalar(r["dtau_decay_e"], ctx.scalar_like[1]))


def window_loss(window: Window, weights, g_leak, tau_decay_e):
    """The window loss as a differentiable function of (weights[0..3], g_leak, tau_decay_e)."""
    return WindowLoss.apply(window, *weights, g_leak, tau_decay_e)
