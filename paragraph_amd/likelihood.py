"""PA log-likelihood of a citation sequence as a differentiable PyTorch op.

``PALogLikelihood`` is a ``torch.autograd.Function`` over the HIP kernels
behind ``pg_pa_loglik`` / ``pg_pa_loglik_grad`` (include/paragraph.h): the
forward pass returns log L; the backward pass calls the gradient entry,
which recomputes the per-year partition sums and every weight instead of
keeping anything from the forward pass.  Parameters stay on the device
({alpha, tau} as a float64 pair, log-fitness phi as float32), kernels run on
torch's current stream and nothing synchronises with the host, so the op can
sit inside an ordinary training loop (``fit_adam``).

``loglik_plain_torch`` restates the same likelihood with stock PyTorch ops
(FP64, autograd for the gradient).  It is the plain comparator the benchmark
times the fused kernels against, and an independent FP64 check in the tests;
``PALogLikelihood`` never calls it.
"""
from __future__ import annotations

import torch

from . import Context


def _stream(device: torch.device) -> int:
    return torch.cuda.current_stream(device).cuda_stream


class PALogLikelihood(torch.autograd.Function):
    """log L(alpha, tau, phi) of the sequence held by a ``Context``.

    alpha_tau: float64 CUDA tensor [2] = (alpha, tau); logfit: float32 CUDA
    tensor [n_papers].  Returns a float64 scalar tensor.
    """

    @staticmethod
    def forward(fctx, pgctx: Context, alpha_tau: torch.Tensor, logfit: torch.Tensor):
        if not (alpha_tau.is_cuda and logfit.is_cuda):
            raise ValueError("PALogLikelihood: alpha_tau and logfit must be CUDA tensors")
        if alpha_tau.dtype != torch.float64 or alpha_tau.numel() != 2:
            raise ValueError("PALogLikelihood: alpha_tau must be float64 with 2 elements")
        if logfit.dtype != torch.float32:
            raise ValueError("PALogLikelihood: logfit must be float32")
        n = pgctx.seq_info().n_papers
        if logfit.numel() < n:
            raise ValueError(f"PALogLikelihood: logfit has {logfit.numel()} < {n} entries")
        at = alpha_tau.detach().contiguous()
        phi = logfit.detach().contiguous()
        out = torch.empty(3, dtype=torch.float64, device=logfit.device)
        pgctx.pa_loglik(0.0, 1.0, phi.data_ptr(), alpha_tau_dev=at.data_ptr(),
                        scalars_dev=out.data_ptr(), stream=_stream(logfit.device), host=False)
        fctx.pgctx = pgctx
        fctx.save_for_backward(at, phi)
        return out[0]

    @staticmethod
    def backward(fctx, g: torch.Tensor):
        at, phi = fctx.saved_tensors
        out = torch.empty(3, dtype=torch.float64, device=phi.device)
        grad = torch.empty(phi.numel(), dtype=torch.float64, device=phi.device)
        n = fctx.pgctx.seq_info().n_papers
        if n < phi.numel():
            grad[n:].zero_()
        fctx.pgctx.pa_loglik_grad(0.0, 1.0, phi.data_ptr(), grad.data_ptr(),
                                  alpha_tau_dev=at.data_ptr(), scalars_dev=out.data_ptr(),
                                  stream=_stream(phi.device), host=False)
        g = g.to(torch.float64)
        d_at = out[1:3] * g if fctx.needs_input_grad[1] else None
        d_phi = (grad * g).to(phi.dtype) if fctx.needs_input_grad[2] else None
        return None, d_at, d_phi


def pa_loglik(pgctx: Context, alpha_tau: torch.Tensor, logfit: torch.Tensor) -> torch.Tensor:
    return PALogLikelihood.apply(pgctx, alpha_tau, logfit)


def fit_adam(pgctx: Context, logfit0: torch.Tensor, alpha0: float = 1.0, tau0: float = 5.0,
             steps: int = 100, lr: float = 0.02, fit_logfit: bool = True,
             per_event: float = 1.0):
    """Maximises log L by Adam from (alpha0, tau0, logfit0).

    The loss is -log L * per_event (pass 1/#events for a mean).  Returns
    (alpha_tau [2] float64, logfit float32, history of log L values as a
    float64 CUDA tensor; nothing is copied to the host inside the loop).
    """
    dev = logfit0.device
    at = torch.tensor([alpha0, tau0], dtype=torch.float64, device=dev, requires_grad=True)
    phi = logfit0.detach().clone().to(torch.float32).requires_grad_(fit_logfit)
    params = [at, phi] if fit_logfit else [at]
    opt = torch.optim.Adam(params, lr=lr)
    hist = torch.empty(steps, dtype=torch.float64, device=dev)
    for i in range(steps):
        opt.zero_grad(set_to_none=True)
        ll = PALogLikelihood.apply(pgctx, at, phi)
        (-ll * per_event).backward()
        opt.step()
        with torch.no_grad():
            at[1].clamp_(min=1e-3)
        hist[i] = ll.detach()
    return at.detach(), phi.detach(), hist


class PlainSequence:
    """The sequence as torch tensors for ``loglik_plain_torch``."""

    def __init__(self, seq: dict, device):
        self.years = []
        m = seq["m"]
        self.birth = torch.as_tensor(seq["birth"], device=device).to(torch.int64)
        slot = torch.arange(max(m, 1), device=device)
        for y in range(len(seq["n_exist"])):
            deg = torch.as_tensor(seq["deg"][y], device=device).to(torch.float64)
            cit = torch.as_tensor(seq["cit"][y], device=device).to(torch.int64)
            kind = torch.as_tensor(seq["kind"][y], device=device)
            cnt = torch.as_tensor(seq["cnt"][y], device=device).to(torch.int64)
            valid = slot[None, :m] < cnt[:, None]
            cit = torch.where(valid, cit, torch.zeros_like(cit))
            event = valid & (kind != 1)
            self.years.append((int(seq["year_of"][y]), deg, cit, valid, event))


def loglik_plain_torch(ps: PlainSequence, alpha_tau: torch.Tensor,
                       logfit: torch.Tensor) -> torch.Tensor:
    """log L with stock PyTorch ops in FP64 (gradient by autograd)."""
    alpha, tau = alpha_tau[0], alpha_tau[1]
    phi = logfit.to(torch.float64)
    ll = torch.zeros((), dtype=torch.float64, device=logfit.device)
    for year, deg, cit, valid, event in ps.years:
        N = deg.shape[0]
        L = torch.log1p(deg)
        age = (year - ps.birth[:N]).to(torch.float64)
        x = alpha * L - age / tau + phi[:N]
        w = torch.exp(x)
        Z = w.sum()
        if cit.numel() == 0:
            continue
        wt = torch.where(valid, w[cit], torch.zeros_like(w[cit]))
        P = torch.cumsum(wt, dim=1) - wt
        term = x[cit] - torch.log(Z - P)
        ll = ll + torch.where(event, term, torch.zeros_like(term)).sum()
    return ll
