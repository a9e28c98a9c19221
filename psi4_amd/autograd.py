"""Differentiable two-electron SCF energy on the GPU JK.

``scf_energy_2e(jk, C, B=None)`` returns E = tr(D(2J - K)), D = C C^T, as a
torch scalar whose backward runs the library's adjoint kernels
(``dfjk_energy_adjoint_device``):

    dE/dC     = 4 (2J - K) C
    dE/dB_Qmn = 4 d_Q D_mn - 2 (D B_Q D)_mn          (d_Q = <B_Q, D>)

dE/dB is the three-index gradient density that psi4's DF-SCF gradient
contracts with derivative integrals (scfgrad/jk_grad.cc:240-262, 868+ form the
same contraction from (A|mn) D_mn and (A|ij) C_pi C_qi).

B, when given, is the dense fitted tensor [naux][nbf][nbf] on the device.  The
energy is then a function of that B alone: it is loaded into a tensor-level JK
object of the same shape (no screening mask, so every element counts) that is
kept beside the caller's JK and reused, and the caller's JK -- its resident
tensor, auxiliary range and precision -- is left as it was.  B must be
symmetric in its last two indices: the JK object stores one triangle, and dE/dB
is the gradient with respect to each dense element at a symmetric point.  All
arithmetic is FP64.
"""
import torch


def _sync():
    torch.cuda.current_stream().synchronize()


def _tensor_twin(jk, nbf, naux):
    """Tensor-level JK of jk's shape that holds a caller-supplied B (made once, kept on jk)."""
    twin = getattr(jk, "_autograd_twin", None)
    if twin is None:
        from .core import JK
        twin = JK.build_tensor(nbf, naux)
        twin.set_print(0)
        twin.initialize()
        jk._autograd_twin = twin
    return twin


class _SCFEnergy2e(torch.autograd.Function):
    @staticmethod
    def forward(ctx, C, B, jk):
        if C.dim() != 2 or C.dtype != torch.float64 or not C.is_cuda:
            raise ValueError("C must be a CUDA float64 nbf x nocc tensor")
        info = jk.info()
        if C.shape[0] != info["nbf"]:
            raise ValueError("C has %d rows, the basis has %d functions" % (C.shape[0], info["nbf"]))
        C = C.contiguous()
        if B is not None:
            n = info["nbf"]
            if tuple(B.shape) != (info["naux"], n, n) or B.dtype != torch.float64 or not B.is_cuda:
                raise ValueError("B must be a CUDA float64 naux x nbf x nbf tensor")
            B = B.contiguous()
            jk = _tensor_twin(jk, n, info["naux"])
            _sync()
            jk.set_B_dense_device(B.data_ptr())
        _sync()
        E = jk.energy_adjoint_device(C.data_ptr(), C.shape[1], 0, 0)
        ctx.jk = jk
        ctx.with_B = B is not None
        if B is not None:
            ctx.save_for_backward(C, B)
        else:
            ctx.save_for_backward(C)
        return torch.tensor(E, dtype=torch.float64, device=C.device)

    @staticmethod
    def backward(ctx, gE):
        C = ctx.saved_tensors[0]
        jk = ctx.jk
        if ctx.with_B:
            # the twin may hold the B of a later forward: load this one again
            _sync()
            jk.set_B_dense_device(ctx.saved_tensors[1].data_ptr())
        want_C = ctx.needs_input_grad[0]
        want_B = ctx.with_B and ctx.needs_input_grad[1]
        dC = torch.empty_like(C) if want_C else None
        dB = None
        if want_B:
            info = jk.info()                 # the twin: its whole auxiliary range is in use
            dB = torch.empty((info["naux"], info["nbf"], info["nbf"]), dtype=torch.float64, device=C.device)
        _sync()
        jk.energy_adjoint_device(C.data_ptr(), C.shape[1], dC.data_ptr() if want_C else 0,
                                 dB.data_ptr() if want_B else 0)
        if want_C:
            dC.mul_(gE)
        if want_B:
            dB.mul_(gE)
        return dC, dB, None


def scf_energy_2e(jk, C, B=None):
    """E = tr(D(2J - K)) with D = C C^T, differentiable in C (and B when given)."""
    return _SCFEnergy2e.apply(C, B, jk)
