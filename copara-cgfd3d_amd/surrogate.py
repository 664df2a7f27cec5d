"""Fourier-neural-operator surrogate for surface peak ground velocity, with
the finite-difference engine as its data source, its fallback and the
reference its errors are measured against.

For a fixed topography and medium, the surrogate maps a source
(x, y, depth, strike, dip, rake, log10 M0) to the three surface PGV maps the
engine's PGV capture produces and to the velocity traces of fixed surface
stations.

- **Linearity in M0.** The engine is linear in the moment, so the network
  predicts log10(PGV / M0) and log10 M0 is added exactly (traces: trace / M0
  and the factor M0 applied exactly).
- **Matrix products.** Every product of the network runs on the BF16 MFMA
  kernel of libcgfd_fno.so (FP32 accumulation), in the forward pass and in
  both gradients of the backward pass: the lifting and projection layers,
  the pointwise channel mixing, the per-mode complex spectral mixing
  (realified as 2x2 real blocks), and the truncated Fourier transforms
  themselves: only the kept modes are ever formed, so each forward and
  inverse transform is a product with a small constant DFT matrix on the
  channels-last activations.  Each layer's bias, residual sum and GELU run
  in the epilogue of its product, the GELU backward and the Adam update are
  kernels of the same library; PyTorch supplies device memory, indexing and
  reductions, and its FFT runs only in the FP32 reference path.
- **Scale.** The PGV maps cover a centred square window of the surface
  (the whole surface by default); with stations, a seismogram head predicts
  their three-component traces, one sample per step, from the trunk's
  latent at each station.
- **Gate.** Two networks trained from different seeds. Their disagreement is
  the error indicator, and scenarios above the threshold are re-run by the
  FD engine.

There is no reference counterpart (the reference has no learned model).
"""
import ctypes
import math
import os
import time

import numpy as np
import torch

import scenario

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


class SurrogateError(RuntimeError):
    pass


def load_library():
    """libcgfd_fno.so; there is no fallback when it is missing."""
    global _LIB
    if _LIB is None:
        path = os.environ.get("CGFD_FNO_LIB", os.path.join(_HERE, "libcgfd_fno.so"))
        if not os.path.exists(path):
            raise SurrogateError(f"{path} is not built (run copara-cgfd3d_amd/build.sh)")
        lib = ctypes.CDLL(path)
        L, P, I = ctypes.c_long, ctypes.c_void_p, ctypes.c_int
        D, LL = ctypes.c_double, ctypes.c_longlong
        lib.cgfd_fno_gemm_bf16.argtypes = [P, L, L, L, P, L, L, L, P, L, L, I, I, I, I, I, P]
        lib.cgfd_fno_gemm_bf16_ex.argtypes = [P, L, L, L, P, L, L, L, P, L, L, I, I, I, I, I, P, P, L, L, P, I,
                                              P, L, L, P]
        lib.cgfd_fno_gelu_bwd.argtypes = [P, P, P, P, L, P]
        lib.cgfd_fno_adam.argtypes = [P, P, I, D, D, D, D, LL, I, P]
        for fn in ("cgfd_fno_gemm_bf16", "cgfd_fno_gemm_bf16_ex", "cgfd_fno_gelu_bwd", "cgfd_fno_adam"):
            getattr(lib, fn).restype = I
        _LIB = lib
    return _LIB


def _span_ok(t, offsets):
    """Every element the kernel addresses from t's data pointer lies inside
    t's storage (all strides here are non-negative)."""
    last = t.storage_offset() + sum(offsets)
    return t.storage_offset() >= 0 and last < t.untyped_storage().nbytes() // t.element_size()


ACT_NONE, ACT_GELU = 0, 1


def _as3(t):
    return t if t is None or t.dim() == 3 else t.unsqueeze(0)


def gemm(a, b, out, accumulate=False, bias=None, addend=None, pre=None, act=ACT_NONE, out_bf16=None):
    """out[z] = act((out[z] +) a[z] @ b[z] + bias + addend[z]) on the matrix
    cores, the value before the activation also stored into pre[z] and the
    result again in bf16 into out_bf16[z] when given.  a: (batch,) M x K
    bf16, b: (batch,) K x N bf16, out: (batch,) M x N fp32 with unit column
    stride; any non-negative strides.  bias: N fp32 (contiguous); addend:
    (batch,) M x N fp32 with unit column stride; pre: strided as out;
    out_bf16: (batch,) M x N bf16 with unit column stride.  Runs on the
    current stream."""
    if a.dim() == 2:
        a, b, out = a.unsqueeze(0), b.unsqueeze(0), out.unsqueeze(0)
    addend, pre, out_bf16 = _as3(addend), _as3(pre), _as3(out_bf16)
    if not (a.is_cuda and b.is_cuda and out.is_cuda):
        raise SurrogateError("gemm operands must be on the GPU")
    if a.dtype != torch.bfloat16 or b.dtype != torch.bfloat16 or out.dtype != torch.float32:
        raise SurrogateError("gemm takes bf16 inputs and an fp32 output")
    Z, M, K = a.shape
    Zb, Kb, N = b.shape
    if (Zb, Kb) != (Z, K) or tuple(out.shape) != (Z, M, N) or out.stride(2) != 1:
        raise SurrogateError(f"gemm shapes {tuple(a.shape)} x {tuple(b.shape)} -> {tuple(out.shape)}")
    if min(a.stride() + b.stride() + out.stride()) < 0:
        raise SurrogateError("gemm strides must be non-negative")
    for t, ext in ((a, (Z, M, K)), (b, (Z, K, N)), (out, (Z, M, N))):
        if not _span_ok(t, [(e - 1) * s for e, s in zip(ext, t.stride())]):
            raise SurrogateError("gemm operand strides address memory outside the tensor")
    if bias is not None and (not bias.is_cuda or bias.dtype != torch.float32 or tuple(bias.shape) != (N,)
                             or not bias.is_contiguous()):
        raise SurrogateError("gemm bias must be N contiguous fp32 values on the GPU")
    if addend is not None:
        if (not addend.is_cuda or addend.dtype != torch.float32 or tuple(addend.shape) != (Z, M, N)
                or addend.stride(2) != 1 or min(addend.stride()) < 0
                or not _span_ok(addend, [(e - 1) * s for e, s in zip((Z, M, N), addend.stride())])):
            raise SurrogateError("gemm addend must be a (batch,) M x N fp32 tensor with unit column stride")
    if pre is not None and (not pre.is_cuda or pre.dtype != torch.float32 or pre.shape != out.shape
                            or pre.stride() != out.stride()
                            or not _span_ok(pre, [(e - 1) * s for e, s in zip((Z, M, N), pre.stride())])):
        raise SurrogateError("gemm pre must be an fp32 tensor shaped and strided as out")
    if out_bf16 is not None and (not out_bf16.is_cuda or out_bf16.dtype != torch.bfloat16
                                 or tuple(out_bf16.shape) != (Z, M, N) or out_bf16.stride(2) != 1
                                 or min(out_bf16.stride()) < 0
                                 or not _span_ok(out_bf16, [(e - 1) * s for e, s in zip((Z, M, N), out_bf16.stride())])):
        raise SurrogateError("gemm out_bf16 must be a (batch,) M x N bf16 tensor with unit column stride")
    sa, sb, so = a.stride(), b.stride(), out.stride()
    lib = load_library()
    stream = torch.cuda.current_stream().cuda_stream
    if bias is None and addend is None and pre is None and act == ACT_NONE and out_bf16 is None:
        rc = lib.cgfd_fno_gemm_bf16(a.data_ptr(), sa[0], sa[1], sa[2], b.data_ptr(), sb[0], sb[1], sb[2],
                                    out.data_ptr(), so[0], so[1], M, N, K, Z, int(accumulate), stream)
    else:
        sd = addend.stride() if addend is not None else (0, 0, 1)
        se = out_bf16.stride() if out_bf16 is not None else (0, 0, 1)
        rc = lib.cgfd_fno_gemm_bf16_ex(
            a.data_ptr(), sa[0], sa[1], sa[2], b.data_ptr(), sb[0], sb[1], sb[2], out.data_ptr(), so[0], so[1],
            M, N, K, Z, int(accumulate), None if bias is None else bias.data_ptr(),
            None if addend is None else addend.data_ptr(), sd[0], sd[1], None if pre is None else pre.data_ptr(),
            int(act), None if out_bf16 is None else out_bf16.data_ptr(), se[0], se[1], stream)
    if rc != 0:
        raise SurrogateError(f"cgfd_fno_gemm_bf16 returned {rc}")
    return out


def gelu_backward(gy, pre, gx=None, want_bf16=True):
    """gy * GELU'(pre) (fp32) and, with want_bf16, the same values in bf16,
    from the project's kernel.  gx may be gy (in place)."""
    if not (gy.is_cuda and pre.is_cuda) or gy.dtype != torch.float32 or pre.dtype != torch.float32:
        raise SurrogateError("gelu_backward takes fp32 tensors on the GPU")
    if gy.shape != pre.shape or not gy.is_contiguous() or not pre.is_contiguous():
        raise SurrogateError("gelu_backward takes contiguous tensors of one shape")
    gx = torch.empty_like(gy) if gx is None else gx
    if gx.shape != gy.shape or not gx.is_contiguous() or gx.dtype != torch.float32:
        raise SurrogateError("gelu_backward output must be contiguous fp32 of the input's shape")
    gxb = torch.empty(gy.shape, dtype=torch.bfloat16, device=gy.device) if want_bf16 else None
    rc = load_library().cgfd_fno_gelu_bwd(gy.data_ptr(), pre.data_ptr(), gx.data_ptr(),
                                          None if gxb is None else gxb.data_ptr(), gy.numel(),
                                          torch.cuda.current_stream().cuda_stream)
    if rc != 0:
        raise SurrogateError(f"cgfd_fno_gelu_bwd returned {rc}")
    return gx, gxb


def _weight_grad(xb, gyb, chunk=512):
    """x^T @ gy for tall x [R, K] and gy [R, N]: R is split into chunks (the
    batch index of the GEMM, enough of them to fill the chip) whose partial
    products are summed."""
    R, K = xb.shape
    N = gyb.shape[1]
    full = R // chunk
    gw = torch.zeros(K, N, device=xb.device, dtype=torch.float32)
    if full:
        xa = xb[: full * chunk].view(full, chunk, K).transpose(1, 2)     # [full, K, chunk]
        ga = gyb[: full * chunk].view(full, chunk, N)                    # [full, chunk, N]
        part = torch.empty(full, K, N, device=xb.device, dtype=torch.float32)
        gemm(xa, ga, part)
        gw += part.sum(0)
    if R - full * chunk:
        gemm(xb[full * chunk:].t(), gyb[full * chunk:], gw, accumulate=True)
    return gw


class _Matmul(torch.autograd.Function):
    """y = x @ w with x [R, K], w [K, N] (fp32 masters, bf16 on the cores).
    K is zero-padded to a multiple of 8 so that x's rows stage with 16-byte
    loads; the padding adds exact zeros and is cut from the gradients."""

    @staticmethod
    def forward(ctx, x, w):
        K = x.shape[1]
        pad = -K % 8
        xb = torch.nn.functional.pad(x.to(torch.bfloat16), (0, pad)).contiguous()
        wb = torch.nn.functional.pad(w.to(torch.bfloat16), (0, 0, 0, pad)).contiguous()
        y = torch.empty(x.shape[0], w.shape[1], device=x.device, dtype=torch.float32)
        gemm(xb, wb, y)
        ctx.save_for_backward(xb, wb)
        ctx.k = K
        return y

    @staticmethod
    def backward(ctx, gy):
        xb, wb = ctx.saved_tensors
        gyb = gy.contiguous().to(torch.bfloat16)
        gx = torch.empty(xb.shape, device=gy.device, dtype=torch.float32)
        gemm(gyb, wb.t(), gx)
        return gx[:, :ctx.k], _weight_grad(xb, gyb)[:ctx.k]


class _Linear(torch.autograd.Function):
    """y = act(x @ w + b (+ d)) with x [R, K], w [K, N], b [N] and an
    optional residual d [R, N]: one product with the bias, the residual and
    the GELU in its epilogue.  The backward runs the GELU derivative as one
    kernel that also writes the bf16 operand of the two gradient products."""

    @staticmethod
    def forward(ctx, x, w, b, d, gelu, xb, want_bf16):
        K = x.shape[1]
        pad = -K % 8
        if xb is None or pad:  # a producer's bf16 copy serves as the operand when no padding is needed
            xb = torch.nn.functional.pad(x.to(torch.bfloat16), (0, pad))
        xb = xb.contiguous()
        wb = torch.nn.functional.pad(w.to(torch.bfloat16), (0, 0, 0, pad)).contiguous()
        y = torch.empty(x.shape[0], w.shape[1], device=x.device, dtype=torch.float32)
        pre = torch.empty_like(y) if gelu else None
        yb = torch.empty(y.shape, device=x.device, dtype=torch.bfloat16) if want_bf16 else None
        gemm(xb, wb, y, bias=b.contiguous(), addend=None if d is None else d.contiguous(), pre=pre,
             act=ACT_GELU if gelu else ACT_NONE, out_bf16=yb)
        ctx.save_for_backward(xb, wb, pre)
        ctx.k, ctx.gelu, ctx.has_d = K, gelu, d is not None
        if yb is None:
            return y
        ctx.mark_non_differentiable(yb)
        ctx.set_materialize_grads(False)
        return y, yb

    @staticmethod
    def backward(ctx, gy, *unused):
        xb, wb, pre = ctx.saved_tensors
        gy = gy.contiguous()
        if ctx.gelu:
            gz, gzb = gelu_backward(gy, pre)
        else:
            gz, gzb = gy, gy.to(torch.bfloat16)
        gx = torch.empty(xb.shape, device=gy.device, dtype=torch.float32)
        gemm(gzb, wb.t(), gx)
        return gx[:, :ctx.k], _weight_grad(xb, gzb)[:ctx.k], gz.sum(0), (gz if ctx.has_d else None), None, None, None


class _Bmm(torch.autograd.Function):
    """y[p] = x[p] @ w[p] with x [P, B, K], w [P, K, N]."""

    @staticmethod
    def forward(ctx, x, w):
        xb = x.contiguous().to(torch.bfloat16)
        wb = w.contiguous().to(torch.bfloat16)
        y = torch.empty(x.shape[0], x.shape[1], w.shape[2], device=x.device, dtype=torch.float32)
        gemm(xb, wb, y)
        ctx.save_for_backward(xb, wb)
        return y

    @staticmethod
    def backward(ctx, gy):
        xb, wb = ctx.saved_tensors
        gyb = gy.contiguous().to(torch.bfloat16)
        gx = torch.empty(xb.shape, device=gy.device, dtype=torch.float32)
        gemm(gyb, wb.transpose(1, 2), gx)
        gw = torch.empty(wb.shape, device=gy.device, dtype=torch.float32)
        gemm(xb.transpose(1, 2), gyb, gw)
        return gx, gw


class _LeftConst(torch.autograd.Function):
    """y[z] = a @ x[z] for a constant matrix a [M, K] (bf16) and x [Z, K, N]:
    the truncated DFTs of the spectral layers."""

    @staticmethod
    def forward(ctx, x, a, xb=None):
        Z, K, N = x.shape
        xb = x.contiguous().to(torch.bfloat16) if xb is None else xb.contiguous()
        y = torch.empty(Z, a.shape[0], N, device=x.device, dtype=torch.float32)
        gemm(a.unsqueeze(0).expand(Z, -1, -1), xb, y)
        ctx.save_for_backward(a)
        return y

    @staticmethod
    def backward(ctx, gy):
        (a,) = ctx.saved_tensors
        Z, M, N = gy.shape
        gyb = gy.contiguous().to(torch.bfloat16)
        gx = torch.empty(Z, a.shape[1], N, device=gy.device, dtype=torch.float32)
        gemm(a.t().unsqueeze(0).expand(Z, -1, -1), gyb, gx)
        return gx, None, None


_DFT = {}


def dft_matrices(H, W, m, device):
    """The truncated transforms of an FNO layer on an H x W grid keeping
    wavenumbers kw < m and kh in [0, m) u [H-m, H) (the kept kh indexed
    0..2m-1), as bf16 matrices (computed in float64):

    fw [2m, W]    rows (re/im, kw): the forward real DFT along w
    gh [4m, 2H]   rows (re/im, kh index), columns (h, re/im): the forward
                  complex DFT along h, realified
    gi [2H, 4m]   rows (h, re/im), columns (kh index, re/im): the inverse
                  complex DFT along h, realified
    fi [W, 2m]    columns (re/im, kw): the inverse real DFT along w with the
                  one-sided weights (1 at kw = 0, else 2) and 1/(H W)

    These reproduce numpy's (and torch's) rfft2 / irfft2 conventions: e^-i
    forward, e^+i inverse, no forward scaling, and the imaginary part of the
    kw = 0 column ignored by the real inverse."""
    key = (H, W, m, str(device))
    if key not in _DFT:
        _DFT[key] = tuple(torch.from_numpy(a).to(device=device, dtype=torch.bfloat16).contiguous()
                          for a in dft_arrays(H, W, m))
    return _DFT[key]


def dft_arrays(H, W, m):
    """The float64 arrays behind dft_matrices (same shapes and layouts)."""
    if not (2 * m <= H and m <= W // 2):
        raise SurrogateError(f"{m} modes do not fit an {H} x {W} grid")
    kw = np.arange(m)[:, None]
    w = np.arange(W)[None, :]
    phi = 2 * np.pi * kw * w / W                                   # [m, W]
    fw = np.concatenate([np.cos(phi), -np.sin(phi)])               # [2m, W]
    kh = np.concatenate([np.arange(m), np.arange(H - m, H)])[:, None]
    h = np.arange(H)[None, :]
    th = 2 * np.pi * kh * h / H                                    # [2m, H]
    c, sn = np.cos(th), np.sin(th)
    gh = np.empty((2, 2 * m, H, 2))
    gh[0, :, :, 0], gh[0, :, :, 1] = c, sn                         # re = cos*re + sin*im
    gh[1, :, :, 0], gh[1, :, :, 1] = -sn, c                        # im = cos*im - sin*re
    gi = np.empty((H, 2, 2 * m, 2))
    gi[:, 0, :, 0], gi[:, 0, :, 1] = c.T, -sn.T                    # re = cos*re - sin*im
    gi[:, 1, :, 0], gi[:, 1, :, 1] = sn.T, c.T                     # im = sin*re + cos*im
    wt = np.where(np.arange(m) == 0, 1.0, 2.0)[None, :] / (H * W)
    fi = np.concatenate([wt * np.cos(phi.T), -wt * np.sin(phi.T)], axis=1)   # [W, 2m]
    return tuple(np.ascontiguousarray(a.reshape(r, -1)) for a, r in ((fw, 2 * m), (gh, 4 * m), (gi, 2 * H), (fi, W)))


class Adam:
    """Adam (torch.optim.Adam's update without weight decay or amsgrad) on
    the project's kernel: one launch updates every parameter.  Gradients
    live in preallocated .grad tensors that backward accumulates into and
    the kernel clears as it reads them, so zero_grad launches nothing and
    the launch's address table is built once (and again only if a .grad
    tensor is replaced)."""

    def __init__(self, params, lr=1e-3, betas=(0.9, 0.999), eps=1e-8):
        self.params = [p for p in params]
        if not self.params or any(not p.is_cuda or p.dtype != torch.float32 for p in self.params):
            raise SurrogateError("Adam takes fp32 parameters on the GPU")
        self.lr, (self.beta1, self.beta2), self.eps = float(lr), (float(betas[0]), float(betas[1])), float(eps)
        self.m = [torch.zeros_like(p) for p in self.params]
        self.v = [torch.zeros_like(p) for p in self.params]
        for p in self.params:
            if p.grad is None or not p.grad.is_contiguous():
                p.grad = torch.zeros_like(p, memory_format=torch.contiguous_format)
        self.steps = 0
        self._table = None

    def _tables(self):
        key = [p.grad.data_ptr() for p in self.params]
        if self._table is None or self._table[0] != key:
            rows, chunks = [], []
            for t, (p, m, v) in enumerate(zip(self.params, self.m, self.v)):
                if not (p.is_contiguous() and p.grad.is_contiguous() and p.grad.shape == p.shape):
                    raise SurrogateError("Adam needs contiguous parameters and gradients")
                rows.append([p.data_ptr(), p.grad.data_ptr(), m.data_ptr(), v.data_ptr(), p.numel()])
                chunks += [[t, o] for o in range(0, p.numel(), ADAM_CHUNK)]
            dev = self.params[0].device
            self._table = (key, torch.tensor(rows, dtype=torch.int64, device=dev),
                           torch.tensor(chunks, dtype=torch.int64, device=dev), len(chunks))
        return self._table[1:]

    def zero_grad(self):
        """The step clears every gradient it reads; only a gradient that was
        replaced or dropped since needs work here."""
        for p in self.params:
            if p.grad is None:
                p.grad = torch.zeros_like(p)

    def step(self, lr=None):
        rows, chunks, n = self._tables()
        self.steps += 1
        rc = load_library().cgfd_fno_adam(rows.data_ptr(), chunks.data_ptr(), n, self.lr if lr is None else float(lr),
                                          self.beta1, self.beta2, self.eps, self.steps, 1,
                                          torch.cuda.current_stream().cuda_stream)
        if rc != 0:
            raise SurrogateError(f"cgfd_fno_adam returned {rc}")


ADAM_CHUNK = 8192  # CGFD_FNO_ADAM_CHUNK


def cosine_lr(lr, epoch, epochs):
    """The cosine schedule (annealed to zero over `epochs`) at an epoch."""
    return 0.5 * lr * (1.0 + math.cos(math.pi * epoch / epochs))


def matmul(x, w, backend):
    return _Matmul.apply(x, w) if backend == "mfma" else x @ w


def linear(x, w, b, backend, d=None, gelu=False, xb=None, want_bf16=False):
    """act(x @ w + b (+ d)): the fused kernel (backend "mfma") or FP32
    PyTorch (backend "torch").  xb: x already in bf16 (a producer's copy);
    with want_bf16 the result comes with its own bf16 copy, (y, yb) (yb is
    None on the torch backend)."""
    if backend == "mfma":
        return _Linear.apply(x, w, b, d, gelu, xb, want_bf16)
    y = x @ w + b
    if d is not None:
        y = y + d
    y = torch.nn.functional.gelu(y) if gelu else y
    return (y, None) if want_bf16 else y


def bmm(x, w, backend):
    return _Bmm.apply(x, w) if backend == "mfma" else torch.bmm(x, w)


# ---------------------------------------------------------------------------
# the network
# ---------------------------------------------------------------------------
class FNO2d(torch.nn.Module):
    """Lifting, `layers` Fourier layers (spectral mixing of the lowest
    `modes` x `modes` frequencies + pointwise mixing, GELU between layers)
    and a two-layer projection, on channels-last [B, H, W, C] activations.
    backend "mfma" runs the products on libcgfd_fno.so; "torch" is the plain
    FP32 reference used to check gradients."""

    def __init__(self, cin, cout, width=64, modes=16, layers=4, proj=128, seed=0):
        super().__init__()
        g = torch.Generator().manual_seed(seed)

        def lin(i, o):
            return torch.nn.Parameter((torch.rand(i, o, generator=g) * 2 - 1) / math.sqrt(i))

        self.width, self.modes, self.layers = width, modes, layers
        self.lift_w, self.lift_b = lin(cin, width), torch.nn.Parameter(torch.zeros(width))
        P = 2 * modes * modes
        scale = 1.0 / (width * width)
        self.spec_re = torch.nn.ParameterList(
            [torch.nn.Parameter(scale * torch.rand(P, width, width, generator=g)) for _ in range(layers)])
        self.spec_im = torch.nn.ParameterList(
            [torch.nn.Parameter(scale * torch.rand(P, width, width, generator=g)) for _ in range(layers)])
        self.pw_w = torch.nn.ParameterList([lin(width, width) for _ in range(layers)])
        self.pw_b = torch.nn.ParameterList([torch.nn.Parameter(torch.zeros(width)) for _ in range(layers)])
        self.p1_w, self.p1_b = lin(width, proj), torch.nn.Parameter(torch.zeros(proj))
        self.p2_w, self.p2_b = lin(proj, cout), torch.nn.Parameter(torch.zeros(cout))

    def _spectral(self, v, l, backend, vb=None):
        if backend == "mfma":
            return self._spectral_dft(v, l, vb)
        B, H, W, C = v.shape
        m = self.modes
        ft = torch.fft.rfft2(v, dim=(1, 2))                                   # [B, H, W/2+1, C]
        sel = torch.cat([ft[:, :m, :m], ft[:, H - m:, :m]], dim=1)            # [B, 2m, m, C]
        x = sel.permute(1, 2, 0, 3).reshape(2 * m * m, B, C)                  # [P, B, C]
        xr = torch.cat([x.real, x.imag], dim=2)                               # [P, B, 2C]
        wre, wim = self.spec_re[l], self.spec_im[l]
        wr = torch.cat([torch.cat([wre, wim], 2), torch.cat([-wim, wre], 2)], 1)   # [P, 2C, 2C]
        yr = bmm(xr, wr, backend)
        y = torch.complex(yr[..., :C], yr[..., C:]).reshape(2 * m, m, B, C).permute(2, 0, 1, 3)
        out = torch.zeros(B, H, W // 2 + 1, C, dtype=ft.dtype, device=v.device)
        out[:, :m, :m] = y[:, :m]
        out[:, H - m:, :m] = y[:, m:]
        return torch.fft.irfft2(out, s=(H, W), dim=(1, 2))

    def _spectral_dft(self, v, l, vb=None):
        """The spectral layer with the truncated transforms as products with
        the dft_matrices, on channels-last activations: no FFT and no
        full-size transposes, only the kept modes are permuted.  vb: v's bf16
        copy from the producing layer, if it made one."""
        B, H, W, C = v.shape
        m = self.modes
        fw, gh, gi, fi = dft_matrices(H, W, m, v.device)
        y1 = _LeftConst.apply(v.reshape(B * H, W, C), fw,
                              None if vb is None else vb.reshape(B * H, W, C))  # [(b,h), (ri,kw), c]
        y2 = _LeftConst.apply(y1.reshape(B, 2 * H, m * C), gh)               # [b, (ri,kh), (kw,c)]
        x = y2.reshape(B, 2, 2 * m, m, C).permute(2, 3, 0, 1, 4).reshape(2 * m * m, B, 2 * C)
        wre, wim = self.spec_re[l], self.spec_im[l]
        wr = torch.cat([torch.cat([wre, wim], 2), torch.cat([-wim, wre], 2)], 1)   # [P, 2C, 2C]
        yr = bmm(x, wr, "mfma")                                               # [(kh,kw), b, (ri,c)]
        z = yr.reshape(2 * m, m, B, 2, C).permute(2, 0, 3, 1, 4).reshape(B, 4 * m, m * C)
        z1 = _LeftConst.apply(z, gi)                                          # [b, (h,ri), (kw,c)]
        return _LeftConst.apply(z1.reshape(B * H, 2 * m, C), fi).reshape(B, H, W, C)

    def trunk(self, x, backend="mfma"):
        """The lifting and the Fourier layers: the latent [B, H, W, width]
        and its bf16 copy [B*H*W, width] (None on the torch backend), which
        every product's epilogue writes for the next products."""
        B, H, W, _ = x.shape
        R = B * H * W
        v, vb = linear(x.reshape(R, -1), self.lift_w, self.lift_b, backend, want_bf16=True)
        for l in range(self.layers):
            sp = self._spectral(v.reshape(B, H, W, self.width), l, backend, vb).reshape(R, self.width)
            v, vb = linear(v, self.pw_w[l], self.pw_b[l], backend, d=sp, gelu=l < self.layers - 1, xb=vb,
                           want_bf16=True)
        return v.reshape(B, H, W, self.width), vb

    def head(self, v, backend="mfma", vb=None):
        """The two-layer projection of a latent to the output channels."""
        B, H, W, _ = v.shape
        h, hb = linear(v.reshape(B * H * W, self.width), self.p1_w, self.p1_b, backend, gelu=True, xb=vb,
                       want_bf16=True)
        return linear(h, self.p2_w, self.p2_b, backend, xb=hb).reshape(B, H, W, -1)

    def forward(self, x, backend="mfma"):
        v, vb = self.trunk(x, backend)
        return self.head(v, backend, vb)


# ---------------------------------------------------------------------------
# scenarios and input encoding
# ---------------------------------------------------------------------------
FC = 1.5  # Ricker dominant frequency of every scenario (Hz)


def base_medium(ni=80, nj=80, nk=48, nt=600, dt=0.006, npml=10, window=None, stations=0, depth=None):
    """The fixed Gaussian-hill medium (h = 1.5 km, sigma = 3 km, three flat
    layers) on which the surrogate is trained.

    window: the side of the centred square of surface points whose PGV the
    surrogate maps (default: the whole surface, which must then be square).
    stations: the number of surface stations (a multiple of 4) whose
    three-component traces the seismogram head predicts, one sample per
    step.  depth: the (lowest, highest + 1) source depth in cells below the
    surface row (default 10 to 35, capped above the bottom PML)."""
    sc = scenario.config_hill(ni=ni, nj=nj, nk=nk, nt=nt, dh=100.0, dt=dt, npml=npml, nrecv_side=2)
    sc["stf"] = scenario.stf_table("Ricker", nt, dt, 1.2 / FC, FC)[None, :]
    n = min(ni, nj) if window is None else int(window)
    if ni != nj or (ni - n) % 2 or not 0 < n <= ni:
        raise SurrogateError(f"the PGV window ({n}) must be centred on a square surface ({ni} x {nj})")
    sc["window"] = ((ni - n) // 2, (nj - n) // 2, n)
    sc["stations"] = station_layout(sc, stations) if stations else np.zeros((0, 2), np.int64)
    sc["recv_locs"] = np.array([[i + scenario.LFD, j + scenario.LFD, nk - 1 + scenario.LFD]
                                for i, j in sc["stations"]], np.int32).reshape(-1, 3)
    lo, hi = (10, 36) if depth is None else depth
    sc["depth_cells"] = (int(lo), int(min(hi, nk - npml - 2)))
    return sc


def station_layout(base, count, seed=7):
    """`count` surface stations (grid indices (i, j)) inside the PGV window,
    closed under SYM: orbits of 4 points (p, its x/y swap, its mirror in
    both axes and their composition), drawn clear of the window's edge, of
    the two diagonals (where an orbit would collapse) and of each other;
    ordered orbit by orbit."""
    if count % 4:
        raise SurrogateError("the station count must be a multiple of 4")
    i0, _, n = base["window"]
    N = base["ni"]
    rng = np.random.default_rng(seed)
    pts = []
    for _ in range(100000):
        if len(pts) == count:
            break
        i, j = (int(v) for v in rng.integers(i0 + 6, i0 + n - 6, 2))
        if abs(i - j) < 3 or abs(i + j - (N - 1)) < 3:
            continue
        orbit = [(i, j), (j, i), (N - 1 - i, N - 1 - j), (N - 1 - j, N - 1 - i)]
        if all(max(abs(a - c), abs(b - d)) >= 4 for a, b in orbit for c, d in pts):
            pts += orbit
    if len(pts) != count:
        raise SurrogateError(f"no layout of {count} stations fits the window")
    return np.array(pts, np.int64)


def draw_sources(n, base, seed):
    """n sources on the base grid: grid-aligned (x, y) inside the PGV window
    and clear of the PML, at base["depth_cells"] below the surface row,
    uniform strike/dip/rake, log-uniform M0 1e14-1e17."""
    rng = np.random.default_rng(seed)
    ni, nj, dh = base["ni"], base["nj"], base["dh"]
    npml = int(base["nabs"][0])
    i0, j0, w = base["window"]
    ilo, ihi = max(npml + 4, i0 + 4), min(ni - npml - 4, i0 + w - 4)
    jlo, jhi = max(npml + 4, j0 + 4), min(nj - npml - 4, j0 + w - 4)
    dlo, dhi = base["depth_cells"]
    out = []
    for _ in range(n):
        i = int(rng.integers(ilo, ihi))
        j = int(rng.integers(jlo, jhi))
        dk = int(rng.integers(dlo, dhi))
        s, d, r = rng.uniform([0.0, 10.0, -180.0], [360.0, 90.0, 180.0])
        lm0 = rng.uniform(14.0, 17.0)
        out.append(np.array([i * dh, j * dh, dk * dh, s, d, r, lm0]))
    return np.array(out)


def source_rows(params):
    """(x, y, depth, mxx, myy, mzz, mxy, mxz, myz, log10 M0) rows: the unit
    moment tensor of each (strike, dip, rake)."""
    params = np.asarray(params, np.float64)
    m = np.array([scenario.angle2moment(*p[3:6]) for p in params]).reshape(-1, 6)
    return np.concatenate([params[:, :3], m, params[:, 6:7]], 1)


# The hill medium is symmetric under the dihedral group of the square about
# its centre (ni = nj, centred hill, flat layers, equal PML): each element is
# (swap x and y, mirror x, mirror y).  The discrete scheme keeps a subgroup:
# x and y share one forward/backward sign per stage, so swapping them is exact
# and mirroring both stays within ~0.1 %, while mirroring one axis alone
# changes the PGV maps by a few percent.  Training uses SYM, so every FD run
# yields 4 samples.
D4 = [(t, fx, fy) for t in (0, 1) for fx in (0, 1) for fy in (0, 1)]
SYM = [(0, 0, 0), (1, 0, 0), (0, 1, 1), (1, 1, 1)]


def transform(base, src, pgv, g):
    """Apply dihedral element g = (swap, mirror x, mirror y) to source rows
    [n, 10] and, if given, PGV maps [n, 3, w, w] of the centred window."""
    t, fx, fy = g
    ni, nj, dh = base["ni"], base["nj"], base["dh"]
    if t and ni != nj:
        raise ValueError("the x/y swap needs a square surface grid")
    s = np.array(src, np.float64, copy=True)
    P = None if pgv is None else np.array(pgv, copy=True)
    if t:
        s[:, [0, 1]] = s[:, [1, 0]]
        s[:, 3:9] = s[:, [4, 3, 5, 6, 8, 7]]            # mxx<->myy, mxz<->myz
        if P is not None:
            P = P[:, [1, 0, 2]].transpose(0, 1, 3, 2)   # |Vx| <-> |Vy|
    if fx:
        s[:, 0] = (ni - 1) * dh - s[:, 0]
        s[:, [6, 7]] *= -1                              # mxy, mxz
        if P is not None:
            P = P[:, :, ::-1, :]
    if fy:
        s[:, 1] = (nj - 1) * dh - s[:, 1]
        s[:, [6, 8]] *= -1                              # mxy, myz
        if P is not None:
            P = P[:, :, :, ::-1]
    return s, (None if P is None else np.ascontiguousarray(P))


def transform_traces(base, tr, g):
    """Apply dihedral element g to station traces [n, S, 3, T]: each
    station's record moves to the image station, with Vx and Vy exchanged
    by the swap and negated by the mirror of their axis."""
    t, fx, fy = g
    st = base["stations"]
    N = base["ni"]
    P = np.array(tr, copy=True)
    pos = st.copy()
    if t:
        P = P[:, :, [1, 0, 2]]
        pos = pos[:, [1, 0]]
    if fx:
        P[:, :, 0] *= -1
        pos[:, 0] = N - 1 - pos[:, 0]
    if fy:
        P[:, :, 1] *= -1
        pos[:, 1] = N - 1 - pos[:, 1]
    where = {(int(a), int(b)): k for k, (a, b) in enumerate(st)}
    idx = [where[(int(a), int(b))] for a, b in pos]
    out = np.empty_like(P)
    out[:, idx] = P
    return out


# The seismogram head works on traces in reduced time: each station's record
# starts at the direct P wave's arrival bound r / V_ALIGN (r the distance from
# the source to the station, V_ALIGN above every P velocity of the medium, so
# no arrival is cut) and is scaled by r / 1 km against geometric spreading.
V_ALIGN = 6000.0


def _station_geometry(base, src):
    """Sample shifts [n, S] and distances [n, S] (m) from source rows to the
    stations (source depth below the surface row, stations on it)."""
    st = np.asarray(base["stations"], np.float64) * base["dh"]
    src = np.asarray(src, np.float64)
    dx = st[None, :, 0] - src[:, None, 0]
    dy = st[None, :, 1] - src[:, None, 1]
    r = np.sqrt(dx * dx + dy * dy + src[:, 2:3] ** 2)
    shift = np.minimum(np.floor(r / V_ALIGN / base["dt"]).astype(np.int64), base["nt"] - 1)
    return shift, np.maximum(r, base["dh"])


def align_traces(base, src, tr, inverse=False):
    """Station traces [n, S, 3, T] to reduced time and spreading-corrected
    amplitude (or back, with inverse=True; the samples before the arrival
    bound come back as zeros)."""
    shift, r = _station_geometry(base, src)
    tr = np.asarray(tr)
    out = np.zeros_like(tr)
    T = tr.shape[-1]
    for n in range(tr.shape[0]):
        for q in range(tr.shape[1]):
            k, f = int(shift[n, q]), r[n, q] / 1000.0
            if inverse:
                out[n, q, :, k:] = tr[n, q, :, :T - k] / f
            else:
                out[n, q, :, :T - k] = tr[n, q, :, k:] * f
    return out


def fd_scenario(base, src):
    """The FD scenario dict of one source row."""
    dh = base["dh"]
    i, j, dk = int(round(src[0] / dh)), int(round(src[1] / dh)), int(round(src[2] / dh))
    sc = dict(base)
    sc["moment_locs"] = np.array([[i + 3, j + 3, base["nk"] - 1 - dk + 3]], np.int32)
    sc["moment_vec"] = (np.asarray(src[3:9]) * 10 ** src[9]).astype(np.float32)[None, :]
    sc["moment_stf"] = base["stf"]
    return sc


def run_fd(base, params, group=32, log=None, rows=False, graph=True, traces=False, share=True):
    """Surface PGV maps [n, 3, w, w] of the window of the scenarios
    (parameter rows, or source rows with rows=True), advanced `group` at a
    time with cgfd_run_batch (replaying captured step cycles when graph is
    set; the results are the same); `log` gets a line per group.  With
    traces=True also the stations' velocity traces [n, S, 3, nt].  With
    share, the grids of a group hold one device copy of the medium
    (cgfd_share_medium) and the batch runs their copies of each tile
    together; the results are the same."""
    import cgfd3d

    src = np.asarray(params, np.float64) if rows else source_rows(params)
    i0, j0, w = base["window"]
    out = np.empty((len(src), 3, w, w), np.float32)
    ns = len(base["stations"])
    trc = np.empty((len(src), ns, 3, base["nt"]), np.float32) if traces else None
    for g0 in range(0, len(src), group):
        batch = []
        for p in src[g0:g0 + group]:
            sc = fd_scenario(base, p)
            if share and batch:
                sc["medium_of"] = batch[0]
            s = cgfd3d.Simulation(sc["ni"], sc["nj"], sc["nk"], sc["dh"], sc["dt"], sc["nt"])
            s.setup(sc)
            s.set_graph(graph)
            batch.append(s)
        cgfd3d.run_batch(batch, base["nt"])
        for k, s in enumerate(batch):
            out[g0 + k] = s.get_pgv()[:, 3 + i0:3 + i0 + w, 3 + j0:3 + j0 + w]
            if traces:
                trc[g0 + k] = s.get_seismo()[:3].transpose(1, 0, 2)
            s.close()
        if log:
            log(f"FD scenarios {g0 + len(batch)}/{len(src)}")
    return (out, trc) if traces else out


NCH = 13


def _window(base):
    """(i0, j0, rows, columns) of the mapped surface points: the centred
    window, or the whole surface of a base without one."""
    if "window" in base:
        i0, j0, w = base["window"]
        return i0, j0, w, w
    return 0, 0, base["ni"], base["nj"]


def encode(base, src):
    """Input fields [n, w, w, 13] on the window's surface points from source
    rows: x and y, the source's relative position and a 1 km bump around
    its epicentre, depth, hypocentral distance and its logarithm, and five
    of the unit moment-tensor components (mzz = -(mxx + myy))."""
    ni, nj, dh = base["ni"], base["nj"], base["dh"]
    i0, j0, wi, wj = _window(base)
    L = max(ni, nj) * dh
    X, Y = np.meshgrid((i0 + np.arange(wi)) * dh, (j0 + np.arange(wj)) * dh, indexing="ij")
    out = np.empty((len(src), wi, wj, NCH), np.float32)
    for n, p in enumerate(src):
        dx, dy = (X - p[0]) / L, (Y - p[1]) / L
        dz = p[2] / L
        r = np.sqrt(dx * dx + dy * dy + dz * dz)
        out[n, ..., 0], out[n, ..., 1] = X / L, Y / L
        out[n, ..., 2], out[n, ..., 3] = dx, dy
        out[n, ..., 4] = np.exp(-((X - p[0]) ** 2 + (Y - p[1]) ** 2) / (2 * 1000.0 ** 2))
        out[n, ..., 5] = dz
        out[n, ..., 6] = r
        out[n, ..., 7] = np.log10(r)
        out[n, ..., 8:13] = p[[3, 4, 6, 7, 8]]
    return out


def target(pgv, src):
    """log10(PGV / M0) [n, ni, nj, 3] of source rows (a floor keeps silent
    points finite)."""
    m0 = 10.0 ** np.asarray(src)[:, 9]
    return np.log10(np.maximum(pgv / m0[:, None, None, None], 1e-30)).transpose(0, 2, 3, 1).astype(np.float32)


def pgv_error(pred, ref):
    """Relative L2 error of each scenario's three PGV maps: [n]."""
    d = (pred - ref).reshape(len(ref), -1)
    return np.linalg.norm(d, axis=1) / np.linalg.norm(ref.reshape(len(ref), -1), axis=1)


def pgv_pointwise_error(pred, ref):
    """Median over the surface points and components of |pred/ref - 1|, per
    scenario: [n]."""
    r = np.abs(pred / np.maximum(ref, 1e-30) - 1.0).reshape(len(ref), -1)
    return np.median(r, axis=1)


def trace_error(pred, ref):
    """Relative L2 error of each scenario's station traces [n, S, 3, T]: [n]."""
    d = (pred - ref).reshape(len(ref), -1)
    return np.linalg.norm(d, axis=1) / np.linalg.norm(ref.reshape(len(ref), -1), axis=1)


class StationHead(torch.nn.Module):
    """Seismogram head: at each station, the trunk's latent and the input
    fields (source offset, distance, mechanism) through a three-layer MLP
    (fused products) to the station's three velocity traces, in reduced
    time and corrected for geometric spreading (align_traces)."""

    def __init__(self, width, cin, nsamp, hidden=256, seed=0):
        super().__init__()
        g = torch.Generator().manual_seed(1000 + seed)

        def lin(i, o):
            return torch.nn.Parameter((torch.rand(i, o, generator=g) * 2 - 1) / math.sqrt(i))

        self.nsamp = nsamp
        self.w1, self.b1 = lin(width + cin, hidden), torch.nn.Parameter(torch.zeros(hidden))
        self.w2, self.b2 = lin(hidden, hidden), torch.nn.Parameter(torch.zeros(hidden))
        self.w3, self.b3 = lin(hidden, 3 * nsamp), torch.nn.Parameter(torch.zeros(3 * nsamp))

    def forward(self, v, x, si, sj, backend="mfma"):
        B, S = v.shape[0], len(si)
        f = torch.cat([v[:, si, sj], x[:, si, sj]], -1).reshape(B * S, -1)
        h = linear(f, self.w1, self.b1, backend, gelu=True)
        h = linear(h, self.w2, self.b2, backend, gelu=True)
        return linear(h, self.w3, self.b3, backend).reshape(B, S, 3, self.nsamp)


class Surrogate:
    """Two FNOs from different seeds (each with a seismogram head when the
    base medium has stations); predictions are their mean, and their
    disagreement on the PGV maps is the gate's error indicator."""

    def __init__(self, base, width=64, modes=16, layers=4, seeds=(0, 1), device="cuda", hidden=256):
        self.base = base
        self.nets = [FNO2d(NCH, 3, width, modes, layers, seed=s).to(device) for s in seeds]
        ns = len(base["stations"])
        self.heads = [StationHead(width, NCH, base["nt"], hidden, seed=s).to(device) for s in seeds] if ns else []
        i0, j0, _ = base["window"]
        st = torch.from_numpy(base["stations"]).to(device)
        self.si, self.sj = (st[:, 0] - i0) if ns else None, (st[:, 1] - j0) if ns else None
        self.device = device
        self.mu = self.sd = self.tsd = None
        self.threshold = None
        # error indicator: "l2" = relative L2 distance of the two networks'
        # PGV maps; "log" = median |difference| of their log10 PGV
        self.indicator = "l2"
        # test-time averaging over the symmetric images of each query
        self.tta = False

    def _outputs(self, k, x):
        v, vb = self.nets[k].trunk(x)
        maps = self.nets[k].head(v, vb=vb)
        return maps, (self.heads[k](v, x, self.si, self.sj) if self.heads else None)

    def fit(self, params, pgv, epochs=60, batch=16, lr=3e-3, log=None, augment=False, traces=None):
        """Train both networks (and their seismogram heads, on `traces`
        [n, S, 3, T] when given) on parameter rows and their FD outputs;
        with augment=True every run is also used in its 3 images under
        SYM.  The loss is the mean squared error of the normalised
        log10(PGV/M0) maps plus that of the traces / M0 scaled to unit
        variance."""
        src = source_rows(params)
        if traces is not None and not self.heads:
            raise SurrogateError("traces given, but the base medium has no stations")
        if augment:
            pairs = [transform(self.base, src, pgv, g) for g in SYM]
            if traces is not None:
                traces = np.concatenate([transform_traces(self.base, traces, g) for g in SYM])
            src = np.concatenate([p[0] for p in pairs])
            pgv = np.concatenate([p[1] for p in pairs])
        x = torch.from_numpy(encode(self.base, src)).to(self.device)
        y = target(pgv, src)
        self.mu, self.sd = float(y.mean()), float(y.std())
        y = torch.from_numpy((y - self.mu) / self.sd).to(self.device)
        yt = None
        if traces is not None:
            tm = align_traces(self.base, src, traces) / (10.0 ** src[:, 9])[:, None, None, None]
            self.tsd = float(np.sqrt(np.mean(np.square(tm, dtype=np.float64))))
            yt = torch.from_numpy((tm / self.tsd).astype(np.float32)).to(self.device)
        losses = []
        for k, net in enumerate(self.nets):
            params_k = list(net.parameters()) + (list(self.heads[k].parameters()) if yt is not None else [])
            opt = Adam(params_k, lr=lr)
            g = torch.Generator().manual_seed(100 + k)
            for ep in range(epochs):
                # the epoch's order on the device and its loss read once per
                # epoch, so the steps queue without waiting for each other
                perm = torch.randperm(len(x), generator=g).to(self.device)
                tot = torch.zeros((), device=self.device)
                for b0 in range(0, len(x), batch):
                    idx = perm[b0:b0 + batch]
                    xb = x[idx]
                    if yt is None:
                        loss = torch.mean((net(xb) - y[idx]) ** 2)
                    else:
                        maps, tr = self._outputs(k, xb)
                        loss = torch.mean((maps - y[idx]) ** 2) + torch.mean((tr - yt[idx]) ** 2)
                    opt.zero_grad()
                    loss.backward()
                    opt.step(cosine_lr(lr, ep, epochs))
                    tot += loss.detach() * len(idx)
                losses.append(float(tot) / len(x))
                if log and (ep % 10 == 0 or ep == epochs - 1):
                    log(f"net {k} epoch {ep}: loss {losses[-1]:.4f}")
        return losses

    @torch.no_grad()
    def _raw(self, src, batch):
        """Each network's normalised log10(PGV/M0) maps [n, w, w, 3] and
        (with heads) its normalised traces [n, S, 3, T] for source rows;
        with self.tta, averaged over the SYM images of each source (mapped
        back: every element of SYM is its own inverse)."""
        images = SYM if self.tta else [(0, 0, 0)]
        outs = [0.0 for _ in self.nets]
        touts = [0.0 for _ in self.nets]
        for g in images:
            sg, _ = transform(self.base, src, None, g)
            x = torch.from_numpy(encode(self.base, sg)).to(self.device)
            for k in range(len(self.nets)):
                parts = [self._outputs(k, x[b:b + batch]) for b in range(0, len(x), batch)]
                o = torch.cat([p[0] for p in parts])
                t = torch.cat([p[1] for p in parts]) if self.heads else None
                if g != (0, 0, 0):
                    o = o.permute(0, 3, 1, 2).cpu().numpy()
                    o = torch.from_numpy(transform(self.base, sg, o, g)[1]).to(self.device).permute(0, 2, 3, 1)
                    if t is not None:
                        t = torch.from_numpy(transform_traces(self.base, t.cpu().numpy(), g)).to(self.device)
                outs[k] = outs[k] + o / len(images)
                if t is not None:
                    touts[k] = touts[k] + t / len(images)
        return outs, (touts if self.heads else None)

    @torch.no_grad()
    def predict(self, params, batch=32, rows=False, traces=False):
        """PGV maps [n, 3, w, w] (the ensemble mean) and the indicator [n]
        for parameter rows (or source rows with rows=True); with
        traces=True also the mean station traces [n, S, 3, T]."""
        src = np.asarray(params, np.float64) if rows else source_rows(params)
        outs, touts = self._raw(src, batch)
        x = outs[0]
        m0 = torch.from_numpy(10.0 ** src[:, 9]).to(self.device).float()
        pgv = [10.0 ** (o * self.sd + self.mu) * m0[:, None, None, None] for o in outs]
        mean = sum(pgv) / len(pgv)
        if self.indicator == "log":
            ind = torch.median((outs[0] - outs[1]).abs().reshape(len(x), -1), dim=1).values * self.sd
        else:
            d = (pgv[0] - pgv[1]).reshape(len(x), -1)
            ind = torch.linalg.vector_norm(d, dim=1) / torch.linalg.vector_norm(mean.reshape(len(x), -1), dim=1)
        res = (mean.permute(0, 3, 1, 2).cpu().numpy(), ind.cpu().numpy())
        if traces:
            if touts is None or self.tsd is None:
                raise SurrogateError("no trained seismogram head")
            tr = sum(touts) / len(touts) * self.tsd * m0[:, None, None, None]
            res = res + (align_traces(self.base, src, tr.cpu().numpy(), inverse=True),)
        return res

    def calibrate(self, params, pgv, target_err=0.05, quantile=0.5):
        """Largest indicator threshold whose accepted scenarios keep the
        `quantile` of their PGV errors (the median by default) at or below
        target_err on a calibration set."""
        pred, ind = self.predict(params)
        err = pgv_error(pred, pgv)
        order = np.argsort(ind)
        best = 0.0
        for n in range(1, len(order) + 1):
            if np.quantile(err[order[:n]], quantile) <= target_err:
                best = float(ind[order[n - 1]])
        self.threshold = best
        return best

    def answer(self, params, fd_group=32):
        """Gated answers: surrogate PGV where the indicator is below the
        threshold, FD runs for the rest.  Returns (pgv, accepted mask)."""
        pred, ind = self.predict(params)
        ok = ind <= self.threshold
        if (~ok).any():
            pred[~ok] = run_fd(self.base, params[~ok], group=fd_group)
        return pred, ok


def timed(fn):
    t0 = time.perf_counter()
    r = fn()
    torch.cuda.synchronize()
    return r, time.perf_counter() - t0
