"""FP32 PyTorch restatement of the denoiser network (the checker of the HIP
kernels): the same layers, parameter layout and loss as
convector_amd/csrc/denoise.hip, with BF16 rounding applied where the GPU
stores BF16 (weights, the input and every hidden activation) when bf16=True.
Runs on the CPU."""
import numpy as np
import torch
import torch.nn.functional as F

# (cin, cout, relu): enc0a enc0b enc1a enc1b bot2a bot2b dec1a dec1b dec0a dec0b out
LAYERS = [(16, 32, 1), (32, 32, 1), (32, 64, 1), (64, 64, 1), (64, 128, 1), (128, 128, 1),
          (192, 64, 1), (64, 64, 1), (96, 32, 1), (32, 32, 1), (32, 3, 0)]


def num_params():
    return sum(co * 9 * ci + co for ci, co, _ in LAYERS)


def unpack(params: np.ndarray):
    """Flat parameters -> [(W [cout][cin][3][3], b [cout])] as torch tensors
    (the flat layout is [cout][ky][kx][cin] then the bias, per layer)."""
    out, off = [], 0
    for ci, co, _ in LAYERS:
        w = params[off:off + co * 9 * ci].reshape(co, 3, 3, ci)
        off += co * 9 * ci
        b = params[off:off + co]
        off += co
        out.append((torch.tensor(np.ascontiguousarray(w.transpose(0, 3, 1, 2))),
                    torch.tensor(np.ascontiguousarray(b))))
    return out


def _bf(x, on):
    return x.to(torch.bfloat16).to(torch.float32) if on else x


def forward(inp: torch.Tensor, wb, bf16=True):
    """inp: (nb, h, w, 10) -> (nb, h, w, 3) log1p radiance."""
    x = inp.permute(0, 3, 1, 2)
    x = torch.cat([x, torch.zeros(x.shape[0], 6, x.shape[2], x.shape[3])], 1)
    x0 = _bf(x, bf16)

    def conv(t, k, relu=True):
        w, b = wb[k]
        y = F.conv2d(t, _bf(w, bf16), b, padding=1)
        return _bf(F.relu(y), bf16) if relu else y

    e0a = conv(x0, 0)
    e0b = conv(e0a, 1)
    p0 = _bf(F.avg_pool2d(e0b, 2), bf16)
    e1a = conv(p0, 2)
    e1b = conv(e1a, 3)
    p1 = _bf(F.avg_pool2d(e1b, 2), bf16)
    b2a = conv(p1, 4)
    b2b = conv(b2a, 5)
    cat1 = torch.cat([F.interpolate(b2b, scale_factor=2, mode="nearest"), e1b], 1)
    d1a = conv(cat1, 6)
    d1b = conv(d1a, 7)
    cat0 = torch.cat([F.interpolate(d1b, scale_factor=2, mode="nearest"), e0b], 1)
    d0a = conv(cat0, 8)
    d0b = conv(d0a, 9)
    out = conv(d0b, 10, relu=False)
    return out.permute(0, 2, 3, 1)


def loss_and_grads(params: np.ndarray, inp: np.ndarray, target: np.ndarray, bf16=True):
    """Mean |out - log1p(max(target, 0))| and its gradient w.r.t. the flat
    parameters (the GPU's layout)."""
    wb = unpack(params)
    for w, b in wb:
        w.requires_grad_(True)
        b.requires_grad_(True)
    out = forward(torch.tensor(inp), wb, bf16)
    tgt = torch.log1p(torch.clamp(torch.tensor(target), min=0.0))
    loss = (out - tgt).abs().mean()
    loss.backward()
    g = []
    for w, b in wb:
        g.append(w.grad.numpy().transpose(0, 2, 3, 1).ravel())
        g.append(b.grad.numpy().ravel())
    return float(loss), np.concatenate(g).astype(np.float32), out.detach().numpy()


def layer_slices():
    """Slices of each layer's weights and bias in the flat vector."""
    sl, off = [], 0
    for ci, co, _ in LAYERS:
        sl.append((slice(off, off + co * 9 * ci), slice(off + co * 9 * ci, off + co * 9 * ci + co)))
        off += co * 9 * ci + co
    return sl
