"""fp32 PyTorch restatement of the learned denoiser, the reference the bf16
MFMA kernels of libpt are checked against (forward values and parameter
gradients).  It reads the same flat parameter array pt_unet_get_params
returns: per layer the weights [9][cout][cin] then the biases [cout]."""
import numpy as np
import torch
import torch.nn.functional as F

C = (32, 48, 64, 96, 128)
IN_C, OUT_C, FEAT = 16, 16, 10


def layer_table():
    """(cin, cout, relu) of the 19 convolutions in network order."""
    t = []
    cin = IN_C
    for c in C[:4]:  # encoder
        t += [(cin, c, True), (c, c, True)]
        cin = c
    t += [(C[3], C[4], True), (C[4], C[4], True)]  # bottleneck
    for lvl in (3, 2, 1, 0):  # decoder: [upsampled | skip]
        t += [(C[lvl + 1] + C[lvl], C[lvl], True), (C[lvl], C[lvl], True)]
    t.append((C[0], OUT_C, False))
    return t


def param_count():
    return sum(9 * co * ci + co for ci, co, _ in layer_table())


def unflatten(flat):
    """Flat fp32 array -> list of (weight [cout][cin][3][3], bias [cout]) tensors."""
    flat = torch.as_tensor(np.asarray(flat, np.float32))
    out, o = [], 0
    for ci, co, _ in layer_table():
        w = flat[o:o + 9 * co * ci].reshape(3, 3, co, ci).permute(2, 3, 0, 1).contiguous()
        o += 9 * co * ci
        b = flat[o:o + co].clone()
        o += co
        out.append((w, b))
    assert o == flat.numel()
    return out


def flatten_grads(params):
    parts = []
    for w, b in params:
        parts.append(w.grad.permute(2, 3, 0, 1).reshape(-1))
        parts.append(b.grad.reshape(-1))
    return torch.cat(parts).numpy()


def pack_input(feat):
    """[n][h][w][10] features -> [n][16][h][w] network input."""
    f = torch.as_tensor(np.asarray(feat, np.float32))
    x = torch.zeros(f.shape[:3] + (IN_C,), dtype=torch.float32)
    x[..., 0:3] = torch.log1p(torch.clamp(f[..., 0:3], min=0.0))
    x[..., 3:9] = f[..., 3:9]
    d = torch.clamp(f[..., 9], min=0.0)
    x[..., 9] = d / (d + 1.0)
    return x.permute(0, 3, 1, 2).contiguous()


def forward(params, x):
    """params: unflatten() output; x: [n][16][h][w]; returns [n][3][h][w] (log space)."""
    table = layer_table()
    li = 0

    def conv(t):
        nonlocal li
        w, b = params[li]
        relu = table[li][2]
        li += 1
        y = F.conv2d(t, w, b, padding=1)
        return F.relu(y) if relu else y

    skips = []
    for _ in range(4):
        x = conv(conv(x))
        skips.append(x)
        x = F.avg_pool2d(x, 2)
    x = conv(conv(x))
    for lvl in (3, 2, 1, 0):
        up = x.repeat_interleave(2, dim=2).repeat_interleave(2, dim=3)
        x = conv(conv(torch.cat([up, skips[lvl]], dim=1)))
    return conv(x)[:, :3]


def loss_and_grads(flat, feat, target):
    """L1 loss on log(1 + x) (mean over pixels and channels) and its gradient
    as a flat array in the same order as the parameters."""
    params = unflatten(flat)
    for w, b in params:
        w.requires_grad_(True)
        b.requires_grad_(True)
    pred = forward(params, pack_input(feat))
    t = torch.as_tensor(np.asarray(target, np.float32)).permute(0, 3, 1, 2)
    loss = (pred - torch.log1p(torch.clamp(t, min=0.0))).abs().mean()
    loss.backward()
    return float(loss.detach()), flatten_grads(params)


def layer_slices():
    """(name, weight slice, bias slice) of each layer in the flat array."""
    names = ["e0a", "e0b", "e1a", "e1b", "e2a", "e2b", "e3a", "e3b", "b4a", "b4b",
             "d3a", "d3b", "d2a", "d2b", "d1a", "d1b", "d0a", "d0b", "out"]
    out, o = [], 0
    for name, (ci, co, _) in zip(names, layer_table()):
        out.append((name, slice(o, o + 9 * co * ci), slice(o + 9 * co * ci, o + 9 * co * ci + co)))
        o += 9 * co * ci + co
    return out


def _bf16(t):
    return t.to(torch.bfloat16).float()


def _e4m3(t):
    """Quantize to OCP e4m3 (round to nearest even, saturating at +-448) and back."""
    return torch.clamp(t, -448.0, 448.0).to(torch.float8_e4m3fn).float()


def forward_fp8(params, feat, act_scales, w_scales, w8):
    """The fp8 inference path as the device runs it: every inner convolution's
    input (bf16) quantized per channel as e4m3(clamp(x * (1 / s_a), +-448)), the
    e4m3 weights (bytes, [9][cout][cin] per layer), the fp32 product times
    s_w[co] plus the bias, ReLU, stored as bf16 (pooling averages bf16 values
    into bf16); the first, the last decoder and the output convolution run on
    bf16 weights and the last layer's output stays fp32.  params: unflatten()
    output (for the biases); returns [n][3][h][w] (log space)."""
    table = layer_table()
    w8 = torch.from_numpy(np.asarray(w8, np.uint8).copy())
    sa = torch.as_tensor(np.asarray(act_scales, np.float32))
    sw = torch.as_tensor(np.asarray(w_scales, np.float32))
    offs = [0, 0, 0]
    li = 0

    def conv(t):
        nonlocal li
        ci, co, relu = table[li]
        inv = 1.0 / sa[offs[0]:offs[0] + ci]
        wq = w8[offs[2]:offs[2] + 9 * co * ci].view(torch.float8_e4m3fn).float()
        wq = wq.reshape(3, 3, co, ci).permute(2, 3, 0, 1).contiguous()
        s = sw[offs[1]:offs[1] + co]
        offs[0] += ci
        offs[1] += co
        offs[2] += 9 * co * ci
        last = li == len(table) - 1
        if li in (0, len(table) - 2) or last:  # the first, the last decoder and the output convolution stay bf16
            y = F.conv2d(t, _bf16(params[li][0]), params[li][1], padding=1)
        else:
            xq = _e4m3(t * inv[None, :, None, None])
            y = F.conv2d(xq, wq, None, padding=1) * s[None, :, None, None] + params[li][1][None, :, None, None]
        li += 1
        if relu:
            y = F.relu(y)
        return y if last else _bf16(y)

    x = _bf16(pack_input(feat))
    skips = []
    for _ in range(4):
        x = conv(conv(x))
        skips.append(x)
        x = _bf16(F.avg_pool2d(x, 2))
    x = conv(conv(x))
    for lvl in (3, 2, 1, 0):
        up = x.repeat_interleave(2, dim=2).repeat_interleave(2, dim=3)
        x = conv(conv(torch.cat([up, skips[lvl]], dim=1)))
    return conv(x)[:, :3]


def fp8_sizes():
    """(input scales, output scales, weight bytes) of the fp8 state."""
    t = layer_table()
    return sum(ci for ci, _, _ in t), sum(co for _, co, _ in t), sum(9 * ci * co for ci, co, _ in t)
