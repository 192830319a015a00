"""fp32 PyTorch restatement of the dispatcher MLP, run on the CPU in a child
interpreter by tests/test_gpu_dispatcher.py (so torch's own HIP runtime never
shares a process with libstc.so).

    python tests/torch_mlp_ref.py train IN.npz OUT.npz
    python tests/torch_mlp_ref.py forward IN.npz OUT.npz

train: the loss the device training kernel minimises -- over samples i,
w_i * sum_s softmax(z_i)_s * r_is, z = W2 relu(W1 x + b1) + b2 -- with
torch.optim.Adam (betas 0.9 / 0.999, eps 1e-8), full batch, autograd for the
gradients; IN holds x (nsamp x 7), r (nsamp x 4), w (nsamp), the flat initial
parameters (w1 16x7, b1 16, w2 4x16, b2 4), iters and lr.

forward: the logits of the model (mean, inv_std, w1, b1, w2, b2) on raw
features f (nbucket x 7): x = (f - mean) * inv_std.
"""
import sys

import numpy as np
import torch

H, I, O = 16, 7, 4


def unpack(p):
    p = torch.as_tensor(p, dtype=torch.float32)
    w1 = p[: H * I].reshape(H, I)
    b1 = p[H * I: H * I + H]
    w2 = p[H * I + H: H * I + H + O * H].reshape(O, H)
    b2 = p[H * I + H + O * H:]
    return w1, b1, w2, b2


def train(z):
    torch.set_num_threads(1)
    x = torch.as_tensor(z["x"], dtype=torch.float32)
    r = torch.as_tensor(z["r"], dtype=torch.float32)
    w = torch.as_tensor(z["w"], dtype=torch.float32)
    params = [t.clone().requires_grad_(True) for t in unpack(z["params"])]
    opt = torch.optim.Adam(params, lr=float(z["lr"]), betas=(0.9, 0.999), eps=1e-8)
    for _ in range(int(z["iters"])):
        opt.zero_grad()
        w1, b1, w2, b2 = params
        h = torch.relu(x @ w1.T + b1)
        logits = h @ w2.T + b2
        p = torch.softmax(logits, dim=1)
        loss = (w * (p * r).sum(dim=1)).sum()
        loss.backward()
        opt.step()
    return {k: t.detach().numpy() for k, t in zip(("w1", "b1", "w2", "b2"), params)}


def forward(z):
    f = torch.as_tensor(z["f"], dtype=torch.float32)
    x = (f - torch.as_tensor(z["mean"])) * torch.as_tensor(z["inv_std"])
    h = torch.relu(x @ torch.as_tensor(z["w1"]).T + torch.as_tensor(z["b1"]))
    logits = h @ torch.as_tensor(z["w2"]).T + torch.as_tensor(z["b2"])
    return {"logits": logits.numpy()}


if __name__ == "__main__":
    mode, src, dst = sys.argv[1:4]
    with np.load(src) as z:
        out = train(z) if mode == "train" else forward(z)
    np.savez(dst, **out)
