"""The surrogate network's matrix-core kernels against the CPU (FP32 torch).

* pywrf_mlp_gemm, every epilogue, at shapes that leave partial tiles in M
  and N and with split-K: equal to the FP32 product of the same BF16 values
  (to FP32 summation order; BF16 outputs to one BF16 rounding);
* one training step (forward, loss, backward: all parameter gradients) equal
  to torch autograd on the CPU in FP32 of the same network, with the same
  BF16 roundings of the operands the kernels round (weights, inputs, hidden
  activations, output and hidden gradients);
* Adam equal to torch.optim.Adam.
"""
import pytest
import torch

pytestmark = pytest.mark.gpu


def _bf(x):
    return x.to(torch.bfloat16)


def _rel(a, b):
    a, b = a.double().cpu(), b.double().cpu()
    return float((a - b).abs().max() / max(b.abs().max(), 1e-30))


@pytest.mark.parametrize("M,N,K,ksplit", [(256, 128, 64, 1), (300, 200, 96, 1), (1000, 288, 512, 1),
                                          (96, 512, 4096, 8), (51529 % 4096 + 7, 384, 384, 1),
                                          (700, 512, 288, 1), (512, 288, 8192, 16)])
def test_gemm_matches_cpu(M, N, K, ksplit):
    from pywrf_amd import surrogate as sg
    gen = torch.Generator().manual_seed(M * 7 + N)
    A = _bf(torch.randn(M, K, generator=gen))
    B = _bf(torch.randn(N, K, generator=gen))
    bias = torch.randn(N, generator=gen)
    ref = A.float() @ B.float().t()
    Ad, Bd, biasd = A.cuda(), B.cuda(), bias.cuda()
    if ksplit == 1:
        C = torch.full((M, N), 7.0, device="cuda")
        sg.gemm(0, Ad, Bd, C, biasd)
        assert _rel(C, ref + bias) < 1e-5
        Cb = torch.zeros(M, N, dtype=torch.bfloat16, device="cuda")
        sg.gemm(1, Ad, Bd, Cb, biasd)
        want = torch.relu(ref + bias)
        assert _rel(Cb.float(), want) < 1e-2
        # one BF16 rounding: |Cb - want| <= 2^-8 |want| (+ summation-order slack)
        assert bool(((Cb.float().cpu() - want).abs() <= 2 ** -8 * want.abs() + 1e-3 * want.abs().max()).all())
        H = _bf(torch.randn(M, N, generator=gen))
        Cd = torch.zeros(M, N, dtype=torch.bfloat16, device="cuda")
        sg.gemm(2, Ad, Bd, Cd, H=H.cuda())
        want = ref * (H.float() > 0)
        assert _rel(Cd.float(), want) < 1e-2
        Ct = torch.full((N, M + 5), 3.0, device="cuda")
        sg.gemm(4, Ad, Bd, Ct, biasd)
        assert _rel(Ct[:, :M].t(), ref + bias) < 1e-5
        assert bool((Ct[:, M:] == 3.0).all())
    C = torch.zeros(M, N, device="cuda")
    sg.gemm(3, Ad, Bd, C, ksplit=ksplit)
    assert _rel(C, ref) < 1e-5


def test_transpose_and_rowsum():
    from pywrf_amd import surrogate as sg
    x = _bf(torch.randn(300, 200)).cuda()
    t = torch.zeros(200, 300, dtype=torch.bfloat16, device="cuda")
    sg.transpose(x, t)
    assert torch.equal(t.cpu(), x.cpu().t())
    r = torch.zeros(300, device="cuda")
    sg.rowsum(x, r)
    assert _rel(r, x.float().cpu().sum(1)) < 1e-5


def _cpu_reference_step(net, X, T, w):
    """Forward / backward of the network in FP32 torch autograd on the CPU,
    rounding to BF16 where the kernels store BF16: weights, inputs, the
    hidden activations after ReLU, the output gradient and the hidden
    gradients."""
    W = [torch.nn.Parameter(_bf(x.cpu()).float()) for x in net.W]
    b = [torch.nn.Parameter(x.cpu().clone()) for x in net.b]

    class Round(torch.autograd.Function):
        @staticmethod
        def forward(ctx, x):
            return _bf(x).float()

        @staticmethod
        def backward(ctx, g):
            return _bf(g).float()

    x = X.cpu().float()
    h1 = Round.apply(torch.relu(x @ W[0].t() + b[0]))
    h2 = Round.apply(torch.relu(h1 @ W[1].t() + b[1]))
    y = h2 @ W[2].t() + b[2]
    wc = w.cpu()
    M = x.shape[0]
    loss = ((y - T.cpu()) ** 2 * wc).sum() / (M * float(wc.sum()))
    y.register_hook(lambda g: _bf(g).float())
    loss.backward()
    return float(loss), [p.grad for p in W], [p.grad for p in b]


def test_training_step_gradients_match_cpu_autograd():
    from pywrf_amd import surrogate as sg
    M = 512
    gen = torch.Generator().manual_seed(3)
    net = sg.MLP(seed=5)
    for n in range(3):                       # non-zero biases
        net.b[n].copy_(0.1 * torch.randn(net.b[n].shape[0], generator=gen))
    net.b[2][sg.FOUT:] = 0.0
    X = torch.zeros(M, sg.KIN, dtype=torch.bfloat16)
    X[:, :sg.FIN] = _bf(torch.randn(M, sg.FIN, generator=gen))
    T = torch.zeros(M, sg.NO)
    T[:, :sg.FOUT] = torch.randn(M, sg.FOUT, generator=gen)
    w = sg.loss_weights("cpu")
    ws = sg.Workspace(M)
    Xd, Td, wd = X.cuda(), T.cuda(), w.cuda()
    sg.forward(net, Xd, ws)
    loss = float(sg.backward(net, Xd, Td, wd, ws, float(w.sum()), ksplit=4)[0])
    ref_loss, gW, gb = _cpu_reference_step(net, X, T, w)
    print("loss", loss, ref_loss)
    assert abs(loss - ref_loss) <= 1e-4 * ref_loss
    for n in range(3):
        eW, eb = _rel(ws.gW[n], gW[n]), _rel(ws.gb[n], gb[n])
        print("layer", n + 1, "dW %.2e  db %.2e" % (eW, eb))
        assert eW < 2e-3 and eb < 2e-3
    # padded input columns and the padded output row get zero gradient
    assert float(ws.gW[0][:, sg.FIN:].abs().max()) == 0.0
    assert float(ws.gW[2][sg.FOUT:].abs().max()) == 0.0


def test_adam_matches_torch():
    from pywrf_amd import surrogate as sg
    net = sg.MLP(seed=1)
    ws = sg.Workspace(32)
    p = [torch.nn.Parameter(x.cpu().clone()) for x in net.W + net.b]
    opt = torch.optim.Adam(p, lr=1e-3)
    gen = torch.Generator().manual_seed(0)
    for _ in range(3):
        grads = [torch.randn(x.shape, generator=gen) for x in p]
        for q, g in zip(p, grads):
            q.grad = g.clone()
        opt.step()
        for n in range(3):
            ws.gW[n].copy_(grads[n])
            ws.gb[n].copy_(grads[3 + n])
        sg.adam_step(net, ws, 1e-3)
    for mine, ref in zip(net.W + net.b, p):
        assert _rel(mine, ref.detach()) < 1e-5
    assert torch.equal(net.Wb[1].cpu(), _bf(net.W[1].cpu()))
    assert torch.equal(net.Wt[2].cpu(), _bf(net.W[2].cpu()).t())
