"""The differentiable-degridding workload at its stated size (diff1024:
1024^2 image of 50 Gaussian blobs, 1e6 visibilities with w in +-500,
w-kernels per 20-wavelength w-bin with supports 16..32, Q = 8, far field 64
pixels): the forward model and the gradient against the FP64 C oracle on
all 1e6 visibilities, the dot-product adjoint test, central finite
differences on 64 pixels, and 200 gradient-descent steps (the operator the
bench's diff1024 workload times)."""
import numpy
import pytest
import torch

from crocodile_amd import workloads as W
from crocodile_amd.autograd import WProjectionOperator, degrid_image
from oracle import cgrid
from oracle import synthesis_ref as R

from test_gpu_core import assert_close

pytestmark = pytest.mark.gpu

C = W.SPEC['diff1024']
N, THETA, WSTEP, FF, Q = C['N'], C['theta'], 20.0, 64, C['Q']
LAM = N / THETA


def support(w):
    return W.wproj_support(w, C['wmax'], 16, 32)


@pytest.fixture(scope='module')
def problem(gpu):
    print('generating diff1024 (1e6 visibilities, 1024^2)', flush=True)
    image, uvw = W.diff1024()
    op = WProjectionOperator(THETA, LAM, uvw, WSTEP, FF, support, Q)
    wb = numpy.round(uvw[:, 2] / WSTEP).astype(int)
    b0 = wb.min()
    tables = [R.w_kernel(THETA, WSTEP * b, FF, support(WSTEP * b), Q, invert=False) / FF ** 2
              for b in range(b0, wb.max() + 1)]
    ref = dict(bank=cgrid.Bank(tables), vplane=(wb - b0).astype(numpy.int32), p=uvw / LAM)
    return image, uvw, op, ref


def np_forward(ref, image):
    return cgrid.convdegrid(ref['bank'], R.fft(image.astype(complex)), ref['p'], ref['vplane'])


def np_gradient(ref, image, vobs):
    """d/dI ||F I - vobs||^2 = 2 Re(F^H (F I - vobs)), F^H = N^2 ifft conj grid conj."""
    res = np_forward(ref, image) - vobs
    acc = numpy.zeros((N, N), dtype=complex)
    cgrid.convgrid(ref['bank'], acc, ref['p'], numpy.conj(res), ref['vplane'])
    return 2 * numpy.real(R.ifft(numpy.conj(acc)) * N * N)


def test_c4_forward_and_adjoint(problem):
    image, uvw, op, ref = problem
    assert op.nv == 1_000_000 and N == 1024
    assert min(op.bank.plane_s_h) == 16 and max(op.bank.plane_s_h) == 32
    v = degrid_image(torch.as_tensor(image, device='cuda'), op)
    assert_close(v.cpu().numpy(), np_forward(ref, image), what='C4 forward')
    rng = numpy.random.default_rng(1)
    x = torch.as_tensor(rng.standard_normal((N, N)).astype(numpy.float32), device='cuda')
    g = torch.as_tensor(rng.standard_normal(op.nv) + 1j * rng.standard_normal(op.nv),
                        device='cuda').to(torch.complex64)
    lhs = torch.sum(torch.conj(g).to(torch.complex128) * op.apply(x).to(torch.complex128)).real
    rhs = torch.sum(x.to(torch.float64) * op.adjoint(g).real.to(torch.float64))
    rel = abs(float(lhs - rhs)) / abs(float(lhs))
    print('C4 adjoint dot product rel %.3e' % rel, flush=True)
    assert rel < 1e-5


def test_c4_gradient_vs_oracle_and_fd(problem):
    image, uvw, op, ref = problem
    rng = numpy.random.default_rng(2)
    vobs = np_forward(ref, image)
    x0 = (image * rng.uniform(0.5, 1.5, image.shape)).astype(numpy.float32)
    x = torch.as_tensor(x0, device='cuda').requires_grad_(True)
    vo = torch.as_tensor(vobs, device='cuda').to(torch.complex64)
    loss = torch.sum(torch.abs(degrid_image(x, op) - vo) ** 2)
    loss.backward()
    grad = x.grad.cpu().numpy()
    assert_close(grad, np_gradient(ref, x0, vobs), what='C4 gradient')
    h = 0.25
    pix = rng.integers(0, N, (64, 2))

    def L(img):
        v = op.apply(torch.as_tensor(img, device='cuda')).cpu().numpy().astype(complex)
        return float(numpy.sum(numpy.abs(v - vobs) ** 2))

    errs = []
    for py, px in pix:
        xp, xm = x0.copy(), x0.copy()
        xp[py, px] += h
        xm[py, px] -= h
        errs.append(abs((L(xp) - L(xm)) / (2 * h) - grad[py, px]) / abs(grad).max())
    print('C4 finite differences: max error / max|grad| = %.3e' % max(errs), flush=True)
    assert max(errs) < 1e-3


def test_c4_gradient_descent_200_steps(problem):
    image, uvw, op, ref = problem
    vobs = np_forward(ref, image)
    vo = torch.as_tensor(vobs, device='cuda').to(torch.complex64)
    x = torch.zeros((N, N), dtype=torch.float32, device='cuda', requires_grad=True)
    v = torch.randn(N, N, device='cuda')
    for _ in range(20):
        v = op.adjoint(op.apply(v)).real
        v = v / torch.linalg.norm(v)
    lip = float(torch.sum(v * op.adjoint(op.apply(v)).real))
    eta = 1.0 / (2 * lip)
    losses = []
    for step in range(200):
        loss = torch.sum(torch.abs(degrid_image(x, op) - vo) ** 2)
        x.grad = None
        loss.backward()
        if step in (0, 199):
            assert_close(x.grad.cpu().numpy(), np_gradient(ref, x.detach().cpu().numpy(), vobs),
                         rms_tol=1e-4, max_tol=1e-3, what='C4 gradient at step %d' % step)
        losses.append(float(loss.detach()))
        with torch.no_grad():
            x -= eta * x.grad
    print('C4 loss %.4e -> %.4e' % (losses[0], losses[-1]), flush=True)
    assert all(b <= a * (1 + 1e-6) for a, b in zip(losses, losses[1:]))
    assert losses[-1] < 0.2 * losses[0]
