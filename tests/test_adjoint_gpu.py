"""The transposes of the transforms (zgl_backward_adjoint, zgl_forward_adjoint)
and the autograd test binding built on them.

Pins, independent of the adjoint formula:
  * dense transposes: at small orders B and F are assembled column by column
    from the oracle (oracle/zgl_oracle.c, a restatement of zest's
    zt::/st::GLQTransformer) and B^T y, F^T d are compared with the GPU;
  * the inner-product identities <B c, y> = <c, B^T y>, <F g, d> = <g, F^T d>
    at larger orders;
  * gradients of a weighted least-squares loss (the differentiation workload:
    fits of g = r exp(-(a.x)^2) on the ball grid at order 64) against central
    finite differences, and against the closed form 2 diag(1/nu) F(B c - g)
    evaluated with compiled zest on the CPU (the oracle when zest is absent).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ZT_CONV = [(zn, sn, sp) for zn in (0, 1) for sn in (0, 1) for sp in (-1, 1)]
ST_CONV = [(sn, sp, lay) for sn in (0, 1) for sp in (-1, 1) for lay in (0, 1)]


def _rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def _dense_zt(order, zn, sn, sp):
    """B (grid x coeff) and F (coeff x grid) as dense matrices from the oracle,
    over the real coefficient array (ncoeff * 2)."""
    import oracle
    nc = oracle.zt_num_coeffs(order) * 2
    ng = int(np.prod(oracle.zt_shape(order)))
    B = np.zeros((ng, nc))
    for i in range(nc):
        e = np.zeros(nc)
        e[i] = 1.0
        B[:, i] = oracle.zt_backward(e.reshape(-1, 2), order, zernike_norm=zn, sh_norm=sn,
                                     sh_phase=sp).ravel()
    F = np.zeros((nc, ng))
    for i in range(ng):
        e = np.zeros(ng)
        e[i] = 1.0
        F[:, i] = oracle.zt_forward(e.reshape(oracle.zt_shape(order)), order, zernike_norm=zn,
                                    sh_norm=sn, sh_phase=sp).ravel()
    return B, F


def _dense_st(order, sn, sp, layout):
    import oracle
    nc = oracle.st_num_coeffs(order) * 2
    shape = oracle.st_shape(order, layout)
    ng = int(np.prod(shape))
    B = np.zeros((ng, nc))
    for i in range(nc):
        e = np.zeros(nc)
        e[i] = 1.0
        B[:, i] = oracle.st_backward(e.reshape(-1, 2), order, sh_norm=sn, sh_phase=sp,
                                     layout=layout).ravel()
    F = np.zeros((nc, ng))
    for i in range(ng):
        e = np.zeros(ng)
        e[i] = 1.0
        F[:, i] = oracle.st_forward(e.reshape(shape), order, sh_norm=sn, sh_phase=sp,
                                    layout=layout).ravel()
    return B, F


@pytest.mark.parametrize("order", [1, 3, 5])
@pytest.mark.parametrize("conv", ZT_CONV)
def test_zernike_adjoints_equal_dense_transposes(device, order, conv):
    import torch
    import zest_amd
    zn, sn, sp = conv
    B, F = _dense_zt(order, zn, sn, sp)
    rng = np.random.default_rng(order * 17 + zn * 4 + sn * 2 + (sp > 0))
    batch = 2
    y = rng.standard_normal((batch, B.shape[0]))
    d = rng.standard_normal((batch, B.shape[1]))
    plan = zest_amd.Plan(zest_amd.ZERNIKE, order, zernike_norm=zn, sh_norm=sn, sh_phase=sp)
    shape = zest_amd.zt_shape(order)
    dy = torch.from_numpy(y.reshape((batch,) + shape)).to(device)
    bt = torch.zeros((batch, zest_amd.zt_num_coeffs(order), 2), dtype=torch.float64, device=device)
    plan.backward_adjoint(dy, bt, order, batch=batch)
    dd = torch.from_numpy(d.reshape(batch, -1, 2)).to(device)
    ft = torch.empty((batch,) + shape, dtype=torch.float64, device=device)
    plan.forward_adjoint(dd, ft, order, batch=batch)
    torch.cuda.synchronize()
    for b in range(batch):
        assert _rel(bt[b].cpu().numpy().ravel(), B.T @ y[b]) < 1e-12
        assert _rel(ft[b].cpu().numpy().ravel(), F.T @ d[b]) < 1e-12


@pytest.mark.parametrize("order", [2, 5])
@pytest.mark.parametrize("conv", ST_CONV)
def test_sh_adjoints_equal_dense_transposes(device, order, conv):
    import torch
    import zest_amd
    sn, sp, layout = conv
    B, F = _dense_st(order, sn, sp, layout)
    rng = np.random.default_rng(order * 13 + sn * 4 + (sp > 0) * 2 + layout)
    y = rng.standard_normal(B.shape[0])
    d = rng.standard_normal(B.shape[1])
    plan = zest_amd.Plan(zest_amd.SH, order, sh_norm=sn, sh_phase=sp, layout=layout)
    shape = zest_amd.st_shape(order, layout)
    bt = torch.zeros((zest_amd.st_num_coeffs(order), 2), dtype=torch.float64, device=device)
    plan.backward_adjoint(torch.from_numpy(y.reshape(shape)).to(device), bt, order)
    ft = torch.empty(shape, dtype=torch.float64, device=device)
    plan.forward_adjoint(torch.from_numpy(d.reshape(-1, 2)).to(device), ft, order)
    torch.cuda.synchronize()
    assert _rel(bt.cpu().numpy().ravel(), B.T @ y) < 1e-12
    assert _rel(ft.cpu().numpy().ravel(), F.T @ d) < 1e-12


def _zero_m0_sin(c, order, kind):
    from zest_amd import synth
    m = (synth.zernike_nlm(order)[2] if kind == "zt" else synth.sh_lm(order)[1])
    c[..., m == 0, 1] = 0.0
    return c


@pytest.mark.parametrize("kind,order,precision,tol", [
    ("zt", 64, 0, 1e-12), ("zt", 128, 0, 1e-12), ("st", 200, 0, 1e-12),
    ("zt", 64, 1, 1e-5), ("st", 200, 1, 1e-5)])
def test_inner_product_identities(device, kind, order, precision, tol):
    """<B c, y> = <c, B^T y> and <F g, d> = <g, F^T d> (float64 sums; the FP32
    arithmetic of precision 1 is checked at FP32 accuracy)."""
    import torch
    import zest_amd
    rng = np.random.default_rng(order + 7 * precision)
    if kind == "zt":
        plan = zest_amd.Plan(zest_amd.ZERNIKE, order, precision=precision)
        shape, nc = zest_amd.zt_shape(order), zest_amd.zt_num_coeffs(order)
    else:
        plan = zest_amd.Plan(zest_amd.SH, order, precision=precision)
        shape, nc = zest_amd.st_shape(order), zest_amd.st_num_coeffs(order)
    c = _zero_m0_sin(rng.standard_normal((nc, 2)), order, kind)
    d = _zero_m0_sin(rng.standard_normal((nc, 2)), order, kind)
    y = rng.standard_normal(shape)
    g = rng.standard_normal(shape)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(device)
    Bc = torch.empty(shape, dtype=torch.float64, device=device)
    plan.backward(t(c), Bc, order)
    Bty = torch.zeros((nc, 2), dtype=torch.float64, device=device)
    plan.backward_adjoint(t(y), Bty, order)
    Fg = torch.zeros((nc, 2), dtype=torch.float64, device=device)
    plan.forward(t(g), Fg, order)
    Ftd = torch.empty(shape, dtype=torch.float64, device=device)
    plan.forward_adjoint(t(d), Ftd, order)
    torch.cuda.synchronize()
    lhs1, rhs1 = float((Bc.cpu().numpy() * y).sum()), float((c * Bty.cpu().numpy()).sum())
    lhs2, rhs2 = float((Fg.cpu().numpy() * d).sum()), float((g * Ftd.cpu().numpy()).sum())
    scale1 = np.abs(Bc.cpu().numpy() * y).sum()
    scale2 = np.abs(g * Ftd.cpu().numpy()).sum()
    print(f"{kind} order {order}: <Bc,y> rel {abs(lhs1 - rhs1) / scale1:.2e}, "
          f"<Fg,d> rel {abs(lhs2 - rhs2) / scale2:.2e}")
    assert abs(lhs1 - rhs1) <= tol * scale1
    assert abs(lhs2 - rhs2) <= tol * scale2


def _fit_problem(order, seed, nfit):
    """Targets g = r exp(-(a.x)^2) on the ball grid (x Cartesian, a ~ U(0.5, 2)^3
    per fit), the quadrature weights W_x = w_j r_k^2 v_k, and nu_n (the forward
    prefactor times kappa_n) for conventions (unnormed, geo, none)."""
    import oracle
    from zest_amd import synth
    nphi, nth, nr = oracle.zt_shape(order)
    z, wz = oracle.gl_nodes_weights(nth)
    x, wx = oracle.gl_nodes_weights(nr)
    r = 0.5 * (1.0 + x)
    phi = 2.0 * np.pi * np.arange(nphi) / nphi
    st = np.sqrt(np.maximum(0.0, 1.0 - z * z))
    X = r[None, None, :] * st[None, :, None] * np.cos(phi)[:, None, None]
    Y = r[None, None, :] * st[None, :, None] * np.sin(phi)[:, None, None]
    Z = r[None, None, :] * z[None, :, None] * np.ones(nphi)[:, None, None]
    u = synth.uniform(seed, 3 * nfit).reshape(nfit, 3) * 1.5 + 0.5
    g = np.stack([r[None, None, :] * np.exp(-(a[0] * X + a[1] * Y + a[2] * Z) ** 2) for a in u])
    W = wz[None, :, None] * (r * r * wx)[None, None, :] * np.ones(nphi)[:, None, None]
    n = synth.zernike_nlm(order)[0]
    nu = (1.0 / (4.0 * np.pi)) * 0.5 * (2.0 * np.pi / nphi) * (2.0 * n + 3.0)
    return g, W, nu


def test_fit_gradient_fd_and_zest(device):
    import torch
    import oracle
    import zest_amd
    from zest_amd.autograd import ZernikeBackward
    from zest_amd import synth
    order, nfit = 64, 4
    g, W, nu = _fit_problem(order, 3, nfit)
    plan = zest_amd.Plan(zest_amd.ZERNIKE, order)
    c0 = np.stack([synth.zernike_coeffs(order, 300 + i) for i in range(nfit)]) * 0.1
    dW = torch.from_numpy(W).to(device)
    dg = torch.from_numpy(g).to(device)

    def loss(c):
        r = ZernikeBackward.apply(c, plan, order) - dg
        return (dW * r * r).sum(dim=(1, 2, 3))

    c = torch.from_numpy(c0).to(device).requires_grad_(True)
    L = loss(c)
    L.sum().backward()
    grad = c.grad.detach().cpu().numpy()

    # closed form 2 diag(1/nu) F(B c - g), with compiled zest when it is built
    closed = np.zeros_like(grad)
    for i in range(nfit):
        if oracle.REF is not None:
            z = oracle.ZestZernike(order, 1, 0, -1)
            res = z.backward(c0[i]) - g[i]
            closed[i] = 2.0 * z.forward(res) / nu[:, None]
        else:
            res = oracle.zt_backward(c0[i], order) - g[i]
            closed[i] = 2.0 * oracle.zt_forward(res, order) / nu[:, None]
    m = synth.zernike_nlm(order)[2]
    closed[:, m == 0, 1] = 0.0
    err_closed = _rel(grad, closed)
    print(f"gradient vs closed form ({'zest' if oracle.REF is not None else 'oracle'}): {err_closed:.2e}")
    assert err_closed < 1e-10

    # central differences on a few components of fit 0, h = 1e-6
    rng = np.random.default_rng(5)
    idx = rng.choice(grad.shape[1], 12, replace=False)
    h = 1e-6
    with torch.no_grad():
        for k in idx:
            for part in (0, 1):
                if part == 1 and m[k] == 0:
                    continue
                cp = c.detach().clone()
                cm = c.detach().clone()
                cp[0, k, part] += h
                cm[0, k, part] -= h
                fd = float((loss(cp)[0] - loss(cm)[0]) / (2 * h))
                an = float(grad[0, k, part])
                assert abs(fd - an) <= 1e-6 * max(1.0, abs(an)) + 1e-5 * np.abs(grad[0]).max(), (k, part, fd, an)


def test_autograd_forward_matches_forward_adjoint(device):
    """d/dg <F g, d> = F^T d through the autograd binding."""
    import torch
    import zest_amd
    from zest_amd.autograd import ZernikeForward
    order = 32
    plan = zest_amd.Plan(zest_amd.ZERNIKE, order)
    rng = np.random.default_rng(11)
    g = torch.from_numpy(rng.standard_normal((2,) + zest_amd.zt_shape(order))).to(device).requires_grad_(True)
    d = torch.from_numpy(rng.standard_normal((2, zest_amd.zt_num_coeffs(order), 2))).to(device)
    (ZernikeForward.apply(g, plan, order) * d).sum().backward()
    ft = torch.empty((2,) + zest_amd.zt_shape(order), dtype=torch.float64, device=device)
    plan.forward_adjoint(d, ft, order, batch=2)
    torch.cuda.synchronize()
    assert _rel(g.grad.cpu().numpy(), ft.cpu().numpy()) < 1e-13
