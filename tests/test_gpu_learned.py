"""Learned preconditioner on the GPU: network kernels against the CPU oracle, training,
the gain over plain MGPCG and the fallback (needs an MI355X).

Tolerances (BF16 storage of activations and weights, FP32 accumulation on the device, the
oracle in float64 with BF16 rounding at the same points):
  * forward output: 5e-3 relative L2; brick scales 1e-5 relative
  * loss 1e-3 relative; weight/bias gradients 2e-2 relative L2 per layer
  * preconditioner correction z - M_mg(r): 1e-2 relative L2
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def rel_l2(a, ref):
    return float(np.linalg.norm(a - ref) / max(np.linalg.norm(ref), 1e-300))


@pytest.fixture(scope="module")
def lib():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from paracfd_amd import _lib as L
    return L.load()


def sphere(cells_x=64, levels=2):
    from paracfd_amd import _lib as L, cases
    g, d = cases.sphere_box(cells_x=cells_x, levels=levels)
    return L.Topology(g, d)


def learned_core(topo, **kw):
    from paracfd_amd import ExternalAeroCore, _lib as L, default_options
    base = dict(rho=1.0, nu=1e-3, speed=1.0, cfl=0.5, rtol=1e-8, max_iterations=200, preconditioner=L.PC_LEARNED,
                field_precision=L.FP32, vcycle_precision=L.FP32)
    base.update(kw)
    core = ExternalAeroCore(topo, default_options(**base))
    core.set_reference(topo.info.reference_area)
    return core


def random_weights(seed, scale=0.2):
    from oracle import nn_oracle as N
    rng = np.random.default_rng(seed)
    W = rng.normal(0.0, scale, size=(N.SETS, 4, 27, N.C, N.C)).astype(np.float32)
    B = rng.normal(0.0, 0.05, size=(N.SETS, 4, N.C)).astype(np.float32)
    return W, B


def network_bricks(topo, mode):
    """The bricks learned_bricks = mode selects: bricks holding special cells (1) or cut cells
    (2), and their face neighbours."""
    nbr = topo.array("brick_nbr").reshape(-1, 6)
    sd, sp = topo.array("special_dof"), topo.array("special_patch")
    hit = np.unique(sd[sp >= 0] // 512) if mode == 2 else np.unique(sd // 512)
    sel = np.zeros(topo.ndof // 512, dtype=bool)
    sel[hit] = True
    for b in hit:
        sel[nbr[b][nbr[b] >= 0]] = True
    return np.nonzero(sel)[0]


def test_forward_matches_oracle(lib):
    from oracle import nn_oracle as N
    topo = sphere()
    print(f"[learned] forward parity, dofs={topo.ndof}", flush=True)
    core = learned_core(topo)
    W, B = random_weights(1)
    core.set_learned_weights(W, B)
    rng = np.random.default_rng(2)
    r = rng.normal(size=topo.ndof)
    y, s = core.learned_forward(r)
    mask = topo.array("mask")
    y_ref, s_ref = N.network_output(W.astype(np.float64), B.astype(np.float64), mask, r, topo.array("brick_nbr"),
                                    topo.array("brick_level"))
    assert rel_l2(s, s_ref) < 1e-5
    assert np.all(y[(mask & N.ACTIVE) == 0] == 0.0)
    err = rel_l2(y, y_ref)
    print(f"[learned] forward rel L2 {err:.2e}", flush=True)
    assert err < 5e-3


@pytest.mark.parametrize("mode", [0, 2])
def test_gradient_matches_oracle(lib, oracle_built, mode):
    """Loss and every weight-set / layer gradient of the training kernels against the oracle,
    over every brick (mode 0) and over the cut-cell bricks with their neighbours (mode 2: the
    halo of a brick at the set's edge is zero where the neighbour is outside the set)."""
    from oracle import nn_oracle as N
    from oracle.oracle import Oracle
    topo = sphere()
    print(f"[learned] gradient parity, dofs={topo.ndof}, learned_bricks={mode}", flush=True)
    core = learned_core(topo, learned_bricks=mode)
    orc = Oracle(topo, core.options)
    W, B = random_weights(3)
    core.set_learned_weights(W, B)
    mask = topo.array("mask")
    rng = np.random.default_rng(4)
    rp = rng.normal(size=topo.ndof) * ((mask & N.ACTIVE) != 0)
    loss, gw, gb = core.learned_gradient(rp)
    bricks = None if mode == 0 else network_bricks(topo, mode)
    args = (mask, rp, lambda x: orc.mg_primitive(0, 0, x), topo.array("brick_nbr"), topo.array("level_h"),
            topo.array("brick_level"), bricks)
    lref, dW, dB = N.energy_loss_and_grad(W.astype(np.float64), B.astype(np.float64), *args)
    print(f"[learned] loss {loss:.6e} oracle {lref:.6e}", flush=True)
    assert loss == pytest.approx(lref, rel=1e-3)
    sets = np.unique(N.weight_set(topo.array("brick_level")[bricks if bricks is not None else slice(None)]))
    for st in range(N.SETS):
        if st not in sets:
            assert not gw[st].any() and not gb[st].any()
            continue
        for l in range(4):
            ew, eb = rel_l2(gw[st, l], dW[st, l]), rel_l2(gb[st, l], dB[st, l])
            print(f"[learned] set {st} layer {l}: dW rel L2 {ew:.2e}, db rel L2 {eb:.2e}", flush=True)
            assert ew < 2e-2 and eb < 2e-2
    # the float64 network without BF16 rounding is close as well
    _, dW64, _ = N.energy_loss_and_grad(W.astype(np.float64), B.astype(np.float64), *args, bf16_round=False)
    for st in sets:
        for l in range(4):
            assert rel_l2(gw[st, l], dW64[st, l]) < 1e-1
    # the gradient pass leaves the weights unchanged
    W2, B2 = core.learned_weights()
    assert np.array_equal(W2, W) and np.array_equal(B2, B)


def test_preconditioner_matches_oracle(lib, oracle_built):
    """z = M_mg(r) + (s_b/h_b) N(r - A M_mg(r)) through the solver's own preconditioner path
    (V-cycle kernels, in-kernel post-V-cycle residual, network, correction)."""
    from oracle import nn_oracle as N
    from oracle.oracle import Oracle
    from paracfd_amd import _lib as L
    topo = sphere()
    print(f"[learned] preconditioner parity, dofs={topo.ndof}", flush=True)
    core = learned_core(topo, vcycle_precision=L.FP64, field_precision=L.FP64)
    orc = Oracle(topo, core.options)
    W, B = random_weights(5, scale=0.1)
    core.set_learned_weights(W, B)
    mask = topo.array("mask")
    rng = np.random.default_rng(6)
    r = rng.normal(size=topo.ndof) * ((mask & N.ACTIVE) != 0)
    z = core.mg_primitive(3, 0, b=r)
    zmg = orc.mg_primitive(3, 0, b=r)
    rp = r - orc.mg_primitive(0, 0, zmg)
    dz, _ = N.correction(W.astype(np.float64), B.astype(np.float64), mask, rp, topo.array("brick_nbr"),
                         topo.array("level_h"), topo.array("brick_level"))
    err = rel_l2(z - zmg, dz)
    print(f"[learned] correction rel L2 {err:.2e}, |dz|/|z| {np.linalg.norm(dz) / np.linalg.norm(zmg):.2e}", flush=True)
    assert err < 1e-2
    core.learned_enable(False)
    assert rel_l2(core.mg_primitive(3, 0, b=r), zmg) < 1e-12


def trained_sphere_core(topo):
    core = learned_core(topo, max_iterations=300)
    core.initialize()
    for _ in range(2):
        core.step()
    for _ in range(3):
        st = core.learned_collect(max_pairs=12)
        core.step()
    assert st["pairs"] > 0
    tr = core.learned_train(steps=200, batch=4, lr=2e-3)
    print(f"[learned] {st['pairs']} residuals, loss {tr['loss_first']:.4f} -> {tr['loss_last']:.4f} "
          f"in {tr['train_ms']:.0f} ms", flush=True)
    return core, tr


def test_training_reduces_energy_loss(lib):
    topo = sphere(cells_x=32, levels=2)
    print(f"[learned] training, dofs={topo.ndof}", flush=True)
    _, tr = trained_sphere_core(topo)
    # the untrained network adds nothing (loss 0); training removes error energy (loss < 0)
    assert tr["loss_first"] == pytest.approx(0.0, abs=1e-12)
    assert np.isfinite(tr["loss_last"]) and tr["loss_last"] < -0.2


def test_learned_preconditioner_gains_over_plain_mgpcg(lib):
    """After training on the solver's own residuals, the learned flexible PCG needs fewer
    iterations than plain MGPCG on the next projections, to the same residual, with the same
    lift and drag."""
    topo = sphere(cells_x=32, levels=2)
    core, _ = trained_sphere_core(topo)
    v = core.download_cell_momentum_state()
    f = core.download_faces()
    p = core.download_pressure()
    result = {}
    for on in (True, False):
        core.upload_state(*v)
        core.upload_faces(*f)
        core.upload_pressure(p)
        core.mark_initialized(0.0)
        core.learned_enable(on)
        its, loads = 0, None
        for _ in range(4):
            st = core.step()
            assert st["converged"] == 1 and st["relative_residual"] <= 1e-8 and st["fallback_used"] == 0
            its += st["iterations"]
            loads = core.pressure_loads()
        result[on] = (its, loads)
        print(f"[learned] learned={on}: {its} iterations over 4 steps, CD {loads['cd']:.8f}", flush=True)
    assert result[True][0] < result[False][0]
    for key in ("cd", "cl"):
        a, b = result[True][1][key], result[False][1][key]
        assert abs(a - b) <= 1e-3 * max(abs(b), 1e-3)


def test_fallback_on_broken_network(lib):
    """A network whose output overflows makes the learned solve break down; the projection
    restarts with plain MGPCG, converges and reports the fallback."""
    topo = sphere(cells_x=32, levels=1)
    print(f"[learned] fallback, dofs={topo.ndof}", flush=True)
    core = learned_core(topo)
    core.initialize()
    from oracle import nn_oracle as N
    W = np.full((N.SETS, 4, 27, N.C, N.C), 1e30, dtype=np.float32)
    B = np.zeros((N.SETS, 4, N.C), dtype=np.float32)
    core.set_learned_weights(W, B)
    st = core.step()
    assert st["fallback_used"] == 1
    assert st["converged"] == 1 and st["relative_residual"] <= 1e-8
    core.learned_enable(False)
    st = core.step()
    assert st["fallback_used"] == 0 and st["converged"] == 1


def test_validation_rejects_a_network_that_does_not_help(lib):
    """After training, plain and learned projections alternate (learned first); a network whose
    correction is negligible (learning rate 1e-30 leaves the zero output layer at ~0) does the
    same iterations plus its own cost, so the validation switches it off."""
    topo = sphere(cells_x=32, levels=2)
    print(f"[learned] validation, dofs={topo.ndof}", flush=True)
    core = learned_core(topo)
    core.initialize()
    core.step()
    core.learned_collect(max_pairs=8)
    core.step()
    core.learned_train(steps=2, batch=1, lr=1e-30)
    used, its = [], []
    for _ in range(6):
        st = core.step()
        used.append(st["learned_used"])
        its.append(st["iterations"])
        assert st["converged"] == 1
    assert used == [1, 0, 1, 0, 1, 0]
    print(f"[learned] validation iterations {its}", flush=True)
    assert st["learned_rejected"] == 1
    st = core.step()
    assert st["learned_used"] == 0 and st["learned_rejected"] == 1
