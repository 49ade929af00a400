"""GPU wizard parity: exhaustive trial sizes and sample statistics against the CPU oracle, MLP forward/backward
against the oracle's rounding model, and what the trained selector gains over the rule without the network."""
import numpy as np
import pytest

import fastlanes_amd as fa
from fastlanes_amd import workloads as wl
from oracle import wizard_oracle as wo

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_cuda():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


@pytest.fixture(scope="module")
def wiz():
    from fastlanes_amd import wizard

    return wizard


def rowgroups():
    out = []
    for fam in wl.WIZARD_FAMILIES:
        for seed in range(3):
            out.append((fam, wl.wizard_column(fam, wo.RG, seed=seed + 20)))
    from tests.test_wizard_oracle import edge_rowgroups

    out += list(edge_rowgroups())
    return out


def test_candidates_match_oracle_pool(wiz):
    assert tuple(wiz.candidates()) == wo.CANDIDATES


def test_exhaustive_sizes_and_features_match_oracle(torch_cuda, wiz):
    torch = torch_cuda
    rgs = rowgroups()
    host = np.concatenate([x for _, x in rgs])
    n = len(rgs)
    d_x = torch.from_numpy(host).cuda()
    K = len(wo.CANDIDATES)
    d_s = torch.zeros((n, K), dtype=torch.int64, device="cuda")
    d_b = torch.zeros(n, dtype=torch.int32, device="cuda")
    d_f = torch.zeros((n, 64), dtype=torch.float32, device="cuda")
    for rep in range(2):  # the second call reuses the hash tables the first one restored
        print(f"wizard sizes call {rep}", flush=True)
        wiz.exhaustive_sizes(d_x.data_ptr(), n, d_s.data_ptr())
        wiz.get_best(d_s.data_ptr(), n, d_b.data_ptr())
        torch.cuda.synchronize()
        sizes, best = d_s.cpu().numpy(), d_b.cpu().numpy()
        for r, (name, x) in enumerate(rgs):
            want = wo.candidate_sizes(x).astype(np.int64)
            np.testing.assert_array_equal(sizes[r], want, err_msg=f"row-group {r} ({name}) call {rep}")
            assert best[r] == wo.best(want), name
    print("wizard features", flush=True)
    wiz.features(d_x.data_ptr(), n, d_f.data_ptr())
    torch.cuda.synchronize()
    feats = d_f.cpu().numpy()
    for r, (name, x) in enumerate(rgs):
        want = wo.features(x)
        np.testing.assert_allclose(feats[r], want, rtol=2e-6, atol=1e-6, err_msg=f"features of {name}")


def _rand_batch(seed, B=64):
    g = np.random.default_rng(seed)
    X = g.standard_normal((B, 64)).astype(np.float32) * 0.5
    y = g.integers(0, 8, B).astype(np.int32)
    return X, y


def test_mlp_forward_backward_matches_oracle(torch_cuda, wiz):
    torch = torch_cuda
    mlp = wiz.MLP(8, max_batch=256, seed=3)
    p = mlp.params()
    W3t = wo.unpack_params(p)[4]
    assert np.all(W3t[8:] == 0)  # padded classes start at zero
    # non-zero biases exercise the epilogue
    g = np.random.default_rng(1)
    p = p + np.concatenate([np.zeros(256 * 64), g.standard_normal(256) * 0.1, np.zeros(256 * 256),
                            g.standard_normal(256) * 0.1, np.zeros(32 * 256),
                            np.r_[g.standard_normal(8) * 0.1, np.zeros(24)]]).astype(np.float32)
    mlp.set_params(p)
    for B in (64, 96):
        X, y = _rand_batch(B, B)
        d_x, d_y = torch.from_numpy(X).cuda(), torch.from_numpy(y).cuda()
        print(f"mlp forward/backward B={B}", flush=True)
        loss = mlp.forward_backward(d_x.data_ptr(), d_y.data_ptr(), B)
        grads = mlp.grads()
        z = mlp.logits(B)
        want_loss, want_g, want_z = wo.mlp_forward_backward(p, X, y, 8)
        np.testing.assert_allclose(z, want_z, rtol=2e-2, atol=2e-3)
        assert abs(loss - want_loss) < 1e-3 * max(1.0, abs(want_loss))
        for name, a, b in zip(("W1t", "b1", "W2t", "b2", "W3t", "b3"), wo.unpack_params(grads),
                              wo.unpack_params(want_g)):
            scale = max(float(np.abs(b).max()), 1e-6)
            err = float(np.abs(a - b).max()) / scale
            assert err < 3e-2, (name, B, err)


def test_mlp_predict_is_argmax_of_logits(torch_cuda, wiz):
    torch = torch_cuda
    mlp = wiz.MLP(8, max_batch=128, seed=5)
    X, _ = _rand_batch(9, 300)  # several max_batch chunks and a ragged tail
    d_x = torch.from_numpy(X).cuda()
    d_c = torch.full((300,), -1, dtype=torch.int32, device="cuda")
    mlp.predict(d_x.data_ptr(), 300, d_c.data_ptr())
    torch.cuda.synchronize()
    got = d_c.cpu().numpy()
    _, _, Z = wo.mlp_forward_backward(mlp.params(), X, np.zeros(300, dtype=np.int64), 8)
    want = np.argmax(Z[:, :8], axis=1)
    top2 = np.sort(Z[:, :8], axis=1)
    clear = (top2[:, -1] - top2[:, -2]) > 1e-2  # ignore near-ties decided by rounding
    assert clear.sum() > 250
    np.testing.assert_array_equal(got[clear], want[clear])


def test_training_gains_over_rule_without_network(torch_cuda, wiz):
    """Train on the exhaustive labels of a small data set; on held-out columns the trained selector must have lower
    size regret than argmin of the sample estimates alone, and its choice decodes bit-exact."""
    torch = torch_cuda
    print("wizard training study", flush=True)
    res, mlp = wiz.run(torch.device("cuda:0"), torch.cuda.current_stream(), n_columns=60, n=1 << 18, steps=600,
                       batch=256, reps=2, return_model=True)
    print(res, flush=True)
    assert res["loss_curve"][-1][1] < res["loss_curve"][0][1]
    assert res["test_regret"] < res["no_network_test_regret"]
    assert res["test_regret"] < 0.05
    # the selector only picks which bit-exact codec runs: held-out row-groups of every family are encoded with the
    # expression the trained network predicts from their features, decoded on the GPU and compared
    for fam in wl.WIZARD_FAMILIES:
        x = wl.wizard_column(fam, wo.RG, seed=77)
        d_x = torch.from_numpy(x).cuda()
        d_f = torch.zeros((1, 64), dtype=torch.float32, device="cuda")
        d_c = torch.full((1,), -1, dtype=torch.int32, device="cuda")
        wiz.features(d_x.data_ptr(), 1, d_f.data_ptr())
        mlp.predict(d_f.data_ptr(), 1, d_c.data_ptr())
        torch.cuda.synchronize()
        k = int(d_c.cpu().numpy()[0])
        assert 0 <= k < len(wo.CANDIDATES)
        eid = wo.CANDIDATES[k]
        print(f"learned choice for {fam}: expression {eid}", flush=True)
        enc = fa.encode_from_memory(x, eid)
        assert enc.size == int(wo.candidate_sizes(x)[k])
        plan = fa.Plan(enc, np.int64)
        d_enc = torch.from_numpy(np.concatenate([enc, np.zeros(16, np.uint8)])).cuda()
        d_out = torch.empty(x.nbytes, dtype=torch.uint8, device="cuda")
        plan.decode(d_enc.data_ptr(), d_out.data_ptr(), x.nbytes)
        plan.check()
        np.testing.assert_array_equal(d_out.cpu().numpy().view(np.int64), x, err_msg=f"{fam}, expression {eid}")


def test_exhaustive_sizes_many_rowgroups_per_block(torch_cuda, wiz):
    """More row-groups than the launch has blocks (one per compute unit): each block sizes several row-groups in a
    row on one hash table, restored between them; every row-group's sizes equal the oracle's."""
    torch = torch_cuda
    rgs = rowgroups()
    n_cu = torch.cuda.get_device_properties(0).multi_processor_count
    n = 2 * n_cu + 7
    order = [(3 * i) % len(rgs) for i in range(n)]  # neighbours differ: a block meets different families in turn
    host = np.concatenate([rgs[j][1] for j in order])
    want = {j: wo.candidate_sizes(rgs[j][1]).astype(np.int64) for j in set(order)}
    d_x = torch.from_numpy(host).cuda()
    d_s = torch.zeros((n, len(wo.CANDIDATES)), dtype=torch.int64, device="cuda")
    print(f"wizard sizes over {n} row-groups", flush=True)
    wiz.exhaustive_sizes(d_x.data_ptr(), n, d_s.data_ptr())
    torch.cuda.synchronize()
    sizes = d_s.cpu().numpy()
    for r, j in enumerate(order):
        np.testing.assert_array_equal(sizes[r], want[j], err_msg=f"row-group {r} ({rgs[j][0]})")
