"""The denoiser's training loop on the CPU (no network: a stand-in records
what train() hands it)."""
import pytest


@pytest.mark.parametrize("halves", [False, True])
def test_training_batches_have_the_network_layout(halves):
    """denoise_train.train on a CPU stand-in of a crop dataset: every batch
    handed to the network is [2][batch][h][w][3] beauty for a halves network
    ([batch][h][w][3] otherwise) with [batch][h][w][3] albedo, normal and
    target, each element at its own input level."""
    import torch
    from gartus_amd import denoise_train
    n, L, s = 5, 3, 8
    data = {k: torch.zeros((n, L, s, s, 3)) for k in ("albedo", "normal")}
    data["beauty"] = torch.zeros((n, L, 2, s, s, 3) if halves else (n, L, s, s, 3))
    data["target"] = torch.zeros((n, s, s, 3))
    for i in range(n):
        for j in range(L):
            data["albedo"][i, j] = 10 * i + j  # marks the element and its level
            data["beauty"][i, j] = 10 * i + j
        data["target"][i] = 10 * i

    class Stub:
        shapes = []

        def train_step(self, b, a, nr, t, lr):
            self.shapes.append((tuple(b.shape), tuple(a.shape), tuple(t.shape)))
            for e in range(a.shape[0]):  # beauty, albedo and target of an element belong together
                tag = float(a[e, 0, 0, 0])
                assert float((b[:, e] if halves else b[e])[..., 0, 0, 0].max()) == tag
                assert float(t[e, 0, 0, 0]) == 10 * (tag // 10)
            return 0.0
    stub = Stub()
    denoise_train.train(stub, data, steps=4, batch=3, log=lambda m: None)
    want_b = (2, 3, s, s, 3) if halves else (3, s, s, 3)
    assert stub.shapes == [(want_b, (3, s, s, 3), (3, s, s, 3))] * 4
