"""The trained correction against the method without it: a network trained
on phantom / plain-FDK pairs made by the pipeline lowers the RMSE against the
phantom on held-out phantoms, and the plain-FDK fallback takes over where the
correction is implausibly large."""
import pytest
import torch

from analyticalreconbigpatient_amd import cnn, learned

pytestmark = pytest.mark.gpu

SPEC = learned.PairSpec(nx=96, nz=32, dx=1.0, radius=44.0)


@pytest.fixture(scope="module")
def trained():
    train_pairs = [learned.make_pair(1000 + i, SPEC) for i in range(12)]
    held_out = [learned.make_pair(2000 + i, SPEC) for i in range(3)]
    net = cnn.Net(seed=0)
    hist = learned.train(net, train_pairs, steps=250, batch=4, patch=32, lr=1e-3)
    return net, train_pairs, held_out, hist


def test_pairs_are_plain_fdk_and_phantom(gpu, trained):
    _, pairs, _, _ = trained
    x0, t, pitch = pairs[0]
    assert x0.shape == t.shape == (96, 96, 32) and 0.5 <= pitch <= 1.5
    m = learned.fov_mask(SPEC, x0.device)
    err = float(((x0 - t)[m] ** 2).mean().sqrt()) * 1000
    assert 5 < err < 500                        # a reconstruction, not a copy


def test_trained_network_beats_plain_fdk_on_held_out_phantoms(gpu, trained):
    net, _, held_out, hist = trained
    print("loss", hist[0], hist[-1], flush=True)
    assert hist[-1][1] < hist[0][1]
    r = learned.evaluate(net, held_out, SPEC, bound=10.0)
    print(r, flush=True)
    assert r["rmse_cnn_hu"] < r["rmse_fdk_hu"]
    assert r["reduction"] > 0.05


def test_fallback_keeps_plain_fdk_when_the_correction_is_too_large(gpu, trained):
    net, _, held_out, _ = trained
    x0 = held_out[0][0][None]
    y, frac = net.correct(x0, bound=0.0)        # every block exceeds a zero bound
    assert frac == 1.0 and torch.equal(y, x0)


def test_clinical_crop_pairs_are_plain_fdk_and_phantom(gpu):
    """A crop of a random body on the 888x64 scanner: the plain reconstruction
    differs from the phantom by a reconstruction's error, not by a copy's or a
    misplaced grid's."""
    spec = learned.clinical_spec(nx=64, nz=32)
    x0, t, pitch = learned.make_pair(3000, spec)
    assert x0.shape == t.shape == (64, 64, 32) and 0.5 <= pitch <= 1.5
    err = float(((x0 - t) ** 2).mean().sqrt()) * 1000
    print(f"888x64 crop: RMSE vs phantom {err:.1f} HU", flush=True)
    assert 5 < err < 500
    assert float(t.abs().max()) > 0.05                    # structure inside the body
