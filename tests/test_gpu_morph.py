"""Device morphology (tc_morph_device): the binary image it produces equals
the host morphology's (itself pinned to the oracle's steps) pixel for pixel
on random thumbnails, nested holes, a maze of long flood paths, a
checkerboard of tiny components, uniform images and non-square widths."""
import numpy as np
import pytest
import torch

import taichi_amd as T

from helpers import morph_images

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("case", morph_images(), ids=lambda c: c[0])
def test_device_morph_equals_host(case, dev):
    _, a = case
    H, W = a.shape
    want = T.morph_from_image(a)
    d = torch.from_numpy(np.ascontiguousarray(a).ravel()).to(dev)
    for _ in range(2):  # a second call gives the same image
        got = T.morph_device(d, H, W).cpu().numpy().reshape(H, W)
        assert np.array_equal(got, want)
