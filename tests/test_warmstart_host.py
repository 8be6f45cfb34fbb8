"""CPU tests of the warm-start host logic: training labels, the network's
input features and the candidate-start score."""

import numpy as np
import torch

from lpe import warmstart
from lpe.workloads import consortium


def test_bound_labels_ignore_infinite_bounds():
    """A free column (both bounds infinite) is never labelled at a bound; a
    one-sided column only at its finite bound."""
    inf = float("inf")
    x = torch.tensor([[0.0, 5.0, -3.0, 2.0, 2.0]], dtype=torch.float64)
    lo = torch.tensor([[-inf, -inf, -3.0, 0.0, 2.0]], dtype=torch.float64)
    up = torch.tensor([[inf, 5.0, inf, 4.0, 4.0]], dtype=torch.float64)
    np.testing.assert_array_equal(warmstart.bound_labels(x, lo, up).numpy(), [[0, 1, 1, 0, 1]])


def test_consortium_features_are_240_and_carry_biomass():
    md = consortium(2)
    B = 3
    conc = torch.rand(B, md.n_medium, dtype=torch.float64) * 5 + 0.1
    lower = torch.as_tensor(md.lower).expand(B, -1).clone()
    lower[:, md.term_cols] = -torch.rand(B, len(md.term_cols), dtype=torch.float64) * 10 - 0.5
    bio = torch.tensor([[0.1, 0.2, 0.3], [1.0, 2.0, 3.0], [0.5, 0.5, 0.5]], dtype=torch.float64)
    f = warmstart.consortium_features(md, conc, lower, bio)
    assert f.shape == (B, 240) and f.dtype == torch.float32
    s_idx, _ = md.terms()
    # the biomass of each species is recovered from a term's scaled bound and the bound itself
    rec = f[:, 60:].double() / lower[:, md.term_cols]
    np.testing.assert_allclose(rec.numpy(), bio[:, s_idx].numpy(), rtol=1e-6)


def test_start_score_gates_and_nonfinite():
    m = torch.zeros(4, 8, dtype=torch.float64)
    m[:, 5] = 1.0
    m[0, 0] = 1e-5           # exactly at the primal gate
    m[1, 1] = 3e-7           # three times the dual gate
    m[2, 2] = 5e-8           # half the gap gate
    m[3, 5] = 0.0            # non-finite candidate
    np.testing.assert_allclose(warmstart.start_score(m)[:3].numpy(), [1.0, 3.0, 0.5])
    assert warmstart.start_score(m)[3] == float("inf")
