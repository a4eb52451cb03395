"""The learned screening path on the GPU: the committed surrogate weights screen seeded
candidates; the batched FDTD verifies the surrogate's picks and an equal number of randomly
chosen designs. The network must pick designs whose verified -10 dB bandwidth beats the random
picks (what the method gains over screening without it), and its S11 must track the FDTD."""
import numpy as np
import pytest

from em_lab_amd import surrogate as S

pytestmark = pytest.mark.gpu

N_CAND = 4096
TOP = 32


def test_surrogate_picks_beat_random_designs():
    print("case: surrogate screening", N_CAND, "candidates, verify", TOP, "+", TOP, flush=True)
    geoms, sizes, X = S.candidates(N_CAND)
    m = S.load_surrogate(max_batch=N_CAND)
    idx, s_pred, bw_pred = S.screen(m, X, TOP)
    m.close()
    rng = np.random.default_rng(0)
    rand = rng.choice(N_CAND, TOP, replace=False)
    both = np.concatenate([idx, rand])
    s_fdtd, _, res = S.verify(geoms, sizes, both)
    assert np.all(res < 1e-2)  # rang down
    nu = S.sweep_nu()
    bw = S.bandwidth_10db(s_fdtd, nu)
    bw_top, bw_rand = bw[:TOP], bw[TOP:]
    print("verified bandwidth: surrogate picks", bw_top.mean(), "random", bw_rand.mean(), flush=True)
    assert bw_top.mean() > 1.5 * bw_rand.mean()
    assert np.mean(bw_top > 0) >= 0.9
    err = S.s11_error_db(s_pred[both], s_fdtd)
    print("S11 error (dB): median", np.median(err), "max", err.max(), flush=True)
    assert np.median(err) < 2.0
