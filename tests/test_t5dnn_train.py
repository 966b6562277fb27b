"""T5 DNN training support on the CPU: parameter files, labels, and the plain
PyTorch restatement of the network checked against the oracle's forward pass
(SDL/NeuralNetwork.h:116-162)."""
import ctypes as C

import numpy as np

from tracklooper_amd import t5dnn


def test_parameter_files_round_trip(tmp_path):
    p = t5dnn.load_params()
    assert p.shape == (t5dnn.N_PARAMS,) and p.dtype == np.float32
    assert t5dnn.N_PARAMS == 2337
    t5dnn.write_params(p, str(tmp_path))
    np.testing.assert_array_equal(t5dnn.load_params(str(tmp_path)).view(np.uint32), p.view(np.uint32))
    q = p.copy()
    q[1000] = np.float32(-1.25e-7)
    t5dnn.write_params(q, str(tmp_path))
    assert t5dnn.load_params(str(tmp_path))[1000] == np.float32(-1.25e-7)


def test_labels_need_all_ten_hits_from_one_track():
    trk = np.array([3, 3, 3, 3, 3, 3, 3, 3, 3, 3, 4, -1])
    rows = np.array([list(range(10)), list(range(1, 11)), [0] * 9 + [11], [11] * 10])
    np.testing.assert_array_equal(t5dnn.labels(rows, trk), [1, 0, 0, 0])


def test_reference_logits_match_oracle_forward():
    import torch
    from oracle import orc
    rng = np.random.default_rng(3)
    p = t5dnn.load_params()
    x = rng.normal(0.0, 1.5, (64, 38)).astype(np.float32)
    logit = t5dnn.reference_logits(torch.as_tensor(x, dtype=torch.float64),
                                   torch.as_tensor(p, dtype=torch.float64)).numpy()
    want = 1.0 / (1.0 + np.exp(-logit))
    got = np.array([orc.lib().orc_t5dnn_forward(np.ascontiguousarray(r).ctypes.data_as(C.c_void_p)) for r in x])
    np.testing.assert_allclose(got, want, rtol=2e-5, atol=2e-6)
