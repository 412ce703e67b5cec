"""Host-side logic that needs no GPU: the params.pkl writer of the run() driver."""
import pickle

import numpy as np


def test_lambda_source_and_params_pkl(tmp_path):
    from pypty_amd.iterative import lambda_source, save_params

    params = dict(update_probe=lambda epoch: epoch > 5,
                  epoch_max=7, dataset=np.zeros((2, 2)), output_folder=str(tmp_path))
    assert lambda_source(params["update_probe"]) == "lambda epoch: epoch > 5"
    assert lambda_source(3) == 3
    path = tmp_path / "params.pkl"
    save_params(str(path), params, strip_dataset=True)
    with open(path, "rb") as fh:  # a file this test wrote
        back = pickle.load(fh)
    assert "dataset" not in back
    assert back["update_probe"] == "lambda epoch: epoch > 5"
    assert back["epoch_max"] == 7
    save_params(str(path), params, strip_dataset=False)
    with open(path, "rb") as fh:
        assert "dataset" in pickle.load(fh)


def test_params_pkl_converts_nested_step_lambdas(tmp_path):
    from pypty_amd.iterative import save_params

    path = tmp_path / "params.pkl"
    save_params(str(path), dict(wolfe_c1_constant=[0.1, lambda e: 1 if e < 3 else 0.5]))
    with open(path, "rb") as fh:  # a file this test wrote
        assert pickle.load(fh)["wolfe_c1_constant"] == [0.1, "lambda e: 1 if e < 3 else 0.5"]


def test_get_window_closed_form():
    import numpy as np

    from pypty_amd.utils import get_window

    w = get_window(64, 0.4, 0.8)  # inverted: 0 in the flat centre, 1 beyond r_max
    half = 31.5
    x = np.arange(64) - half
    assert w[32, 32] == 0.0 and w[0, 0] == 1.0
    inside = np.abs(x) <= 0.4 * half
    assert np.all(w[np.ix_(inside, inside)] == 0.0)
    beyond = np.abs(x) > 0.8 * half
    assert np.all(w[beyond, :] == 1.0) and np.all(w[:, beyond] == 1.0)
    assert np.allclose(get_window(64, 0.4, 0.8, inverted=False), 1 - w)
