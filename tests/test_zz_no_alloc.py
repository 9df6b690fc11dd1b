"""gmk_item_predict_static, the entry point that allocates nothing: its scores
equal gmk_item_predict's (the same kernel; the model staged in the library's
static device array instead of an allocation) on two calls in a row, the
second smaller, and a count above the array's capacity is refused."""
import numpy as np
import pytest

from gmkbe_amd import engine, gen
from gmkbe_amd.engine import Engine, GmkError

pytestmark = pytest.mark.gpu


def test_static_predict_matches_allocating_predict():
    g = gen.chung_lu(90_000, 60_000, 400_000, 2.1, 31)
    with Engine(0) as e:
        e.set_graph(g)
        X = e.item_features(0, 4096, 2)
        model, _ = e.mlp_train(X, (X[:, 0] + np.sin(X[:, 2])).astype(np.float32), epochs=60, lr=1e-2, seed=9)
        e.set_item_model(model)
        for first, count in ((0, engine.STATIC_ITEMS), (g.n_left + 1000, 777)):
            print(f"[gpu] static predict {first}+{count}", flush=True)
            got = e.item_predict_static(first, count, 2)
            want = e.item_predict(first, count, 2)
            np.testing.assert_array_equal(got, want)
            assert np.isfinite(got).all() and np.unique(got).size > 1
        with pytest.raises(GmkError, match="at most 65536"):
            e.item_predict_static(0, engine.STATIC_ITEMS + 1, 2)
