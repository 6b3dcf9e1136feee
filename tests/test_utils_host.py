"""Host-side utilities of the reference that need no GPU: get_glx
(src/utilities.jl:417-426), pinned by closed forms."""
from types import SimpleNamespace

import numpy as np


def _model(haf, dx=1000.0):
    nx, ny = haf.shape
    xx = np.asfortranarray(np.broadcast_to(dx * (np.arange(nx)[:, None] + 0.5), (nx, ny)))
    return SimpleNamespace(fields=SimpleNamespace(gh=SimpleNamespace(haf=np.asfortranarray(haf))),
                           grid=SimpleNamespace(xxh=xx))


def test_glx_linear_haf_is_exact_and_column_major(W):
    # haf = X_j - x in each row: the crossing is X_j itself.  Rows 0 and 2
    # cross between cells 1 and 2, row 1 between cells 3 and 4, so
    # column-major order (x fastest, then y) gives X_0, X_1, X_2 while a
    # row-major walk would give X_0, X_2, X_1.
    X = np.array([2300.0, 3700.0, 1800.0])
    xx = 1000.0 * (np.arange(6)[:, None] + 0.5)
    haf = X[None, :] - xx
    glx = W.get_glx(_model(haf))
    np.testing.assert_allclose(glx, X, rtol=1e-14)


def test_glx_several_none_and_exact_zero(W):
    nx, ny = 8, 3
    haf = np.ones((nx, ny))
    # row 0: +, -, +, - : two crossings, at the midpoints of cells 1-2 and 5-6
    haf[:, 0] = [1, 1, -1, -1, 1, 1, -1, -1]
    # row 1 stays grounded: no crossing
    # row 2 touches zero exactly: sign 1 -> 0 -> -1 is never a drop of 2, so
    # the reference reports no grounding line there
    haf[:, 2] = [2, 1, 0, -1, -2, -3, -4, -5]
    glx = W.get_glx(_model(haf))
    np.testing.assert_allclose(glx, [2000.0, 6000.0], rtol=1e-15)
    assert W.get_glx(_model(np.ones((nx, ny)))).size == 0
