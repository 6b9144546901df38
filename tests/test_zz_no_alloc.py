"""h4h_update_acc_jrk_static, the update with no context and no allocation
(its work buffers are one static device array), against
h4h_update_acc_jrk on a context (Hermite4CPU::update_acc_jrk,
cpu/Hermite4CPU.cpp:309-336): the same rows, bit for bit, for every
precision, on two calls in a row of which the second is smaller (the static
array keeps the first call's contents)."""
import numpy as np
import pytest

from gravidy_amd import Hermite4HIP, _lib, ic, softening_e2

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

SENTINEL = 7.25  # rows the call must leave alone


def _system(n, seed):
    r, v = ic.plummer(n, seed)
    p6 = np.ascontiguousarray(np.concatenate([r[:, :3], v[:, :3]], axis=1))
    pm = r[:, 3].astype(np.float32)
    return r, v, p6, pm


def _static(p6, pm, move, e2, precision):
    """One h4h_update_acc_jrk_static call on device copies; returns f6."""
    lib = _lib.load()
    n = p6.shape[0]
    dev = torch.device("cuda:0")
    d_p6 = torch.from_numpy(p6).to(dev)
    d_pm = torch.from_numpy(pm).to(dev)
    d_move = torch.from_numpy(np.ascontiguousarray(move, np.uint32).view(np.int32)).to(dev)
    d_f6 = torch.full((n, 6), SENTINEL, dtype=torch.float64, device=dev)
    stream = torch.cuda.current_stream(dev).cuda_stream
    torch.cuda.synchronize()
    print(f"static call n={n} nact={move.size} precision={precision}", flush=True)
    rc = lib.h4h_update_acc_jrk_static(d_p6.data_ptr(), d_pm.data_ptr(), n, d_move.data_ptr(),
                                       int(move.size), float(e2), _lib.PRECISIONS[precision],
                                       d_f6.data_ptr(), stream)
    _lib.check("h4h_update_acc_jrk_static", rc)
    torch.cuda.synchronize()
    return d_f6.cpu().numpy()


def _context(r, v, p6, pm, move, e2, precision):
    with Hermite4HIP(r, v, e2, precision=precision) as h:
        h.upload_predictor(p6, pm)
        h.set_active(move)
        h.update_acc_jrk(move.size)
        return h.download()["f"]


def _check(r, v, p6, pm, move, e2, precision):
    f_static = _static(p6, pm, move, e2, precision)
    f_ctx = _context(r, v, p6, pm, move, e2, precision)
    act = np.zeros(p6.shape[0], bool)
    act[move] = True
    assert np.isfinite(f_static[act]).all()
    assert f_static[act].tobytes() == f_ctx[act].tobytes()
    assert (f_static[~act] == SENTINEL).all()


@pytest.mark.parametrize("precision", ["fp64", "mixed", "exact"])
def test_static_update_matches_context_twice(precision):
    e2 = softening_e2(1e-4)
    rng = np.random.default_rng(11)
    # the largest system and the active count with the most j-split partials
    # of its plan, then a smaller system with a small active set
    for n, nact, seed in ((65536, 65281, 3), (5000, 17, 4)):
        if precision == "exact" and nact > 4096:
            nact = 4096  # one lane per particle over all j: keep it short
        r, v, p6, pm = _system(n, seed)
        move = rng.permutation(n)[:nact].astype(np.uint32)
        _check(r, v, p6, pm, move, e2, precision)


def test_static_update_small_active_sets():
    """Active counts around the lane-group widths and the j-split limits, on
    a ragged n (the last j tile padded)."""
    e2 = softening_e2(1e-4)
    r, v, p6, pm = _system(3001, 9)
    rng = np.random.default_rng(5)
    for nact in (1, 4, 5, 64, 300, 3001):
        move = np.sort(rng.permutation(3001)[:nact]).astype(np.uint32)
        _check(r, v, p6, pm, move, e2, "fp64")


def test_static_update_flags_bad_indices():
    """An index >= n is caught on the device: the call succeeds and every
    row it can address is NaN."""
    e2 = softening_e2(1e-4)
    _, _, p6, pm = _system(1000, 2)
    move = np.array([3, 1000, 999, 4000000000], np.uint32)
    f = _static(p6, pm, move, e2, "fp64")
    assert np.isnan(f[[3, 999]]).all()
    rest = np.ones(1000, bool)
    rest[[3, 999]] = False
    assert (f[rest] == SENTINEL).all()
    # the next call with valid indices is clean again
    move = np.array([3, 999], np.uint32)
    f = _static(p6, pm, move, e2, "fp64")
    assert np.isfinite(f[[3, 999]]).all()
