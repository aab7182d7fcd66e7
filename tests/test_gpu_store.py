"""The device segment store packs the same windows as the host loader of the reference's
npz schema (scripts/eval.py:101-171), bit for bit."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_store_windows_equal_host_loader(gpu, tmp_path):
    from spi_amd.data import load_npz_dataset, save_npz_trajectory
    from spi_amd.store import SegmentStore
    rng = np.random.default_rng(0)
    files = []
    for k, (T, kp) in enumerate(((30, 20.0), (17, 25.0))):
        st = rng.normal(size=(T, 37)).astype(np.float32)
        ac = rng.normal(size=(T, 12)).astype(np.float32)
        f = str(tmp_path / f"t{k}.npz")
        save_npz_trajectory(f, st, ac, 0.02, kp, 0.5)
        files.append(f)
    H = 6
    host = load_npz_dataset(files, H)
    store = SegmentStore()
    for f in files:
        store.add_npz(f)
    b = store.windows(H)
    assert np.array_equal(b.segments.cpu().numpy(), host.packed())
    assert np.array_equal(b.motion_ends, host.motion_ends)
    assert np.array_equal(b.kp, host.kp[:, 0]) and b.segments.shape[0] == (30 - H) + (17 - H)
    from spi_amd.data import reference_segment_weights
    assert np.array_equal(b.reference_weights(8).cpu().numpy(), reference_segment_weights(host.motion_ends, 8))


def test_pack_windows_rows_one_launch_and_bounds(gpu):
    """spi_pack_windows_rows: windows from a row table over trajectories held back to back equal
    spi_pack_windows on each trajectory; a window whose rows reach past the buffer (or start
    below 0) comes back as NaN, and the windows around it are unaffected.  A store that gained
    a trajectory after a first call packs the new one too (the joined buffer is rebuilt)."""
    import torch
    from spi_amd import _lib
    from spi_amd.data import segment_stride
    from spi_amd.store import SegmentStore
    rng = np.random.default_rng(3)
    H = 5
    lens = (23, 9, 14)
    store = SegmentStore()
    for T in lens[:2]:
        store.add_arrays(rng.normal(size=(T, 37)).astype(np.float32), rng.normal(size=(T, 12)).astype(np.float32))
    first_call = store.windows(H).segments.shape[0]
    store.add_arrays(rng.normal(size=(lens[2], 37)).astype(np.float32),
                     rng.normal(size=(lens[2], 12)).astype(np.float32))
    b = store.windows(H)
    assert first_call == sum(T - H for T in lens[:2]) and b.segments.shape[0] == sum(T - H for T in lens)
    L = _lib.lib()
    ref = []
    for tr in store.trajectories:
        o = torch.empty((tr.n_steps - H, segment_stride(H)), dtype=torch.float32, device="cuda")
        _lib.check(L.spi_pack_windows(_lib.ptr(tr.states), _lib.ptr(tr.actions), tr.n_steps, H, 0, tr.n_steps - H,
                                      _lib.ptr(o), _lib.stream()), "spi_pack_windows")
        ref.append(o)
    assert torch.equal(b.segments, torch.cat(ref))
    states, actions, _ = store._joined_buffers()
    R = states.shape[0]
    rows = torch.tensor([0, R - H - 1, R - H, -1, 7], dtype=torch.int32, device="cuda")
    out = torch.empty((5, segment_stride(H)), dtype=torch.float32, device="cuda")
    _lib.check(L.spi_pack_windows_rows(_lib.ptr(states), _lib.ptr(actions), R, _lib.ptr(rows, torch.int32), 5, H,
                                       _lib.ptr(out), _lib.stream()), "spi_pack_windows_rows")
    o = out.cpu().numpy()
    assert np.isnan(o[2]).all() and np.isnan(o[3]).all()
    joined = torch.cat(ref).cpu().numpy()
    assert np.array_equal(o[0], joined[0]) and np.array_equal(o[4], joined[7])
    last = R - H - 1                             # last window of the last trajectory
    assert np.array_equal(o[1], joined[-1]) and last == sum(lens) - H - 1
