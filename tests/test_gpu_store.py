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
