"""Device command generators vs the numpy restatement of the reference's loops, and the
Optuna-shaped study over the device CEM."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

from test_commands_oracle import DEFAULT, IDX, RANGES  # noqa: E402


@pytest.mark.parametrize("mode", ["constant", "polynomial", "bezier"])
def test_command_sequences_match_oracle(gpu, mode):
    import spio_commands as oc
    from spi_amd.agents.commands import CommandGenerator
    gen = CommandGenerator(RANGES, DEFAULT, IDX, mode, rollout_length=25, horizon_length=5, dt=0.02)
    assert gen.total_steps == 1250
    names = gen.param_names()
    assert names == oc.param_names(mode, IDX, 5)
    rng = np.random.default_rng(1)
    b = gen.param_bounds()
    P = rng.uniform(b[:, 0], b[:, 1], size=(6, len(names)))
    out = gen.generate(torch.tensor(P, dtype=torch.float32)).cpu().numpy()
    for k in range(6):
        ref = oc.commands_from_params(P[k].astype(np.float32), mode, RANGES, DEFAULT, IDX, 5, 250)
        err = np.abs(out[k] - ref).max()
        assert err < 2e-6, (mode, k, err)
    best = dict(zip(names, P[2]))
    assert np.abs(gen.from_params(best) - out[2]).max() == 0.0


def test_device_study_ask_tell_converges_and_resumes(gpu):
    from spi_amd.agents.study import DeviceCEMSampler, create_study
    target = {"dim_0_update_0": 0.3, "dim_2_update_0": -0.6, "dim_5_update_0": 0.8}
    study = create_study(direction="maximize", sampler=DeviceCEMSampler(batch_size=256, seed=3))
    for it in range(10):
        trials = [study.ask() for _ in range(256)]
        vals = []
        for t in trials:
            x = {k: t.suggest_float(k, -1.0, 1.0) for k in target}
            vals.append(-sum((x[k] - target[k]) ** 2 for k in target))
        for t, v in zip(trials, vals):
            study.tell(t, v)
        if it == 4:
            state = study.state_dict()
    bp = study.best_params
    assert all(abs(bp[k] - target[k]) < 0.05 for k in target), bp
    assert study.best_value > -5e-3
    # resume from iteration 5 and continue: same trajectory as the uninterrupted study
    s2 = create_study(direction="maximize", sampler=DeviceCEMSampler(batch_size=256, seed=3))
    s2.load_state_dict(state)
    assert s2._iteration == 5 and s2._names == list(target)
    t = s2.ask()
    assert set(target) == {n for n in s2._names}
    assert -1.0 <= t.suggest_float("dim_0_update_0", -1.0, 1.0) <= 1.0
