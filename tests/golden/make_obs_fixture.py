"""Writes go2_omni_actor_obs.json: the actor observation layout of the go2_omni policy as the
reference's config files state it (config/obs/loco/go2_omni.yaml obs_dict.actor_obs,
obs_auxiliary.short_history, obs_scales, obs_dims; config/robot/go2/go2.yaml dof_obs_size;
config/env/legged_base.yaml clip_observations).  Reads YAML data only.

    python tests/golden/make_obs_fixture.py /root/reference/spigym
"""
import json
import os
import sys

import yaml


def main(root: str) -> None:
    with open(os.path.join(root, "config/obs/loco/go2_omni.yaml")) as f:
        obs = yaml.safe_load(f)["obs"]
    with open(os.path.join(root, "config/robot/go2/go2.yaml")) as f:
        dof_obs = int(yaml.safe_load(f)["robot"]["dof_obs_size"])
    with open(os.path.join(root, "config/env/legged_base.yaml")) as f:
        env = yaml.safe_load(f)
    clip = float(env["env"]["config"]["normalization"]["clip_observations"])
    dims = {}
    for d in obs["obs_dims"]:
        for k, v in d.items():
            dims[k] = dof_obs if isinstance(v, str) and "dof_obs_size" in v else int(v)
    keys = [k for k in obs["obs_dict"]["actor_obs"]]
    hist = obs["obs_auxiliary"]["short_history"]
    out = {"actor_obs": keys, "short_history": {k: int(v) for k, v in hist.items()},
           "scales": {k: float(obs["obs_scales"][k]) for k in keys if k != "short_history"},
           "dims": {k: dims[k] for k in keys if k != "short_history"}, "clip_observations": clip}
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "go2_omni_actor_obs.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)
    print("wrote", path)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/spigym")
