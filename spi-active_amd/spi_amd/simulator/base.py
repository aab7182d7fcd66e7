"""The simulator interface the sysID envs program against.

Same method names, arguments and meaning as spigym/simulator/base_simulator/base_simulator.py:6-171
(the drop-in boundary): the env constructs ``SimulatorClass(config=..., device=...)`` from
``config.simulator._target_`` (spigym/envs/base_task/base_task.py:27-28) and then calls
these methods.
"""


class BaseSimulator:
    def __init__(self, config, device):
        self.config = config
        self.sim_device = device
        self.headless = False

    def set_headless(self, headless):
        self.headless = headless

    def setup(self):
        raise NotImplementedError

    def setup_terrain(self, mesh_type):
        raise NotImplementedError

    def load_assets(self):
        raise NotImplementedError

    def create_envs(self, num_envs, env_origins, base_init_state):
        raise NotImplementedError

    def get_dof_limits_properties(self):
        raise NotImplementedError

    def find_rigid_body_indice(self, body_name):
        raise NotImplementedError

    def prepare_sim(self):
        raise NotImplementedError

    def refresh_sim_tensors(self):
        raise NotImplementedError

    def apply_torques_at_dof(self, torques):
        raise NotImplementedError

    def set_actor_root_state_tensor(self, set_env_ids, root_states):
        raise NotImplementedError

    def set_dof_state_tensor(self, set_env_ids, dof_states):
        raise NotImplementedError

    def simulate_at_each_physics_step(self):
        raise NotImplementedError

    def setup_viewer(self):
        raise NotImplementedError

    def render(self, sync_frame_time=True):
        raise NotImplementedError
