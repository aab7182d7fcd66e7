"""Optimizers over the device rollout engine (the reference's spigym/agents layer)."""
