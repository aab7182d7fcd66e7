"""Minimal composition of Hydra-style YAML configs (Hydra/OmegaConf are optional).

Supports what the sysID configs use: ``defaults`` lists (``- name``, ``- group: name``,
``- /group: name``, ``_self_``), ``# @package _global_`` files, recursive merging,
``${a.b.c}`` interpolation, and ``_target_`` class lookup.  Values are returned as
``AttrDict`` (attribute access like OmegaConf's DictConfig).
"""
from __future__ import annotations

import copy
import importlib
import os
import re
from typing import Any, Dict, Iterable, List, Optional

import yaml

PKG_CONFIG_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "configs")


class AttrDict(dict):
    def __getattr__(self, k):
        try:
            return self[k]
        except KeyError as e:
            raise AttributeError(k) from e

    def __setattr__(self, k, v):
        self[k] = v

    def get(self, k, default=None):
        return super().get(k, default)

    @staticmethod
    def wrap(x):
        if isinstance(x, dict) and not isinstance(x, AttrDict):
            return AttrDict({k: AttrDict.wrap(v) for k, v in x.items()})
        if isinstance(x, list):
            return [AttrDict.wrap(v) for v in x]
        return x


def merge(a: Dict, b: Dict) -> Dict:
    out = copy.deepcopy(a)
    for k, v in b.items():
        if isinstance(v, dict) and isinstance(out.get(k), dict):
            out[k] = merge(out[k], v)
        else:
            out[k] = copy.deepcopy(v)
    return out


def _find(config_dir: str, group: Optional[str], name: str, here: str) -> str:
    cands = []
    if group:
        cands.append(os.path.join(config_dir, group.strip("/"), name + ".yaml"))
    cands.append(os.path.join(here, name + ".yaml"))
    cands.append(os.path.join(config_dir, name + ".yaml"))
    for c in cands:
        if os.path.exists(c):
            return c
    raise FileNotFoundError(f"config {group + '/' if group else ''}{name} not found under {config_dir}")


def load_file(config_dir: str, path: str) -> Dict:
    with open(path) as f:
        data = yaml.safe_load(f) or {}
    defaults = data.pop("defaults", []) or []
    here = os.path.dirname(path)
    merged: Dict = {}
    self_done = False
    for d in defaults:
        if d == "_self_":
            merged = merge(merged, data)
            self_done = True
            continue
        if isinstance(d, str):
            group, name = None, d
        else:
            (group, name), = d.items()
            if name is None:
                continue
        merged = merge(merged, load_file(config_dir, _find(config_dir, group, name, here)))
    if not self_done:
        merged = merge(merged, data)
    return merged


_INTERP = re.compile(r"\$\{([^}]+)\}")


def _lookup(root: Dict, dotted: str):
    cur: Any = root
    for part in dotted.split("."):
        cur = cur[part]
    return cur


def resolve(cfg: Dict) -> Dict:
    def walk(x, root):
        if isinstance(x, dict):
            return {k: walk(v, root) for k, v in x.items()}
        if isinstance(x, list):
            return [walk(v, root) for v in x]
        if isinstance(x, str):
            m = _INTERP.fullmatch(x)
            if m and not m.group(1).startswith(("eval:", "oc.", "sqrt:", "sum:")):
                try:
                    return walk(_lookup(root, m.group(1)), root)
                except (KeyError, TypeError):
                    return x
            return _INTERP.sub(lambda mm: str(_safe(root, mm.group(1), mm.group(0))), x) if "${" in x else x
        return x

    def _safe(root, key, default):
        try:
            return _lookup(root, key)
        except (KeyError, TypeError):
            return default

    for _ in range(4):  # nested interpolations
        cfg = walk(cfg, cfg)
    return cfg


def compose(name: str, config_dir: str = PKG_CONFIG_DIR, overrides: Iterable[Dict] = ()) -> AttrDict:
    cfg = load_file(config_dir, _find(config_dir, None, name, config_dir))
    for o in overrides:
        cfg = merge(cfg, o)
    return AttrDict.wrap(resolve(cfg))


def get_class(path: str):
    mod, _, cls = path.rpartition(".")
    return getattr(importlib.import_module(mod), cls)
